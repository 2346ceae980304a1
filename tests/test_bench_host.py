"""Host-side behaviour of bench.py that needs no GPU: a request for more than one GPU ends at once
with an explanation and a non-zero status (the hot path runs on one GPU; sharding is design only)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "single GPU" in (r.stdout + r.stderr)
    assert '"metric"' not in r.stdout   # no result line
