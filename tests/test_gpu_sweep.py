"""The 8-point batched sweep at its full boxes in one solver group (24 x 24 lattice, U in {0.5, 1, 1.5, 2}
at β = 5 and {2.05, 2.09, 2.5, 2.9} at β = 10, K3 boxes (12, 24) .. (24, 48), mΠν = 1, FP32-MFMA
contraction), three fdPA iterations, checked two ways that need no oracle at these sizes:
  * every member ends bit for bit where the same point iterated alone ends (grouped GEMM launches,
    member streams and all; the solo runs are made one point at a time after the group is freed: the
    group holds about 210 GB of HBM);
  * the members run with k-independent Green functions, so every momentum component must equal the
    local solver with the same boxes (the exact reduction of tests/test_gpu_fullsize.py), to 1e-5 (the
    two paths round different but equal operand sums to FP32).
Starting from F = 0 with G0 = Σ0 = 0 and RefVertex(U), K1 is nonzero after the first iteration and
K2, K3 after the second, so the third exercises every term."""
import os
import sys

import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import fdga  # noqa: E402

ITERS = 3
P_SAMPLE = (0, 1, 25, 300, 575)


def _point(k):
    import bench
    U, beta, f = bench.C4_POINTS[k]
    return dict(L=bench.C4_L, T=1.0 / beta, U=U, nK1=2 * f, nK2=(f, f), nK3=(f // 2, f), mPinu=1, nG=4 * f)


def _gloc(w):
    return fdga.siam_bare_Green(w["nG"], w["T"], e=0.1, Δ=0.8, D=None)


def _lattice(w):
    NP = w["L"] ** 2
    Gk = np.repeat(_gloc(w)[:, None], NP, axis=1)
    z = np.zeros_like(Gk)
    return fdga.NL_ParquetSolver(w["nK1"], w["nK2"], w["nK3"], w["L"], Gk, z, z, T=w["T"], U=w["U"],
                                 mΠν_factor=w["mPinu"], precision="fp32")


def _samples(x, w):
    """the P-sampled vertex classes of a flattened lattice state: {(ch, cls): [n, len(P_SAMPLE)]}"""
    NP = w["L"] ** 2
    n1 = 2 * w["nK1"] - 1
    n2 = (2 * w["nK2"][0] - 1) * 2 * w["nK2"][1]
    n3 = (2 * w["nK3"][0] - 1) * (2 * w["nK3"][1]) ** 2
    out, off = {}, 0
    for ch in range(3):
        for cls, n in ((1, n1), (2, n2), (3, n3)):
            a = x[off:off + n * NP].reshape(n, NP, order="F")
            out[(ch, cls)] = a[:, P_SAMPLE].copy()
            off += n * NP
    return out


@pytest.mark.timeout(900)
def test_sweep_group_full_boxes():
    import bench
    npts = len(bench.C4_POINTS)
    W = [_point(k) for k in range(npts)]
    print("sweep: building the 8 lattice members", flush=True)
    members = [_lattice(w) for w in W]
    G = fdga.SolverGroup(members)
    for it in range(ITERS):
        G.iterate("fdPA")
        G.sync()
        print(f"sweep: group iteration {it} done", flush=True)
    G.close()
    grouped = []
    for k, S in enumerate(members):   # smallest first; each member is freed once read
        grouped.append(fdga.flatten(S))
        S.close()
    del members
    print("sweep: group results read back; solo runs", flush=True)
    for k, w in enumerate(W):
        S = _lattice(w)
        for it in range(ITERS):
            fdga.iterate_solver(S, "fdPA")
        x = fdga.flatten(S)
        S.close()
        assert np.array_equal(x, grouped[k]), f"point {k}: grouped != alone ({relerr(grouped[k], x):.3e})"
        # the k-independent reduction: each sampled momentum component equals the local solver
        Gl = _gloc(w)
        zl = np.zeros_like(Gl)
        Sl = fdga.ParquetSolver(w["nK1"], w["nK2"], w["nK3"], Gl, zl, zl, T=w["T"], U=w["U"], mΠν_factor=w["mPinu"],
                                precision="fp32")
        for it in range(ITERS):
            fdga.iterate_solver(Sl, "fdPA")
        smp = _samples(x, w)
        for ch in range(3):
            for cls in (1, 2, 3):
                loc = Sl.vertex(0, ch, cls)
                e = max(relerr(smp[(ch, cls)][:, j], loc) for j in range(len(P_SAMPLE)))
                assert e < 1e-5, f"point {k} ch={ch} cls={cls}: {e:.3e}"
        Sl.close()
        grouped[k] = None
        print(f"sweep: point {k} (K3 {w['nK3']}) equal alone and reduced", flush=True)
