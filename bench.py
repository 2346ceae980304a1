#!/usr/bin/env python3
"""Benchmark: parquet iterations/s and BSE TFLOP/s of the FP32-MFMA and FP64 GPU paths on one MI355X.

Headline workload (the configuration the metric is quoted on): the nonlocal 2D Hubbard fdPA solver
(NL_ParquetSolver) on a 16 x 16 lattice, t = 1, t' = 0, U = 2, β = 5, half filling, nK1 = 64,
nK2 = (32, 32), nK3 = (16, 32), mΠν_factor = 32, nG = 256, around a local reference system: a
synthetic reference vertex with K1 = K2 = 0 and K3 a complex Gaussian (numpy PCG64, seed 0) times
0.1 U / (1 + |ν| + |ν'|) (ν, ν' in frequency units), G0 the momentum-averaged bare lattice G, Σ0 = 0.  One "step" is one iterate_solver! pass (src/solve.jl:4-116, strategy :fdPA) with all
inputs resident in HBM, replayed from the captured hipGraph of one iteration; `warmup` passes run
first, then `steps` passes are timed with HIP events on the solver's stream.

Both precision paths run from the same initial state: the FP64 path and the FP32-MFMA form (the BSE
contractions on FP32 MFMA with FP32 accumulation, the K3 arrays of FL, F and Fbuff stored in FP32, the
lattice bubble's products and momentum transform in FP32; Dyson, the Π tail sums, the SDE and
everything else FP64).  The max relative error of
Σ, χ_sp(Ω, q), χ_ch(Ω, q) and the three K3 classes of the FP32 path against FP64 is taken after
`--acc-iters` iterations from that state (the lattice fixed point is not solved for: the host
Anderson loop over 1.6 GB state vectors is outside the timed path).  `value` is the FP32-MFMA path's
iterations/s when every error is <= 1e-6, else the FP64 path's (`config.value_path`).

`roofline` is the BSE contraction, the iteration's largest kernel family (BSE_L_K3! + BSE_K3!, four
launches per iteration): on the FP32 path HBM-bound, algorithmic 168 B per K3 element of the three
channels per iteration (gemm_roofline) over the launches' HIP-event time, with `traffic` the rocprofv3
FETCH + WRITE bytes per launch (profiles/c2_bse_gemm_traffic.txt); on the FP64 path MFMA-bound
(executed flops 8 BM BN Nν3 per computed tile).  `roofline_bubble` is the lattice bubble kernel
(k_bubble_nl, bubbles_real_space!): G(R) read once, Πpp and Πph written on the band chunks and the
chunk sums, over its HIP-event time; `traffic` from profiles/c2_bubble_nl_traffic.txt.  `cpu_baseline` times the CPU
restatement of the reference loops (oracle/) on a stated bounded sample of the same workload, at one
thread and at all cores.

Other workloads print their own JSON line: --workload c1 (local SIAM fdPA, nK3 = (32, 64)), c0 (the
small SIAM scPA solve: hipGraph replay, Newton–Krylov vs mixing), c4 (the 8-point lattice sweep in one
solver group), c3 (the tangent and adjoint of build_K3_cache! + BSE_K3! on the headline lattice).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fddgasolver.jl_amd"))

WORKLOAD = dict(T=0.05, U=2.0, Delta=3.141592653589793 / 5, nK1=128, nK2=(64, 64), nK3=(32, 64), mPinu=6, nG=256)
FP64_MFMA_PEAK = 78.6e12
FP32_MFMA_PEAK = 157.3e12
HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def flops_bse_k3(w, NP=1):
    NW3 = 2 * w["nK3"][0] - 1
    NV3 = 2 * w["nK3"][1]
    return 8.0 * NW3 * NP * NV3 ** 3   # one channel, one GEMM


def build_solver(w, precision="fp64"):
    import fdga
    S = fdga.parquet_solver_siam_parquet_approximation(w["nG"], w["nK1"], w["nK2"], w["nK3"], e=0.0, Δ=w["Delta"],
                                                       D=None, T=w["T"], U=w["U"], mΠν_factor=w["mPinu"],
                                                       precision=precision)
    return S


def _threads_all():
    """the CPU threads this process may use (the box's share, or OMP_NUM_THREADS when set)"""
    env = os.environ.get("OMP_NUM_THREADS")
    n = len(os.sched_getaffinity(0))
    return min(n, int(env)) if env and env.isdigit() and int(env) > 0 else n


def _time_oracle(o, strategy, reps):
    """the median of `reps` timed iterate_solver! passes of the oracle solver o (the pass's cost does not
    depend on the state it starts from)"""
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        o.iterate(strategy)
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def _time_both(po, make, strategy, reps_all, reps_one):
    """one warm-up pass at all cores, then `reps_all` passes at all cores and `reps_one` at one thread"""
    nall = _threads_all()
    po.set_threads(nall)
    o = make()
    o.iterate(strategy)
    tn = _time_oracle(o, strategy, reps_all)
    po.set_threads(1)
    t1 = _time_oracle(o, strategy, reps_one)
    po.set_threads(nall)
    return nall, tn, t1


def cpu_baseline_c1(sym=True, reps=5):
    """The CPU restatement of the reference loops (oracle/) on the local SIAM fdPA workload: full-box
    iterate_solver! passes from the solver's initial state with the same symmetry groups as the GPU run,
    at one thread and at all cores (1 warm-up, median of `reps`)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    w = WORKLOAD

    def make():
        o = po.Oracle(nG=w["nG"], nK1=w["nK1"], nK2=w["nK2"], nK3=w["nK3"], mPinu=w["mPinu"], T=w["T"], U=w["U"])
        o.siam(0.0, w["Delta"], None)
        o.init()
        if sym:
            o.init_sym_grp()
        return o
    nall, tn, t1 = _time_both(po, make, po.FDPA, reps, reps)
    return {"value": 1.0 / tn, "unit": "iters/s", "cores": nall, "kind": "port",
            "single_thread": {"value": 1.0 / t1, "cores": 1},
            "sample": f"full fdPA iterate_solver! passes of the workload box on the CPU restatement "
                      f"(oracle/fdga_oracle.cpp{', init_sym_grp!' if sym else ''}): 1 warm-up, median of {reps} "
                      f"each: {tn:.2f} s at {nall} threads, {t1:.2f} s at 1 thread"}


def cpu_baseline_c2(reps=5, reps_one=3):
    """The CPU restatement of the reference loops (oracle/) on a bounded sample of the lattice workload:
    its frequency boxes, bubble mesh, G mesh size per momentum and synthetic reference vertex on a 1 x 1
    lattice (NP = 1), fdPA iterate_solver! passes timed at all cores (median of `reps`) and at one thread
    (median of `reps_one`) after one warm-up, scaled by 1/256 to the 256 momenta of the 16 x 16 lattice (the restated loops
    run every stage once per momentum; the momentum FFTs, O(L² log L), are a small part)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle as po
    w = C2
    NP = w["L"] ** 2

    def make():
        o = po.Oracle(nonlocal_=True, nG=w["nG"], nK1=w["nK1"], nK2=w["nK2"], nK3=w["nK3"], mPinu=w["mPinu"], T=w["T"],
                      U=w["U"], LG=1, L=1, ref_kind=1, ref_boxes=(w["nK1"], w["nK2"], w["nK3"]))
        o.hubbard(w["mu"], w["t1"])
        Gb = o.get(po.GBARE)
        o.set(po.G0, Gb)
        for (ch, cls), a in c2_reference_vertex(w, seed=0).items():
            o.set_vertex(po.VTX_F0, ch, cls, a)
        o.init()
        return o
    nall, tn, t1 = _time_both(po, make, po.FDPA, reps, reps_one)
    return {"value": 1.0 / (tn * NP), "unit": "iters/s", "cores": nall, "kind": "port",
            "single_thread": {"value": 1.0 / (t1 * NP), "cores": 1},
            "sample": f"the workload's frequency boxes, bubble mesh and synthetic reference vertex on a 1x1 lattice "
                      f"(NP = 1, the Hubbard G at k = 0), fdPA iterate_solver! passes of the CPU restatement "
                      f"(oracle/fdga_oracle.cpp), 1 warm-up: median of {reps} at {nall} threads {tn:.2f} s, median of "
                      f"{reps_one} at 1 thread {t1:.2f} s; scaled x{NP} to the 16x16 lattice's {NP} momenta "
                      f"(value = 1 / ({NP} t))"}


def gemm_traffic(path, tile):
    """HBM bytes per launch of the BSE_K3! GEMM kernel with tile shape `tile` (FETCH + WRITE,
    calibrated against a plain copy) as measured by rocprofv3 --pmc and recorded in
    profiles/c1_bse_k3_gemm_traffic.txt; None when the file has no entry for the path and shape."""
    import re
    f = os.path.join(ROOT, "profiles", "c1_bse_k3_gemm_traffic.txt")
    if not os.path.exists(f):
        return None
    m = re.search(rf"^{path}_hbm_bytes_per_launch_{tile}\s*=\s*([0-9.eE+]+)", open(f).read(), re.M)
    return float(m.group(1)) if m else None


def observables(S):
    """Σ, χ_sp = K1a/U², χ_ch = -(2 K1t - K1a)/U² (pSp storage of the connected K1) and the three
    channels' K3, read back from the device."""
    import numpy as np
    import fdga
    chi_sp, chi_ch = fdga.susceptibilities(S)
    return {"Sigma": S.Σ, "chi_sp": chi_sp, "chi_ch": chi_ch,
            "K3": np.concatenate([S.vertex(0, ch, 3) for ch in range(3)])}


def max_relerr(a, b):
    import numpy as np
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def time_path(S, args, names):
    """warmup + timed iterations (HIP events on the solver's stream; replayed from the captured
    hipGraph of one iteration unless --eager), then a per-stage pass (eager, one stage at a time).
    The per-stage pass runs the contractions with one workgroup per work item: a persistent grid
    tuned to leave compute units to the iteration's other chain has no other chain to serve there,
    and the kernels alone are what the stage times and the roofline describe."""
    import fdga
    S.set_graph(not args.eager)
    for _ in range(args.warmup):
        fdga.iterate_solver(S, "fdPA")
    S.sync()
    ms = S.time_iterations("fdPA", args.steps)
    S.set_graph(False)
    grids = S.gemm_grids()
    S.set_gemm_grid(True, 0)
    S.set_gemm_grid(False, 0)
    nst = max(1, min(args.steps, 5))
    S.set_stage_timing(True)
    for _ in range(nst):
        fdga.iterate_solver(S, "fdPA")
    S.sync()
    st = [x / nst for x in S.stage_ms()]
    S.set_stage_timing(False)
    S.set_gemm_grid(True, grids[0])
    S.set_gemm_grid(False, grids[1])
    return ms / args.steps, dict(zip(names, st))


C2 = dict(L=16, T=0.2, U=2.0, mu=0.0, t1=1.0, nK1=64, nK2=(32, 32), nK3=(16, 32), mPinu=32, nG=256)


STAGES = ["dyson+bubbles", "k3_cache", "L_K2", "L_K3", "K1", "K2", "K3", "update", "SDE", "gemm_L_K3", "gemm_K3",
          "bubble_kernel"]


def c2_reference_vertex(w, seed=0):
    """The synthetic local reference vertex of the lattice workloads: K1 = K2 = 0, K3 of each channel a
    complex Gaussian (numpy PCG64, `seed`) times 0.1 U / (1 + |ν| + |ν'|) with ν = πT(2n+1) in frequency
    units.  Returns {(ch, cls): array in the reference layout}."""
    import numpy as np
    nb, nf = w["nK3"]
    rng = np.random.default_rng(seed)
    NW, NV = 2 * nb - 1, 2 * nf
    nu = np.pi * w["T"] * (2 * np.arange(-nf, nf) + 1)
    amp = 0.1 * w["U"] / (1 + np.abs(nu)[:, None] + np.abs(nu)[None, :])   # [ν, ν']
    F0 = {}
    for ch in range(3):
        k3 = (rng.standard_normal((NW, NV, NV)) + 1j * rng.standard_normal((NW, NV, NV))) * amp[None, :, :]
        F0[(ch, 1)] = np.zeros(2 * w["nK1"] - 1, complex)
        F0[(ch, 2)] = np.zeros((2 * w["nK2"][0] - 1) * 2 * w["nK2"][1], complex)
        F0[(ch, 3)] = k3.reshape(-1, order="F")
    return F0


def build_c2(precision, w=C2, seed=0):
    """Nonlocal 2D Hubbard fdPA solver (NL_ParquetSolver, src/nonlocal/ParquetSolver.jl:77-152) on an
    L x L lattice around a local reference system: the synthetic local reference vertex of
    c2_reference_vertex, G0(ν, k) = the momentum average of the bare lattice G (the same at every k, as
    the local reference systems of fdDΓA, cf. test/test_nonlocal_symmetry.jl:144-155) and Σ0 = 0.  (With
    G0 equal to the lattice G the bubbles of both systems coincide and the fdPA correction vanishes.)"""
    import numpy as np
    import fdga
    L = w["L"]
    Gb = fdga.hubbard_bare_Green(w["nG"], w["T"], L, μ=w["mu"], t1=w["t1"])
    G0 = np.repeat(Gb.mean(axis=1, keepdims=True), L * L, axis=1)
    F0 = c2_reference_vertex(w, seed)
    S = fdga.NL_ParquetSolver(w["nK1"], w["nK2"], w["nK3"], L, Gb, G0, np.zeros_like(Gb), F0, T=w["T"], U=w["U"],
                              mΠν_factor=w["mPinu"], F0_boxes=(w["nK1"], w["nK2"], w["nK3"]), precision=precision)
    return S


def recorded_traffic(name, key="hbm_bytes_per_launch"):
    """HBM bytes per launch recorded in profiles/<name> (rocprofv3 --pmc FETCH_SIZE x 2 + WRITE_SIZE x 1, the
    factors calibrated against a 512 MiB copy); None if absent."""
    import re
    f = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(f):
        return None
    m = re.search(rf"^{key}\s*=\s*([0-9.eE+]+)", open(f).read(), re.M)
    return float(m.group(1)) if m else None


def c2_band(w):
    """(wb, number of 16-ω chunks the lattice bubble stores) of the lattice workload: the band half-width
    wb = min(nVP, NK + max(nK1 - 1, nK2b - 1 + nK2f) + 3) with NK = 1 + the largest box of F and its
    reference vertex, and the chunks of 16 bubble frequencies that meet [nVP - wb, nVP + wb)."""
    nVP = w["nK1"] * w["mPinu"]
    NK = 1 + max(w["nK1"], w["nK2"][0], w["nK2"][1], w["nK3"][0], w["nK3"][1])
    wb = min(nVP, NK + max(w["nK1"] - 1, w["nK2"][0] - 1 + w["nK2"][1]) + 3)
    return wb, (nVP + wb + 15) // 16 - (nVP - wb) // 16


def tune_gemm_grid(S, grids=(0, 256, 512, 768, 1024), lk3_grids=(0, 160, 192, 224), reps=4):
    """Workgroup counts of the contraction launches by whole-iteration time (hipGraph replay): a persistent
    grid leaves compute units to the iteration's other chain (BSE_L_K2!, weights, BSE_K1!, BSE_K2!),
    which the launch-by-launch autotuner cannot see.  First one count for both contractions, then
    BSE_L_K3!'s alone (the FP32 lattice form runs one workgroup per compute unit: counts below 256 leave
    some free).  The state is restored afterwards; results do not depend on the grids.  Returns
    (BSE_L_K3!, BSE_K3!) counts."""
    import fdga
    x = fdga.flatten(S)
    S.set_graph(True)

    def timed(gl, gk):
        S.set_gemm_grid(True, gl)
        S.set_gemm_grid(False, gk)
        fdga.unflatten(S, x)
        fdga.iterate_solver(S, "fdPA")   # capture
        S.sync()
        return S.time_iterations("fdPA", reps)
    best = min((timed(g, g), g) for g in grids)
    gk = best[1]
    best = min((timed(g, gk), g) for g in lk3_grids if g != gk or g == 0)
    gl = best[1]
    S.set_gemm_grid(True, gl)
    S.set_gemm_grid(False, gk)
    S.set_graph(False)
    fdga.unflatten(S, x)
    return gl, gk


def gemm_roofline(head, P, gemm_peak, traffic, n3P):
    """Roofline object of the BSE contraction GEMMs.  FP32 path: HBM-bound (their epilogues and FP32 operands
    move 168 B per K3 element of the three channels per iteration, every K3 array in FP32: FL.K3 and the
    FP32 right operand written by BSE_L_K3! (24 + 24), FL.a read for its t post-processing (8); BSE_K3! reads
    A32, R32, FL.K3, FL.a for the t self term and Fbuff.a for the t post-processing (24 + 24 + 24 + 8 + 8) and
    writes Fbuff.K3 (24)); the MFMA rate is stated beside.  FP64 path: MFMA-bound (executed flops)."""
    t = P["t_gemm"]
    mfma = P["f_exec"] / t
    kern = ("k_lk3_pb64 (BSE_L_K3!) + zgemm_tiles_t (BSE_K3!), FP32 MFMA" if head == "fp32"
            else "zgemm_tiles_t (FP64 MFMA, BSE_L_K3! + BSE_K3!)")
    common = {"kernel": kern + ", 4 launches per iteration",
              "launches_per_iteration": 4, "ms_per_iteration": t * 1e3, "traffic": traffic,
              "executed_flops_per_iteration": P["f_exec"], "mfma_tflops": mfma / 1e12, "mfma_frac": mfma / gemm_peak}
    if head == "fp32":
        b = 168.0 * n3P / 4
        return dict({"bound": "hbm", "achieved": b / (t / 4) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": b / (t / 4) / HBM_PEAK, "algorithmic_bytes_per_launch": b}, **common)
    return dict({"bound": "mfma", "achieved": mfma / 1e12, "peak": gemm_peak / 1e12, "unit": "TFLOP/s",
                 "frac": mfma / gemm_peak}, **common)


def main_c2(args):
    """The headline line: the lattice fdPA iteration in both precision paths."""
    import fdga
    w = C2
    NW3, NV3, NP = 2 * w["nK3"][0] - 1, 2 * w["nK3"][1], w["L"] ** 2
    NWP, NVP = 2 * w["nK1"] - 1, 2 * w["nK1"] * w["mPinu"]
    f_dense = 2 * 3 * 8.0 * NW3 * NP * NV3 ** 3          # BSE_L_K3! + BSE_K3!, three channels, dense
    paths, x0 = {}, None
    run_paths = [q for q in ("fp64", "fp32") if q in args.paths.split(",")]
    for prec in run_paths:
        t0 = time.perf_counter()
        S = build_c2(prec)
        log(f"bench c2[{prec}]: solver built ({time.perf_counter() - t0:.1f} s)")
        if x0 is None:
            x0 = fdga.flatten(S)
        else:
            fdga.unflatten(S, x0)
        for _ in range(args.acc_iters):
            fdga.iterate_solver(S, "fdPA")
        obs = observables(S)
        fdga.unflatten(S, x0)
        S.autotune(reps=3)   # tile shapes of the contraction launches (results do not depend on them)
        grid = tune_gemm_grid(S)
        log(f"bench c2[{prec}]: autotuned tile shapes (BSE_L_K3!, BSE_K3!): {S.gemm_tiles()}, grid {grid}")
        ms, st = time_path(S, args, STAGES)
        f_exec = S.gemm_flops(True) + S.gemm_flops(False)
        t_gemm = (st["gemm_L_K3"] + st["gemm_K3"]) * 1e-3
        log(f"bench c2[{prec}]: {ms:.3f} ms/iteration; stages " + ", ".join(f"{k}={v:.3f}" for k, v in st.items()))
        paths[prec] = dict(ms=ms, st=st, obs=obs, f_exec=f_exec, t_gemm=t_gemm, bse_tflops=f_exec / t_gemm / 1e12,
                           bse_tflops_dense_equiv=f_dense / t_gemm / 1e12)
        S.close()
    if len(run_paths) < 2:   # profiling run of one path: per-stage line only
        q = run_paths[0]
        print(json.dumps({"path": q, "ms_per_step": paths[q]["ms"], "stage_ms": paths[q]["st"]}), flush=True)
        return
    errs = {k: max_relerr(paths["fp32"]["obs"][k], paths["fp64"]["obs"][k]) for k in paths["fp64"]["obs"]}
    log("bench c2: FP32-MFMA vs FP64 max rel.err after %d iterations: " % args.acc_iters
        + ", ".join(f"{k}={v:.2e}" for k, v in errs.items()))
    head = "fp32" if max(errs.values()) <= 1e-6 else "fp64"
    P = paths[head]
    # k_bubble_nl as built: G(R) read once (FP32 on the FP32 path), Πpp and Πph stored on the band
    # chunks, the chunk sums of all 16-ω chunks written (the far regions enter the later stages only
    # through those sums)
    wb, nchunk = c2_band(w)
    b_bubble = ((8.0 if head == "fp32" else 16.0) * 2 * w["nG"] * NP
                + 16.0 * (2 * nchunk * 16 * NWP * NP + 2 * (NVP // 16) * NWP * NP))
    t_bub = P["st"]["bubble_kernel"] * 1e-3
    gemm_peak = FP32_MFMA_PEAK if head == "fp32" else FP64_MFMA_PEAK
    gemm_traffic = recorded_traffic("c2_bse_gemm_traffic.txt", f"{head}_hbm_bytes_per_launch")
    out = {
        "metric": "parquet iters/s + BSE TFLOP/s, FP32-MFMA vs FP64 GPU @ rel.err<=1e-6 vs FP64",
        "value": 1000.0 / P["ms"],
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": P["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32-mfma contractions, f32 K3 storage and f32 lattice-bubble interior, f64 elsewhere"
                  if head == "fp32" else "f64"),
        "data": "synthetic (Hubbard model inputs, local reference system: Gaussian reference K3 seed 0, k-averaged G0)",
        "config": {"workload": "nonlocal 2D Hubbard fdPA (NL_ParquetSolver), 16x16 lattice, t=1, t'=0, U=2, beta=5, "
                               "half filling, nK1=64, nK2=(32,32), nK3=(16,32), mPinu=32, nG=256, local reference system "
                               "(Gaussian K3 of amplitude 0.1 U/(1+|nu|+|nu'|), k-averaged bare G0, Sigma0=0)",
                   "value_path": head,
                   "iteration_launch": "eager" if args.eager else "hipGraph replay of one captured iteration",
                   "accuracy": {"iterations_from_one_state": args.acc_iters, "max_relerr_fp32_vs_fp64": errs},
                   "fp32_mfma": {"iters_per_s": 1000.0 / paths["fp32"]["ms"], "bse_tflops": paths["fp32"]["bse_tflops"],
                                 "bse_tflops_dense_equiv": paths["fp32"]["bse_tflops_dense_equiv"],
                                 "stage_ms": paths["fp32"]["st"]},
                   "fp64": {"iters_per_s": 1000.0 / paths["fp64"]["ms"], "bse_tflops": paths["fp64"]["bse_tflops"],
                            "bse_tflops_dense_equiv": paths["fp64"]["bse_tflops_dense_equiv"],
                            "stage_ms": paths["fp64"]["st"]},
                   "speedup_fp32_over_fp64": paths["fp64"]["ms"] / paths["fp32"]["ms"]},
        # the dominant kernel: the BSE contraction GEMMs (four launches per iteration: BSE_L_K3! p+a and t,
        # BSE_K3! p+a and t), over their summed durations (HIP events, one kernel at a time)
        "roofline": gemm_roofline(head, P, gemm_peak, gemm_traffic, NW3 * NV3 * NV3 * NP),
        "roofline_bubble": {"bound": "hbm", "achieved": b_bubble / t_bub / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                            "frac": b_bubble / t_bub / HBM_PEAK,
                            "traffic": recorded_traffic("c2_bubble_nl_traffic.txt"),
                            "kernel": "k_bubble_nl (bubbles_real_space!: Πpp, Πph stored on the band chunks, chunk sums)",
                            "algorithmic_bytes_per_launch": b_bubble, "band_halfwidth": wb, "stored_chunks": nchunk,
                            "launch_ms": t_bub * 1e3},
    }
    if not args.no_cpu_baseline:
        log("bench c2: timing the CPU baseline (oracle, bounded sample, 1 thread and all cores)")
        out["cpu_baseline"] = cpu_baseline_c2()
    print(json.dumps(out), flush=True)


# batched sweep: 8 Hubbard points (U, beta, nK3f) on a 24 x 24 lattice, mΠν = 1; per point
# K3 = (nK3f/2, nK3f), K2 = (nK3f, nK3f), K1 = 2 nK3f, nG = 4 nK3f (beta 5 for U <= 2, 10 above)
C4_POINTS = [(0.5, 5.0, 24), (1.0, 5.0, 28), (1.5, 5.0, 32), (2.0, 5.0, 36),
             (2.05, 10.0, 40), (2.09, 10.0, 44), (2.5, 10.0, 48), (2.9, 10.0, 48)]
C4_L = 24


def build_c4_point(k, precision):
    U, beta, f = C4_POINTS[k]
    w = dict(L=C4_L, T=1.0 / beta, U=U, mu=0.0, t1=1.0, nK1=2 * f, nK2=(f, f), nK3=(f // 2, f), mPinu=1, nG=4 * f)
    return build_c2(precision, w, seed=k), w


def main_c4(args):
    """The batched sweep: one parquet iteration of all 8 points per step, every member's stages on
    its own stream and the BSE contractions of all points in one grouped persistent launch each,
    tile shapes and grids from the autotuner, the group iteration replayed from a hipGraph; FP32
    MFMA contraction.  Compared with the same points iterated one after the other."""
    import fdga
    t0 = time.perf_counter()
    S, ws = [], []
    for k in range(len(C4_POINTS)):
        Sk, w = build_c4_point(k, "fp32")
        S.append(Sk)
        ws.append(w)
        log(f"bench c4: point {k} U={w['U']} beta={1 / w['T']:g} K3={w['nK3']} built ({time.perf_counter() - t0:.1f} s)")
    G = fdga.SolverGroup(S)
    G.iterate("fdPA")   # first iteration eagerly (caches each point's reference SDE)
    G.sync()
    t1 = time.perf_counter()
    G.autotune(reps=2)
    tuned = {("L_K3", "K3")[wh]: dict(zip(("tile", "grid", "computed_tiles"), G.tiles(wh))) for wh in (0, 1)}
    log(f"bench c4: autotuned in {time.perf_counter() - t1:.1f} s: {tuned}")
    G.set_graph(True)
    for _ in range(max(1, args.warmup)):
        G.iterate("fdPA")
    G.sync()
    ms = G.time_iterations("fdPA", args.steps) / args.steps
    G.set_timing(True)
    G.iterate("fdPA")
    G.sync()
    gl, gk = G.gemm_ms()
    G.set_timing(False)
    # algorithmic flops of one grouped launch: 3 channels x 8 flop per complex MAC, dense per point
    f_launch = sum(3 * 8.0 * (2 * w["nK3"][0] - 1) * w["L"] ** 2 * (2 * w["nK3"][1]) ** 3 for w in ws)
    tf_l, tf_k = f_launch / (gl * 1e-3) / 1e12, f_launch / (gk * 1e-3) / 1e12
    log(f"bench c4: {ms:.1f} ms per sweep iteration; grouped GEMMs {gl:.2f} ms ({tf_l:.1f} TFLOP/s) and "
        f"{gk:.2f} ms ({tf_k:.1f} TFLOP/s)")
    G.close()
    solo = []
    for Sk in S:   # the same points one after the other (eager, each alone)
        solo.append(Sk.time_iterations("fdPA", 1))
    log(f"bench c4: points one after the other {sum(solo):.1f} ms ({', '.join(f'{x:.1f}' for x in solo)})")
    peak = 157.3
    print(json.dumps({
        "workload": "c4: 8 NL Hubbard points, 24x24 lattice, U in {0.5,1,1.5,2 (beta 5), 2.05,2.09,2.5,2.9 (beta 10)}, "
                    "K3 (12,24)..(24,48), mPinu=1, fdPA, FP32-MFMA contraction, synthetic local reference vertices",
        "steps": args.steps, "warmup": args.warmup,
        "sweep_ms_per_iteration": ms, "sweep_iters_per_s": 1000.0 / ms,
        "points_one_after_another_ms": sum(solo),
        "grouped_gemm": {"L_K3_ms": gl, "K3_ms": gk, "flops_per_launch": f_launch,
                         "L_K3_tflops": tf_l, "K3_tflops": tf_k, "peak_tflops_fp32": peak,
                         "roofline_frac": max(tf_l, tf_k) / peak},
        "autotuned": tuned}), flush=True)


def main_c3(args):
    """C3 on the headline lattice (FP64): the tangent and adjoint of build_K3_cache! followed by BSE_K3! of
    the three channels (fdga_bse_k3_jvp / fdga_bse_k3_vjp), linearised at the state after --acc-iters fdPA
    iterations.  Times of the primal composite, the JVP and the VJP (the first VJP also builds the
    gather's transposed map), each also in parquet-iteration equivalents; the JVP against central
    differences with h = 1e-4 along a vertex direction (dF, dFL) and a bubble direction dΠ (on the K3
    box), each exact for a central difference (the composite is quadratic in (F, FL) and affine in Π);
    and the identity <ō, J v> = <J^T ō, v> for a random cotangent."""
    import numpy as np
    import fdga
    from fdga import _lib as L
    S = build_c2("fp64")
    for _ in range(args.acc_iters):
        fdga.iterate_solver(S, "fdPA")
    S.sync()
    ms_iter = S.time_iterations("fdPA", 3) / 3
    log(f"bench c3: solver built, {args.acc_iters} iterations; one iteration {ms_iter:.2f} ms")
    chs = (0, 1, 2)
    x = ({(ch, cls): S.vertex(0, ch, cls) for ch in chs for cls in (1, 2, 3)},
         {ch: S.vertex(2, ch, 3) for ch in chs}, [S.get(L.PIPP), S.get(L.PIPH)])
    rng = np.random.default_rng(7)

    def rnd(n, sc):
        return sc * (rng.standard_normal(n) + 1j * rng.standard_normal(n))

    def load(T_, y):
        F, FL, Pi = y
        for (ch, cls), a in F.items():
            T_.set_vertex(0, ch, cls, a)
        for ch, a in FL.items():
            T_.set_vertex(2, ch, 3, a)
        T_.set(L.PIPP, Pi[0])
        T_.set(L.PIPH, Pi[1])

    def composite(y):
        load(S, y)
        t0 = time.perf_counter()
        fdga.build_K3_cache(S)
        for ch in (0, 2, 1):
            fdga.BSE_K3(S, ch)
        S.sync()
        ms = (time.perf_counter() - t0) * 1e3
        return np.concatenate([S.vertex(1, ch, 3) for ch in chs]), ms

    def axpy(y, v, h):
        return ({k: y[0][k] + h * v[0][k] for k in y[0]}, {k: y[1][k] + h * v[1][k] for k in y[1]},
                [a + h * b for a, b in zip(y[2], v[2])])

    zPi = [np.zeros_like(x[2][0]), np.zeros_like(x[2][1])]
    vF = ({k: rnd(a.size, np.abs(a).max() + 1e-3) for k, a in x[0].items()},
          {k: rnd(a.size, np.abs(a).max() + 1e-3) for k, a in x[1].items()}, zPi)
    # dΠ on the K3 box only (the entries the composite reads): reference layout [Ω, ν, P]
    NWP, NVP, NP = 2 * C2["nK1"] - 1, 2 * C2["nK1"] * C2["mPinu"], C2["L"] ** 2
    nb, nf = C2["nK3"]
    dPi = []
    for a in x[2]:
        d = np.zeros((NWP, NVP, NP), complex)
        sub = d[C2["nK1"] - nb:C2["nK1"] + nb - 1, NVP // 2 - nf:NVP // 2 + nf, :]
        sub[...] = rnd(sub.size, np.abs(a).max()).reshape(sub.shape)
        dPi.append(d.reshape(-1, order="F"))
    vP = ({k: np.zeros_like(a) for k, a in x[0].items()}, {k: np.zeros_like(a) for k, a in x[1].items()}, dPi)
    Tn = fdga.tangent_solver(S)
    h = 1e-4
    out = {"workload": "c3: tangent and adjoint of build_K3_cache! + BSE_K3! (p, a, t) on the 16x16 Hubbard lattice "
                       "(nK3=(16,32), FP64), linearised after %d fdPA iterations" % args.acc_iters,
           "iteration_ms_fp64": ms_iter}
    errs, jv = {}, []
    t_primal = []
    for name, v in (("vertex", vF), ("bubble", vP)):
        fp, tp = composite(axpy(x, v, h))
        fm, tm = composite(axpy(x, v, -h))
        t_primal += [tp, tm]
        fd = (fp - fm) / (2 * h)
        load(S, x)
        load(Tn, v)
        S.sync()
        t0 = time.perf_counter()
        fdga.BSE_K3_jvp(S, Tn)
        t_jvp = (time.perf_counter() - t0) * 1e3
        j = np.concatenate([Tn.vertex(1, ch, 3) for ch in chs])
        errs[name] = max_relerr(j, fd)
        jv.append(j)
        log(f"bench c3: JVP along the {name} direction: {t_jvp:.1f} ms, rel. err. against the central difference "
            f"{errs[name]:.2e}")
    # timings at the linearisation point (three of each after the runs above)
    load(S, x)
    load(Tn, vF)
    S.sync()
    t0 = time.perf_counter()
    for _ in range(3):
        fdga.BSE_K3_jvp(S, Tn)
    t_jvp = (time.perf_counter() - t0) * 1e3 / 3
    ob = rnd(jv[0].size, 1.0)
    n3 = ob.size // 3

    def vjp():
        for k, ch in enumerate(chs):
            Tn.set_vertex(1, ch, 3, ob[k * n3:(k + 1) * n3])
        Tn.sync()
        t0 = time.perf_counter()
        fdga.BSE_K3_vjp(S, Tn)
        return (time.perf_counter() - t0) * 1e3
    t_vjp_first = vjp()
    t_vjp = min(vjp() for _ in range(3))
    lhs = np.sum(ob * (jv[0] + jv[1]))
    rhs = sum(np.sum(Tn.vertex(0, ch, cls) * vF[0][ch, cls]) for ch in chs for cls in (1, 2, 3))
    rhs += sum(np.sum(Tn.vertex(2, ch, 3) * vF[1][ch]) for ch in chs)
    rhs += np.sum(Tn.get(L.PIPP) * dPi[0]) + np.sum(Tn.get(L.PIPH) * dPi[1])
    dot = float(abs(lhs - rhs) / np.sum(np.abs(ob * (jv[0] + jv[1]))))
    primal = min(t_primal)
    out.update({"primal_composite_ms": primal, "jvp_ms": t_jvp, "vjp_ms": t_vjp, "vjp_first_call_ms": t_vjp_first,
                "in_iterations": {"primal": primal / ms_iter, "jvp": t_jvp / ms_iter, "vjp": t_vjp / ms_iter},
                "jvp_vs_central_difference_relerr": errs, "h": h, "dot_identity_relerr": dot,
                "timing": "host clock around synchronised calls (array uploads excluded)"})
    log(f"bench c3: primal composite {primal:.1f} ms, JVP {t_jvp:.1f} ms, VJP {t_vjp:.1f} ms (first call "
        f"{t_vjp_first:.1f} ms); <ō, J v> vs <J^T ō, v>: {dot:.2e}")
    Tn.close()
    S.close()
    print(json.dumps(out), flush=True)


C0 = dict(T=0.1, U=1.0, Delta=3.141592653589793 / 5, nK1=33, nK2=(16, 16), nK3=(8, 8), mPinu=6, nG=96)


def main_c0(args):
    """The small SIAM scPA solve: one iteration launched eagerly vs replayed from a hipGraph, and the
    fixed point to 1e-8 with the Anderson mixing of solve! vs Newton–Krylov."""
    import numpy as np
    import fdga
    from fdga import newton
    w = C0

    def mk():
        S = fdga.parquet_solver_siam_parquet_approximation(w["nG"], w["nK1"], w["nK2"], w["nK3"], e=0.0, Δ=w["Delta"],
                                                           D=None, T=w["T"], U=w["U"], mΠν_factor=w["mPinu"])
        S.init_sym_grp()
        return S
    out = {"workload": "c0: local SIAM scPA, U=1, beta=10, nK1=33, nK2=(16,16), nK3=(8,8), mPinu=6, nG=96, symmetry groups"}
    S = mk()
    for _ in range(args.warmup):
        fdga.iterate_solver(S, "scPA")
    n = max(args.steps, 50)
    ms_eager = S.time_iterations("scPA", n) / n
    S.set_graph(True)
    fdga.iterate_solver(S, "scPA")   # capture
    ms_graph = S.time_iterations("scPA", n) / n
    out["iteration_ms"] = {"eager": ms_eager, "hipgraph": ms_graph}
    log(f"bench c0: iteration {ms_eager * 1e3:.1f} us eager, {ms_graph * 1e3:.1f} us hipGraph")
    A = mk()
    t0 = time.perf_counter()
    it_a, res_a = fdga.solve(A, strategy="scPA", tol=1e-8, maxiter=200)
    t_a = time.perf_counter() - t0
    N = mk()
    t0 = time.perf_counter()
    it_n, res_n, nev = newton.solve_newton(N, strategy="scPA", tol=1e-8)
    t_n = time.perf_counter() - t0
    d = float(np.max(np.abs(fdga.flatten(A) - fdga.flatten(N))) / np.max(np.abs(fdga.flatten(A))))
    out["solve_to_1e-8"] = {"anderson": {"iterations": it_a, "parquet_iterations": it_a, "residual": res_a, "seconds": t_a},
                            "newton_krylov": {"newton_iterations": it_n, "parquet_iterations": nev, "residual": res_n,
                                              "seconds": t_n},
                            "solution_relerr": d}
    log(f"bench c0: Anderson {it_a} iterations ({t_a:.2f} s); Newton-Krylov {it_n} Newton steps, {nev} parquet "
        f"iterations ({t_n:.2f} s); solutions differ by {d:.1e}")
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tol", type=float, default=1e-8, help="Anderson tolerance of the untimed FP64 set-up solve")
    ap.add_argument("--no-sym", action="store_true",
                    help="identity symmetry groups (every element evaluated) instead of init_sym_grp!")
    ap.add_argument("--paths", default="fp64,fp32", help="precision paths to run (profiling runs select one)")
    ap.add_argument("--no-dense", action="store_true", help="skip the identity-group FP64 timing")
    ap.add_argument("--eager", action="store_true", help="launch the timed iterations kernel by kernel instead of "
                                                         "replaying the captured hipGraph of one iteration")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 512 MiB device-to-device copy first (calibration of the rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE counters against a copy of known size)")
    ap.add_argument("--tol32", type=float, default=1e-7,
                    help="Anderson tolerance of the FP32-MFMA solve (its residual floor is ~5e-8 at this workload)")
    ap.add_argument("--workload", default="c2", choices=["c0", "c1", "c2", "c3", "c4"],
                    help="c2: the headline NL Hubbard lattice workload (the metric's JSON line); c1: the local SIAM fdPA "
                         "workload at its fixed point (separate JSON line); c0: the small SIAM solve (hipGraph replay and "
                         "Newton–Krylov vs mixing, separate JSON line); c4: the 8-point lattice sweep in one solver group "
                         "(grouped GEMM launches, autotuned tiles, hipGraph; separate JSON line); c3: the tangent and adjoint of "
                         "build_K3_cache! + BSE_K3! on the c2 lattice (separate JSON line)")
    ap.add_argument("--acc-iters", type=int, default=5,
                    help="c2: iterations from the initial state after which the FP32 path is compared with FP64")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: the parquet hot path is run on a single GPU in this build "
              "(multi-GPU sharding is design-only); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        import torch
        a = torch.ones(64 * 2**20, dtype=torch.float64, device="cuda")   # 512 MiB
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.workload == "c2":
        return main_c2(args)
    if args.workload == "c0":
        return main_c0(args)
    if args.workload == "c4":
        return main_c4(args)
    if args.workload == "c3":
        return main_c3(args)
    return main_c1(args)


def main_c1(args):
    """The local SIAM fdPA workload at its fixed point (secondary line)."""
    import fdga
    w = WORKLOAD
    names = STAGES
    f_gemm = flops_bse_k3(w)   # algorithmic flops of one dense channel GEMM
    paths = {}
    x64 = None
    run_paths = [q for q in ("fp64", "fp32") if q in args.paths.split(",")]
    for prec in run_paths:
        log(f"bench[{prec}]: building SIAM fdPA solver nK1={w['nK1']} nK2={w['nK2']} nK3={w['nK3']} nG={w['nG']}")
        S = build_solver(w, prec)
        if not args.no_sym:
            S.init_sym_grp()
        if x64 is not None:
            fdga.unflatten(S, x64)   # the FP32 path starts from the FP64 fixed point
        t0 = time.perf_counter()
        it, res = fdga.solve(S, strategy="fdPA", tol=args.tol if prec == "fp64" else args.tol32, maxiter=80)
        log(f"bench[{prec}]: fixed point after {it} Anderson iterations (max|f| = {res:.2e}, "
            f"{time.perf_counter() - t0:.1f} s)")
        obs = observables(S)
        if x64 is None:
            x64 = fdga.flatten(S)
        S.autotune(reps=5)   # tile shapes of the two contraction launches (results do not depend on them)
        tiles = S.gemm_tiles()
        log(f"bench[{prec}]: autotuned tile shapes (BSE_L_K3!, BSE_K3!): {tiles}")
        ms_step, st = time_path(S, args, names)
        log(f"bench[{prec}]: {ms_step:.3f} ms/iteration; stages " + ", ".join(f"{n}={v:.3f}" for n, v in st.items()))
        # the two batched (p, a, t) GEMM kernels of one fdPA iteration, timed alone with HIP events
        t_bse = (st["gemm_L_K3"] + st["gemm_K3"]) * 1e-3
        # flops the GEMM launches execute: computed output tiles over the full contraction length
        f_exec = S.gemm_flops(True) + S.gemm_flops(False)
        paths[prec] = dict(ms=ms_step, st=st, obs=obs, it=it, res=res, gemm_avg_s=t_bse / 2.0, tiles=tiles,
                           f_exec_launch=f_exec / 2.0, bse_tflops=f_exec / t_bse / 1e12,
                           bse_tflops_dense_equiv=6 * f_gemm / t_bse / 1e12)
        if prec == "fp64" and not args.no_sym and not args.no_dense:
            # the same state with identity groups: every element evaluated
            S.reset_sym_grp()
            ms_d, st_d = time_path(S, args, names)
            log(f"bench[fp64, dense]: {ms_d:.3f} ms/iteration")
            paths["fp64_dense"] = dict(ms=ms_d, st=st_d)
        S.close()
    if len(run_paths) < 2:
        # profiling run of one path: per-stage line only
        q = run_paths[0]
        print(json.dumps({"path": q, "ms_per_step": paths[q]["ms"], "stage_ms": paths[q]["st"]}), flush=True)
        return
    errs = {k: max_relerr(paths["fp32"]["obs"][k], paths["fp64"]["obs"][k]) for k in paths["fp64"]["obs"]}
    log("bench: FP32-MFMA vs FP64 max rel.err " + ", ".join(f"{k}={v:.2e}" for k, v in errs.items()))
    fp32_ok = max(errs.values()) <= 1e-6
    head = "fp32" if fp32_ok else "fp64"
    P = paths[head]
    peak = FP32_MFMA_PEAK if head == "fp32" else FP64_MFMA_PEAK
    achieved = P["f_exec_launch"] / P["gemm_avg_s"]
    traffic = gemm_traffic(head, P["tiles"][1])
    out = {
        "workload": "c1 (secondary line)",
        "metric": "parquet iters/s + BSE TFLOP/s, FP32-MFMA vs FP64 GPU @ rel.err<=1e-6 vs FP64",
        "value": 1000.0 / P["ms"],
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": P["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mfma contraction, f64 elsewhere" if head == "fp32" else "f64",
        "data": "synthetic (SIAM model inputs; solver state at its fdPA fixed point)",
        "config": {"workload": "local SIAM fdPA, U=2, beta=20, nK1=128, nK2=(64,64), nK3=(32,64), mPinu=6, nG=256",
                   "symmetry_groups": "identity" if args.no_sym else "init_sym_grp! (frequency symmetries)",
                   "value_path": head,
                   "iteration_launch": "eager" if args.eager else "hipGraph replay of one captured iteration",
                   "fp32_mfma": {"iters_per_s": 1000.0 / paths["fp32"]["ms"], "bse_tflops": paths["fp32"]["bse_tflops"],
                                 "bse_tflops_dense_equiv": paths["fp32"]["bse_tflops_dense_equiv"],
                                 "stage_ms": paths["fp32"]["st"], "anderson_iterations": paths["fp32"]["it"],
                                 "gemm_tiles_L_K3_K3": paths["fp32"]["tiles"], "max_relerr_vs_fp64": errs},
                   "fp64": {"iters_per_s": 1000.0 / paths["fp64"]["ms"], "bse_tflops": paths["fp64"]["bse_tflops"],
                            "bse_tflops_dense_equiv": paths["fp64"]["bse_tflops_dense_equiv"],
                            "stage_ms": paths["fp64"]["st"], "anderson_iterations": paths["fp64"]["it"],
                            "gemm_tiles_L_K3_K3": paths["fp64"]["tiles"]},
                   "speedup_fp32_over_fp64": paths["fp64"]["ms"] / paths["fp32"]["ms"]},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                     "kernel": f"zgemm_tiles_t ({'FP32' if head == 'fp32' else 'FP64'} MFMA BSE_K3!/BSE_L_K3! contraction)"},
    }
    if not args.no_cpu_baseline:
        log("bench: timing the CPU baseline (oracle, 1 thread and all cores)")
        out["cpu_baseline"] = cpu_baseline_c1(sym=not args.no_sym)
    if "fp64_dense" in paths:
        out["config"]["fp64_dense"] = {"iters_per_s": 1000.0 / paths["fp64_dense"]["ms"],
                                       "stage_ms": paths["fp64_dense"]["st"]}
    print(json.dumps(out), flush=True)





if __name__ == "__main__":
    main()
