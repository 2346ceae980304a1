#!/usr/bin/env python3
"""Benchmark: parquet iterations/s and BSE TFLOP/s of the FP32-MFMA and FP64 GPU paths on one MI355X.

Workload (one "step" = one iterate_solver! pass, src/solve.jl:4-116, strategy :fdPA):
SIAM with G0(iν) = 1/(iν + iΔ sign ν), Δ = π/5, U = 2, β = 20 (T = 0.05), nK1 = 128,
nK2 = (64, 64), nK3 = (32, 64), mΠν_factor = 6, nG = 256.  The solver is first brought to
its fixed point with the Anderson loop (untimed), then `warmup` passes run and `steps` passes
are timed with HIP events on the solver's stream, all inputs resident in HBM.

Both precision paths are run: the FP64 solver to its fixed point, then the FP32-MFMA solver
(the BSE_L_K3!/BSE_K3! GEMMs in FP32 on MFMA, everything else FP64) from that point to its own
fixed point; the max relative error of Σ, χ_sp, χ_ch and K3 between the two is reported.
`value` is the FP32-MFMA path's iterations/s when that error is <= 1e-6, else the FP64 path's
(`config.value_path` says which).

Prints one JSON line.  `roofline` is for the BSE contraction kernel of the headline path
(BSE_K3! + BSE_L_K3!, batched complex GEMM on MFMA, tile shapes chosen by the autotuner): flops per
launch = 8 * BM * BN * Nν3 per computed BM x BN output tile (all tiles with identity groups; the
tiles that hold a class representative under init_sym_grp!) over the measured average launch time.
`bse_tflops_dense_equiv` divides the dense count 8 NΩ3 Nν3^3 per channel by the same time.  `cpu_baseline` times the CPU restatement
(oracle/, one thread) on one iteration of the same workload.
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


def cpu_baseline(sym=True, reps=3):
    """The CPU restatement of the reference loops (oracle/, one thread) on the same workload:
    fdPA iterate_solver! passes of the full box from the solver's initial state, with the same
    symmetry groups as the GPU run; one warm-up pass, then the median of `reps` passes."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    w = WORKLOAD
    o = po.Oracle(nG=w["nG"], nK1=w["nK1"], nK2=w["nK2"], nK3=w["nK3"], mPinu=w["mPinu"], T=w["T"], U=w["U"])
    o.siam(0.0, w["Delta"], None)
    o.init()
    if sym:
        o.init_sym_grp()
    o.iterate(po.FDPA)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        o.iterate(po.FDPA)
        ts.append(time.perf_counter() - t0)
    dt = sorted(ts)[len(ts) // 2]
    return {"value": 1.0 / dt, "unit": "iters/s", "cores": 1, "kind": "port",
            "sample": f"fdPA iterate_solver! passes of the full workload box on the CPU restatement "
                      f"(oracle/fdga_oracle.cpp, 1 thread{', init_sym_grp!' if sym else ''}): 1 warm-up, "
                      f"median of {reps} = {dt:.2f} s"}


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
    hipGraph of one iteration unless --eager), then a per-stage pass (eager, one stage at a time)."""
    import fdga
    S.set_graph(not args.eager)
    for _ in range(args.warmup):
        fdga.iterate_solver(S, "fdPA")
    S.sync()
    ms = S.time_iterations("fdPA", args.steps)
    S.set_graph(False)
    nst = max(1, min(args.steps, 5))
    S.set_stage_timing(True)
    for _ in range(nst):
        fdga.iterate_solver(S, "fdPA")
    S.sync()
    st = [x / nst for x in S.stage_ms()]
    S.set_stage_timing(False)
    return ms / args.steps, dict(zip(names, st))


C2 = dict(L=16, T=0.2, U=2.0, mu=0.0, t1=1.0, nK1=64, nK2=(32, 32), nK3=(16, 32), mPinu=32, nG=256)


def build_c2(precision, w=C2, seed=0):
    """Nonlocal 2D Hubbard fdPA solver (NL_ParquetSolver, src/nonlocal/ParquetSolver.jl:77-152) on an
    L x L lattice with a synthetic local reference Vertex: K1 = K2 = 0 and K3 a complex Gaussian
    (numpy PCG64, `seed`) times 0.1 U / (1 + |ν| + |ν'|) (ν in units of πT), G0 the bare lattice G,
    Σ0 = 0."""
    import numpy as np
    import fdga
    L = w["L"]
    Gb = fdga.hubbard_bare_Green(w["nG"], w["T"], L, μ=w["mu"], t1=w["t1"])
    nb, nf = w["nK3"]
    rng = np.random.default_rng(seed)
    NW, NV = 2 * nb - 1, 2 * nf
    n = np.arange(-nf, nf)
    amp = 0.1 * w["U"] / (1 + np.abs(2 * n[None, :] + 1) + np.abs(2 * n[:, None] + 1))   # [ν', ν] -> (ν, ν') symmetric
    F0 = {}
    for ch in range(3):
        k3 = (rng.standard_normal((NW, NV, NV)) + 1j * rng.standard_normal((NW, NV, NV))) * amp[None, :, :]
        F0[(ch, 1)] = np.zeros(2 * w["nK1"] - 1, complex)
        F0[(ch, 2)] = np.zeros((2 * w["nK2"][0] - 1) * 2 * w["nK2"][1], complex)
        F0[(ch, 3)] = k3.reshape(-1, order="F")
    S = fdga.NL_ParquetSolver(w["nK1"], w["nK2"], w["nK3"], L, Gb, Gb, np.zeros_like(Gb), F0, T=w["T"], U=w["U"],
                              mΠν_factor=w["mPinu"], F0_boxes=(w["nK1"], w["nK2"], w["nK3"]), precision=precision)
    return S


def main_c2(args):
    """One-GPU timing of the C2 lattice iteration in both precisions, from the same state; the
    FP32-MFMA path's relative error against FP64 after one iteration."""
    import numpy as np
    import fdga
    names = ["dyson+bubbles", "k3_cache", "L_K2", "L_K3", "K1", "K2", "K3", "update", "SDE", "gemm_L_K3", "gemm_K3"]
    w = C2
    NW3, NV3, NP = 2 * w["nK3"][0] - 1, 2 * w["nK3"][1], w["L"] ** 2
    f_dense = 2 * 3 * 8.0 * NW3 * NP * NV3 ** 3
    res, x0, first = {}, None, {}
    for prec in ("fp64", "fp32"):
        t0 = time.perf_counter()
        S = build_c2(prec)
        log(f"bench c2[{prec}]: solver built ({time.perf_counter() - t0:.1f} s)")
        if x0 is None:
            x0 = fdga.flatten(S)
        else:
            fdga.unflatten(S, x0)
        fdga.iterate_solver(S, "fdPA")
        S.sync()
        first[prec] = fdga.flatten(S)
        fdga.unflatten(S, x0)
        ms, st = time_path(S, args, names)
        log(f"bench c2[{prec}]: {ms:.2f} ms/iteration; stages " + ", ".join(f"{k}={v:.2f}" for k, v in st.items()))
        t_bse = (st["gemm_L_K3"] + st["gemm_K3"]) * 1e-3
        res[prec] = {"ms_per_step": ms, "iters_per_s": 1000.0 / ms, "stage_ms": st,
                     "bse_tflops": f_dense / t_bse / 1e12}
        if prec == "fp32" and not args.no_sym:
            # the same iteration with init_sym_grp!'s lattice groups installed (src/nonlocal/ParquetSolver.jl:
            # 198-300): a different result on this workload, whose synthetic reference K3 is not symmetric
            t0 = time.perf_counter()
            S.init_sym_grp()
            log(f"bench c2[{prec}, lattice groups]: groups built ({time.perf_counter() - t0:.1f} s)")
            fdga.unflatten(S, x0)
            ms_g, st_g = time_path(S, args, names)
            log(f"bench c2[{prec}, lattice groups]: {ms_g:.2f} ms/iteration; stages "
                + ", ".join(f"{k}={v:.2f}" for k, v in st_g.items()))
            res["fp32_groups"] = {"ms_per_step": ms_g, "iters_per_s": 1000.0 / ms_g, "stage_ms": st_g,
                                  "bse_gemm_flops_executed": S.gemm_flops(1) + S.gemm_flops(0)}
        S.close()
    a, b = first["fp32"], first["fp64"]
    rel = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    print(json.dumps({"workload": "c2: NL Hubbard fdPA, L=16, beta=5, U=2, nK1=64, nK2=(32,32), nK3=(16,32), mPinu=32, "
                                  "nG=256, synthetic local reference vertex", "steps": args.steps, "warmup": args.warmup,
                      "fp64": res["fp64"], "fp32_mfma": res["fp32"],
                      "fp32_mfma_with_lattice_groups": res.get("fp32_groups"),
                      "fp32_vs_fp64_state_relerr_after_one_iteration": rel}), flush=True)


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
    ap.add_argument("--workload", default="c1", choices=["c0", "c1", "c2", "c4"],
                    help="c1: the headline SIAM workload (one JSON line of the metric); c2: the NL Hubbard lattice "
                         "iteration (timing and FP32 error, separate JSON line); c0: the small SIAM solve (hipGraph "
                         "replay and Newton–Krylov vs mixing, separate JSON line); c4: the 8-point lattice sweep in one "
                         "solver group (grouped GEMM launches, autotuned tiles, hipGraph; separate JSON line)")
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
    import fdga
    w = WORKLOAD
    names = ["dyson+bubbles", "k3_cache", "L_K2", "L_K3", "K1", "K2", "K3", "update", "SDE", "gemm_L_K3", "gemm_K3"]
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
        log("bench: timing the CPU baseline (oracle, one thread)")
        out["cpu_baseline"] = cpu_baseline(sym=not args.no_sym)
    if "fp64_dense" in paths:
        out["config"]["fp64_dense"] = {"iters_per_s": 1000.0 / paths["fp64_dense"]["ms"],
                                       "stage_ms": paths["fp64_dense"]["st"]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
