"""Shared set-up for the parity tests: one problem instance built identically in the HIP
solver (through the C-ABI) and in the CPU oracle."""
import numpy as np

import pyoracle as po

CASES = {
    # name: (nonlocal, LG, L, ref_kind, tables)
    "local_ref": (False, 1, 1, 0, False),
    "local_vertexref": (False, 1, 1, 1, True),
    "nl_L2": (True, 2, 2, 1, True),
    "nl_L3": (True, 3, 3, 0, True),
    "nl_LG4_L2": (True, 4, 2, 1, False),
}

# lattices used by the bubble, band and group tests only: L = 16 takes the radix-4 FFT of the
# fused lattice bubble, L = 24 the 3 x 8 FFT, L = 6 (even, half weights at |R| = L/2) and L = 5
# with LG = 10 the generic DFT
EXTRA_CASES = {
    "nl_L16": (True, 16, 16, 0, False),
    "nl_L24": (True, 24, 24, 0, False),
    "nl_L6": (True, 6, 6, 1, True),
    "nl_LG10_L5": (True, 10, 5, 0, False),
    # 16 x 16 lattice with a reference vertex: non-zero FL, and enough (Ω, P) batches for the
    # lattice contraction's two-launch form
    "nl_L16_ref": (True, 16, 16, 1, True),
}

BOX = dict(nG=10, nK1=6, nK2=(4, 4), nK3=(3, 3), mPinu=2, T=0.3, U=1.3)
REF_BOX = (7, (5, 4), (3, 3))
TAB = (2, 2)


def rand_c(rng, n, scale=0.2):
    return scale * (rng.standard_normal(n) + 1j * rng.standard_normal(n))


def make_pair(name, seed=0, fdpa_h0=True, box=None, precision="fp64"):
    """Returns (hip_solver, oracle) with identical random inputs for case `name`."""
    import fdga
    from fdga import _lib as L
    nonlocal_, LG, Lv, ref_kind, tables = CASES[name] if name in CASES else EXTRA_CASES[name]
    b = dict(BOX if box is None else box)
    rng = np.random.default_rng(seed)
    tab = TAB if tables else (1, 1)
    S = fdga.Solver(nG=b["nG"], nK1=b["nK1"], nK2=b["nK2"], nK3=b["nK3"], T=b["T"], U=b["U"], mPinu=b["mPinu"],
                    nonlocal_=nonlocal_, LG=LG, L_=Lv, ref_kind=ref_kind, ref_boxes=REF_BOX if ref_kind else None,
                    tab=tab, fdpa_h0=fdpa_h0, precision=precision)
    O = po.Oracle(nonlocal_=nonlocal_, nG=b["nG"], nK1=b["nK1"], nK2=b["nK2"], nK3=b["nK3"], mPinu=b["mPinu"],
                  T=b["T"], U=b["U"], LG=LG, L=Lv, ref_kind=ref_kind, ref_boxes=REF_BOX if ref_kind else None,
                  tab=tab, fdpa_h0=fdpa_h0)
    NG = 2 * b["nG"]
    if nonlocal_:
        Gb = fdga.hubbard_bare_Green(b["nG"], b["T"], LG, μ=0.3, t1=1.0, t2=-0.2)
    else:
        Gb = fdga.siam_bare_Green(b["nG"], b["T"], e=0.2, Δ=0.6, D=None)[:, None]
    G0 = Gb / (1.0 + Gb * rand_c(rng, Gb.shape, 0.3)) if ref_kind else np.zeros_like(Gb)
    S0 = rand_c(rng, Gb.shape, 0.1) if ref_kind else np.zeros_like(Gb)
    pairs = [(L.GBARE, po.GBARE, Gb), (L.G0, po.G0, G0), (L.SIGMA0, po.SIGMA0, S0)]
    for hw, ow, arr in pairs:
        S.set(hw, arr.reshape(-1, order="F"))
        O.set(ow, arr.reshape(-1, order="F"))
    if tables:
        n = (2 * TAB[0] - 1) * (2 * TAB[1]) ** 2
        for hw, ow in ((L.TAB_FPP, po.TAB_FPP), (L.TAB_FPX, po.TAB_FPX), (L.TAB_FTP, po.TAB_FTP), (L.TAB_FTX, po.TAB_FTX)):
            t = rand_c(rng, n, 0.3)
            S.set(hw, t)
            O.set(ow, t)
    if ref_kind:
        for ch in range(3):
            for cls in (1, 2, 3):
                n = O.get(po.vertex_id(po.VTX_F0, ch, cls)).size
                v = rand_c(rng, n, 0.3)
                S.set_vertex(3, ch, cls, v)
                O.set_vertex(po.VTX_F0, ch, cls, v)
    S.init()
    O.init()
    # a random target vertex and self-energy so that every term of the iteration is exercised
    for ch in range(3):
        for cls in (1, 2, 3):
            n = O.get(po.vertex_id(po.VTX_F, ch, cls)).size
            v = rand_c(rng, n, 0.2)
            S.set_vertex(0, ch, cls, v)
            O.set_vertex(po.VTX_F, ch, cls, v)
    sig = rand_c(rng, NG * (LG * LG), 0.05)
    S.set(L.SIGMA, sig)
    O.set(po.SIGMA, sig)
    return S, O


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale) if a.size else 0.0
