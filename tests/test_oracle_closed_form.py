"""Pins of the CPU oracle by closed forms, special cases and known answers.

Each test states what fixes the expected value independently of the oracle's own code:
  * Hubbard G0 at special k and occupations      test/test_hubbard.jl:25-33, 82-88
  * Dyson                                         test/test_hubbard.jl:46-62
  * bubble = product of Green functions          test/test_hubbard.jl:71-78 (+ the 1/ν tail
                                                  of src/bubble.jl:21-25)
  * real-space bubble = momentum convolution      convolution theorem (brute force, tiny L)
  * channel evaluator                             test/test_channel.jl:47-60
  * K1-only channel conversions (bosonic parts)   test/test_nonlocal_vertex.jl:148-154, s-wave form
  * first iteration from F = 0                    K1_r = s T Σ_ω U (Π - Π0) U (src/BSEa/BSEa_K1.jl:38-43)
  * NL solver with a 1x1 k-mesh and NL solver with k-independent G reduce to the local solver
  * symmetry classes of init_sym_grp! (local and lattice)   the group orbits by union-find from the
                                                  generators restated on value tuples; lowest-index
                                                  representatives; every operation that all paths
                                                  through a class agree on
Parity unpinned: the operation of a class member reached along paths with different operations
(the depth-first exploration order of MatsubaraFunctions' SymmetryGroup, not vendored here) is taken
from the first path found, the same hypothesis as the representative order.
"""
import numpy as np
import pytest

import pyoracle as po


def fval(T, n):
    return np.pi * T * (2 * n + 1)


# ------------------------------------------------------------------------------------------
def test_hubbard_bare_green_special_k():
    T, t1, mu, nG, LG = 0.5, 1.3, 0.2, 5, 8
    o = po.Oracle(nonlocal_=True, nG=nG, nK1=2, nK2=(1, 1), nK3=(1, 1), mPinu=1, T=T, U=1.0, LG=LG, L=1)
    o.hubbard(mu, t1)
    Gb = o.get(po.GBARE).reshape(2 * nG, LG, LG, order="F")
    for n in range(-2, 3):
        nu = fval(T, n)
        i = n + nG
        # k = (0,0), (0,π/2), (0,π), (0,3π/2), (π,π): indices (j1, j2) with k = 2π j / 8
        for (j1, j2), eps in [((0, 0), -4 * t1), ((0, 2), -2 * t1), ((0, 4), 0.0), ((0, 6), -2 * t1), ((4, 4), 4 * t1)]:
            np.testing.assert_allclose(-1j * Gb[i, j1, j2], 1 / (1j * nu + mu - eps), rtol=1e-13)


@pytest.mark.parametrize("mu,occ", [(-4.0, 0.0502663698543071), (-2.0, 0.2057188296739284), (0.0, 0.5),
                                    (4.0, 1 - 0.0502663698543071), (2.0, 1 - 0.2057188296739284)])
def test_hubbard_occupation_golden(mu, occ):
    o = po.Oracle(nonlocal_=True, nG=20, nK1=2, nK2=(1, 1), nK3=(1, 1), mPinu=1, T=0.5, U=1.0, LG=8, L=1)
    o.hubbard(mu, 1.0)
    assert abs(o.occupation(po.GBARE) - occ) < 1e-12


def test_dyson_closed_form():
    T, nG = 0.5, 5
    o = po.Oracle(nonlocal_=True, nG=nG, nK1=2, nK2=(1, 1), nK3=(1, 1), mPinu=1, T=T, U=1.0, LG=8, L=1)
    o.hubbard(0.2, 1.3)
    Gb = o.get(po.GBARE)
    o.set(po.SIGMA, np.full(Gb.size, -0.5 + 0.2j))
    o.dyson()
    G = o.get(po.G)
    # -i G satisfies G^-1 = Gbare^-1 - Σ for the physical (-i) quantities
    np.testing.assert_allclose(1 / (-1j * G), 1 / (-1j * Gb) - (-1j * (-0.5 + 0.2j)), rtol=1e-13)


# ------------------------------------------------------------------------------------------
def test_local_bubble_is_product_with_tail():
    T, nG, nK1, mp = 0.1, 6, 4, 3
    o = po.Oracle(nG=nG, nK1=nK1, nK2=(2, 2), nK3=(2, 2), mPinu=mp, T=T, U=1.0)
    o.siam(0.1, np.pi / 5, 10.0)
    o.init()
    G = o.get(po.G)
    Nf = nK1 * mp

    def g(n):
        return G[n + nG] if -nG <= n < nG else 1 / fval(T, n)
    pp = o.get(po.PIPP).reshape(2 * nK1 - 1, 2 * Nf, order="F")
    ph = o.get(po.PIPH).reshape(2 * nK1 - 1, 2 * Nf, order="F")
    for m in range(-(nK1 - 1), nK1):
        for n in range(-Nf, Nf):
            # Ω - ν is fermionic index m - n - 1, Ω + ν is m + n
            assert pp[m + nK1 - 1, n + Nf] == pytest.approx(g(n) * g(m - n - 1), rel=1e-14)
            assert ph[m + nK1 - 1, n + Nf] == pytest.approx(g(n) * g(m + n), rel=1e-14)


@pytest.mark.parametrize("L", [3, 4])
def test_nl_bubble_real_space_is_momentum_convolution(L):
    """Π(Ω,ν,P) = (1/N) Σ_k G(ν,k) G(Ω∓ν, P∓k) for frequencies inside the G mesh
    (the 1/ν tail enters only the R = 0 term and only out of the mesh)."""
    T, nG, nK1, mp = 0.5, 6, 3, 1
    o = po.Oracle(nonlocal_=True, nG=nG, nK1=nK1, nK2=(1, 1), nK3=(1, 1), mPinu=mp, T=T, U=1.0, LG=L, L=L)
    o.hubbard(0.2, 1.0, -0.3)
    o.init()
    G = o.get(po.G).reshape(2 * nG, L, L, order="F")
    Nf = nK1 * mp
    NW = 2 * nK1 - 1
    pp = o.get(po.PIPP).reshape(NW, 2 * Nf, L, L, order="F")
    ph = o.get(po.PIPH).reshape(NW, 2 * Nf, L, L, order="F")
    for m in range(-(nK1 - 1), nK1):
        for n in range(-Nf, Nf):
            a, b = m - n - 1, m + n
            for p1 in range(L):
                for p2 in range(L):
                    if -nG <= a < nG and -nG <= n < nG:
                        ref = np.mean([G[n + nG, k1, k2] * G[a + nG, (p1 - k1) % L, (p2 - k2) % L]
                                       for k1 in range(L) for k2 in range(L)])
                        assert abs(pp[m + nK1 - 1, n + Nf, p1, p2] - ref) < 1e-13
                    if -nG <= b < nG and -nG <= n < nG:
                        ref = np.mean([G[n + nG, k1, k2] * G[b + nG, (p1 + k1) % L, (p2 + k2) % L]
                                       for k1 in range(L) for k2 in range(L)])
                        assert abs(ph[m + nK1 - 1, n + Nf, p1, p2] - ref) < 1e-13


# ------------------------------------------------------------------------------------------
def _random_solver(nonlocal_=False, L=1, seed=0, **kw):
    rng = np.random.default_rng(seed)
    args = dict(nG=8, nK1=5, nK2=(4, 3), nK3=(3, 3), mPinu=2, T=0.5, U=3.0)
    args.update(kw)
    o = po.Oracle(nonlocal_=nonlocal_, LG=L, L=L, **args)
    for ch in range(3):
        for cls in (1, 2, 3):
            n = o.get(po.vertex_id(po.VTX_F, ch, cls)).size
            o.set_vertex(po.VTX_F, ch, cls, rng.standard_normal(n) + 1j * rng.standard_normal(n))
    return o, args


def test_channel_evaluator_identities():
    """test/test_channel.jl:47-60, through F restricted to one channel (same channel: no conversion)."""
    o, a = _random_solver()
    nK1 = a["nK1"]
    nK2b, nK2f = a["nK2"]
    nK3b, nK3f = a["nK3"]
    for ch, flag in [(po.PCH, 2), (po.TCH, 4), (po.ACH, 8)]:
        K1 = o.vertex(po.VTX_F, ch, 1)
        K2 = o.vertex(po.VTX_F, ch, 2).reshape(2 * nK2b - 1, 2 * nK2f, order="F")
        K3 = o.vertex(po.VTX_F, ch, 3).reshape(2 * nK3b - 1, 2 * nK3f, 2 * nK3f, order="F")
        for m in range(-(nK3b - 1), nK3b):
            for n in range(-nK3f, nK3f):
                for npp in range(-nK3f, nK3f):
                    v = o.eval(po.VTX_F, m, n, npp, 0, ch, po.PSP, flag)
                    ref = (K1[m + nK1 - 1] + K2[m + nK2b - 1, n + nK2f] + K2[m + nK2b - 1, npp + nK2f]
                           + K3[m + nK3b - 1, n + nK3f, npp + nK3f])
                    assert abs(v - ref) < 1e-13
                    assert abs(o.eval(po.VTX_F, m, po.INF, npp, 0, ch, po.PSP, flag)
                               - (K1[m + nK1 - 1] + K2[m + nK2b - 1, npp + nK2f])) < 1e-13
                    assert abs(o.eval(po.VTX_F, m, n, po.INF, 0, ch, po.PSP, flag)
                               - (K1[m + nK1 - 1] + K2[m + nK2b - 1, n + nK2f])) < 1e-13
            assert abs(o.eval(po.VTX_F, m, po.INF, po.INF, 0, ch, po.PSP, flag) - K1[m + nK1 - 1]) < 1e-13
        # outside every box only K1 (if in its box) survives; far outside: zero
        for m in (-100, 100):
            assert o.eval(po.VTX_F, m, 0, -1, 0, ch, po.PSP, flag) == 0


def test_nl_k1_only_channel_conversions():
    """test/test_nonlocal_vertex.jl:148-154 in the s-wave form used on the path:
    F(Ω,ν,ω,P,kSW,kSW,r,pSp) = U + K1_r(Ω,P) + mean_q K1_r'(Ω_r'(Ω,ν,ω), q) + mean_q K1_r''(...)."""
    L, T, U = 3, 0.5, 3.0
    rng = np.random.default_rng(1)
    o = po.Oracle(nonlocal_=True, nG=4, nK1=10, nK2=(4, 3), nK3=(2, 1), mPinu=1, T=T, U=U, LG=L, L=L)
    K1 = {}
    for ch in range(3):
        k = rng.standard_normal(19 * L * L) + 1j * rng.standard_normal(19 * L * L)
        o.set_vertex(po.VTX_F, ch, 1, k)
        K1[ch] = k.reshape(19, L * L, order="F")
    nK1 = 10
    k1 = lambda ch, m, P: K1[ch][m + nK1 - 1, P] if abs(m) < nK1 else 0
    k1avg = lambda ch, m: K1[ch][m + nK1 - 1].mean() if abs(m) < nK1 else 0
    for (m, n, w) in [(1, 2, -1), (0, 5, 3), (-3, -2, 4), (2, -7, -6)]:
        for P in (0, 4):
            # conversions (bosonic components): p→t Ω-ν-ω = m-n-w-1, p→a ν-ω = n-w,
            # t→p Ω+ν+ω = m+n+w+1, t→a ω-ν, a→p Ω+ν+ω, a→t ν-ω
            ref_p = U + k1(po.PCH, m, P) + k1avg(po.TCH, m - n - w - 1) + k1avg(po.ACH, n - w)
            ref_t = U + k1(po.TCH, m, P) + k1avg(po.PCH, m + n + w + 1) + k1avg(po.ACH, w - n)
            ref_a = U + k1(po.ACH, m, P) + k1avg(po.PCH, m + n + w + 1) + k1avg(po.TCH, n - w)
            for ch, ref in [(po.PCH, ref_p), (po.TCH, ref_t), (po.ACH, ref_a)]:
                assert abs(o.eval(po.VTX_F, m, n, w, P, ch, po.PSP) - ref) < 1e-12
            # νInf: only the same channel (src/vertex.jl:246-284)
            assert abs(o.eval(po.VTX_F, m, po.INF, po.INF, P, po.PCH, po.PSP) - (U + k1(po.PCH, m, P))) < 1e-12


# ------------------------------------------------------------------------------------------
def test_first_iteration_k1_closed_form():
    """From F = 0: K1_a = T U² Σ_ω (Πph-Π0ph), K1_p = T U² Σ_ω (Πpp-Π0pp) and K1_t = 0
    (the dSp BSE gives -K1_a, and γt = (γt_d + γa)/2)."""
    T, U, nG, nK1, mp = 0.1, 1.3, 10, 6, 2
    o = po.Oracle(nG=nG, nK1=nK1, nK2=(3, 3), nK3=(2, 2), mPinu=mp, T=T, U=U)
    o.siam(0.2, np.pi / 5, None)
    o.init()
    o.build_k3_cache()
    for ch in (po.PCH, po.ACH, po.TCH):
        o.bse_k1(ch)
    NW, NF = 2 * nK1 - 1, 2 * nK1 * mp
    pp = (o.get(po.PIPP) - o.get(po.PI0PP)).reshape(NW, NF, order="F")
    ph = (o.get(po.PIPH) - o.get(po.PI0PH)).reshape(NW, NF, order="F")
    np.testing.assert_allclose(o.vertex(po.VTX_FBUFF, po.ACH, 1), T * U * U * ph.sum(axis=1), rtol=1e-12)
    np.testing.assert_allclose(o.vertex(po.VTX_FBUFF, po.PCH, 1), T * U * U * pp.sum(axis=1), rtol=1e-12)
    assert np.max(np.abs(o.vertex(po.VTX_FBUFF, po.TCH, 1))) < 1e-15


def _local_and_nl(L, t_zero):
    """A local SIAM-like solver and an NL solver whose G is k-independent; both start from the
    same k-independent G0 = local G and random P-independent vertex."""
    T, U = 0.2, 1.5
    box = dict(nG=12, nK1=6, nK2=(4, 4), nK3=(3, 3), mPinu=2, T=T, U=U)
    loc = po.Oracle(**box)
    nl = po.Oracle(nonlocal_=True, LG=L, L=L, **box)
    loc.siam(0.3, 0.7, None)
    gb = loc.get(po.GBARE)
    nl.set(po.GBARE, np.tile(gb, L * L))
    rng = np.random.default_rng(5)
    for ch in range(3):
        for cls in (1, 2, 3):
            n = loc.get(po.vertex_id(po.VTX_F, ch, cls)).size
            v = 0.1 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
            loc.set_vertex(po.VTX_F, ch, cls, v)
            nl.set_vertex(po.VTX_F, ch, cls, np.tile(v, L * L))
    loc.init()
    nl.init()
    return loc, nl


@pytest.mark.parametrize("L", [1, 2, 3])
def test_nl_with_k_independent_data_is_local(L):
    loc, nl = _local_and_nl(L, True)
    for strategy in (po.SCPA, po.FDPA):
        loc.iterate(strategy)
        nl.iterate(strategy)
        for ch in range(3):
            for cls in (1, 2, 3):
                a = loc.vertex(po.VTX_F, ch, cls)
                b = nl.vertex(po.VTX_F, ch, cls).reshape(a.size, L * L, order="F")
                assert np.max(np.abs(b - a[:, None])) < 1e-12, (strategy, ch, cls)
        S = nl.get(po.SIGMA).reshape(-1, L * L, order="F")
        assert np.max(np.abs(S - loc.get(po.SIGMA)[:, None])) < 1e-12


def test_symmetrized_solution_obeys_symmetries():
    """init_sym_grp! symmetries hold exactly in the symmetrized iteration (src/symmetries.jl)."""
    nmax, T = 4, 0.1
    o = po.Oracle(nG=6 * nmax, nK1=4 * nmax, nK2=(nmax, nmax), nK3=(nmax, nmax), mPinu=2, T=T, U=1.0)
    o.siam(0.4, np.pi / 5, 10.0)
    o.init()
    o.init_sym_grp()
    for _ in range(3):
        o.iterate(po.SCPA)
    nK1 = 4 * nmax
    for ch in range(3):
        K1 = o.vertex(po.VTX_F, ch, 1)
        np.testing.assert_allclose(K1[::-1], np.conj(K1), atol=1e-14)   # sK1: K1(-Ω) = K1(Ω)*
        K2 = o.vertex(po.VTX_F, ch, 2).reshape(2 * nmax - 1, 2 * nmax, order="F")
        np.testing.assert_allclose(K2[::-1, ::-1], np.conj(K2), atol=1e-14)   # sK2pp1/ph1
    S = o.get(po.SIGMA)
    np.testing.assert_allclose(S[::-1], -np.conj(S), atol=1e-14)   # sΣ: iΣ(-ν) = -(iΣ(ν))*



# ---------------------------------------------------------------- symmetry classes
# init_sym_grp! (src/ParquetSolver.jl:203-231; NL src/nonlocal/ParquetSolver.jl:198-300) with the
# generators of src/symmetries.jl and src/nonlocal/symmetries.jl, restated here on value tuples
# (momenta as BrillouinPoint pairs (j1, j2); linear index j1 + L j2, first mesh index fastest).
# The class partition is checked against the group's orbits computed by union-find (independent
# of the exploration order), the representative against the lowest linear index, and each
# member's operation wherever all paths through the class give the same one.
SYM_BOX = dict(nG=8, nK1=5, nK2=(3, 3), nK3=(2, 2), mPinu=2, T=0.25, U=1.2)
SGN, CON = 1, 2


def _sym_groups(L, nonlocal_):
    b = SYM_BOX
    F = lambda N: ("f", N)
    B = lambda N: ("b", N)
    fm = lambda n: -n - 1                  # -ν
    neg = lambda k: ((-k[0]) % L, (-k[1]) % L)
    ref = lambda k: (k[1], k[0])           # _ref: (x, y) -> (y, x)
    rot = lambda k: (k[1], (-k[0]) % L)    # _rot: (x, y) -> fold_back((y, -x))

    def g(freq, kmap):
        if not nonlocal_:
            return lambda w: freq(w)
        return lambda w: (lambda r: (r[0] + (kmap(w[-1]),), r[1]))(freq(w[:-1]))

    spatial = [g(lambda w: (w, 0), ref), g(lambda w: (w, 0), rot)] if nonlocal_ else []
    sS = g(lambda w: ((fm(w[0]),), SGN | CON), neg)
    sK1 = g(lambda w: ((-w[0],), CON), neg)
    sK2pp1 = g(lambda w: ((-w[0], fm(w[1])), CON), neg)
    sK2pp2 = g(lambda w: ((w[0], w[0] - w[1] - 1), 0), lambda k: k)
    sK2ph2 = g(lambda w: ((-w[0], w[0] + w[1]), 0), neg)
    sK3pp1 = g(lambda w: ((-w[0], fm(w[1]), fm(w[2])), CON), neg)
    sK3pp2 = g(lambda w: ((w[0], w[2], w[1]), 0), lambda k: k)
    sK3pp3 = g(lambda w: ((w[0], w[0] - w[1] - 1, w[0] - w[2] - 1), 0), lambda k: k)
    sK3ph3 = g(lambda w: ((-w[0], w[0] + w[1], w[0] + w[2]), 0), neg)
    dS, d1 = [F(b["nG"])], [B(b["nK1"])]
    d2 = [B(b["nK2"][0]), F(b["nK2"][1])]
    d3 = [B(b["nK3"][0]), F(b["nK3"][1]), F(b["nK3"][1])]
    return {0: (dS, [sS]), 1: (d1, [sK1]), 2: (d2, [sK2pp1, sK2pp2]), 3: (d3, [sK3pp1, sK3pp2, sK3pp3]),
            4: (d3, [sK3pp1, sK3pp3]), 5: (d2, [sK2pp1, sK2ph2]), 6: (d3, [sK3pp1, sK3pp2, sK3ph3]),
            7: (d3, [sK3pp1, sK3ph3]), 8: (d2, [sK2pp1, sK2pp2]), 9: (d2, [sK2pp1, sK2ph2])}, spatial


def _mesh_points(dims, L, nonlocal_):
    axes = []
    for kind, N in dims:
        axes.append(list(range(-(N - 1), N)) if kind == "b" else list(range(-N, N)))
    if nonlocal_:
        axes.append([(j % L, j // L) for j in range(L * L)])
    pts = [()]
    for ax in axes:                        # first index fastest
        pts = [p + (v,) for v in ax for p in pts]
    index = {p: i for i, p in enumerate(pts)}
    return pts, index


@pytest.mark.parametrize("nonlocal_", [False, True])
def test_symmetry_classes_are_the_group_orbits(nonlocal_):
    L = 4 if nonlocal_ else 1
    o = po.Oracle(nonlocal_=nonlocal_, LG=L, L=L, **SYM_BOX)
    if nonlocal_:
        o.hubbard(0.3, 1.0, -0.2)
    else:
        o.siam(0.3, 0.7, None)
    o.init()
    o.init_sym_grp()
    groups, spatial = _sym_groups(L, nonlocal_)
    for gid, (dims, gens) in groups.items():
        gens = gens + spatial
        pts, index = _mesh_points(dims, L, nonlocal_)   # Σ's momentum mesh is G's: LG = L here
        n = len(pts)
        edges = []
        for i, w in enumerate(pts):
            for gen in gens:
                w2, op = gen(w)
                j = index.get(w2)
                if j is not None:
                    edges.append((i, j, op))
        parent = list(range(n))

        def find(x):
            while parent[x] != x:
                parent[x] = parent[parent[x]]
                x = parent[x]
            return x
        for i, j, _ in edges:
            ri, rj = find(i), find(j)
            if ri != rj:
                parent[max(ri, rj)] = min(ri, rj)
        root = np.array([find(i) for i in range(n)])
        rep, op = o.sym_classes(gid)
        assert rep.size == n, gid
        # representatives: the lowest linear index of the orbit (find() returns the orbit minimum)
        np.testing.assert_array_equal(rep, root, err_msg=f"group {gid}")
        # operations: forced wherever every path in the class gives the same one
        adj = [[] for _ in range(n)]
        for i, j, o_ in edges:
            adj[i].append((j, o_))
            adj[j].append((i, o_))
        forced = np.full(n, -1)
        conflict = set()
        for r in np.unique(root):
            forced[r] = 0
            stack = [r]
            while stack:
                x = stack.pop()
                for y, o_ in adj[x]:
                    want = forced[x] ^ o_
                    if forced[y] < 0:
                        forced[y] = want
                        stack.append(y)
                    elif forced[y] != want:
                        conflict.add(r)
        ok = ~np.isin(root, list(conflict))
        assert ok.sum() > n // 4, f"group {gid}: too few classes with a unique operation"
        np.testing.assert_array_equal(op[ok], forced[ok], err_msg=f"group {gid}")
