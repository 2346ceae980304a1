"""The gather of build_K3_cache! as the linear map that the adjoint's segmented sums use
(fdga_k3_gather_map, built on the host): applied to a random vertex it reproduces the oracle's caches.

With U = 0 and no reference vertex the caches are linear in F, so the map is pinned entry by entry: the
local solver's seven caches (Γpx, Γpp, Γa, Γt(d), Fp, Fa, Ft(d), whole), and on a 2 x 2 lattice the
caches of the momentum-averaged vertex (the F caches less their same-channel K2 + K3 at each momentum,
which the lattice contraction adds itself).  Boxes with unequal sizes catch a transposed index.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import pyoracle as po

CASES = {
    "local": dict(nonlocal_=False, L=1, nK1=5, nK2=(4, 3), nK3=(2, 3)),
    "lattice": dict(nonlocal_=True, L=2, nK1=5, nK2=(3, 4), nK3=(3, 2)),
}
SLOT_CACHE = (0, 4, 5, 6, 7, 8, 9)


def _k3_dev(ref, nb, nf):
    """K3 [Ω, ν, νp] (reference layout) -> [ν, νp, Ω] (device layout), flattened."""
    return np.transpose(ref.reshape(2 * nb - 1, 2 * nf, 2 * nf, order="F"), (1, 2, 0)).reshape(-1, order="F")


@pytest.mark.parametrize("case", list(CASES))
def test_gather_map_reproduces_the_oracle_caches(case):
    import fdga
    c = CASES[case]
    nK1, nK2, nK3, NP = c["nK1"], c["nK2"], c["nK3"], c["L"] ** 2
    O = po.Oracle(nonlocal_=c["nonlocal_"], nG=8, nK1=nK1, nK2=nK2, nK3=nK3, mPinu=1, T=0.4, U=0.0,
                  LG=c["L"], L=c["L"])
    rng = np.random.default_rng(3)
    avg = []
    own = {}
    for ch in (po.PCH, po.TCH, po.ACH):
        parts = []
        for cls in (1, 2, 3):
            n = O.vertex(po.VTX_F, ch, cls).size
            v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
            O.set_vertex(po.VTX_F, ch, cls, v)
            a = v.reshape(-1, NP, order="F")
            if cls == 3:
                own[ch, 3] = a
                parts.append(_k3_dev(a.mean(axis=1), *nK3))
            else:
                if cls == 2:
                    own[ch, 2] = a
                parts.append(a.mean(axis=1))
        avg.append(np.concatenate(parts))
    avg = np.concatenate(avg)
    O.build_k3_cache()

    off, src, cf = fdga.k3_gather_map(nK1, nK2, nK3, lattice=c["nonlocal_"])
    NW, NV = 2 * nK3[0] - 1, 2 * nK3[1]
    n3 = NW * NV * NV
    assert off.size == avg.size + 1 and off[-1] == src.size
    dst = np.repeat(np.arange(avg.size), np.diff(off))
    M = sp.csr_matrix((cf.astype(np.float64), (src, dst)), shape=(7 * n3, avg.size))
    caches = M @ avg

    NW2, NV2 = 2 * nK2[0] - 1, 2 * nK2[1]
    for s, k in enumerate(SLOT_CACHE):
        ref = O.get(po.CACHE0 + k).reshape(NW, NV, NV, NP, order="F")   # [Ω, a, b, P]
        mine = np.transpose(caches[s * n3:(s + 1) * n3].reshape(NV, NV, NW, order="F"), (2, 0, 1))
        for p in range(NP):
            want = ref[..., p].copy()
            if c["nonlocal_"] and k >= 7:
                # the same-channel K2(Ω, ν) + K3(Ω, ν, ω) at this momentum (t: 2 (t) - (a))
                def own_of(ch):
                    k2 = own[ch, 2][:, p].reshape(NW2, NV2, order="F")
                    k3 = own[ch, 3][:, p].reshape(NW, NV, NV, order="F")
                    o2 = nK2[0] - nK3[0]
                    f2 = nK2[1] - nK3[1]
                    return k2[o2:o2 + NW, f2:f2 + NV][:, :, None] + k3
                o = {7: own_of(po.PCH), 8: own_of(po.ACH), 9: 2 * own_of(po.TCH) - own_of(po.ACH)}[k]
                want = want - o
            err = np.abs(mine - want).max() / np.abs(want).max()
            assert err < 1e-13, f"{case} cache {po.CACHE_NAMES[k]} P={p}: {err:.3e}"


def test_gather_map_is_destination_sorted_and_bounded():
    """Every cache entry reads each vertex element with a small integer coefficient, and the
    segments cover the entries exactly once."""
    import fdga
    off, src, cf = fdga.k3_gather_map(6, (5, 4), (3, 3), lattice=True)
    assert np.all(np.diff(off) >= 0) and off[0] == 0 and off[-1] == src.size
    assert src.min() >= 0 and src.max() < 7 * 5 * 6 * 6
    assert set(np.unique(cf)).issubset({-2.0, -1.0, 1.0, 2.0})
