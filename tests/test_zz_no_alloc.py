"""fdga_fixed_point_noalloc (the fixed_point! residual staged through the library's file-scope device
scratch array, with no device allocation) against fdga_fixed_point on the same solver and x, called
through the C-ABI: two calls in a row, the second on a smaller state vector (the scratch keeps the
first call's contents), both also checked against the oracle; a state vector larger than the scratch
is refused with an error before any device work."""
import ctypes

import numpy as np
import pytest

import pyoracle as po
from helpers import make_pair, relerr

pytestmark = pytest.mark.gpu

import fdga  # noqa: E402
from fdga import _lib as L  # noqa: E402


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _noalloc(S, x, strategy, update):
    n = S.lib.fdga_flat_size(S.h) if update else S.lib.fdga_flat_size_F(S.h)
    x = np.ascontiguousarray(x, dtype=np.complex128)
    assert x.size == n
    R = np.empty(n, dtype=np.complex128)
    S._check(S.lib.fdga_fixed_point_noalloc(S.h, _dp(R.view(np.float64)), _dp(x.view(np.float64)), strategy,
                                            int(update), 1))
    return R


@pytest.mark.parametrize("first,second", [(("nl_L3", True), ("local_ref", False)),
                                          (("nl_L2", True), ("nl_L2", False))])
def test_noalloc_matches_fixed_point(first, second):
    sizes = []
    for k, (case, update) in enumerate((first, second)):
        print(f"noalloc: call {k} on {case} update_Σ={update}", flush=True)
        S, O = make_pair(case, seed=51 + k)
        n = S.lib.fdga_flat_size(S.h) if update else S.lib.fdga_flat_size_F(S.h)
        sizes.append(n)
        rng = np.random.default_rng(61 + k)
        x = O.flatten()[:n] + 0.05 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
        R_ref = fdga.fixed_point(np.empty(n, complex), x, S, "fdPA", update_Σ=update)
        R = _noalloc(S, x, L.FDPA, update)
        Ro = O.fixed_point(x, po.FDPA, update_sigma=update)
        e_dev, e_orc = relerr(R, R_ref), relerr(R, Ro)
        print(f"noalloc: call {k} n={n}: vs fdga_fixed_point {e_dev:.2e}, vs oracle {e_orc:.2e}", flush=True)
        assert e_dev < 1e-13 and e_orc < 1e-12
        S.close()
    assert sizes[1] < sizes[0]


def test_noalloc_refuses_a_state_beyond_the_scratch():
    # 12 x 12 lattice, K3 box (8, 16): 3 * 15 * 32^2 * 144 K3 elements of 16 bytes, above 64 MiB
    S = fdga.Solver(nG=32, nK1=16, nK2=(8, 16), nK3=(8, 16), T=0.3, U=1.0, mPinu=1, nonlocal_=True, LG=12, L_=12)
    n = S.lib.fdga_flat_size(S.h)
    assert 16 * n > 64 << 20
    R = np.empty(n, dtype=np.complex128)
    x = np.zeros(n, dtype=np.complex128)
    rc = S.lib.fdga_fixed_point_noalloc(S.h, _dp(R.view(np.float64)), _dp(x.view(np.float64)), L.FDPA, 1, 1)
    assert rc != 0
    with pytest.raises(fdga.FdgaError, match="exceeds the scratch"):
        S._check(rc)
    S.close()
