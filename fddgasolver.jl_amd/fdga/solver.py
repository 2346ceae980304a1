"""Host side of the MI355X parquet solver, mirroring fdDGAsolver.jl's solver API.

Names follow the reference (src/fdDGAsolver.jl:180-205) with Julia's `!` dropped:
parquet_solver_siam_parquet_approximation, parquet_solver_hubbard_parquet_approximation,
ParquetSolver, NL_ParquetSolver, iterate_solver, fixed_point, solve, Dyson, bubbles,
build_K3_cache, build_K3_cache_mfRG, BSE_L_K2, BSE_L_K3, BSE_K1, BSE_K2, BSE_K3, BSE_K1_new, BSE_K2_new,
SDE, flatten, unflatten, mfRGLinearMap.
All numerical work runs on the GPU through libfdga.so; this module only allocates the
context, converts arguments and drives the fixed-point loop.
"""
import ctypes

import numpy as np

from . import _lib as L

pCh, tCh, aCh = L.PCH, L.TCH, L.ACH
_STRATEGIES = {"scPA": L.SCPA, "fdPA": L.FDPA, "scPA_new": L.SCPA_NEW, "fdPA_new": L.FDPA_NEW}
# a strategy iterate_solver! accepts (src/solve.jl:10) whose BSE forms are not on the device path: the
# one-loop K3 of :fdPA_1loop (src/solve.jl:47-58)
_STRATEGIES_NOT_BUILT = ("fdPA_1loop",)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class FdgaError(RuntimeError):
    pass


class Solver:
    """Device-resident ParquetSolver (local) or NL_ParquetSolver (s-wave lattice).

    Constructor arguments follow ParquetSolver (src/ParquetSolver.jl:81-97) and
    NL_ParquetSolver (src/nonlocal/ParquetSolver.jl:77-92): boxes nK1, nK2, nK3, the fermionic
    G mesh nG, the bubble factor mΠν_factor, the k-meshes LG (G, Σ) and L (vertex, bubble)."""

    def __init__(self, *, nG, nK1, nK2, nK3, T, U, mPinu, nonlocal_=False, LG=1, L_=1,
                 ref_kind=0, ref_boxes=None, tab=(1, 1), fdpa_h0=True, precision="fp64"):
        # the Channel constructor's box assertions (src/channel.jl:33-34,40)
        if nK1 < nK2[0]:
            raise AssertionError("No. bosonic frequencies in K1 must be larger than no. bosonic frequencies in K2")
        if nK1 < nK2[1]:
            raise AssertionError("No. bosonic frequencies in K1 must be larger than no. fermionic frequencies in K2")
        if any(a < b for a, b in zip(nK2, nK3)):
            raise AssertionError("Number of frequencies in K2 must be larger than in K3")
        d = L.Dims()
        d.nonlocal_ = int(nonlocal_)
        d.nG, d.nK1 = nG, nK1
        d.nK2b, d.nK2f = nK2
        d.nK3b, d.nK3f = nK3
        d.mPinu = mPinu
        d.LG, d.L = (LG, L_) if nonlocal_ else (1, 1)
        d.T, d.U = T, U
        d.ref_kind = ref_kind
        rb = ref_boxes or (1, (1, 1), (1, 1))
        d.r_nK1 = rb[0]
        d.r_nK2b, d.r_nK2f = rb[1]
        d.r_nK3b, d.r_nK3f = rb[2]
        d.tab_nb, d.tab_nf = tab
        d.fdpa_h0 = int(fdpa_h0)
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        d.precision = L.FP64 if precision == "fp64" else L.FP32
        self.precision = precision
        self.dims = d
        self.T, self.U = T, U
        self.nonlocal_ = bool(nonlocal_)
        self.nG, self.nK1, self.nK2, self.nK3, self.mPinu = nG, nK1, tuple(nK2), tuple(nK3), mPinu
        self.LG, self.L = d.LG, d.L
        self.NP, self.NkG = d.L * d.L, d.LG * d.LG
        self.lib = L.load()
        h = ctypes.c_void_p()
        rc = self.lib.fdga_create(ctypes.byref(h), ctypes.byref(d))
        self.h = h
        if rc != 0:
            msg = self.lib.fdga_last_error(h).decode() if h.value else "fdga_create failed"
            if h.value:
                self.lib.fdga_destroy(h)
            self.h = None
            raise FdgaError(msg)

    def close(self):
        if getattr(self, "h", None) is not None:
            self.lib.fdga_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise FdgaError(self.lib.fdga_last_error(self.h).decode())

    # ------------------------------------------------------------------ arrays
    def size(self, which):
        return self.lib.fdga_size(self.h, which)

    def get(self, which):
        n = self.size(which)
        if n < 0:
            raise KeyError(which)
        out = np.empty(n, dtype=np.complex128)
        self._check(self.lib.fdga_get(self.h, which, _dp(out.view(np.float64))))
        return out

    def set(self, which, data):
        n = self.size(which)
        a = np.ascontiguousarray(np.asarray(data, dtype=np.complex128).reshape(-1, order="F"))
        if a.size != n:
            raise ValueError(f"array {which}: expected {n} elements, got {a.size}")
        self._check(self.lib.fdga_set(self.h, which, _dp(a.view(np.float64))))

    def vertex(self, vtx, ch, cls):
        return self.get(L.vertex_id(vtx, ch, cls))

    def set_vertex(self, vtx, ch, cls, data):
        self.set(L.vertex_id(vtx, ch, cls), data)

    @property
    def Σ(self):
        return self.get(L.SIGMA).reshape(2 * self.nG, self.NkG, order="F")

    def init(self):
        """Constructor tail: Π0 = bubbles(G0), G = Dyson(Σ0, Gbare), Π = bubbles(G)."""
        self._check(self.lib.fdga_init(self.h))

    # ------------------------------------------------------------------ timing
    def init_sym_grp(self):
        """init_sym_grp!(S): frequency symmetry groups of the local solver (src/ParquetSolver.jl:203-231);
        for the s-wave lattice solver with their momentum forms plus the square lattice's reflection and
        rotation (src/nonlocal/ParquetSolver.jl:198-300)."""
        self._check(self.lib.fdga_init_sym_grp(self.h))

    def reset_sym_grp(self):
        """reset_sym_grp!(S) (src/ParquetSolver.jl:233-243)."""
        self._check(self.lib.fdga_reset_sym_grp(self.h))

    def time_iterations(self, strategy, n):
        ms = ctypes.c_double(0.0)
        self._check(self.lib.fdga_time_iterations(self.h, _strategy(strategy), n, ctypes.byref(ms)))
        return ms.value

    def set_graph(self, on):
        """Replay each fixed-point iteration from a captured hipGraph (True) or launch eagerly."""
        self._check(self.lib.fdga_set_graph(self.h, int(on)))

    _TILE_NAMES = ("64x64", "32x32", "64x32", "32x64")

    def autotune(self, reps=3):
        """Time every tile shape of the two contraction launches and keep the fastest (the state is
        left as it was; results do not depend on the tile shape)."""
        self._check(self.lib.fdga_autotune(self.h, reps))

    def gemm_tiles(self):
        """Tile shapes of the BSE_L_K3! and BSE_K3! launches."""
        return tuple(self._TILE_NAMES[self.lib.fdga_gemm_tile(self.h, left)] for left in (1, 0))

    def set_gemm_tile(self, left, tile):
        self._check(self.lib.fdga_set_gemm_tile(self.h, int(left), self._TILE_NAMES.index(tile)))

    def gemm_grids(self):
        """Workgroup counts of the BSE_L_K3! and BSE_K3! launches (0: one per output tile)."""
        return tuple(self.lib.fdga_gemm_grid(self.h, left) for left in (1, 0))

    def set_gemm_grid(self, left, grid):
        """A persistent grid of `grid` workgroups (a multiple of 8) for that launch, or 0; results do
        not depend on it."""
        self._check(self.lib.fdga_set_gemm_grid(self.h, int(left), int(grid)))

    def gemm_flops(self, left):
        """Executed flops of one batched BSE_L_K3! (left) or BSE_K3! launch."""
        return self.lib.fdga_gemm_flops(self.h, int(left))

    def set_stage_timing(self, on):
        self._check(self.lib.fdga_set_stage_timing(self.h, int(on)))

    def stage_ms(self):
        """Stage times 0..8, the two BSE GEMM kernels (9: BSE_L_K3!, 10: BSE_K3!) and the bubble
        kernel (11), ms."""
        out = []
        for s in range(12):
            ms = ctypes.c_double(0.0)
            self._check(self.lib.fdga_stage_ms(self.h, s, ctypes.byref(ms)))
            out.append(ms.value)
        return out

    def sync(self):
        self._check(self.lib.fdga_sync(self.h))


class SolverGroup:
    """Several solvers on one device iterated together (a parameter sweep): each member's stages
    run on its own stream, and the BSE_L_K3!/BSE_K3! contractions of all members run as one
    grouped GEMM launch each.  Members keep their own boxes and lattices; they must share the
    contraction precision and be all local or all lattice solvers, without K3 symmetry groups.
    No counterpart in the reference (one solver per parameter point); each member ends where
    `iterate_solver` on the member alone would.  Close the group before its members."""

    TILES = {"64x64": L.TILE_64x64, "32x32": L.TILE_32x32, "64x32": L.TILE_64x32, "32x64": L.TILE_32x64}

    def __init__(self, members):
        self.members = list(members)
        self.lib = L.load()
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = ctypes.c_void_p()
        rc = self.lib.fdga_group_create(ctypes.byref(h), arr, len(self.members))
        self.h = h
        if rc != 0:
            msg = self.lib.fdga_group_last_error(h).decode() if h.value else "fdga_group_create failed"
            if h.value:
                self.lib.fdga_group_destroy(h)
            self.h = None
            raise FdgaError(msg)

    def close(self):
        if getattr(self, "h", None) is not None:
            self.lib.fdga_group_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise FdgaError(self.lib.fdga_group_last_error(self.h).decode())

    def iterate(self, strategy="fdPA", update_Σ=True):
        """iterate_solver! (src/solve.jl:4-116) of every member."""
        self._check(self.lib.fdga_group_iterate(self.h, _strategy(strategy), int(update_Σ)))

    def sync(self):
        self._check(self.lib.fdga_group_sync(self.h))

    def time_iterations(self, strategy, n):
        ms = ctypes.c_double(0.0)
        self._check(self.lib.fdga_group_time_iterations(self.h, _strategy(strategy), n, ctypes.byref(ms)))
        return ms.value

    def set_graph(self, on):
        self._check(self.lib.fdga_group_set_graph(self.h, int(on)))

    def set_timing(self, on):
        self._check(self.lib.fdga_group_set_timing(self.h, int(on)))

    def gemm_ms(self):
        """Accumulated times of the grouped BSE_L_K3! and BSE_K3! launches while timing is on."""
        out = []
        for w in (0, 1):
            ms = ctypes.c_double(0.0)
            self._check(self.lib.fdga_group_gemm_ms(self.h, w, ctypes.byref(ms)))
            out.append(ms.value)
        return out

    def tiles(self, which):
        """(tile shape name, grid, computed output tiles) of contraction `which` (0 BSE_L_K3!,
        1 BSE_K3!); grid 0 = one workgroup per work item."""
        t, g, n = ctypes.c_int32(0), ctypes.c_int64(0), ctypes.c_int64(0)
        self._check(self.lib.fdga_group_tiles(self.h, which, ctypes.byref(t), ctypes.byref(g), ctypes.byref(n)))
        name = {v: k for k, v in self.TILES.items()}[t.value]
        return name, g.value, n.value

    def set_tiles(self, which, tile, grid=0):
        self._check(self.lib.fdga_group_set_tiles(self.h, which, self.TILES[tile], grid))

    def autotune(self, reps=3):
        """Time every tile shape and grid of both contractions and keep the fastest."""
        self._check(self.lib.fdga_group_autotune(self.h, reps))


def _strategy(strategy):
    if isinstance(strategy, int):
        return strategy
    if strategy in _STRATEGIES_NOT_BUILT:
        raise NotImplementedError(f"strategy {strategy} (a reference strategy) is not on the device path; use scPA, "
                                  "fdPA, scPA_new or fdPA_new")
    if strategy not in _STRATEGIES:
        raise AssertionError("Calculation strategy unknown")
    return _STRATEGIES[strategy]


# ------------------------------------------------------------------------------ models
def siam_bare_Green(nG, T, *, e, Δ, D):
    """i*G0 of the SIAM on MatsubaraMesh(T, nG, Fermion) (src/models/siam.jl:9-29); D=None: D = ∞."""
    n = np.arange(-nG, nG)
    v = np.pi * T * (2 * n + 1)
    if D is None or D == np.inf:
        return 1.0 / (v + 1j * e + Δ * np.sign(v))
    return 1.0 / (v + 1j * e + 2 * Δ / np.pi * np.arctan(D / v))


def hubbard_band(k1, k2, *, t1, t2=0.0, t3=0.0):
    """src/models/hubbard.jl:31-44"""
    return (-2 * t1 * (np.cos(k1) + np.cos(k2)) - 4 * t2 * np.cos(k1) * np.cos(k2)
            - 2 * t3 * (np.cos(2 * k1) + np.cos(2 * k2)))


def hubbard_bare_Green(nG, T, LG, *, μ, t1, t2=0.0, t3=0.0):
    """i*G0 of the square-lattice Hubbard model, [ν, k] with k = 2π(j1, j2)/LG, j1 fastest
    (src/models/hubbard.jl:8-29)."""
    n = np.arange(-nG, nG)
    v = np.pi * T * (2 * n + 1)
    j = np.arange(LG)
    k1 = np.tile(2 * np.pi * j / LG, LG)
    k2 = np.repeat(2 * np.pi * j / LG, LG)
    ek = hubbard_band(k1, k2, t1=t1, t2=t2, t3=t3)
    return 1j / (1j * v[:, None] + μ - ek[None, :])


def compute_occupation(G, T, Nk=1):
    """src/dyson.jl:33-41"""
    return 0.5 + np.imag(np.sum(G)) * T / Nk


# ------------------------------------------------------------------------------ constructors
def ParquetSolver(nK1, nK2, nK3, Gbare, G0, Σ0, F0=None, *, T, U, mΠν_factor=6, fdpa_h0=True,
                  F0_boxes=None, F0_tables=None, precision="fp64"):
    """ParquetSolver(nK1, nK2, nK3, Gbare, G0, Σ0, F0) (src/ParquetSolver.jl:81-156).

    F0 = None: RefVertex(T, U) with null tables.  F0 = dict {(ch, cls): array} of a local
    reference Vertex around RefVertex(T, U) with boxes F0_boxes = (nK1, nK2, nK3).
    F0_tables = (tb, tf, {'Fp_p': .., 'Fp_x': .., 'Ft_p': .., 'Ft_x': ..}) optional RefVertex tables."""
    nG = len(Gbare) // 2
    return _make(False, nG, nK1, nK2, nK3, T, U, mΠν_factor, 1, 1, Gbare, G0, Σ0, F0, F0_boxes, F0_tables, fdpa_h0,
                 precision)


def NL_ParquetSolver(nK1, nK2, nK3, L, Gbare, G0, Σ0, F0=None, *, T, U, mΠν_factor=32, fdpa_h0=True,
                     F0_boxes=None, F0_tables=None, precision="fp64"):
    """NL_ParquetSolver(nK1, nK2, nK3, mK_Γ, Gbare, G0, Σ0, F0) (src/nonlocal/ParquetSolver.jl:77-152);
    Gbare, G0, Σ0 are [ν, k] on an LG x LG mesh, the vertex lives on an L x L mesh."""
    Gbare = np.asarray(Gbare)
    nG = Gbare.shape[0] // 2
    LG = int(round(np.sqrt(Gbare.shape[1])))
    return _make(True, nG, nK1, nK2, nK3, T, U, mΠν_factor, LG, L, Gbare, G0, Σ0, F0, F0_boxes, F0_tables, fdpa_h0,
                 precision)


def _make(nonlocal_, nG, nK1, nK2, nK3, T, U, mp, LG, L_, Gbare, G0, Σ0, F0, F0_boxes, F0_tables, fdpa_h0, precision):
    tab = (1, 1)
    if F0_tables is not None:
        tab = (F0_tables[0], F0_tables[1])
    S = Solver(nG=nG, nK1=nK1, nK2=nK2, nK3=nK3, T=T, U=U, mPinu=mp, nonlocal_=nonlocal_, LG=LG, L_=L_,
               ref_kind=0 if F0 is None else 1, ref_boxes=F0_boxes, tab=tab, fdpa_h0=fdpa_h0, precision=precision)
    S.set(L.GBARE, Gbare)
    S.set(L.G0, G0)
    S.set(L.SIGMA0, Σ0)
    if F0 is not None:
        for (ch, cls), arr in F0.items():
            S.set_vertex(3, ch, cls, arr)
    if F0_tables is not None:
        for name, which in (("Fp_p", L.TAB_FPP), ("Fp_x", L.TAB_FPX), ("Ft_p", L.TAB_FTP), ("Ft_x", L.TAB_FTX)):
            if name in F0_tables[2]:
                S.set(which, F0_tables[2][name])
    S.init()
    return S


def parquet_solver_siam_parquet_approximation(nG, nK1, nK2, nK3, *, e, Δ, D, T, U, mΠν_factor=6, fdpa_h0=True,
                                              precision="fp64"):
    """src/ParquetSolver.jl:170-200: SIAM with G0 = Σ0 = 0 and F0 = RefVertex(T, U)."""
    Gbare = siam_bare_Green(nG, T, e=e, Δ=Δ, D=D)
    z = np.zeros_like(Gbare)
    return ParquetSolver(nK1, nK2, nK3, Gbare, z, z, T=T, U=U, mΠν_factor=mΠν_factor, fdpa_h0=fdpa_h0,
                         precision=precision)


def parquet_solver_hubbard_parquet_approximation(nG, nK1, nK2, nK3, LG, L, *, T, U, μ, t1, t2=0.0, t3=0.0,
                                                 mΠν_factor=32, fdpa_h0=True, precision="fp64"):
    """src/nonlocal/ParquetSolver.jl:166-196: Hubbard with G0 = Σ0 = 0 and F0 = RefVertex(T, U)."""
    Gbare = hubbard_bare_Green(nG, T, LG, μ=μ, t1=t1, t2=t2, t3=t3)
    z = np.zeros_like(Gbare)
    return NL_ParquetSolver(nK1, nK2, nK3, L, Gbare, z, z, T=T, U=U, mΠν_factor=mΠν_factor, fdpa_h0=fdpa_h0,
                            precision=precision)


# ------------------------------------------------------------------------------ stages
def Dyson(S):
    S._check(S.lib.fdga_dyson(S.h))


def bubbles(S):
    S._check(S.lib.fdga_bubbles(S.h))


def build_K3_cache(S):
    S._check(S.lib.fdga_build_k3_cache(S.h))


def BSE_L_K2(S, Ch):
    S._check(S.lib.fdga_bse_l_k2(S.h, Ch))


def BSE_L_K3(S, Ch):
    S._check(S.lib.fdga_bse_l_k3(S.h, Ch))


def BSE_K1(S, Ch, is_mfRG=False):
    """BSE_K1!(S, Ch, is_mfRG) (src/BSE_templates.jl:12-35)."""
    S._check(S.lib.fdga_bse_k1(S.h, Ch, int(is_mfRG)))


def BSE_K2(S, Ch, is_mfRG=False):
    """BSE_K2!(S, Ch, is_mfRG) (src/BSE_templates.jl:77-110)."""
    S._check(S.lib.fdga_bse_k2(S.h, Ch, int(is_mfRG)))


def BSE_K3(S, Ch, is_mfRG=False):
    """BSE_K3!(S, Ch, is_mfRG) (src/BSE_templates.jl:151-180)."""
    S._check(S.lib.fdga_bse_k3(S.h, Ch, int(is_mfRG)))


def tangent_solver(S):
    """A context with S's dimensions and precision, U = 0 and no reference vertex: the carrier of the
    tangent and cotangent arrays of BSE_K3_jvp / BSE_K3_vjp (F, FL.K3, Πpp/Πph and Fbuff.K3)."""
    return Solver(nG=S.nG, nK1=S.nK1, nK2=S.nK2, nK3=S.nK3, T=S.T, U=0.0, mPinu=S.mPinu, nonlocal_=S.nonlocal_,
                  LG=S.LG, L_=S.L, precision=S.precision)


def BSE_K3_jvp(S, Tn):
    """Tangent of build_K3_cache!(S) followed by BSE_K3!(S, Ch) for Ch = p, a, t (src/build_K3_cache.jl:20-95,
    src/BSEa/BSEa_K3.jl:43-128, src/BSE_templates.jl:151-180) at S's current F, FL and Π: reads the tangent
    F, FL.K3, Πpp, Πph from Tn (a tangent_solver(S)) and writes the tangent Fbuff.K3 into Tn."""
    S._check(S.lib.fdga_bse_k3_jvp(S.h, Tn.h))


def BSE_K3_vjp(S, Tn):
    """Adjoint of the same map (pairing sum(u * v), no conjugation): reads the cotangent of Fbuff.K3 from Tn
    (overwritten) and writes the cotangents of F (K1, K2, K3), FL.K3, Πpp and Πph into Tn."""
    S._check(S.lib.fdga_bse_k3_vjp(S.h, Tn.h))


def k3_gather_map(nK1, nK2, nK3, lattice=False):
    """The gather of build_K3_cache! as a transposed, destination-sorted linear map (host only, no
    device): (off, src, coef) with the cache entries src[off[d]:off[d + 1]] summed into vertex element d
    of the momentum-averaged vertex (lattice) or of the vertex (local); see fdga_k3_gather_map."""
    d = L.Dims()
    d.nonlocal_ = int(lattice)
    d.nK1 = nK1
    d.nK2b, d.nK2f = nK2
    d.nK3b, d.nK3f = nK3
    d.L = d.LG = 2 if lattice else 1
    lib = L.load()
    n = lib.fdga_k3_gather_map(ctypes.byref(d), None, None, None)
    if n < 0:
        raise FdgaError("fdga_k3_gather_map failed")
    NV3, NW3 = 2 * nK3[1], 2 * nK3[0] - 1
    nd = 3 * ((2 * nK1 - 1) + (2 * nK2[0] - 1) * 2 * nK2[1] + NW3 * NV3 * NV3)
    off = np.empty(nd + 1, dtype=np.int64)
    src = np.empty(max(n, 1), dtype=np.int32)
    cf = np.empty(max(n, 1), dtype=np.float32)
    lib.fdga_k3_gather_map(ctypes.byref(d), off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                           src.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                           cf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return off, src[:n], cf[:n]


def BSE_K1_new(S, Ch, is_mfRG=False):
    """BSE_K1_new!(S, Ch, is_mfRG) (src/BSE_templates.jl:188-211): K1 = (U + K1 + K2') Π U."""
    S._check(S.lib.fdga_bse_k1_new(S.h, Ch, int(is_mfRG)))


def BSE_K2_new(S, Ch, is_mfRG=False):
    """BSE_K2_new!(S, Ch, is_mfRG) (src/BSE_templates.jl:217-252): K2 = (K2 + K3 + (I - U)) Π U."""
    S._check(S.lib.fdga_bse_k2_new(S.h, Ch, int(is_mfRG)))


def build_K3_cache_mfRG(S, is_first_iteration=True):
    """build_K3_cache_mfRG!(S, is_first_iteration) (src/nonlocal_2/build_K3_cache.jl:97-164); the
    reference defines it for NL_ParquetSolver only."""
    S._check(S.lib.fdga_build_k3_cache_mfrg(S.h, int(is_first_iteration)))


def update_F(S):
    S._check(S.lib.fdga_update_F(S.h))


def SDE(S, strategy="scPA", include_Hartree=True):
    S._check(S.lib.fdga_sde(S.h, _strategy(strategy), int(include_Hartree)))


def init_sym_grp(S):
    S.init_sym_grp()


def reset_sym_grp(S):
    S.reset_sym_grp()


def iterate_solver(S, strategy="fdPA", update_Σ=True, compute_Hartree=True):
    """iterate_solver!(S; strategy, update_Σ, compute_Hartree) (src/solve.jl:4-116)."""
    S._check(S.lib.fdga_iterate(S.h, _strategy(strategy), int(update_Σ), int(compute_Hartree)))


def flatten(S):
    n = S.lib.fdga_flat_size(S.h)
    x = np.empty(n, dtype=np.complex128)
    S._check(S.lib.fdga_flatten(S.h, _dp(x.view(np.float64))))
    return x


def flatten_vertex(S):
    """flatten(S.F): the vertex part [γp; γt; γa] of the solver vector (src/vertex.jl:169-177)."""
    return flatten(S)[: S.lib.fdga_flat_size_F(S.h)]


def _vector(S, x, n, name):
    """x as a C-contiguous complex128 array of exactly n elements (the C side reads and writes 2 n doubles)."""
    a = np.asarray(x)
    if a.dtype != np.complex128:
        a = a.astype(np.complex128)
    a = np.ascontiguousarray(a).reshape(-1)
    if a.size != n:
        raise ValueError(f"{name}: expected {n} elements, got {a.size}")
    return a


def _out_vector(R, n, name):
    if not isinstance(R, np.ndarray) or R.dtype != np.complex128 or not R.flags.c_contiguous or R.size != n:
        raise ValueError(f"{name}: expected a C-contiguous complex128 array of {n} elements")
    return R


def unflatten(S, x):
    x = _vector(S, x, S.lib.fdga_flat_size(S.h), "unflatten")
    S._check(S.lib.fdga_unflatten(S.h, _dp(x.view(np.float64))))


def susceptibilities(S):
    """Spin and charge susceptibilities from the connected K1 of the full vertex in its pSp storage.
    The density/magnetic relations K1_D = -U² χ_D and K1_M = -U² χ_M with K1a = -K1_M and
    K1t = (K1_D - K1_M) / 2 (src/utility/load_triqs.jl:73-75,106-115) give
    χ_sp = K1a / U² and χ_ch = -(2 K1t - K1a) / U².  Returns (χ_sp, χ_ch) over the K1 box's bosonic
    frequencies Ω = -(nK1 - 1) .. nK1 - 1, with a second axis over the lattice momenta (the vertex's
    P order) for the lattice solver."""
    NP = S.NP if S.nonlocal_ else 1
    return chi_from_k1(S.vertex(0, aCh, 1), S.vertex(0, tCh, 1), S.U, NP)


def chi_from_k1(k1a, k1t, U, NP=1):
    """(χ_sp, χ_ch) from the a- and t-channel K1 arrays of one vertex (Ω fastest, then P)."""
    k1a = np.asarray(k1a, dtype=np.complex128)
    k1t = np.asarray(k1t, dtype=np.complex128)
    if k1a.shape != k1t.shape:
        raise ValueError(f"K1a and K1t differ in shape: {k1a.shape} vs {k1t.shape}")
    U2 = float(U) ** 2
    sp, ch = k1a / U2, -(2.0 * k1t - k1a) / U2
    if NP > 1:
        sp, ch = sp.reshape(-1, NP, order="F"), ch.reshape(-1, NP, order="F")
    return sp, ch


def fixed_point(R, x, S, strategy="fdPA", update_Σ=True, compute_Hartree=True):
    """fixed_point!(R, x, S; update_Σ, compute_Hartree) (src/solve.jl:119-157): R = G(x) - x.  With
    update_Σ = False x and R hold the vertex only (flatten(S.F)), the form src/mfRG.jl:111-173 uses."""
    n = S.lib.fdga_flat_size(S.h) if update_Σ else S.lib.fdga_flat_size_F(S.h)
    x = _vector(S, x, n, "fixed_point: x")
    _out_vector(R, n, "fixed_point: R")
    S._check(S.lib.fdga_fixed_point(S.h, _dp(R.view(np.float64)), _dp(x.view(np.float64)), _strategy(strategy),
                                    int(update_Σ), int(compute_Hartree)))
    return R


class mfRGLinearMap:
    """mfRGLinearMap(S, strategy) (src/mfRG.jl:20-89): the linearised BSE around the reference vertex
    acting on vertex-only vectors, y = x - G_lin(x / 100) * 100.  Applying it overwrites S.F, S.Fbuff,
    S.FL and the caches, as in the reference; the first product also builds the F caches
    (is_first_iteration)."""

    def __init__(self, S, strategy="fdPA"):
        if strategy not in ("fdPA", "fdPA_new", "fdPA_1loop"):
            raise ValueError(f"Invalid strategy {strategy}. Must be fdPA or fdPA_new or fdPA_1loop.")
        self.S, self.strategy, self.is_first_iteration = S, strategy, True
        n = S.lib.fdga_flat_size_F(S.h)
        self.shape = (n, n)

    def __matmul__(self, x):
        return self.matvec(x)

    def matvec(self, x):
        S = self.S
        x = _vector(S, x, self.shape[0], "mfRGLinearMap: x")
        y = np.empty_like(x)
        S._check(S.lib.fdga_mfrg_apply(S.h, _dp(y.view(np.float64)), _dp(x.view(np.float64)), _strategy(self.strategy),
                                       int(self.is_first_iteration)))
        self.is_first_iteration = False
        return y


def solve(S, *, strategy="fdPA", maxiter=100, tol=1e-4, δ=0.85, mem=8, verbose=False, update_Σ=True,
          compute_Hartree=True):
    """solve!(S; maxiter, tol, δ, mem, update_Σ, compute_Hartree) (src/solve.jl:160-196): Anderson acceleration
    as NLsolve 4.5.1 method=:anderson (beta = δ, m = mem, ftol = tol, xtol = 0, aa_start = 1).
    The iterate is x_{k+1} = g_k - Σ_i γ_i Δg_i with g = x + δ f(x) and
    γ = argmin ||f_k - ΔF γ||; converged when max|f| < tol.  update_Σ = False iterates the vertex only
    (vertex-only vectors, see fixed_point).  Returns (iterations, residual)."""
    strategy = _strategy(strategy)
    x = flatten(S) if update_Σ else flatten_vertex(S)
    f = np.empty_like(x)
    dG, dF = [], []
    fold = gold = None
    res = np.inf
    it = 0
    for it in range(1, maxiter + 1):
        fixed_point(f, x, S, strategy, update_Σ, compute_Hartree)
        g = x + δ * f
        res = float(np.max(np.abs(f)))
        if verbose:
            print(f"iter {it:4d}  max|f| = {res:.3e}")
        if res < tol:
            break
        x = g
        if mem == 0:
            continue
        if it == 1:
            fold, gold = f.copy(), g.copy()
            continue
        dG.append(g - gold)
        dF.append(f - fold)
        gold, fold = g.copy(), f.copy()
        if len(dG) > min(mem, x.size):
            dG.pop(0)
            dF.pop(0)
        Fm = np.stack(dF, axis=1)
        Q, R = np.linalg.qr(Fm)
        gam = np.linalg.solve(R, Q.conj().T @ f)
        for i in range(len(dG)):
            x = x - gam[i] * dG[i]
    return it, res
