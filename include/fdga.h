/*
 * fdga.h -- C-ABI of the MI355X (gfx950) Bethe-Salpeter hot path of fdDGAsolver.jl.
 *
 * The reference has no FFI; its seam is Julia multiple dispatch on the solver and the
 * `mode` symbol (src/solve.jl:173-175).  Each entry point below replaces one reference
 * function for a solver whose state lives on the device; a Julia method for `S.mode === :hip`
 * binds it with `ccall` (fddgasolver.jl_amd/julia/FdgaHIP.jl).  Host arrays use the reference's MeshFunction
 * storage (column-major, first index fastest, ComplexF64 = two interleaved doubles) and are
 * borrowed only for the duration of a call.  Every function returns 0 on success and a
 * negative code on error; fdga_last_error() returns the message.  A context is not
 * re-entrant; separate contexts may be used from separate threads.
 */
#ifndef FDGA_H
#define FDGA_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdga_ctx fdga_ctx;

/* Solver dimensions: the constructor arguments of ParquetSolver (src/ParquetSolver.jl:81-156)
 * and NL_ParquetSolver (src/nonlocal/ParquetSolver.jl:77-152). */
typedef struct fdga_dims {
    int32_t nonlocal;   /* 0 ParquetSolver (local), 1 NL_ParquetSolver (s-wave, momentum P) */
    int32_t nG;         /* fermionic mesh of G, Σ (MatsubaraMesh(T, nG, Fermion)) */
    int32_t nK1, nK2b, nK2f, nK3b, nK3f;
    int32_t mPinu;      /* mΠν_factor: bubble ν-mesh = nK1 * mPinu */
    int32_t LG, L;      /* k-mesh of G/Σ and of the vertex/bubble (L x L); 1, 1 when local */
    double  T;
    double  U;          /* bare interaction of the innermost RefVertex */
    int32_t ref_kind;   /* F0 = RefVertex (0) or local Vertex around the RefVertex (1) */
    int32_t r_nK1, r_nK2b, r_nK2f, r_nK3b, r_nK3f;  /* boxes of the local reference Vertex */
    int32_t tab_nb, tab_nf;                         /* RefVertex table box; (1, 1) = null tables */
    int32_t fdpa_h0;    /* 1: fdPA SDE subtracts (n0-1/2) U i after Σ0 (src/SDE.jl:26-31) */
    int32_t precision;  /* FDGA_FP64, or FDGA_FP32: BSE_L_K3!/BSE_K3! contraction on FP32 MFMA (operands
                           rounded to FP32, FP32 accumulation); everything else stays FP64 */
} fdga_dims;

enum { FDGA_FP64 = 0, FDGA_FP32 = 1 };
/* tile shapes of the contraction GEMMs (workgroups of 2x2, 1x1, 2x1, 1x2 waves of 32x32 each) */
enum { FDGA_TILE_64x64 = 0, FDGA_TILE_32x32 = 1, FDGA_TILE_64x32 = 2, FDGA_TILE_32x64 = 3 };
enum { FDGA_PCH = 0, FDGA_TCH = 1, FDGA_ACH = 2 };
/* strategies of iterate_solver! (src/solve.jl:10); :fdPA_1loop is not on the device path */
enum { FDGA_SCPA = 0, FDGA_FDPA = 1, FDGA_SCPA_NEW = 2, FDGA_FDPA_NEW = 3 };

/* array identifiers for fdga_set / fdga_get (reference layouts) */
enum {
    FDGA_GBARE = 0, FDGA_G0, FDGA_SIGMA0, FDGA_G, FDGA_SIGMA,   /* [NG, NkG]        S.Gbare S.G0 S.Σ0 S.G S.Σ */
    FDGA_PI0PP, FDGA_PI0PH, FDGA_PIPP, FDGA_PIPH,               /* [NΩ1, NνΠ, NP]   S.Π0pp S.Π0ph S.Πpp S.Πph */
    FDGA_LPP, FDGA_LPH,                                         /* [NΩ2, Nν2, NP]   S.Lpp S.Lph (get only) */
    FDGA_TAB_FPP, FDGA_TAB_FPX, FDGA_TAB_FTP, FDGA_TAB_FTX,     /* RefVertex tables [tb, tf, tf] */
    FDGA_CACHE0 = 32,   /* + c: S.cache_Γpx F0p F0a F0t Γpp Γa Γt Fp Fa Ft [NΩ3,Nν3,Nν3,NP] (get only) */
    FDGA_VERTEX0 = 64   /* + 16*v + 3*ch + (cls-1); v: 0 S.F, 1 S.Fbuff, 2 S.FL, 3 S.F0 (local Vertex) */
};
#define FDGA_VERTEX(v, ch, cls) (FDGA_VERTEX0 + 16 * (v) + 3 * (ch) + ((cls) - 1))

/* lifetime (ParquetSolver constructor, src/ParquetSolver.jl:81-156) */
int         fdga_create(fdga_ctx** ctx, const fdga_dims* dims);
int         fdga_destroy(fdga_ctx* ctx);
const char* fdga_last_error(const fdga_ctx* ctx);
int64_t     fdga_size(const fdga_ctx* ctx, int32_t which);        /* complex elements, -1 if unknown */
int         fdga_set(fdga_ctx* ctx, int32_t which, const double* host);
int         fdga_get(fdga_ctx* ctx, int32_t which, double* host);
/* Π0 = bubbles(G0); Σ = Σ0; G = Dyson(Σ, Gbare); Π = bubbles(G)   (src/ParquetSolver.jl:97-114) */
int         fdga_init(fdga_ctx* ctx);

/* symmetry groups: init_sym_grp!(S) (src/ParquetSolver.jl:203-231) installs the frequency symmetry
 * groups of src/symmetries.jl:1-94; for NL_ParquetSolver (src/nonlocal/ParquetSolver.jl:198-300)
 * their momentum forms plus the square lattice's reflection and rotation
 * (src/nonlocal/symmetries.jl:1-133).  Afterwards every stage's result holds op(value of the class
 * representative) at the other members, as SG(f, InitFunction) and SG(f) give.  The K3 contractions
 * compute only the tiles that hold representatives.  reset_sym_grp!(S) (src/ParquetSolver.jl:233-243)
 * restores the identity groups (SG0pp2 and SG0ph2 are left installed, as there).  Fails when a K3
 * array has 2^29 elements or more. */
int fdga_init_sym_grp(fdga_ctx* ctx);
int fdga_reset_sym_grp(fdga_ctx* ctx);

/* stages of iterate_solver! (src/solve.jl:4-116) */
int fdga_dyson(fdga_ctx* ctx);                          /* Dyson!(S)          src/dyson.jl:1-31 */
int fdga_bubbles(fdga_ctx* ctx);                        /* bubbles!(S)        src/bubble.jl:1-36, src/nonlocal/bubble.jl:87-158 */
int fdga_build_k3_cache(fdga_ctx* ctx);                 /* build_K3_cache!(S) src/build_K3_cache.jl:20-95 */
int fdga_bse_l_k2(fdga_ctx* ctx, int32_t ch);           /* BSE_L_K2!(S, Ch)   src/BSE_templates.jl:40-71 */
int fdga_bse_l_k3(fdga_ctx* ctx, int32_t ch);           /* BSE_L_K3!(S, Ch)   src/BSE_templates.jl:116-145 */
/* is_mfrg: the is_mfRG = Val(true) argument of the reference wrappers (src/BSEa/BSEa_K1.jl:29-35,
 * BSEa_K2.jl:70-78, BSEa_K3.jl:83-97; lattice src/nonlocal/BSEa/BSEa_K{1,2,3}.jl) */
int fdga_bse_k1(fdga_ctx* ctx, int32_t ch, int32_t is_mfrg);   /* BSE_K1!(S, Ch, is_mfRG)  src/BSE_templates.jl:12-35 */
int fdga_bse_k2(fdga_ctx* ctx, int32_t ch, int32_t is_mfrg);   /* BSE_K2!(S, Ch, is_mfRG)  src/BSE_templates.jl:77-110 */
int fdga_bse_k3(fdga_ctx* ctx, int32_t ch, int32_t is_mfrg);   /* BSE_K3!(S, Ch, is_mfRG)  src/BSE_templates.jl:151-180 */
/* the alternative forms K1 = (U + K1 + K2') Π U and K2 = (K2 + K3 + (I - U)) Π U (their lattice
 * is_mfRG form: (F - F0) Π U; the local forms have no is_mfRG branch) */
int fdga_bse_k1_new(fdga_ctx* ctx, int32_t ch, int32_t is_mfrg);   /* BSE_K1_new!(S, Ch, is_mfRG) src/BSE_templates.jl:188-211 */
int fdga_bse_k2_new(fdga_ctx* ctx, int32_t ch, int32_t is_mfrg);   /* BSE_K2_new!(S, Ch, is_mfRG) src/BSE_templates.jl:217-252 */
/* build_K3_cache_mfRG!(S, is_first_iteration) (src/nonlocal_2/build_K3_cache.jl:97-164; NL_ParquetSolver
 * only, as in the reference; not under lattice symmetry groups) */
int fdga_build_k3_cache_mfrg(fdga_ctx* ctx, int32_t is_first);
int fdga_update_F(fdga_ctx* ctx);                       /* set!(S.F, S.Fbuff) src/solve.jl:87 */
int fdga_sde(fdga_ctx* ctx, int32_t strategy, int32_t include_hartree);   /* SDE!(S) src/SDE.jl:3-33 */
/* iterate_solver!(S; strategy, update_Σ, compute_Hartree) src/solve.jl:4-116 */
int fdga_iterate(fdga_ctx* ctx, int32_t strategy, int32_t update_sigma, int32_t compute_hartree);

/* fixed-point layer: x = [γp(K1,K2,K3); γt; γa; Σ] (src/ParquetSolver.jl:277-306) */
int64_t fdga_flat_size(const fdga_ctx* ctx);
int64_t fdga_flat_size_F(const fdga_ctx* ctx);   /* the vertex part [γp; γt; γa] (src/vertex.jl:153-196) */
int     fdga_flatten(fdga_ctx* ctx, double* x);
int     fdga_unflatten(fdga_ctx* ctx, const double* x);
/* fixed_point!(R, x, S; update_Σ, compute_Hartree) (src/solve.jl:119-157): R = G(x) - x.  update_sigma = 0:
 * x and R are vertex-only vectors of fdga_flat_size_F elements (the form src/mfRG.jl:111-173 uses;
 * src/solve.jl:138 passes the whole solver vector to unflatten!(S.F, x), whose length assertion,
 * src/vertex.jl:195, it fails) */
int     fdga_fixed_point(fdga_ctx* ctx, double* R, const double* x, int32_t strategy, int32_t update_sigma,
                         int32_t compute_hartree);
/* fdga_fixed_point without any device allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The state vector is staged through a file-scope device array of FDGA_SCRATCH_BYTES
 * (64 MiB, held by each device the library runs on until the process ends); a state vector larger than
 * that (fdga_flat_size or fdga_flat_size_F elements of 16 bytes) is refused with an error.  The iteration
 * runs its launches eagerly (no hipGraph capture).  Every call on one device shares that array: calls
 * that overlap on one device, from any context, race. */
#define FDGA_SCRATCH_BYTES ((int64_t)64 << 20)
int     fdga_fixed_point_noalloc(fdga_ctx* ctx, double* R, const double* x, int32_t strategy,
                                 int32_t update_sigma, int32_t compute_hartree);
/* mfRGLinearMap(S, strategy) * x (src/mfRG.jl:20-89): y = x - G_lin(x / 100) * 100 with G_lin the
 * linearised BSE (build_K3_cache_mfRG!, for :fdPA BSE_L_K2! and the is_mfRG BSE_K1!/BSE_K2!, for :fdPA_new
 * the is_mfRG BSE_K1_new!/BSE_K2_new!, then BSE_L_K3! and the is_mfRG BSE_K3!); x, y vertex-only
 * (fdga_flat_size_F).  is_first: the map's is_first_iteration.  NL_ParquetSolver only. */
int     fdga_mfrg_apply(fdga_ctx* ctx, double* y, const double* x, int32_t strategy, int32_t is_first);

/* Tangent (JVP) and adjoint (VJP) of build_K3_cache! followed by BSE_K3! of the three channels p, a, t
 * (src/build_K3_cache.jl:20-95, src/BSEa/BSEa_K3.jl:43-128, src/BSE_templates.jl:151-180; NL
 * src/nonlocal/build_K3_cache.jl:18-94, src/nonlocal/BSEa/BSEa_K3.jl:1-128), as a map of S.F, S.FL.K3 and
 * Πpp/Πph to S.Fbuff.K3, linearised at ctx's current state (its caches and right operands are rebuilt).
 * No counterpart in the reference, which differentiates nothing: for Newton-Krylov, sensitivity and
 * gradient callers; the reference's own linearisation is mfRGLinearMap (src/mfRG.jl:20-89).
 * `tan` is a second context of the same dimensions with U = 0 and no reference vertex (ref_kind 0, zero
 * tables); both FP64, no K3 symmetry groups on ctx.
 *   JVP: reads tan's F (K1, K2, K3), FL.K3 and Πpp/Πph as the tangent; writes tan's Fbuff.K3.
 *   VJP: reads tan's Fbuff.K3 as the cotangent (overwritten); writes tan's F (K1, K2, K3), FL.K3 and
 *        Πpp/Πph (zero outside the K3 box).  Pairing <u, v> = sum of u v (no conjugation).
 * The VJP's gather transpose is a segmented sum over the map fdga_k3_gather_map describes. */
int     fdga_bse_k3_jvp(fdga_ctx* ctx, fdga_ctx* tan);
int     fdga_bse_k3_vjp(fdga_ctx* ctx, fdga_ctx* tan);
/* The gather of build_K3_cache! as a linear map of the momentum-averaged vertex (the vertex itself when
 * local), transposed and sorted by destination; host only, no device needed.  Destinations: per channel
 * p, t, a the arrays K1 [NΩ1], K2 [NΩ2, Nν2], K3 [Nν3, Nν3, NΩ3] in a row; sources: s n3 + o over the
 * cache slots s = Γpx, Γpp, Γa, Γt(d), Fp, Fa, Ft(d) (the F caches' momentum-independent part on the
 * lattice), entry o = ν + Nν3 (ν' + Nν3 iΩ).  Returns the number of entries (< 0 on error); fills
 * off [3 (NΩ1 + NΩ2 Nν2 + NΩ3 Nν3²) + 1], src and coef [entries] when they are not NULL. */
int64_t fdga_k3_gather_map(const fdga_dims* dims, int64_t* off, int32_t* src, float* coef);

/* synchronisation and timing on the context's stream */
int fdga_sync(fdga_ctx* ctx);
/* runs `n` iterate_solver! passes (update_Σ, compute_Hartree) and returns the elapsed milliseconds
 * measured with HIP events */
int fdga_time_iterations(fdga_ctx* ctx, int32_t strategy, int32_t n, double* ms);
/* per-stage device time accumulated by fdga_iterate while stage timing is on (ms, summed over
 * passes); stage: 0 dyson+bubbles, 1 cache, 2 L_K2, 3 L_K3, 4 K1, 5 K2, 6 K3, 7 update, 8 SDE;
 * 9 and 10: the BSE_L_K3! and BSE_K3! GEMM kernels alone (inside stages 3 and 6); 11: the bubble
 * kernel alone (inside stage 0) */
int fdga_stage_ms(const fdga_ctx* ctx, int32_t stage, double* ms);
/* enable (1) / disable (0) per-stage event timing inside fdga_iterate (adds a host sync per
 * stage); resets the accumulated stage times */
int fdga_set_stage_timing(fdga_ctx* ctx, int32_t on);
/* Replay fdga_iterate(ctx, strategy, 1, 1) from hipGraphs captured on first use (1), or launch the
 * kernels one by one (0, the default).  Results are identical; the graphs are recaptured after
 * fdga_init, fdga_init_sym_grp / fdga_reset_sym_grp and after fdga_set of G0, Π0, the RefVertex
 * tables or the reference Vertex.  No counterpart in the reference (execution option). */
int fdga_set_graph(fdga_ctx* ctx, int32_t on);
/* number of 64 x 64 output tiles (each over the full contraction length) that the batched (p, a, t)
 * BSE_L_K3! (left = 1) or BSE_K3! (left = 0) GEMM launch computes; fewer than the dense count
 * when symmetry groups are installed.  No counterpart in the reference (measurement). */
int64_t fdga_gemm_blocks(const fdga_ctx* ctx, int32_t left);
/* executed flops of that launch (8 BM BN Nν3 per computed BM x BN tile), its tile shape (FDGA_TILE_*),
 * and a manual choice of the shape.  No counterpart in the reference (measurement / execution option). */
double  fdga_gemm_flops(const fdga_ctx* ctx, int32_t left);
int     fdga_gemm_tile(const fdga_ctx* ctx, int32_t left);
int     fdga_set_gemm_tile(fdga_ctx* ctx, int32_t left, int32_t tile);
/* workgroup count of that launch: 0 = one workgroup per output tile (the default), otherwise a
 * persistent grid of `grid` workgroups (a multiple of 8, so that every workgroup stays on one XCD)
 * looping over the tiles (the FP32 lattice BSE_L_K3!: over its (channel, Ω, momentum group) tasks); a
 * smaller grid leaves compute units to the other chain of the iteration.
 * Results do not depend on it.  No counterpart in the reference (execution option). */
int     fdga_gemm_grid(const fdga_ctx* ctx, int32_t left);
int     fdga_set_gemm_grid(fdga_ctx* ctx, int32_t left, int32_t grid);
/* autotuner: times every tile shape of the batched (p, a, t) BSE_L_K3! and BSE_K3! launches on the
 * current operands (`reps` launches each) and keeps the fastest; the solver state is left as it was
 * (the launches write into scratch).  Results do not depend on the tile shape.  No counterpart in
 * the reference. */
int     fdga_autotune(fdga_ctx* ctx, int32_t reps);

/* Self-test of the batched complex GEMM engine used by BSE_L_K3!/BSE_K3!: C = A B for dense
 * column-major M x K and K x N ComplexF64 matrices on the FP64 (precision 0) or FP32 (1) MFMA path.  No counterpart in the reference; it exists
 * so that the library's tests can check the MFMA tiling directly. */
int fdga_selftest_zgemm(int32_t M, int32_t N, int32_t K, int32_t precision, const double* A, const double* B, double* C);
/* The same through tile shape `tile` (FDGA_TILE_*) over `grid` workgroups (0: one per output tile;
 * otherwise a multiple of 8, the workgroups looping over the tiles).  No counterpart in the reference. */
int fdga_selftest_zgemm_tiles(int32_t M, int32_t N, int32_t K, int32_t precision, int32_t tile, int64_t grid,
                              const double* A, const double* B, double* C);
/* The same for `nbatch` independent products, A, B and C holding the batches one after the other
 * (the BSE contractions are batches of small products).  No counterpart in the reference. */
int fdga_selftest_zgemm_batched(int32_t M, int32_t N, int32_t K, int32_t nbatch, int32_t precision, int32_t tile,
                                int64_t grid, const double* A, const double* B, double* C);

/* ---- groups of solver contexts: a parameter sweep on one device ----------------------------
 * The reference solves one parameter point per solver; a sweep calls solve! per point.  A group
 * iterates several contexts together: each member's stages run on the member's stream, and the
 * BSE_L_K3!/BSE_K3! contractions of all members run as one grouped GEMM launch each (the (Ω, P,
 * channel) blocks of every member as work items, each with the member's own box sizes).  Each
 * member's results equal those of fdga_iterate on the member alone.  Members must share the
 * precision, be all local or all lattice solvers and have no K3 symmetry groups; they must not be
 * iterated alone while in a group, and the group must be destroyed before its members.  No
 * counterpart in the reference. */
typedef struct fdga_group fdga_group;
int         fdga_group_create(fdga_group** g, fdga_ctx* const* members, int32_t n);
int         fdga_group_destroy(fdga_group* g);
const char* fdga_group_last_error(const fdga_group* g);
/* iterate_solver! (src/solve.jl:4-116) of every member */
int fdga_group_iterate(fdga_group* g, int32_t strategy, int32_t update_sigma);
int fdga_group_sync(fdga_group* g);
int fdga_group_time_iterations(fdga_group* g, int32_t strategy, int32_t n, double* ms);
/* replay fdga_group_iterate(strategy, 1) from captured hipGraphs (as fdga_set_graph) */
int fdga_group_set_graph(fdga_group* g, int32_t on);
/* event timing of the two grouped contractions (which: 0 BSE_L_K3!, 1 BSE_K3!), ms summed */
int fdga_group_set_timing(fdga_group* g, int32_t on);
int fdga_group_gemm_ms(const fdga_group* g, int32_t which, double* ms);
/* tile shape and grid of contraction `which`; `computed`: non-idle output tiles of one launch */
int fdga_group_tiles(fdga_group* g, int32_t which, int32_t* tile, int64_t* grid, int64_t* computed);
int fdga_group_set_tiles(fdga_group* g, int32_t which, int32_t tile, int64_t grid);
/* autotuner: times every tile shape and grid of both contractions (`reps` launches each) on the
 * members' current operands and keeps the fastest; the solver state is left as it was */
int fdga_group_autotune(fdga_group* g, int32_t reps);

#ifdef __cplusplus
}
#endif
#endif
