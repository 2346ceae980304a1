/*
 * fdga_oracle.h -- CPU FP64 restatement of the fdDGAsolver.jl parquet iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in fddgasolver.jl_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it.  The product never links it.
 *
 * Layouts follow the reference's MeshFunction storage: column-major (first
 * index fastest), ComplexF64 = two interleaved doubles.  Frequency meshes:
 *   fermionic MatsubaraMesh(T, N): n = -N .. N-1        (2N points)
 *   bosonic   MatsubaraMesh(T, N): m = -(N-1) .. N-1    (2N-1 points)
 * Momentum meshes are L x L with the first coordinate fastest.
 */
#ifndef FDGA_ORACLE_H
#define FDGA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int32_t nonlocal;   /* 0: ParquetSolver (src/ParquetSolver.jl), 1: NL_ParquetSolver (src/nonlocal/ParquetSolver.jl) */
    int32_t nG;         /* G/Sigma fermionic mesh N */
    int32_t nK1, nK2b, nK2f, nK3b, nK3f;
    int32_t mPinu;      /* mΠν_factor */
    int32_t LG, L;      /* k-mesh of G/Sigma and of the vertex/bubble (1, 1 for local) */
    double  T;
    double  U;          /* bare vertex of the innermost RefVertex */
    int32_t ref_kind;   /* 0: F0 = RefVertex, 1: F0 = local Vertex whose F0 is the RefVertex */
    int32_t r_nK1, r_nK2b, r_nK2f, r_nK3b, r_nK3f;  /* boxes of the local reference Vertex (ref_kind 1) */
    int32_t tab_nb, tab_nf;                         /* RefVertex table box (1,1 = null tables) */
    int32_t fdpa_h0;    /* 1: fdPA SDE subtracts (n0-1/2)U i after Σ0 (src/SDE.jl:26-31, current code);
                           0: omit that line (the convention the reference's test goldens were made with) */
} orc_params;

/* array identifiers for orc_set / orc_get */
enum {
    ORC_GBARE = 0, ORC_G0, ORC_SIGMA0, ORC_G, ORC_SIGMA,        /* [NG, NkG] */
    ORC_PI0PP, ORC_PI0PH, ORC_PIPP, ORC_PIPH,                   /* [NΩΠ, NνΠ, NP] */
    ORC_LPP, ORC_LPH,                                           /* [NΩ2, Nν2, NP] */
    ORC_TAB_FPP, ORC_TAB_FPX, ORC_TAB_FTP, ORC_TAB_FTX,         /* RefVertex tables [tb, tf, tf] */
    ORC_CACHE0 = 32,   /* + c, c in 0..9 : Γpx F0p F0a F0t Γpp Γa Γt Fp Fa Ft, [NΩ3,Nν3,Nν3,NP] */
    ORC_VERTEX0 = 64   /* + 16*vtx + 3*ch + (cls-1); vtx: 0 F, 1 Fbuff, 2 FL, 3 F0 (local Vertex) */
};
#define ORC_VERTEX(vtx, ch, cls) (ORC_VERTEX0 + 16 * (vtx) + 3 * (ch) + ((cls) - 1))

/* channel / spin / strategy codes */
enum { ORC_PCH = 0, ORC_TCH = 1, ORC_ACH = 2 };
enum { ORC_PSP = 0, ORC_XSP = 1, ORC_DSP = 2 };
enum { ORC_SCPA = 0, ORC_FDPA = 1, ORC_SCPA_NEW = 2, ORC_FDPA_NEW = 3 };
#define ORC_INF (-2147483647 - 1)   /* νInf */

typedef struct orc_solver orc_solver;

orc_solver* orc_create(const orc_params* p);
void        orc_destroy(orc_solver* s);
int64_t     orc_size(const orc_solver* s, int32_t which);          /* complex elements */
int         orc_set(orc_solver* s, int32_t which, const double* data);
int         orc_get(const orc_solver* s, int32_t which, double* data);
void        orc_set_U(orc_solver* s, double U_re, double U_im);

/* model inputs (src/models/siam.jl:9-29, src/models/hubbard.jl:8-44) */
void orc_siam_bare_green(orc_solver* s, double e, double Delta, double D);   /* D <= 0 : D = Inf */
void orc_hubbard_bare_green(orc_solver* s, double mu, double t1, double t2, double t3);
/* constructor tail: Π0 = bubbles(G0); G = Dyson(Σ0, Gbare); Π = bubbles(G)  (src/ParquetSolver.jl:97-114) */
void orc_init(orc_solver* s);
/* symmetry groups of init_sym_grp! (src/ParquetSolver.jl:203-231; NL src/nonlocal/ParquetSolver.jl:198-300) */
int  orc_init_sym_grp(orc_solver* s);
/* class table of one installed group, for tests: id 0 SGΣ, 1 SGpp[1] (= SGph[1]), 2 SGpp[2], 3 SGpp[3],
 * 4 SGppL[3], 5 SGph[2], 6 SGph[3], 7 SGphL[3], 8 SG0pp2, 9 SG0ph2.  orc_sym_size returns the number
 * of elements (-1: identity group or unknown id); orc_sym_classes writes per element the
 * representative's linear index and the operation (bit 0 sgn, bit 1 con) */
int64_t orc_sym_size(const orc_solver* s, int32_t id);
int  orc_sym_classes(const orc_solver* s, int32_t id, int64_t* rep, int32_t* op);

/* stages of iterate_solver! (src/solve.jl:4-116) */
void orc_dyson(orc_solver* s);
void orc_bubbles(orc_solver* s);
void orc_build_k3_cache(orc_solver* s);
void orc_bse_l_k2(orc_solver* s, int32_t ch);
void orc_bse_l_k3(orc_solver* s, int32_t ch);
/* is_mfrg: the is_mfRG = Val(true) forms */
void orc_bse_k1(orc_solver* s, int32_t ch, int32_t is_mfrg);
void orc_bse_k2(orc_solver* s, int32_t ch, int32_t is_mfrg);
void orc_bse_k3(orc_solver* s, int32_t ch, int32_t is_mfrg);
void orc_bse_k1_new(orc_solver* s, int32_t ch, int32_t is_mfrg);
void orc_bse_k2_new(orc_solver* s, int32_t ch, int32_t is_mfrg);
void orc_update_F(orc_solver* s);
void orc_sde(orc_solver* s, int32_t strategy, int32_t include_hartree);
void orc_iterate(orc_solver* s, int32_t strategy, int32_t update_sigma, int32_t compute_hartree);
/* build_K3_cache_mfRG!(S, is_first_iteration) and mfRGLinearMap * x (vertex-only vectors) */
void orc_build_k3_cache_mfrg(orc_solver* s, int32_t is_first);
void orc_mfrg_apply(orc_solver* s, double* y, const double* x, int32_t strategy, int32_t is_first);

/* fixed-point layer (src/solve.jl:119-196) */
int64_t orc_flat_size(const orc_solver* s);
int64_t orc_flat_size_F(const orc_solver* s);   /* vertex part only */
void    orc_flatten(const orc_solver* s, double* x);
void    orc_unflatten(orc_solver* s, const double* x);
/* update_sigma = 0: x and R are vertex-only vectors (orc_flat_size_F) */
void    orc_fixed_point(orc_solver* s, double* R, const double* x, int32_t strategy, int32_t update_sigma,
                        int32_t compute_hartree);
/* Anderson acceleration as NLsolve 4.5.1 method=:anderson; returns iterations, writes final max|f| */
int     orc_solve(orc_solver* s, int32_t strategy, int32_t maxiter, double tol, double beta,
                  int32_t mem, int32_t update_sigma, double* final_res);

/* point evaluator of a vertex: vtx 0 F, 1 Fbuff, 2 FL, 3 F0; flags bit0 F0, bit1 γp, bit2 γt, bit3 γa */
void orc_vertex_eval(const orc_solver* s, int32_t vtx, int32_t m, int32_t n, int32_t np, int32_t iP,
                     int32_t ch, int32_t sp, int32_t flags, double* out);
double orc_occupation(const orc_solver* s, int32_t which);
/* OpenMP threads of the parallel loops (the CPU baseline's "all cores" leg); default 1 thread is not
 * forced: the OpenMP runtime's default applies until this is called */
void    orc_set_threads(int32_t n);
int32_t orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
