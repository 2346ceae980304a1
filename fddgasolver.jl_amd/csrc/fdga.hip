// fdga.hip -- C-ABI of the MI355X parquet hot path (include/fdga.h).
//
// A context owns every device array of one ParquetSolver / NL_ParquetSolver
// (src/ParquetSolver.jl:5-157, src/nonlocal/ParquetSolver.jl:1-153) and one HIP stream.
// Host arrays cross the boundary only in fdga_set / fdga_get / fdga_(un)flatten.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <future>
#include <mutex>
#include <vector>

#include "../../include/fdga.h"
#include "fdga_kernels.h"
#include "fdga_adjoint.h"
#include "fdga_forms.h"
#include "fdga_sym.h"

using namespace fdga;

namespace {

struct DevChan {
    double2 *K1 = nullptr, *K2 = nullptr, *K3 = nullptr;
    double2 *A1 = nullptr, *A2 = nullptr, *A3 = nullptr;   // averages (alias K* unless own_avg)
    bool own_avg = false;                                   // A* allocated separately
};

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// device copy of one symmetry group's classes (fdga_sym.h); ncls = 0: identity group
struct SymDev {
    int* off = nullptr;
    int* mem = nullptr;
    int ncls = 0;
    std::vector<int> tiles[TILE_N];   // K3 groups, per tile shape: GEMM blocks (iΩ * tiles + tile) holding a representative
    int* repop = nullptr;     // K3 groups: per element (representative << 2) | op
};
// groups of init_sym_grp! (src/ParquetSolver.jl:203-231): SGΣ; SGpp[1] = SGph[1] (same generator);
// SGpp[2] = SGppL[2]; SGph[2] = SGphL[2]; SGpp[3], SGph[3], SGppL[3], SGphL[3]; SG0pp2, SG0ph2
enum { SG_S, SG_K1, SG_PP2, SG_PH2, SG_PP3, SG_PH3, SG_PPL3, SG_PHL3, SG_0PP2, SG_0PH2, SG_N };

}  // namespace

// Scratch of fdga_fixed_point_noalloc: the staging of the host state vector for its layout changes, in a
// file-scope device array so that the entry point allocates nothing.  Each device the library runs on
// holds one copy until the process ends.  In host code `scratch` is only a symbol: its device address is
// taken once per device with hipGetSymbolAddress at context creation (g_scratch_ptr), and only that
// address is handed to copies and kernels.
constexpr size_t kScratchBytes = (size_t)FDGA_SCRATCH_BYTES;
static __device__ __align__(256) double2 scratch[kScratchBytes / sizeof(double2)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double2* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

struct fdga_ctx {
    fdga_dims d{};
    hipStream_t st = nullptr;
    int device = 0;
    double2* scratch = nullptr;   // this device's copy of the file-scope scratch array
    // Πpp (bit 0) / Πph (bit 1) stored only on the band by the last lattice bubble; pi_band_G: G is still
    // the one that bubble read (fdga_get can recompute the whole array)
    int pi_band = 0, pi_band_G = 0;
    // channels (bit ch) whose FP32 BSE_K3! right operand was written by the last BSE_L_K3! epilogue
    // from the current Π, caches and FL (cleared by anything that changes them)
    int r32_fresh = 0;
    // FL.K3 of the three channels holds FP32 values in the first half of its bytes (the FP32 lattice
    // iteration's form; fl_widen restores FP64 for other readers)
    int fl32 = 0;
    // F.K3 and Fbuff.K3 of the three channels hold FP32 values in the same way (the FP32 lattice
    // iteration's form, taken and left with FL.K3's; fk_widen restores FP64)
    int fk32 = 0;
    // the stream the stage kernels are launched on: st, or st2 for the second chain of an iteration
    hipStream_t run = nullptr;
    hipStream_t st2 = nullptr;
    hipEvent_t evfork = nullptr, evjoin = nullptr;
    std::string err;
    int NP = 1, NkG = 1, NG = 0;
    int NW1 = 0, NW2 = 0, NV2 = 0, NW3 = 0, NV3 = 0, NWP = 0, NVP = 0, nVP = 0;
    int wb = 0;   // half-width of the band of the bubble mesh that the gathers read (W, tail sums)
    size_t nGel = 0, nPiel = 0, n1 = 0, n2 = 0, n3 = 0, nL = 0;
    double2 *Gbare = nullptr, *G0 = nullptr, *Sig0 = nullptr, *G = nullptr, *Sig = nullptr;
    double2 *Pi0pp = nullptr, *Pi0ph = nullptr, *Pipp = nullptr, *Piph = nullptr;
    double2 *Lpp = nullptr, *Lph = nullptr;
    double2* Wc[3] = {nullptr, nullptr, nullptr};                  // BSE_K1!/BSE_K2! weights per channel
    double2* Wfar[3] = {nullptr, nullptr, nullptr};                // their far factors per row (WTail)
    double2* Ptail[4] = {nullptr, nullptr, nullptr, nullptr};      // tail sums of Π0pp, Π0ph, Πpp, Πph
    size_t ntail = 0;                                               // elements of one tail-sum array
    double2* pchunk[2] = {nullptr, nullptr};   // lattice bubble: chunk sums of the last Πpp / Πph [P, chunk, Ω]
    int w_fresh = 0;                                                // Wc valid for the current iteration
    // lattice BSE_K2! / BSE_L_K2! tables and GEMM result (s-wave solver only)
    double2 *k2x = nullptr, *k2xf = nullptr, *k2y = nullptr, *lk2x = nullptr;
    double2 *sde0 = nullptr;
    int sde0_valid = 0, sde0_h = -1;
    double2 *tab[4] = {nullptr, nullptr, nullptr, nullptr};
    int tabs_nonzero = 0;
    double2 U{};
    DevChan F[3], Fb[3], FL[3], R0[3];
    double2 *cache[12] = {};   // build_K3_cache! caches; [10], [11] the a and t Γ caches transposed
    // the F caches [7..9] hold build_K3_cache_mfRG!'s complete, momentum-independent F0 differences (the plain
    // lattice build stores only their momentum-independent part, the BSE_K3! loader adds K2r + K3r)
    int caches_mfrg = 0;
    double2 *R[3] = {nullptr, nullptr, nullptr};                    // BSE_K3! right operands [ω, νp, Ω, P]
    double2 *csum[3] = {nullptr, nullptr, nullptr};                 // their column sums [νp, Ω, P]
    double2 *tmpG1 = nullptr, *tmpG2 = nullptr, *tmpG3 = nullptr;   // [NG, NkG] scratch
    double2 *tmpL1 = nullptr, *tmpL2 = nullptr;                     // [NW2*NV2, NP] scratch
    double2 *u2C = nullptr;                                         // [4nG-1, NkG]
    double *scal = nullptr;                                         // device scalars (+ occupation partial sums)
    BubbleFold* folds = nullptr;                                    // lattice bubble: R -> R' folding table
    double2 *stage = nullptr;                                       // staging for layout changes
    size_t stage_n = 0;
    SymDev sg[SG_N];                                                // installed symmetry groups
    // GEMM block lists under the K3 groups: [0..3] one channel of SG_PP3, SG_PH3, SG_PPL3, SG_PHL3;
    // [4] BSE_L_K3! and [5] BSE_K3! over the channels (p, a, t) in one launch
    int* blist[6][TILE_N] = {};
    int blist_n[6][TILE_N] = {};      // workgroups of the list (computed tiles and idle padding)
    int blist_tiles[6][TILE_N] = {};  // computed tiles
    int ktile[2] = {TILE_64x64, TILE_64x64};   // tile shape of the BSE_L_K3! / BSE_K3! launches
    int kgrid[2] = {0, 0};   // their workgroup count (0: one per work item; else a persistent grid)
    int* finmap = nullptr;    // k_k3_finish chunk map (XCD-paired Ω blocks)
    int finmap_n = 0, fin_nchunk = 0;
    double2* k3raw = nullptr;                                       // raw t-channel K3 under groups [n3, NP]
    double2* k2raw = nullptr;                                       // lattice BSE_K2! raw values under groups [2 n2 NP]
    int* k3_pl = nullptr;     // lattice BSE_K3! groups: momenta whose batches hold representatives
    int k3_npl = 0;
    // hipGraph replay of fdga_iterate(strategy, 1): one graph per strategy and per F/Fbuff pointer
    // state (set!(S.F, S.Fbuff) swaps the two buffers every iteration)
    int graph_on = 0;
    int swap_state = 0;
    hipGraphExec_t gexec[4][2] = {};   // strategies :scPA, :fdPA, :scPA_new, :fdPA_new
    // timing
    int timing = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double stage_ms[12] = {};   // 0..8 stages of fdga_iterate, 9 / 10 the BSE_L_K3! / BSE_K3! GEMM kernels, 11 the bubble kernel
    // a tangent context's scratch for the adjoint of BSE_K3! (first fdga_bse_k3_vjp on it): the cache
    // cotangents [7 n3], the averaged vertex cotangent [3 (n1 + n2 + n3)], one channel's Ā [n3 NP] and
    // the gather's transposed map (destination offsets, sources, coefficients)
    double2 *adj_cbar = nullptr, *adj_avg = nullptr, *adj_Ab = nullptr;
    int64_t* adj_off = nullptr;
    int* adj_src = nullptr;
    float* adj_cf = nullptr;
    size_t adj_nd = 0;
};

// drops the captured hipGraphs of the iteration (defined with the graph replay below)
void graphs_invalidate(fdga_ctx* c);

namespace {

int fail(fdga_ctx* c, const std::string& msg, int code = -1) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(c, call)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) return fail((c), std::string(#call) + ": " + hipGetErrorString(e_), -2); \
    } while (0)

#define LAUNCHCHK(c)                                                                       \
    do {                                                                                   \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess) return fail((c), std::string("kernel launch: ") + hipGetErrorString(e_), -3); \
    } while (0)

SymCls symcls(const fdga_ctx* c, int id) { return SymCls{c->sg[id].off, c->sg[id].mem, c->sg[id].ncls}; }
// number of waves of a one-wave-per-output-class launch
size_t nclass(const SymCls& s, size_t nout) { return s.ncls ? (size_t)s.ncls : nout; }
// class members of the n-element array x from their representatives (per-element table rep); no-op
// without a table
int sym_fill(fdga_ctx* c, double2* x, const int* rep, size_t n) {
    if (!rep) return 0;
    k_sym_fill<<<cdiv(n, 256), 256, 0, c->run>>>(x, rep, n);
    LAUNCHCHK(c);
    return 0;
}
void free_sym(SymDev& g) {
    if (g.off) (void)hipFree(g.off);
    if (g.mem) (void)hipFree(g.mem);
    if (g.repop) (void)hipFree(g.repop);
    g = SymDev{};
}

void free_blists(fdga_ctx* c) {
    for (int t = 0; t < TILE_N; ++t)
        for (int k = 0; k < 6; ++k) {
            if (c->blist[k][t]) (void)hipFree(c->blist[k][t]);
            c->blist[k][t] = nullptr;
            c->blist_n[k][t] = 0;
            c->blist_tiles[k][t] = 0;
        }
    if (c->finmap) (void)hipFree(c->finmap);
    c->finmap = nullptr;
    c->finmap_n = 0;
}

// Orders work items for XCD locality: item i goes to XCD key(i) % 8, and workgroup 8 r + x runs
// the r-th item of XCD x (workgroups are dispatched round-robin over the 8 XCDs); -1 pads.
template <class Key>
std::vector<int> xcd_order(const std::vector<int>& items, Key key) {
    std::vector<std::vector<int>> q(8);
    for (int it : items) q[key(it) & 7].push_back(it);
    size_t rows = 0;
    for (auto& v : q) rows = std::max(rows, v.size());
    std::vector<int> out(rows * 8, -1);
    for (int x = 0; x < 8; ++x)
        for (size_t r = 0; r < q[x].size(); ++r) out[r * 8 + x] = q[x][r];
    return out;
}

int dalloc(fdga_ctx* c, double2** p, size_t n) {
    if (n == 0) n = 1;
    HIPCHK(c, hipMalloc((void**)p, n * sizeof(double2)));
    k_zero<<<cdiv(n, 256), 256, 0, c->run>>>(*p, n);
    LAUNCHCHK(c);
    return 0;
}

ChanV chan_view(const fdga_ctx* c, const DevChan& dc, int n1, int n2b, int n2f, int n3b, int n3f, int NP) {
    ChanV v;
    v.K1 = dc.K1; v.K2 = dc.K2; v.K3 = dc.K3;
    v.A1 = dc.A1; v.A2 = dc.A2; v.A3 = dc.A3;
    v.n1 = n1; v.n2b = n2b; v.n2f = n2f; v.n3b = n3b; v.n3f = n3f;
    v.NP = NP;
    v.k3f = 0;   // K3 in FP32: F and Fbuff under fk32, FL under fl32
    for (int r = 0; r < 3; ++r)
        if (((&dc == &c->F[r] || &dc == &c->Fb[r]) && c->fk32) || (&dc == &c->FL[r] && c->fl32)) v.k3f = 1;
    return v;
}
ChanV chanF(const fdga_ctx* c, const DevChan& dc) {
    return chan_view(c, dc, c->d.nK1, c->d.nK2b, c->d.nK2f, c->d.nK3b, c->d.nK3f, c->NP);
}

RefV ref_view(const fdga_ctx* c, bool zero) {
    RefV r;
    if (zero) {
        r.U = cmk(0.0, 0.0);
        r.tb = 0; r.tf = 0;
        for (auto& t : r.tab) t = nullptr;
        return r;
    }
    r.U = c->U;
    r.tb = c->tabs_nonzero ? c->d.tab_nb : 0;
    r.tf = c->d.tab_nf;
    for (int k = 0; k < 4; ++k) r.tab[k] = c->tab[k];
    return r;
}

F0V f0_view(const fdga_ctx* c) {
    F0V f;
    f.kind = c->d.ref_kind;
    f.ref = ref_view(c, false);
    for (int r = 0; r < 3; ++r)
        f.g[r] = (c->d.ref_kind == 1) ? chan_view(c, c->R0[r], c->d.r_nK1, c->d.r_nK2b, c->d.r_nK2f, c->d.r_nK3b, c->d.r_nK3f, 1)
                                      : ChanV{};
    return f;
}
F0V zero_f0() {
    F0V f{};
    f.kind = 0;
    f.ref.U = cmk(0.0, 0.0);
    f.ref.tb = 0;
    return f;
}

VtxV vtx_F(const fdga_ctx* c) {
    VtxV v;
    for (int r = 0; r < 3; ++r) v.g[r] = chanF(c, c->F[r]);
    v.f0 = f0_view(c);
    return v;
}
VtxV vtx_FL(const fdga_ctx* c) {
    VtxV v;
    for (int r = 0; r < 3; ++r) v.g[r] = chanF(c, c->FL[r]);
    v.f0 = zero_f0();
    return v;
}

int alloc_chan(fdga_ctx* c, DevChan& dc, size_t s1, size_t s2, size_t s3, int NP, bool avg) {
    if (dalloc(c, &dc.K1, s1 * NP) || dalloc(c, &dc.K2, s2 * NP) || dalloc(c, &dc.K3, s3 * NP)) return -2;
    if (avg && NP > 1) {
        if (dalloc(c, &dc.A1, s1) || dalloc(c, &dc.A2, s2) || dalloc(c, &dc.A3, s3)) return -2;
        dc.own_avg = true;
    } else {
        dc.A1 = dc.K1; dc.A2 = dc.K2; dc.A3 = dc.K3;
    }
    return 0;
}
void free_chan(DevChan& dc) {
    if (dc.own_avg) {
        (void)hipFree(dc.A1);
        (void)hipFree(dc.A2);
        (void)hipFree(dc.A3);
    }
    (void)hipFree(dc.K1); (void)hipFree(dc.K2); (void)hipFree(dc.K3);
    dc = DevChan{};
}

// P averages of the target vertex F (s-wave cross-channel lookups)
int refresh_avg(fdga_ctx* c) {
    if (c->NP == 1) return 0;
    // the K1, K2 and K3 averages of the three channels in one launch
    PavgJobs jobs{};
    unsigned blk = 0;
    int j = 0;
    for (int r = 0; r < 3; ++r) {
        const std::pair<double2*, const double2*> ak[3] = {{c->F[r].A1, c->F[r].K1}, {c->F[r].A2, c->F[r].K2},
                                                           {c->F[r].A3, c->F[r].K3}};
        const size_t nn[3] = {c->n1, c->n2, c->n3};
        for (int k = 0; k < 3; ++k, ++j) {
            jobs.A[j] = ak[k].first;
            jobs.K[j] = ak[k].second;
            jobs.Kf[j] = (k == 2 && c->fk32) ? reinterpret_cast<const float2*>(c->F[r].K3) : nullptr;
            jobs.n[j] = nn[k];
            jobs.blk0[j] = blk;
            blk += cdiv(nn[k], 16);
        }
    }
    jobs.blk0[j] = blk;
    jobs.njob = j;
    k_pavg_multi<<<blk, 256, 0, c->run>>>(jobs, c->NP);
    LAUNCHCHK(c);
    return 0;
}

int ensure_stage(fdga_ctx* c, size_t n) {
    if (c->stage_n >= n) return 0;
    if (c->stage) (void)hipFree(c->stage);
    c->stage = nullptr;
    c->stage_n = 0;
    HIPCHK(c, hipMalloc((void**)&c->stage, n * sizeof(double2)));
    c->stage_n = n;
    return 0;
}

// the 2-D transform of src's rows into data (src = nullptr: in place)
int dft2(fdga_ctx* c, double2* data, size_t nrows, int L, int sign, double scale, const double2* src = nullptr) {
    if (!src) src = data;
    if (L == 1) {
        if (src != data)
            HIPCHK(c, hipMemcpyAsync(data, src, nrows * sizeof(double2), hipMemcpyDeviceToDevice, c->run));
        if (scale != 1.0) {
            k_axpby<<<cdiv(nrows, 256), 256, 0, c->run>>>(data, data, scale, 0.0, nrows);
            LAUNCHCHK(c);
        }
        return 0;
    }
    const int LL = L * L;
    int RB = std::max(1, std::min(16, 1024 / LL));
    while (RB > 1 && cdiv(nrows, (size_t)RB) < 512) RB /= 2;   // at least two workgroups per compute unit
    size_t smem = (size_t)(2 * RB * LL + L) * sizeof(double2);
    k_dft2_rows<<<cdiv(nrows, RB), 256, smem, c->run>>>(data, src, nrows, L, RB, sign, scale);
    LAUNCHCHK(c);
    return 0;
}

// the R in [-L/2, L/2]^2 folding onto each R' = mod(R, L) (src/nonlocal/bubble.jl:118-134)
std::vector<BubbleFold> bubble_folds(int L, int LG) {
    std::vector<BubbleFold> f((size_t)L * L);
    const int h = L / 2;
    auto md = [](int a, int n) { return ((a % n) + n) % n; };
    for (int r2p = 0; r2p < L; ++r2p)
        for (int r1p = 0; r1p < L; ++r1p) {
            int t[4] = {0, 0, 0, 0}, n = 0;
            for (int R2 = -h; R2 <= h; ++R2) {
                if (md(R2, L) != r2p) continue;
                for (int R1 = -h; R1 <= h; ++R1) {
                    if (md(R1, L) != r1p) continue;
                    int halves = 0;   // weight 2^-halves
                    if (LG % 2 == 0) halves = (std::abs(R1) == LG / 2) + (std::abs(R2) == LG / 2);
                    const int kp = md(R1, LG) + LG * md(R2, LG), km = md(-R1, LG) + LG * md(-R2, LG);
                    t[n++] = kp | (km << 12) | ((R1 == 0 && R2 == 0) << 24) | ((1 + halves) << 25);
                }
            }
            f[(size_t)r1p + (size_t)L * r2p] = BubbleFold{t[0], t[1], t[2], t[3]};
        }
    return f;
}

// band tail sums of one array on the bubble mesh (rows of 2 nVP entries)
// ω entries per block of the lattice bubble (k_bubble_nl): 16 at L = 16 (16 x 256 complex in LDS);
// 6 at L = 24 (6 x 576 complex + 24 twiddles, 55.7 KB); otherwise as many L x L tiles as fit 64 KB
int bubble_nb(int L) {
    if (L == 16) return 16;
    if (L == 24) return 6;
    return std::max(1, std::min(16, (65536 / (int)sizeof(double2) - L) / (L * L)));
}

int tail_sums(fdga_ctx* c, const double2* X, double2* T) {
    const size_t nrows = (size_t)c->NWP * c->NP;
    k_tail_sums<<<(unsigned)nrows, 256, 0, c->run>>>(X, T, c->nVP, c->wb, nrows);
    LAUNCHCHK(c);
    return 0;
}
double2* pi_tail(fdga_ctx* c, const double2* Pi) {
    if (Pi == c->Pi0pp) return c->Ptail[0];
    if (Pi == c->Pi0ph) return c->Ptail[1];
    if (Pi == c->Pipp) return c->Ptail[2];
    return c->Ptail[3];
}

void kernel_begin(fdga_ctx* c);
void kernel_end(fdga_ctx* c, int s);

// bubbles for G (target, into Πpp/Πph) or G0 (reference, into Π0pp/Π0ph), with their tail sums.
// band: the lattice bubble stores Π only on the 16-ω chunks that hold the band ω ∈ [-wb, wb) (the
// gathers, contractions, weights and SDE sums read nothing else; the far regions enter through the
// chunk sums); fdga_get of Πpp/Πph then recomputes the whole array from the same G.
int bubbles_into(fdga_ctx* c, double2* pp, double2* ph, const double2* Gsrc, bool band = false) {
    const fdga_dims& d = c->d;
    c->r32_fresh = 0;
    if (!d.nonlocal) {
        dim3 g(cdiv(c->NVP, 256), c->NWP);
        kernel_begin(c);
        k_bubble_local<<<g, 256, 0, c->run>>>(pp, ph, Gsrc, d.nG, d.nK1, c->nVP, d.T);
        kernel_end(c, 11);
        LAUNCHCHK(c);
        const size_t nrows = (size_t)c->NWP * c->NP;
        k_tail_sums2<<<(unsigned)(2 * nrows), 256, 0, c->run>>>(pp, pi_tail(c, pp), ph, pi_tail(c, ph), c->nVP, c->wb, nrows);
        LAUNCHCHK(c);
        return 0;
    }
    // G(R) = fft(G)/LG^2 (src/nonlocal/bubble.jl:110)
    if (dft2(c, c->tmpG1, c->NG, d.LG, -1, 1.0 / ((double)d.LG * d.LG), Gsrc)) return -3;
    // the Π_R products and the backward DFT over R' (src/nonlocal/bubble.jl:110-155) in one pass
    const int L = d.L, NB = bubble_nb(L);
    const dim3 grid(cdiv(c->NVP, NB), c->NWP);
    double2 *cpp = c->pchunk[0], *cph = c->pchunk[1];
    // the chunks [wlo, whi) that k_tail_sums_chunked reads entry by entry: those that meet the band
    const int wlo = band ? (c->nVP - c->wb) / NB : 0;
    const int whi = band ? (c->nVP + c->wb + NB - 1) / NB : (int)grid.x;
    kernel_begin(c);
    if (L == 16 && d.precision == FDGA_FP32) {
        // the FP32 form: G(R) rounded to FP32 (into tmpG2), the products and the transform in FP32
        float2* GR32 = reinterpret_cast<float2*>(c->tmpG2);
        k_narrow32<<<cdiv(c->nGel, 256), 256, 0, c->run>>>(GR32, c->tmpG1, c->nGel);
        k_bubble_nl<16, 16, 512, float2><<<grid, 512, (size_t)NB * 256 * sizeof(float2), c->run>>>(
            pp, ph, GR32, c->folds, d.nG, L, d.nK1, c->nVP, d.T, NB, cpp, cph, wlo, whi);
    } else if (L == 16) {
        k_bubble_nl<16, 16, 512><<<grid, 512, (size_t)NB * 256 * sizeof(double2), c->run>>>(
            pp, ph, c->tmpG1, c->folds, d.nG, L, d.nK1, c->nVP, d.T, NB, cpp, cph, wlo, whi);
    } else if (L == 24) {
        k_bubble_nl<24, 6><<<grid, 256, (size_t)(NB * 576 + 24) * sizeof(double2), c->run>>>(
            pp, ph, c->tmpG1, c->folds, d.nG, L, d.nK1, c->nVP, d.T, NB, cpp, cph, wlo, whi);
    } else {
        k_bubble_nl<0, 0><<<grid, 256, (size_t)(NB * L * L + L) * sizeof(double2), c->run>>>(
            pp, ph, c->tmpG1, c->folds, d.nG, L, d.nK1, c->nVP, d.T, NB, cpp, cph, wlo, whi);
    }
    if (pp == c->Pipp) {
        c->pi_band = band ? 3 : 0;
        c->pi_band_G = 1;
    }
    kernel_end(c, 11);
    LAUNCHCHK(c);
    // tail sums from the bubble's chunk sums
    const unsigned nblk = (unsigned)c->NWP * (unsigned)cdiv(c->NP, TAIL_NPB);
    k_tail_sums_chunked<<<2 * nblk, 256, 0, c->run>>>(pp, cpp, pi_tail(c, pp), ph, cph, pi_tail(c, ph), nblk, c->nVP, c->wb,
                                                      NB, (int)grid.x, c->NWP, c->NP);
    LAUNCHCHK(c);
    return 0;
}

void stage_begin(fdga_ctx* c) {
    if (c->timing) (void)hipEventRecord(c->ev[0], c->st);
}
void stage_end(fdga_ctx* c, int s) {
    if (!c->timing) return;
    (void)hipEventRecord(c->ev[1], c->st);
    (void)hipEventSynchronize(c->ev[1]);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    c->stage_ms[s] += ms;
}
// event timing of one kernel launch inside a stage (ev[2], ev[3])
void kernel_begin(fdga_ctx* c) {
    if (c->timing) (void)hipEventRecord(c->ev[2], c->st);
}
void kernel_end(fdga_ctx* c, int s) {
    if (!c->timing) return;
    (void)hipEventRecord(c->ev[3], c->st);
    (void)hipEventSynchronize(c->ev[3]);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev[2], c->ev[3]);
    c->stage_ms[s] += ms;
}

const double kSign[3] = {+1.0, -1.0, +1.0};

double2* pi_of(fdga_ctx* c, int ch, bool zero) {
    if (zero) return ch == PCH ? c->Pi0pp : c->Pi0ph;
    return ch == PCH ? c->Pipp : c->Piph;
}

// bound of every box of F, F0 and the RefVertex tables (frequency window of the gather-GEMVs)
// Window bound of the SDE L-vertex sums: their channel evaluations keep the bosonic argument Ω
// fixed (src/SDE.jl:79-217), so K1 is a constant there and the summand depends on ω only through the
// K2 and K3 boxes (target and reference Vertex) and the RefVertex tables.
int sde_box_bound(const fdga_ctx* c) {
    const fdga_dims& d = c->d;
    int nk = std::max({d.nK2f, d.nK3f, d.tab_nf});
    if (d.ref_kind == 1) nk = std::max({nk, d.r_nK2f, d.r_nK3f});
    return nk + 1;
}
int box_bound(const fdga_ctx* c) {
    const fdga_dims& d = c->d;
    int nk = std::max({d.nK1, d.nK2b, d.nK2f, d.nK3b, d.nK3f, d.tab_nb, d.tab_nf});
    if (d.ref_kind == 1) nk = std::max({nk, d.r_nK1, d.r_nK2b, d.r_nK2f, d.r_nK3b, d.r_nK3f});
    return nk + 1;
}

// ---------------------------------------------------------------------------------------
// SDE levels (src/SDE.jl:35-48): F, F.F0, F.F0.F0 ... accumulated into S
struct Level { int kind; ChanV g[3]; int chanNP; double scale; };

int sde_sigma(fdga_ctx* c, double2* S, const double2* G, double scale) {
    const fdga_dims& d = c->d;
    if (!d.nonlocal) {
        const SymCls ss = symcls(c, SG_S);
        k_sde_sigma_local<<<cdiv(nclass(ss, c->NG) * WAVE, 256), 256, 0, c->run>>>(S, G, c->Lpp, c->Lph, d.nG, d.nK2b, d.nK2f,
                                                                                  d.T, scale, ss);
        LAUNCHCHK(c);
        return 0;
    }
    // real-space evaluation (src/nonlocal/SDE.jl:209-280)
    if (dft2(c, c->tmpG1, c->NG, d.LG, -1, 1.0 / ((double)d.LG * d.LG), G)) return -3;
    size_t nrow = (size_t)c->NW2 * c->NV2;
    if (dft2(c, c->tmpL1, nrow, d.L, -1, 1.0 / ((double)d.L * d.L), c->Lpp)) return -3;
    if (dft2(c, c->tmpL2, nrow, d.L, -1, 1.0 / ((double)d.L * d.L), c->Lph)) return -3;
    k_sde_sigma_nl_R<<<cdiv((size_t)c->NG * c->NkG * WAVE, 256), 256, 0, c->run>>>(c->tmpG2, c->tmpG1, c->tmpL1, c->tmpL2,
                                                                                 d.nG, d.nK2b, d.nK2f, d.LG, d.L, d.T);
    LAUNCHCHK(c);
    if (dft2(c, c->tmpG2, c->NG, d.LG, +1, 1.0)) return -3;
    if (sym_fill(c, c->tmpG2, c->sg[SG_S].repop, c->nGel)) return -3;   // SGΣ(Σ) (src/nonlocal/SDE.jl:280)
    k_axpby<<<cdiv(c->nGel, 256), 256, 0, c->run>>>(S, c->tmpG2, scale, 1.0, c->nGel);
    LAUNCHCHK(c);
    return 0;
}

// Σ += U² T² Σ G G G  (src/SDE.jl:328-364; NL src/nonlocal/SDE.jl:400-447)
int sde_u2(fdga_ctx* c, double2* S, const double2* G, bool c_ready = false) {
    const fdga_dims& d = c->d;
    const double2 fac = cscale(cmul(c->U, c->U), d.T * d.T);
    const int NB = 4 * d.nG - 1;
    if (!d.nonlocal) {
        if (!c_ready) k_u2_C<<<cdiv((size_t)NB * WAVE, 256), 256, 0, c->run>>>(c->u2C, G, G, d.nG, 1);
        const SymCls ss = symcls(c, SG_S);
        k_u2_S<<<cdiv(nclass(ss, c->NG) * WAVE, 256), 256, 0, c->run>>>(S, c->u2C, G, d.nG, 1, fac, 1, ss);
        LAUNCHCHK(c);
        return 0;
    }
    const int LG = d.LG;
    if (dft2(c, c->tmpG1, c->NG, LG, -1, 1.0 / ((double)LG * LG), G)) return -3;   // G(+R)
    if (dft2(c, c->tmpG2, c->NG, LG, +1, 1.0 / ((double)LG * LG), G)) return -3;   // G(-R)
    // four outputs per wave (NG = 2 nG is even; the last C wave may hold fewer than four m)
    k_u2_C4<<<cdiv((size_t)cdiv(NB, 4) * c->NkG * WAVE, 256), 256, 0, c->run>>>(c->u2C, c->tmpG2, c->tmpG1, d.nG, c->NkG);
    if (c->NG % 4 == 0)
        k_u2_S4<<<cdiv((size_t)(c->NG / 4) * c->NkG * WAVE, 256), 256, 0, c->run>>>(c->tmpG3, c->u2C, c->tmpG1, d.nG, c->NkG,
                                                                                 fac);
    else
        k_u2_S<<<cdiv((size_t)c->NG * c->NkG * WAVE, 256), 256, 0, c->run>>>(c->tmpG3, c->u2C, c->tmpG1, d.nG, c->NkG, fac,
                                                                             0, SymCls{nullptr, nullptr, 0});
    LAUNCHCHK(c);
    if (dft2(c, c->tmpG3, c->NG, LG, +1, 1.0)) return -3;
    if (sym_fill(c, c->tmpG3, c->sg[SG_S].repop, c->nGel)) return -3;   // SGΣ(Σ_U²) (src/nonlocal/SDE.jl:444)
    k_axpby<<<cdiv(c->nGel, 256), 256, 0, c->run>>>(S, c->tmpG3, 1.0, 1.0, c->nGel);
    LAUNCHCHK(c);
    return 0;
}

int sde_hartree(fdga_ctx* c, double2* S, const double2* G, double sign) {
    if (c->nGel <= 16384) {   // one block: the sum and the update without two more launches
        k_hartree_small<<<1, 256, 0, c->run>>>(S, G, c->nGel, c->scal, c->d.T, (double)c->NkG, c->U, sign);
        LAUNCHCHK(c);
        return 0;
    }
    k_occupation_part<<<OCC_BLOCKS, 256, 0, c->run>>>(c->scal + 8, G, c->nGel);
    k_occupation<<<1, WAVE, 0, c->run>>>(c->scal, c->scal + 8, c->d.T, (double)c->NkG);
    k_hartree<<<cdiv(c->nGel, 256), 256, 0, c->run>>>(S, c->nGel, c->scal, c->U, sign);
    LAUNCHCHK(c);
    return 0;
}

// Lpp/Lph of one level (lattice solver: plus a second Vertex level `lv2` of the same scale, summed
// in the same pass); `zero` selects the reference system's groups SG0pp2/SG0ph2
int sde_level_L(fdga_ctx* c, const Level& lv, const double2* Pp, const double2* Ph, bool zero,
                const Level* lv2 = nullptr) {
    // the bubbles' tail sums are kept current by bubbles_into and fdga_set
    const fdga_dims& d = c->d;
    const size_t nout = (size_t)c->NW2 * c->NV2 * c->NP;
    RefV ref = ref_view(c, false);
    const SymCls spp = symcls(c, zero ? SG_0PP2 : SG_PP2), sph = symcls(c, zero ? SG_0PH2 : SG_PH2);
    const double2 TU = cscale(c->U, d.T);
    const int NK = sde_box_bound(c);
    const double2* Tpp = pi_tail(c, Pp);
    const double2* Tph = pi_tail(c, Ph);
    if (d.nonlocal && lv.kind == 0) {   // lattice solver, Vertex level: one wave per (Ω, P)
        const unsigned nb = (unsigned)cdiv((size_t)c->NW2 * c->NP * WAVE, 256);
        const SdeLevelNL l0{lv.g[0], lv.g[1], lv.g[2], lv.chanNP};
        if (lv2) {
            const SdeLevelNL l1{lv2->g[0], lv2->g[1], lv2->g[2], lv2->chanNP};
            // a momentum-independent second level (the local reference vertex): its K3 terms as GEMMs
            // over the momenta after the pass
            const bool refgemm = lv2->chanNP == 1 && c->NP > 1;
            // (the target level's K3 form fixed at compile time: one load form in the loops)
            // (with it fixed, the target level's a/t K3 sum runs in k_sde_h3 after the pass)
            if (refgemm && l0.gp.k3f && l0.ga.k3f && l0.gt.k3f) {
                k_sde_L_nl<2, true, 1><<<nb, 256, 0, c->run>>>(c->Lpp, c->Lph, Pp, Ph, Tpp, Tph, l0, l1, d.nK2b, d.nK2f,
                                                               d.nK1, c->nVP, c->wb, c->NP, TU);
                k_sde_h3<true><<<nb, 256, 0, c->run>>>(c->Lph, Ph, l0.ga, l0.gt, d.nK2b, d.nK2f, d.nK1, c->nVP, c->NP, TU);
            } else if (refgemm && !l0.gp.k3f && !l0.ga.k3f && !l0.gt.k3f) {
                k_sde_L_nl<2, true, 0><<<nb, 256, 0, c->run>>>(c->Lpp, c->Lph, Pp, Ph, Tpp, Tph, l0, l1, d.nK2b, d.nK2f,
                                                               d.nK1, c->nVP, c->wb, c->NP, TU);
                k_sde_h3<false><<<nb, 256, 0, c->run>>>(c->Lph, Ph, l0.ga, l0.gt, d.nK2b, d.nK2f, d.nK1, c->nVP, c->NP, TU);
            } else if (refgemm) {   // (the reference level's K3 terms still in the GEMMs below)
                k_sde_L_nl<2, true><<<nb, 256, 0, c->run>>>(c->Lpp, c->Lph, Pp, Ph, Tpp, Tph, l0, l1, d.nK2b, d.nK2f,
                                                            d.nK1, c->nVP, c->wb, c->NP, TU);
            } else {
                k_sde_L_nl<2><<<nb, 256, 0, c->run>>>(c->Lpp, c->Lph, Pp, Ph, Tpp, Tph, l0, l1, d.nK2b, d.nK2f, d.nK1,
                                                      c->nVP, c->wb, c->NP, TU);
            }
            LAUNCHCHK(c);
            if (refgemm) {
                const ChanV& gp = lv2->g[PCH];
                const int NV = 2 * gp.n3f, NW = 2 * gp.n3b - 1;
                const SdeRefA<0> ap{gp.K3, nullptr, NV};
                const SdeRefA<1> ah{lv2->g[ACH].K3, lv2->g[TCH].K3, NV};
                const SdeRefB<0> bp{Pp, gp.n3b, gp.n3f, d.nK1, c->nVP, 0};
                const SdeRefB<1> bh{Ph, gp.n3b, gp.n3f, d.nK1, c->nVP, 0};
                const SdeRefE ep{c->Lpp, TU, d.nK2b, d.nK2f, gp.n3b, gp.n3f, 0};
                const SdeRefE eh{c->Lph, TU, d.nK2b, d.nK2f, gp.n3b, gp.n3f, 0};
                const int tm = cdiv(NV, GBM), tn = cdiv(c->NP, GBN);
                if (zgemm_launch(FDGA_FP64, NW, c->run, NV, c->NP, NV, tm, tn, nullptr, 0, ap, bp, ep) != hipSuccess ||
                    zgemm_launch(FDGA_FP64, NW, c->run, NV, c->NP, NV, tm, tn, nullptr, 0, ah, bh, eh) != hipSuccess)
                    return fail(c, "SDE reference-level GEMM launch failed", -3);
            }
        } else {
            k_sde_L_nl<1><<<nb, 256, 0, c->run>>>(c->Lpp, c->Lph, Pp, Ph, Tpp, Tph, l0, l0, d.nK2b, d.nK2f, d.nK1, c->nVP,
                                                  c->wb, c->NP, TU);
        }
        LAUNCHCHK(c);
        // SGpp2 / SGph2 (SG0pp2 / SG0ph2 for the reference system): members from representatives
        const int idp = zero ? SG_0PP2 : SG_PP2, idh = zero ? SG_0PH2 : SG_PH2;
        if (sym_fill(c, c->Lpp, spp.ncls ? c->sg[idp].repop : nullptr, nout)) return -3;
        return sym_fill(c, c->Lph, sph.ncls ? c->sg[idh].repop : nullptr, nout);
    }
    if (!spp.ncls && !sph.ncls) {
        k_sde_L<<<cdiv(nout * WAVE, 256), 256, 0, c->run>>>(c->Lpp, c->Lph, Pp, Ph, Tpp, Tph, lv.g[0], lv.g[1], lv.g[2], ref,
                                                           lv.kind, d.nK2b, d.nK2f, d.nK1, c->nVP, c->wb, c->NP,
                                                           lv.chanNP, TU, NK, spp);
    } else {
        const size_t nw = nclass(spp, nout) + nclass(sph, nout);
        k_sde_L_pair<<<cdiv(nw * WAVE, 256), 256, 0, c->run>>>(c->Lpp, c->Lph, Pp, Ph, Tpp, Tph, lv.g[0], lv.g[1], lv.g[2], ref,
                                                               lv.kind, d.nK2b, d.nK2f, d.nK1, c->nVP, c->wb, c->NP,
                                                               lv.chanNP, TU, NK, spp, sph);
    }
    LAUNCHCHK(c);
    return 0;
}

// SDE!(Σ, G, Πpp, Πph, Lpp, Lph, F; include_U², include_Hartree) over the levels starting at F
// (target: start = 0) or at F0 (reference system: start = 1)
int sde_levels(fdga_ctx* c, double2* S, const double2* G, const double2* Pp, const double2* Ph, int start, int incH) {
    std::vector<Level> lv;
    if (start == 0) {
        Level l{0, {}, c->NP, 1.0};
        for (int r = 0; r < 3; ++r) l.g[r] = chanF(c, c->F[r]);
        lv.push_back(l);
    }
    if (c->d.ref_kind == 1) {
        Level l{0, {}, 1, 1.0};
        F0V f = f0_view(c);
        for (int r = 0; r < 3; ++r) l.g[r] = f.g[r];
        lv.push_back(l);
    }
    Level lr{1, {}, 1, 1.0 / 3.0};
    lv.push_back(lr);
    // lattice solver: the target and the local reference vertex (two Vertex levels of scale 1) in one
    // L pass and one Σ sum (Σ is linear in the L-vertices)
    const bool merge01 = c->d.nonlocal && lv.size() >= 2 && lv[0].kind == 0 && lv[1].kind == 0 &&
                         lv[0].scale == lv[1].scale;
    for (size_t i = 0; i < lv.size(); ++i) {
        if (merge01 && i == 1) continue;
        const bool refl = lv[i].kind == 1;
        // a RefVertex with null tables contributes (U + 0) - U = 0 exactly
        bool u2c_ready = false;
        if (!(refl && !c->tabs_nonzero)) {
            if (sde_level_L(c, lv[i], Pp, Ph, start == 1, merge01 && i == 0 ? &lv[1] : nullptr)) return -3;
            if (i == 0 && !c->d.nonlocal) {   // this level's Σ sum and the U² term's C(m) in one launch
                const SymCls ss = symcls(c, SG_S);
                const size_t nsig = nclass(ss, c->NG), nu = (size_t)(4 * c->d.nG - 1);
                k_sigma_u2c<<<cdiv((nsig + nu) * WAVE, 256), 256, 0, c->run>>>(
                    S, G, c->Lpp, c->Lph, c->d.nG, c->d.nK2b, c->d.nK2f, c->d.T, lv[i].scale, ss, c->u2C, (int)nsig);
                LAUNCHCHK(c);
                u2c_ready = true;
            } else if (sde_sigma(c, S, G, lv[i].scale)) {
                return -3;
            }
        }
        if (i == 0) {
            if (sde_u2(c, S, G, u2c_ready)) return -3;
            if (incH && sde_hartree(c, S, G, +1.0)) return -3;
        }
    }
    return 0;
}

// K3-class geometry
K3Geom k3geom(const fdga_ctx* c) {
    K3Geom g;
    g.n3b = c->d.nK3b; g.n3f = c->d.nK3f; g.NP = c->NP;
    g.nWP = c->d.nK1; g.nVP = c->nVP;
    g.n2b = c->d.nK2b; g.n2f = c->d.nK2f;
    return g;
}

int idx_gamma(int ch) { return ch == PCH ? 4 : ch == ACH ? 5 : 6; }
int idx_f0(int ch) { return ch == PCH ? 1 : ch == ACH ? 2 : 3; }
int idx_fbase(int ch) { return ch == PCH ? 7 : ch == ACH ? 8 : 9; }

}  // namespace

// classes of the channel's K2 (SGpp[2] = SGppL[2] for p, SGph[2] = SGphL[2] for a and t)
static SymCls sym_k2(const fdga_ctx* c, int ch) { return symcls(c, ch == PCH ? SG_PP2 : SG_PH2); }

template <int CH>
static int launch_lk2(fdga_ctx* c) {
    if (c->d.nonlocal) {
        // table Γ F0 over the K2 mesh, then one GEMM per Ω against Π0 (rows of the bubble mesh)
        const int NV2 = c->NV2, NW2 = c->NW2;
        const size_t nx = (size_t)NV2 * NV2 * NW2;
        k_lk2nl_table<CH><<<cdiv(nx, 256), 256, 0, c->run>>>(c->lk2x, vtx_F(c), c->d.nK2b, c->d.nK2f);
        LAUNCHCHK(c);
        const K2nl_A a{c->lk2x, NV2, NV2};
        const K2nl_B b{pi_of(c, CH, true), c->NVP, c->nVP - c->d.nK2f, c->d.nK1 - c->d.nK2b, c->NWP};
        const LK2nl_E e{c->FL[CH].K2, CH == TCH ? c->FL[ACH].K2 : nullptr, NW2, NV2, c->d.T * kSign[CH], CH == TCH};
        if (zgemm_launch(0, NW2, c->run, NV2, c->NP, NV2, cdiv(NV2, GBM), cdiv(c->NP, GBN), nullptr, 0, a, b, e) != hipSuccess)
            return fail(c, "BSE_L_K2 GEMM launch failed", -3);
        // SGppL[2] / SGphL[2]: class members from their representative (the t channel's
        // post-processing commutes with the fill: its a channel is already symmetric)
        return sym_fill(c, c->FL[CH].K2, sym_k2(c, CH).ncls ? c->sg[CH == PCH ? SG_PP2 : SG_PH2].repop : nullptr,
                        c->n2 * c->NP);
    }
    const SymCls sym = sym_k2(c, CH);
    const size_t nw = nclass(sym, c->n2 * c->NP);
    k_gemv_lk2<CH><<<cdiv(nw * WAVE, 256), 256, 0, c->run>>>(c->FL[CH].K2, c->FL[ACH].K2, pi_of(c, CH, true), vtx_F(c),
                                                            c->d.nK2b, c->d.nK2f, c->d.nK1, c->nVP, c->NP,
                                                            c->d.T * kSign[CH], box_bound(c), sym);
    LAUNCHCHK(c);
    return 0;
}

// BSE_L_K2! of the p and a channels: one launch for the local solver
static int launch_lk2_pa(fdga_ctx* c) {
    if (c->d.nonlocal) return launch_lk2<PCH>(c) || launch_lk2<ACH>(c);
    const SymCls sp = sym_k2(c, PCH), sa = sym_k2(c, ACH);
    const size_t np = nclass(sp, c->n2 * c->NP), na = nclass(sa, c->n2 * c->NP);
    k_gemv_lk2_pa<<<cdiv((np + na) * WAVE, 256), 256, 0, c->run>>>(
        c->FL[PCH].K2, pi_of(c, PCH, true), sp, c->d.T * kSign[PCH], c->FL[ACH].K2, pi_of(c, ACH, true), sa,
        c->d.T * kSign[ACH], vtx_F(c), c->d.nK2b, c->d.nK2f, c->d.nK1, c->nVP, c->NP, box_bound(c), (int)np);
    LAUNCHCHK(c);
    return 0;
}

// the arrays of channel ch's weight: W on the band and its far sums from Π and Π0 and their tails
static WArgs w_args(fdga_ctx* c, int ch) {
    double2* P = pi_of(c, ch, false);
    double2* P0 = pi_of(c, ch, true);
    return WArgs{c->Wc[ch], c->Wfar[ch], P, P0};
}
// the tail sums of channel ch's weight, formed from those of Π and Π0 and the far factors
static WTail wtail(fdga_ctx* c, int ch) {
    return WTail{pi_tail(c, pi_of(c, ch, false)), pi_tail(c, pi_of(c, ch, true)), c->Wfar[ch]};
}
// K1/K2 weights of channel CH on the band into Wc[CH], their far factors into Wfar[CH]
template <int CH>
static int launch_w(fdga_ctx* c) {
    const size_t nrows = (size_t)c->NWP * c->NP;
    k_bse_w<CH><<<cdiv(nrows * (2 * c->wb), 256), 256, 0, c->run>>>(w_args(c, CH), vtx_FL(c), f0_view(c), c->d.nK1,
                                                                       c->nVP, c->wb, box_bound(c), nrows);
    LAUNCHCHK(c);
    return 0;
}
// the weights of the three channels in one launch
static int launch_w_all(fdga_ctx* c) {
    const size_t nrows = (size_t)c->NWP * c->NP;
    // lattice solver: one thread per chunk of W_PCH momenta (the P-independent reference factor once)
    const int pchunk = c->NP >= 2 * W_PCH;
    const size_t nth = pchunk ? (size_t)c->NWP * cdiv(c->NP, W_PCH) * (2 * c->wb) : nrows * (2 * c->wb);
    const unsigned nb = cdiv(nth, 256);
    k_bse_w3<<<3 * nb, 256, 0, c->run>>>(w_args(c, PCH), w_args(c, TCH), w_args(c, ACH), vtx_FL(c), f0_view(c), c->d.nK1,
                                         c->nVP, c->wb, box_bound(c), nrows, nb, pchunk, c->NP);
    LAUNCHCHK(c);
    return 0;
}

template <int CH>
static int launch_k1(fdga_ctx* c) {
    if (!c->w_fresh && launch_w<CH>(c)) return -3;
    const SymCls sym = symcls(c, SG_K1);
    const size_t nw = nclass(sym, c->n1 * c->NP);
    k_gemv_k1<CH><<<(unsigned)nw, 256, 0, c->run>>>(c->Fb[CH].K1, c->Fb[ACH].K1, c->Wc[CH], wtail(c, CH), vtx_F(c), c->d.nK1,
                                                           c->d.nK1, c->nVP, c->wb, c->NP, c->d.T * kSign[CH], box_bound(c), sym);
    LAUNCHCHK(c);
    return 0;
}

// BSE_K1! of the p and a channels in one launch (t separately: its post-processing reads the
// finished a channel)
static int launch_k1_pa(fdga_ctx* c) {
    if (!c->w_fresh && (launch_w<PCH>(c) || launch_w<ACH>(c))) return -3;
    const SymCls sym = symcls(c, SG_K1);
    const int nw = (int)nclass(sym, c->n1 * c->NP);
    k_gemv_k1_pa<<<(unsigned)(2 * nw), 256, 0, c->run>>>(c->Fb[PCH].K1, c->Wc[PCH], wtail(c, PCH), c->d.T * kSign[PCH],
                                                          c->Fb[ACH].K1, c->Wc[ACH], wtail(c, ACH), c->d.T * kSign[ACH],
                                                          vtx_F(c), c->d.nK1, c->d.nK1, c->nVP, c->wb, c->NP,
                                                          box_bound(c), sym, nw);
    LAUNCHCHK(c);
    return 0;
}

// lattice BSE_K2!: X table, GEMM X W per Ω over the band, finish per (Ω, P)
template <int CH>
static int launch_k2_nl(fdga_ctx* c) {
    if (!c->w_fresh && launch_w<CH>(c)) return -3;
    const int NV2 = c->NV2, NW2 = c->NW2, K = 2 * c->wb;
    const size_t nx = (size_t)NV2 * K * NW2 + (size_t)NV2 * 2 * NW2;
    k_k2nl_table<CH><<<cdiv(nx, 256), 256, 0, c->run>>>(c->k2x, c->k2xf, vtx_F(c), c->d.nK2b, c->d.nK2f, c->wb);
    LAUNCHCHK(c);
    const K2nl_A a{c->k2x, NV2, K};
    const K2nl_B b{c->Wc[CH], K, 0, c->d.nK1 - c->d.nK2b, c->NWP};
    const K2nl_E e{c->k2y, NV2, NW2};
    if (zgemm_launch(0, NW2, c->run, NV2, c->NP, K, cdiv(NV2, GBM), cdiv(c->NP, GBN), nullptr, 0, a, b, e) != hipSuccess)
        return fail(c, "BSE_K2 GEMM launch failed", -3);
    // under SGpp[2] / SGph[2] the diagram values go to k2raw first: members take op(representative)
    // and then their own FL.K2 term (src/BSEa/BSEa_K2.jl:60-100)
    const int* rp = sym_k2(c, CH).ncls ? c->sg[CH == PCH ? SG_PP2 : SG_PH2].repop : nullptr;
    const size_t n2 = c->n2 * c->NP;
    double2* raw = rp ? c->k2raw + (CH == PCH ? 0 : n2) : nullptr;
    k_k2nl_finish<CH><<<cdiv((size_t)NW2 * c->NP * WAVE, 256), 256, 0, c->run>>>(
        c->Fb[CH].K2, c->Fb[ACH].K2, c->k2y, c->k2xf, c->Wc[CH], wtail(c, CH), vtx_F(c), chanF(c, c->FL[CH]),
        chanF(c, c->FL[ACH]), c->d.nK2b, c->d.nK2f, c->d.nK1, c->wb, c->NP, c->d.T * kSign[CH], raw);
    LAUNCHCHK(c);
    {   // the own-K3 product term (T s Σ_ω K3 W; halved for a finished t, whose post-processing halves)
        const VtxV F = vtx_F(c);
        double2* out = raw ? raw : c->Fb[CH].K2;
        const double f = c->d.T * kSign[CH] * ((CH == TCH && !raw) ? 0.5 : 1.0);
        const unsigned nb = (unsigned)cdiv((size_t)NW2 * c->NP * WAVE, 256);
        if (F.g[CH].k3f)
            k_k2nl_k3<CH, true><<<nb, 256, 0, c->run>>>(out, c->Wc[CH], F, c->d.nK2b, c->d.nK2f, c->d.nK1, c->wb, c->NP, f);
        else
            k_k2nl_k3<CH, false><<<nb, 256, 0, c->run>>>(out, c->Wc[CH], F, c->d.nK2b, c->d.nK2f, c->d.nK1, c->wb, c->NP, f);
        LAUNCHCHK(c);
    }
    if (rp) {
        k_k2nl_symfinish<CH><<<cdiv(n2, 256), 256, 0, c->run>>>(c->Fb[CH].K2, c->Fb[ACH].K2, raw, rp, c->FL[CH].K2,
                                                              c->FL[ACH].K2, n2);
        LAUNCHCHK(c);
    }
    return 0;
}

template <int CH>
static int launch_k2(fdga_ctx* c) {
    if (c->d.nonlocal) return launch_k2_nl<CH>(c);
    if (!c->w_fresh && launch_w<CH>(c)) return -3;
    const SymCls sym = sym_k2(c, CH);
    const size_t nw = nclass(sym, c->n2 * c->NP);
    k_gemv_k2<CH><<<cdiv(nw * WAVE, 256), 256, 0, c->run>>>(c->Fb[CH].K2, c->Fb[ACH].K2, c->Wc[CH], wtail(c, CH), vtx_F(c),
                                                           chanF(c, c->FL[CH]),
                                                           chanF(c, c->FL[ACH]), c->d.nK2b, c->d.nK2f, c->d.nK1, c->nVP,
                                                           c->wb, c->NP, c->d.T * kSign[CH], box_bound(c), sym);
    LAUNCHCHK(c);
    return 0;
}

// BSE_K2! of the a and t channels in one kernel (t needs the finished a channel)
static int launch_k2_at(fdga_ctx* c) {
    if (c->d.nonlocal) return launch_k2_nl<ACH>(c) || launch_k2_nl<TCH>(c);
    if (!c->w_fresh && (launch_w<ACH>(c) || launch_w<TCH>(c))) return -3;
    const SymCls sym = symcls(c, SG_PH2);   // a and t share SGph[2]
    const size_t nw = nclass(sym, c->n2 * c->NP);
    k_gemv_k2_at<<<cdiv(nw * WAVE, 256), 256, 0, c->run>>>(c->Fb[ACH].K2, c->Fb[TCH].K2, c->Wc[ACH], c->Wc[TCH],
                                                          wtail(c, ACH), wtail(c, TCH), vtx_F(c),
                                                          chanF(c, c->FL[ACH]), chanF(c, c->FL[TCH]), c->d.nK2b, c->d.nK2f,
                                                          c->d.nK1, c->nVP, c->wb, c->NP, c->d.T, box_bound(c), sym);
    LAUNCHCHK(c);
    return 0;
}

// BSE_K2! of the three channels: one launch for the local solver
static int launch_k2_all(fdga_ctx* c) {
    if (c->d.nonlocal) return launch_k2<PCH>(c) || launch_k2_at(c);
    if (!c->w_fresh && (launch_w<PCH>(c) || launch_w<ACH>(c) || launch_w<TCH>(c))) return -3;
    const SymCls sp = sym_k2(c, PCH), sat = symcls(c, SG_PH2);
    const size_t np = nclass(sp, c->n2 * c->NP), nat = nclass(sat, c->n2 * c->NP);
    k_gemv_k2_all<<<cdiv((np + nat) * WAVE, 256), 256, 0, c->run>>>(
        c->Fb[PCH].K2, c->Wc[PCH], wtail(c, PCH), chanF(c, c->FL[PCH]), sp, (int)np, c->Fb[ACH].K2, c->Fb[TCH].K2,
        c->Wc[ACH], c->Wc[TCH], wtail(c, ACH), wtail(c, TCH), vtx_F(c), chanF(c, c->FL[ACH]), chanF(c, c->FL[TCH]),
        c->d.nK2b, c->d.nK2f, c->d.nK1, c->nVP, c->wb, c->NP, c->d.T, box_bound(c), sat);
    LAUNCHCHK(c);
    return 0;
}

// classes of the channel's K3 (BSE_K3!: SGpp[3] / SGph[3]; BSE_L_K3!: SGppL[3] / SGphL[3]),
// stored with device (K3 [ν, νp, Ω]) indices
static SymCls sym_k3(const fdga_ctx* c, int ch, bool left) {
    return symcls(c, ch == PCH ? (left ? SG_PPL3 : SG_PP3) : (left ? SG_PHL3 : SG_PH3));
}

// GEMM block list of a K3 launch over `n` channels with tile shape `tile`: the precomputed list when
// every slot has a group and the channels are one channel or (p, a, t); null (all blocks) otherwise
static const int* k3_blocks(const fdga_ctx* c, const int* chans, int n, bool left, int tile, size_t& nblocks,
                            size_t* ntiles = nullptr) {
    for (int s = 0; s < n; ++s)
        if (!sym_k3(c, chans[s], left).ncls) return nullptr;
    int k = -1;
    if (n == 1) k = (chans[0] == PCH ? 0 : 1) + (left ? 2 : 0);
    else if (n == 3 && chans[0] == PCH && chans[1] == ACH && chans[2] == TCH) k = left ? 4 : 5;
    if (k < 0 || !c->blist[k][tile]) return nullptr;
    nblocks = (size_t)c->blist_n[k][tile];
    if (ntiles) *ntiles = (size_t)c->blist_tiles[k][tile];
    return c->blist[k][tile];
}

// Finish of a K3 launch: with symmetry groups the class fill and the t post-processing in one
// pass (k_k3_finish; the t slot's GEMM output went to k3raw); without, the t post-processing
// unless the epilogue did it.
static int k3_finish(fdga_ctx* c, DevChan* out, const int* chans, int n, bool left, bool sym, bool t_in_epilogue) {
    if (!sym) {
        if (!t_in_epilogue)
            for (int s = 0; s < n; ++s)
                if (chans[s] == TCH) k_post_t<<<cdiv(c->n3 * c->NP, 256), 256, 0, c->run>>>(out[TCH].K3, out[ACH].K3, c->n3 * c->NP);
        LAUNCHCHK(c);
        return 0;
    }
    K3Fin f{};
    double2* o[3] = {nullptr, nullptr, nullptr};
    const double2* r[3] = {nullptr, nullptr, nullptr};
    const int* rp[3] = {nullptr, nullptr, nullptr};
    f.n = n;
    f.tslot = f.aslot = -1;
    for (int s = 0; s < n; ++s) {
        const int ch = chans[s];
        o[s] = out[ch].K3;
        r[s] = (ch == TCH) ? c->k3raw : out[ch].K3;
        rp[s] = c->sg[ch == PCH ? (left ? SG_PPL3 : SG_PP3) : (left ? SG_PHL3 : SG_PH3)].repop;
        if (ch == TCH) f.tslot = s;
        if (ch == ACH) f.aslot = s;
    }
    f.out0 = o[0]; f.out1 = o[1]; f.out2 = o[2];
    f.raw0 = r[0]; f.raw1 = r[1]; f.raw2 = r[2];
    f.rep0 = rp[0]; f.rep1 = rp[1]; f.rep2 = rp[2];
    f.afinal = out[ACH].K3;
    k_k3_finish<<<(unsigned)c->finmap_n, 256, 0, c->run>>>(f, c->finmap, c->fin_nchunk, c->NV3 * c->NV3);
    LAUNCHCHK(c);
    return 0;
}

// true when every channel of the launch has a K3 symmetry group (local solver, NP = 1)
static bool k3_sym(const fdga_ctx* c, const int* chans, int n, bool left) {
    for (int s = 0; s < n; ++s)
        if (!sym_k3(c, chans[s], left).ncls) return false;
    return n > 0;
}

// BSE_L_K3! for `n` channels in one launch.  A single t channel without a symmetry group is
// post-processed in the epilogue (its a channel is complete); otherwise t is post-processed
// after the launch (and after the symmetry fill).
struct LK3Set {
    LK3_Aop a;
    LK3_Bop b;
    LK3_Ep e;
    LK3R_Ep er;   // the epilogue that also writes BSE_K3!'s FP32 right operand (when ro)
    bool ro;
};
// the lattice solver's FP32 contraction without K3 symmetry groups: BSE_L_K3!'s epilogue, which
// holds every element of FL.K3, also writes BSE_K3!'s right operand (no separate pass over FL.K3)
static bool lk3_writes_r32(const fdga_ctx* c, bool sym) {
    return c->d.precision == FDGA_FP32 && c->NP > 1 && !sym;
}
// operand loaders and epilogue of a BSE_L_K3! launch (batch = (slot NW3 + iΩ) NP... see K3Geom)
static LK3Set lk3_set(fdga_ctx* c, const int* chans, int n, bool sym, bool t_epi, bool f32 = false) {
    LK3Chan cc[3] = {};
    for (int s = 0; s < n; ++s) {
        const int ch = chans[s];
        cc[s].G = c->cache[idx_gamma(ch)];
        cc[s].F0r = c->cache[idx_f0(ch)];
        cc[s].Pi0 = pi_of(c, ch, true);
        cc[s].out = (sym && ch == TCH) ? c->k3raw : c->FL[ch].K3;
        cc[s].outa = c->FL[ACH].K3;
        cc[s].scale = c->d.T * ((ch == ACH) ? 1.0 : -1.0);   // src/BSE_templates.jl:122,130,138
        cc[s].tpost = (ch == TCH && t_epi);
        if (lk3_writes_r32(c, sym)) {   // BSE_K3!'s FP32 right operand from the same epilogue
            cc[s].R32 = reinterpret_cast<float2*>(c->R[ch]);
            cc[s].Pi = pi_of(c, ch, false);
            cc[s].Gam = c->cache[(ch == PCH) ? 0 : (ch == ACH) ? 10 : 11];
            cc[s].s1 = (ch == ACH) ? 1.0 : -1.0;   // src/BSE_templates.jl:156,164,172
            cc[s].s2 = (ch == TCH) ? -1.0 : 1.0;
            cc[s].ch = ch;
            if (f32) {
                cc[s].out32 = reinterpret_cast<float2*>(c->FL[ch].K3);
                cc[s].outa32 = reinterpret_cast<const float2*>(c->FL[ACH].K3);
            }
        }
    }
    const K3Geom g = k3geom(c);
    return LK3Set{LK3_Aop{g, cc[0], cc[1], cc[2]}, LK3_Bop{g, cc[0], cc[1], cc[2]}, LK3_Ep{g, cc[0], cc[1], cc[2]},
                  LK3R_Ep{g, cc[0], cc[1], cc[2]}, lk3_writes_r32(c, sym)};
}
// one K3-class GEMM launch of `n` channels with the context's tile shape for it (which: 0 BSE_L_K3!,
// 1 BSE_K3!), over the symmetry block list when there is one
template <class AL, class BL, class EP>
static hipError_t k3_gemm(fdga_ctx* c, const int* chans, int n, int which, const AL& a, const BL& b, const EP& e);
// BSE_L_K3! launch with the set's epilogue
static hipError_t lk3_gemm(fdga_ctx* c, const int* chans, int n, const LK3Set& ls) {
    return ls.ro ? k3_gemm(c, chans, n, 0, ls.a, ls.b, ls.er) : k3_gemm(c, chans, n, 0, ls.a, ls.b, ls.e);
}
template <class AL, class BL, class EP>
static hipError_t k3_gemm(fdga_ctx* c, const int* chans, int n, int which, const AL& a, const BL& b, const EP& e) {
    const int NV = c->NV3, tile = c->ktile[which];
    size_t nl = 0;
    const int* bl = k3_blocks(c, chans, n, which == 0, tile, nl);
    const GemmSchedOne sc = sched_one(NV, NV, NV, cdiv(NV, tile_bm(tile)), cdiv(NV, tile_bn(tile)), n * c->NW3 * c->NP, bl, nl,
                                      c->NW3, c->NP);
    const size_t grid = c->kgrid[which] ? std::min((size_t)c->kgrid[which], (size_t)sc.nitems) : (size_t)sc.nitems;
    return zgemm_run(c->d.precision, tile, grid, c->run, sc, a, b, e);
}
// Without symmetry groups a (p, a, t) contraction over many (Ω, P) batches runs as two launches, (p, a)
// and then t with its post-processing (src/BSE_templates.jl:176-177) in the epilogue, where its a
// channel is complete: one K3-sized pass (k_post_t) less.  Few batches (the local solver) keep one
// launch, where a second, small launch would not fill the GPU.
static bool split_t(const fdga_ctx* c, const int* chans, int n, bool sym) {
    return !sym && n == 3 && chans[0] == PCH && chans[1] == ACH && chans[2] == TCH && (size_t)c->NW3 * c->NP >= 512;
}
static const int kChansPA[2] = {PCH, ACH};
static const int kChansT[1] = {TCH};
static const int kChansAll[3] = {PCH, ACH, TCH};

// The FP32 lattice iteration (its three-channel BSE_L_K3! in two launches, no K3 symmetry groups)
// keeps FL.K3 in FP32, in the first half of each FL.K3 array's bytes: BSE_L_K3!'s epilogue writes it
// rounded and BSE_K3! reads it (1.6 GB less traffic per iteration at L = 16).  fl_widen / fl_narrow
// convert the three arrays through the right-operand buffers (so the fresh right operands are lost);
// they run eagerly, never inside a captured graph (fl_prepare before each iteration).
static bool fl32_mode(const fdga_ctx* c) {
    return lk3_writes_r32(c, false) && split_t(c, kChansAll, 3, k3_sym(c, kChansAll, 3, true));
}
static int fl_widen(fdga_ctx* c) {
    if (!c->fl32) return 0;
    const size_t n = (size_t)c->n3 * c->NP;
    for (int r = 0; r < 3; ++r) {
        float2* tmp = reinterpret_cast<float2*>(c->R[r]);
        HIPCHK(c, hipMemcpyAsync(tmp, c->FL[r].K3, n * sizeof(float2), hipMemcpyDeviceToDevice, c->run));
        k_widen32<<<cdiv(n, 256), 256, 0, c->run>>>(c->FL[r].K3, tmp, n);
        LAUNCHCHK(c);
    }
    c->fl32 = 0;
    c->r32_fresh = 0;
    return 0;
}
static int fl_narrow(fdga_ctx* c) {
    if (c->fl32) return 0;
    const size_t n = (size_t)c->n3 * c->NP;
    for (int r = 0; r < 3; ++r) {
        float2* tmp = reinterpret_cast<float2*>(c->R[r]);
        k_narrow32<<<cdiv(n, 256), 256, 0, c->run>>>(tmp, c->FL[r].K3, n);
        LAUNCHCHK(c);
        HIPCHK(c, hipMemcpyAsync(c->FL[r].K3, tmp, n * sizeof(float2), hipMemcpyDeviceToDevice, c->run));
    }
    c->fl32 = 1;
    c->r32_fresh = 0;
    return 0;
}
// The same for F.K3 and Fbuff.K3 (six arrays): the FP32 lattice iteration reads F.K3 (BSE_K3!'s left
// operand, the K2 finish, the SDE, the momentum averages) and writes Fbuff.K3 in FP32: some 17
// passes over a channel's K3 at half the bytes, about 4.4 GB less traffic per iteration at L = 16.
// Flatten, unflatten and the array accessors convert on the fly; anything else that reads or writes
// them raw widens first.
static int fk_convert(fdga_ctx* c, bool narrow) {
    if ((c->fk32 != 0) == narrow) return 0;
    const size_t n = (size_t)c->n3 * c->NP;
    float2* tmp = reinterpret_cast<float2*>(c->R[0]);
    for (DevChan* v : {c->F, c->Fb})
        for (int r = 0; r < 3; ++r) {
            if (narrow) {
                k_narrow32<<<cdiv(n, 256), 256, 0, c->run>>>(tmp, v[r].K3, n);
                LAUNCHCHK(c);
                HIPCHK(c, hipMemcpyAsync(v[r].K3, tmp, n * sizeof(float2), hipMemcpyDeviceToDevice, c->run));
            } else {
                HIPCHK(c, hipMemcpyAsync(tmp, v[r].K3, n * sizeof(float2), hipMemcpyDeviceToDevice, c->run));
                k_widen32<<<cdiv(n, 256), 256, 0, c->run>>>(v[r].K3, tmp, n);
                LAUNCHCHK(c);
            }
        }
    c->fk32 = narrow;
    c->r32_fresh = 0;
    // the momentum averages of the rounded values (as unflatten would form them)
    return narrow ? refresh_avg(c) : 0;
}
static int fk_widen(fdga_ctx* c) { return fk_convert(c, false); }
// the FL.K3, F.K3 and Fbuff.K3 forms an iteration of this context works in
static int fl_prepare(fdga_ctx* c) {
    const bool f32 = fl32_mode(c);
    return (f32 ? fl_narrow(c) : fl_widen(c)) || fk_convert(c, f32);
}

static int launch_lk3(fdga_ctx* c, const int* chans, int n) {
    const bool sym = k3_sym(c, chans, n, true);
    int bits = 0;
    for (int s = 0; s < n; ++s) bits |= 1 << chans[s];
    c->r32_fresh &= ~bits;
    if (split_t(c, chans, n, sym)) {
        // (all three channels are written: FL.K3 takes the FP32 form on the FP32 lattice path)
        const bool f32 = fl32_mode(c);
        if (!f32 && fl_widen(c)) return -3;
        const LK3Set l1 = lk3_set(c, kChansPA, 2, false, false, f32), l2 = lk3_set(c, kChansT, 1, false, true, f32);
        kernel_begin(c);
        if (f32 && c->NV3 <= 64) {   // momentum-batched, the operands of one Ω on chip (same results)
            const K3Geom g = k3geom(c);
            const size_t npg = cdiv(c->NP, LK3P_G);
            // one workgroup per task (slot, Ω, momentum group), or the persistent grid set for BSE_L_K3!
            auto grid = [&](size_t ntask) {
                const size_t full = 8 * (size_t)cdiv(ntask, 8);
                return (unsigned)(c->kgrid[0] ? std::min(full, (size_t)c->kgrid[0]) : full);
            };
            if (c->NV3 == 64) {   // the momenta software-pipelined (Nν3 = 64)
                k_lk3_pb64<false><<<grid(2 * c->NW3 * npg), 512, 0, c->run>>>(g, l1.er.c0, l1.er.c1, 2);
                k_lk3_pb64<true><<<grid(c->NW3 * npg), 512, 0, c->run>>>(g, l2.er.c0, l2.er.c0, 1);
            } else {
                k_lk3_pbatch<<<grid(2 * c->NW3 * npg), 256, 0, c->run>>>(g, l1.er.c0, l1.er.c1, 2);
                k_lk3_pbatch<<<grid(c->NW3 * npg), 256, 0, c->run>>>(g, l2.er.c0, l2.er.c0, 1);
            }
            LAUNCHCHK(c);
        } else if (lk3_gemm(c, kChansPA, 2, l1) != hipSuccess || lk3_gemm(c, kChansT, 1, l2) != hipSuccess) {
            return fail(c, "BSE_L_K3 GEMM launch failed", -3);
        }
        kernel_end(c, 9);
        c->fl32 = f32;
        if (lk3_writes_r32(c, false)) c->r32_fresh |= bits;
        return 0;
    }
    if (fl_widen(c)) return -3;   // a subset of the channels, or symmetry groups: FP64 FL.K3
    const bool t_epi = (n == 1 && chans[0] == TCH && !sym);
    const LK3Set ls = lk3_set(c, chans, n, sym, t_epi);
    kernel_begin(c);
    if (lk3_gemm(c, chans, n, ls) != hipSuccess)
        return fail(c, "BSE_L_K3 GEMM launch failed", -3);
    kernel_end(c, 9);
    if (k3_finish(c, c->FL, chans, n, true, sym, t_epi)) return -3;
    // a t channel post-processed after the launch has its R from the unpost-processed values
    if (lk3_writes_r32(c, sym)) c->r32_fresh |= bits;
    return 0;
}

// BSE_K3! operands: right-operand builders (k_k3_R) and the GEMM's loaders and epilogue
struct K3Set {
    K3Geom g;
    K3RChan rc[3];
    K3Chan kc[3];
    double T;
    K3_Bop b() const { return K3_Bop{g, T, kc[0], kc[1], kc[2]}; }
    template <int FORM = 2> K3_EpT<FORM> e() const { return K3_EpT<FORM>{g, T, kc[0], kc[1], kc[2]}; }
    template <bool FULL> K3_AopT<FULL> a() const { return K3_AopT<FULL>{g, T, kc[0], kc[1], kc[2]}; }
    bool op32() const { return kc[0].A32 != nullptr; }   // FP32 A and R (lattice solver, FP32 contraction)
    bool r32() const { return kc[0].R32 != nullptr; }    // FP32 R (any solver, FP32 contraction)
    K3_Aop32 a32() const { return K3_Aop32{g, kc[0], kc[1], kc[2]}; }
    K3_Bop32 b32() const { return K3_Bop32{g, kc[0], kc[1], kc[2]}; }
};
// mfrg: the is_mfRG = Val(true) form (src/BSEa/BSEa_K3.jl:83-97): R = Π0 (s1 Γ~ + s2 L~) (the right-operand
// builder with Π0 in place of Π), A = the F caches of build_K3_cache_mfRG! as they are (no K2/K3 of the
// target vertex added)
// fl32: FL.K3 in its FP32 form (-1: as the context holds it now; solver groups fix it per member)
static K3Set k3_set(fdga_ctx* c, const int* chans, int n, bool sym, bool t_epi, bool mfrg = false, int fl32 = -1) {
    const bool f32 = fl32 < 0 ? c->fl32 != 0 : fl32 != 0;
    const bool fk = fl32 < 0 ? c->fk32 != 0 : fl32 != 0;   // F.K3 / Fbuff.K3 in FP32
    K3Set ks{};
    ks.g = k3geom(c);
    ks.T = c->d.T;
    K3RChan* rc = ks.rc;
    K3Chan* kc = ks.kc;
    for (int s = 0; s < n; ++s) {
        const int ch = chans[s];
        rc[s].ch = ch;
        rc[s].s1 = (ch == ACH) ? 1.0 : -1.0;   // src/BSE_templates.jl:156,164,172
        rc[s].s2 = (ch == TCH) ? -1.0 : 1.0;
        rc[s].Pi = pi_of(c, ch, mfrg);
        rc[s].Pi0 = pi_of(c, ch, true);
        rc[s].F0r = c->cache[idx_f0(ch)];
        rc[s].Gam = c->cache[(ch == PCH) ? 0 : (ch == ACH) ? 10 : 11];
        rc[s].Lr = c->FL[ch].K3;
        rc[s].La = c->FL[ACH].K3;
        rc[s].Lr32 = f32 ? reinterpret_cast<const float2*>(c->FL[ch].K3) : nullptr;
        rc[s].La32 = f32 ? reinterpret_cast<const float2*>(c->FL[ACH].K3) : nullptr;
        rc[s].R = c->R[ch];
        rc[s].csum = c->csum[ch];
        kc[s].ch = ch;
        kc[s].Fbase = c->cache[idx_fbase(ch)];
        kc[s].full = (c->NP == 1) || c->caches_mfrg;
        kc[s].K3r = c->F[ch].K3;
        kc[s].K3a = c->F[ACH].K3;
        kc[s].K2r = c->F[ch].K2;
        kc[s].K2a = c->F[ACH].K2;
        kc[s].R = c->R[ch];
        kc[s].csum = c->csum[ch];
        kc[s].Lr = c->FL[ch].K3;
        kc[s].La = c->FL[ACH].K3;
        kc[s].l32 = f32;
        kc[s].out = (sym && ch == TCH) ? c->k3raw : c->Fb[ch].K3;
        kc[s].outa = c->Fb[ACH].K3;
        kc[s].k32 = fk;   // (never with symmetry groups: those are the local solver's)
        kc[s].tpost = (ch == TCH && t_epi);
        if (c->d.precision == FDGA_FP32) {
            // FP32 operands in the right-operand buffer: R32 in its first half, and for the lattice
            // solver A32 in its second (the local solver's A is the full FP64 cache)
            float2* buf = reinterpret_cast<float2*>(c->R[ch]);
            rc[s].R32 = buf;
            kc[s].R32 = buf;
            if (!kc[s].full) {
                // A32 carries the rank-1 K2r(ν) term, so R needs no column sums
                kc[s].A32 = buf + c->n3 * c->NP;
                kc[s].K2row = c->csum[ch];   // the column-sum buffer is free on this path
                kc[s].csum = nullptr;
                rc[s].csum = nullptr;
            }
        }
    }
    return ks;
}
// the right operands R and their column sums of a BSE_K3! launch over `n` channels
// With `sym` (BSE_K3! symmetry groups installed) the lattice solver builds the operands only at the
// momenta whose batches hold class representatives: the contraction reads no other.
// skipR: the FP32 right operands are already there (written by BSE_L_K3!'s epilogue)
static int k3_rhs(fdga_ctx* c, const K3Set& ks, int n, bool sym = false, bool skipR = false) {
    const int* pl = (sym && c->k3_pl) ? c->k3_pl : nullptr;
    const int npl = pl ? c->k3_npl : c->NP;
    if (skipR) {
    } else if (c->NV3 <= WAVE && c->NP > 1) {   // lattice: P-independent operands loaded once per momentum chunk
        const size_t nwave = (size_t)cdiv(c->NV3, K3R_COLS) * c->NW3 * cdiv(npl, K3_PCH) * n;
        k_k3_R_pchunk<1, K3R_COLS><<<cdiv(nwave * WAVE, 256), 256, 0, c->run>>>(ks.g, ks.rc[0], ks.rc[1], ks.rc[2], n, pl,
                                                                                npl);
    } else if (c->NV3 <= 2 * WAVE && c->NP > 1) {   // two ω per lane, two columns per wave
        const size_t nwave = (size_t)cdiv(c->NV3, 2) * c->NW3 * cdiv(npl, K3_PCH) * n;
        k_k3_R_pchunk<2, 2><<<cdiv(nwave * WAVE, 256), 256, 0, c->run>>>(ks.g, ks.rc[0], ks.rc[1], ks.rc[2], n, pl, npl);
    } else {
        const size_t nwave = (size_t)cdiv(c->NV3, K3R_COLS) * c->NW3 * c->NP * n;
        k_k3_R<<<cdiv(nwave * WAVE, 256), 256, 0, c->run>>>(ks.g, ks.rc[0], ks.rc[1], ks.rc[2], n);
    }
    LAUNCHCHK(c);
    if (!skipR)
        for (int s = 0; s < n; ++s) c->r32_fresh &= ~(1 << ks.kc[s].ch);
    if (ks.op32()) {
        k_k2_rows<<<cdiv((size_t)c->NV3 * c->NW3 * c->NP * n, 256), 256, 0, c->run>>>(ks.g, ks.kc[0], ks.kc[1], ks.kc[2], n);
        LAUNCHCHK(c);
        // the three slots (p, a, t) in one pass when the launch holds them in that order
        const K3Chan* kc = ks.kc;
        const bool pat = n == 3 && kc[0].ch == PCH && kc[1].ch == ACH && kc[2].ch == TCH && kc[0].k32 == kc[1].k32 &&
                         kc[1].k32 == kc[2].k32;
        const dim3 grid((unsigned)(c->NW3 * cdiv(npl, K3_PCH) * (pat ? 1 : n)), cdiv((size_t)c->NV3 * c->NV3, 256));
        if (pat && kc[0].k32) k_k3_A32_pat<true><<<grid, 256, 0, c->run>>>(ks.g, kc[0], kc[1], kc[2], pl, npl);
        else if (pat) k_k3_A32_pat<false><<<grid, 256, 0, c->run>>>(ks.g, kc[0], kc[1], kc[2], pl, npl);
        else k_k3_A32<<<grid, 256, 0, c->run>>>(ks.g, kc[0], kc[1], kc[2], pl, npl);
        LAUNCHCHK(c);
    }
    return 0;
}
// the BSE_K3! GEMM launch of `n` channels with the loaders of the solver kind and precision
static hipError_t k3_gemm_set(fdga_ctx* c, const int* chans, int n, const K3Set& ks) {
    // the FP64 solver's arrays are all FP64; the FP32 lattice iteration's all FP32
    if (c->d.precision == FDGA_FP64)
        return ks.kc[0].full ? k3_gemm(c, chans, n, 1, ks.a<true>(), ks.b(), ks.e<0>())
                             : k3_gemm(c, chans, n, 1, ks.a<false>(), ks.b(), ks.e<0>());
    bool all32 = true, any32 = false;
    for (int s = 0; s < n; ++s) {
        all32 = all32 && ks.kc[s].l32 && ks.kc[s].k32;
        any32 = any32 || ks.kc[s].l32 || ks.kc[s].k32;
    }
    if (ks.kc[0].full)   // the local solver (FP64 forms), or build_K3_cache_mfRG!'s caches
        return any32 ? k3_gemm(c, chans, n, 1, ks.a<true>(), ks.b32(), ks.e<2>())
                     : k3_gemm(c, chans, n, 1, ks.a<true>(), ks.b32(), ks.e<0>());
    return all32 ? k3_gemm(c, chans, n, 1, ks.a32(), ks.b32(), ks.e<1>()) : k3_gemm(c, chans, n, 1, ks.a32(), ks.b32(), ks.e<2>());
}
// BSE_K3! for `n` channels: right operands and column sums, then the GEMM launch(es)
static int launch_k3(fdga_ctx* c, const int* chans, int n, bool mfrg = false) {
    if (mfrg && c->NP > 1 && !c->caches_mfrg)
        return fail(c, "BSE_K3!(S, Ch, Val(true)) of an NL_ParquetSolver reads the F caches of build_K3_cache_mfRG!(S, true); "
                       "the device keeps only the momentum-independent part of build_K3_cache!'s");
    const bool sym = k3_sym(c, chans, n, false);
    // the right operands from BSE_L_K3!'s epilogue when it wrote them for every channel of the launch
    bool fresh = !mfrg && !sym && lk3_writes_r32(c, false) && !c->caches_mfrg;
    for (int s = 0; s < n; ++s) fresh = fresh && (c->r32_fresh >> chans[s] & 1);
    if (split_t(c, chans, n, sym)) {
        if (k3_rhs(c, k3_set(c, chans, n, false, false, mfrg), n, false, fresh)) return -3;
        kernel_begin(c);
        if (k3_gemm_set(c, kChansPA, 2, k3_set(c, kChansPA, 2, false, false, mfrg)) != hipSuccess ||
            k3_gemm_set(c, kChansT, 1, k3_set(c, kChansT, 1, false, true, mfrg)) != hipSuccess)
            return fail(c, "BSE_K3 GEMM launch failed", -3);
        kernel_end(c, 10);
        return 0;
    }
    if (c->fk32) {   // a subset of the channels, or symmetry groups: FP64 F.K3 and Fbuff.K3
        if (fk_widen(c)) return -3;
        fresh = false;
    }
    const bool t_epi = (n == 1 && chans[0] == TCH && !sym);
    const K3Set ks = k3_set(c, chans, n, sym, t_epi, mfrg);
    if (k3_rhs(c, ks, n, sym, fresh)) return -3;
    kernel_begin(c);
    if (k3_gemm_set(c, chans, n, ks) != hipSuccess) return fail(c, "BSE_K3 GEMM launch failed", -3);
    kernel_end(c, 10);
    return k3_finish(c, c->Fb, chans, n, false, sym, t_epi);
}

// ---------------------------------------------------------------------------------------
// the is_mfRG forms of BSE_K1!/BSE_K2! and the _new forms (fdga_forms.h): one wave per output class
template <int FORM, int CH>
static int launch_form(fdga_ctx* c, int is_mfrg) {
    constexpr bool K2 = (FORM == FORM_K2_MFRG || FORM == FORM_K2_NEW);
    FormArgs a{};
    a.F = vtx_F(c);
    a.FL = vtx_FL(c);
    a.f0 = f0_view(c);
    const bool zero = (FORM == FORM_K1_MFRG || FORM == FORM_K2_MFRG);   // the mfRG forms sum against Π0 only
    a.Pi = pi_of(c, CH, zero);
    a.Pi0 = pi_of(c, CH, true);
    a.Pt = pi_tail(c, a.Pi);
    a.P0t = pi_tail(c, a.Pi0);
    a.out = K2 ? c->Fb[CH].K2 : c->Fb[CH].K1;
    a.outa = K2 ? c->Fb[ACH].K2 : c->Fb[ACH].K1;
    a.Lr = c->FL[CH].K2;
    a.La = c->FL[ACH].K2;
    a.U = c->U;   // bare_vertex(F, Sp) for the pSp (p, a) and dSp (t) components
    a.scale = c->d.T * kSign[CH];
    a.nl = c->d.nonlocal;
    a.mfrg = is_mfrg;
    a.n1 = c->d.nK1; a.n2b = c->d.nK2b; a.n2f = c->d.nK2f;
    a.nWP = c->d.nK1; a.nVP = c->nVP; a.wb = c->wb; a.NP = c->NP;
    a.NK = box_bound(c);
    a.sym = K2 ? sym_k2(c, CH) : symcls(c, SG_K1);
    const size_t nw = nclass(a.sym, (K2 ? c->n2 : c->n1) * c->NP);
    k_form<FORM, CH><<<cdiv(nw * WAVE, 256), 256, 0, c->run>>>(a);
    LAUNCHCHK(c);
    return 0;
}
template <int FORM>
static int launch_form_ch(fdga_ctx* c, int ch, int is_mfrg) {
    if (ch == PCH) return launch_form<FORM, PCH>(c, is_mfrg);
    if (ch == TCH) return launch_form<FORM, TCH>(c, is_mfrg);
    if (ch == ACH) return launch_form<FORM, ACH>(c, is_mfrg);
    return fail(c, "bad channel");
}

// build_K3_cache_mfRG!(S, is_first_iteration) (src/nonlocal_2/build_K3_cache.jl:97-164): the caches
// evaluated over the K3 groups SGpp[3] / SGph[3] (a class fill when groups are installed)
static int build_k3_cache_mfrg(fdga_ctx* c, int first) {
    if (first) c->caches_mfrg = 1;
    CachePtrs cp;
    for (int k = 0; k < 12; ++k) cp.c[k] = c->cache[k];
    const size_t n = c->n3;
    k_cache_mfrg<<<cdiv(n, 256), 256, 0, c->run>>>(cp, vtx_F(c), c->d.nK3b, c->d.nK3f, first);
    LAUNCHCHK(c);
    const int* rpp = c->sg[SG_PP3].ncls ? c->sg[SG_PP3].repop : nullptr;
    const int* rph = c->sg[SG_PH3].ncls ? c->sg[SG_PH3].repop : nullptr;
    const int fills[7][2] = {{0, 0}, {4, 0}, {5, 1}, {6, 1}, {7, 0}, {8, 1}, {9, 1}};
    for (const auto& f : fills) {
        if (f[0] >= 7 && !first) continue;
        if (sym_fill(c, c->cache[f[0]], f[1] ? rph : rpp, n)) return -3;
    }
    k_cache_transpose<<<cdiv(n, 256), 256, 0, c->run>>>(c->cache[10], c->cache[11], c->cache[5], c->cache[6], c->NV3,
                                                         c->NW3);
    LAUNCHCHK(c);
    return 0;
}

// =========================================================================================
extern "C" {

const char* fdga_last_error(const fdga_ctx* c) { return c ? c->err.c_str() : "null context"; }

int fdga_create(fdga_ctx** out, const fdga_dims* dims) {
    if (!out || !dims) return -1;
    *out = nullptr;
    auto* c = new fdga_ctx();
    c->d = *dims;
    fdga_dims& d = c->d;
    if (!d.nonlocal) { d.L = 1; d.LG = 1; }
    // constructor assertions (src/channel.jl:33-40) and box relations the kernels rely on
    if (d.nK1 < d.nK2b || d.nK1 < d.nK2f || d.nK2b < d.nK3b || d.nK2f < d.nK3f || d.nK3b < 1 || d.nK3f < 1 ||
        d.mPinu < 1 || d.nG < 1 || d.L < 1 || d.LG < 1 || d.T <= 0.0) {
        fail(c, "invalid dimensions (need nK1 >= nK2 >= nK3 >= 1, mPinu >= 1, nG >= 1, T > 0)");
        *out = c;
        return -1;
    }
    if (d.L > 32 || d.LG > 64) {
        fail(c, "the lattice bubble transforms L x L planes in LDS and indexes the LG x LG mesh of G(R) in 12 bits: "
                "L must be at most 32 and LG at most 64");
        *out = c;
        return -1;
    }
    if (d.nK1 * d.mPinu < d.nK3f) {
        fail(c, "bubble mesh nK1*mPinu must contain the K3 fermionic box");
        *out = c;
        return -1;
    }
    if (d.ref_kind == 1 && (d.r_nK1 < d.r_nK2b || d.r_nK1 < d.r_nK2f || d.r_nK2b < d.r_nK3b || d.r_nK2f < d.r_nK3f || d.r_nK3b < 1)) {
        fail(c, "invalid reference Vertex boxes");
        *out = c;
        return -1;
    }
    if (d.precision != FDGA_FP64 && d.precision != FDGA_FP32) {
        fail(c, "unknown precision (FDGA_FP64 = 0, FDGA_FP32 = 1)");
        *out = c;
        return -1;
    }
    if (d.tab_nb < 1) d.tab_nb = 1;
    if (d.tab_nf < 1) d.tab_nf = 1;
    (void)hipGetLastError();
    if (hipStreamCreate(&c->st) != hipSuccess || hipStreamCreate(&c->st2) != hipSuccess ||
        hipEventCreateWithFlags(&c->evfork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->evjoin, hipEventDisableTiming) != hipSuccess) {
        fail(c, "hipStreamCreate failed (no HIP device?)", -2);
        *out = c;
        return -2;
    }
    c->run = c->st;
    // the scratch array's address on this device, taken once per device with the device current
    if (hipGetDevice(&c->device) != hipSuccess || c->device < 0 || c->device >= kMaxDevices) {
        fail(c, "hipGetDevice failed", -2);
        *out = c;
        return -2;
    }
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[c->device]) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                fail(c, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e), -2);
                *out = c;
                return -2;
            }
            g_scratch_ptr[c->device] = static_cast<double2*>(p);
        }
        c->scratch = g_scratch_ptr[c->device];
    }
    c->NP = d.L * d.L;
    c->NkG = d.LG * d.LG;
    c->NG = 2 * d.nG;
    c->NW1 = 2 * d.nK1 - 1;
    c->NW2 = 2 * d.nK2b - 1;
    c->NV2 = 2 * d.nK2f;
    c->NW3 = 2 * d.nK3b - 1;
    c->NV3 = 2 * d.nK3f;
    c->NWP = 2 * d.nK1 - 1;
    c->nVP = d.nK1 * d.mPinu;
    c->NVP = 2 * c->nVP;
    // the gathers read W and the tail sums only for |ω| <= Wc = NK + |Ω| + |ν| + 2 with Ω, ν in the
    // K1 or K2 box (BSE_K1!, BSE_K2!, SDE L-vertices)
    c->wb = std::min(c->nVP, box_bound(c) + std::max(d.nK1 - 1, d.nK2b - 1 + d.nK2f) + 3);
    c->nGel = (size_t)c->NG * c->NkG;
    c->nPiel = (size_t)c->NVP * c->NWP * c->NP;
    c->n1 = (size_t)c->NW1;
    c->n2 = (size_t)c->NW2 * c->NV2;
    c->n3 = (size_t)c->NV3 * c->NV3 * c->NW3;
    c->nL = c->n2 * c->NP;
    c->U = cmk(d.U, 0.0);
    int rc = 0;
    for (double2** p : {&c->Gbare, &c->G0, &c->Sig0, &c->G, &c->Sig, &c->tmpG1, &c->tmpG2, &c->tmpG3})
        rc |= dalloc(c, p, c->nGel);
    for (double2** p : {&c->Pi0pp, &c->Pi0ph, &c->Pipp, &c->Piph}) rc |= dalloc(c, p, c->nPiel);
    for (double2** p : {&c->Wc[0], &c->Wc[1], &c->Wc[2]}) rc |= dalloc(c, p, (size_t)c->NWP * c->NP * (2 * (size_t)c->wb));
    c->ntail = (size_t)c->NWP * c->NP * (2 * (size_t)c->wb + 2);
    for (double2** p : {&c->Ptail[0], &c->Ptail[1], &c->Ptail[2], &c->Ptail[3]}) rc |= dalloc(c, p, c->ntail);
    for (double2** p : {&c->Wfar[0], &c->Wfar[1], &c->Wfar[2]}) rc |= dalloc(c, p, (size_t)4 * c->NWP * c->NP);
    for (double2** p : {&c->Lpp, &c->Lph, &c->tmpL1, &c->tmpL2}) rc |= dalloc(c, p, c->nL);
    if (d.nonlocal) {
        const size_t nch = (size_t)c->NP * cdiv(c->NVP, bubble_nb(d.L)) * c->NWP;
        rc |= dalloc(c, &c->pchunk[0], nch);
        rc |= dalloc(c, &c->pchunk[1], nch);
        rc |= dalloc(c, &c->k2x, (size_t)c->NV2 * (2 * (size_t)c->wb) * c->NW2);
        rc |= dalloc(c, &c->k2xf, (size_t)c->NV2 * 2 * c->NW2);
        rc |= dalloc(c, &c->k2y, (size_t)c->NV2 * c->NW2 * c->NP);
        rc |= dalloc(c, &c->lk2x, (size_t)c->NV2 * c->NV2 * c->NW2);
    }
    rc |= dalloc(c, &c->sde0, c->nGel);
    rc |= dalloc(c, &c->u2C, (size_t)(4 * d.nG - 1) * c->NkG);
    size_t ntab = (size_t)(2 * d.tab_nb - 1) * (2 * d.tab_nf) * (2 * d.tab_nf);
    for (auto& t : c->tab) rc |= dalloc(c, &t, ntab);
    for (int r = 0; r < 3; ++r) {
        rc |= alloc_chan(c, c->F[r], c->n1, c->n2, c->n3, c->NP, true);
        rc |= alloc_chan(c, c->Fb[r], c->n1, c->n2, c->n3, c->NP, false);
        rc |= alloc_chan(c, c->FL[r], c->n1, c->n2, c->n3, c->NP, false);
        if (d.ref_kind == 1) {
            size_t r1 = (size_t)(2 * d.r_nK1 - 1), r2 = (size_t)(2 * d.r_nK2b - 1) * (2 * d.r_nK2f);
            size_t r3 = (size_t)(2 * d.r_nK3f) * (2 * d.r_nK3f) * (2 * d.r_nK3b - 1);
            rc |= alloc_chan(c, c->R0[r], r1, r2, r3, 1, false);
        }
    }
    for (auto& p : c->cache) rc |= dalloc(c, &p, c->n3);
    for (int r = 0; r < 3; ++r) {
        rc |= dalloc(c, &c->R[r], c->n3 * c->NP);
        rc |= dalloc(c, &c->csum[r], (size_t)c->NV3 * c->NW3 * c->NP);
    }
    if (hipMalloc((void**)&c->scal, (8 + OCC_BLOCKS) * sizeof(double)) != hipSuccess) rc = -2;
    if (d.nonlocal && !rc) {
        const std::vector<BubbleFold> f = bubble_folds(d.L, d.LG);
        if (hipMalloc((void**)&c->folds, f.size() * sizeof(BubbleFold)) != hipSuccess ||
            hipMemcpy(c->folds, f.data(), f.size() * sizeof(BubbleFold), hipMemcpyHostToDevice) != hipSuccess)
            rc = -2;
    }
    for (auto& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) rc = -2;
    if (rc) {
        if (c->err.empty()) c->err = "device allocation failed";
        *out = c;
        return -2;
    }
    if (hipStreamSynchronize(c->st) != hipSuccess) {
        *out = c;
        return fail(c, "stream synchronisation failed", -2);
    }
    *out = c;
    return 0;
}

int fdga_destroy(fdga_ctx* c) {
    if (!c) return 0;
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (double2* p : {c->Gbare, c->G0, c->Sig0, c->G, c->Sig, c->tmpG1, c->tmpG2, c->tmpG3, c->Pi0pp, c->Pi0ph,
                       c->Pipp, c->Piph, c->Wc[0], c->Wc[1], c->Wc[2], c->Lpp, c->Lph, c->tmpL1, c->tmpL2, c->sde0, c->u2C,
                       c->stage, c->Wfar[0], c->Wfar[1], c->Wfar[2], c->Ptail[0], c->Ptail[1], c->Ptail[2], c->Ptail[3],
                       c->k2x, c->k2xf, c->k2y, c->lk2x, c->pchunk[0], c->pchunk[1]})
        if (p) (void)hipFree(p);
    for (auto* t : c->tab) if (t) (void)hipFree(t);
    for (auto* p : c->cache) if (p) (void)hipFree(p);
    for (int r = 0; r < 3; ++r) {
        if (c->R[r]) (void)hipFree(c->R[r]);
        if (c->csum[r]) (void)hipFree(c->csum[r]);
    }
    for (int r = 0; r < 3; ++r) {
        free_chan(c->F[r]); free_chan(c->Fb[r]); free_chan(c->FL[r]);
        if (c->R0[r].K1) free_chan(c->R0[r]);
    }
    if (c->scal) (void)hipFree(c->scal);
    if (c->folds) (void)hipFree(c->folds);
    graphs_invalidate(c);
    for (auto& g : c->sg) free_sym(g);
    free_blists(c);
    if (c->k3raw) (void)hipFree(c->k3raw);
    if (c->k2raw) (void)hipFree(c->k2raw);
    if (c->k3_pl) (void)hipFree(c->k3_pl);
    for (void* q : {(void*)c->adj_cbar, (void*)c->adj_avg, (void*)c->adj_Ab, (void*)c->adj_off, (void*)c->adj_src,
                    (void*)c->adj_cf})
        if (q) (void)hipFree(q);
    for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
    if (c->st2) (void)hipStreamSynchronize(c->st2);
    if (c->evfork) (void)hipEventDestroy(c->evfork);
    if (c->evjoin) (void)hipEventDestroy(c->evjoin);
    if (c->st2) (void)hipStreamDestroy(c->st2);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return 0;
}

// array table: pointer, element count, layout kind (0 plain, 1 K3, 2 bubble, 3 cache, 4 F-cache)
struct ArrInfo { double2* p; size_t n; int kind; int NW, NV, NP; int ch; };

static int arr_info(fdga_ctx* c, int which, ArrInfo& a) {
    const fdga_dims& d = c->d;
    a = ArrInfo{nullptr, 0, 0, 0, 0, 1, 0};
    switch (which) {
        case FDGA_GBARE: a.p = c->Gbare; a.n = c->nGel; return 0;
        case FDGA_G0: a.p = c->G0; a.n = c->nGel; return 0;
        case FDGA_SIGMA0: a.p = c->Sig0; a.n = c->nGel; return 0;
        case FDGA_G: a.p = c->G; a.n = c->nGel; return 0;
        case FDGA_SIGMA: a.p = c->Sig; a.n = c->nGel; return 0;
        case FDGA_PI0PP: case FDGA_PI0PH: case FDGA_PIPP: case FDGA_PIPH: {
            double2* ps[4] = {c->Pi0pp, c->Pi0ph, c->Pipp, c->Piph};
            a.p = ps[which - FDGA_PI0PP]; a.n = c->nPiel; a.kind = 2; a.NW = c->NWP; a.NV = c->NVP; a.NP = c->NP;
            return 0;
        }
        case FDGA_LPP: a.p = c->Lpp; a.n = c->nL; return 0;
        case FDGA_LPH: a.p = c->Lph; a.n = c->nL; return 0;
        case FDGA_TAB_FPP: case FDGA_TAB_FPX: case FDGA_TAB_FTP: case FDGA_TAB_FTX:
            a.p = c->tab[which - FDGA_TAB_FPP];
            a.n = (size_t)(2 * d.tab_nb - 1) * (2 * d.tab_nf) * (2 * d.tab_nf);
            return 0;
        default: break;
    }
    if (which >= FDGA_CACHE0 && which < FDGA_CACHE0 + 10) {
        int k = which - FDGA_CACHE0;
        a.p = c->cache[k]; a.n = c->n3 * c->NP; a.kind = (k >= 7) ? 4 : 3;
        a.NW = c->NW3; a.NV = c->NV3; a.NP = c->NP; a.ch = k - 7;
        return 0;
    }
    if (which >= FDGA_VERTEX0 && which < FDGA_VERTEX0 + 64) {
        int k = which - FDGA_VERTEX0, v = k / 16, ch = (k % 16) / 3, cls = (k % 16) % 3 + 1;
        if (ch > 2) return -1;
        if (v == 3 && d.ref_kind != 1) return -1;
        DevChan& dc = v == 0 ? c->F[ch] : v == 1 ? c->Fb[ch] : v == 2 ? c->FL[ch] : c->R0[ch];
        int NP = (v == 3) ? 1 : c->NP;
        if (cls == 1) { a.p = dc.K1; a.n = (size_t)(v == 3 ? 2 * d.r_nK1 - 1 : c->NW1) * NP; return 0; }
        if (cls == 2) {
            a.p = dc.K2;
            a.n = (v == 3 ? (size_t)(2 * d.r_nK2b - 1) * (2 * d.r_nK2f) : c->n2) * NP;
            return 0;
        }
        a.p = dc.K3; a.kind = 1; a.NP = NP;
        a.NW = v == 3 ? 2 * d.r_nK3b - 1 : c->NW3;
        a.NV = v == 3 ? 2 * d.r_nK3f : c->NV3;
        a.n = (size_t)a.NW * a.NV * a.NV * NP;
        return 0;
    }
    return -1;
}

int64_t fdga_size(const fdga_ctx* c, int32_t which) {
    ArrInfo a;
    if (!c || arr_info(const_cast<fdga_ctx*>(c), which, a)) return -1;
    return (int64_t)a.n;
}

static void invalidate(fdga_ctx* c, int which) {
    if (which == FDGA_G0 || which == FDGA_SIGMA0 || which == FDGA_PI0PP || which == FDGA_PI0PH ||
        (which >= FDGA_TAB_FPP && which <= FDGA_TAB_FTX) ||
        (which >= FDGA_VERTEX(3, 0, 1) && which <= FDGA_VERTEX(3, 2, 3))) {
        c->sde0_valid = 0;
        graphs_invalidate(c);   // a captured fdPA iteration reuses the cached SDE of the reference
    }
}

// F.K3 or Fbuff.K3 of a channel while they hold FP32 values
static bool fk_array(const fdga_ctx* c, int which) {
    const int k = which - FDGA_VERTEX0;
    return c->fk32 && k >= 0 && k < 32 && (k % 16) < 9 && (k % 16) % 3 == 2;
}

int fdga_set(fdga_ctx* c, int32_t which, const double* host) {
    if (!c || !host) return -1;
    c->r32_fresh = 0;
    ArrInfo a;
    if (arr_info(c, which, a)) return fail(c, "unknown array id " + std::to_string(which));
    if (which >= FDGA_VERTEX(2, 0, 1) && which <= FDGA_VERTEX(2, 2, 3) && fl_widen(c)) return -3;
    if (which == FDGA_LPP || which == FDGA_LPH || a.kind == 3 || a.kind == 4) return fail(c, "array is read-only");
    const size_t bytes = a.n * sizeof(double2);
    if (a.kind == 0) {
        HIPCHK(c, hipMemcpyAsync(a.p, host, bytes, hipMemcpyHostToDevice, c->st));
    } else {
        if (ensure_stage(c, a.n)) return -2;
        HIPCHK(c, hipMemcpyAsync(c->stage, host, bytes, hipMemcpyHostToDevice, c->st));
        if (a.kind == 1 && fk_array(c, which))
            k_k3_to_dev32<<<cdiv(a.n, 256), 256, 0, c->run>>>(reinterpret_cast<float2*>(a.p), c->stage, a.NW, a.NV, a.NP);
        else if (a.kind == 1) k_k3_to_dev<<<cdiv(a.n, 256), 256, 0, c->run>>>(a.p, c->stage, a.NW, a.NV, a.NP);
        else k_pi_to_dev<<<cdiv(a.n, 256), 256, 0, c->run>>>(a.p, c->stage, a.NW, a.NV, a.NP);
        LAUNCHCHK(c);
        if (a.kind == 2 && tail_sums(c, a.p, pi_tail(c, a.p))) return -3;
    }
    if (which >= FDGA_TAB_FPP && which <= FDGA_TAB_FTX) {
        for (size_t i = 0; i < 2 * a.n; ++i)
            if (host[i] != 0.0) { c->tabs_nonzero = 1; break; }
    }
    if (which == FDGA_G) c->pi_band_G = 0;
    if (which == FDGA_PIPP) c->pi_band &= ~1;
    if (which == FDGA_PIPH) c->pi_band &= ~2;
    invalidate(c, which);
    if (which >= FDGA_VERTEX(0, 0, 1) && which <= FDGA_VERTEX(0, 2, 3) && refresh_avg(c)) return -3;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int fdga_get(fdga_ctx* c, int32_t which, double* host) {
    if (!c || !host) return -1;
    ArrInfo a;
    if (arr_info(c, which, a)) return fail(c, "unknown array id " + std::to_string(which));
    if (which >= FDGA_VERTEX(2, 0, 1) && which <= FDGA_VERTEX(2, 2, 3) && fl_widen(c)) return -3;   // FL.K3 in FP64
    if ((which == FDGA_PIPP && (c->pi_band & 1)) || (which == FDGA_PIPH && (c->pi_band & 2))) {
        if (!c->pi_band_G)
            return fail(c, "Πpp/Πph are stored on the band only and G has changed since the bubbles were computed: "
                           "call fdga_bubbles first");
        if (bubbles_into(c, c->Pipp, c->Piph, c->G, false)) return -3;
    }
    const size_t bytes = a.n * sizeof(double2);
    if (a.kind == 0) {
        HIPCHK(c, hipMemcpyAsync(host, a.p, bytes, hipMemcpyDeviceToHost, c->st));
    } else {
        if (ensure_stage(c, a.n)) return -2;
        if (a.kind == 1 && fk_array(c, which))
            k_k3_to_ref32<<<cdiv(a.n, 256), 256, 0, c->run>>>(c->stage, reinterpret_cast<const float2*>(a.p), a.NW, a.NV, a.NP);
        else if (a.kind == 1) k_k3_to_ref<<<cdiv(a.n, 256), 256, 0, c->run>>>(c->stage, a.p, a.NW, a.NV, a.NP);
        else if (a.kind == 2) k_pi_to_ref<<<cdiv(a.n, 256), 256, 0, c->run>>>(c->stage, a.p, a.NW, a.NV, a.NP);
        else if (a.kind == 3) k_cache_to_ref<<<cdiv(a.n, 256), 256, 0, c->run>>>(c->stage, a.p, a.NW, a.NV, a.NP);
        else {
            int ch = a.ch == 0 ? PCH : a.ch == 1 ? ACH : TCH;
            k_cache_F_julia<<<cdiv(a.n, 256), 256, 0, c->run>>>(c->stage, a.p, chanF(c, c->F[ch]), chanF(c, c->F[ACH]),
                                                               ch, c->d.nK3b, c->d.nK3f, c->NP, c->NP == 1 || c->caches_mfrg);
        }
        LAUNCHCHK(c);
        HIPCHK(c, hipMemcpyAsync(host, c->stage, bytes, hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int fdga_init(fdga_ctx* c) {
    if (!c) return -1;
    c->r32_fresh = 0;
    if (bubbles_into(c, c->Pi0pp, c->Pi0ph, c->G0)) return -3;
    HIPCHK(c, hipMemcpyAsync(c->Sig, c->Sig0, c->nGel * sizeof(double2), hipMemcpyDeviceToDevice, c->st));
    k_dyson<<<cdiv(c->nGel, 256), 256, 0, c->run>>>(c->G, c->Sig, c->Gbare, c->nGel);
    LAUNCHCHK(c);
    if (bubbles_into(c, c->Pipp, c->Piph, c->G)) return -3;
    if (refresh_avg(c)) return -3;
    c->sde0_valid = 0;
    graphs_invalidate(c);
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

}  // extern "C"

namespace {
// uploads a group's classes; with `n` > 0 also the per-element table (representative << 2) | op
// over the n elements of the output
int upload_sym(fdga_ctx* c, int id, const fdga_sym::Classes& cl, size_t n = 0) {
    free_sym(c->sg[id]);
    SymDev& g = c->sg[id];
    HIPCHK(c, hipMalloc((void**)&g.off, cl.off.size() * sizeof(int)));
    HIPCHK(c, hipMalloc((void**)&g.mem, cl.mem.size() * sizeof(int)));
    HIPCHK(c, hipMemcpy(g.off, cl.off.data(), cl.off.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(g.mem, cl.mem.data(), cl.mem.size() * sizeof(int), hipMemcpyHostToDevice));
    g.ncls = cl.ncls();
    if (n) {
        std::vector<int> repop(n);
        for (int k = 0; k < g.ncls; ++k) {
            const int64_t dv = cl.mem[cl.off[k]] >> 2;
            for (int q = cl.off[k]; q < cl.off[k + 1]; ++q) repop[cl.mem[q] >> 2] = (int)((dv << 2) | (cl.mem[q] & 3));
        }
        HIPCHK(c, hipMalloc((void**)&g.repop, repop.size() * sizeof(int)));
        HIPCHK(c, hipMemcpy(g.repop, repop.data(), repop.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    return 0;
}
}  // namespace

extern "C" {

int fdga_init_sym_grp(fdga_ctx* c) {
    if (!c) return -1;
    const fdga_dims& d = c->d;
    const bool nl = d.nonlocal;
    using namespace fdga_sym;
    HIPCHK(c, hipStreamSynchronize(c->st));
    graphs_invalidate(c);
    const int64_t NW = c->NW3, NV = c->NV3, NP = c->NP;
    // K3 element counts beyond 2^29 do not fit the (index << 2) | op class encoding
    if ((int64_t)c->n3 * NP >= ((int64_t)1 << 29))
        return fail(c, "init_sym_grp!: the K3 mesh is too large for the symmetry class tables");
    // meshes: the frequency axes, for the lattice solver followed by the momentum mesh (the L x L
    // transfer momenta; Σ: the LG x LG momenta of G)
    auto mesh = [nl](std::vector<Axis> ax, int L) {
        if (nl) ax.push_back(Axis::momentum(L));
        return ax;
    };
    // generators in the order of init_sym_grp!; each frequency map carries the momentum map of its
    // lattice form (src/nonlocal/symmetries.jl), and the lattice groups end with the reflection and
    // the rotation (src/nonlocal/ParquetSolver.jl:198-300).  Momentum maps are unused without a
    // momentum axis.
    auto gens = [nl](std::vector<Generator> g) {
        if (nl) {
            g.push_back(Generator(nullptr, K_REF));
            g.push_back(Generator(nullptr, K_ROT));
        }
        return g;
    };
    const Generator Gs(g_sigma, K_NEG), G1(g_k1, K_NEG), G21(g_k2_1, K_NEG), G2pp(g_k2pp_2, K_ID),
        G2ph(g_k2ph_2, K_NEG), G31(g_k3_1, K_NEG), G32(g_k3_2, K_ID), G3pp(g_k3pp_3, K_ID), G3ph(g_k3ph_3, K_NEG);
    const std::vector<Axis> mS = mesh({Axis(false, d.nG)}, d.LG), m1 = mesh({Axis(true, d.nK1)}, d.L),
                            m2 = mesh({Axis(true, d.nK2b), Axis(false, d.nK2f)}, d.L),
                            m3 = mesh({Axis(true, d.nK3b), Axis(false, d.nK3f), Axis(false, d.nK3f)}, d.L);
    // reference K3 index iΩ + NW (iν + NV (iν' + NV P)) -> device index iν + NV (iν' + NV (iΩ + NW P))
    auto k3dev = [NW, NV](int64_t r) {
        const int64_t iW = r % NW, iv = (r / NW) % NV, rest = r / (NW * NV), ivp = rest % NV, p = rest / NV;
        return iv + NV * (ivp + NV * (iW + NW * p));
    };
    const Classes pp2 = build(m2, gens({G21, G2pp}));
    const Classes ph2 = build(m2, gens({G21, G2ph}));
    // the lattice solver evaluates Σ, K2, L_K2 and the SDE L-vertices at every element and fills the
    // class members afterwards (per-element tables); K1 and the local gathers evaluate by class
    const size_t nS = nl ? c->nGel : 0, n2 = nl ? c->n2 * c->NP : 0;
    int rc = 0;
    rc |= upload_sym(c, SG_S, build(mS, gens({Gs})), nS);
    rc |= upload_sym(c, SG_K1, build(m1, gens({G1})));
    rc |= upload_sym(c, SG_PP2, pp2, n2);
    rc |= upload_sym(c, SG_PH2, ph2, n2);
    const size_t nk3 = c->n3 * c->NP;
    // the four K3 tables (tens of millions of elements on a 16 x 16 lattice) are built concurrently
    const std::vector<Generator> k3gens[4] = {gens({G31, G32, G3pp}), gens({G31, G32, G3ph}), gens({G31, G3pp}),
                                              gens({G31, G3ph})};
    std::future<Classes> k3cls[4];
    for (int k = 0; k < 4; ++k)
        k3cls[k] = std::async(std::launch::async, [&m3, &k3gens, k, k3dev] { return build(m3, k3gens[k], k3dev); });
    const int k3ids[4] = {SG_PP3, SG_PH3, SG_PPL3, SG_PHL3};
    for (int k = 0; k < 4; ++k) rc |= upload_sym(c, k3ids[k], k3cls[k].get(), nk3);
    rc |= upload_sym(c, SG_0PP2, pp2, n2);
    rc |= upload_sym(c, SG_0PH2, ph2, n2);
    if (rc) return -2;
    // GEMM blocks holding a representative (block = (iΩ + NW P) * tiles + ti + tm * tj, rows ν,
    // columns ν'), for every tile shape
    const int64_t NB = NW * NP;   // GEMM batches of one channel
    for (int id : {SG_PP3, SG_PH3, SG_PPL3, SG_PHL3}) {
        SymDev& g = c->sg[id];
        std::vector<int> off(g.ncls + 1), mem;
        HIPCHK(c, hipMemcpy(off.data(), g.off, off.size() * sizeof(int), hipMemcpyDeviceToHost));
        mem.resize(off.back());
        HIPCHK(c, hipMemcpy(mem.data(), g.mem, mem.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (int t = 0; t < TILE_N; ++t) {
            const int tm = cdiv(NV, tile_bm(t)), tn = cdiv(NV, tile_bn(t)), tiles = tm * tn;
            std::vector<char> hit((size_t)NB * tiles, 0);
            for (int k = 0; k < g.ncls; ++k) {
                const int64_t dv = mem[off[k]] >> 2;
                const int iv = (int)(dv % NV), ivp = (int)((dv / NV) % NV);
                const int64_t bt = dv / (NV * NV);
                hit[(size_t)bt * tiles + iv / tile_bm(t) + tm * (ivp / tile_bn(t))] = 1;
            }
            g.tiles[t].clear();
            for (size_t b = 0; b < hit.size(); ++b)
                if (hit[b]) g.tiles[t].push_back((int)b);
        }
    }
    // lattice: the momenta of the BSE_K3! representatives (the operand builders skip the others)
    if (c->k3_pl) (void)hipFree(c->k3_pl);
    c->k3_pl = nullptr;
    c->k3_npl = 0;
    if (nl) {
        std::vector<char> used((size_t)NP, 0);
        for (int id : {SG_PP3, SG_PH3})
            for (int b : c->sg[id].tiles[0]) used[(size_t)(b / (cdiv(NV, tile_bm(0)) * cdiv(NV, tile_bn(0))) / NW)] = 1;
        std::vector<int> pl;
        for (int64_t p = 0; p < NP; ++p)
            if (used[(size_t)p]) pl.push_back((int)p);
        HIPCHK(c, hipMalloc((void**)&c->k3_pl, std::max<size_t>(1, pl.size()) * sizeof(int)));
        HIPCHK(c, hipMemcpy(c->k3_pl, pl.data(), pl.size() * sizeof(int), hipMemcpyHostToDevice));
        c->k3_npl = (int)pl.size();
    }
    free_blists(c);
    if (!c->k3raw && dalloc(c, &c->k3raw, c->n3 * c->NP)) return -2;
    // lattice BSE_K2!: raw class values before the FL.K2 term, p in the first half, a and t in the second
    if (nl && !c->k2raw && dalloc(c, &c->k2raw, 2 * c->n2 * c->NP)) return -2;
    // block lists in XCD order: all tiles of one batch (channel slot, Ω) on one XCD, batch b on XCD b % 8
    for (int t = 0; t < TILE_N; ++t) {
        const int tiles = cdiv(NV, tile_bm(t)) * cdiv(NV, tile_bn(t));
        auto upload_list = [&](int k, const std::vector<int>& v) -> int {
            const std::vector<int> o = xcd_order(v, [tiles](int b) { return b / tiles; });
            HIPCHK(c, hipMalloc((void**)&c->blist[k][t], std::max<size_t>(1, o.size()) * sizeof(int)));
            HIPCHK(c, hipMemcpy(c->blist[k][t], o.data(), o.size() * sizeof(int), hipMemcpyHostToDevice));
            c->blist_n[k][t] = (int)o.size();
            c->blist_tiles[k][t] = (int)v.size();
            return 0;
        };
        const int ids1[4] = {SG_PP3, SG_PH3, SG_PPL3, SG_PHL3};
        for (int k = 0; k < 4; ++k) rc |= upload_list(k, c->sg[ids1[k]].tiles[t]);
        for (int left = 1; left >= 0; --left) {
            // slots (p, a, t) of the batched launches
            const int ids3[3] = {left ? SG_PPL3 : SG_PP3, left ? SG_PHL3 : SG_PH3, left ? SG_PHL3 : SG_PH3};
            std::vector<int> all;
            for (int sl = 0; sl < 3; ++sl)
                for (int b : c->sg[ids3[sl]].tiles[t]) all.push_back(b + sl * (int)NB * tiles);
            rc |= upload_list(left ? 4 : 5, all);
        }
    }
    // k_k3_finish chunks: 256 elements of one (Ω, P) block.  Local solver: the blocks Ω and -Ω on
    // one XCD; lattice solver (representatives in other momentum blocks too): blocks in order
    if (nl) {
        const int nchunk = cdiv((size_t)NV * NV, 256);
        std::vector<int> items((size_t)NB * nchunk);
        for (size_t i = 0; i < items.size(); ++i) items[i] = (int)i;
        const std::vector<int> o = xcd_order(items, [nchunk](int it) { return it / nchunk; });
        HIPCHK(c, hipMalloc((void**)&c->finmap, o.size() * sizeof(int)));
        HIPCHK(c, hipMemcpy(c->finmap, o.data(), o.size() * sizeof(int), hipMemcpyHostToDevice));
        c->finmap_n = (int)o.size();
        c->fin_nchunk = nchunk;
    } else {
        const int nchunk = cdiv((size_t)NV * NV, 256);
        // per pair (Ω, -Ω) the chunks of both blocks alternate, so that an XCD works on both
        // blocks of a pair at the same time
        std::vector<int> items;
        const int n3b0 = d.nK3b - 1;
        for (int m = 0; m <= n3b0; ++m)
            for (int ch = 0; ch < nchunk; ++ch) {
                items.push_back((n3b0 - m) * nchunk + ch);
                if (m > 0) items.push_back((n3b0 + m) * nchunk + ch);
            }
        const int n3b = d.nK3b;
        const std::vector<int> o = xcd_order(items, [nchunk, n3b](int it) { return std::abs(it / nchunk - (n3b - 1)); });
        HIPCHK(c, hipMalloc((void**)&c->finmap, o.size() * sizeof(int)));
        HIPCHK(c, hipMemcpy(c->finmap, o.data(), o.size() * sizeof(int), hipMemcpyHostToDevice));
        c->finmap_n = (int)o.size();
        c->fin_nchunk = nchunk;
    }
    c->sde0_valid = 0;
    return rc ? -2 : 0;
}

int fdga_reset_sym_grp(fdga_ctx* c) {
    if (!c) return -1;
    HIPCHK(c, hipStreamSynchronize(c->st));
    graphs_invalidate(c);
    // SGΣ, SGpp, SGph, SGppL, SGphL become identity groups; SG0pp2 / SG0ph2 are left as they are
    for (int id : {SG_S, SG_K1, SG_PP2, SG_PH2, SG_PP3, SG_PH3, SG_PPL3, SG_PHL3}) free_sym(c->sg[id]);
    free_blists(c);
    if (c->k3_pl) (void)hipFree(c->k3_pl);
    c->k3_pl = nullptr;
    c->k3_npl = 0;
    c->sde0_valid = 0;
    return 0;
}

int fdga_dyson(fdga_ctx* c) {
    c->pi_band_G = 0;
    k_dyson<<<cdiv(c->nGel, 256), 256, 0, c->run>>>(c->G, c->Sig, c->Gbare, c->nGel);
    LAUNCHCHK(c);
    return 0;
}

int fdga_bubbles(fdga_ctx* c) { return bubbles_into(c, c->Pipp, c->Piph, c->G, true); }

int fdga_build_k3_cache(fdga_ctx* c) {
    c->r32_fresh = 0;
    c->caches_mfrg = 0;
    CachePtrs cp;
    for (int k = 0; k < 12; ++k) cp.c[k] = c->cache[k];
    const dim3 g(cdiv(c->NV3, 16), cdiv(c->NV3, 4), c->NW3);
    k_cache1<<<g, 64, 0, c->run>>>(cp, vtx_F(c), c->d.nK3b, c->d.nK3f);
    k_cache2<<<g, 64, 0, c->run>>>(cp, vtx_F(c), c->d.nK3b, c->d.nK3f, c->NP == 1);
    LAUNCHCHK(c);
    return 0;
}


int fdga_bse_l_k2(fdga_ctx* c, int32_t ch) {
    if (ch == PCH) return launch_lk2<PCH>(c);
    if (ch == TCH) return launch_lk2<TCH>(c);
    if (ch == ACH) return launch_lk2<ACH>(c);
    return fail(c, "bad channel");
}

int fdga_bse_l_k3(fdga_ctx* c, int32_t ch) {
    if (ch < 0 || ch > 2) return fail(c, "bad channel");
    const int chans[1] = {ch};
    return launch_lk3(c, chans, 1);
}

int fdga_bse_k1(fdga_ctx* c, int32_t ch, int32_t is_mfrg) {
    if (is_mfrg) return launch_form_ch<FORM_K1_MFRG>(c, ch, 1);
    if (ch == PCH) return launch_k1<PCH>(c);
    if (ch == TCH) return launch_k1<TCH>(c);
    if (ch == ACH) return launch_k1<ACH>(c);
    return fail(c, "bad channel");
}

int fdga_bse_k2(fdga_ctx* c, int32_t ch, int32_t is_mfrg) {
    if (is_mfrg) return launch_form_ch<FORM_K2_MFRG>(c, ch, 1);
    if (ch == PCH) return launch_k2<PCH>(c);
    if (ch == TCH) return launch_k2<TCH>(c);
    if (ch == ACH) return launch_k2<ACH>(c);
    return fail(c, "bad channel");
}

int fdga_bse_k3(fdga_ctx* c, int32_t ch, int32_t is_mfrg) {
    if (ch < 0 || ch > 2) return fail(c, "bad channel");
    const int chans[1] = {ch};
    return launch_k3(c, chans, 1, is_mfrg != 0);
}

int fdga_bse_k1_new(fdga_ctx* c, int32_t ch, int32_t is_mfrg) { return launch_form_ch<FORM_K1_NEW>(c, ch, is_mfrg); }

int fdga_bse_k2_new(fdga_ctx* c, int32_t ch, int32_t is_mfrg) { return launch_form_ch<FORM_K2_NEW>(c, ch, is_mfrg); }

int fdga_build_k3_cache_mfrg(fdga_ctx* c, int32_t is_first) {
    if (!c) return -1;
    c->r32_fresh = 0;
    if (!c->d.nonlocal)
        return fail(c, "build_K3_cache_mfRG! has no method for ParquetSolver (src/nonlocal_2/build_K3_cache.jl:97-99 "
                       "defines it for NL_ParquetSolver and NL2_ParquetSolver)");
    if (c->sg[SG_PP3].ncls || c->sg[SG_PH3].ncls)
        return fail(c, "build_K3_cache_mfRG! under the lattice symmetry groups is not supported: the class fill would "
                       "make the caches depend on the momentum");
    return build_k3_cache_mfrg(c, is_first != 0);
}

}  // extern "C"

namespace {
// host side of set!(S.F, S.Fbuff): exchange the buffers (an involution)
void swap_F_host(fdga_ctx* c) {
    for (int r = 0; r < 3; ++r) {
        std::swap(c->F[r].K1, c->Fb[r].K1);
        std::swap(c->F[r].K2, c->Fb[r].K2);
        std::swap(c->F[r].K3, c->Fb[r].K3);
        if (!c->F[r].own_avg) { c->F[r].A1 = c->F[r].K1; c->F[r].A2 = c->F[r].K2; c->F[r].A3 = c->F[r].K3; }
        c->Fb[r].A1 = c->Fb[r].K1; c->Fb[r].A2 = c->Fb[r].K2; c->Fb[r].A3 = c->Fb[r].K3;
    }
    c->swap_state ^= 1;
}
}  // namespace

// drop the captured iterations (their kernel arguments no longer describe the context)
void graphs_invalidate(fdga_ctx* c) {
    for (auto& row : c->gexec)
        for (auto& g : row)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
}

extern "C" {

int fdga_update_F(fdga_ctx* c) {
    // Fbuff is overwritten in full by every iteration, so set!(S.F, S.Fbuff) is a buffer swap
    swap_F_host(c);
    return refresh_avg(c);
}

int fdga_sde(fdga_ctx* c, int32_t strategy, int32_t include_hartree) {
    // the _new strategies update Σ as their base strategy (src/SDE.jl:4-28)
    if (strategy == FDGA_FDPA_NEW) strategy = FDGA_FDPA;
    if (strategy == FDGA_SCPA_NEW) strategy = FDGA_SCPA;
    if (strategy != FDGA_SCPA && strategy != FDGA_FDPA) return fail(c, "Calculation strategy unknown");
    // fdPA (src/SDE.jl:7-31): Σ = SDE(F, G) - SDE(F0, G0, Π0) + Σ0 [- (n0 - 1/2) U i].  Everything
    // but SDE(F, G) is the same every iteration: formed once into sde0, before the target levels so
    // that Lpp/Lph end up holding the target's (S.Lpp, S.Lph)
    if (strategy == FDGA_FDPA && (!c->sde0_valid || c->sde0_h != include_hartree)) {
        k_zero<<<cdiv(c->nGel, 256), 256, 0, c->run>>>(c->sde0, c->nGel);
        if (sde_levels(c, c->sde0, c->G0, c->Pi0pp, c->Pi0ph, 1, include_hartree)) return -3;
        k_axpby<<<cdiv(c->nGel, 256), 256, 0, c->run>>>(c->sde0, c->Sig0, 1.0, -1.0, c->nGel);   // Σ0 - SDE(F0, ..)
        LAUNCHCHK(c);
        if (include_hartree && c->d.fdpa_h0 && sde_hartree(c, c->sde0, c->G0, -1.0)) return -3;
        c->sde0_valid = 1;
        c->sde0_h = include_hartree;
    }
    // the target levels accumulate onto the constant part (fdPA) or onto zero (scPA)
    if (strategy == FDGA_FDPA) k_copy<<<cdiv(c->nGel, 256), 256, 0, c->run>>>(c->Sig, c->sde0, c->nGel);
    else k_zero<<<cdiv(c->nGel, 256), 256, 0, c->run>>>(c->Sig, c->nGel);
    LAUNCHCHK(c);
    return sde_levels(c, c->Sig, c->G, c->Pipp, c->Piph, 0, include_hartree) ? -3 : 0;
}

static int iterate_eager(fdga_ctx* c, int32_t strategy, int32_t update_sigma, int32_t hartree);

// One iteration with its two independent chains on two streams.  After Dyson and the bubbles,
// chain A (caches -> BSE_L_K3! -> BSE_K3!) runs on st and chain B (BSE_L_K2! -> weights ->
// BSE_K1! -> BSE_K2!) on st2; they write disjoint arrays (FL.K3, Fbuff.K3 vs FL.K2, Fbuff.K1/K2)
// and B reads FL only at ν' = ∞ (K1 and K2 classes).  The join precedes set!(S.F, S.Fbuff) and SDE!.
static int iterate_two_chains(fdga_ctx* c, int32_t strategy, int32_t update_sigma, int32_t hartree) {
    if (update_sigma && (fdga_dyson(c) || fdga_bubbles(c))) return -3;
    HIPCHK(c, hipEventRecord(c->evfork, c->st));
    HIPCHK(c, hipStreamWaitEvent(c->st2, c->evfork, 0));
    const int all[3] = {PCH, ACH, TCH};
    int rc = 0;
    // chain B on st2
    c->run = c->st2;
    if (strategy == FDGA_FDPA) rc = launch_lk2_pa(c) || fdga_bse_l_k2(c, TCH);
    rc = rc ? rc : launch_w_all(c);
    c->w_fresh = 1;
    rc = rc ? rc : (launch_k1_pa(c) || fdga_bse_k1(c, TCH, 0));
    rc = rc ? rc : launch_k2_all(c);
    c->w_fresh = 0;
    c->run = c->st;
    if (rc) return -3;
    // chain A on st
    if (fdga_build_k3_cache(c)) return -3;
    if (strategy == FDGA_FDPA && launch_lk3(c, all, 3)) return -3;
    if (launch_k3(c, all, 3)) return -3;
    // join
    HIPCHK(c, hipEventRecord(c->evjoin, c->st2));
    HIPCHK(c, hipStreamWaitEvent(c->st, c->evjoin, 0));
    if (fdga_update_F(c)) return -3;
    if (update_sigma && fdga_sde(c, strategy, hartree)) return -3;
    return 0;
}

// :scPA_new / :fdPA_new (src/solve.jl:26-46): caches, [BSE_L_K3!], BSE_K3!, then the alternative K1 and
// K2 forms, which read only F, F0, Π and Π0
static int iterate_new(fdga_ctx* c, int32_t strategy, int32_t update_sigma, int32_t hartree) {
    const int all[3] = {PCH, ACH, TCH};
    stage_begin(c);
    if (update_sigma && (fdga_dyson(c) || fdga_bubbles(c))) return -3;
    stage_end(c, 0);
    stage_begin(c);
    if (fdga_build_k3_cache(c)) return -3;
    stage_end(c, 1);
    if (strategy == FDGA_FDPA_NEW) {
        stage_begin(c);
        if (launch_lk3(c, all, 3)) return -3;
        stage_end(c, 3);
    }
    stage_begin(c);
    if (launch_k3(c, all, 3)) return -3;
    stage_end(c, 6);
    stage_begin(c);
    for (int ch : all)
        if (launch_form_ch<FORM_K1_NEW>(c, ch, 0)) return -3;
    stage_end(c, 4);
    stage_begin(c);
    for (int ch : all)
        if (launch_form_ch<FORM_K2_NEW>(c, ch, 0)) return -3;
    stage_end(c, 5);
    stage_begin(c);
    if (fdga_update_F(c)) return -3;
    stage_end(c, 7);
    stage_begin(c);
    if (update_sigma && fdga_sde(c, strategy, hartree)) return -3;
    stage_end(c, 8);
    return 0;
}

// Captures fdga_iterate(strategy, 1, 1) for both F/Fbuff pointer states.  Capture records the launches
// without running them, so the host-side buffer swap done while recording is undone after each.
static int capture_iteration(fdga_ctx* c, int strategy) {
    for (int k = 0; k < 2; ++k) {
        const int state = c->swap_state;
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        const int rc = iterate_eager(c, strategy, 1, 1);
        const hipError_t ec = hipStreamEndCapture(c->st, &g);
        if (rc || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            return fail(c, "hipGraph capture of the iteration failed", -3);
        }
        hipGraphExec_t ex = nullptr;
        const hipError_t ei = hipGraphInstantiateWithFlags(&ex, g, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) return fail(c, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei), -3);
        c->gexec[strategy][state] = ex;
        swap_F_host(c);   // undo the swap recorded by update_F ...
        swap_F_host(c);   // ... and move to the other pointer state for the next capture
    }
    return 0;
}

int fdga_iterate(fdga_ctx* c, int32_t strategy, int32_t update_sigma, int32_t compute_hartree) {
    if (!c) return -1;
    if (strategy < FDGA_SCPA || strategy > FDGA_FDPA_NEW) return fail(c, "Calculation strategy unknown");
    const bool fd = (strategy == FDGA_FDPA || strategy == FDGA_FDPA_NEW);
    c->caches_mfrg = 0;
    if (fl_prepare(c)) return -3;   // eagerly: a captured iteration finds FL.K3 and F.K3 in their form
    // graph replay: the fixed-point iteration with Σ update and the Hartree term, stage timing off, and
    // for the fdPA strategies once the reference system's SDE is cached (the first such iteration
    // computes it eagerly)
    if (c->graph_on && update_sigma && compute_hartree && !c->timing && (!fd || (c->sde0_valid && c->sde0_h == 1))) {
        if (!c->gexec[strategy][c->swap_state] && capture_iteration(c, strategy)) return -3;
        HIPCHK(c, hipGraphLaunch(c->gexec[strategy][c->swap_state], c->st));
        swap_F_host(c);
        if (c->d.nonlocal) c->pi_band = 3;   // the replayed Dyson and band-stored bubbles
        c->pi_band_G = 1;
        return 0;
    }
    return iterate_eager(c, strategy, update_sigma, compute_hartree);
}

static int iterate_eager(fdga_ctx* c, int32_t strategy, int32_t update_sigma, int32_t hartree) {
    if (strategy == FDGA_SCPA_NEW || strategy == FDGA_FDPA_NEW) return iterate_new(c, strategy, update_sigma, hartree);
    if (!c->timing) return iterate_two_chains(c, strategy, update_sigma, hartree);
    stage_begin(c);
    if (update_sigma) {
        if (fdga_dyson(c) || fdga_bubbles(c)) return -3;
    }
    stage_end(c, 0);
    stage_begin(c);
    if (fdga_build_k3_cache(c)) return -3;
    stage_end(c, 1);
    if (strategy == FDGA_FDPA) {
        stage_begin(c);
        if (launch_lk2_pa(c) || fdga_bse_l_k2(c, TCH)) return -3;
        stage_end(c, 2);
        stage_begin(c);
        const int all[3] = {PCH, ACH, TCH};
        if (launch_lk3(c, all, 3)) return -3;
        stage_end(c, 3);
    }
    stage_begin(c);
    // the K1/K2 weights depend on Π, Π0, FL and F0 only: computed once for both stages
    if (launch_w_all(c)) return -3;
    c->w_fresh = 1;
    int rc = launch_k1_pa(c) || fdga_bse_k1(c, TCH, 0);
    stage_end(c, 4);
    stage_begin(c);
    rc = rc ? rc : launch_k2_all(c);
    c->w_fresh = 0;
    if (rc) return -3;
    stage_end(c, 5);
    stage_begin(c);
    {
        const int all[3] = {PCH, ACH, TCH};
        if (launch_k3(c, all, 3)) return -3;
    }
    stage_end(c, 6);
    stage_begin(c);
    if (fdga_update_F(c)) return -3;
    stage_end(c, 7);
    stage_begin(c);
    if (update_sigma && fdga_sde(c, strategy, hartree)) return -3;
    stage_end(c, 8);
    return 0;
}

int64_t fdga_flat_size(const fdga_ctx* c) {
    return (int64_t)(3 * (c->n1 + c->n2 + c->n3) * c->NP + c->nGel);
}
int64_t fdga_flat_size_F(const fdga_ctx* c) { return (int64_t)(3 * (c->n1 + c->n2 + c->n3) * c->NP); }

}  // extern "C"

namespace {
// x (host, reference layout [γp; γt; γa] of K1, K2, K3, then Σ when `sigma`) into the device arrays;
// the vertex part is multiplied by `scale`
// `buf`: the staging buffer (nullptr: the context's, allocated on first use)
int unflatten_into(fdga_ctx* c, const double* x, bool sigma, double scale, double2* buf = nullptr) {
    const size_t n = (size_t)(sigma ? fdga_flat_size(c) : fdga_flat_size_F(c));
    if (!buf) {
        if (ensure_stage(c, n)) return -2;
        buf = c->stage;
    }
    HIPCHK(c, hipMemcpyAsync(buf, x, n * sizeof(double2), hipMemcpyHostToDevice, c->st));
    if (scale != 1.0) {
        const size_t nf = (size_t)fdga_flat_size_F(c);
        k_axpby<<<cdiv(nf, 256), 256, 0, c->run>>>(buf, buf, scale, 0.0, nf);
        LAUNCHCHK(c);
    }
    size_t off = 0;
    for (int r = 0; r < 3; ++r) {
        HIPCHK(c, hipMemcpyAsync(c->F[r].K1, buf + off, c->n1 * c->NP * sizeof(double2), hipMemcpyDeviceToDevice, c->st));
        off += c->n1 * c->NP;
        HIPCHK(c, hipMemcpyAsync(c->F[r].K2, buf + off, c->n2 * c->NP * sizeof(double2), hipMemcpyDeviceToDevice, c->st));
        off += c->n2 * c->NP;
        if (c->fk32)
            k_k3_to_dev32<<<cdiv(c->n3 * c->NP, 256), 256, 0, c->run>>>(reinterpret_cast<float2*>(c->F[r].K3), buf + off,
                                                                          c->NW3, c->NV3, c->NP);
        else k_k3_to_dev<<<cdiv(c->n3 * c->NP, 256), 256, 0, c->run>>>(c->F[r].K3, buf + off, c->NW3, c->NV3, c->NP);
        LAUNCHCHK(c);
        off += c->n3 * c->NP;
    }
    if (sigma) HIPCHK(c, hipMemcpyAsync(c->Sig, buf + off, c->nGel * sizeof(double2), hipMemcpyDeviceToDevice, c->st));
    if (refresh_avg(c)) return -3;
    return 0;
}
// the device arrays into x (host, the same layout); synchronises the stream
int flatten_from(fdga_ctx* c, double* x, bool sigma, double2* buf = nullptr) {
    const size_t n = (size_t)(sigma ? fdga_flat_size(c) : fdga_flat_size_F(c));
    if (!buf) {
        if (ensure_stage(c, n)) return -2;
        buf = c->stage;
    }
    size_t off = 0;
    for (int r = 0; r < 3; ++r) {
        HIPCHK(c, hipMemcpyAsync(buf + off, c->F[r].K1, c->n1 * c->NP * sizeof(double2), hipMemcpyDeviceToDevice, c->st));
        off += c->n1 * c->NP;
        HIPCHK(c, hipMemcpyAsync(buf + off, c->F[r].K2, c->n2 * c->NP * sizeof(double2), hipMemcpyDeviceToDevice, c->st));
        off += c->n2 * c->NP;
        if (c->fk32)
            k_k3_to_ref32<<<cdiv(c->n3 * c->NP, 256), 256, 0, c->run>>>(buf + off, reinterpret_cast<const float2*>(c->F[r].K3),
                                                                          c->NW3, c->NV3, c->NP);
        else k_k3_to_ref<<<cdiv(c->n3 * c->NP, 256), 256, 0, c->run>>>(buf + off, c->F[r].K3, c->NW3, c->NV3, c->NP);
        LAUNCHCHK(c);
        off += c->n3 * c->NP;
    }
    if (sigma) HIPCHK(c, hipMemcpyAsync(buf + off, c->Sig, c->nGel * sizeof(double2), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(x, buf, n * sizeof(double2), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}
}  // namespace

extern "C" {

int fdga_unflatten(fdga_ctx* c, const double* x) { return unflatten_into(c, x, true, 1.0); }
int fdga_flatten(fdga_ctx* c, double* x) { return flatten_from(c, x, true); }

int fdga_fixed_point(fdga_ctx* c, double* R, const double* x, int32_t strategy, int32_t update_sigma,
                     int32_t compute_hartree) {
    if (!c || !R || !x) return -1;
    if (unflatten_into(c, x, update_sigma != 0, 1.0)) return -3;
    if (fdga_iterate(c, strategy, update_sigma, compute_hartree)) return -3;
    if (flatten_from(c, R, update_sigma != 0)) return -3;
    const size_t n = 2 * (size_t)(update_sigma ? fdga_flat_size(c) : fdga_flat_size_F(c));
    for (size_t i = 0; i < n; ++i) R[i] -= x[i];
    return 0;
}

int fdga_fixed_point_noalloc(fdga_ctx* c, double* R, const double* x, int32_t strategy, int32_t update_sigma,
                             int32_t compute_hartree) {
    if (!c || !R || !x) return -1;
    if (!c->scratch) return fail(c, "no scratch address for this device");
    const size_t n = (size_t)(update_sigma ? fdga_flat_size(c) : fdga_flat_size_F(c));
    if (n * sizeof(double2) > kScratchBytes)
        return fail(c, "fdga_fixed_point_noalloc: the state vector (" + std::to_string(n * sizeof(double2)) +
                           " bytes) exceeds the scratch of " + std::to_string(kScratchBytes) +
                           " bytes; use fdga_fixed_point");
    if (strategy < FDGA_SCPA || strategy > FDGA_FDPA_NEW) return fail(c, "Calculation strategy unknown");
    if (unflatten_into(c, x, update_sigma != 0, 1.0, c->scratch)) return -3;
    // always the eager launches: capturing and instantiating a hipGraph allocates
    c->caches_mfrg = 0;
    if (fl_prepare(c) || iterate_eager(c, strategy, update_sigma, compute_hartree)) return -3;
    if (flatten_from(c, R, update_sigma != 0, c->scratch)) return -3;
    for (size_t i = 0; i < 2 * n; ++i) R[i] -= x[i];
    return 0;
}

int fdga_mfrg_apply(fdga_ctx* c, double* y, const double* x, int32_t strategy, int32_t is_first) {
    if (!c || !y || !x) return -1;
    if (strategy != FDGA_FDPA && strategy != FDGA_FDPA_NEW)
        return fail(c, "Invalid strategy. Must be fdPA or fdPA_new (fdPA_1loop is not on the device path).");
    if (!c->d.nonlocal)
        return fail(c, "build_K3_cache_mfRG! has no method for ParquetSolver (src/nonlocal_2/build_K3_cache.jl:97-99 "
                       "defines it for NL_ParquetSolver and NL2_ParquetSolver)");
    const double factor = 1e-2;   // src/mfRG.jl:39
    const int all[3] = {PCH, ACH, TCH};
    if (unflatten_into(c, x, false, factor)) return -3;
    if (fdga_build_k3_cache_mfrg(c, is_first)) return -3;
    if (strategy == FDGA_FDPA) {
        for (int ch : all)
            if (fdga_bse_l_k2(c, ch)) return -3;
        for (int ch : all)
            if (launch_form_ch<FORM_K1_MFRG>(c, ch, 1)) return -3;
        for (int ch : all)
            if (launch_form_ch<FORM_K2_MFRG>(c, ch, 1)) return -3;
    } else {
        for (int ch : all)
            if (launch_form_ch<FORM_K1_NEW>(c, ch, 1)) return -3;
        for (int ch : all)
            if (launch_form_ch<FORM_K2_NEW>(c, ch, 1)) return -3;
    }
    if (launch_lk3(c, all, 3)) return -3;
    if (launch_k3(c, all, 3, true)) return -3;
    if (fdga_update_F(c)) return -3;
    if (flatten_from(c, y, false)) return -3;
    const size_t n = 2 * (size_t)fdga_flat_size_F(c);
    for (size_t i = 0; i < n; ++i) y[i] = x[i] - y[i] / factor;
    return 0;
}

int fdga_sync(fdga_ctx* c) {
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int fdga_time_iterations(fdga_ctx* c, int32_t strategy, int32_t n, double* ms) {
    if (!c || !ms || n < 1) return -1;
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipEventRecord(a, c->st));
    for (int i = 0; i < n; ++i)
        if (fdga_iterate(c, strategy, 1, 1)) return -3;
    HIPCHK(c, hipEventRecord(b, c->st));
    HIPCHK(c, hipEventSynchronize(b));
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, a, b));
    *ms = t;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return 0;
}

int fdga_stage_ms(const fdga_ctx* c, int32_t stage, double* ms) {
    if (!c || !ms || stage < 0 || stage > 11) return -1;
    *ms = c->stage_ms[stage];
    return 0;
}

// C[b] = A[b] B[b] for nbatch dense column-major complex matrices through the BSE GEMM engine
// (tile < 0: the single-problem launch with 64 x 64 tiles)
static int selftest_zgemm(int32_t M, int32_t N, int32_t K, int32_t nbatch, int32_t precision, int32_t tile, int64_t grid,
                          const double* A, const double* B, double* C) {
    if (M < 1 || N < 1 || K < 1 || nbatch < 1 || !A || !B || !C || tile >= TILE_N || grid < 0 || grid % 8) return -1;
    double2 *dA = nullptr, *dB = nullptr, *dC = nullptr;
    const size_t na = (size_t)M * K, nb = (size_t)K * N, nc = (size_t)M * N;
    if (hipMalloc((void**)&dA, na * nbatch * sizeof(double2)) != hipSuccess) return -2;
    if (hipMalloc((void**)&dB, nb * nbatch * sizeof(double2)) != hipSuccess) { (void)hipFree(dA); return -2; }
    if (hipMalloc((void**)&dC, nc * nbatch * sizeof(double2)) != hipSuccess) { (void)hipFree(dA); (void)hipFree(dB); return -2; }
    int rc = 0;
    if (hipMemcpy(dA, A, na * nbatch * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dB, B, nb * nbatch * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) rc = -2;
    if (!rc) {
        const Dense_A a{dA, M, na};
        const Dense_B b{dB, K, nb};
        const Dense_C e{dC, M, nc};
        hipError_t err;
        if (tile < 0) {
            err = zgemm_launch(precision, nbatch, 0, M, N, K, cdiv(M, GBM), cdiv(N, GBN), nullptr, 0, a, b, e);
        } else {
            const GemmSchedOne sc = sched_one(M, N, K, cdiv(M, tile_bm(tile)), cdiv(N, tile_bn(tile)), nbatch, nullptr, 0);
            err = zgemm_run(precision, tile, grid ? (size_t)grid : (size_t)sc.nitems, 0, sc, a, b, e);
        }
        if (err != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = -3;
    }
    if (!rc && hipMemcpy(C, dC, nc * nbatch * sizeof(double2), hipMemcpyDeviceToHost) != hipSuccess) rc = -2;
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC);
    return rc;
}

int fdga_selftest_zgemm(int32_t M, int32_t N, int32_t K, int32_t precision, const double* A, const double* B, double* C) {
    return selftest_zgemm(M, N, K, 1, precision, -1, 0, A, B, C);
}
int fdga_selftest_zgemm_tiles(int32_t M, int32_t N, int32_t K, int32_t precision, int32_t tile, int64_t grid, const double* A,
                              const double* B, double* C) {
    if (tile < 0) return -1;
    return selftest_zgemm(M, N, K, 1, precision, tile, grid, A, B, C);
}
int fdga_selftest_zgemm_batched(int32_t M, int32_t N, int32_t K, int32_t nbatch, int32_t precision, int32_t tile, int64_t grid,
                                const double* A, const double* B, double* C) {
    if (tile < 0) return -1;
    return selftest_zgemm(M, N, K, nbatch, precision, tile, grid, A, B, C);
}

// GEMM blocks (GBM x GBN output tiles over the full contraction length) of the batched (p, a, t)
// BSE_L_K3! (left = 1) or BSE_K3! (left = 0) launch of one iteration
int64_t fdga_gemm_blocks(const fdga_ctx* c, int32_t left) {
    if (!c) return -1;
    const int chans[3] = {PCH, ACH, TCH};
    const int tile = c->ktile[left ? 0 : 1];
    size_t nl = 0, nt = (size_t)3 * c->NW3 * c->NP * cdiv(c->NV3, tile_bm(tile)) * cdiv(c->NV3, tile_bn(tile));
    (void)k3_blocks(c, chans, 3, left != 0, tile, nl, &nt);
    return (int64_t)nt;
}
// executed flops of that launch: 8 BM BN Nν3 per computed tile of the current tile shape
double fdga_gemm_flops(const fdga_ctx* c, int32_t left) {
    if (!c) return -1.0;
    const int tile = c->ktile[left ? 0 : 1];
    return (double)fdga_gemm_blocks(c, left) * 8.0 * tile_bm(tile) * tile_bn(tile) * c->NV3;
}
// tile shape of the BSE_L_K3! (left = 1) or BSE_K3! (left = 0) launches
int fdga_gemm_tile(const fdga_ctx* c, int32_t left) { return c ? c->ktile[left ? 0 : 1] : -1; }
int fdga_gemm_grid(const fdga_ctx* c, int32_t left) { return c ? c->kgrid[left ? 0 : 1] : -1; }
int fdga_set_gemm_grid(fdga_ctx* c, int32_t left, int32_t grid) {
    if (!c) return -1;
    if (grid < 0 || grid % 8) return fail(c, "the workgroup count must be 0 (one per work item) or a multiple of 8");
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->kgrid[left ? 0 : 1] = grid;
    graphs_invalidate(c);
    return 0;
}
int fdga_set_gemm_tile(fdga_ctx* c, int32_t left, int32_t tile) {
    if (!c) return -1;
    if (tile < 0 || tile >= TILE_N) return fail(c, "unknown tile shape");
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->ktile[left ? 0 : 1] = tile;
    graphs_invalidate(c);
    return 0;
}

// Times every tile shape of the two contraction launches (p, a, t batched) on the current operands,
// `reps` launches each, and keeps the fastest.  The caches and right operands are rebuilt first; the
// BSE_L_K3! timing writes into the right-operand scratch instead of FL, the BSE_K3! timing into
// Fbuff (overwritten by every iteration), so the solver state is left as it was.
int fdga_autotune(fdga_ctx* c, int32_t reps) {
    if (!c || reps < 1) return -1;
    const int all[3] = {PCH, ACH, TCH};
    c->run = c->st;
    c->r32_fresh = 0;   // the right-operand buffers serve as scratch outputs below
    if (fdga_build_k3_cache(c)) return -3;
    const bool symL = k3_sym(c, all, 3, true), symK = k3_sym(c, all, 3, false);
    LK3Set ls = lk3_set(c, all, 3, symL, false);
    ls.e.c0.out = ls.er.c0.out = c->R[0];   // scratch outputs
    ls.e.c1.out = ls.er.c1.out = c->R[1];
    ls.e.c2.out = ls.er.c2.out = c->R[2];
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    auto timeit = [&](auto&& launch) -> float {
        (void)launch();
        (void)hipEventRecord(a, c->st);
        for (int r = 0; r < reps; ++r) (void)launch();
        (void)hipEventRecord(b, c->st);
        (void)hipEventSynchronize(b);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        return ms;
    };
    int rc = 0;
    float best = 1e30f;
    int bt = c->ktile[0];
    for (int t = 0; t < TILE_N; ++t) {
        c->ktile[0] = t;
        const float ms = timeit([&] { return lk3_gemm(c, all, 3, ls); });
        if (ms < best) { best = ms; bt = t; }
    }
    c->ktile[0] = bt;
    const K3Set ks = k3_set(c, all, 3, symK, false);
    rc |= k3_rhs(c, ks, 3);
    best = 1e30f;
    bt = c->ktile[1];
    for (int t = 0; t < TILE_N; ++t) {
        c->ktile[1] = t;
        const float ms = timeit([&] { return k3_gemm_set(c, all, 3, ks); });
        if (ms < best) { best = ms; bt = t; }
    }
    c->ktile[1] = bt;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (hipGetLastError() != hipSuccess || rc) return fail(c, "autotune launch failed", -3);
    HIPCHK(c, hipStreamSynchronize(c->st));
    graphs_invalidate(c);
    return 0;
}

// replay fdga_iterate(strategy, 1) from captured hipGraphs (1) or launch eagerly (0)
int fdga_set_graph(fdga_ctx* c, int32_t on) {
    if (!c) return -1;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (!on) graphs_invalidate(c);
    c->graph_on = on ? 1 : 0;
    return 0;
}

// enable (1) / disable (0) per-stage event timing; resets the accumulated stage times
int fdga_set_stage_timing(fdga_ctx* c, int32_t on) {
    if (!c) return -1;
    c->timing = on;
    for (auto& s : c->stage_ms) s = 0.0;
    return 0;
}

}  // extern "C"

// =========================================================================================
// Groups of solver contexts (a parameter sweep on one device).  Each member runs its own stages
// on its own stream, concurrently with the other members; the BSE_L_K3! and BSE_K3! contractions
// of all members run as one grouped GEMM launch each on the group's stream (the (Ω, P, channel)
// blocks of every member as work items, with the member's own box sizes), persistent over a
// chosen grid and with a chosen tile shape.  Each member's results are those of fdga_iterate on
// the member alone.  No counterpart in the reference, which solves one parameter point per solver.
struct GroupItems {   // work items of one tile shape
    int* items = nullptr;     // lattice groups: the p and a channels' items; else all three channels'
    int* items_t = nullptr;   // lattice groups: the t channel's items (a second launch)
    GemmGroupGeom* geo = nullptr;
    int nitems = 0, nitems_t = 0;
    int64_t computed = 0;   // non-idle items (both launches)
};

struct fdga_group {
    std::vector<fdga_ctx*> m;
    int n = 0, precision = 0, nonlocal = 0;
    hipStream_t st = nullptr;
    std::vector<hipEvent_t> ev_m;    // member -> group joins
    hipEvent_t ev_g = nullptr;       // group -> members forks
    std::string err;
    int* boff = nullptr;             // first batch of each member in the launch's batch space
    std::vector<int> boff_h;
    GroupItems gi[TILE_N];
    // per-member functors on the device: BSE_L_K3! (independent of the F/Fbuff state) and BSE_K3!
    // for each F/Fbuff pointer state
    LK3_Aop* lk3a = nullptr;
    LK3_Bop* lk3b = nullptr;
    LK3_Ep* lk3e = nullptr;
    LK3R_Ep* lk3er = nullptr;   // used instead of lk3e when lk3_ro (every member writes its FP32 R there)
    int lk3_ro = 0;
    void* k3a[2] = {nullptr, nullptr};   // K3_AopT<true> (local), K3_AopT<false> or K3_Aop32 (lattice FP64 / FP32)
    void* k3b[2] = {nullptr, nullptr};   // K3_Bop or K3_Bop32
    K3_Ep* k3e[2] = {nullptr, nullptr};
    int tile[2] = {TILE_64x64, TILE_64x64};   // [0] BSE_L_K3!, [1] BSE_K3!
    size_t grid[2] = {0, 0};                  // workgroups; 0: one per work item
    int graph_on = 0;
    hipGraphExec_t gexec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    int timing = 0;
    hipEvent_t tev[2] = {nullptr, nullptr};
    double gemm_ms[2] = {0.0, 0.0};
};

namespace {

int gfail(fdga_group* g, const std::string& msg, int code = -1) {
    g->err = msg;
    return code;
}
#define GHIPCHK(g, call)                                                                   \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) return gfail((g), std::string(#call) + ": " + hipGetErrorString(e_), -2); \
    } while (0)

bool has_k3_groups(const fdga_ctx* c) {
    for (int k : {SG_PP3, SG_PH3, SG_PPL3, SG_PHL3})
        if (c->sg[k].ncls) return true;
    return false;
}

// work items of one tile shape: members in order of decreasing box (the larger items first), each
// member's batches in chunks of 8 with the tiles of a batch on one XCD (item index mod 8)
int group_build_items(fdga_group* g, int tile) {
    GroupItems& gi = g->gi[tile];
    if (gi.items) return 0;
    const int BM = tile_bm(tile), BN = tile_bn(tile);
    std::vector<GemmGroupGeom> geo((size_t)g->n);
    std::vector<int> order((size_t)g->n);
    for (int j = 0; j < g->n; ++j) order[(size_t)j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return g->m[(size_t)a]->NV3 > g->m[(size_t)b]->NV3; });
    std::vector<int> items, items_t;
    int64_t computed = 0;
    for (int j = 0; j < g->n; ++j) {
        const fdga_ctx* c = g->m[(size_t)j];
        const int NV = c->NV3;
        geo[(size_t)j] = GemmGroupGeom{NV, NV, NV, (int)cdiv(NV, BM), (int)cdiv(NV, BN), g->boff_h[(size_t)j]};
    }
    // batches [b0, b1) of member j in chunks of 8, the tiles of a batch on one XCD
    auto add = [&](std::vector<int>& v, int j, int tiles, int b0, int b1) {
        for (int c0 = b0; c0 < b1; c0 += 8)
            for (int t = 0; t < tiles; ++t)
                for (int x = 0; x < 8; ++x) {
                    const int b = c0 + x;
                    v.push_back(b < b1 ? (j << 24) | (b * tiles + t) : -1);
                    computed += b < b1;
                }
    };
    for (int j : order) {
        const GemmGroupGeom& q = geo[(size_t)j];
        const int tiles = q.tiles_m * q.tiles_n;
        const int nb = g->boff_h[(size_t)j + 1] - g->boff_h[(size_t)j];
        if ((int64_t)nb * tiles >= (1 << 24)) return gfail(g, "grouped GEMM: too many blocks in one member", -1);
        if (g->nonlocal) {   // slots p, a in the first launch, t (batches from 2 nb / 3 on) in the second
            add(items, j, tiles, 0, 2 * nb / 3);
            add(items_t, j, tiles, 2 * nb / 3, nb);
        } else {
            add(items, j, tiles, 0, nb);
        }
    }
    GHIPCHK(g, hipMalloc((void**)&gi.items, items.size() * sizeof(int)));
    GHIPCHK(g, hipMalloc((void**)&gi.geo, geo.size() * sizeof(GemmGroupGeom)));
    GHIPCHK(g, hipMemcpy(gi.items, items.data(), items.size() * sizeof(int), hipMemcpyHostToDevice));
    GHIPCHK(g, hipMemcpy(gi.geo, geo.data(), geo.size() * sizeof(GemmGroupGeom), hipMemcpyHostToDevice));
    if (!items_t.empty()) {
        GHIPCHK(g, hipMalloc((void**)&gi.items_t, items_t.size() * sizeof(int)));
        GHIPCHK(g, hipMemcpy(gi.items_t, items_t.data(), items_t.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    gi.nitems = (int)items.size();
    gi.nitems_t = (int)items_t.size();
    gi.computed = computed;
    return 0;
}

template <class T>
int group_upload(fdga_group* g, T** dst, const std::vector<T>& src) {
    GHIPCHK(g, hipMalloc((void**)dst, src.size() * sizeof(T)));
    GHIPCHK(g, hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

// per-member functors of the two contractions (the three channels p, a, t, no symmetry groups;
// lattice groups: the t channel post-processed in its epilogue, in a launch after p and a);
// BSE_K3! for both F/Fbuff pointer states
int group_build_functors(fdga_group* g) {
    const int all[3] = {PCH, ACH, TCH};
    std::vector<LK3_Aop> la;
    std::vector<LK3_Bop> lb;
    std::vector<LK3_Ep> le;
    std::vector<LK3R_Ep> ler;
    std::vector<K3_AopT<true>> af[2];
    std::vector<K3_AopT<false>> an[2];
    std::vector<K3_Aop32> a32[2];
    std::vector<K3_Bop> kb[2];
    std::vector<K3_Bop32> b32[2];
    std::vector<K3_Ep> ke[2];
    for (fdga_ctx* c : g->m) {
        const bool f32 = fl32_mode(c);   // the FL.K3 form of the member's iterations (fl_prepare)
        const LK3Set ls = lk3_set(c, all, 3, false, g->nonlocal, f32);
        la.push_back(ls.a);
        lb.push_back(ls.b);
        le.push_back(ls.e);
        ler.push_back(ls.er);
        g->lk3_ro = (c == g->m[0] ? 1 : g->lk3_ro) && ls.ro;
        for (int k = 0; k < 2; ++k) {
            const int s = c->swap_state;
            const K3Set ks = k3_set(c, all, 3, false, g->nonlocal, false, f32 ? 1 : 0);
            af[s].push_back(ks.a<true>());
            an[s].push_back(ks.a<false>());
            a32[s].push_back(ks.a32());
            kb[s].push_back(ks.b());
            b32[s].push_back(ks.b32());
            ke[s].push_back(ks.e());
            swap_F_host(c);   // the other state; two swaps restore the context
        }
    }
    if (group_upload(g, &g->lk3a, la) || group_upload(g, &g->lk3b, lb) ||
        (g->lk3_ro ? group_upload(g, &g->lk3er, ler) : group_upload(g, &g->lk3e, le)))
        return -2;
    const bool op32 = g->nonlocal && g->precision == FDGA_FP32;
    for (int s = 0; s < 2; ++s) {
        int rc = op32 ? group_upload(g, (K3_Aop32**)&g->k3a[s], a32[s])
                 : g->nonlocal ? group_upload(g, (K3_AopT<false>**)&g->k3a[s], an[s])
                               : group_upload(g, (K3_AopT<true>**)&g->k3a[s], af[s]);
        rc = rc ? rc
                : g->precision == FDGA_FP32 ? group_upload(g, (K3_Bop32**)&g->k3b[s], b32[s])
                                            : group_upload(g, (K3_Bop**)&g->k3b[s], kb[s]);
        if (rc || group_upload(g, &g->k3e[s], ke[s])) return -2;
    }
    return 0;
}

void group_graphs_invalidate(fdga_group* g) {
    for (auto& row : g->gexec)
        for (auto& x : row)
            if (x) {
                (void)hipGraphExecDestroy(x);
                x = nullptr;
            }
}

size_t group_grid(const fdga_group* g, int which, bool t_launch = false) {
    const GroupItems& gi = g->gi[g->tile[which]];
    const size_t n = (size_t)(t_launch ? gi.nitems_t : gi.nitems);
    return g->grid[which] ? std::min(g->grid[which], n) : n;
}

// one launch of the grouped contraction `which` (0 BSE_L_K3!, 1 BSE_K3!) over the item list `sc`
hipError_t group_gemm_launch(fdga_group* g, int which, const GemmSchedGroup& sc, size_t grid) {
    const int tile = g->tile[which];
    if (which == 0 && g->lk3_ro)
        return zgemm_run(g->precision, tile, grid, g->st, sc, GroupOp<LK3_Aop>{g->lk3a, g->boff, g->n},
                         GroupOp<LK3_Bop>{g->lk3b, g->boff, g->n}, GroupOp<LK3R_Ep>{g->lk3er, g->boff, g->n});
    if (which == 0)
        return zgemm_run(g->precision, tile, grid, g->st, sc, GroupOp<LK3_Aop>{g->lk3a, g->boff, g->n},
                         GroupOp<LK3_Bop>{g->lk3b, g->boff, g->n}, GroupOp<LK3_Ep>{g->lk3e, g->boff, g->n});
    const int s = g->m[0]->swap_state;
    const GroupOp<K3_Bop> b{(const K3_Bop*)g->k3b[s], g->boff, g->n};
    const GroupOp<K3_Ep> ep{g->k3e[s], g->boff, g->n};
    if (g->nonlocal && g->precision == FDGA_FP32)
        return zgemm_run(g->precision, tile, grid, g->st, sc, GroupOp<K3_Aop32>{(const K3_Aop32*)g->k3a[s], g->boff, g->n},
                         GroupOp<K3_Bop32>{(const K3_Bop32*)g->k3b[s], g->boff, g->n}, ep);
    if (g->nonlocal)
        return zgemm_run(g->precision, tile, grid, g->st, sc, GroupOp<K3_AopT<false>>{(K3_AopT<false>*)g->k3a[s], g->boff, g->n},
                         b, ep);
    if (g->precision == FDGA_FP32)
        return zgemm_run(g->precision, tile, grid, g->st, sc, GroupOp<K3_AopT<true>>{(K3_AopT<true>*)g->k3a[s], g->boff, g->n},
                         GroupOp<K3_Bop32>{(const K3_Bop32*)g->k3b[s], g->boff, g->n}, ep);
    return zgemm_run(g->precision, tile, grid, g->st, sc, GroupOp<K3_AopT<true>>{(K3_AopT<true>*)g->k3a[s], g->boff, g->n},
                     b, ep);
}

// the grouped contraction `which` (0 BSE_L_K3!, 1 BSE_K3!) on the group's stream: one launch, or for
// lattice groups the p and a channels and then the t channel
int group_gemm(fdga_group* g, int which) {
    const GroupItems& gi = g->gi[g->tile[which]];
    if (which == 1) {
        const int s = g->m[0]->swap_state;
        for (const fdga_ctx* c : g->m)
            if (c->swap_state != s) return gfail(g, "group members are not at the same iteration (F/Fbuff state)");
    }
    if (g->timing) GHIPCHK(g, hipEventRecord(g->tev[0], g->st));
    hipError_t e = group_gemm_launch(g, which, GemmSchedGroup{gi.items, gi.geo, gi.nitems}, group_grid(g, which));
    if (e == hipSuccess && gi.nitems_t)
        e = group_gemm_launch(g, which, GemmSchedGroup{gi.items_t, gi.geo, gi.nitems_t}, group_grid(g, which, true));
    if (e != hipSuccess) return gfail(g, std::string("grouped GEMM launch: ") + hipGetErrorString(e), -3);
    if (g->timing) {
        GHIPCHK(g, hipEventRecord(g->tev[1], g->st));
        GHIPCHK(g, hipEventSynchronize(g->tev[1]));
        float ms = 0.f;
        GHIPCHK(g, hipEventElapsedTime(&ms, g->tev[0], g->tev[1]));
        g->gemm_ms[which] += ms;
    }
    return 0;
}

// group stream -> every member stream
int group_fork(fdga_group* g) {
    GHIPCHK(g, hipEventRecord(g->ev_g, g->st));
    for (fdga_ctx* c : g->m) GHIPCHK(g, hipStreamWaitEvent(c->st, g->ev_g, 0));
    return 0;
}
// every member stream -> group stream
int group_join(fdga_group* g) {
    for (int j = 0; j < g->n; ++j) {
        GHIPCHK(g, hipEventRecord(g->ev_m[(size_t)j], g->m[(size_t)j]->st));
        GHIPCHK(g, hipStreamWaitEvent(g->st, g->ev_m[(size_t)j], 0));
    }
    return 0;
}
int member_err(fdga_group* g, const fdga_ctx* c) { return gfail(g, std::string("member: ") + c->err, -3); }

// one iterate_solver! of every member (src/solve.jl:4-116), the contractions grouped
int group_iterate_eager(fdga_group* g, int strategy, int update_sigma) {
    const int all[3] = {PCH, ACH, TCH};
    for (fdga_ctx* c : g->m) {
        if (has_k3_groups(c)) return gfail(g, "group members must not have K3 symmetry groups installed");
        c->run = c->st;
    }
    if (group_fork(g)) return -2;
    for (fdga_ctx* c : g->m) {
        if (update_sigma && (fdga_dyson(c) || fdga_bubbles(c))) return member_err(g, c);
        if (fdga_build_k3_cache(c)) return member_err(g, c);
        if (strategy == FDGA_FDPA)
            for (int ch : {PCH, ACH, TCH})
                if (fdga_bse_l_k2(c, ch)) return member_err(g, c);
    }
    if (strategy == FDGA_FDPA) {
        if (group_join(g) || group_gemm(g, 0) || group_fork(g)) return -3;
        for (fdga_ctx* c : g->m)
            if (k3_finish(c, c->FL, all, 3, true, false, g->nonlocal)) return member_err(g, c);
    }
    for (fdga_ctx* c : g->m) {
        if (launch_w_all(c)) return member_err(g, c);
        c->w_fresh = 1;
        int rc = 0;
        for (int ch : {PCH, ACH, TCH}) rc = rc ? rc : fdga_bse_k1(c, ch, 0);
        rc = rc ? rc : launch_k2_all(c);
        c->w_fresh = 0;
        if (rc) return member_err(g, c);
        // fdPA: the grouped BSE_L_K3! epilogue wrote the FP32 lattice right operands
        const bool skipR = strategy == FDGA_FDPA && g->lk3_ro;
        if (k3_rhs(c, k3_set(c, all, 3, false, false), 3, false, skipR)) return member_err(g, c);
    }
    if (group_join(g) || group_gemm(g, 1) || group_fork(g)) return -3;
    for (fdga_ctx* c : g->m) {
        if (k3_finish(c, c->Fb, all, 3, false, false, g->nonlocal)) return member_err(g, c);
        if (fdga_update_F(c)) return member_err(g, c);
        if (update_sigma && fdga_sde(c, strategy, 1)) return member_err(g, c);
    }
    return group_join(g);
}

int group_capture(fdga_group* g, int strategy) {
    for (int k = 0; k < 2; ++k) {
        const int state = g->m[0]->swap_state;
        hipGraph_t gr = nullptr;
        GHIPCHK(g, hipStreamBeginCapture(g->st, hipStreamCaptureModeThreadLocal));
        const int rc = group_iterate_eager(g, strategy, 1);
        const hipError_t ec = hipStreamEndCapture(g->st, &gr);
        if (rc || ec != hipSuccess) {
            if (gr) (void)hipGraphDestroy(gr);
            (void)hipGetLastError();
            return gfail(g, "hipGraph capture of the group iteration failed" + (g->err.empty() ? "" : ": " + g->err), -3);
        }
        hipGraphExec_t ex = nullptr;
        const hipError_t ei = hipGraphInstantiateWithFlags(&ex, gr, 0);
        (void)hipGraphDestroy(gr);
        if (ei != hipSuccess) return gfail(g, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei), -3);
        g->gexec[strategy][state] = ex;
        for (fdga_ctx* c : g->m) {   // undo the recorded swap, then move to the other state
            swap_F_host(c);
            swap_F_host(c);
        }
    }
    return 0;
}

}  // namespace

extern "C" {

int fdga_group_create(fdga_group** out, fdga_ctx* const* members, int32_t n) {
    if (!out) return -1;
    *out = nullptr;
    auto* g = new fdga_group();
    *out = g;
    if (!members || n < 1 || n > 127) return gfail(g, "a group needs 1 to 127 members");
    for (int j = 0; j < n; ++j) {
        fdga_ctx* c = members[j];
        if (!c) return gfail(g, "null member");
        for (fdga_ctx* o : g->m)
            if (o == c) return gfail(g, "a context appears twice in the group");
        c->caches_mfrg = 0;   // the group's functors take the plain build_K3_cache! layout
        g->m.push_back(c);
    }
    g->n = n;
    g->precision = g->m[0]->d.precision;
    g->nonlocal = g->m[0]->d.nonlocal;
    for (fdga_ctx* c : g->m) {
        if (c->d.precision != g->precision) return gfail(g, "group members must share the contraction precision");
        if (c->d.nonlocal != g->nonlocal) return gfail(g, "group members must be all local or all lattice solvers");
        if (has_k3_groups(c)) return gfail(g, "group members must not have K3 symmetry groups installed");
    }
    GHIPCHK(g, hipStreamCreate(&g->st));
    GHIPCHK(g, hipEventCreateWithFlags(&g->ev_g, hipEventDisableTiming));
    g->ev_m.assign((size_t)n, nullptr);
    for (auto& e : g->ev_m) GHIPCHK(g, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& e : g->tev) GHIPCHK(g, hipEventCreate(&e));
    g->boff_h.assign((size_t)n + 1, 0);
    for (int j = 0; j < n; ++j) {
        const fdga_ctx* c = g->m[(size_t)j];
        const int64_t nb = (int64_t)3 * c->NW3 * c->NP;
        if ((int64_t)g->boff_h[(size_t)j] + nb > (int64_t)1 << 30) return gfail(g, "grouped GEMM: too many batches");
        g->boff_h[(size_t)j + 1] = g->boff_h[(size_t)j] + (int)nb;
    }
    if (group_upload(g, &g->boff, g->boff_h)) return -2;
    for (fdga_ctx* c : g->m) GHIPCHK(g, hipStreamSynchronize(c->st));
    if (group_build_items(g, TILE_64x64) || group_build_functors(g)) return -2;
    return 0;
}

int fdga_group_destroy(fdga_group* g) {
    if (!g) return 0;
    if (g->st) (void)hipStreamSynchronize(g->st);
    group_graphs_invalidate(g);
    for (auto& gi : g->gi) {
        if (gi.items) (void)hipFree(gi.items);
        if (gi.items_t) (void)hipFree(gi.items_t);
        if (gi.geo) (void)hipFree(gi.geo);
    }
    for (void* p : {(void*)g->boff, (void*)g->lk3a, (void*)g->lk3b, (void*)g->lk3e, (void*)g->lk3er, g->k3a[0], g->k3a[1], g->k3b[0],
                    g->k3b[1], (void*)g->k3e[0], (void*)g->k3e[1]})
        if (p) (void)hipFree(p);
    for (auto e : g->ev_m)
        if (e) (void)hipEventDestroy(e);
    for (auto e : g->tev)
        if (e) (void)hipEventDestroy(e);
    if (g->ev_g) (void)hipEventDestroy(g->ev_g);
    if (g->st) (void)hipStreamDestroy(g->st);
    delete g;
    return 0;
}

const char* fdga_group_last_error(const fdga_group* g) { return g ? g->err.c_str() : "null group"; }

int fdga_group_iterate(fdga_group* g, int32_t strategy, int32_t update_sigma) {
    if (!g) return -1;
    if (strategy == FDGA_SCPA_NEW || strategy == FDGA_FDPA_NEW)
        return gfail(g, "the :scPA_new / :fdPA_new strategies are not on the solver-group path; iterate the members alone");
    if (strategy != FDGA_SCPA && strategy != FDGA_FDPA) return gfail(g, "Calculation strategy unknown");
    bool sde0 = true;
    for (fdga_ctx* c : g->m) {
        sde0 = sde0 && c->sde0_valid;
        c->run = c->st;
        if (fl_prepare(c)) return member_err(g, c);
    }
    if (g->graph_on && update_sigma && !g->timing && (strategy == FDGA_SCPA || sde0)) {
        const int s = g->m[0]->swap_state;
        for (const fdga_ctx* c : g->m)
            if (c->swap_state != s || has_k3_groups(c)) return gfail(g, "group members changed since the group was made");
        if (!g->gexec[strategy][s] && group_capture(g, strategy)) return -3;
        GHIPCHK(g, hipGraphLaunch(g->gexec[strategy][s], g->st));
        for (fdga_ctx* c : g->m) {
            swap_F_host(c);
            if (c->d.nonlocal) c->pi_band = 3;   // the replayed Dyson and band-stored bubbles
            c->pi_band_G = 1;
        }
        return 0;
    }
    return group_iterate_eager(g, strategy, update_sigma);
}

int fdga_group_sync(fdga_group* g) {
    if (!g) return -1;
    GHIPCHK(g, hipStreamSynchronize(g->st));
    for (fdga_ctx* c : g->m) GHIPCHK(g, hipStreamSynchronize(c->st));
    return 0;
}

int fdga_group_time_iterations(fdga_group* g, int32_t strategy, int32_t n, double* ms) {
    if (!g || !ms || n < 1) return -1;
    if (fdga_group_sync(g)) return -2;
    GHIPCHK(g, hipEventRecord(g->tev[0], g->st));
    for (int i = 0; i < n; ++i)
        if (fdga_group_iterate(g, strategy, 1)) return -3;
    GHIPCHK(g, hipEventRecord(g->tev[1], g->st));
    GHIPCHK(g, hipEventSynchronize(g->tev[1]));
    float t = 0.f;
    GHIPCHK(g, hipEventElapsedTime(&t, g->tev[0], g->tev[1]));
    *ms = t;
    return 0;
}

int fdga_group_set_graph(fdga_group* g, int32_t on) {
    if (!g) return -1;
    if (fdga_group_sync(g)) return -2;
    if (!on) group_graphs_invalidate(g);
    g->graph_on = on ? 1 : 0;
    return 0;
}

// event timing of the grouped contractions inside fdga_group_iterate (adds a host sync per
// contraction); resets the accumulated times
int fdga_group_set_timing(fdga_group* g, int32_t on) {
    if (!g) return -1;
    g->timing = on ? 1 : 0;
    g->gemm_ms[0] = g->gemm_ms[1] = 0.0;
    return 0;
}
int fdga_group_gemm_ms(const fdga_group* g, int32_t which, double* ms) {
    if (!g || !ms || which < 0 || which > 1) return -1;
    *ms = g->gemm_ms[which];
    return 0;
}

// tile shape and grid of contraction `which` (0 BSE_L_K3!, 1 BSE_K3!); grid 0 = one workgroup per
// work item; the number of computed output tiles of that shape in `computed`
int fdga_group_tiles(fdga_group* g, int32_t which, int32_t* tile, int64_t* grid, int64_t* computed) {
    if (!g || which < 0 || which > 1) return -1;
    if (tile) *tile = g->tile[which];
    if (grid) *grid = (int64_t)g->grid[which];
    if (computed) *computed = g->gi[g->tile[which]].computed;
    return 0;
}
int fdga_group_set_tiles(fdga_group* g, int32_t which, int32_t tile, int64_t grid) {
    if (!g || which < 0 || which > 1) return -1;
    if (tile < 0 || tile >= TILE_N) return gfail(g, "unknown tile shape");
    if (grid != 0 && (grid < 8 || grid % 8)) return gfail(g, "grid must be 0 or a positive multiple of 8");
    if (fdga_group_sync(g) || group_build_items(g, tile)) return -2;
    g->tile[which] = tile;
    g->grid[which] = (size_t)grid;
    group_graphs_invalidate(g);
    return 0;
}

// Times every tile shape and grid (one workgroup per item; 4, 8 or 16 workgroups per CU) of both
// contractions on the members' current operands and keeps the fastest.  The contractions'
// outputs are scratch until the next iteration recomputes them; F, Fbuff's inputs and Σ are left
// as they are.
int fdga_group_autotune(fdga_group* g, int32_t reps) {
    if (!g || reps < 1) return -1;
    const int all[3] = {PCH, ACH, TCH};
    for (int t = 0; t < TILE_N; ++t)
        if (group_build_items(g, t)) return -2;
    if (fdga_group_sync(g) || group_fork(g)) return -2;
    for (fdga_ctx* c : g->m) {
        c->run = c->st;
        // the operand forms the group's launches were made for (those of the members' iterations)
        if (fl_prepare(c) || fdga_build_k3_cache(c) || k3_rhs(c, k3_set(c, all, 3, false, false), 3))
            return member_err(g, c);
    }
    if (group_join(g) || fdga_group_sync(g)) return -2;
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t pr;
        if (hipGetDeviceProperties(&pr, dev) == hipSuccess && pr.multiProcessorCount > 0) ncu = pr.multiProcessorCount;
    }
    const int keep_timing = g->timing;
    g->timing = 0;
    for (int which = 0; which < 2; ++which) {
        float best = 1e30f;
        int bt = g->tile[which];
        size_t bg = g->grid[which];
        for (int t = 0; t < TILE_N; ++t)
            for (int per : {0, 4, 8, 16}) {
                const size_t grid = per ? (size_t)((ncu * per + 7) / 8 * 8) : 0;
                if (grid && grid >= (size_t)g->gi[t].nitems) continue;
                g->tile[which] = t;
                g->grid[which] = grid;
                if (group_gemm(g, which)) return -3;   // warm-up
                GHIPCHK(g, hipEventRecord(g->tev[0], g->st));
                for (int r = 0; r < reps; ++r)
                    if (group_gemm(g, which)) return -3;
                GHIPCHK(g, hipEventRecord(g->tev[1], g->st));
                GHIPCHK(g, hipEventSynchronize(g->tev[1]));
                float ms = 0.f;
                GHIPCHK(g, hipEventElapsedTime(&ms, g->tev[0], g->tev[1]));
                if (ms < best) { best = ms; bt = t; bg = grid; }
            }
        g->tile[which] = bt;
        g->grid[which] = bg;
    }
    g->timing = keep_timing;
    group_graphs_invalidate(g);
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Tangent and adjoint of build_K3_cache! followed by BSE_K3! of the three channels (fdga_adjoint.h).
// The tangent context `t` holds the tangent inputs (JVP: F, FL.K3, Πpp/Πph; output Fbuff.K3) or the
// cotangent (VJP: Fbuff.K3; outputs F, FL.K3, Πpp/Πph).  Everything runs on the solver's stream.
namespace {

adj::Boxes adj_boxes(const fdga_dims& d) { return adj::Boxes{d.nK1, d.nK2b, d.nK2f, d.nK3b, d.nK3f}; }

int adj_check(fdga_ctx* c, fdga_ctx* t) {
    const fdga_dims &a = c->d, &b = t->d;
    if (a.nonlocal != b.nonlocal || a.nG != b.nG || a.nK1 != b.nK1 || a.nK2b != b.nK2b || a.nK2f != b.nK2f ||
        a.nK3b != b.nK3b || a.nK3f != b.nK3f || a.mPinu != b.mPinu || a.L != b.L || a.LG != b.LG)
        return fail(c, "tangent context: its dimensions differ from the solver's");
    if (a.precision != FDGA_FP64 || b.precision != FDGA_FP64)
        return fail(c, "the tangent and adjoint of BSE_K3! run on FP64 contexts only");
    if (b.ref_kind != 0 || b.U != 0.0 || t->tabs_nonzero)
        return fail(c, "tangent context: needs U = 0 and no reference vertex (ref_kind 0, zero RefVertex tables)");
    if (has_k3_groups(c)) return fail(c, "the tangent and adjoint of BSE_K3! need the identity K3 symmetry groups");
    c->run = c->st;
    t->run = c->st;   // the tangent context's kernels on the solver's stream (restored by adj_done)
    HIPCHK(c, hipStreamSynchronize(t->st));
    return 0;
}
int adj_done(fdga_ctx* c, fdga_ctx* t) {
    t->run = t->st;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}
AdjAChan adj_chan_a(const fdga_ctx* c, int ch) {
    return AdjAChan{ch, c->NP == 1, c->cache[idx_fbase(ch)], c->F[ch].K2, c->F[ACH].K2, c->F[ch].K3, c->F[ACH].K3};
}
const double2* adj_gam(const fdga_ctx* c, int ch) { return c->cache[ch == PCH ? 0 : ch == ACH ? 10 : 11]; }
// the linearisation point: the solver's caches and FP64 right operands from its current F, FL and Π
int adj_point(fdga_ctx* c) {
    const int all[3] = {PCH, ACH, TCH};
    if (fdga_build_k3_cache(c)) return -3;
    return k3_rhs(c, k3_set(c, all, 3, false, false), 3) ? -3 : 0;
}
template <class AL, class BL, class EP>
int adj_gemm(fdga_ctx* c, int K, const AL& a, const BL& b, const EP& e) {
    const int NV = c->NV3;
    const GemmSchedOne sc = sched_one(NV, NV, K, cdiv(NV, 64), cdiv(NV, 64), c->NW3 * c->NP, nullptr, 0);
    if (zgemm_run(FDGA_FP64, TILE_64x64, (size_t)sc.nitems, c->st, sc, a, b, e) != hipSuccess)
        return fail(c, "tangent / adjoint GEMM launch failed", -3);
    return 0;
}
AdjRT adj_rt(fdga_ctx* c, int ch, const double2* Rb) {
    return AdjRT{ch, (ch == ACH) ? 1.0 : -1.0, (ch == TCH) ? -1.0 : 1.0, pi_of(c, ch, false), c->cache[idx_f0(ch)],
                      adj_gam(c, ch), c->FL[ch].K3, c->FL[ACH].K3, Rb};
}
// the transposed gather map of a tangent context, built and uploaded once
int adj_ensure(fdga_ctx* c, fdga_ctx* t) {
    if (t->adj_off) return 0;
    adj::GatherMap gm;
    if (!adj::gather_map(adj_boxes(t->d), t->NP == 1, gm)) return fail(c, "the gather map of these boxes does not fit");
    const adj::Boxes B = adj_boxes(t->d);
    t->adj_nd = B.ndst();
    if (dalloc(t, &t->adj_cbar, (size_t)adj::kSlots * t->n3) || dalloc(t, &t->adj_avg, t->adj_nd) ||
        dalloc(t, &t->adj_Ab, t->n3 * t->NP))
        return fail(c, "tangent context: " + t->err, -2);
    HIPCHK(c, hipMalloc((void**)&t->adj_off, gm.off.size() * sizeof(int64_t)));
    HIPCHK(c, hipMalloc((void**)&t->adj_src, std::max<size_t>(gm.src.size(), 1) * sizeof(int)));
    HIPCHK(c, hipMalloc((void**)&t->adj_cf, std::max<size_t>(gm.cf.size(), 1) * sizeof(float)));
    HIPCHK(c, hipMemcpy(t->adj_off, gm.off.data(), gm.off.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    if (!gm.src.empty()) {
        HIPCHK(c, hipMemcpy(t->adj_src, gm.src.data(), gm.src.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(t->adj_cf, gm.cf.data(), gm.cf.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return 0;
}

}  // namespace

extern "C" {

int64_t fdga_k3_gather_map(const fdga_dims* d, int64_t* off, int32_t* src, float* coef) {
    if (!d || d->nK1 < 1 || d->nK2b < 1 || d->nK2f < 1 || d->nK3b < 1 || d->nK3f < 1) return -1;
    adj::GatherMap gm;
    if (!adj::gather_map(adj_boxes(*d), d->nonlocal == 0 || d->L == 1, gm)) return -1;
    if (off) std::memcpy(off, gm.off.data(), gm.off.size() * sizeof(int64_t));
    if (src) std::memcpy(src, gm.src.data(), gm.src.size() * sizeof(int32_t));
    if (coef) std::memcpy(coef, gm.cf.data(), gm.cf.size() * sizeof(float));
    return (int64_t)gm.src.size();
}

int fdga_bse_k3_jvp(fdga_ctx* c, fdga_ctx* t) {
    if (!c || !t) return -1;
    if (int rc = adj_check(c, t)) return rc;
    if (adj_point(c) || refresh_avg(t) || fdga_build_k3_cache(t)) return -3;   // tangent caches: F0 = 0 in t
    const K3Geom g = k3geom(c);
    const size_t n = c->n3 * c->NP;
    for (int ch : {PCH, ACH, TCH}) {
        const AdjR r{ch, (ch == ACH) ? 1.0 : -1.0, (ch == TCH) ? -1.0 : 1.0, pi_of(c, ch, false), pi_of(t, ch, false),
                          c->cache[idx_f0(ch)], adj_gam(c, ch), adj_gam(t, ch), c->FL[ch].K3, c->FL[ACH].K3,
                          t->FL[ch].K3, t->FL[ACH].K3, t->R[ch]};
        k_adj_dR<<<cdiv(n, 256), 256, 0, c->st>>>(g, r);
        LAUNCHCHK(c);
        const AdjJvpA A{g, adj_chan_a(t, ch), adj_chan_a(c, ch)};
        const AdjJvpB B{g, c->R[ch], t->R[ch]};
        const AdjJvpE E{g, c->d.T, ch == TCH, t->FL[ch].K3, t->FL[ACH].K3, t->Fb[ch].K3};
        if (adj_gemm(c, 2 * c->NV3, A, B, E)) return -3;
    }
    k_adj_tpost<<<cdiv(n, 256), 256, 0, c->st>>>(t->Fb[TCH].K3, t->Fb[ACH].K3, n);
    LAUNCHCHK(c);
    return adj_done(c, t);
}

int fdga_bse_k3_vjp(fdga_ctx* c, fdga_ctx* t) {
    if (!c || !t) return -1;
    if (int rc = adj_check(c, t)) return rc;
    if (adj_ensure(c, t) || adj_point(c)) return -3;
    const K3Geom g = k3geom(c);
    const size_t n = c->n3 * c->NP, n3 = c->n3;
    const double T = c->d.T;
    // raw cotangents: the t post-processing transposed, in place
    k_adj_tpost_t<<<cdiv(n, 256), 256, 0, c->st>>>(t->Fb[TCH].K3, t->Fb[ACH].K3, n);
    LAUNCHCHK(c);
    // outputs start at zero: the vertex cotangent accumulates, Π̄ is written on the K3 box only
    for (int r = 0; r < 3; ++r)
        for (double2* q : {t->F[r].K1, t->F[r].K2, t->F[r].K3}) {
            const size_t m = q == t->F[r].K1 ? c->n1 * c->NP : q == t->F[r].K2 ? c->n2 * c->NP : n;
            k_zero<<<cdiv(m, 256), 256, 0, c->st>>>(q, m);
        }
    k_zero<<<cdiv(c->nPiel, 256), 256, 0, c->st>>>(t->Pipp, c->nPiel);
    k_zero<<<cdiv(c->nPiel, 256), 256, 0, c->st>>>(t->Piph, c->nPiel);
    k_zero<<<cdiv((size_t)adj::kSlots * n3, 256), 256, 0, c->st>>>(t->adj_cbar, (size_t)adj::kSlots * n3);
    LAUNCHCHK(c);
    // R̄_r = T A_r^T Ḡ_r (into t's right-operand buffers)
    for (int ch : {PCH, ACH, TCH})
        if (adj_gemm(c, c->NV3, AdjAT{g, adj_chan_a(c, ch)}, AdjDense<false>{g, t->Fb[ch].K3},
                     AdjOut{g, t->R[ch], T}))
            return -3;
    const AdjRT rp = adj_rt(c, PCH, t->R[PCH]), ra = adj_rt(c, ACH, t->R[ACH]), rt = adj_rt(c, TCH, t->R[TCH]);
    k_adj_pibar<<<cdiv((size_t)c->NV3 * c->NW3 * c->NP, 256), 256, 0, c->st>>>(g, rp, ra, rt, t->Pipp, t->Piph);
    k_adj_gambar<<<cdiv(n3, 256), 256, 0, c->st>>>(g, rp, ra, rt, t->adj_cbar);
    k_adj_flbar<<<cdiv(n, 256), 256, 0, c->st>>>(g, rp, ra, rt, t->Fb[PCH].K3, t->Fb[ACH].K3, t->Fb[TCH].K3,
                                                      t->FL[PCH].K3, t->FL[ACH].K3, t->FL[TCH].K3);
    LAUNCHCHK(c);
    // Ā_r = T Ḡ_r R_r^T: its momentum sum is the Fbase cotangent; on the lattice the same-channel
    // K2r + K3r terms of A take it per momentum
    for (int ch : {PCH, ACH, TCH}) {
        if (adj_gemm(c, c->NV3, AdjDense<false>{g, t->Fb[ch].K3}, AdjDense<true>{g, c->R[ch]},
                     AdjOut{g, t->adj_Ab, T}))
            return -3;
        const int slot = ch == PCH ? 4 : ch == ACH ? 5 : 6;
        k_adj_fbase<<<cdiv(n3, 256), 256, 0, c->st>>>(g, t->adj_Ab, t->adj_cbar + (size_t)slot * n3);
        LAUNCHCHK(c);
        if (c->NP > 1) {
            const unsigned nb = cdiv((size_t)c->NV3 * c->NW3 * c->NP, 256);
            k_adj_own<<<nb, 256, 0, c->st>>>(g, t->adj_Ab, ch == TCH ? 2.0 : 1.0, t->F[ch].K2, t->F[ch].K3);
            if (ch == TCH) k_adj_own<<<nb, 256, 0, c->st>>>(g, t->adj_Ab, -1.0, t->F[ACH].K2, t->F[ACH].K3);
            LAUNCHCHK(c);
        }
    }
    // the gather's transpose onto the averaged vertex, then the average's transpose (1 / NP per momentum)
    k_adj_gather_t<<<cdiv(t->adj_nd * WAVE, 256), 256, 0, c->st>>>(t->adj_avg, t->adj_off, t->adj_src, t->adj_cf,
                                                                         t->adj_cbar, t->adj_nd);
    LAUNCHCHK(c);
    const size_t per = c->n1 + c->n2 + c->n3;
    for (int r = 0; r < 3; ++r) {
        const double2* a = t->adj_avg + (size_t)r * per;
        k_adj_bcast<<<cdiv(c->n1 * c->NP, 256), 256, 0, c->st>>>(t->F[r].K1, a, c->n1, c->NP, 1.0 / c->NP);
        k_adj_bcast<<<cdiv(c->n2 * c->NP, 256), 256, 0, c->st>>>(t->F[r].K2, a + c->n1, c->n2, c->NP, 1.0 / c->NP);
        k_adj_bcast<<<cdiv(n, 256), 256, 0, c->st>>>(t->F[r].K3, a + c->n1 + c->n2, c->n3, c->NP, 1.0 / c->NP);
    }
    LAUNCHCHK(c);
    t->pi_band = 0;
    t->r32_fresh = 0;
    return adj_done(c, t);
}

}  // extern "C"
