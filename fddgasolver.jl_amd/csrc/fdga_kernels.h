// fdga_kernels.h -- gfx950 kernels of the parquet iteration (FP64 path).
//
// Kernel map (reference function -> kernel):
//   Dyson!                 k_dyson
//   bubbles! (local)       k_bubble_local        elementwise, HBM bound
//   bubbles_real_space!    k_dft2_rows (G -> G(R)) + k_bubble_nl (products and bfft over R' in LDS)
//   s-wave averages        k_pavg_multi
//   build_K3_cache!        k_cache               gather: Γ_irr / F0 / F-base caches
//   BSE_L_K3!, BSE_K3!     zgemm_tiles<...>      batched complex GEMM on v_mfma_f64_16x16x4f64
//   BSE_K1!, BSE_K2!,      k_bse_w + k_gemv_*    gather-GEMV over the bubble mesh
//   BSE_L_K2!              (k_bse_w and k_tail_sums keep W and the far sums on the band the gathers read)
//   SDE!                   k_sde_L, k_sde_sigma_local, k_u2_*, k_occupation
#pragma once
#include <type_traits>

#include "fdga_device.h"

namespace fdga {

constexpr int WAVE = 64;

// spin component of the channel's BSE: p, a parallel spin; t density (src/BSE_templates.jl:12-35)
template <int CH> struct ChSpin { static constexpr int SP = (CH == TCH) ? DSP : PSP; };


// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// ---------------------------------------------------------------------------------------
// Symmetry classes of an output array (SymmetryGroup of init_sym_grp!, src/ParquetSolver.jl:203-231).
// Class k holds mem[off[k] .. off[k+1]) with mem = (linear index << 2) | op; its first member is
// the representative (op 0).  op bit 0 negates, bit 1 conjugates.  ncls == 0 is the identity
// group: every output is its own class, evaluated densely.
struct SymCls {
    const int* off;
    const int* mem;
    int ncls;
    __device__ __forceinline__ int count(int nout) const { return ncls ? ncls : nout; }
    __device__ __forceinline__ int rep(int w) const { return ncls ? (mem[off[w]] >> 2) : w; }
    __device__ __forceinline__ void range(int w, int& b, int& e) const {
        if (ncls) { b = off[w]; e = off[w + 1]; } else { b = 0; e = 1; }
    }
    __device__ __forceinline__ int member(int w, int k) const { return ncls ? mem[k] : (w << 2); }
};
__device__ __forceinline__ double2 sym_apply(int op, double2 v) {
    if (op & 1) v = cmk(-v.x, -v.y);
    if (op & 2) v.y = -v.y;
    return v;
}

// Finish of a K3 launch under symmetry groups (one thread per element, coalesced stores):
//   x_s[j] = op(raw_s[rep(j)]) for each slot s (rep_s[j] = (rep << 2) | op; null: x_s[j] = raw_s[j]),
//   then for the t slot x_t[j] = (x_t[j] + x_a[j]) / 2 (src/BSE_templates.jl:176-177).
// The p and a slots are finished in place (raw_s = out_s: representatives keep their value, and
// only representatives are read); the t slot's raw values come from a separate array, because
// its post-processing also rewrites representatives.  Representatives of a K3 element lie in the
// element's own Ω block or in the -Ω block: workgroup w works on chunk wgmap[w] = iΩ * nchunk +
// chunk, and the map places both blocks of a pair (Ω, -Ω) on one XCD (workgroups are dispatched
// round-robin over the 8 XCDs), so the gathers hit that XCD's L2.  -1 entries are idle.
struct K3Fin {
    double2* out0; double2* out1; double2* out2;
    const double2* raw0; const double2* raw1; const double2* raw2;
    const int* rep0; const int* rep1; const int* rep2;
    int n;                    // slots in the launch
    int tslot, aslot;         // slot of t and of a (-1: absent)
    const double2* afinal;    // final a-channel K3 when a is not in the launch
};
__device__ __forceinline__ double2 fin_load(const double2* raw, const int* rep, size_t j) {
    if (!rep) return raw[j];
    const int code = rep[j];
    return sym_apply(code & 3, raw[code >> 2]);
}
// class members of x from their representatives, in place: x[j] = op(x[rep(j)]).  Representatives
// map to themselves and are not written, so every read sees the representative's own value.
__global__ void k_sym_fill(double2* __restrict__ x, const int* __restrict__ rep, size_t n) {
    const size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int code = rep[j];
    if ((size_t)(code >> 2) != j) x[j] = sym_apply(code & 3, x[code >> 2]);
}

__global__ __launch_bounds__(256) void k_k3_finish(K3Fin f, const int* __restrict__ wgmap, int nchunk, int blk) {
    const int code = wgmap[blockIdx.x];
    if (code < 0) return;
    const int iW = code / nchunk, e = (code - iW * nchunk) * 256 + (int)threadIdx.x;
    if (e >= blk) return;
    const size_t j = (size_t)e + (size_t)blk * iW;
    double2 v0 = fin_load(f.raw0, f.rep0, j), v1 = czero(), v2 = czero();
    if (f.n > 1) v1 = fin_load(f.raw1, f.rep1, j);
    if (f.n > 2) v2 = fin_load(f.raw2, f.rep2, j);
    if (f.tslot >= 0) {
        double2 a;
        if (f.aslot == 0) a = v0;
        else if (f.aslot == 1) a = v1;
        else if (f.aslot == 2) a = v2;
        else a = f.afinal[j];
        if (f.tslot == 0) v0 = cscale(cadd(v0, a), 0.5);
        else if (f.tslot == 1) v1 = cscale(cadd(v1, a), 0.5);
        else v2 = cscale(cadd(v2, a), 0.5);
    }
    f.out0[j] = v0;
    if (f.n > 1) f.out1[j] = v1;
    if (f.n > 2) f.out2[j] = v2;
}

// Dyson!: G = 1 / (1/Gbare + Σ)  (src/dyson.jl:5-31)
__global__ void k_dyson(double2* __restrict__ G, const double2* __restrict__ S, const double2* __restrict__ Gb, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) G[i] = cinv(cadd(cinv(Gb[i]), S[i]));
}

// bubbles! local with 1/ν tail (src/bubble.jl:9-36); output [ω, Ω] (ω fastest)
__global__ void k_bubble_local(double2* __restrict__ pp, double2* __restrict__ ph, const double2* __restrict__ G,
                               int nG, int nW, int nV, double T) {
    int iw = blockIdx.x * blockDim.x + threadIdx.x;
    int iW = blockIdx.y;
    if (iw >= 2 * nV) return;
    int m = iW - (nW - 1), n = iw - nV;
    auto g = [&](int k) -> double2 {
        return in_f(k, nG) ? G[k + nG] : cmk(1.0 / (M_PI * T * (2.0 * k + 1.0)), 0.0);
    };
    double2 gv = g(n);
    size_t o = (size_t)iw + (size_t)(2 * nV) * iW;
    pp[o] = cmul(gv, g(bmf(m, n)));
    ph[o] = cmul(gv, g(bpf(m, n)));
}

// ---------------------------------------------------------------------------------------
// 2D DFT over the trailing L x L plane of rows: data[r + nrows*(j1 + L*j2)], in place.
// One workgroup per block of RB rows (fewer rows per block when that leaves the GPU short of
// workgroups; every output is the same sum in the same order either way), two LDS planes (in/out)
// plus the L twiddles; sign -1 = fft, +1 = bfft (FFTW conventions of src/nonlocal/bubble.jl:110,154);
// `scale` multiplies the result.  The input rows are read from src, which is data itself (in place)
// or another array of the same shape: a workgroup reads all of its rows into LDS before it writes
// them, and no two workgroups share a row.
__global__ void k_dft2_rows(double2* data, const double2* src, size_t nrows, int L, int RB, int sign, double scale) {
    extern __shared__ double2 sm[];   // 2 * RB * L * L + L
    const int LL = L * L;
    double2* s0 = sm;
    double2* s1 = sm + RB * LL;
    double2* tw = sm + 2 * RB * LL;
    const size_t r0 = (size_t)blockIdx.x * RB;
    const int nrb = (int)min((size_t)RB, nrows - r0);
    for (int k = threadIdx.x; k < L; k += blockDim.x) {
        double sn, cs;
        sincospi(sign * 2.0 * k / L, &sn, &cs);
        tw[k] = cmk(cs, sn);
    }
    for (int e0 = threadIdx.x; e0 < RB * LL; e0 += 4 * blockDim.x) {   // four loads in flight per thread
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = e0 + u * blockDim.x, r = e % RB, j = e / RB;
            v[u] = (e < RB * LL && r < nrb) ? src[r0 + r + nrows * (size_t)j] : czero();
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (e0 + u * blockDim.x < RB * LL) s0[e0 + u * blockDim.x] = v[u];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < RB * LL; e += blockDim.x) {   // along j1
        int r = e % RB, j = e / RB, k1 = j % L, j2 = j / L;
        double2 s = czero();
        for (int j1 = 0; j1 < L; ++j1) s = cadd(s, cmul(s0[r + RB * (j1 + L * j2)], tw[(j1 * k1) % L]));
        s1[e] = s;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < RB * LL; e += blockDim.x) {   // along j2
        int r = e % RB, j = e / RB, j1 = j % L, k2 = j / L;
        double2 s = czero();
        for (int j2 = 0; j2 < L; ++j2) s = cadd(s, cmul(s1[r + RB * (j1 + L * j2)], tw[(j2 * k2) % L]));
        s0[e] = s;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < RB * LL; e += blockDim.x) {
        int r = e % RB, j = e / RB;
        if (r < nrb) data[r0 + r + nrows * (size_t)j] = cscale(s0[e], scale);
    }
}

// bubbles_real_space! (src/nonlocal/bubble.jl:87-158) fused for a tile of NB bubble frequencies ω
// at one Ω: the real-space products Π_R(ω, Ω, R') over the R in [-L/2, L/2]^2 that fold onto
// R' = mod(R, L) (half weight at |R_i| = LG/2 for even LG; the 1/ν tail of G only at R = 0) are
// formed in LDS, transformed by the unnormalised backward 2D DFT over R' (FFTW bfft, sign +1) in
// LDS, and written once as Π(ω, Ω, q), q = q1 + L q2.  Πpp and Πph use the LDS plane in turn.
// LDS plane: s[R' * NB + ω - ω0] (the NB frequencies of one R' are contiguous).
template <int L>
struct BubbleDft {
    // y = Σ_j x[j] e^{+2πi jk/L} over j < L, twiddles tw[k] = e^{+2πi k/L}
    __device__ static void run(double2* x, const double2* tw, int Lr) {
        double2 y[L];
#pragma unroll
        for (int k = 0; k < L; ++k) {
            double2 acc = czero();
#pragma unroll
            for (int j = 0; j < L; ++j) acc = cadd(acc, cmul(x[j], tw[(j * k) % L]));
            y[k] = acc;
        }
#pragma unroll
        for (int k = 0; k < L; ++k) x[k] = y[k];
    }
};
// length 16: radix-4 in two stages, j = 4 j1 + j2, k = k1 + 4 k2:
//   X[k1 + 4 k2] = Σ_j2 e^{2πi j2 k2/4} [e^{2πi j2 k1/16} Σ_j1 x[4 j1 + j2] e^{2πi j1 k1/4}]
template <class CT>
__device__ __forceinline__ void dft4p(CT& a, CT& b, CT& c, CT& d) {
    // sign +1: X0 = a+b+c+d, X1 = a+ib-c-id, X2 = a-b+c-d, X3 = a-ib-c+id
    const CT s0 = cadd(a, c), s1 = csub(a, c), s2 = cadd(b, d), s3 = csub(b, d);
    CT is3;
    is3.x = -s3.y;
    is3.y = s3.x;
    a = cadd(s0, s2);
    b = cadd(s1, is3);
    c = csub(s0, s2);
    d = csub(s1, is3);
}
template <>
struct BubbleDft<16> {
    template <class CT>   // double2, or float2 for the FP32 form
    __device__ static void run(CT* x, const CT*, int) {
        constexpr double C1 = 0.92387953251128675613, S1 = 0.38268343236508977173, C2 = 0.70710678118654752440;
#pragma unroll
        for (int j2 = 0; j2 < 4; ++j2) dft4p(x[j2], x[4 + j2], x[8 + j2], x[12 + j2]);
        // x[4 k1 + j2] now holds A[j2][k1]; twiddle e^{2πi j2 k1/16}
        const CT w1 = cmk_as<CT>(C1, S1), w2 = cmk_as<CT>(C2, C2), w3 = cmk_as<CT>(S1, C1), w6 = cmk_as<CT>(-C2, C2),
                 w9 = cmk_as<CT>(-C1, -S1);
        x[4 + 1] = cmul(x[4 + 1], w1);
        x[4 + 2] = cmul(x[4 + 2], w2);
        x[4 + 3] = cmul(x[4 + 3], w3);
        x[8 + 1] = cmul(x[8 + 1], w2);
        {   // e^{iπ/2}
            const CT z = x[8 + 2];
            x[8 + 2].x = -z.y;
            x[8 + 2].y = z.x;
        }
        x[8 + 3] = cmul(x[8 + 3], w6);
        x[12 + 1] = cmul(x[12 + 1], w3);
        x[12 + 2] = cmul(x[12 + 2], w6);
        x[12 + 3] = cmul(x[12 + 3], w9);
        CT y[16];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            CT a = x[4 * k1], b = x[4 * k1 + 1], c = x[4 * k1 + 2], d = x[4 * k1 + 3];
            dft4p(a, b, c, d);
            y[k1] = a; y[k1 + 4] = b; y[k1 + 8] = c; y[k1 + 12] = d;
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = y[k];
    }
};

// The R in [-L/2, L/2]^2 that fold onto R' = mod(R, L) (up to two per axis), with their
// k-mesh indices on the LG mesh (+R and -R), weights (1/2 per axis at |R_i| = LG/2 for even LG)
// and whether R = 0; built once per solver on the host (bubble_folds in the library).
// The R in [-L/2, L/2]^2 that fold onto one R' = mod(R, L), at most 4, packed one per int of an
// int4 so that a product loads its whole fold with one 16-byte load:
//   bits 0-11 index of R in G(R) (LG x LG), bits 12-23 index of -R, bit 24 R is the origin,
//   bits 25-26 weight code (0: no term, 1: weight 1, 2: 1/2, 3: 1/4; half weights at |R| = LG/2)
typedef int4 BubbleFold;
__device__ __forceinline__ int fold_kp(int t) { return t & 0xFFF; }
__device__ __forceinline__ int fold_km(int t) { return (t >> 12) & 0xFFF; }
__device__ __forceinline__ bool fold_origin(int t) { return (t >> 24) & 1; }
__device__ __forceinline__ int fold_wcode(int t) { return (t >> 25) & 3; }
// length 24 = 3 x 8, j = 8 j1 + j2, k = k1 + 3 k2:
//   X[k1 + 3 k2] = Σ_j2 e^{2πi j2 k2/8} [e^{2πi j2 k1/24} Σ_j1 x[8 j1 + j2] e^{2πi j1 k1/3}]
// (3-point DFTs, twiddles from tw[k] = e^{2πi k/24}, 8-point FFTs as 4 x 2)
__device__ __forceinline__ void dft3p(double2& a, double2& b, double2& c) {
    constexpr double S3 = 0.86602540378443864676;   // sin(2π/3)
    const double2 s = cadd(b, c), d = csub(b, c);
    const double2 m = csub(a, cscale(s, 0.5));
    const double2 id = cmk(-S3 * d.y, S3 * d.x);   // i sin(2π/3) (b - c)
    a = cadd(a, s);
    b = cadd(m, id);
    c = csub(m, id);
}
__device__ __forceinline__ void dft8p(double2* y) {
    constexpr double C2 = 0.70710678118654752440;
    dft4p(y[0], y[2], y[4], y[6]);
    dft4p(y[1], y[3], y[5], y[7]);
    // odd half times e^{2πi k/8}, k = 0..3 (now at y[1], y[3], y[5], y[7])
    y[3] = cmul(y[3], cmk(C2, C2));
    y[5] = cmk(-y[5].y, y[5].x);
    y[7] = cmul(y[7], cmk(-C2, C2));
    double2 o[8];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        o[k] = cadd(y[2 * k], y[2 * k + 1]);
        o[k + 4] = csub(y[2 * k], y[2 * k + 1]);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) y[k] = o[k];
}
template <>
struct BubbleDft<24> {
    __device__ static void run(double2* x, const double2* tw, int) {
        double2 A[24];   // A[3 j2 + k1]
#pragma unroll
        for (int j2 = 0; j2 < 8; ++j2) {
            double2 a = x[j2], b = x[8 + j2], c = x[16 + j2];
            dft3p(a, b, c);
            A[3 * j2] = a;
            A[3 * j2 + 1] = j2 ? cmul(b, tw[j2]) : b;
            A[3 * j2 + 2] = j2 ? cmul(c, tw[2 * j2]) : c;
        }
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1) {
            double2 y[8];
#pragma unroll
            for (int j2 = 0; j2 < 8; ++j2) y[j2] = A[3 * j2 + k1];
            dft8p(y);
#pragma unroll
            for (int k2 = 0; k2 < 8; ++k2) x[k1 + 3 * k2] = y[k2];
        }
    }
};

// Π_R(ω, Ω, R') for one R' (pp: which = 0, ph: 1; see k_bubble_nl)
template <class CT>   // double2, or float2 (the FP32 form: G(R) in FP32, products in FP32)
__device__ __forceinline__ CT bubble_R(const CT* __restrict__ GR, const BubbleFold& f, int nG, double T, int m, int n,
                                       int which) {
    const int NG = 2 * nG;
    CT acc = cmk_as<CT>(0.0, 0.0);
    const int n2nd = which == 0 ? bmf(m, n) : bpf(m, n);
    const int terms[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int t = terms[c], wc = fold_wcode(t);
        if (wc == 0) break;
        const bool origin = fold_origin(t);
        auto gat = [&](int k, int kk) -> CT {
            if (in_f(k, nG)) return GR[(size_t)(k + nG) + (size_t)NG * kk];
            return cmk_as<CT>(origin ? 1.0 / (M_PI * T * (2.0 * k + 1.0)) : 0.0, 0.0);
        };
        const CT x = cmul(gat(n2nd, fold_kp(t)), gat(n, which == 0 ? fold_kp(t) : fold_km(t)));
        acc = cadd(acc, wc == 1 ? x : cscale(x, wc == 2 ? 0.5 : 0.25));
    }
    return acc;
}

// Blocks (16 ω chunks) with blockIdx.x in [wlo, whi) store their Π values; every block stores its chunk
// sums.  The iteration passes the chunks that hold the band ω ∈ [-wb, wb) (every Π entry a later
// stage reads outside the band enters only through the chunk sums); [0, gridDim.x) stores all of Π.
// LT: L fixed at compile time (16: radix-4 FFT; 24: unrolled direct DFT), or 0 for a run-time
// L <= 32 (direct DFT through private arrays, for small test lattices).  Block: NT threads (512 at
// L = 16, held to 128 VGPRs: two blocks per CU, twice the waves on the same LDS tiles, to hide the
// latency of the G(R) gathers).
// CT: double2, or float2 for the FP32 form (G(R), the products, the transform and the LDS tiles in
// FP32; Π stored and the chunk sums accumulated in FP64): 32 KB of LDS per block and at most 80 VGPRs,
// three blocks per CU.
template <int LT, int NBT, int NT = 256, class CT = double2>
__global__ __launch_bounds__(NT, NT >= 512 ? (sizeof(CT) == 8 ? 6 : 4) : 1) void k_bubble_nl(double2* __restrict__ pp, double2* __restrict__ ph,
                                                   const CT* __restrict__ GR, const BubbleFold* __restrict__ folds,
                                                   int nG, int Lr, int nW, int nV, double T, int NBr,
                                                   double2* __restrict__ cpp, double2* __restrict__ cph, int wlo,
                                                   int whi) {
    extern __shared__ double2 sm[];   // NB * L * L (+ L twiddles unless L = 16) of CT
    const int L = LT ? LT : Lr, LL = L * L, NB = NBT ? NBT : NBr;
    CT* s = reinterpret_cast<CT*>(sm);
    CT* tw = s + NB * LL;
    const int tid = threadIdx.x;
    const int iw0 = blockIdx.x * NB, iW = blockIdx.y;
    const int m = iW - (nW - 1);
    const int nb = min(NB, 2 * nV - iw0);
    if (LT != 16)
        for (int k = tid; k < L; k += blockDim.x) {
            double sn, cs;
            sincospi(2.0 * k / L, &sn, &cs);
            tw[k] = cmk_as<CT>(cs, sn);
        }
    for (int which = 0; which < 2; ++which) {
        // Far from ω = 0 one of each product's two frequencies lies outside G's mesh for every ω of
        // the block: then every product vanishes except the origin's (R = 0, the asymptotic G on both
        // sides), which folds onto R' = 0 alone, and the transform of that delta is the constant
        // Π_R(ω, R' = 0) at every momentum (exactly: the radix-4 steps add zeros and multiply zeros).
        // Such blocks store that constant and its chunk sums without the products and transforms:
        // the same values (up to the sign of zeros).  (L = 16, the radix-4 form.)
        bool far = LT == 16;
        if (LT == 16)
            for (int wl = 0; wl < nb; ++wl) {
                const int n = iw0 + wl - nV, n2nd = which == 0 ? bmf(m, n) : bpf(m, n);
                far = far && (!in_f(n, nG) || !in_f(n2nd, nG));
            }
        if (far) {
            if (tid < NB)
                s[tid] = tid < nb ? bubble_R(GR, folds[0], nG, T, m, iw0 + tid - nV, which) : cmk_as<CT>(0.0, 0.0);
            __syncthreads();
            if ((int)blockIdx.x >= wlo && (int)blockIdx.x < whi) {
                double2* out = which == 0 ? pp : ph;
                const size_t qs = (size_t)(2 * nV) * (size_t)(2 * nW - 1);
                for (int e = tid; e < NB * LL; e += blockDim.x) {
                    const int wl = e % NB, P = e / NB;
                    if (wl < nb) out[(size_t)(iw0 + wl) + (size_t)(2 * nV) * (size_t)iW + qs * (size_t)P] = cmk(s[wl].x, s[wl].y);
                }
            }
            double2* cs = which == 0 ? cpp : cph;
            if (cs)
                for (int P = tid; P < LL; P += blockDim.x) {
                    double sr = 0.0, si = 0.0;
                    for (int j = 0; j < NB; ++j) {   // the order of the transformed block's sums below
                        const int wl = (j + P) % NB;
                        if (wl < nb) { const CT v = s[wl]; sr += v.x; si += v.y; }
                    }
                    cs[(size_t)P + (size_t)LL * (blockIdx.x + (size_t)gridDim.x * iW)] = cmk(sr, si);
                }
            __syncthreads();   // s is reused by the next products
            continue;
        }
        // products (a compile-time trip count when L and NB are fixed, so that the loads of
        // several products are in flight together)
        if (LT && NBT) {
            constexpr int NQ = (NBT * (LT ? LT : 1) * (LT ? LT : 1) + NT - 1) / NT;
#pragma unroll 8
            for (int q = 0; q < NQ; ++q) {
                const int e = tid + NT * q;
                if (e < NB * LL) {
                    const int wl = e % NB, R = e / NB;
                    s[e] = wl < nb ? bubble_R(GR, folds[R], nG, T, m, iw0 + wl - nV, which) : cmk_as<CT>(0.0, 0.0);
                }
            }
        } else {
            for (int e = tid; e < NB * LL; e += blockDim.x) {
                const int wl = e % NB, R = e / NB;
                s[e] = wl < nb ? bubble_R(GR, folds[R], nG, T, m, iw0 + wl - nV, which) : cmk_as<CT>(0.0, 0.0);
            }
        }
        __syncthreads();
        // along R1' (thread: ω, R2'), then along R2' (thread: ω, q1) with the write-out
        for (int pass = 0; pass < 2; ++pass) {
            for (int e = tid; e < NB * L; e += blockDim.x) {
                const int wl = e % NB, r = e / NB;
                const int base = pass == 0 ? L * r : r, step = pass == 0 ? 1 : L;
                CT x[LT ? LT : 32];
                for (int j = 0; j < L; ++j) x[j] = s[(base + step * j) * NB + wl];
                if (LT) BubbleDft<LT ? LT : 1>::run(x, tw, L);
                else {
                    CT y[32];
                    for (int k = 0; k < L; ++k) {
                        CT acc = cmk_as<CT>(0.0, 0.0);
                        for (int j = 0; j < L; ++j) acc = cadd(acc, cmul(x[j], tw[(j * k) % L]));
                        y[k] = acc;
                    }
                    for (int k = 0; k < L; ++k) x[k] = y[k];
                }
                if (pass == 0) {
                    for (int k = 0; k < L; ++k) s[(base + k) * NB + wl] = x[k];
                } else {
                    // back into the thread's own column, for the chunk sums below
                    for (int k = 0; k < L; ++k) s[(r + L * k) * NB + wl] = x[k];
                    if (wl < nb && (int)blockIdx.x >= wlo && (int)blockIdx.x < whi) {
                        double2* out = which == 0 ? pp : ph;
                        const size_t o = (size_t)(iw0 + wl) + (size_t)(2 * nV) * (size_t)iW;
                        const size_t qs = (size_t)(2 * nV) * (size_t)(2 * nW - 1);
                        for (int k = 0; k < L; ++k) out[o + qs * (size_t)(r + L * k)] = cmk(x[k].x, x[k].y);
                    }
                }
            }
            __syncthreads();
        }
        // chunk sums C[P + LL (chunk + chunks iΩ)] = Σ_{ω in this block's chunk} Π(ω, Ω, P) for the
        // tail sums (k_tail_sums_chunked); each thread a momentum, the ω order rotated by P so that
        // the lanes' LDS reads spread over the banks
        double2* cs = which == 0 ? cpp : cph;
        if (cs) {
            for (int P = tid; P < LL; P += blockDim.x) {
                double sr = 0.0, si = 0.0;
                for (int j = 0; j < NB; ++j) {
                    const int wl = (j + P) % NB;
                    if (wl < nb) { const CT v = s[P * NB + wl]; sr += v.x; si += v.y; }
                }
                cs[(size_t)P + (size_t)LL * (blockIdx.x + (size_t)gridDim.x * iW)] = cmk(sr, si);
            }
            __syncthreads();   // s is reused by the next products
        }
    }
}

// s-wave average over P of one array of `n` elements per P: a block takes 16 consecutive
// elements, its 16 thread groups sum every 16th momentum, and the groups' sums are added in a
// fixed order
// (Kf: the array held in FP32, read instead of K)
__device__ __forceinline__ void pavg_block(double2* __restrict__ A, const double2* __restrict__ K, const float2* __restrict__ Kf,
                                           size_t n, int NP, size_t blk) {
    __shared__ double2 red[16][17];
    const int el = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const size_t i = blk * 16 + el;
    double2 s = czero();
    if (i < n) {
        if (Kf)
            for (int p = grp; p < NP; p += 16) { const float2 v = Kf[i + n * p]; s = cadd(s, cmk(v.x, v.y)); }
        else
            for (int p = grp; p < NP; p += 16) s = cadd(s, K[i + n * p]);
    }
    red[grp][el] = s;
    __syncthreads();
    if (grp == 0 && i < n) {
        double2 t = czero();
#pragma unroll
        for (int g = 0; g < 16; ++g) t = cadd(t, red[g][el]);
        A[i] = cscale(t, 1.0 / NP);
    }
}
// several averages in one launch: job j takes blocks [blk0[j], blk0[j + 1])
constexpr int PAVG_JOBS = 9;
struct PavgJobs {
    double2* A[PAVG_JOBS];
    const double2* K[PAVG_JOBS];
    const float2* Kf[PAVG_JOBS];
    size_t n[PAVG_JOBS];
    unsigned blk0[PAVG_JOBS + 1];
    int njob;
};
__global__ __launch_bounds__(256) void k_pavg_multi(PavgJobs jobs, int NP) {
    int j = 0;
    while (j + 1 < jobs.njob && blockIdx.x >= jobs.blk0[j + 1]) ++j;
    pavg_block(jobs.A[j], jobs.K[j], jobs.Kf[j], jobs.n[j], NP, blockIdx.x - jobs.blk0[j]);
}

// ---------------------------------------------------------------------------------------
// build_K3_cache! (src/build_K3_cache.jl:20-95; NL src/nonlocal/build_K3_cache.jl:18-94).
// The P-independent caches are stored once; the P-dependent part of cache_F* (the same-channel
// K2 + K3 of the target vertex) is added inside the BSE GEMM operand loader.
//   c[0] Γpx [ω, νp]   c[1] F0p [ω, νp]   c[2] F0a [ω, νp]   c[3] F0t(d) [ω, νp]
//   c[4] Γpp [ν, ω]    c[5] Γa  [ν, ω]    c[6] Γt(d) [ν, ω]
//   c[7] Fbase_p [ν, ω]  c[8] Fbase_a [ν, ω]  c[9] Fbase_t(d) [ν, ω]
// c[0..9] the caches of build_K3_cache! (see k_cache2); c[10], c[11] Γa and Γt (t: 2 Γt - Γa) again,
// transposed to [ω, ν, Ω] for the right operands of BSE_K3! (written by k_cache2 through LDS)
struct CachePtrs { double2* c[12]; };

// loop 1 of build_K3_cache! at (Ω, ω, νp): Γpx and the F0 caches
__device__ __forceinline__ void cache1_elem(const CachePtrs& out, const VtxV& F, int n3b, int n3f, int i, int j, int iW) {
    const int NV = 2 * n3f;
    if (i >= NV) return;
    const int m = iW - (n3b - 1);
    const int a = i - n3f, b = j - n3f;
    const size_t o = (size_t)i + (size_t)NV * ((size_t)j + (size_t)NV * iW);
    out.c[0][o] = F.eval(m, a, b, -1, PCH, XSP, FL_GT | FL_GA);
    out.c[1][o] = csub(F.f0.eval(m, a, b, PCH, XSP), F.f0.eval(m, a, FDGA_INF, PCH, XSP));
    double2 f0a = csub(F.f0.eval(m, a, b, ACH, PSP), F.f0.eval(m, a, FDGA_INF, ACH, PSP));
    double2 f0t = csub(F.f0.eval(m, a, b, TCH, PSP), F.f0.eval(m, a, FDGA_INF, TCH, PSP));
    out.c[2][o] = f0a;
    out.c[3][o] = csub(cscale(f0t, 2.0), f0a);
}

// loop 2 of build_K3_cache! at (Ω, ν, ω): Γ_irr of each channel and the F caches.  With `full`
// (one momentum) the F caches are the reference's cache_Fr = F0 part + Γr + K2r(Ω,ν) + K3r(Ω,ν,ω)
// (t: 2 (t) - (a)); otherwise (NL) they hold only the P-independent part Fbase_r and the same-channel
// K2r + K3r(P) are added by the BSE_K3! loader and epilogue.
// (i, j) inside the box; returns the Γa and Γt (2 Γt - Γa) entries for the transposed copies
__device__ __forceinline__ void cache2_elem(const CachePtrs& out, const VtxV& F, int n3b, int n3f, int full, int i, int j,
                                            int iW, double2& ga_out, double2& gt_out) {
    const int NV = 2 * n3f;
    const int m = iW - (n3b - 1);
    const int a = i - n3f, b = j - n3f;
    const size_t o = (size_t)i + (size_t)NV * ((size_t)j + (size_t)NV * iW);
    double2 gpp = F.eval_psp(m, a, b, -1, PCH, FL_GT | FL_GA);
    double2 ga = F.eval_psp(m, a, b, -1, ACH, FL_GP | FL_GT);
    double2 gt = F.eval_psp(m, a, b, -1, TCH, FL_GP | FL_GA);
    double2 dp = csub(F.f0.eval_psp(m, a, b, PCH), F.f0.eval_psp(m, FDGA_INF, b, PCH));
    double2 da = csub(F.f0.eval_psp(m, a, b, ACH), F.f0.eval_psp(m, FDGA_INF, b, ACH));
    double2 dt = csub(F.f0.eval_psp(m, a, b, TCH), F.f0.eval_psp(m, FDGA_INF, b, TCH));
    ga_out = ga;
    gt_out = csub(cscale(gt, 2.0), ga);
    out.c[4][o] = gpp;
    out.c[5][o] = ga_out;
    out.c[6][o] = gt_out;
    double2 fbp = cadd(dp, gpp), fba = cadd(da, ga), fbt = cadd(dt, gt);
    if (full) {
        // within the K3 box F(ν) - F(νInf) of the channel's own γ is K2(Ω, ν) + K3(Ω, ν, ω)
        fbp = cadd(fbp, cadd(F.g[PCH].k2(m, a, 0), F.g[PCH].k3(m, a, b, 0)));
        fba = cadd(fba, cadd(F.g[ACH].k2(m, a, 0), F.g[ACH].k3(m, a, b, 0)));
        fbt = cadd(fbt, cadd(F.g[TCH].k2(m, a, 0), F.g[TCH].k3(m, a, b, 0)));
    }
    out.c[7][o] = fbp;
    out.c[8][o] = fba;
    out.c[9][o] = csub(cscale(fbt, 2.0), fba);
}
// loop 1 over tiles of 16 ω x 4 νp (one 64-thread block, blockIdx.z = iΩ)
__global__ __launch_bounds__(64) void k_cache1(CachePtrs out, VtxV F, int n3b, int n3f) {
    const int j = blockIdx.y * 4 + (threadIdx.x >> 4);
    if (j < 2 * n3f) cache1_elem(out, F, n3b, n3f, blockIdx.x * 16 + (threadIdx.x & 15), j, blockIdx.z);
}
// loop 2 over tiles of 16 ν x 4 ω (one 64-thread block, blockIdx.z = iΩ); the Γa and Γt entries
// also go through LDS to their transposed copies, four consecutive ω per store run
__global__ __launch_bounds__(64) void k_cache2(CachePtrs out, VtxV F, int n3b, int n3f, int full) {
    __shared__ double2 t[2][16][5];
    const int NV = 2 * n3f;
    const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 4, iW = blockIdx.z;
    {
        const int il = threadIdx.x & 15, jl = threadIdx.x >> 4;
        if (i0 + il < NV && j0 + jl < NV)
            cache2_elem(out, F, n3b, n3f, full, i0 + il, j0 + jl, iW, t[0][il][jl], t[1][il][jl]);
    }
    __syncthreads();
    const int il = threadIdx.x >> 2, jl = threadIdx.x & 3;
    if (i0 + il < NV && j0 + jl < NV) {
        const size_t o = (size_t)(j0 + jl) + (size_t)NV * ((size_t)(i0 + il) + (size_t)NV * iW);
        out.c[10][o] = t[0][il][jl];
        out.c[11][o] = t[1][il][jl];
    }
}

// materialise cache_F*[Ω, ν, ω, P] in the reference layout (ABI read-back only)
__global__ void k_cache_F_julia(double2* __restrict__ out, const double2* __restrict__ base, ChanV gr, ChanV ga, int ch,
                                int n3b, int n3f, int NP, int full) {
    const int NV = 2 * n3f, NW = 2 * n3b - 1;
    size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t tot = (size_t)NW * NV * NV * NP;
    if (e >= tot) return;
    int iW = (int)(e % NW);
    size_t r = e / NW;
    int i = (int)(r % NV); r /= NV;
    int j = (int)(r % NV);
    int p = (int)(r / NV);
    const int m = iW - (n3b - 1), v = i - n3f, w = j - n3f;
    double2 x = base[(size_t)i + (size_t)NV * ((size_t)j + (size_t)NV * iW)];
    if (full) { out[e] = x; return; }
    double2 dr = cadd(gr.k2(m, v, p), gr.k3(m, v, w, p));
    if (ch == TCH) dr = csub(cscale(dr, 2.0), cadd(ga.k2(m, v, p), ga.k3(m, v, w, p)));
    out[e] = cadd(x, dr);
}

// ---------------------------------------------------------------------------------------
// Batched complex GEMM on the matrix cores.  C[b](m, n) = Σ_k A[b](m, k) B[b](k, n), operands
// produced element-wise by loader functors (the gather is fused into the load), results consumed
// by an epilogue functor.  256 threads = 4 waves, 64 x 64 complex tile per workgroup, each wave a
// 32 x 32 sub-tile = 2 x 2 MFMA tiles with separate real/imag accumulators
// (C_r = A_r B_r - A_i B_i, C_i = A_r B_i + A_i B_r).
//   F32 = false: v_mfma_f64_16x16x4f64 (FP64 in, FP64 accumulate)
//   F32 = true : v_mfma_f32_16x16x4f32 (operands rounded to FP32, FP32 accumulate)
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int GBM = 64, GBN = 64, GBK = 16, GPAD = 1;

template <bool F32> struct MfmaT;
template <> struct MfmaT<false> {
    typedef double S;
    typedef v4d Acc;
    __device__ __forceinline__ static Acc mfma(S a, S b, Acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    // accumulator element r of lane l is D(row = l/16 + 4 r, col = l % 16); D is the transposed
    // output tile (operands swapped), so row() gives the output column
    __device__ __forceinline__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct MfmaT<true> {
    typedef float S;
    typedef v4f Acc;
    __device__ __forceinline__ static Acc mfma(S a, S b, Acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    // accumulator element r of lane l is D(row = 4 (l/16) + r, col = l % 16)
    __device__ __forceinline__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// Operand loaders may opt out of the register prefetch (a loader with many loads per element
// raises the register count to one wave per SIMD): a loader type with `static constexpr bool
// prefetch = false` is staged synchronously.
template <class L> struct Prefetch { static constexpr bool on = true; };

// A bound loader whose element combines several loads may expose them separately: a `Raw` type,
// `raw(i, k)` issuing the loads and `combine(Raw)` the arithmetic.  The register prefetch then holds
// the raw loads and combines them only when the tile is written to LDS, after the MFMAs of the
// current tile, so that the loads stay in flight during the MFMAs.  Other loaders are their own raw
// value.  A zero-filled Raw (the padding outside the matrix) must combine to zero.
template <class L, class = void>
struct LoaderRaw {
    typedef double2 Raw;
    __device__ __forceinline__ static Raw load(const L& l, int a, int b) { return l(a, b); }
    __device__ __forceinline__ static Raw zero() { return czero(); }
    __device__ __forceinline__ static double2 combine(const L&, const Raw& r) { return r; }
};
template <class L>
struct LoaderRaw<L, std::void_t<typename L::Raw>> {
    typedef typename L::Raw Raw;
    __device__ __forceinline__ static Raw load(const L& l, int a, int b) { return l.raw(a, b); }
    __device__ __forceinline__ static Raw zero() { return Raw{}; }
    __device__ __forceinline__ static double2 combine(const L& l, const Raw& r) { return l.combine(r); }
};

// One work item of a GEMM launch: an output tile of one batch.
struct GemmWork { int batch, m0, n0, M, N, K; };

// Work items of a launch with one problem shape.  Without a block list and with at least 8
// batches, item it runs batch 8 (it / 8 / tiles) + it % 8, tile (it / 8) % tiles: items go
// round-robin to the 8 XCDs, and the tiles of one batch (which share their A rows and B columns)
// stay on one XCD; the item count covers a multiple of 8 batches, items past the last batch are
// idle.  With fewer batches the items are the tiles in order (spread over all XCDs).  A block list
// (built in the XCD order, -1 = idle) selects the computed tiles instead.
struct GemmSchedOne {
    int M, N, K, tiles_m, tiles_n, nbatch, nitems;
    const int* blist;
    // pnp > 1: the batches are b = (slot pnw + iΩ) pnp... in memory order (slot pnp + P) pnw + iΩ, and
    // they are visited momentum-fastest, so that the workgroups an XCD runs at a time share a few Ω
    // and their momentum-independent operands stay in that XCD's L2
    int pnw, pnp;
    __device__ __forceinline__ bool item(int it, int BM, int BN, GemmWork& w) const {
        const int tiles = tiles_m * tiles_n;
        int bid;
        if (blist) {
            bid = blist[it];
        } else if (nbatch < 8) {
            bid = it;
        } else {
            const int r = it >> 3;
            int bt = (r / tiles) * 8 + (it & 7);
            if (bt < nbatch && pnp > 1) {
                const int per = pnw * pnp, sl = bt / per, q = bt - sl * per;
                bt = sl * per + (q % pnp) * pnw + q / pnp;
            }
            bid = bt < nbatch ? bt * tiles + r % tiles : -1;
        }
        if (bid < 0) return false;
        w.batch = bid / tiles;
        const int t = bid - w.batch * tiles;
        w.m0 = (t % tiles_m) * BM;
        w.n0 = (t / tiles_m) * BN;
        w.M = M; w.N = N; w.K = K;
        return true;
    }
};

// Work items of a grouped launch over several problems (solver contexts) with their own shapes:
// items[it] = member << 24 | local block (batch * tiles + tile of that member), -1 = idle; the
// member's batches are numbered from geo[member].boff in the launch's batch space.
struct GemmGroupGeom { int M, N, K, tiles_m, tiles_n, boff; };
struct GemmSchedGroup {
    const int* items;
    const GemmGroupGeom* geo;
    int nitems;
    __device__ __forceinline__ bool item(int it, int BM, int BN, GemmWork& w) const {
        const int e = items[it];
        if (e < 0) return false;
        const GemmGroupGeom g = geo[e >> 24];
        const int lb = e & 0xFFFFFF, tiles = g.tiles_m * g.tiles_n;
        const int b = lb / tiles, t = lb - b * tiles;
        w.batch = g.boff + b;
        w.m0 = (t % g.tiles_m) * BM;
        w.n0 = (t / g.tiles_m) * BN;
        w.M = g.M; w.N = g.N; w.K = g.K;
        return true;
    }
};

// A functor over the members of a grouped launch: bind(batch) binds the member's own functor to
// its local batch (members in batch order, boff[j] = first batch of member j).
template <class F>
struct GroupOp {
    const F* f;
    const int* boff;
    int n;
    __device__ __forceinline__ auto bind(int batch) const -> decltype(f->bind(0)) {
        int j = 0;
        while (j + 1 < n && batch >= boff[j + 1]) ++j;
        return f[j].bind(batch - boff[j]);
    }
};
template <class F> struct Prefetch<GroupOp<F>> { static constexpr bool on = Prefetch<F>::on; };

// Tile shapes: a workgroup of WGM x WGN waves, each wave a 32 x 32 sub-tile (2 x 2 MFMA tiles), so
// the workgroup tile is 32 WGM x 32 WGN.  The workgroups loop over the work items (it = blockIdx,
// blockIdx + gridDim, ...): a grid smaller than the item count makes the launch persistent; every
// workgroup leaves the loop when the items run out.  Each output element accumulates over k in
// the same order (k-steps of GBK, then 4 at a time) whatever the tile shape or grid, so the results
// do not depend on them.
template <bool F32, int WGM, int WGN, class SC, class AL, class BL, class EP>
__global__ __launch_bounds__(64 * WGM * WGN) void zgemm_tiles_t(SC sc, AL al, BL bl, EP ep) {
    typedef MfmaT<F32> MT;
    typedef typename MT::S S;
    typedef typename MT::Acc Acc;
    constexpr int NT = 64 * WGM * WGN, BM = 32 * WGM, BN = 32 * WGN;
    __shared__ S As[2][GBK][BM + GPAD];
    __shared__ S Bs[2][GBK][BN + GPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave % WGM) * 32, wn = (wave / WGM) * 32;
    constexpr int QA = (BM * GBK) / NT, QB = (BN * GBK) / NT;
    constexpr bool PF = Prefetch<AL>::on && Prefetch<BL>::on;
    for (int it = blockIdx.x; it < sc.nitems; it += gridDim.x) {
        GemmWork w;
        if (!sc.item(it, BM, BN, w)) continue;
        const int M = w.M, N = w.N, K = w.K, m0 = w.m0, n0 = w.n0;
        const auto A = al.bind(w.batch);
        const auto B = bl.bind(w.batch);
        const auto E = ep.bind(w.batch);
        Acc accr[2][2], acci[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                accr[i][j] = Acc{0, 0, 0, 0};
                acci[i][j] = Acc{0, 0, 0, 0};
            }
        // software pipeline: the operands of the next K-tile are loaded into registers while the
        // current one is multiplied from LDS
        typedef LoaderRaw<std::decay_t<decltype(A)>> RA;
        typedef LoaderRaw<std::decay_t<decltype(B)>> RB;
        typename RA::Raw pa[QA];
        typename RB::Raw pb[QB];
        auto fetch = [&](int k0) {
#pragma unroll
            for (int q = 0; q < QA; ++q) {
                const int e = tid + NT * q;
                const int gm = m0 + e % BM, gk = k0 + e / BM;
                pa[q] = (gm < M && gk < K) ? RA::load(A, gm, gk) : RA::zero();
            }
#pragma unroll
            for (int q = 0; q < QB; ++q) {
                const int e = tid + NT * q;
                // B operands are stored with k (ω) fastest: consecutive threads take consecutive k
                const int gn = n0 + e / GBK, gk = k0 + e % GBK;
                pb[q] = (gn < N && gk < K) ? RB::load(B, gk, gn) : RB::zero();
            }
        };
        if (PF) fetch(0);
        for (int k0 = 0; k0 < K; k0 += GBK) {
            if (PF) {
#pragma unroll
                for (int q = 0; q < QA; ++q) {
                    const int e = tid + NT * q;
                    const double2 v = RA::combine(A, pa[q]);
                    As[0][e / BM][e % BM] = (S)v.x;
                    As[1][e / BM][e % BM] = (S)v.y;
                }
#pragma unroll
                for (int q = 0; q < QB; ++q) {
                    const int e = tid + NT * q;
                    const double2 v = RB::combine(B, pb[q]);
                    Bs[0][e % GBK][e / GBK] = (S)v.x;
                    Bs[1][e % GBK][e / GBK] = (S)v.y;
                }
            } else {
                // element by element: load and store to LDS without holding the tile in registers
#pragma unroll
                for (int q = 0; q < QA; ++q) {
                    const int e = tid + NT * q;
                    const int gm = m0 + e % BM, gk = k0 + e / BM;
                    const double2 v = (gm < M && gk < K) ? A(gm, gk) : czero();
                    As[0][e / BM][e % BM] = (S)v.x;
                    As[1][e / BM][e % BM] = (S)v.y;
                }
#pragma unroll
                for (int q = 0; q < QB; ++q) {
                    const int e = tid + NT * q;
                    const int gn = n0 + e / GBK, gk = k0 + e % GBK;
                    const double2 v = (gn < N && gk < K) ? B(gk, gn) : czero();
                    Bs[0][e % GBK][e / GBK] = (S)v.x;
                    Bs[1][e % GBK][e / GBK] = (S)v.y;
                }
            }
            __syncthreads();
            if (PF && k0 + GBK < K) fetch(k0 + GBK);
#pragma unroll
            for (int kk = 0; kk < GBK; kk += 4) {
                const int kr = kk + (lane >> 4), c = lane & 15;
                S ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    ar[i] = As[0][kr][wm + 16 * i + c];
                    ai[i] = As[1][kr][wm + 16 * i + c];
                    nai[i] = -ai[i];
                }
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    br[j] = Bs[0][kr][wn + 16 * j + c];
                    bi[j] = Bs[1][kr][wn + 16 * j + c];
                }
                // B fragment as the MFMA's left operand: the accumulator is the transposed tile, so
                // lane % 16 runs along the output rows (contiguous in memory).  The eight first terms
                // are independent; each accumulator's second term follows eight MFMAs later.
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        accr[i][j] = MT::mfma(br[j], ar[i], accr[i][j]);
                        acci[i][j] = MT::mfma(bi[j], ar[i], acci[i][j]);
                    }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        accr[i][j] = MT::mfma(bi[j], nai[i], accr[i][j]);
                        acci[i][j] = MT::mfma(br[j], ai[i], acci[i][j]);
                    }
            }
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = m0 + wm + 16 * i + (lane & 15);
                    const int col = n0 + wn + 16 * j + MT::row(lane, r);
                    if (row < M && col < N) E(row, col, cmk((double)accr[i][j][r], (double)acci[i][j][r]));
                }
    }
}

// tile shapes of the GEMM launches: workgroups of 2 x 2, 1 x 1, 2 x 1 and 1 x 2 waves
enum { TILE_64x64 = 0, TILE_32x32 = 1, TILE_64x32 = 2, TILE_32x64 = 3, TILE_N = 4 };
inline int tile_bm(int tile) { return tile == TILE_64x64 || tile == TILE_64x32 ? 64 : 32; }
inline int tile_bn(int tile) { return tile == TILE_64x64 || tile == TILE_32x64 ? 64 : 32; }

// launch of one scheduler over `grid` workgroups (persistent when grid < sc.nitems) in the given
// precision (0 FP64, 1 FP32 MFMA) and tile shape
template <int WGM, int WGN, class SC, class AL, class BL, class EP>
inline void zgemm_go(int precision, unsigned grid, hipStream_t st, const SC& sc, const AL& a, const BL& b, const EP& e) {
    if (precision == 1) zgemm_tiles_t<true, WGM, WGN><<<grid, 64 * WGM * WGN, 0, st>>>(sc, a, b, e);
    else zgemm_tiles_t<false, WGM, WGN><<<grid, 64 * WGM * WGN, 0, st>>>(sc, a, b, e);
}
template <class SC, class AL, class BL, class EP>
inline hipError_t zgemm_run(int precision, int tile, size_t grid, hipStream_t st, const SC& sc, const AL& a, const BL& b,
                            const EP& e) {
    if (grid == 0 || sc.nitems == 0) return hipSuccess;
    switch (tile) {
    case TILE_32x32: zgemm_go<1, 1>(precision, (unsigned)grid, st, sc, a, b, e); break;
    case TILE_64x32: zgemm_go<2, 1>(precision, (unsigned)grid, st, sc, a, b, e); break;
    case TILE_32x64: zgemm_go<1, 2>(precision, (unsigned)grid, st, sc, a, b, e); break;
    default: zgemm_go<2, 2>(precision, (unsigned)grid, st, sc, a, b, e); break;
    }
    return hipGetLastError();
}

// launch helper of the single-problem GEMMs (64 x 64 tiles, one workgroup per item): without a
// block list the items cover nbatch batches of tm x tn tiles (rounded up to a multiple of 8
// batches); with one, nlist workgroups run the listed blocks.
inline GemmSchedOne sched_one(int M, int N, int K, int tm, int tn, int nbatch, const int* blist, size_t nlist,
                              int pnw = 1, int pnp = 1) {
    const size_t nitems = blist ? nlist : nbatch < 8 ? (size_t)nbatch * tm * tn : (size_t)((nbatch + 7) / 8) * 8 * tm * tn;
    return GemmSchedOne{M, N, K, tm, tn, nbatch, (int)nitems, blist, pnw, pnp};
}
template <class AL, class BL, class EP>
inline hipError_t zgemm_launch(int precision, int nbatch, hipStream_t st, int M, int N, int K, int tm, int tn,
                               const int* blist, size_t nlist, const AL& a, const BL& b, const EP& e) {
    const GemmSchedOne sc = sched_one(M, N, K, tm, tn, nbatch, blist, nlist);
    const size_t nitems = (size_t)sc.nitems;
    if (nitems == 0) return hipSuccess;
    if (precision == 1) zgemm_tiles_t<true, 2, 2><<<(unsigned)nitems, 256, 0, st>>>(sc, a, b, e);
    else zgemm_tiles_t<false, 2, 2><<<(unsigned)nitems, 256, 0, st>>>(sc, a, b, e);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// K3-class BSE (src/BSEa/BSEa_K3.jl; NL src/nonlocal/BSEa/BSEa_K3.jl).  One GEMM launch covers
// up to three channels; batch b = ((chan slot) * NP + P) * NW3 + iΩ.  The right operand R of the
// full BSE is materialised by k_k3_R (with its column sums), so both loaders are plain loads.
struct K3Geom {
    int n3b, n3f, NP;
    int nWP, nVP;   // bubble mesh N's
    int n2b, n2f;   // K2 box (rank-1 same-channel K2 term)
    __device__ __forceinline__ int NV() const { return 2 * n3f; }
    __device__ __forceinline__ int NW() const { return 2 * n3b - 1; }
    __device__ __forceinline__ size_t pidx(int m, int w, int p) const {
        return (size_t)(w + nVP) + (size_t)(2 * nVP) * ((size_t)(m + nWP - 1) + (size_t)(2 * nWP - 1) * p);
    }
};

// --- BSE_L_K3!: A = Γr[ν, ω, Ω], B = Π0(Ω, ω, P) F0r[ω, νp, Ω], FL.γr.K3 = T s A B
struct LK3Chan {
    const double2* G;      // Γr cache
    const double2* F0r;    // F0r cache
    const double2* Pi0;    // Π0pp or Π0ph
    double2* out;          // FL.γr.K3
    const double2* outa;   // FL.γa.K3 (t channel, in-launch post-processing)
    double scale;          // T * s
    int tpost;             // t channel post-processing inside the epilogue
    // FP32 contraction of the lattice solver: the epilogue also writes BSE_K3!'s right operand
    //   R(ω, νp) = s1 [(Π-Π0)(ω) F0r[ω, νp] + Π(ω) Γ~[ω, νp]] + s2 Π(ω) L~(ω, νp) 1(ω~ in box)
    // rounded to FP32 (L~ = the unpost-processed result at row ω~, t: 2 γt - γa of the finished
    // values); R32 null: FL.γr.K3 only
    float2* R32;
    const double2* Pi;     // Πpp or Πph
    const double2* Gam;    // p: Γpx [ω, νp, Ω]; a, t: Γr transposed [ω, ν, Ω]
    double s1, s2;
    int ch;
    // FL.K3 kept in FP32 (the first half of the FL.K3 array's bytes): out32 replaces out, and the t
    // post-processing reads the finished a channel from outa32
    float2* out32;
    const float2* outa32;
};
// Loaders and epilogues resolve their batch (channel slot, Ω, P) once per workgroup with bind(b),
// which returns base pointers for that batch; the per-element calls do only the in-matrix offset.
// The slot is selected by branches, never by indexing the kernel-argument descriptors.
struct LK3_Aop {
    K3Geom g; LK3Chan c0, c1, c2;
    struct Bound {
        const double2* G; int NV;
        __device__ __forceinline__ double2 operator()(int i, int k) const { return G[(size_t)i + (size_t)NV * k]; }
    };
    __device__ __forceinline__ Bound one(const LK3Chan& c, int bb) const {
        const int NV = g.NV();
        return Bound{c.G + (size_t)NV * NV * (bb % g.NW()), NV};
    }
    __device__ __forceinline__ Bound bind(int b) const {
        const int nb = g.NW() * g.NP, s = b / nb, bb = b - s * nb;
        return s == 0 ? one(c0, bb) : s == 1 ? one(c1, bb) : one(c2, bb);
    }
};
// B element of BSE_L_K3!, Π0(ω) F0r[ω, νp], without FMA contraction (the momentum-batched kernel
// forms the same products and must round them alike)
__device__ __forceinline__ double2 lk3_b(double2 p, double2 f) {
#pragma clang fp contract(off)
    return cmk(p.x * f.x - p.y * f.y, p.x * f.y + p.y * f.x);
}
struct LK3_Bop {
    K3Geom g; LK3Chan c0, c1, c2;
    struct Bound {
        const double2* pi0; const double2* F0r; int NV;   // pi0[k] = Π0(Ω, ω = k - n3f, P)
        __device__ __forceinline__ double2 operator()(int k, int j) const {
            return lk3_b(pi0[k], F0r[(size_t)k + (size_t)NV * j]);
        }
        // the two loads held raw by the prefetch (see LoaderRaw)
        struct Raw { double2 p, f; };
        __device__ __forceinline__ Raw raw(int k, int j) const { return Raw{pi0[k], F0r[(size_t)k + (size_t)NV * j]}; }
        __device__ __forceinline__ double2 combine(const Raw& x) const { return lk3_b(x.p, x.f); }
    };
    __device__ __forceinline__ Bound one(const LK3Chan& c, int bb) const {
        const int NV = g.NV(), NW = g.NW();
        const int iW = bb % NW, p = bb / NW, m = iW - (g.n3b - 1);
        return Bound{c.Pi0 + g.pidx(m, -g.n3f, p), c.F0r + (size_t)NV * NV * iW, NV};
    }
    __device__ __forceinline__ Bound bind(int b) const {
        const int nb = g.NW() * g.NP, s = b / nb, bb = b - s * nb;
        return s == 0 ? one(c0, bb) : s == 1 ? one(c1, bb) : one(c2, bb);
    }
};
// BSE_L_K3! epilogue.  RO: also BSE_K3!'s FP32 right operand (LK3Chan::R32); the plain form keeps
// the epilogue's registers low for the other paths.
// (no FMA contraction in these epilogues: the t post-processing (γt + γa)/2 of the scaled value must
// round as the separate post-processing pass does)
struct LK3BoundPlain {
    double2* out; const double2* outa; double scale; int tpost; int NV;
    __device__ __forceinline__ void operator()(int i, int j, double2 v) const {
#pragma clang fp contract(off)
        const size_t o = (size_t)i + (size_t)NV * j;
        double xr = v.x * scale, xi = v.y * scale;
        if (tpost) {
            const double2 a = outa[o];
            xr = (xr + a.x) * 0.5;
            xi = (xi + a.y) * 0.5;
        }
        out[o] = cmk(xr, xi);
    }
};
// BSE_L_K3! epilogue arithmetic of the right-operand form for one element: FL = T s v (t: then the
// post-processing with the finished a channel), and the right operand R at the element's R row
struct LK3RIn { double2 f0, gm, P1, P0, oa; };
__device__ __forceinline__ void lk3r_elem(double2 v, double scale, const LK3RIn& q, bool hasL, int tpost, double s1,
                                          double s2, float2& r32, double& xr, double& xi) {
#pragma clang fp contract(off)
    xr = v.x * scale;
    xi = v.y * scale;
    const double dr = q.P1.x - q.P0.x, di = q.P1.y - q.P0.y;
    double rr = (dr * q.f0.x - di * q.f0.y) + (q.P1.x * q.gm.x - q.P1.y * q.gm.y);
    double ri = (dr * q.f0.y + di * q.f0.x) + (q.P1.x * q.gm.y + q.P1.y * q.gm.x);
    rr *= s1;
    ri *= s1;
    if (hasL) {
        rr += s2 * (q.P1.x * xr - q.P1.y * xi);
        ri += s2 * (q.P1.x * xi + q.P1.y * xr);
    }
    r32 = make_float2((float)rr, (float)ri);
    if (tpost) {
        xr = (xr + q.oa.x) * 0.5;
        xi = (xi + q.oa.y) * 0.5;
    }
}
struct LK3BoundR {
    double2* out; const double2* outa; double scale; int tpost; int NV;
    float2* out32; const float2* outa32;
    float2* R32; const double2* pi; const double2* pi0; const double2* F0r; const double2* Gam;
    double s1, s2; int m, n3f, ch;
    // The element at row i is L~ at the R row whose crossing is row i (p: ω = Ω - ν - 1, an
    // involution; a, t: ω = ν).  For p, the R row ω = ν itself has no L~ term when its crossing
    // lies outside the box; that row is written here instead (every R row exactly once).
    __device__ __forceinline__ int r_row(int i, bool& hasL) const {
        hasL = true;
        if (ch != PCH) return i;
        const int wc = m - (i - n3f) - 1;
        hasL = in_f(wc, n3f);
        return hasL ? wc + n3f : i;
    }
    __device__ __forceinline__ void operator()(int i, int j, double2 v) const {
        bool hasL;
        const int k = r_row(i, hasL);
        const size_t kj = (size_t)k + (size_t)NV * j;
        LK3RIn q;
        q.f0 = F0r[kj];
        q.gm = Gam[kj];
        q.P1 = pi[k];
        q.P0 = pi0[k];
        if (tpost) {
            const size_t o = (size_t)i + (size_t)NV * j;
            if (outa32) { const float2 a = outa32[o]; q.oa = cmk(a.x, a.y); }
            else q.oa = outa[o];
        } else {
            q.oa = czero();
        }
        float2 r32;
        double xr, xi;
        lk3r_elem(v, scale, q, hasL, tpost, s1, s2, r32, xr, xi);
        R32[(size_t)k + (size_t)NV * j] = r32;
        if (out32) out32[(size_t)i + (size_t)NV * j] = make_float2((float)xr, (float)xi);
        else out[(size_t)i + (size_t)NV * j] = cmk(xr, xi);
    }
};
template <bool RO>
struct LK3_EpT {
    K3Geom g; LK3Chan c0, c1, c2;
    typedef typename std::conditional<RO, LK3BoundR, LK3BoundPlain>::type Bound;
    __device__ __forceinline__ Bound one(const LK3Chan& c, int bb) const {
        const int NV = g.NV(), NW = g.NW();
        const size_t nn = (size_t)NV * NV, off = nn * bb;
        Bound r;
        r.out = c.out + off;
        r.outa = c.tpost ? c.outa + off : nullptr;
        r.scale = c.scale; r.tpost = c.tpost; r.NV = NV;
        if constexpr (RO) {
            r.out32 = c.out32 ? c.out32 + off : nullptr;
            r.outa32 = (c.tpost && c.outa32) ? c.outa32 + off : nullptr;
            const int iW = bb % NW, p = bb / NW, m = iW - (g.n3b - 1);
            r.R32 = c.R32 + off;
            r.pi = c.Pi + g.pidx(m, -g.n3f, p);
            r.pi0 = c.Pi0 + g.pidx(m, -g.n3f, p);
            r.F0r = c.F0r + nn * iW;
            r.Gam = c.Gam + nn * iW;
            r.s1 = c.s1; r.s2 = c.s2; r.m = m; r.n3f = g.n3f; r.ch = c.ch;
        }
        return r;
    }
    __device__ __forceinline__ Bound bind(int b) const {
        const int nb = g.NW() * g.NP, s = b / nb, bb = b - s * nb;
        return s == 0 ? one(c0, bb) : s == 1 ? one(c1, bb) : one(c2, bb);
    }
};
typedef LK3_EpT<false> LK3_Ep;
typedef LK3_EpT<true> LK3R_Ep;

// BSE_L_K3! of the FP32 lattice contraction with its momentum-independent operands kept on chip: one
// workgroup per (channel slot, Ω, LK3P_G momenta).  Γr(Ω), the left operand rounded to FP32, stays in
// LDS; each thread keeps F0r(Ω) at its output positions in registers, so a momentum reads only its Π0
// and Π rows (and Γ~(Ω), from L2, in the epilogue) and forms B = Π0 F0r into LDS.  The tile is zgemm_tiles_t's 64 x 64 (four
// waves of 32 x 32, the same MFMA sequence and K order) and the epilogue is LK3BoundR's, so every
// element rounds exactly as in the GEMM engine.  FL.K3 in FP32 (out32) and the FP32 right operand
// (R32) are written; Nν3 <= 64.
constexpr int LK3P_G = 8;
__global__ __launch_bounds__(256, 2) void k_lk3_pbatch(K3Geom g, LK3Chan c0, LK3Chan c1, int nslot) {
    typedef MfmaT<true> MT;
    typedef MT::Acc Acc;
    __shared__ float Gs[2][64][65];   // [re/im][k][m] = Γr(m, k) in FP32
    __shared__ float Bs[2][64][65];   // [re/im][k][n] = B(k, n) of the current momentum
    const int NV = g.NV(), NW = g.NW(), NP = g.NP, n3f = g.n3f;
    const int npg = (NP + LK3P_G - 1) / LK3P_G, ntask = nslot * NW * npg;
    // tasks in contiguous ranges per XCD (blockIdx % 8), momentum group fastest: the blocks of one XCD
    // share Ω and its momentum-independent operands in that XCD's L2; a smaller grid loops over them
    const int perx = (ntask + 7) / 8, nbx = (int)(gridDim.x / 8);
    for (int kt = (int)(blockIdx.x / 8); kt < perx; kt += nbx) {
    const int task = (int)(blockIdx.x % 8) * perx + kt;
    if (task >= ntask) break;   // uniform over the block
    const int pg = task % npg, sw = task / npg, iW = sw % NW, s = sw / NW;
    const int m = iW - (g.n3b - 1);
    const size_t nn = (size_t)NV * NV;
    const double2* G = (s == 0 ? c0.G : c1.G) + nn * iW;
    const double2* F0r = (s == 0 ? c0.F0r : c1.F0r) + nn * iW;
    const double2* Gam = (s == 0 ? c0.Gam : c1.Gam) + nn * iW;
    const double2* Pi0 = s == 0 ? c0.Pi0 : c1.Pi0;
    const double2* Pi = s == 0 ? c0.Pi : c1.Pi;
    float2* out32 = s == 0 ? c0.out32 : c1.out32;
    const float2* outa32 = s == 0 ? c0.outa32 : c1.outa32;
    float2* R32 = s == 0 ? c0.R32 : c1.R32;
    const double scale = s == 0 ? c0.scale : c1.scale, s1 = s == 0 ? c0.s1 : c1.s1, s2 = s == 0 ? c0.s2 : c1.s2;
    const int tpost = s == 0 ? c0.tpost : c1.tpost, ch = s == 0 ? c0.ch : c1.ch;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave % 2) * 32, wn = (wave / 2) * 32;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int mm = e % 64, k = e / 64;
        const double2 v = (mm < NV && k < NV) ? G[(size_t)mm + (size_t)NV * k] : czero();
        Gs[0][k][mm] = (float)v.x;
        Gs[1][k][mm] = (float)v.y;
    }
    for (int e = tid; e < 64 * 65; e += 256) {   // entries outside the box stay zero
        (&Bs[0][0][0])[e] = 0.f;
        (&Bs[1][0][0])[e] = 0.f;
    }
    // output positions: rows wm + 16 i + (lane & 15), columns wn + 16 j + 4 (lane >> 4) + r; the R row of
    // row i is its crossing (p: Ω - ω - 1, an involution of the box rows; a, t: the row itself)
    int row[2], krow[2];
    bool hasL[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        row[i] = wm + 16 * i + (lane & 15);
        const int wc = m - (row[i] - n3f) - 1;
        hasL[i] = ch != PCH || in_f(wc, n3f);
        krow[i] = (ch == PCH && hasL[i]) ? wc + n3f : row[i];
    }
    double2 f0[2][2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = wn + 16 * j + 4 * (lane >> 4) + r;
                f0[i][j][r] = (row[i] < NV && col < NV) ? F0r[(size_t)krow[i] + (size_t)NV * col] : czero();
            }
    const int pbeg = pg * LK3P_G, pend = min(NP, pbeg + LK3P_G);
    // the Π and Π0 entries of the thread's two R rows, one momentum ahead
    double2 P1[2], P0[2], P1n[2], P0n[2];
    auto prow = [&](int p, double2* a, double2* b) {
        const size_t pr = g.pidx(m, -n3f, p);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            a[i] = row[i] < NV ? Pi[pr + krow[i]] : czero();
            b[i] = row[i] < NV ? Pi0[pr + krow[i]] : czero();
        }
    };
    prow(pbeg, P1n, P0n);
    for (int p = pbeg; p < pend; ++p) {
        const size_t off = nn * ((size_t)iW + (size_t)NW * p);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            P1[i] = P1n[i];
            P0[i] = P0n[i];
        }
        if (p + 1 < pend) prow(p + 1, P1n, P0n);
        __syncthreads();   // Gs written (first momentum) / Bs read by the previous momentum's MFMAs
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = wn + 16 * j + 4 * (lane >> 4) + r;
                    if (row[i] < NV && col < NV) {
                        const double2 b = lk3_b(P0[i], f0[i][j][r]);
                        Bs[0][krow[i]][col] = (float)b.x;
                        Bs[1][krow[i]][col] = (float)b.y;
                    }
                }
        __syncthreads();
        Acc accr[2][2], acci[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                accr[i][j] = Acc{0, 0, 0, 0};
                acci[i][j] = Acc{0, 0, 0, 0};
            }
        // 4-deep steps over the K range padded to GBK, as the GEMM engine runs them; each step's LDS
        // fragments are read one step ahead so that their latency overlaps the MFMAs
        const int nst = ((NV + GBK - 1) / GBK) * (GBK / 4), cc = lane & 15;
        float fa[2][4], fb[2][4];   // [buffer][ar0, ar1, ai0, ai1] / [br0, br1, bi0, bi1]
        auto frag = [&](int st, float* a, float* b) {
            const int kr = 4 * st + (lane >> 4);
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                a[i] = Gs[0][kr][wm + 16 * i + cc];
                a[2 + i] = Gs[1][kr][wm + 16 * i + cc];
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                b[j] = Bs[0][kr][wn + 16 * j + cc];
                b[2 + j] = Bs[1][kr][wn + 16 * j + cc];
            }
        };
        auto step = [&](const float* a, const float* b) {
            const float nai[2] = {-a[2], -a[3]};
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    accr[i][j] = MT::mfma(b[j], a[i], accr[i][j]);
                    acci[i][j] = MT::mfma(b[2 + j], a[i], acci[i][j]);
                }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    accr[i][j] = MT::mfma(b[2 + j], nai[i], accr[i][j]);
                    acci[i][j] = MT::mfma(b[j], a[2 + i], acci[i][j]);
                }
        };
        frag(0, fa[0], fb[0]);
        for (int st = 0; st < nst; st += 2) {   // nst is a multiple of 4
            frag(st + 1, fa[1], fb[1]);
            step(fa[0], fb[0]);
            if (st + 2 < nst) frag(st + 2, fa[0], fb[0]);
            step(fa[1], fb[1]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = wn + 16 * j + 4 * (lane >> 4) + r;
                    if (row[i] >= NV || col >= NV) continue;
                    const size_t o = off + (size_t)row[i] + (size_t)NV * col;
                    LK3RIn q;
                    q.f0 = f0[i][j][r];
                    q.gm = Gam[(size_t)krow[i] + (size_t)NV * col];   // momentum-independent: an L2 hit
                    q.P1 = P1[i];
                    q.P0 = P0[i];
                    if (tpost) { const float2 a = outa32[o]; q.oa = cmk(a.x, a.y); }
                    else q.oa = czero();
                    float2 r32;
                    double xr, xi;
                    lk3r_elem(cmk((double)accr[i][j][r], (double)acci[i][j][r]), scale, q, hasL[i], tpost, s1, s2,
                              r32, xr, xi);
                    R32[off + (size_t)krow[i] + (size_t)NV * col] = r32;
                    out32[o] = make_float2((float)xr, (float)xi);
                }
    }
    __syncthreads();   // the next task rewrites Gs and Bs
    }
}

// Nν3 = 64: the same tile with the momenta software-pipelined inside each workgroup.  While the MFMAs
// of momentum p run (sixteen 4-deep steps), every step also forms B elements of momentum p + 1 (into
// the second LDS buffer) and finishes epilogue elements of momentum p - 1 (from its saved
// accumulators), so that their VALU work and stores overlap the matrix core's.  Eight waves of 16 x 32
// (two per SIMD); F0r and Γ~ at a thread's 8 positions stay in registers, its Π row and the t
// post-processing's FL.a entries are loaded one momentum ahead.  Every element sees the MFMA sequence
// and epilogue arithmetic of k_lk3_pbatch and the GEMM engine: the results are equal bit for bit.
// TP: the launch of the t channel (post-processing with the finished a channel)
template <bool TP>
__global__ __launch_bounds__(512, 1) void k_lk3_pb64(K3Geom g, LK3Chan c0, LK3Chan c1, int nslot) {
    typedef MfmaT<true> MT;
    typedef MT::Acc Acc;
    constexpr int NV = 64, NST = 16, NE = 8;
    __shared__ float Gs[2][64][65];      // [re/im][k][m] = Γr(m, k) in FP32
    __shared__ float Bs[2][2][64][65];   // [buffer][re/im][k][n]
    const int NW = g.NW(), NP = g.NP, n3f = g.n3f;
    const int npg = (NP + LK3P_G - 1) / LK3P_G, ntask = nslot * NW * npg;
    // tasks (slot, Ω, momentum group) in contiguous ranges per XCD (blockIdx % 8), momentum group fastest,
    // so that the blocks of one XCD share Ω and its momentum-independent operands in its L2; a grid of
    // fewer blocks than tasks loops over them
    const int perx = (ntask + 7) / 8, nbx = (int)(gridDim.x / 8);
    for (int k = (int)(blockIdx.x / 8); k < perx; k += nbx) {
    const int task = (int)(blockIdx.x % 8) * perx + k;
    if (task >= ntask) break;   // uniform over the block
    const int pg = task % npg, sw = task / npg, iW = sw % NW, s = sw / NW;
    const int m = iW - (g.n3b - 1);
    const size_t nn = (size_t)NV * NV;
    const double2* G = (s == 0 ? c0.G : c1.G) + nn * iW;
    const double2* F0r = (s == 0 ? c0.F0r : c1.F0r) + nn * iW;
    const double2* Gam = (s == 0 ? c0.Gam : c1.Gam) + nn * iW;
    const double2* Pi0 = s == 0 ? c0.Pi0 : c1.Pi0;
    const double2* Pi = s == 0 ? c0.Pi : c1.Pi;
    float2* out32 = s == 0 ? c0.out32 : c1.out32;
    const float2* outa32 = s == 0 ? c0.outa32 : c1.outa32;
    float2* R32 = s == 0 ? c0.R32 : c1.R32;
    const double scale = s == 0 ? c0.scale : c1.scale, s1 = s == 0 ? c0.s1 : c1.s1, s2 = s == 0 ? c0.s2 : c1.s2;
    constexpr int tpost = TP;
    const int ch = s == 0 ? c0.ch : c1.ch;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave % 4) * 16, wn = (wave / 4) * 32, cc = lane & 15;
    for (int e = tid; e < 64 * 64; e += 512) {
        const double2 v = G[e];   // e = m + 64 k
        Gs[0][e / 64][e % 64] = (float)v.x;
        Gs[1][e / 64][e % 64] = (float)v.y;
    }
    // the thread's row and its R row (p: the crossing Ω - ω - 1, an involution of the box rows)
    const int row = wm + (lane & 15);
    const int wcr = m - (row - n3f) - 1;
    const bool hasL = ch != PCH || in_f(wcr, n3f);
    const int krow = (ch == PCH && hasL) ? wcr + n3f : row;
    // element e of the thread: column wn + 16 (e / 4) + 4 (lane / 16) + e % 4
    auto ecol = [&](int e) { return wn + 16 * (e >> 2) + 4 * (lane >> 4) + (e & 3); };
    double2 f0[NE], gm[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const size_t kj = (size_t)krow + (size_t)NV * ecol(e);
        f0[e] = F0r[kj];
        gm[e] = Gam[kj];
    }
    const int pbeg = pg * LK3P_G, n = min(NP, pbeg + LK3P_G) - pbeg;
    // Π1, Π0 of the R row per momentum: p - 1 (epilogue), p, p + 1 (B staging), p + 2 (in flight)
    double2 Pp1, Pp0, Pc1, Pc0, Pn1, Pn0, Pq1, Pq0;
    auto prow = [&](int p, double2& a, double2& b) {
        const size_t pr = g.pidx(m, -n3f, p);
        a = Pi[pr + krow];
        b = Pi0[pr + krow];
    };
    prow(pbeg, Pc1, Pc0);
    if (n > 1) prow(pbeg + 1, Pn1, Pn0);
    auto stage = [&](int buf, int e, const double2& P0) {
        const double2 b = lk3_b(P0, f0[e]);
        Bs[buf][0][krow][ecol(e)] = (float)b.x;
        Bs[buf][1][krow][ecol(e)] = (float)b.y;
    };
#pragma unroll
    for (int e = 0; e < NE; ++e) stage(0, e, Pc0);
    float2 oac[TP ? NE : 1], oan[TP ? NE : 1];   // FL.a entries of the epilogue's / the current momentum
    Acc accr[2], acci[2], par[2], pai[2];   // [column block]
    auto epilogue = [&](int e, size_t off, const double2& P1, const double2& P0) {
        const int j = e >> 2, r = e & 3, col = ecol(e);
        LK3RIn q;
        q.f0 = f0[e];
        q.gm = gm[e];
        q.P1 = P1;
        q.P0 = P0;
        if constexpr (TP) q.oa = cmk(oac[e].x, oac[e].y);
        else q.oa = czero();
        float2 r32;
        double xr, xi;
        lk3r_elem(cmk((double)par[j][r], (double)pai[j][r]), scale, q, hasL, tpost, s1, s2, r32, xr, xi);
        R32[off + (size_t)krow + (size_t)NV * col] = r32;
        out32[off + (size_t)row + (size_t)NV * col] = make_float2((float)xr, (float)xi);
    };
    __syncthreads();
    // one momentum: its MFMAs, with B of the next (STAGE) and the epilogue of the previous (EPI) in the
    // same straight-line steps (one element of each every other step)
    auto momentum = [&](int t, auto stage_c, auto epi_c) {
        constexpr bool STAGE = decltype(stage_c)::value, EPI = decltype(epi_c)::value;
        const int p = pbeg + t, cur = t & 1;
        const size_t offc = nn * ((size_t)iW + (size_t)NW * p), offp = offc - nn * (size_t)NW;
        if (t + 2 < n) prow(p + 2, Pq1, Pq0);
        if constexpr (TP) {
#pragma unroll
            for (int e = 0; e < NE; ++e) oan[e] = outa32[offc + (size_t)row + (size_t)NV * ecol(e)];
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            accr[j] = Acc{0, 0, 0, 0};
            acci[j] = Acc{0, 0, 0, 0};
        }
        float fa[2][2], fb[2][4];   // [buffer][ar, ai] / [br0, br1, bi0, bi1]
        auto frag = [&](int st, float* a, float* b) {
            const int kr = 4 * st + (lane >> 4);
            a[0] = Gs[0][kr][wm + cc];
            a[1] = Gs[1][kr][wm + cc];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                b[j] = Bs[cur][0][kr][wn + 16 * j + cc];
                b[2 + j] = Bs[cur][1][kr][wn + 16 * j + cc];
            }
        };
        frag(0, fa[0], fb[0]);
#pragma unroll
        for (int st = 0; st < NST; ++st) {
            if (st + 1 < NST) frag(st + 1, fa[(st + 1) & 1], fb[(st + 1) & 1]);
            const float* a = fa[st & 1];
            const float* b = fb[st & 1];
            const float nai = -a[1];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                accr[j] = MT::mfma(b[j], a[0], accr[j]);
                acci[j] = MT::mfma(b[2 + j], a[0], acci[j]);
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                accr[j] = MT::mfma(b[2 + j], nai, accr[j]);
                acci[j] = MT::mfma(b[j], a[1], acci[j]);
            }
            if (st % 2 == 0) {
                if constexpr (STAGE) stage(cur ^ 1, st / 2, Pn0);
                if constexpr (EPI) epilogue(st / 2, offp, Pp1, Pp0);
            }
        }
        // rotate: this momentum's accumulators and Π rows become the next epilogue's
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            par[j] = accr[j];
            pai[j] = acci[j];
        }
        Pp1 = Pc1; Pp0 = Pc0;
        Pc1 = Pn1; Pc0 = Pn0;
        Pn1 = Pq1; Pn0 = Pq0;
        if constexpr (TP) {
#pragma unroll
            for (int e = 0; e < NE; ++e) oac[e] = oan[e];
        }
        __syncthreads();   // B of p + 1 complete; the reads of Bs[cur] done before it is restaged
    };
    typedef std::integral_constant<bool, true> Yes;
    typedef std::integral_constant<bool, false> No;
    if (n == 1) {
        momentum(0, No{}, No{});
    } else {
        momentum(0, Yes{}, No{});
        for (int t = 1; t < n - 1; ++t) momentum(t, Yes{}, Yes{});
        momentum(n - 1, No{}, Yes{});
    }
    const size_t offl = nn * ((size_t)iW + (size_t)NW * (pbeg + n - 1));
#pragma unroll
    for (int e = 0; e < NE; ++e) epilogue(e, offl, Pp1, Pp0);
    }
}

// --- BSE_K3!:
//   A(ν, ω)  = cache_Fr = Fbase_r[ν, ω, Ω] + K2r(Ω, ν, P) + K3r(Ω, ν, ω, P)   (t: 2·(t) - (a))
//   R(ω, νp) = s1 [(Π-Π0)(ω) F0r[ω, νp] + Π(ω) Γ~(ω, νp)] + s2 Π(ω) L~(ω, νp) 1(ω~ in box)
//   K3 = T (A_3 R + K2r(ν) Σ_ω R(ω, νp)) + L_self      (A_3: A without the ω-independent K2 part)
struct K3Chan {
    int ch;
    int full;              // Fbase is the whole cache_Fr (one momentum): no K2r/K3r terms here
    const double2* Fbase;  // [ν, ω, Ω]
    const double2* K3r;    // target vertex K3 of this channel [ν, νp, Ω, P]
    const double2* K3a;    // target vertex K3 of the a channel (t channel combination)
    const double2* K2r;    // target vertex K2 [Ω, ν, P]
    const double2* K2a;
    const double2* R;      // [ω, νp, Ω, P]
    const double2* csum;   // [νp, Ω, P]
    const double2* Lr;     // FL.γr.K3
    const double2* La;     // FL.γa.K3
    double2* out;          // Fbuff.γr.K3
    const double2* outa;   // Fbuff.γa.K3 (in-launch t post-processing)
    int tpost;
    // lattice solver, FP32 contraction: both operands materialised in FP32 (A32 = the A-loader's
    // element, R32 = the right operand), each rounded as the FP32 GEMM would round it
    float2* A32;
    const float2* R32;
    // A32 holds the K2r(ν) term as well (csum null: no rank-1 term in the epilogue), read from
    // K2row[i + Nν3 (iΩ + NΩ3 P)] = K2r(Ω, ν = i - n3f, P) (t: 2 K2t - K2a), written by k_k2_rows
    double2* K2row;
    // l32: Lr / La hold FP32 values (FL.K3 in its FP32 form, LK3Chan::out32); k32: so do K3r / K3a
    // and out / outa (the vertex and Fbuff of the FP32 lattice iteration).  Flags rather than second
    // pointers keep the launch arguments (scalar registers) of the three channels small.
    int l32, k32;
};
// FULL: the cache holds the whole cache_Fr (one momentum, a single load per element)
template <bool FULL>
struct K3_AopT {
    K3Geom g; double T; K3Chan c0, c1, c2;
    struct Bound {
        const double2* Fb; const double2* K3r; const double2* K3a; int t; int NV;
        __device__ __forceinline__ double2 operator()(int i, int k) const {
            const size_t o = (size_t)i + (size_t)NV * k;
            const double2 x = Fb[o];
            if (FULL) return x;
            if (t) return cadd(x, csub(cscale(K3r[o], 2.0), K3a[o]));
            return cadd(x, K3r[o]);
        }
    };

    __device__ __forceinline__ Bound one(const K3Chan& c, int bb) const {
        const int NV = g.NV();
        const size_t nn = (size_t)NV * NV;
        const int t = (c.ch == TCH);
        return Bound{c.Fbase + nn * (bb % g.NW()), FULL ? nullptr : c.K3r + nn * bb,
                     (FULL || !t) ? nullptr : c.K3a + nn * bb, t, NV};
    }
    // (the lattice form's three loads are combined in the prefetch: holding them raw costs more
    // occupancy than it gains)
    __device__ __forceinline__ Bound bind(int b) const {
        const int nb = g.NW() * g.NP, s = b / nb, bb = b - s * nb;
        return s == 0 ? one(c0, bb) : s == 1 ? one(c1, bb) : one(c2, bb);
    }
};
struct K3_Bop {
    K3Geom g; double T; K3Chan c0, c1, c2;
    struct Bound {
        const double2* R; int NV;
        __device__ __forceinline__ double2 operator()(int k, int j) const { return R[(size_t)k + (size_t)NV * j]; }
    };
    __device__ __forceinline__ Bound one(const K3Chan& c, int bb) const {
        const int NV = g.NV();
        return Bound{c.R + (size_t)NV * NV * bb, NV};
    }
    __device__ __forceinline__ Bound bind(int b) const {
        const int nb = g.NW() * g.NP, s = b / nb, bb = b - s * nb;
        return s == 0 ? one(c0, bb) : s == 1 ? one(c1, bb) : one(c2, bb);
    }
};
// FP32 operands of the lattice BSE_K3! (FP32 contraction): single 8-byte loads, held raw by the
// prefetch (LoaderRaw) and widened only to be rounded back at the LDS store (exact)
struct K3_Aop32 {
    K3Geom g; K3Chan c0, c1, c2;
    struct Bound {
        const float2* A; int NV;
        typedef float2 Raw;
        __device__ __forceinline__ Raw raw(int i, int k) const { return A[(size_t)i + (size_t)NV * k]; }
        __device__ __forceinline__ double2 combine(const Raw& x) const { return cmk(x.x, x.y); }
        __device__ __forceinline__ double2 operator()(int i, int k) const { return combine(raw(i, k)); }
    };
    __device__ __forceinline__ Bound one(const K3Chan& c, int bb) const {
        const int NV = g.NV();
        return Bound{c.A32 + (size_t)NV * NV * bb, NV};
    }
    __device__ __forceinline__ Bound bind(int b) const {
        const int nb = g.NW() * g.NP, s = b / nb, bb = b - s * nb;
        return s == 0 ? one(c0, bb) : s == 1 ? one(c1, bb) : one(c2, bb);
    }
};
struct K3_Bop32 {
    K3Geom g; K3Chan c0, c1, c2;
    struct Bound {
        const float2* R; int NV;
        typedef float2 Raw;
        __device__ __forceinline__ Raw raw(int k, int j) const { return R[(size_t)k + (size_t)NV * j]; }
        __device__ __forceinline__ double2 combine(const Raw& x) const { return cmk(x.x, x.y); }
        __device__ __forceinline__ double2 operator()(int k, int j) const { return combine(raw(k, j)); }
    };
    __device__ __forceinline__ Bound one(const K3Chan& c, int bb) const {
        const int NV = g.NV();
        return Bound{c.R32 + (size_t)NV * NV * bb, NV};
    }
    __device__ __forceinline__ Bound bind(int b) const {
        const int nb = g.NW() * g.NP, s = b / nb, bb = b - s * nb;
        return s == 0 ? one(c0, bb) : s == 1 ? one(c1, bb) : one(c2, bb);
    }
};
// A32 of the lattice BSE_K3!: the K3_AopT<false> element rounded to FP32.  Grid: x over (channel
// slot, Ω, chunk of K3_PCH momenta), y over chunks of 256 of the Nν3 x Nν3 elements; a thread
// loads its P-independent Fbase element once and takes it through the chunk's momenta, four
// momenta's loads in flight together.
// With a momentum list `pl` (npl entries) only the listed momenta are built (the batches that hold
// symmetry-class representatives); otherwise all NP.
constexpr int K3_PCH = 8;
__global__ __launch_bounds__(256) void k_k3_A32(K3Geom g, K3Chan c0, K3Chan c1, K3Chan c2, const int* __restrict__ pl,
                                                int npl) {
    const int NV = g.NV(), NW = g.NW();
    const int nn = NV * NV, npc = (npl + K3_PCH - 1) / K3_PCH;
    const int o = blockIdx.y * 256 + threadIdx.x;
    if (o >= nn) return;
    const int s = (int)blockIdx.x / (NW * npc), rest = (int)blockIdx.x - s * NW * npc;
    const int iW = rest % NW, p0 = (rest / NW) * K3_PCH;
    const K3Chan& c = s == 0 ? c0 : s == 1 ? c1 : c2;
    const bool t = c.ch == TCH;
    const double2 x = c.Fbase[(size_t)o + (size_t)nn * iW];
    const int pend = min(p0 + K3_PCH, npl);
    const int i = o % NV;   // the element's row ν = i - n3f
    for (int pq = p0; pq < pend; pq += 4) {
        double2 k3[4], ka[4], kr[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int ip = pq + u;
            if (ip < pend) {
                const int p = pl ? pl[ip] : ip;
                const size_t r = (size_t)o + (size_t)nn * (iW + (size_t)NW * p);
                if (c.k32) {   // the vertex K3 held in FP32
                    const float2 a = reinterpret_cast<const float2*>(c.K3r)[r];
                    k3[u] = cmk(a.x, a.y);
                    if (t) { const float2 b = reinterpret_cast<const float2*>(c.K3a)[r]; ka[u] = cmk(b.x, b.y); }
                } else {
                    k3[u] = c.K3r[r];
                    if (t) ka[u] = c.K3a[r];
                }
                kr[u] = c.K2row ? c.K2row[(size_t)i + (size_t)NV * (iW + (size_t)NW * p)] : czero();
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int ip = pq + u;
            if (ip < pend) {
                const int p = pl ? pl[ip] : ip;
                double2 v = t ? cadd(x, csub(cscale(k3[u], 2.0), ka[u])) : cadd(x, k3[u]);
                if (c.K2row) v = cadd(v, kr[u]);
                c.A32[(size_t)o + (size_t)nn * (iW + (size_t)NW * p)] = make_float2((float)v.x, (float)v.y);
            }
        }
    }
}
// The same for the three slots (p, a, t) of one launch in one pass: a thread builds the element of
// all three, so the a channel's K3 is read once for its own A and for t's 2 K3t - K3a (each element
// computed exactly as above).  K32: the vertex K3 held in FP32.
template <bool K32>
__global__ __launch_bounds__(256) void k_k3_A32_pat(K3Geom g, K3Chan cp, K3Chan ca, K3Chan ct, const int* __restrict__ pl,
                                                    int npl) {
    typedef typename std::conditional<K32, float2, double2>::type KT;
    const int NV = g.NV(), NW = g.NW();
    const int nn = NV * NV;
    const int o = blockIdx.y * 256 + threadIdx.x;
    if (o >= nn) return;
    const int iW = (int)blockIdx.x % NW, p0 = ((int)blockIdx.x / NW) * K3_PCH;
    const size_t fo = (size_t)o + (size_t)nn * iW;
    const double2 xp = cp.Fbase[fo], xa = ca.Fbase[fo], xt = ct.Fbase[fo];
    const KT* kp = reinterpret_cast<const KT*>(cp.K3r);
    const KT* ka = reinterpret_cast<const KT*>(ca.K3r);
    const KT* kt = reinterpret_cast<const KT*>(ct.K3r);
    const int pend = min(p0 + K3_PCH, npl);
    const int i = o % NV;   // the element's row ν = i - n3f
    for (int pq = p0; pq < pend; pq += 2) {
        KT lp[2], la[2], lt[2];
        double2 rp[2], ra[2], rt[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (pq + u < pend) {
                const int p = pl ? pl[pq + u] : pq + u;
                const size_t r = (size_t)o + (size_t)nn * (iW + (size_t)NW * p);
                lp[u] = kp[r];
                la[u] = ka[r];
                lt[u] = kt[r];
                const size_t q = (size_t)i + (size_t)NV * (iW + (size_t)NW * p);
                rp[u] = cp.K2row ? cp.K2row[q] : czero();
                ra[u] = ca.K2row ? ca.K2row[q] : czero();
                rt[u] = ct.K2row ? ct.K2row[q] : czero();
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (pq + u < pend) {
                const int p = pl ? pl[pq + u] : pq + u;
                const size_t w = (size_t)o + (size_t)nn * (iW + (size_t)NW * p);
                const double2 vp0 = cmk(lp[u].x, lp[u].y), va0 = cmk(la[u].x, la[u].y), vt0 = cmk(lt[u].x, lt[u].y);
                double2 vp = cadd(xp, vp0), va = cadd(xa, va0), vt = cadd(xt, csub(cscale(vt0, 2.0), va0));
                if (cp.K2row) vp = cadd(vp, rp[u]);
                if (ca.K2row) va = cadd(va, ra[u]);
                if (ct.K2row) vt = cadd(vt, rt[u]);
                cp.A32[w] = make_float2((float)vp.x, (float)vp.y);
                ca.A32[w] = make_float2((float)va.x, (float)va.y);
                ct.A32[w] = make_float2((float)vt.x, (float)vt.y);
            }
        }
    }
}
// K2row of the FP32 lattice A operand (K3Chan::K2row) for every channel slot: one thread per
// (ν, Ω, P) of the K3 box, coalesced writes along ν
__global__ __launch_bounds__(256) void k_k2_rows(K3Geom g, K3Chan c0, K3Chan c1, K3Chan c2, int nchan) {
    const int NV = g.NV(), NW = g.NW(), NW2 = 2 * g.n2b - 1, NV2 = 2 * g.n2f;
    const size_t per = (size_t)NV * NW * g.NP;
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= per * nchan) return;
    const int s = (int)(e / per);
    const size_t r = e - (size_t)s * per;
    const int i = (int)(r % NV), iW = (int)((r / NV) % NW), p = (int)(r / ((size_t)NV * NW));
    const K3Chan& c = s == 0 ? c0 : s == 1 ? c1 : c2;
    const size_t q = (size_t)(iW - (g.n3b - 1) + g.n2b - 1) + (size_t)NW2 * ((size_t)(i - g.n3f + g.n2f) + (size_t)NV2 * p);
    double2 v = c.K2r[q];
    if (c.ch == TCH) v = csub(cscale(v, 2.0), c.K2a[q]);
    c.K2row[r] = v;
}
// FORM: 0 FL.K3, F.K3 and Fbuff.K3 all FP64; 1 all FP32 (the FP32 lattice iteration); 2 as each
// channel's l32 / k32 flags say.  The fixed forms compile one load and one store form per array.
template <int FORM>
struct K3_EpT {
    K3Geom g; double T; K3Chan c0, c1, c2;
    struct Bound {
        const double2* k2r; const double2* k2a;    // k2r[NW2 * i] = K2r(Ω, ν = i - n3f, P)
        const double2* csum;
        // FL.K3 of the channel and of a (FP32 values when l32); Fbuff.K3 of the channel and the
        // finished a channel (FP32 when o32): one address per array whatever its form
        const void* L; const void* La;
        void* out; const void* outa;
        double T; int t, tpost, full, NV, NW2, l32, o32;
        __device__ __forceinline__ static double2 ld(const void* p, size_t o, int f32) {
            if (f32) { const float2 v = static_cast<const float2*>(p)[o]; return cmk(v.x, v.y); }
            return static_cast<const double2*>(p)[o];
        }
        // the element's loads (FL.K3, and for t FL.a and the finished a channel) issued together,
        // before any arithmetic
        struct Pre { double2 self, oa; };
        __device__ __forceinline__ Pre pre(int i, int j) const {
            const size_t o = (size_t)i + (size_t)NV * j;
            Pre q;
            q.self = ld(L, o, l32);
            if (t) q.self = csub(cscale(q.self, 2.0), ld(La, o, l32));
            q.oa = tpost ? ld(outa, o, o32) : czero();
            return q;
        }
        __device__ __forceinline__ void post(int i, int j, double2 v, const Pre& q) const {
            if (csum) {
                double2 k2v = k2r[(size_t)NW2 * i];
                if (t) k2v = csub(cscale(k2v, 2.0), k2a[(size_t)NW2 * i]);
                v = cadd(v, cmul(k2v, csum[j]));
            }
            const size_t o = (size_t)i + (size_t)NV * j;
            double2 x = cadd(cscale(v, T), q.self);
            if (tpost) x = cscale(cadd(x, q.oa), 0.5);
            if (o32) static_cast<float2*>(out)[o] = make_float2((float)x.x, (float)x.y);
            else static_cast<double2*>(out)[o] = x;
        }
        __device__ __forceinline__ void operator()(int i, int j, double2 v) const { post(i, j, v, pre(i, j)); }
    };
    __device__ __forceinline__ Bound one(const K3Chan& c, int bb) const {
        const int NV = g.NV(), NW = g.NW();
        const int iW = bb % NW, p = bb / NW, m = iW - (g.n3b - 1);
        const int NW2 = 2 * g.n2b - 1, NV2 = 2 * g.n2f;
        const size_t nn = (size_t)NV * NV, off = nn * bb;
        const int t = (c.ch == TCH);
        // K2r(Ω, ν, P) at ν = i - n3f: index (m + n2b - 1) + NW2 (i - n3f + n2f + NV2 p)
        const size_t k2o = (size_t)(m + g.n2b - 1) + (size_t)NW2 * ((size_t)(g.n2f - g.n3f) + (size_t)NV2 * p);
        Bound r;
        const bool cs = !c.full && c.csum;
        r.k2r = cs ? c.K2r + k2o : nullptr;
        r.k2a = (cs && t) ? c.K2a + k2o : nullptr;
        r.csum = cs ? c.csum + (size_t)NV * bb : nullptr;
        // element offsets in the array's own form (8-byte FP32 or 16-byte FP64 elements)
        r.l32 = FORM == 2 ? c.l32 : FORM;
        r.o32 = FORM == 2 ? c.k32 : FORM;
        const size_t lo = off * (r.l32 ? 8 : 16), oo = off * (r.o32 ? 8 : 16);
        r.L = reinterpret_cast<const char*>(c.Lr) + lo;
        r.La = t ? reinterpret_cast<const char*>(c.La) + lo : nullptr;
        r.out = reinterpret_cast<char*>(c.out) + oo;
        r.outa = c.tpost ? reinterpret_cast<const char*>(c.outa) + oo : nullptr;
        r.T = T; r.t = t; r.tpost = c.tpost; r.full = c.full; r.NV = NV; r.NW2 = NW2;
        return r;
    }
    __device__ __forceinline__ Bound bind(int b) const {
        const int nb = g.NW() * g.NP, s = b / nb, bb = b - s * nb;
        return s == 0 ? one(c0, bb) : s == 1 ? one(c1, bb) : one(c2, bb);
    }
};
typedef K3_EpT<2> K3_Ep;   // the solver groups' (one functor type for every member)

// right operand of BSE_K3! and its column sums; one wave per (νp, Ω, P), lanes over ω
struct K3RChan {
    int ch;
    double s1, s2;
    const double2* Pi;
    const double2* Pi0;
    const double2* F0r;   // [ω, νp, Ω]
    const double2* Gam;   // p: Γpx [ω, νp, Ω]; a, t: Γr transposed, [ω, ν, Ω]
    const double2* Lr;    // FL.γr.K3 [ν, νp, Ω, P]
    const double2* La;    // FL.γa.K3
    double2* R;           // [ω, νp, Ω, P]
    double2* csum;        // [νp, Ω, P]
    float2* R32;          // when set: R rounded to FP32 goes here instead of R
    const float2* Lr32;   // FL.K3 in FP32 (its FP32 form), read instead of Lr / La when set
    const float2* La32;
};
__device__ __forceinline__ double2 k3r_fl(const K3RChan& c, bool a, size_t i) {
    if (c.Lr32) { const float2 v = (a ? c.La32 : c.Lr32)[i]; return cmk(v.x, v.y); }
    return (a ? c.La : c.Lr)[i];
}
// K3R_COLS columns νp = j0 .. j0 + K3R_COLS - 1 per wave: the Π loads are shared and the columns'
// loads issued together
constexpr int K3R_COLS = 4;
template <int CH>
__device__ __forceinline__ void k3_R_columns(const K3Geom& g, const K3RChan& c, int j0, int bb, int lane) {
    constexpr int NC = K3R_COLS;
    const int NV = g.NV(), NW = g.NW();
    const int iW = bb % NW, p = bb / NW;
    const int m = iW - (g.n3b - 1);
    double cr[NC], ci[NC];
#pragma unroll
    for (int u = 0; u < NC; ++u) cr[u] = ci[u] = 0.0;
    for (int k = lane; k < NV; k += WAVE) {
        const int w = k - g.n3f;
        const size_t pi = g.pidx(m, w, p);
        const double2 P1 = c.Pi[pi], P0 = c.Pi0[pi];
        const int w_ = crossing(m, w, CH);
        const bool inb = in_f(w_, g.n3f);
        double2 f0[NC], gm[NC], L[NC], La[NC];
#pragma unroll
        for (int u = 0; u < NC; ++u) {
            const int j = j0 + u;
            if (j < NV) {
                const size_t kj = (size_t)k + (size_t)NV * ((size_t)j + (size_t)NV * iW);
                f0[u] = c.F0r[kj];
                gm[u] = c.Gam[kj];
                if (inb) {
                    const size_t lo = (size_t)(w_ + g.n3f) + (size_t)NV * ((size_t)j + (size_t)NV * bb);
                    L[u] = k3r_fl(c, false, lo);
                    if (CH == TCH) La[u] = k3r_fl(c, true, lo);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < NC; ++u) {
            const int j = j0 + u;
            if (j >= NV) continue;
            double2 x = cscale(cadd(cmul(csub(P1, P0), f0[u]), cmul(P1, gm[u])), c.s1);
            if (inb) {
                double2 Lx = L[u];
                if (CH == TCH) Lx = csub(cscale(Lx, 2.0), La[u]);
                x = cadd(x, cscale(cmul(P1, Lx), c.s2));
            }
            const size_t ro = (size_t)k + (size_t)NV * ((size_t)j + (size_t)NV * bb);
            if (c.R32) c.R32[ro] = make_float2((float)x.x, (float)x.y);
            else c.R[ro] = x;
            cr[u] += x.x; ci[u] += x.y;
        }
    }
    if (c.csum) {
#pragma unroll
        for (int u = 0; u < NC; ++u) {
            const double sr = wave_sum(cr[u]), si = wave_sum(ci[u]);
            if (lane == 0 && j0 + u < NV) c.csum[(size_t)(j0 + u) + (size_t)NV * bb] = cmk(sr, si);
        }
    }
}
__device__ __forceinline__ void k3_R_dispatch(const K3Geom& g, const K3RChan& c, int j0, int bb, int lane) {
    if (c.ch == PCH) k3_R_columns<PCH>(g, c, j0, bb, lane);
    else if (c.ch == TCH) k3_R_columns<TCH>(g, c, j0, bb, lane);
    else k3_R_columns<ACH>(g, c, j0, bb, lane);
}
// Nν3 <= 64 KPL (KPL ω per lane): a wave takes NC columns νp of one Ω through a chunk of K3_PCH
// momenta, loading the P-independent F0r and Γ columns once
template <int CH, int KPL, int NC>
__device__ __forceinline__ void k3_R_columns_pchunk(const K3Geom& g, const K3RChan& c, int j0, int iW, int p0, int lane,
                                                    const int* pl, int npl) {
    const int NV = g.NV(), NW = g.NW();
    const int m = iW - (g.n3b - 1);
    int w_[KPL];
    bool kin[KPL], inb[KPL];
    double2 f0[KPL][NC], gm[KPL][NC];
#pragma unroll
    for (int q = 0; q < KPL; ++q) {
        const int k = lane + WAVE * q;
        kin[q] = k < NV;
        w_[q] = crossing(m, k - g.n3f, CH);
        inb[q] = kin[q] && in_f(w_[q], g.n3f);
#pragma unroll
        for (int u = 0; u < NC; ++u) {
            const int j = j0 + u;
            if (kin[q] && j < NV) {
                const size_t kj = (size_t)k + (size_t)NV * ((size_t)j + (size_t)NV * iW);
                f0[q][u] = c.F0r[kj];
                gm[q][u] = c.Gam[kj];
            } else {
                f0[q][u] = gm[q][u] = czero();
            }
        }
    }
    const int pend = min(p0 + K3_PCH, npl);
    for (int ip = p0; ip < pend; ++ip) {
        const int p = pl ? pl[ip] : ip;
        const size_t bb = (size_t)iW + (size_t)NW * p;
        double2 P1[KPL], P0[KPL], L[KPL][NC], La[KPL][NC];
#pragma unroll
        for (int q = 0; q < KPL; ++q) {
            P1[q] = P0[q] = czero();
            if (kin[q]) {
                const size_t pi = g.pidx(m, lane + WAVE * q - g.n3f, p);
                P1[q] = c.Pi[pi];
                P0[q] = c.Pi0[pi];
            }
#pragma unroll
            for (int u = 0; u < NC; ++u) {
                L[q][u] = La[q][u] = czero();
                if (inb[q] && j0 + u < NV) {
                    const size_t lo = (size_t)(w_[q] + g.n3f) + (size_t)NV * ((size_t)(j0 + u) + (size_t)NV * bb);
                    L[q][u] = k3r_fl(c, false, lo);
                    if (CH == TCH) La[q][u] = k3r_fl(c, true, lo);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < NC; ++u) {
            const int j = j0 + u;
            if (j >= NV) continue;   // uniform over the wave
            double cr = 0.0, ci = 0.0;
#pragma unroll
            for (int q = 0; q < KPL; ++q) {
                if (!kin[q]) continue;
                double2 x = cscale(cadd(cmul(csub(P1[q], P0[q]), f0[q][u]), cmul(P1[q], gm[q][u])), c.s1);
                if (inb[q]) {
                    double2 Lx = L[q][u];
                    if (CH == TCH) Lx = csub(cscale(Lx, 2.0), La[q][u]);
                    x = cadd(x, cscale(cmul(P1[q], Lx), c.s2));
                }
                const size_t ro = (size_t)(lane + WAVE * q) + (size_t)NV * ((size_t)j + (size_t)NV * bb);
                if (c.R32) c.R32[ro] = make_float2((float)x.x, (float)x.y);
                else c.R[ro] = x;
                cr += x.x; ci += x.y;
            }
            if (c.csum) {
                const double sr = wave_sum(cr), si = wave_sum(ci);
                if (lane == 0) c.csum[(size_t)j + (size_t)NV * bb] = cmk(sr, si);
            }
        }
    }
}
// (momenta as in k_k3_A32: the list pl, or all NP when pl is null)
template <int KPL, int NC>
__global__ __launch_bounds__(256, 4) void k_k3_R_pchunk(K3Geom g, K3RChan c0, K3RChan c1, K3RChan c2, int nchan,
                                                     const int* __restrict__ pl, int npl) {
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    const int NV = 2 * g.n3f, NW = 2 * g.n3b - 1, nq = (NV + NC - 1) / NC;
    const int npc = (npl + K3_PCH - 1) / K3_PCH;
    if (wid >= nq * NW * npc * nchan) return;
    int r = wid;
    const int q = r % nq; r /= nq;
    const int iW = r % NW; r /= NW;
    const int pc = r % npc, s = r / npc;
    const K3RChan& c = s == 0 ? c0 : s == 1 ? c1 : c2;
    if (c.ch == PCH) k3_R_columns_pchunk<PCH, KPL, NC>(g, c, q * NC, iW, pc * K3_PCH, lane, pl, npl);
    else if (c.ch == TCH) k3_R_columns_pchunk<TCH, KPL, NC>(g, c, q * NC, iW, pc * K3_PCH, lane, pl, npl);
    else k3_R_columns_pchunk<ACH, KPL, NC>(g, c, q * NC, iW, pc * K3_PCH, lane, pl, npl);
}
// one wave per K3R_COLS columns (νp, Ω, P) of a channel slot
__global__ __launch_bounds__(256) void k_k3_R(K3Geom g, K3RChan c0, K3RChan c1, K3RChan c2, int nchan) {
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    const int NV = 2 * g.n3f, nb = (2 * g.n3b - 1) * g.NP, nq = (NV + K3R_COLS - 1) / K3R_COLS;
    if (wid >= nq * nb * nchan) return;
    const int q = wid % nq, r = wid / nq, bb = r % nb, s = r / nb;
    if (s == 0) k3_R_dispatch(g, c0, q * K3R_COLS, bb, lane);
    else if (s == 1) k3_R_dispatch(g, c1, q * K3R_COLS, bb, lane);
    else k3_R_dispatch(g, c2, q * K3R_COLS, bb, lane);
}

// FL.K3 between its FP32 form (in the first half of the array's bytes) and FP64, through a temporary
__global__ void k_widen32(double2* __restrict__ dst, const float2* __restrict__ src, size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) { const float2 v = src[i]; dst[i] = cmk(v.x, v.y); }
}
__global__ void k_narrow32(float2* __restrict__ dst, const double2* __restrict__ src, size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) { const double2 v = src[i]; dst[i] = make_float2((float)v.x, (float)v.y); }
}

// t channel post-processing after a multi-channel launch: γt = (γt_d + γa) / 2
__global__ void k_post_t(double2* __restrict__ t, const double2* __restrict__ a, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) t[i] = cscale(cadd(t[i], a[i]), 0.5);
}

// dense column-major operands, batch b at offset b * stride (library self-test of the GEMM engine)
struct Dense_A {
    const double2* A; int lda; size_t stride;
    __device__ Dense_A bind(int b) const { Dense_A r = *this; r.A += stride * b; return r; }
    __device__ double2 operator()(int i, int k) const { return A[(size_t)i + (size_t)lda * k]; }
};
struct Dense_B {
    const double2* B; int ldb; size_t stride;
    __device__ Dense_B bind(int b) const { Dense_B r = *this; r.B += stride * b; return r; }
    __device__ double2 operator()(int k, int j) const { return B[(size_t)k + (size_t)ldb * j]; }
};
struct Dense_C {
    double2* C; int ldc; size_t stride;
    __device__ Dense_C bind(int b) const { Dense_C r = *this; r.C += stride * b; return r; }
    __device__ void operator()(int i, int j, double2 v) const { C[(size_t)i + (size_t)ldc * j] = v; }
};

// ---------------------------------------------------------------------------------------
// BSE weight on the bubble mesh for K1/K2 (src/BSEa/BSEa_K1.jl:38-43, BSEa_K2.jl:77-85):
//   W(ω, Ω, P) = (Π - Π0) F0(Ω, ω~, ∞, Ch, Sp) + Π FL(Ω, ω~, ∞, P, Ch, Sp),  ω~ = crossing(Ω, ω)
// W is stored on the band of 2 wb frequencies around ω = 0 that the gathers read (iw in
// [nVP - wb, nVP + wb), row stride 2 wb), and its far sums (the tail layout of k_tail_sums) are
// formed from the tail sums of Π and Π0: for |ω| > NK + |Ω| + 2 the factors F0(Ω, ω~, ∞) and
// FL(Ω, ω~, ∞) no longer depend on ω (|ω~| lies beyond every box, so only the bare vertex and the
// channel's K1(Ω) remain), so Σ_far W = F0_far Σ_far (Π - Π0) + FL_far Σ_far Π.  That holds for the
// tails the gathers use (lo <= nVP - NK - |Ω| - 2); entries nearer ω = 0 are not used.
// One thread per band entry and per tail entry of a row (2 wb + 2 per row, rows (Ω, P) in order).
struct WArgs {   // the arrays of one channel's weight
    double2* W;      // W on the band [2 wb, rows]
    double2* far;    // per row: F0 and FL at the left far point, then at the right one [4, rows]
    const double2* Pi; const double2* Pi0;
};
// The tail sums of W are not stored: at row r and tail entry k (the layout of k_tail_sums) they are
// (Πt - Π0t)·F0_far + Πt·FL_far with the tail sums Πt, Π0t of Π and Π0 and the row's far factors of
// that side (WTail::at).  One thread per band entry; the row's first also evaluates the far factors.
template <int CH>
__device__ __forceinline__ void bse_w_elem(size_t e, const WArgs& a, const VtxV& FL, const F0V& f0, int nWP, int nVP,
                                           int wb, int NK, size_t nrows) {
    constexpr int ch = CH, sp = ChSpin<CH>::SP;
    double2* __restrict__ W = a.W;
    const double2* __restrict__ Pi = a.Pi;
    const double2* __restrict__ Pi0 = a.Pi0;
    const int per = 2 * wb;
    const size_t row = e / per;
    if (row >= nrows) return;
    const int t = (int)(e - row * per);
    const int iW = (int)(row % (size_t)(2 * nWP - 1)), p = (int)(row / (size_t)(2 * nWP - 1));
    const int m = iW - (nWP - 1);
    {
        const int iw = nVP - wb + t, w = iw - nVP;
        const size_t o = (size_t)iw + (size_t)(2 * nVP) * row;
        const int wc = crossing(m, w, ch);
        const double2 f0r = f0.eval(m, wc, FDGA_INF, ch, sp);
        const double2 flr = FL.eval(m, wc, FDGA_INF, p, ch, sp, FL_ALL);
        const double2 pi = Pi[o];
        W[(size_t)t + (size_t)(2 * wb) * row] = cadd(cmul(csub(pi, Pi0[o]), f0r), cmul(pi, flr));
    }
    for (int side = t; side < 2; side += per) {   // the row's far factors, one side per thread
        const int wf = side ? NK + abs(m) + 3 : -(NK + abs(m) + 3);
        const int wc = crossing(m, wf, ch);
        a.far[4 * row + 2 * side] = f0.eval(m, wc, FDGA_INF, ch, sp);
        a.far[4 * row + 2 * side + 1] = FL.eval(m, wc, FDGA_INF, p, ch, sp, FL_ALL);
    }
}
template <int CH>
__global__ __launch_bounds__(256) void k_bse_w(WArgs a, VtxV FL, F0V f0, int nWP, int nVP, int wb, int NK, size_t nrows) {
    bse_w_elem<CH>(blockIdx.x * (size_t)blockDim.x + threadIdx.x, a, FL, f0, nWP, nVP, wb, NK, nrows);
}
// The same weights with one thread per (band entry, Ω) and a chunk of W_PCH momenta: the local
// reference vertex's factor F0(Ω, ω~, ∞) does not depend on P and is evaluated once for the chunk
// (the lattice solver; identical values to bse_w_elem).
constexpr int W_PCH = 8;
template <int CH>
__device__ __forceinline__ void bse_w_pchunk(size_t e, const WArgs& a, const VtxV& FL, const F0V& f0, int nWP, int nVP,
                                             int wb, int NK, int NP) {
    constexpr int ch = CH, sp = ChSpin<CH>::SP;
    double2* __restrict__ W = a.W;
    const double2* __restrict__ Pi = a.Pi;
    const double2* __restrict__ Pi0 = a.Pi0;
    const int per = 2 * wb, NWP = 2 * nWP - 1, npc = (NP + W_PCH - 1) / W_PCH;
    const size_t col = e / per;
    if (col >= (size_t)NWP * npc) return;
    const int t = (int)(e - col * per);
    const int iW = (int)(col % NWP), pc = (int)(col / NWP);
    const int m = iW - (nWP - 1);
    const int iw = nVP - wb + t, w = iw - nVP;
    const int wc = crossing(m, w, ch);
    const double2 f0r = f0.eval(m, wc, FDGA_INF, ch, sp);
    const int p0 = pc * W_PCH, p1 = min(p0 + W_PCH, NP);
#pragma unroll 4
    for (int p = p0; p < p1; ++p) {
        const size_t row = (size_t)iW + (size_t)NWP * p;
        const size_t o = (size_t)iw + (size_t)(2 * nVP) * row;
        const double2 flr = FL.eval(m, wc, FDGA_INF, p, ch, sp, FL_ALL);
        const double2 pi = Pi[o];
        W[(size_t)t + (size_t)(2 * wb) * row] = cadd(cmul(csub(pi, Pi0[o]), f0r), cmul(pi, flr));
    }
    // the rows' far factors (F0 and FL at the first ω beyond every box, on each side): one (side,
    // momentum) pair per thread of the column, so that no thread evaluates them in a chain
    for (int j = t; j < 2 * W_PCH; j += per) {
        const int side = j / W_PCH, p = p0 + j % W_PCH;
        if (p >= p1) continue;
        const int wf = side ? NK + abs(m) + 3 : -(NK + abs(m) + 3);
        const int wcf = crossing(m, wf, ch);
        const size_t row = (size_t)iW + (size_t)NWP * p;
        a.far[4 * row + 2 * side] = f0.eval(m, wcf, FDGA_INF, ch, sp);
        a.far[4 * row + 2 * side + 1] = FL.eval(m, wcf, FDGA_INF, p, ch, sp, FL_ALL);
    }
}
// the three channels' weights in one launch (nb blocks each, in the order p, t, a); pchunk: one
// thread per (band entry, Ω, chunk of W_PCH momenta)
__global__ __launch_bounds__(256) void k_bse_w3(WArgs ap, WArgs at, WArgs aa, VtxV FL, F0V f0, int nWP, int nVP, int wb,
                                                int NK, size_t nrows, unsigned nb, int pchunk, int NP) {
    const unsigned ch = blockIdx.x / nb;
    const size_t e = (blockIdx.x - ch * nb) * (size_t)blockDim.x + threadIdx.x;
    if (pchunk) {
        if (ch == 0) bse_w_pchunk<PCH>(e, ap, FL, f0, nWP, nVP, wb, NK, NP);
        else if (ch == 1) bse_w_pchunk<TCH>(e, at, FL, f0, nWP, nVP, wb, NK, NP);
        else bse_w_pchunk<ACH>(e, aa, FL, f0, nWP, nVP, wb, NK, NP);
        return;
    }
    if (ch == 0) bse_w_elem<PCH>(e, ap, FL, f0, nWP, nVP, wb, NK, nrows);
    else if (ch == 1) bse_w_elem<TCH>(e, at, FL, f0, nWP, nVP, wb, NK, nrows);
    else bse_w_elem<ACH>(e, aa, FL, f0, nWP, nVP, wb, NK, nrows);
}

// Tail sums of rows on the bubble ν-mesh (2 nVP entries, ω index iw = ω + nVP), kept on the band
// of half-width wb around ω = 0 that the gathers can ask for: per row r,
//   left[i]  = Σ_{k < nVP - wb + i} X[k]    (i = 0 .. wb)
//   right[j] = Σ_{k >= nVP + j} X[k]        (j = 0 .. wb)
// stored as T[r * 2 (wb + 1) + i] and T[r * 2 (wb + 1) + wb + 1 + j].
__device__ __forceinline__ void wave_incl_scan(double& r, double& i, int lane) {
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const double yr = __shfl_up(r, o, WAVE), yi = __shfl_up(i, o, WAVE);
        if (lane >= o) { r += yr; i += yi; }
    }
}
// One 256-thread block per row.  The parts outside the band are summed with a strided
// (coalesced) loop and a block reduction; the band is scanned in chunks of 256 entries (wave
// scans, then the preceding waves' totals), the left side from its outer end inwards and the
// right side from ω = wb - 1 down to 0.
__device__ __forceinline__ void tail_sums_row(const double2* __restrict__ X, double2* __restrict__ T, int nVP, int wb,
                                              size_t nrows, size_t row) {
    constexpr int NWV = 256 / WAVE;
    __shared__ double sh[4][NWV];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
    if (row >= nrows) return;
    const double2* x = X + row * (size_t)(2 * nVP);
    double2* tl = T + row * (size_t)(2 * (wb + 1));
    double2* tr = tl + (wb + 1);
    const int nout = nVP - wb;   // entries of each side outside the band
    double lr = 0.0, li = 0.0, rr = 0.0, ri = 0.0;
    int q = tid;
    for (; q + 3 * 256 < nout; q += 4 * 256) {   // four loads of each side in flight
        double2 a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            a[u] = x[q + 256 * u];
            b[u] = x[2 * nVP - 1 - q - 256 * u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) { lr += a[u].x; li += a[u].y; rr += b[u].x; ri += b[u].y; }
    }
    for (; q < nout; q += 256) {
        const double2 a = x[q], b = x[2 * nVP - 1 - q];
        lr += a.x; li += a.y; rr += b.x; ri += b.y;
    }
    lr = wave_sum(lr); li = wave_sum(li); rr = wave_sum(rr); ri = wave_sum(ri);
    if (lane == 0) { sh[0][wv] = lr; sh[1][wv] = li; sh[2][wv] = rr; sh[3][wv] = ri; }
    __syncthreads();
    double cl_r = 0.0, cl_i = 0.0, cr_r = 0.0, cr_i = 0.0;   // running sums carried across chunks
#pragma unroll
    for (int k = 0; k < NWV; ++k) { cl_r += sh[0][k]; cl_i += sh[1][k]; cr_r += sh[2][k]; cr_i += sh[3][k]; }
    if (tid == 0) { tl[0] = cmk(cl_r, cl_i); tr[wb] = cmk(cr_r, cr_i); }
    for (int c0 = 0; c0 < wb; c0 += 256) {
        __syncthreads();   // sh is reused
        const int i = c0 + tid;
        const double2 a = i < wb ? x[nout + i] : czero();
        const double2 b = i < wb ? x[nVP + wb - 1 - i] : czero();
        double ar = a.x, ai = a.y, br = b.x, bi = b.y;
        wave_incl_scan(ar, ai, lane);
        wave_incl_scan(br, bi, lane);
        if (lane == WAVE - 1) { sh[0][wv] = ar; sh[1][wv] = ai; sh[2][wv] = br; sh[3][wv] = bi; }
        __syncthreads();
        double pl_r = cl_r, pl_i = cl_i, pr_r = cr_r, pr_i = cr_i;
        for (int k = 0; k < wv; ++k) { pl_r += sh[0][k]; pl_i += sh[1][k]; pr_r += sh[2][k]; pr_i += sh[3][k]; }
        if (i < wb) {
            tl[i + 1] = cmk(pl_r + ar, pl_i + ai);
            tr[wb - 1 - i] = cmk(pr_r + br, pr_i + bi);
        }
#pragma unroll
        for (int k = 0; k < NWV; ++k) { cl_r += sh[0][k]; cl_i += sh[1][k]; cr_r += sh[2][k]; cr_i += sh[3][k]; }
    }
}
__global__ __launch_bounds__(256) void k_tail_sums(const double2* __restrict__ X, double2* __restrict__ T, int nVP, int wb,
                                                   size_t nrows) {
    tail_sums_row(X, T, nVP, wb, nrows, blockIdx.x);
}
// two arrays in one launch (blocks [0, nrows) the first)
__global__ __launch_bounds__(256) void k_tail_sums2(const double2* __restrict__ X0, double2* __restrict__ T0,
                                                    const double2* __restrict__ X1, double2* __restrict__ T1, int nVP, int wb,
                                                    size_t nrows) {
    if (blockIdx.x < nrows) tail_sums_row(X0, T0, nVP, wb, nrows, blockIdx.x);
    else tail_sums_row(X1, T1, nVP, wb, nrows, blockIdx.x - nrows);
}

// k_tail_sums for an array whose chunk sums C[P + NP (chunk + NC iΩ)] over NB consecutive ω
// were written by the lattice bubble: the parts outside the band are the sums of the chunks that
// lie wholly outside it plus the few entries of the chunks it cuts; the band as in k_tail_sums.
// Reads NC + 2 wb + O(NB) entries per row instead of 2 nVP.  A block takes the rows of 8
// consecutive momenta at one Ω: its chunk-sum reads cover whole 128-byte lines (8 momenta per
// line), the 32 thread groups' partial sums are added in a fixed order, and each wave then scans
// the band of two rows.
constexpr int TAIL_NPB = 8;
// Both arrays of a bubble (Πpp, Πph) in one launch: blocks [0, nblk) the first, the rest the second.
__global__ __launch_bounds__(256) void k_tail_sums_chunked(const double2* __restrict__ X0, const double2* __restrict__ C0,
                                                           double2* __restrict__ T0, const double2* __restrict__ X1,
                                                           const double2* __restrict__ C1, double2* __restrict__ T1,
                                                           unsigned nblk, int nVP, int wb, int NB, int NC, int NWP,
                                                           int NP) {
    __shared__ double2 red[2][32][TAIL_NPB + 1];
    __shared__ double2 far[2][TAIL_NPB];
    const bool second = blockIdx.x >= nblk;
    const double2* __restrict__ X = second ? X1 : X0;
    const double2* __restrict__ C = second ? C1 : C0;
    double2* __restrict__ T = second ? T1 : T0;
    const unsigned bid = blockIdx.x - (second ? nblk : 0u);
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
    const int iW = (int)(bid % NWP), p0 = (int)(bid / NWP) * TAIL_NPB;
    const int nout = nVP - wb;                 // far left: ω index < nout; far right: >= nVP + wb
    const int cl = nout / NB;                  // chunks [0, cl) lie wholly left of the band
    const int cr = (nVP + wb + NB - 1) / NB;   // chunks [cr, NC) lie wholly right of it
    {
        const int pl = tid % TAIL_NPB, cg = tid / TAIL_NPB, p = p0 + pl;
        double lr = 0.0, li = 0.0, rr = 0.0, ri = 0.0;
        if (p < NP) {
            const double2* cp = C + (size_t)p + (size_t)NP * (size_t)NC * iW;
            for (int c = cg; c < cl; c += 32) { const double2 v = cp[(size_t)NP * c]; lr += v.x; li += v.y; }
            for (int c = cr + cg; c < NC; c += 32) { const double2 v = cp[(size_t)NP * c]; rr += v.x; ri += v.y; }
        }
        red[0][cg][pl] = cmk(lr, li);
        red[1][cg][pl] = cmk(rr, ri);
    }
    __syncthreads();
    if (tid < 2 * TAIL_NPB) {
        const int side = tid / TAIL_NPB, pl = tid % TAIL_NPB;
        double sr = 0.0, si = 0.0;
        for (int g = 0; g < 32; ++g) { sr += red[side][g][pl].x; si += red[side][g][pl].y; }
        far[side][pl] = cmk(sr, si);
    }
    __syncthreads();
    for (int pl = wv; pl < TAIL_NPB; pl += 256 / WAVE) {
        const int p = p0 + pl;
        if (p >= NP) break;
        const size_t row = (size_t)iW + (size_t)NWP * p;
        const double2* x = X + row * (size_t)(2 * nVP);
        double2* tl = T + row * (size_t)(2 * (wb + 1));
        double2* tr = tl + (wb + 1);
        // the entries of the two cut chunks that lie outside the band
        double lr = 0.0, li = 0.0, rr = 0.0, ri = 0.0;
        for (int k = cl * NB + lane; k < nout; k += WAVE) { const double2 v = x[k]; lr += v.x; li += v.y; }
        for (int k = nVP + wb + lane; k < min(cr * NB, 2 * nVP); k += WAVE) { const double2 v = x[k]; rr += v.x; ri += v.y; }
        double cl_r = far[0][pl].x + wave_sum(lr), cl_i = far[0][pl].y + wave_sum(li);
        double cr_r = far[1][pl].x + wave_sum(rr), cr_i = far[1][pl].y + wave_sum(ri);
        if (lane == 0) { tl[0] = cmk(cl_r, cl_i); tr[wb] = cmk(cr_r, cr_i); }
        // the band: left from its outer end inwards, right from ω = wb - 1 down to 0
        for (int c0 = 0; c0 < wb; c0 += WAVE) {
            const int i = c0 + lane;
            const double2 a = i < wb ? x[nout + i] : czero();
            const double2 b = i < wb ? x[nVP + wb - 1 - i] : czero();
            double ar = a.x, ai = a.y, br = b.x, bi = b.y;
            wave_incl_scan(ar, ai, lane);
            wave_incl_scan(br, bi, lane);
            if (i < wb) {
                tl[i + 1] = cmk(cl_r + ar, cl_i + ai);
                tr[wb - 1 - i] = cmk(cr_r + br, cr_i + bi);
            }
            cl_r += __shfl(ar, WAVE - 1, WAVE); cl_i += __shfl(ai, WAVE - 1, WAVE);
            cr_r += __shfl(br, WAVE - 1, WAVE); cr_i += __shfl(bi, WAVE - 1, WAVE);
        }
    }
}
// far sums of one row outside the window [lo, hi]: Σ_{k < lo} and Σ_{k > hi}
// (nVP - wb <= lo <= nVP < hi + 1 <= nVP + wb)
__device__ __forceinline__ double2 tail_left(const double2* T, size_t row, int nVP, int wb, int lo) {
    return T[row * (size_t)(2 * (wb + 1)) + (lo - (nVP - wb))];
}
__device__ __forceinline__ double2 tail_right(const double2* T, size_t row, int nVP, int wb, int hi) {
    return T[row * (size_t)(2 * (wb + 1)) + (wb + 1) + (hi + 1 - nVP)];
}
// tail sums of a BSE weight W = (Π - Π0) F0 + Π FL, formed from the tail sums of Π and Π0 and the
// row's far factors (k_bse_w): Σ_far W = F0_far Σ_far (Π - Π0) + FL_far Σ_far Π
struct WTail {
    const double2* Pt; const double2* P0t; const double2* far;
    __device__ __forceinline__ double2 at(size_t row, int k, int wb) const {   // entry k of the row's tail layout
        const size_t o = row * (size_t)(2 * (wb + 1)) + k;
        const int side = k > wb ? 1 : 0;
        const double2 pt = Pt[o];
        return cadd(cmul(csub(pt, P0t[o]), far[4 * row + 2 * side]), cmul(pt, far[4 * row + 2 * side + 1]));
    }
    __device__ __forceinline__ double2 left(size_t row, int nVP, int wb, int lo) const { return at(row, lo - (nVP - wb), wb); }
    __device__ __forceinline__ double2 right(size_t row, int nVP, int wb, int hi) const {
        return at(row, (wb + 1) + (hi + 1 - nVP), wb);
    }
};

// Frequency window of the gather-GEMVs.  Every vertex term that depends on the summed
// frequency ω is supported on |ω| <= Wc = NK + |m| + |n| + 2, where NK bounds every box of the
// vertex (K1 boxes reach furthest: their bosonic argument is a sum or difference of Ω, ν, ω).
// Beyond the window the summand is T(ω_far) W(ω) with T(ω_far) the same value for every ω on
// that side, so Σ_far T W = T(ω_far) Σ_far W: only the weights are summed there.
__device__ __forceinline__ int window(int NK, int m, int n) { return NK + abs(m) + abs(n) + 2; }

// the two far-side summands (ω = -Wc-1 for lane 0, ω = Wc+1 for lane 1) are evaluated by one
// inlined evaluation on every lane and gathered with shuffles
__device__ __forceinline__ void far_pair(double2 t, double2& left, double2& right) {
    left = cmk(__shfl(t.x, 0, WAVE), __shfl(t.y, 0, WAVE));
    right = cmk(__shfl(t.x, 1, WAVE), __shfl(t.y, 1, WAVE));
}

// BSE_K1!: K1(Ω, P) = T s Σ_ω F(Ω, ∞, ω, P, Ch, Sp) W(ω).  One 256-thread block per output
// class (the K1 mesh has few outputs; four waves per output shorten each one's chain of gathers).
template <int CH>
__device__ __forceinline__ void gemv_k1_block(int wid, double2* __restrict__ K1, const double2* __restrict__ K1a,
                                              const double2* __restrict__ W, const WTail& Wtail,
                                              const VtxV& F, int n1, int nWP, int nVP, int wb, int NP, double scale, int NK,
                                              const SymCls& sym) {
    constexpr int SP = ChSpin<CH>::SP;
    constexpr int NT = 256;
    __shared__ double red[NT / WAVE][2];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
    const int NW1 = 2 * n1 - 1;
    if (wid >= sym.count(NW1 * NP)) return;
    const int o0 = sym.rep(wid);
    const int iW = o0 % NW1, p = o0 / NW1;
    const int m = iW - (n1 - 1);
    const int Wc = window(NK, m, 0);
    const size_t row = (size_t)(m + nWP - 1) + (size_t)(2 * nWP - 1) * p;
    const double2* Wr = W + (size_t)(2 * wb) * row + wb;   // Wr[ω], |ω| within the band
    const int lo = max(0, nVP - Wc), hi = min(2 * nVP - 1, nVP + Wc);
    double sr = 0.0, si = 0.0;
    for (int iw = lo + tid; iw <= hi; iw += NT) {
        double2 x = cmul(F.eval(m, FDGA_INF, iw - nVP, p, CH, SP, FL_ALL), Wr[iw - nVP]);
        sr += x.x; si += x.y;
    }
    sr = wave_sum(sr); si = wave_sum(si);
    if (lane == 0) { red[wv][0] = sr; red[wv][1] = si; }
    __syncthreads();
    if (wv != 0) return;
    const double tsr = red[0][0] + red[1][0] + red[2][0] + red[3][0];
    const double tsi = red[0][1] + red[1][1] + red[2][1] + red[3][1];
    double2 tl, tr;
    far_pair(F.eval(m, FDGA_INF, lane == 0 ? -Wc - 1 : Wc + 1, p, CH, SP, FL_ALL), tl, tr);
    // every lane of wave 0 holds the same sums; lanes write the class members
    double2 acc = cmk(tsr, tsi);
    if (lo > 0) acc = cadd(acc, cmul(tl, Wtail.left(row, nVP, wb, lo)));
    if (hi < 2 * nVP - 1) acc = cadd(acc, cmul(tr, Wtail.right(row, nVP, wb, hi)));
    const double2 raw = cscale(acc, scale);
    int b, e;
    sym.range(wid, b, e);
    for (int k = b + lane; k < e; k += WAVE) {
        const int code = sym.member(wid, k), o = code >> 2;
        double2 v = sym_apply(code & 3, raw);
        if (CH == TCH) v = cscale(cadd(v, K1a[o]), 0.5);
        K1[o] = v;
    }
}
template <int CH>
__global__ __launch_bounds__(256) void k_gemv_k1(double2* __restrict__ K1, const double2* __restrict__ K1a,
                                                 const double2* __restrict__ W, WTail Wtail, VtxV F,
                                                 int n1, int nWP, int nVP, int wb, int NP, double scale, int NK, SymCls sym) {
    gemv_k1_block<CH>(blockIdx.x, K1, K1a, W, Wtail, F, n1, nWP, nVP, wb, NP, scale, NK, sym);
}
// the p and a channels in one launch (blocks [0, nb) p, the rest a); t needs the finished a channel
__global__ __launch_bounds__(256) void k_gemv_k1_pa(double2* __restrict__ K1p, const double2* __restrict__ Wp,
                                                    WTail Wtp, double scalep, double2* __restrict__ K1a_,
                                                    const double2* __restrict__ Wa, WTail Wta, double scalea, VtxV F, int n1, int nWP, int nVP, int wb, int NP, int NK,
                                                    SymCls sym, int nb) {
    if ((int)blockIdx.x < nb) gemv_k1_block<PCH>(blockIdx.x, K1p, nullptr, Wp, Wtp, F, n1, nWP, nVP, wb, NP, scalep, NK, sym);
    else gemv_k1_block<ACH>(blockIdx.x - nb, K1a_, nullptr, Wa, Wta, F, n1, nWP, nVP, wb, NP, scalea, NK, sym);
}

// BSE_K2!: K2(Ω, ν, P) = T s Σ_ω [F(Ω,ν,ω) - F(Ω,∞,ω)] W(ω) + FL.K2 (t: 2 FLt - FLa).
// In the difference the channel's own K1(Ω) and K2(Ω, ω) cancel (the νInf evaluation keeps
// exactly those two classes, src/channel.jl:220-339), the channel's K2(Ω, ν) does not depend on
// ω and factors out as K2(Ω, ν) Σ_ω W(ω), and what depends on ω is the reference vertex part,
// the channel's own K3 and the two other channels at the converted arguments.
// the part without the channel's own K3: reference vertex and the two other channels, which the
// s-wave lattice vertex evaluates P-averaged (src/nonlocal/vertex.jl:216-379), so it does not
// depend on P
template <int CH>
__device__ __forceinline__ double2 k2x_psp(const VtxV& F, int m, int v, int w) {
    double2 x = csub(F.f0.eval_psp(m, v, w, CH), F.f0.eval_psp(m, FDGA_INF, w, CH));
    int a, b, c;
    if (CH != PCH) { convert(CH, PCH, m, v, w, a, b, c); x = cadd(x, F.g[PCH].eval(a, b, c, -1)); }
    if (CH != TCH) { convert(CH, TCH, m, v, w, a, b, c); x = cadd(x, F.g[TCH].eval(a, b, c, -1)); }
    if (CH != ACH) { convert(CH, ACH, m, v, w, a, b, c); x = cadd(x, F.g[ACH].eval(a, b, c, -1)); }
    return x;
}
template <int CH>
__device__ __forceinline__ double2 k2d_psp(const VtxV& F, int m, int v, int w, int p) {
    double2 x = k2x_psp<CH>(F, m, v, w);
    const ChanV& g = F.g[CH];
    if (in_b(m, g.n3b) && in_f(v, g.n3f) && in_f(w, g.n3f)) x = cadd(x, g.k3(m, v, w, p));
    return x;
}
// the ω-dependent summand in the channel's spin component (t: density, 2 F(t) - F(a))
template <int CH>
__device__ __forceinline__ double2 k2d_term(const VtxV& F, int m, int v, int w, int p) {
    if (CH == TCH) return csub(cscale(k2d_psp<TCH>(F, m, v, w, p), 2.0), k2d_psp<ACH>(F, m, v, w, p));
    return k2d_psp<CH>(F, m, v, w, p);
}
template <int CH>
__device__ __forceinline__ double2 k2d_rank1(const VtxV& F, int m, int v, int p) {
    if (CH == TCH) return csub(cscale(F.g[TCH].k2(m, v, p), 2.0), F.g[ACH].k2(m, v, p));
    return F.g[CH].k2(m, v, p);
}
template <int CH>
__device__ __forceinline__ void gemv_k2_wave(int wid, int lane, double2* __restrict__ K2, const double2* __restrict__ K2a,
                                             const double2* __restrict__ W, const WTail& Wtail,
                                             const VtxV& F, const ChanV& Lr, const ChanV& La, int n2b, int n2f, int nWP,
                                             int nVP, int wb, int NP, double scale, int NK, const SymCls& sym) {
    const int NW2 = 2 * n2b - 1, NV2 = 2 * n2f;
    if (wid >= sym.count(NW2 * NV2 * NP)) return;
    const int o0 = sym.rep(wid);
    const int iW = o0 % NW2, iv = (o0 / NW2) % NV2, p = o0 / (NW2 * NV2);
    const int m = iW - (n2b - 1), v = iv - n2f;
    const int Wc = window(NK, m, v);
    const size_t row = (size_t)(m + nWP - 1) + (size_t)(2 * nWP - 1) * p;
    const double2* Wr = W + (size_t)(2 * wb) * row + wb;   // Wr[ω], |ω| within the band
    // only the window is iterated for the summand; the far weights come from the tail sums
    const int lo = max(0, nVP - Wc), hi = min(2 * nVP - 1, nVP + Wc);
    double sr = 0.0, si = 0.0, mr = 0.0, mi = 0.0;
    for (int iw = lo + lane; iw <= hi; iw += WAVE) {
        const double2 wv = Wr[iw - nVP];
        double2 x = cmul(k2d_term<CH>(F, m, v, iw - nVP, p), wv);
        sr += x.x; si += x.y;
        mr += wv.x; mi += wv.y;
    }
    double2 tl, tr;
    far_pair(k2d_term<CH>(F, m, v, lane == 0 ? -Wc - 1 : Wc + 1, p), tl, tr);
    sr = wave_sum(sr); si = wave_sum(si);
    mr = wave_sum(mr); mi = wave_sum(mi);
    const double2 wl = lo > 0 ? Wtail.left(row, nVP, wb, lo) : czero();
    const double2 wr = hi < 2 * nVP - 1 ? Wtail.right(row, nVP, wb, hi) : czero();
    double2 acc = cadd(cadd(cmk(sr, si), cmul(tl, wl)), cmul(tr, wr));
    acc = cadd(acc, cmul(k2d_rank1<CH>(F, m, v, p), cadd(cmk(mr, mi), cadd(wl, wr))));
    const double2 raw = cscale(acc, scale);
    int b, e;
    sym.range(wid, b, e);
    for (int k = b + lane; k < e; k += WAVE) {
        const int code = sym.member(wid, k), o = code >> 2;
        double2 val = sym_apply(code & 3, raw);
        if (CH == TCH) {
            val = cadd(val, cscale(Lr.K2[o], 2.0));
            val = cadd(val, cscale(La.K2[o], -1.0));
            val = cscale(cadd(val, K2a[o]), 0.5);
        } else {
            val = cadd(val, Lr.K2[o]);
        }
        K2[o] = val;
    }
}
template <int CH>
__global__ __launch_bounds__(256) void k_gemv_k2(double2* __restrict__ K2, const double2* __restrict__ K2a,
                                                 const double2* __restrict__ W, WTail Wtail, VtxV F,
                                                 ChanV Lr, ChanV La, int n2b, int n2f, int nWP, int nVP, int wb, int NP,
                                                 double scale, int NK, SymCls sym) {
    gemv_k2_wave<CH>((blockIdx.x * blockDim.x + threadIdx.x) / WAVE, threadIdx.x & (WAVE - 1), K2, K2a, W, Wtail, F, Lr, La,
                     n2b, n2f, nWP, nVP, wb, NP, scale, NK, sym);
}

// BSE_K2! of the a and t channels in one pass.  The t channel runs in the density component,
// F(t, dSp) = 2 F(t, pSp) + F(t, xSp) = 2 F(t, pSp) - F(a, pSp) (src/vertex.jl:288-336), so the
// a-channel evaluation is shared:  K2a = T Σ fa Wa + FL.a.K2,
// K2t = (T (-1) Σ (2 ft - fa) Wt + 2 FL.t.K2 - FL.a.K2 + K2a) / 2   (src/BSE_templates.jl:95-110)
__device__ __forceinline__ void gemv_k2_at_wave(int wid, int lane, double2* __restrict__ K2a, double2* __restrict__ K2t,
                                                const double2* __restrict__ Wa, const double2* __restrict__ Wt,
                                                const WTail& Wat, const WTail& Wtt,
                                                const VtxV& F, const ChanV& La, const ChanV& Lt, int n2b, int n2f, int nWP,
                                                int nVP, int wb, int NP, double T, int NK, const SymCls& sym) {
    const int NW2 = 2 * n2b - 1, NV2 = 2 * n2f;
    if (wid >= sym.count(NW2 * NV2 * NP)) return;
    const int o0 = sym.rep(wid);
    const int iW = o0 % NW2, iv = (o0 / NW2) % NV2, p = o0 / (NW2 * NV2);
    const int m = iW - (n2b - 1), v = iv - n2f;
    const int Wc = window(NK, m, v);
    const size_t row = (size_t)(m + nWP - 1) + (size_t)(2 * nWP - 1) * p;
    const size_t roff = (size_t)(2 * wb) * row + wb;   // band rows, indexed by ω
    const double2* wa = Wa + roff;
    const double2* wt = Wt + roff;
    const int lo = max(0, nVP - Wc), hi = min(2 * nVP - 1, nVP + Wc);
    double ar = 0.0, ai = 0.0, tr = 0.0, ti = 0.0;
    double amr = 0.0, ami = 0.0, tmr = 0.0, tmi = 0.0;
    for (int iw = lo + lane; iw <= hi; iw += WAVE) {
        const int w = iw - nVP;
        const double2 fa = k2d_psp<ACH>(F, m, v, w, p);
        const double2 ft = k2d_psp<TCH>(F, m, v, w, p);
        const double2 wav = wa[w], wtv = wt[w];
        const double2 xa = cmul(fa, wav);
        const double2 xt = cmul(csub(cscale(ft, 2.0), fa), wtv);
        ar += xa.x; ai += xa.y; tr += xt.x; ti += xt.y;
        amr += wav.x; ami += wav.y; tmr += wtv.x; tmi += wtv.y;
    }
    // far summands: lane 0 at ω = -Wc-1, lane 1 at ω = Wc+1
    const int wf = lane == 0 ? -Wc - 1 : Wc + 1;
    const double2 fa = k2d_psp<ACH>(F, m, v, wf, p);
    const double2 ft = k2d_psp<TCH>(F, m, v, wf, p);
    double2 fal, far_, fdl, fdr;
    far_pair(fa, fal, far_);
    far_pair(csub(cscale(ft, 2.0), fa), fdl, fdr);
    ar = wave_sum(ar); ai = wave_sum(ai); tr = wave_sum(tr); ti = wave_sum(ti);
    amr = wave_sum(amr); ami = wave_sum(ami); tmr = wave_sum(tmr); tmi = wave_sum(tmi);
    const bool hl = lo > 0, hr = hi < 2 * nVP - 1;
    const double2 awl = hl ? Wat.left(row, nVP, wb, lo) : czero();
    const double2 awr = hr ? Wat.right(row, nVP, wb, hi) : czero();
    const double2 twl = hl ? Wtt.left(row, nVP, wb, lo) : czero();
    const double2 twr = hr ? Wtt.right(row, nVP, wb, hi) : czero();
    double2 A = cadd(cadd(cmk(ar, ai), cmul(fal, awl)), cmul(far_, awr));
    double2 D = cadd(cadd(cmk(tr, ti), cmul(fdl, twl)), cmul(fdr, twr));
    A = cadd(A, cmul(k2d_rank1<ACH>(F, m, v, p), cadd(cmk(amr, ami), cadd(awl, awr))));
    D = cadd(D, cmul(k2d_rank1<TCH>(F, m, v, p), cadd(cmk(tmr, tmi), cadd(twl, twr))));
    const double2 rawa = cscale(A, T), rawt = cscale(D, -T);
    int b, e;
    sym.range(wid, b, e);
    for (int k = b + lane; k < e; k += WAVE) {
        const int code = sym.member(wid, k), o = code >> 2;
        const double2 ka = cadd(sym_apply(code & 3, rawa), La.K2[o]);
        double2 kt = sym_apply(code & 3, rawt);
        kt = cadd(kt, cscale(Lt.K2[o], 2.0));
        kt = cadd(kt, cscale(La.K2[o], -1.0));
        K2a[o] = ka;
        K2t[o] = cscale(cadd(kt, ka), 0.5);
    }
}
// BSE_K2! of all three channels in one launch: waves [0, wp) the p classes (SGpp[2]), then the a and
// t classes (SGph[2], t finished from a in the same wave)
__global__ __launch_bounds__(256, 4) void k_gemv_k2_all(double2* __restrict__ K2p, const double2* __restrict__ Wp,
                                                     WTail Wtp, ChanV Lp, SymCls symp, int wp,
                                                     double2* __restrict__ K2a, double2* __restrict__ K2t,
                                                     const double2* __restrict__ Wa, const double2* __restrict__ Wt,
                                                     WTail Wat, WTail Wtt,
                                                     VtxV F, ChanV La, ChanV Lt, int n2b, int n2f, int nWP, int nVP, int wb,
                                                     int NP, double T, int NK, SymCls symat) {
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    if (wid < wp)
        gemv_k2_wave<PCH>(wid, lane, K2p, nullptr, Wp, Wtp, F, Lp, La, n2b, n2f, nWP, nVP, wb, NP, T, NK, symp);
    else
        gemv_k2_at_wave(wid - wp, lane, K2a, K2t, Wa, Wt, Wat, Wtt, F, La, Lt, n2b, n2f, nWP, nVP, wb, NP, T, NK, symat);
}
__global__ __launch_bounds__(256, 4) void k_gemv_k2_at(double2* __restrict__ K2a, double2* __restrict__ K2t,
                                                    const double2* __restrict__ Wa, const double2* __restrict__ Wt,
                                                    WTail Wat, WTail Wtt,
                                                    VtxV F, ChanV La, ChanV Lt, int n2b, int n2f, int nWP, int nVP,
                                                    int wb, int NP, double T, int NK, SymCls sym) {
    gemv_k2_at_wave((blockIdx.x * blockDim.x + threadIdx.x) / WAVE, threadIdx.x & (WAVE - 1), K2a, K2t, Wa, Wt, Wat, Wtt, F,
                    La, Lt, n2b, n2f, nWP, nVP, wb, NP, T, NK, sym);
}

// BSE_L_K2! of the local solver (src/BSEa/BSEa_K2.jl:1-40): Σ over the bubble mesh of
//   Γ(Ω, ν, ω) Π0(ω) F0(Ω, ω~, ∞).  Γ excludes F0 and the own channel, so it vanishes outside
//   the frequency window.  (The lattice solver's form is k_lk2nl_table + GEMM.)
template <int CH>
__device__ __forceinline__ void gemv_lk2_wave(int wid, int lane, double2* __restrict__ out, const double2* __restrict__ outa,
                                              const double2* __restrict__ Pi0, const VtxV& F, int n2b, int n2f, int nWP,
                                              int nVP, int NP, double scale, int NK, const SymCls& sym) {
    constexpr int SP = ChSpin<CH>::SP;
    constexpr int FLG = FL_ALL & ~FL_F0 & ~(FL_GP << CH);
    const int NW2 = 2 * n2b - 1, NV2 = 2 * n2f;
    if (wid >= sym.count(NW2 * NV2 * NP)) return;
    const int o0 = sym.rep(wid);
    const int iW = o0 % NW2, iv = (o0 / NW2) % NV2, p = o0 / (NW2 * NV2);
    const int m = iW - (n2b - 1), v = iv - n2f;
    const double2* P0 = Pi0 + (size_t)(2 * nVP) * ((size_t)(m + nWP - 1) + (size_t)(2 * nWP - 1) * p);
    double sr = 0.0, si = 0.0;
    const int Wc = window(NK, m, v);
    const int lo = max(0, nVP - Wc), hi = min(2 * nVP - 1, nVP + Wc);
    for (int q = lo + lane; q <= hi; q += WAVE) {
        const int w = q - nVP;
        double2 x = cmul(cmul(F.eval(m, v, w, p, CH, SP, FLG), P0[q]), F.f0.eval(m, crossing(m, w, CH), FDGA_INF, CH, SP));
        sr += x.x; si += x.y;
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    const double2 raw = cscale(cmk(sr, si), scale);
    int b, e;
    sym.range(wid, b, e);
    for (int k = b + lane; k < e; k += WAVE) {
        const int code = sym.member(wid, k), o = code >> 2;
        double2 val = sym_apply(code & 3, raw);
        if (CH == TCH) val = cscale(cadd(val, outa[o]), 0.5);
        out[o] = val;
    }
}
template <int CH>
__global__ __launch_bounds__(256) void k_gemv_lk2(double2* __restrict__ out, const double2* __restrict__ outa,
                                                  const double2* __restrict__ Pi0, VtxV F, int n2b, int n2f, int nWP,
                                                  int nVP, int NP, double scale, int NK, SymCls sym) {
    gemv_lk2_wave<CH>((blockIdx.x * blockDim.x + threadIdx.x) / WAVE, threadIdx.x & (WAVE - 1), out, outa, Pi0, F, n2b, n2f,
                      nWP, nVP, NP, scale, NK, sym);
}
// the p and a channels in one launch (waves [0, wp) p, the rest a); t needs the finished a channel
__global__ __launch_bounds__(256) void k_gemv_lk2_pa(double2* __restrict__ outp, const double2* __restrict__ Pi0p,
                                                     SymCls symp, double scalep, double2* __restrict__ outa_,
                                                     const double2* __restrict__ Pi0a, SymCls syma, double scalea, VtxV F,
                                                     int n2b, int n2f, int nWP, int nVP, int NP, int NK, int wp) {
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    if (wid < wp) gemv_lk2_wave<PCH>(wid, lane, outp, nullptr, Pi0p, F, n2b, n2f, nWP, nVP, NP, scalep, NK, symp);
    else gemv_lk2_wave<ACH>(wid - wp, lane, outa_, nullptr, Pi0a, F, n2b, n2f, nWP, nVP, NP, scalea, NK, syma);
}

// ---------------------------------------------------------------------------------------
// BSE_K2! and BSE_L_K2! of the s-wave lattice solver as tables and GEMMs.  In the s-wave vertex
// the cross channels are P-averaged and F0 is local, so the ω-dependent part of the BSE_K2!
// summand other than the channel's own K3, X(Ω, ν, ω), does not depend on P, and
//   Σ_ω X(Ω, ν, ω) W(ω, Ω, P)
// is, for every Ω, a (ν x ω) by (ω x P) complex GEMM over the band of the bubble mesh (beyond the
// band X is its far value, which multiplies the tail sums of W).  The own-K3 term and the
// rank-1 term K2(Ω, ν, P) Σ_ω W are added per (Ω, P) by k_k2nl_finish.  BSE_L_K2! in the lattice
// solver sums Γ(Ω, ν, ω~) F0(Ω, ω, ∞) Π0(ω, Ω, P) over the K2 mesh with Γ free of F0 and of the
// own channel, which is again a P-independent table times Π0.

// X table of channel CH (t: density, 2 X(t) - X(a)), X[ν + Nν2 (k + 2 wb iΩ)] on the band
// ω = k - wb, and its far values Xf[ν + Nν2 (side + 2 iΩ)] at ω = -wb - 1 and ω = wb
template <int CH>
__device__ __forceinline__ double2 k2x_term(const VtxV& F, int m, int v, int w) {
    if (CH == TCH) return csub(cscale(k2x_psp<TCH>(F, m, v, w), 2.0), k2x_psp<ACH>(F, m, v, w));
    return k2x_psp<CH>(F, m, v, w);
}
template <int CH>
__global__ __launch_bounds__(256) void k_k2nl_table(double2* __restrict__ X, double2* __restrict__ Xf, VtxV F, int n2b,
                                                    int n2f, int wb) {
    const int NV2 = 2 * n2f, NW2 = 2 * n2b - 1;
    const size_t nx = (size_t)NV2 * (2 * wb) * NW2;
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e < nx) {
        const int iv = (int)(e % NV2), k = (int)((e / NV2) % (2 * wb)), iW = (int)(e / ((size_t)NV2 * 2 * wb));
        X[e] = k2x_term<CH>(F, iW - (n2b - 1), iv - n2f, k - wb);
    } else if (e < nx + (size_t)NV2 * 2 * NW2) {
        const size_t f = e - nx;
        const int iv = (int)(f % NV2), side = (int)((f / NV2) % 2), iW = (int)(f / (2 * (size_t)NV2));
        Xf[f] = k2x_term<CH>(F, iW - (n2b - 1), iv - n2f, side ? wb : -wb - 1);
    }
}

// GEMM operands: batch b = iΩ of the K2 mesh
struct K2nl_A {   // X(ν, ω) of one Ω
    const double2* X; int NV, K;
    __device__ K2nl_A bind(int b) const { K2nl_A r = *this; r.X += (size_t)NV * K * b; return r; }
    __device__ double2 operator()(int i, int k) const { return X[(size_t)i + (size_t)NV * k]; }
};
struct K2nl_B {   // W(ω, P) (or Π0) of one Ω: rows of `ld` entries, element `k0 + k` of row (Ω, P)
    const double2* W; int ld, k0, rowoff, nrowP;
    __device__ K2nl_B bind(int b) const { K2nl_B r = *this; r.W += (size_t)ld * (b + rowoff) + k0; return r; }
    __device__ double2 operator()(int k, int j) const { return W[(size_t)k + (size_t)ld * nrowP * j]; }
};
struct K2nl_E {   // Y[ν + Nν2 (iΩ + NΩ2 P)]
    double2* Y; int NV, NW;
    __device__ K2nl_E bind(int b) const { K2nl_E r = *this; r.Y += (size_t)NV * b; return r; }
    __device__ void operator()(int i, int j, double2 v) const { Y[(size_t)i + (size_t)NV * NW * j] = v; }
};
// BSE_L_K2! epilogue: out(Ω, ν, P) = scale Y (t: (scale Y + out_a) / 2)
struct LK2nl_E {
    double2* out; const double2* outa; int NW, NV; double scale; int t;
    __device__ LK2nl_E bind(int b) const { LK2nl_E r = *this; r.out += b; if (r.outa) r.outa += b; return r; }
    __device__ void operator()(int i, int j, double2 v) const {
        const size_t o = (size_t)NW * ((size_t)i + (size_t)NV * j);
        double2 val = cscale(v, scale);
        if (t) val = cscale(cadd(val, outa[o]), 0.5);
        out[o] = val;
    }
};

// BSE_L_K2! table of the lattice solver: XL[ν + Nν2 (q + Nν2 iΩ)] = Γ(Ω, ν, ω~) F0(Ω, ω, ∞) on the
// K2 mesh ω = q - n2f (src/nonlocal/BSEa/BSEa_K2.jl:1-41)
template <int CH>
__global__ __launch_bounds__(256) void k_lk2nl_table(double2* __restrict__ X, VtxV F, int n2b, int n2f) {
    constexpr int SP = ChSpin<CH>::SP;
    constexpr int FLG = FL_ALL & ~FL_F0 & ~(FL_GP << CH);
    const int NV2 = 2 * n2f, NW2 = 2 * n2b - 1;
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)NV2 * NV2 * NW2) return;
    const int iv = (int)(e % NV2), q = (int)((e / NV2) % NV2), iW = (int)(e / ((size_t)NV2 * NV2));
    const int m = iW - (n2b - 1), v = iv - n2f, w = q - n2f;
    X[e] = cmul(F.eval(m, v, crossing(m, w, CH), 0, CH, SP, FLG), F.f0.eval(m, w, FDGA_INF, CH, SP));
}

// K2 = diagram + FL.K2 (t: (diagram + 2 FL.t.K2 - FL.a.K2 + K2a) / 2)  (src/BSEa/BSEa_K2.jl:95-100,
// src/BSE_templates.jl:29-33)
template <int CH>
__device__ __forceinline__ double2 k2_add_fl(double2 val, double2 lr, double2 la, double2 k2a) {
    if (CH == TCH) {
        val = cadd(val, cscale(lr, 2.0));
        val = cadd(val, cscale(la, -1.0));
        return cscale(cadd(val, k2a), 0.5);
    }
    return cadd(val, lr);
}

// BSE_K2! finish of the lattice solver, one wave per (Ω, P):
//   K2 = T s [Y + Xf_L Wt_L + Xf_R Wt_R + K2r(Ω, ν, P) (Σ_band W + Wt_L + Wt_R) + Σ_ω K3r(Ω, ν, ω, P) W(ω)]
//        + FL.K2 (t: (... + 2 FL.t.K2 - FL.a.K2 + K2a) / 2)
// with Wt_L, Wt_R the sums of W beyond the band and K3r the own K3 (t: 2 K3t - K3a).  The K3 product
// term is added afterwards by k_k2nl_k3 (a lean kernel of its own; K2 is linear in it, with weight
// T s, halved for t).
template <int CH>
__global__ __launch_bounds__(256) void k_k2nl_finish(double2* __restrict__ K2, const double2* __restrict__ K2a,
                                                     const double2* __restrict__ Y, const double2* __restrict__ Xf,
                                                     const double2* __restrict__ W, WTail Wtail,
                                                     VtxV F, ChanV Lr, ChanV La, int n2b, int n2f, int nWP, int wb,
                                                     int NP, double scale, double2* __restrict__ raw) {
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    const int NW2 = 2 * n2b - 1, NV2 = 2 * n2f;
    if (wid >= NW2 * NP) return;
    const int iW = wid % NW2, p = wid / NW2;
    const int m = iW - (n2b - 1);
    const size_t row = (size_t)(m + nWP - 1) + (size_t)(2 * nWP - 1) * p;
    const double2* Wr = W + (size_t)(2 * wb) * row + wb;   // Wr[ω], ω in [-wb, wb)
    const double2 wl = Wtail.at(row, 0, wb), wr = Wtail.at(row, (wb + 1) + wb, wb);   // issued first
    double sr = 0.0, si = 0.0;
    for (int k = lane; k < 2 * wb; k += WAVE) { const double2 x = Wr[k - wb]; sr += x.x; si += x.y; }
    sr = wave_sum(sr); si = wave_sum(si);
    const double2 wtot = cadd(cmk(sr, si), cadd(wl, wr));
    for (int iv = lane; iv < NV2; iv += WAVE) {
        const int v = iv - n2f;
        double2 acc = Y[(size_t)iv + (size_t)NV2 * (iW + (size_t)NW2 * p)];
        const double2* xf = Xf + (size_t)iv + (size_t)NV2 * 2 * iW;
        acc = cadd(acc, cadd(cmul(xf[0], wl), cmul(xf[NV2], wr)));
        acc = cadd(acc, cmul(k2d_rank1<CH>(F, m, v, p), wtot));
        const size_t o = (size_t)iW + (size_t)NW2 * ((size_t)iv + (size_t)NV2 * p);
        if (raw) { raw[o] = cscale(acc, scale); continue; }
        K2[o] = k2_add_fl<CH>(cscale(acc, scale), Lr.K2[o], La.K2[o], K2a[o]);
    }
}

// The K3 product term of the lattice BSE_K2! finish, added to its output (K2, or the raw values under a
// symmetry group): out += f Σ_ω K3r(Ω, ν, ω, P) W(ω) with f = T s (halved for t; t: K3r = 2 K3t - K3a).
// One wave per (Ω, P) as the finish, lanes over ν; K32: F.K3 held in FP32.
template <int CH, bool K32>
__global__ __launch_bounds__(256) void k_k2nl_k3(double2* __restrict__ out, const double2* __restrict__ W, VtxV F,
                                                 int n2b, int n2f, int nWP, int wb, int NP, double f) {
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    const int NW2 = 2 * n2b - 1, NV2 = 2 * n2f;
    if (wid >= NW2 * NP) return;
    const int iW = wid % NW2, p = wid / NW2;
    const int m = iW - (n2b - 1);
    const ChanV& g = F.g[CH];
    if (!in_b(m, g.n3b)) return;
    const size_t row = (size_t)(m + nWP - 1) + (size_t)(2 * nWP - 1) * p;
    const double2* Wr = W + (size_t)(2 * wb) * row + wb;   // Wr[ω], ω in [-wb, wb)
    for (int iv = lane; iv < NV2; iv += WAVE) {
        const int v = iv - n2f;
        if (!in_f(v, g.n3f)) continue;
        double2 k3s = czero();
#pragma unroll 8
        for (int w = -g.n3f; w < g.n3f; ++w) {
            double2 k3 = g.template k3p<K32>(m, v, w, p);
            if (CH == TCH) k3 = csub(cscale(k3, 2.0), F.g[ACH].template k3p<K32>(m, v, w, p));
            k3s = cadd(k3s, cmul(k3, Wr[w]));
        }
        const size_t o = (size_t)iW + (size_t)NW2 * ((size_t)iv + (size_t)NV2 * p);
        out[o] = cadd(out[o], cscale(k3s, f));
    }
}

// lattice BSE_K2! under a symmetry group: K2[j] = op(raw[rep(j)]) + the FL.K2 term of j (t: then
// the post-processing with the finished a channel)
template <int CH>
__global__ void k_k2nl_symfinish(double2* __restrict__ K2, const double2* __restrict__ K2a, const double2* __restrict__ raw,
                                 const int* __restrict__ rep, const double2* __restrict__ Lr, const double2* __restrict__ La,
                                 size_t n) {
    const size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int code = rep[j];
    K2[j] = k2_add_fl<CH>(sym_apply(code & 3, raw[code >> 2]), Lr[j], La[j], K2a[j]);
}

// ---------------------------------------------------------------------------------------
// SDE_channel_L_pp! / _ph! (src/SDE.jl:79-217; NL src/nonlocal/SDE.jl:3-142).
// kind 0: Vertex channels g (momentum p if NP > 1); kind 1: RefVertex tables.
template <bool DP, bool DH>
__device__ __forceinline__ void sde_terms(const ChanV& gp, const ChanV& gt, const ChanV& ga, const RefV& ref, int kind,
                                 int m, int v, int w, int pc, double2& xp, double2& xh) {
    xp = czero();
    xh = czero();
    if (kind == 0) {
        if (DP) xp = gp.eval(m, bmf(m, w), v, pc);
        if (DH) xh = cadd(ga.eval(m, v, w, pc), gt.eval(m, v, w, pc));
    } else {
        if (DP) xp = csub(ref.eval(m, bmf(m, w), v, PCH, PSP), ref.bare(PSP));
        if (DH) xh = csub(csub(cadd(ref.eval(m, v, w, ACH, PSP), ref.eval(m, v, w, TCH, PSP)), ref.bare(PSP)), ref.bare(PSP));
    }
}
// WHICH: bit 0 computes Lpp, bit 1 Lph.  Under a symmetry group the two outputs have their own
// classes (SGpp2, SGph2): one wave per class of either (k_sde_L_pair).
template <int WHICH>
__device__ __forceinline__ void sde_L_wave(int wid, int lane, double2* __restrict__ Lpp, double2* __restrict__ Lph,
                                           const double2* __restrict__ Pipp, const double2* __restrict__ Piph,
                                           const double2* __restrict__ Tpp, const double2* __restrict__ Tph,
                                           const ChanV& gp, const ChanV& gt, const ChanV& ga, const RefV& ref, int kind,
                                           int n2b, int n2f, int nWP, int nVP, int wb, int NP, int chanNP, double2 TU,
                                           int NK, const SymCls& sym) {
    constexpr bool DP = WHICH & 1, DH = WHICH & 2;
    const int NW2 = 2 * n2b - 1, NV2 = 2 * n2f;
    if (wid >= sym.count(NW2 * NV2 * NP)) return;
    const int o0 = sym.rep(wid);
    const int iW = o0 % NW2, iv = (o0 / NW2) % NV2, p = o0 / (NW2 * NV2);
    const int m = iW - (n2b - 1), v = iv - n2f;
    const int pc = chanNP > 1 ? p : 0;
    const int Wc = window(NK, m, v);
    const int lo = max(0, nVP - Wc), hi = min(2 * nVP - 1, nVP + Wc);
    const size_t row = (size_t)(m + nWP - 1) + (size_t)(2 * nWP - 1) * p;
    const size_t base = (size_t)(2 * nVP) * row;
    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
    for (int iw = lo + lane; iw <= hi; iw += WAVE) {
        double2 xp, xh;
        sde_terms<DP, DH>(gp, gt, ga, ref, kind, m, v, iw - nVP, pc, xp, xh);
        if (DP) { double2 a = cmul(Pipp[base + iw], xp); ar += a.x; ai += a.y; }
        if (DH) { double2 b = cmul(Piph[base + iw], xh); br += b.x; bi += b.y; }
    }
    double2 fxp, fxh, xpl, xpr, xhl, xhr;
    sde_terms<DP, DH>(gp, gt, ga, ref, kind, m, v, lane == 0 ? -Wc - 1 : Wc + 1, pc, fxp, fxh);
    far_pair(fxp, xpl, xpr);
    far_pair(fxh, xhl, xhr);
    ar = wave_sum(ar); ai = wave_sum(ai); br = wave_sum(br); bi = wave_sum(bi);
    double2 A = cmk(ar, ai), B = cmk(br, bi);
    if (lo > 0) {
        if (DP) A = cadd(A, cmul(tail_left(Tpp, row, nVP, wb, lo), xpl));
        if (DH) B = cadd(B, cmul(tail_left(Tph, row, nVP, wb, lo), xhl));
    }
    if (hi < 2 * nVP - 1) {
        if (DP) A = cadd(A, cmul(tail_right(Tpp, row, nVP, wb, hi), xpr));
        if (DH) B = cadd(B, cmul(tail_right(Tph, row, nVP, wb, hi), xhr));
    }
    const double2 rawp = cmul(A, TU), rawh = cmul(B, TU);
    int b, e;
    sym.range(wid, b, e);
    for (int k = b + lane; k < e; k += WAVE) {
        const int code = sym.member(wid, k), o = code >> 2;
        if (DP) Lpp[o] = sym_apply(code & 3, rawp);
        if (DH) Lph[o] = sym_apply(code & 3, rawh);
    }
}
// both outputs per element (no symmetry groups)
__global__ __launch_bounds__(256) void k_sde_L(double2* __restrict__ Lpp, double2* __restrict__ Lph, const double2* __restrict__ Pipp,
                        const double2* __restrict__ Piph, const double2* __restrict__ Tpp, const double2* __restrict__ Tph,
                        ChanV gp, ChanV gt, ChanV ga, RefV ref, int kind,
                        int n2b, int n2f, int nWP, int nVP, int wb, int NP, int chanNP, double2 TU, int NK, SymCls sym) {
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    sde_L_wave<3>(wid, lane, Lpp, Lph, Pipp, Piph, Tpp, Tph, gp, gt, ga, ref, kind, n2b, n2f, nWP, nVP, wb, NP, chanNP, TU,
                  NK, sym);
}
// the classes of SGpp2 (waves [0, ncls_pp)) and of SGph2 (the waves after them) in one launch
__global__ __launch_bounds__(256) void k_sde_L_pair(double2* __restrict__ Lpp, double2* __restrict__ Lph,
                        const double2* __restrict__ Pipp, const double2* __restrict__ Piph, const double2* __restrict__ Tpp,
                        const double2* __restrict__ Tph, ChanV gp, ChanV gt, ChanV ga, RefV ref, int kind,
                        int n2b, int n2f, int nWP, int nVP, int wb, int NP, int chanNP, double2 TU, int NK, SymCls spp,
                        SymCls sph) {
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    const int npp = spp.count((2 * n2b - 1) * (2 * n2f) * NP);
    if (wid < npp)
        sde_L_wave<1>(wid, lane, Lpp, Lph, Pipp, Piph, Tpp, Tph, gp, gt, ga, ref, kind, n2b, n2f, nWP, nVP, wb, NP, chanNP,
                      TU, NK, spp);
    else
        sde_L_wave<2>(wid - npp, lane, Lpp, Lph, Pipp, Piph, Tpp, Tph, gp, gt, ga, ref, kind, n2b, n2f, nWP, nVP, wb, NP,
                      chanNP, TU, NK, sph);
}

// SDE L-vertices of the s-wave lattice solver at a Vertex level (kind 0), one wave per (Ω, P).
// The channel evaluation (K1 + K2 + K2' + K3 by box, src/nonlocal/channel.jl:90-226) splits the
// sums over the bubble mesh into a rank-1 term, a ν-independent K2 sum and a K3 product:
//   Lpp = T U [(K1p + K2p(ν)) ΣΠpp + Σ_ω Πpp(ω) K2p(Ω-ω-1) + Σ_ω Πpp(ω) K3p(Ω, Ω-ω-1, ν)]
//   Lph = T U Σ_{r=a,t} [(K1r + K2r(ν)) ΣΠph + Σ_ω Πph(ω) K2r(ω) + Σ_ω Πph(ω) K3r(Ω, ν, ω)]
// with the channels at momentum P (chanNP > 1) or the local vertex (chanNP = 1).  ΣΠ over the
// whole bubble mesh is the band sum plus the tail sums at the band edges.
// the classes of a level's channel at its own boxes (src/channel.jl:220-285)
struct ChanBox {
    const ChanV& g;
    int pc;
    __device__ double2 k1(int m) const { return in_b(m, g.n1) ? g.k1(m, pc) : czero(); }
    __device__ double2 k2(int m, int n) const { return in_b(m, g.n2b) && in_f(n, g.n2f) ? g.k2(m, n, pc) : czero(); }
    __device__ double2 k3(int m, int n, int np) const {
        return in_b(m, g.n3b) && in_f(n, g.n3f) && in_f(np, g.n3f) ? g.k3(m, n, np, pc) : czero();
    }
    template <bool K32>
    __device__ double2 k3p(int m, int n, int np) const {   // (pc >= 0)
        return in_b(m, g.n3b) && in_f(n, g.n3f) && in_f(np, g.n3f) ? g.template k3p<K32>(m, n, np, pc) : czero();
    }
};
// the channels of one Vertex level of the lattice SDE: at momentum P (chanNP > 1) or the local
// vertex (chanNP = 1)
struct SdeLevelNL {
    ChanV gp, gt, ga;
    int chanNP;
};
__device__ __forceinline__ double2 shfl_xor2(double2 v, int mask) {
    return cmk(__shfl_xor(v.x, mask, WAVE), __shfl_xor(v.y, mask, WAVE));
}
// The p-channel K3 sum of the lattice SDE for every ν of the level's K3 box (2 n3f <= 64):
//   S(ν) = Σ_n K3p(Ω, n, ν) Πpp(Ω - n - 1)   (n in the box, Ω - n - 1 on the bubble mesh)
// with the lanes over n, so that each load reads a contiguous column of K3 (its first index), and the
// products of four ν at a time reduced across the wave by recursive halving (lanes exchange half
// their partial sums at distances 32 and 16, then sum at 8, 4, 2 and 1).  Returns S(ν) at the lane
// iv = ν + n2f (ν = iv - n2f outside the box: zero).
template <bool K32>
__device__ __forceinline__ double2 sde_p3_lanes(const ChanV& g, int pc, int m, const double2* pp, int nVP, int n2f,
                                                int lane) {
    const int n3 = g.n3f, NV = 2 * n3;
    const int n = lane - n3, wp = m - n - 1;
    const bool ln = lane < NV && wp >= -nVP && wp < nVP;
    const double2 w = ln ? pp[wp] : czero();
    const bool h5 = lane & 32, h4 = lane & 16;
    const int vme = lane - n2f + n3;   // this lane's ν, as an index into the box
    double2 mine = czero();
    for (int c0 = 0; c0 < NV; c0 += 4) {
        double2 x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            x[j] = (ln && c0 + j < NV) ? cmul(g.template k3p<K32>(m, n, c0 + j - n3, pc), w) : czero();
        double2 y[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) y[k] = cadd(h5 ? x[2 + k] : x[k], shfl_xor2(h5 ? x[k] : x[2 + k], 32));
        double2 u = cadd(h4 ? y[1] : y[0], shfl_xor2(h4 ? y[0] : y[1], 16));
        u = cadd(u, shfl_xor2(u, 8));
        u = cadd(u, shfl_xor2(u, 4));
        u = cadd(u, shfl_xor2(u, 2));
        u = cadd(u, shfl_xor2(u, 1));
        // u is S(ν) of box index c0 + ((lane >> 5) & 1) * 2 + ((lane >> 4) & 1)
        const int jv = (vme - c0) & 3;
        const int src = ((jv >> 1) & 1) << 5 | (jv & 1) << 4;
        const double2 s = cmk(__shfl(u.x, src, WAVE), __shfl(u.y, src, WAVE));
        if (vme >= c0 && vme < c0 + 4) mine = s;
    }
    return mine;
}

// The K3 terms of a momentum-independent Vertex level (the local reference vertex) in the lattice SDE
// L-vertices, as FP64 GEMMs per Ω of its K3 box over all momenta at once (n, ν in that box, the Π
// rows of (Ω, P)):
//   Lpp(Ω, ν, P) += T U Σ_n K3p(Ω, n, ν) Πpp(Ω, Ω - n - 1, P)     (SdeRefA<0>, SdeRefB<0>)
//   Lph(Ω, ν, P) += T U Σ_n [K3a + K3t](Ω, ν, n) Πph(Ω, n, P)      (SdeRefA<1>, SdeRefB<1>)
// The level's K3 slab is read once per Ω instead of once per (Ω, P).
template <int PH>
struct SdeRefA {   // A(i = ν, k = n) of one Ω of the box
    const double2* K0; const double2* K1; int NV;
    __device__ SdeRefA bind(int b) const {
        SdeRefA r = *this;
        r.K0 += (size_t)NV * NV * b;
        if (PH) r.K1 += (size_t)NV * NV * b;
        return r;
    }
    __device__ double2 operator()(int i, int k) const {
        if (PH) { const size_t o = (size_t)i + (size_t)NV * k; return cadd(K0[o], K1[o]); }   // K3a + K3t
        return K0[(size_t)k + (size_t)NV * i];                                                // K3p(n, ν)
    }
};
template <int PH>
struct SdeRefB {   // B(k = n, j = P): the Π row of (Ω, P) at ω = Ω - n - 1 (pp) or n (ph), 0 off the mesh
    const double2* Pi; int n3b, n3f, nWP, nVP, m;
    __device__ SdeRefB bind(int b) const {
        SdeRefB r = *this;
        r.m = b - (n3b - 1);
        r.Pi += (size_t)(2 * nVP) * (size_t)(r.m + nWP - 1) + nVP;
        return r;
    }
    __device__ double2 operator()(int k, int j) const {
        const int n = k - n3f, w = PH ? n : m - n - 1;
        if (w < -nVP || w >= nVP) return czero();
        return Pi[(size_t)(2 * nVP) * (size_t)(2 * nWP - 1) * j + w];
    }
};
struct SdeRefE {   // L(Ω, ν, P) += T U C(ν, P) inside the L box
    double2* L; double2 TU; int n2b, n2f, n3b, n3f, m;
    __device__ SdeRefE bind(int b) const { SdeRefE r = *this; r.m = b - (n3b - 1); return r; }
    __device__ void operator()(int i, int j, double2 v) const {
        const int nu = i - n3f;
        if (m <= -n2b || m >= n2b || nu < -n2f || nu >= n2f) return;
        const size_t o = (size_t)(m + n2b - 1) + (size_t)(2 * n2b - 1) * ((size_t)(nu + n2f) + (size_t)(2 * n2f) * j);
        L[o] = cadd(L[o], cmul(v, TU));
    }
};

// NLEV Vertex levels summed into one pair of L-vertices (the target vertex and the local reference
// vertex of the fdPA iteration share ΣΠ, the bubble rows and the launch).  SKIP1: the second level's
// K3 terms are left out (added by the momentum GEMMs SdeRef*, when that level is momentum-independent).
// F0: the first level's K3 form fixed at compile time (0 FP64, 1 FP32; -1: read from the level); with
// a fixed form the first level's a/t K3 sum is left to k_sde_h3 (a lean kernel of its own)
template <int NLEV, bool SKIP1 = false, int F0 = -1>
__global__ __launch_bounds__(256) void k_sde_L_nl(double2* __restrict__ Lpp, double2* __restrict__ Lph,
                                                  const double2* __restrict__ Pipp, const double2* __restrict__ Piph,
                                                  const double2* __restrict__ Tpp, const double2* __restrict__ Tph,
                                                  SdeLevelNL lv0, SdeLevelNL lv1, int n2b, int n2f, int nWP, int nVP,
                                                  int wb, int NP, double2 TU) {
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    const int NW2 = 2 * n2b - 1, NV2 = 2 * n2f;
    if (wid >= NW2 * NP) return;
    // momentum fastest: the waves in flight share Ω, so the local reference vertex's K3(Ω) slab (the
    // same at every momentum) stays in L2 while they read it
    const int p = wid % NP, iW = wid / NP;
    const int m = iW - (n2b - 1);
    const size_t row = (size_t)(m + nWP - 1) + (size_t)(2 * nWP - 1) * p;
    const double2* pp = Pipp + (size_t)(2 * nVP) * row + nVP;   // pp[ω], ω in [-nVP, nVP)
    const double2* ph = Piph + (size_t)(2 * nVP) * row + nVP;
    auto inmesh = [&](int w) { return w >= -nVP && w < nVP; };
    // totals over the bubble mesh (band + tails)
    double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
    for (int w = -wb + lane; w < wb; w += WAVE) {
        const double2 x = pp[w], y = ph[w];
        a0 += x.x; a1 += x.y; b0 += y.x; b1 += y.y;
    }
    a0 = wave_sum(a0); a1 = wave_sum(a1); b0 = wave_sum(b0); b1 = wave_sum(b1);
    const double2* tp = Tpp + (size_t)(2 * (wb + 1)) * row;
    const double2* th = Tph + (size_t)(2 * (wb + 1)) * row;
    const double2 spp = cadd(cmk(a0, a1), cadd(tp[0], tp[2 * wb + 1]));
    const double2 sph = cadd(cmk(b0, b1), cadd(th[0], th[2 * wb + 1]));
    // per level: the ν-independent K2 sums and K1
    double2 k2p[NLEV], k2h[NLEV], k1p[NLEV], k1h[NLEV];
#pragma unroll
    for (int l = 0; l < NLEV; ++l) {
        const SdeLevelNL& L = l == 0 ? lv0 : lv1;
        const int pc = L.chanNP > 1 ? p : 0;
        const ChanBox P{L.gp, pc}, A{L.ga, pc}, T{L.gt, pc};
        double c0 = 0, c1 = 0, d0 = 0, d1 = 0;
        for (int n = -L.gp.n2f + lane; n < L.gp.n2f; n += WAVE) {
            const int wp = m - n - 1;   // Lpp: K2p(Ω, n) at n = Ω - ω - 1
            if (inmesh(wp)) { const double2 x = cmul(P.k2(m, n), pp[wp]); c0 += x.x; c1 += x.y; }
        }
        for (int n = -max(L.ga.n2f, L.gt.n2f) + lane; n < max(L.ga.n2f, L.gt.n2f); n += WAVE) {
            if (inmesh(n)) { const double2 y = cmul(cadd(A.k2(m, n), T.k2(m, n)), ph[n]); d0 += y.x; d1 += y.y; }
        }
        k2p[l] = cmk(wave_sum(c0), wave_sum(c1));
        k2h[l] = cmk(wave_sum(d0), wave_sum(d1));
        k1p[l] = P.k1(m);
        k1h[l] = cadd(A.k1(m), T.k1(m));
    }
    // the p-channel K3 sums with the lanes over the summation index (one pass over ν: 2 n2f <= 64)
    double2 p3[NLEV];
    bool p3l[NLEV];
#pragma unroll
    for (int l = 0; l < NLEV; ++l) {
        const SdeLevelNL& L = l == 0 ? lv0 : lv1;
        p3l[l] = NV2 <= WAVE && 2 * L.gp.n3f <= WAVE;
        p3[l] = czero();
        if (p3l[l] && in_b(m, L.gp.n3b) && !(l == 1 && SKIP1)) {
            const int pc = L.chanNP > 1 ? p : 0;
            const bool f32 = (l == 0 && F0 >= 0) ? F0 == 1 : L.gp.k3f != 0;
            p3[l] = f32 ? sde_p3_lanes<true>(L.gp, pc, m, pp, nVP, n2f, lane)
                             : sde_p3_lanes<false>(L.gp, pc, m, pp, nVP, n2f, lane);
        }
    }
    for (int iv = lane; iv < NV2; iv += WAVE) {
        const int v = iv - n2f;
        double2 xp = czero(), xh = czero();
#pragma unroll
        for (int l = 0; l < NLEV; ++l) {
            const SdeLevelNL& L = l == 0 ? lv0 : lv1;
            const int pc = L.chanNP > 1 ? p : 0;
            const ChanBox P{L.gp, pc}, A{L.ga, pc}, T{L.gt, pc};
            xp = cadd(xp, cadd(cmul(cadd(k1p[l], P.k2(m, v)), spp), k2p[l]));
            xh = cadd(xh, cadd(cmul(cadd(k1h[l], cadd(A.k2(m, v), T.k2(m, v))), sph), k2h[l]));
            const int n3p = L.gp.n3f, n3h = max(L.ga.n3f, L.gt.n3f);
            if (l == 1 && SKIP1) continue;
            // the K3 form is the level's (one vertex): the loops branch on it once
            if (p3l[l]) {
                if (in_b(m, L.gp.n3b) && in_f(v, n3p)) xp = cadd(xp, p3[l]);
            } else if (in_b(m, L.gp.n3b) && in_f(v, n3p)) {
                auto psum = [&](auto form) {   // (wide boxes)
                    constexpr bool K32 = decltype(form)::value;
                    double2 sp = czero();
                    for (int n = -n3p; n < n3p; ++n) {
                        const int wp = m - n - 1;
                        if (inmesh(wp)) sp = cadd(sp, cmul(P.template k3p<K32>(m, n, v), pp[wp]));
                    }
                    return sp;
                };
                const bool f32 = (l == 0 && F0 >= 0) ? F0 == 1 : L.gp.k3f != 0;
                xp = cadd(xp, f32 ? psum(std::true_type{}) : psum(std::false_type{}));
            }
            if (!(l == 0 && F0 >= 0) && (in_b(m, L.ga.n3b) || in_b(m, L.gt.n3b)) && in_f(v, n3h)) {
                auto hsum = [&](auto form) {
                    constexpr bool K32 = decltype(form)::value;
                    double2 sh = czero();
#pragma unroll 8
                    for (int n = -n3h; n < n3h; ++n)
                        if (inmesh(n))
                            sh = cadd(sh, cmul(cadd(A.template k3p<K32>(m, v, n), T.template k3p<K32>(m, v, n)), ph[n]));
                    return sh;
                };
                const bool f32 = (l == 0 && F0 >= 0) ? F0 == 1 : L.ga.k3f != 0;
                xh = cadd(xh, f32 ? hsum(std::true_type{}) : hsum(std::false_type{}));
            }
        }
        const size_t o = (size_t)iW + (size_t)NW2 * ((size_t)iv + (size_t)NV2 * p);
        Lpp[o] = cmul(xp, TU);
        Lph[o] = cmul(xh, TU);
    }
}

// The a/t K3 sum of a lattice Vertex level at momentum P, added to Lph (the part k_sde_L_nl<.., F0>
// leaves out): Lph(Ω, ν, P) += T U Σ_n [K3a + K3t](Ω, ν, n, P) Πph(Ω, n, P), the sum in k_sde_L_nl's
// order.  One wave per (Ω, P), momentum fastest, lanes over ν; few registers, so many waves in flight.
template <bool K32>
__global__ __launch_bounds__(256) void k_sde_h3(double2* __restrict__ Lph, const double2* __restrict__ Piph, ChanV ga,
                                                ChanV gt, int n2b, int n2f, int nWP, int nVP, int NP, double2 TU) {
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    const int NW2 = 2 * n2b - 1, NV2 = 2 * n2f;
    if (wid >= NW2 * NP) return;
    const int p = wid % NP, iW = wid / NP;
    const int m = iW - (n2b - 1);
    if (!in_b(m, ga.n3b) && !in_b(m, gt.n3b)) return;
    const size_t row = (size_t)(m + nWP - 1) + (size_t)(2 * nWP - 1) * p;
    const double2* ph = Piph + (size_t)(2 * nVP) * row + nVP;
    const ChanBox A{ga, p}, T{gt, p};
    const int n3h = max(ga.n3f, gt.n3f);
    for (int iv = lane; iv < NV2; iv += WAVE) {
        const int v = iv - n2f;
        if (!in_f(v, n3h)) continue;
        double2 sh = czero();
#pragma unroll 8
        for (int n = -n3h; n < n3h; ++n)
            if (n >= -nVP && n < nVP)
                sh = cadd(sh, cmul(cadd(A.template k3p<K32>(m, v, n), T.template k3p<K32>(m, v, n)), ph[n]));
        const size_t o = (size_t)iW + (size_t)NW2 * ((size_t)iv + (size_t)NV2 * p);
        Lph[o] = cadd(Lph[o], cmul(sh, TU));
    }
}

// Σ(ν) += scale T Σ_Ω [G(Ω-ν) Lpp(Ω,ν) + G(Ω+ν) Lph(Ω,ν)]  (src/SDE.jl:242-263); one wave per ν
__device__ __forceinline__ void sde_sigma_local_wave(int wid, int lane, double2* __restrict__ S, const double2* __restrict__ G,
                                                     const double2* __restrict__ Lpp, const double2* __restrict__ Lph,
                                                     int nG, int n2b, int n2f, double T, double scale, const SymCls& sym) {
    if (wid >= sym.count(2 * nG)) return;
    const int iv = sym.rep(wid);
    const int v = iv - nG;
    if (!in_f(v, n2f)) return;
    const int NW2 = 2 * n2b - 1;
    double sr = 0.0, si = 0.0;
    for (int iW = lane; iW < NW2; iW += WAVE) {
        const int m = iW - (n2b - 1);
        const size_t k = (size_t)iW + (size_t)NW2 * (v + n2f);
        const int a = bmf(m, v), b = bpf(m, v);
        double2 s = czero();
        if (in_f(a, nG)) s = cadd(s, cmul(G[a + nG], Lpp[k]));
        if (in_f(b, nG)) s = cadd(s, cmul(G[b + nG], Lph[k]));
        sr += s.x; si += s.y;
    }
    sr = wave_sum(sr); si = wave_sum(si);
    // a class of the Σ group ({ν, -ν}) lies inside or outside the K2 box as a whole
    const double2 raw = cscale(cmk(sr, si), T * scale);
    int b, e;
    sym.range(wid, b, e);
    for (int k = b + lane; k < e; k += WAVE) {
        const int code = sym.member(wid, k), o = code >> 2;
        S[o] = cadd(S[o], sym_apply(code & 3, raw));
    }
}
__global__ __launch_bounds__(256) void k_sde_sigma_local(double2* __restrict__ S, const double2* __restrict__ G,
                                                         const double2* __restrict__ Lpp, const double2* __restrict__ Lph,
                                                         int nG, int n2b, int n2f, double T, double scale, SymCls sym) {
    sde_sigma_local_wave((blockIdx.x * blockDim.x + threadIdx.x) / WAVE, threadIdx.x & (WAVE - 1), S, G, Lpp, Lph, nG, n2b,
                         n2f, T, scale, sym);
}

// NL SDE_compute_inner! real-space step (src/nonlocal/SDE.jl:209-276): for output column q of
// Σ_R[ν, q] (q on the LG mesh) collect Σ_R( R) += G(-R) Lpp(R) and Σ_R(-R) += G(-R) Lph(R)
// over R in [-L/2, L/2]^2 (half weights at |R_i| = L/2 for even L); the result is T Σ_R.
// One wave per (ν, q), the lanes over Ω (the L-vertices' Ω index is their fastest: coalesced rows)
// and a wave sum; ν outside the K2 box write zero.
__global__ __launch_bounds__(256) void k_sde_sigma_nl_R(double2* __restrict__ SR, const double2* __restrict__ GR,
                                                        const double2* __restrict__ LppR, const double2* __restrict__ LphR,
                                                        int nG, int n2b, int n2f, int LG, int L, double T) {
    const int NG = 2 * nG;
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    if (wid >= NG * LG * LG) return;
    const int iv = wid % NG, q = wid / NG;
    const int v = iv - nG;
    double sr = 0.0, si = 0.0;
    if (in_f(v, n2f)) {
        const int q1 = q % LG, q2 = q / LG, h = L / 2;
        const int NW2 = 2 * n2b - 1;
        const size_t nrow = (size_t)NW2 * (2 * n2f);
        auto md = [](int a, int n) { return ((a % n) + n) % n; };
        // the R_i in [-h, h] that fold onto q_i or -q_i, in ascending order (at most four per axis when
        // L <= LG; otherwise every R_i is visited)
        int ra[4], rb[4], na = 0, nb = 0;
        const bool lists = L <= LG;
        if (lists)
            for (int R = -h; R <= h; ++R) {
                if ((md(R, LG) == q1 || md(-R, LG) == q1) && na < 4) ra[na++] = R;
                if ((md(R, LG) == q2 || md(-R, LG) == q2) && nb < 4) rb[nb++] = R;
            }
        const int n1 = lists ? na : 2 * h + 1, n2 = lists ? nb : 2 * h + 1;
        for (int i2 = 0; i2 < n2; ++i2)
            for (int i1 = 0; i1 < n1; ++i1) {
                const int R1 = lists ? ra[i1] : i1 - h, R2 = lists ? rb[i2] : i2 - h;
                const bool tpp = md(R1, LG) == q1 && md(R2, LG) == q2;
                const bool tph = md(-R1, LG) == q1 && md(-R2, LG) == q2;
                if (!tpp && !tph) continue;
                double w = 1.0;
                if (L % 2 == 0) {
                    if (abs(R1) == L / 2) w *= 0.5;
                    if (abs(R2) == L / 2) w *= 0.5;
                }
                const size_t iRL = (size_t)(md(R1, L) + L * md(R2, L));
                const size_t imRG = (size_t)(md(-R1, LG) + LG * md(-R2, LG));
                for (int iW = lane; iW < NW2; iW += WAVE) {
                    const int m = iW - (n2b - 1);
                    const size_t k = (size_t)iW + (size_t)NW2 * (v + n2f) + nrow * iRL;
                    double2 s = czero();
                    if (tpp) {
                        const int a = bmf(m, v);
                        if (in_f(a, nG)) s = cadd(s, cscale(cmul(GR[(size_t)(a + nG) + (size_t)NG * imRG], LppR[k]), w));
                    }
                    if (tph) {
                        const int b = bpf(m, v);
                        if (in_f(b, nG)) s = cadd(s, cscale(cmul(GR[(size_t)(b + nG) + (size_t)NG * imRG], LphR[k]), w));
                    }
                    sr += s.x;
                    si += s.y;
                }
            }
        sr = wave_sum(sr);
        si = wave_sum(si);
    }
    if (lane == 0) SR[(size_t)iv + (size_t)NG * q] = cscale(cmk(sr, si), T);
}

// SDE_U2_using_G (src/SDE.jl:328-364) via C(m) = Σ_n G(n) G(n+m):
//   Σ_U²(ν) = U² T² Σ_m C(m) G(m + ν); for the NL solver one column per R of the real-space G.
// One wave per output, lanes over the summed index.
__device__ __forceinline__ void u2_C_wave(int wid, int lane, double2* __restrict__ C, const double2* __restrict__ Gm,
                                          const double2* __restrict__ Gp, int nG, int nR) {
    const int NG = 2 * nG, NB = 4 * nG - 1;
    if (wid >= NB * nR) return;
    const int im = wid % NB, r = wid / NB;
    const int m = im - (2 * nG - 1);
    double sr = 0.0, si = 0.0;
    for (int i2 = lane; i2 < NG; i2 += WAVE) {
        const int n2 = i2 - nG, n1 = n2 + m;
        if (in_f(n1, nG)) {
            double2 x = cmul(Gm[(size_t)(n1 + nG) + (size_t)NG * r], Gp[(size_t)i2 + (size_t)NG * r]);
            sr += x.x; si += x.y;
        }
    }
    sr = wave_sum(sr); si = wave_sum(si);
    if (lane == 0) C[(size_t)im + (size_t)NB * r] = cmk(sr, si);
}
__global__ __launch_bounds__(256) void k_u2_C(double2* __restrict__ C, const double2* __restrict__ Gm,
                                              const double2* __restrict__ Gp, int nG, int nR) {
    u2_C_wave((blockIdx.x * blockDim.x + threadIdx.x) / WAVE, threadIdx.x & (WAVE - 1), C, Gm, Gp, nG, nR);
}
// The lattice solver's C(m) for every R column (nR = NkG): each wave four consecutive m of one column
// (Gp loaded once for the four; every output's lane partial sums and wave sum as in u2_C_wave)
__global__ __launch_bounds__(256) void k_u2_C4(double2* __restrict__ C, const double2* __restrict__ Gm,
                                               const double2* __restrict__ Gp, int nG, int nR) {
    const int NG = 2 * nG, NB = 4 * nG - 1, NB4 = (NB + 3) / 4;
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    if (wid >= NB4 * nR) return;
    const int im0 = 4 * (wid % NB4), r = wid / NB4;
    double sr[4] = {0.0, 0.0, 0.0, 0.0}, si[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i2 = lane; i2 < NG; i2 += WAVE) {
        const int n2 = i2 - nG;
        const double2 gp = Gp[(size_t)i2 + (size_t)NG * r];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int m = im0 + u - (2 * nG - 1), n1 = n2 + m;
            if (im0 + u < NB && in_f(n1, nG)) {
                const double2 x = cmul(Gm[(size_t)(n1 + nG) + (size_t)NG * r], gp);
                sr[u] += x.x; si[u] += x.y;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const double a = wave_sum(sr[u]), b = wave_sum(si[u]);
        if (lane == 0 && im0 + u < NB) C[(size_t)(im0 + u) + (size_t)NB * r] = cmk(a, b);
    }
}
// local solver, first SDE level: the Σ sum over the L-vertices (waves [0, nsig)) and the U² term's
// C(m) (the waves after them) in one launch; they read different arrays and write S and C
__global__ __launch_bounds__(256) void k_sigma_u2c(double2* __restrict__ S, const double2* __restrict__ G,
                                                   const double2* __restrict__ Lpp, const double2* __restrict__ Lph,
                                                   int nG, int n2b, int n2f, double T, double scale, SymCls sym,
                                                   double2* __restrict__ C, int nsig) {
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    if (wid < nsig) sde_sigma_local_wave(wid, lane, S, G, Lpp, Lph, nG, n2b, n2f, T, scale, sym);
    else u2_C_wave(wid - nsig, lane, C, G, G, nG, 1);
}
__global__ __launch_bounds__(256) void k_u2_S(double2* __restrict__ S, const double2* __restrict__ C,
                                              const double2* __restrict__ Gp, int nG, int nR, double2 fac, int accumulate,
                                              SymCls sym) {
    const int NG = 2 * nG, NB = 4 * nG - 1;
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    if (wid >= sym.count(NG * nR)) return;
    const int o0 = sym.rep(wid);
    const int iv = o0 % NG, r = o0 / NG;
    const int v = iv - nG;
    double sr = 0.0, si = 0.0;
    for (int im = lane; im < NB; im += WAVE) {
        const int m = im - (2 * nG - 1), x = m + v;
        if (in_f(x, nG)) {
            double2 y = cmul(C[(size_t)im + (size_t)NB * r], Gp[(size_t)(x + nG) + (size_t)NG * r]);
            sr += y.x; si += y.y;
        }
    }
    sr = wave_sum(sr); si = wave_sum(si);
    const double2 raw = cmul(cmk(sr, si), fac);
    int b, e;
    sym.range(wid, b, e);
    for (int k = b + lane; k < e; k += WAVE) {
        const int code = sym.member(wid, k), o = code >> 2;
        const double2 s = sym_apply(code & 3, raw);
        S[o] = accumulate ? cadd(S[o], s) : s;
    }
}

// The lattice solver's Σ_U² for every R column (no symmetry classes): each wave four consecutive ν of
// one column (C loaded once for the four; every output's sums as in k_u2_S)
__global__ __launch_bounds__(256) void k_u2_S4(double2* __restrict__ S, const double2* __restrict__ C,
                                               const double2* __restrict__ Gp, int nG, int nR, double2 fac) {
    const int NG = 2 * nG, NB = 4 * nG - 1, NG4 = NG / 4;
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    if (wid >= NG4 * nR) return;
    const int iv0 = 4 * (wid % NG4), r = wid / NG4;
    double sr[4] = {0.0, 0.0, 0.0, 0.0}, si[4] = {0.0, 0.0, 0.0, 0.0};
    for (int im = lane; im < NB; im += WAVE) {
        const int m = im - (2 * nG - 1);
        const double2 cm = C[(size_t)im + (size_t)NB * r];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int x = m + iv0 + u - nG;
            if (in_f(x, nG)) {
                const double2 y = cmul(cm, Gp[(size_t)(x + nG) + (size_t)NG * r]);
                sr[u] += y.x; si[u] += y.y;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const double a = wave_sum(sr[u]), b = wave_sum(si[u]);
        if (lane == 0) S[(size_t)(iv0 + u) + (size_t)NG * r] = cmul(cmk(a, b), fac);
    }
}

// occupation n = 1/2 + T Im Σ G / Nk (src/dyson.jl:33-41): single-workgroup reduction
// partial sums of Im G over OCC_BLOCKS blocks, then one block adds them in a fixed order
constexpr int OCC_BLOCKS = 128;
__global__ __launch_bounds__(256) void k_occupation_part(double* __restrict__ part, const double2* __restrict__ G, size_t n) {
    __shared__ double red[256 / WAVE];
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)256 * OCC_BLOCKS) s += G[i].y;
    s = wave_sum(s);
    if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
__global__ __launch_bounds__(WAVE) void k_occupation(double* __restrict__ out, const double* __restrict__ part, double T,
                                                     double nk) {
    double s = 0.0;
    for (int b = threadIdx.x; b < OCC_BLOCKS; b += WAVE) s += part[b];
    s = wave_sum(s);
    if (threadIdx.x == 0) out[0] = 0.5 + s * T / nk;
}
// Σ += sign (n - 1/2) U i
__global__ void k_hartree(double2* __restrict__ S, size_t n, const double* __restrict__ occ, double2 U, double sign) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double f = (occ[0] - 0.5) * sign;
    S[i] = cadd(S[i], cmul(cmk(0.0, f), U));
}

// The three steps above in one block, for small G arrays: occupation n (written to occ[0]), then
// Σ += sign (n - 1/2) U i
__global__ __launch_bounds__(256) void k_hartree_small(double2* __restrict__ S, const double2* __restrict__ G, size_t n,
                                                      double* __restrict__ occ, double T, double nk, double2 U,
                                                      double sign) {
    __shared__ double red[256 / WAVE];
    double s = 0.0;
    for (size_t i = threadIdx.x; i < n; i += 256) s += G[i].y;
    s = wave_sum(s);
    if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = s;
    __syncthreads();
    const double nocc = 0.5 + (red[0] + red[1] + red[2] + red[3]) * T / nk;
    if (threadIdx.x == 0) occ[0] = nocc;
    const double2 h = cmul(cmk(0.0, (nocc - 0.5) * sign), U);
    for (size_t i = threadIdx.x; i < n; i += 256) S[i] = cadd(S[i], h);
}

// zero fill (used instead of hipMemsetAsync so that a captured iteration holds kernel nodes only)
__global__ void k_zero(double2* __restrict__ y, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = czero();
}

// y = x
__global__ void k_copy(double2* __restrict__ y, const double2* __restrict__ x, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}

// y = a x + b y (real scalars a, b)
__global__ void k_axpby(double2* __restrict__ y, const double2* __restrict__ x, double a, double b, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = cadd(cscale(x[i], a), cscale(y[i], b));
}

// layout permutations between the reference layout and the device layout
// K3: ref [Ω, ν, νp, P] <-> dev [ν, νp, Ω, P]
__global__ void k_k3_to_dev(double2* __restrict__ out, const double2* __restrict__ in, int NW, int NV, int NP) {
    size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t tot = (size_t)NW * NV * NV * NP;
    if (e >= tot) return;
    int iW = (int)(e % NW);
    size_t r = e / NW;
    int i = (int)(r % NV); r /= NV;
    int j = (int)(r % NV);
    size_t p = r / NV;
    out[(size_t)i + (size_t)NV * ((size_t)j + (size_t)NV * ((size_t)iW + (size_t)NW * p))] = in[e];
}
// the same, for a K3 array held in FP32 on the device
__global__ void k_k3_to_dev32(float2* __restrict__ out, const double2* __restrict__ in, int NW, int NV, int NP) {
    size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t tot = (size_t)NW * NV * NV * NP;
    if (e >= tot) return;
    int iW = (int)(e % NW);
    size_t r = e / NW;
    int i = (int)(r % NV); r /= NV;
    int j = (int)(r % NV);
    size_t p = r / NV;
    const double2 v = in[e];
    out[(size_t)i + (size_t)NV * ((size_t)j + (size_t)NV * ((size_t)iW + (size_t)NW * p))] = make_float2((float)v.x, (float)v.y);
}
__global__ void k_k3_to_ref32(double2* __restrict__ out, const float2* __restrict__ in, int NW, int NV, int NP) {
    size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t tot = (size_t)NW * NV * NV * NP;
    if (e >= tot) return;
    int iW = (int)(e % NW);
    size_t r = e / NW;
    int i = (int)(r % NV); r /= NV;
    int j = (int)(r % NV);
    size_t p = r / NV;
    const float2 v = in[(size_t)i + (size_t)NV * ((size_t)j + (size_t)NV * ((size_t)iW + (size_t)NW * p))];
    out[e] = cmk(v.x, v.y);
}
__global__ void k_k3_to_ref(double2* __restrict__ out, const double2* __restrict__ in, int NW, int NV, int NP) {
    size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t tot = (size_t)NW * NV * NV * NP;
    if (e >= tot) return;
    int iW = (int)(e % NW);
    size_t r = e / NW;
    int i = (int)(r % NV); r /= NV;
    int j = (int)(r % NV);
    size_t p = r / NV;
    out[e] = in[(size_t)i + (size_t)NV * ((size_t)j + (size_t)NV * ((size_t)iW + (size_t)NW * p))];
}
// bubble: ref [Ω, ν, P] <-> dev [ν, Ω, P]
__global__ void k_pi_to_dev(double2* __restrict__ out, const double2* __restrict__ in, int NW, int NV, int NP) {
    size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t tot = (size_t)NW * NV * NP;
    if (e >= tot) return;
    int iW = (int)(e % NW);
    size_t r = e / NW;
    int iv = (int)(r % NV);
    size_t p = r / NV;
    out[(size_t)iv + (size_t)NV * ((size_t)iW + (size_t)NW * p)] = in[e];
}
__global__ void k_pi_to_ref(double2* __restrict__ out, const double2* __restrict__ in, int NW, int NV, int NP) {
    size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t tot = (size_t)NW * NV * NP;
    if (e >= tot) return;
    int iW = (int)(e % NW);
    size_t r = e / NW;
    int iv = (int)(r % NV);
    size_t p = r / NV;
    out[e] = in[(size_t)iv + (size_t)NV * ((size_t)iW + (size_t)NW * p)];
}
// P-independent cache [i, j, Ω] -> ref [Ω, i, j, P] (replicated over P)
__global__ void k_cache_to_ref(double2* __restrict__ out, const double2* __restrict__ in, int NW, int NV, int NP) {
    size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t tot = (size_t)NW * NV * NV * NP;
    if (e >= tot) return;
    int iW = (int)(e % NW);
    size_t r = e / NW;
    int i = (int)(r % NV); r /= NV;
    int j = (int)(r % NV);
    out[e] = in[(size_t)i + (size_t)NV * ((size_t)j + (size_t)NV * iW)];
}

}  // namespace fdga
