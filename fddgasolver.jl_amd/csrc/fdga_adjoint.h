// fdga_adjoint.h -- tangent (JVP) and adjoint (VJP) of build_K3_cache! followed by BSE_K3! of the three
// channels (src/build_K3_cache.jl:20-95, src/BSEa/BSEa_K3.jl:43-128, src/BSE_templates.jl:151-180; NL
// src/nonlocal/build_K3_cache.jl:18-94, src/nonlocal/BSEa/BSEa_K3.jl:1-128).
//
// The composite maps (F, FL.K3, Π) to Fbuff.K3.  Per channel r and batch (Ω, P), sums over the K3 box:
//   raw_r = T A_r R_r + self_r
//   A_r   = cache_Fr(F)                                          [ν, ω]
//   R_r   = s1 [(Π - Π0) F0r + Π Γ~_r(F)] + s2 Π L~_r(ω~) 1(ω~ in box)   [ω, ν']
//   self_r = FL_r, L~_r = FL_r at row ω~ (t: 2 FL_t - FL_a);  out_p = raw_p, out_a = raw_a,
//   out_t = (raw_t + raw_a) / 2
// The caches are affine in F (the F0 parts are constants), so the gather's tangent is the gather of
// the tangent vertex with F0 = 0: the cache kernels run on a second context with U = 0 and no
// reference vertex.  The contraction's tangent follows the product rule,
//   d raw = T (dA R + A dR) + d self,  dR = s1 [dΠ (F0r + Γ~) + Π dΓ~] + s2 (dΠ L~ + Π dL~) 1(...),
// one GEMM over the concatenated inner range [dA | A] [R; dR].  The adjoint (bilinear pairing
// <u, v> = Σ u v, no conjugation: the gradient of a real loss is its conjugate) runs the transposed
// products Ā = T Ḡ R^T and R̄ = T A^T Ḡ, the right-operand builder's transpose element by element,
// and the gather's transpose as a segmented sum over a destination-sorted map traced on the host
// from the same evaluation rules as the cache kernels (one wave per vertex element, no atomics).
#pragma once
#include <cstdint>
#include <vector>

#include "fdga_kernels.h"

namespace fdga {
namespace adj {

// ---------------------------------------------------------------------------------------
// host: the gather as a linear map of the momentum-averaged vertex (NL: the cross-channel terms read
// the P-averages, src/nonlocal/swave.jl:32-88; local: the vertex itself)
struct Boxes {
    int n1, n2b, n2f, n3b, n3f;
    size_t s1() const { return (size_t)(2 * n1 - 1); }
    size_t s2() const { return (size_t)(2 * n2b - 1) * (2 * n2f); }
    size_t s3() const { return (size_t)(2 * n3f) * (2 * n3f) * (2 * n3b - 1); }
    size_t per_ch() const { return s1() + s2() + s3(); }
    size_t ndst() const { return 3 * per_ch(); }
    // destinations: channel blocks p, t, a (device channel order), each [K1; K2; K3] in device layout
    size_t d1(int ch, int m) const { return ch * per_ch() + (size_t)(m + n1 - 1); }
    size_t d2(int ch, int m, int n) const {
        return ch * per_ch() + s1() + (size_t)(m + n2b - 1) + (size_t)(2 * n2b - 1) * (n + n2f);
    }
    size_t d3(int ch, int m, int n, int np) const {
        return ch * per_ch() + s1() + s2() + (size_t)(n + n3f) +
               (size_t)(2 * n3f) * ((size_t)(np + n3f) + (size_t)(2 * n3f) * (m + n3b - 1));
    }
};

// the vertex elements one cache entry sums, with their coefficients
struct Terms {
    static constexpr int CAP = 32;
    int n = 0;
    size_t dst[CAP];
    float cf[CAP];
    void add(size_t d, float c) {
        if (n < CAP) { dst[n] = d; cf[n] = c; }
        ++n;
    }
};

// γ_ch(Ω, ν, ν') at finite arguments (ChanV::eval)
inline void gamma_terms(const Boxes& B, int ch, int m, int n, int np, float cf, Terms& t) {
    if (!in_b(m, B.n1)) return;
    t.add(B.d1(ch, m), cf);
    if (!in_b(m, B.n2b)) return;
    const bool i1 = in_f(n, B.n2f), i2 = in_f(np, B.n2f);
    if (i1 && i2) {
        t.add(B.d2(ch, m, n), cf);
        t.add(B.d2(ch, m, np), cf);
        if (in_b(m, B.n3b) && in_f(n, B.n3f) && in_f(np, B.n3f)) t.add(B.d3(ch, m, n, np), cf);
    } else if (i1) {
        t.add(B.d2(ch, m, n), cf);
    } else if (i2) {
        t.add(B.d2(ch, m, np), cf);
    }
}
// the γ part of the parallel-spin evaluation in channel ch with flags fl (gamma_psp)
inline void gpsp_terms(const Boxes& B, int m, int n, int np, int ch, int fl, float cf, Terms& t) {
    int a, b, c;
    if (fl & FL_GP) { convert(ch, PCH, m, n, np, a, b, c); gamma_terms(B, PCH, a, b, c, cf, t); }
    if (fl & FL_GT) { convert(ch, TCH, m, n, np, a, b, c); gamma_terms(B, TCH, a, b, c, cf, t); }
    if (fl & FL_GA) { convert(ch, ACH, m, n, np, a, b, c); gamma_terms(B, ACH, a, b, c, cf, t); }
}
// the same-channel K2(Ω, ν) + K3(Ω, ν, ω) of a whole F cache (one momentum)
inline void own_terms(const Boxes& B, int ch, int m, int a, int b, float cf, Terms& t) {
    t.add(B.d2(ch, m, a), cf);
    t.add(B.d3(ch, m, a, b), cf);
}
// cache slot s at (Ω = m, a, b): 0 Γpx (cache1_elem: F(pCh, xSp; γt, γa), the xSp evaluation at
// ν' -> Ω - ν'), 1 Γpp, 2 Γa, 3 Γt(d) = 2 Γt - Γa, 4..6 Fbase_p, Fbase_a, Fbase_t(d) (cache2_elem);
// `full` (one momentum): the F caches hold the same-channel K2 + K3 as well
inline void cache_terms(const Boxes& B, int s, int m, int a, int b, bool full, Terms& t) {
    switch (s) {
    case 0: gpsp_terms(B, m, a, bmf(m, b), PCH, FL_GT | FL_GA, -1.f, t); break;
    case 1:
    case 4:
        gpsp_terms(B, m, a, b, PCH, FL_GT | FL_GA, 1.f, t);
        if (s == 4 && full) own_terms(B, PCH, m, a, b, 1.f, t);
        break;
    case 2:
    case 5:
        gpsp_terms(B, m, a, b, ACH, FL_GP | FL_GT, 1.f, t);
        if (s == 5 && full) own_terms(B, ACH, m, a, b, 1.f, t);
        break;
    default:
        gpsp_terms(B, m, a, b, TCH, FL_GP | FL_GA, 2.f, t);
        gpsp_terms(B, m, a, b, ACH, FL_GP | FL_GT, -1.f, t);
        if (s == 6 && full) {
            own_terms(B, TCH, m, a, b, 2.f, t);
            own_terms(B, ACH, m, a, b, -1.f, t);
        }
    }
}
constexpr int kSlots = 7;
// cache slot -> cache array of the context (c[0], c[4..9])
inline int slot_cache(int s) { return s == 0 ? 0 : s + 3; }

// The transposed gather map sorted by destination: for vertex element d, the cache entries
// src[off[d] .. off[d + 1]) (slot s, entry o as s n3 + o, o = i + Nν3 (j + Nν3 iΩ)) and coefficients.
struct GatherMap {
    std::vector<int64_t> off;
    std::vector<int32_t> src;
    std::vector<float> cf;
};
// Returns false when a cache entry has more terms than Terms holds (not reached by the formulas above)
// or the entries do not fit int32.
inline bool gather_map(const Boxes& B, bool full, GatherMap& g) {
    const int NV = 2 * B.n3f, NW = 2 * B.n3b - 1;
    const size_t n3 = (size_t)NV * NV * NW, nd = B.ndst();
    if ((size_t)kSlots * n3 >= ((size_t)1 << 31)) return false;
    bool ok = true;
    auto each = [&](auto&& f) {
        Terms t;
        for (int s = 0; s < kSlots; ++s)
            for (int iW = 0; iW < NW; ++iW)
                for (int j = 0; j < NV; ++j)
                    for (int i = 0; i < NV; ++i) {
                        t.n = 0;
                        cache_terms(B, s, iW - (B.n3b - 1), i - B.n3f, j - B.n3f, full, t);
                        if (t.n > Terms::CAP) { ok = false; continue; }
                        const int32_t src = (int32_t)((size_t)s * n3 + (size_t)i + (size_t)NV * ((size_t)j + (size_t)NV * iW));
                        for (int q = 0; q < t.n; ++q) f(t.dst[q], src, t.cf[q]);
                    }
    };
    g.off.assign(nd + 1, 0);
    each([&](size_t d, int32_t, float) { ++g.off[d + 1]; });
    for (size_t d = 0; d < nd; ++d) g.off[d + 1] += g.off[d];
    g.src.resize((size_t)g.off[nd]);
    g.cf.resize((size_t)g.off[nd]);
    std::vector<int64_t> pos(g.off.begin(), g.off.end() - 1);
    each([&](size_t d, int32_t s, float c) {
        const int64_t p = pos[d]++;
        g.src[(size_t)p] = s;
        g.cf[(size_t)p] = c;
    });
    return ok;
}

}  // namespace adj

// ---------------------------------------------------------------------------------------
// device: the gather's transpose, out[d] = Σ_q cf[q] cbar[src[q]] over d's segment, one wave per d
__global__ __launch_bounds__(256) void k_adj_gather_t(double2* __restrict__ out, const int64_t* __restrict__ off,
                                                      const int* __restrict__ src, const float* __restrict__ cf,
                                                      const double2* __restrict__ cbar, size_t nd) {
    const size_t wid = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    if (wid >= nd) return;   // uniform over the wave
    double sr = 0.0, si = 0.0;
    for (int64_t q = off[wid] + lane; q < off[wid + 1]; q += WAVE) {
        const double2 x = cbar[src[q]];
        const double c = cf[q];
        sr += c * x.x;
        si += c * x.y;
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    if (lane == 0) out[wid] = cmk(sr, si);
}
// out[e + n p] += w a[e] for every momentum p (the transpose of the momentum average, w = 1 / NP)
__global__ void k_adj_bcast(double2* __restrict__ out, const double2* __restrict__ a, size_t n, int NP, double w) {
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n * NP) return;
    out[e] = cadd(out[e], cscale(a[e % n], w));
}
// forward t post-processing out_t = (raw_t + out_a) / 2, and its transpose on the cotangents
// (ḡ_t -> ḡ_t / 2, ḡ_a -> ḡ_a + ḡ_t / 2)
__global__ void k_adj_tpost(double2* __restrict__ t, const double2* __restrict__ a, size_t n) {
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e < n) t[e] = cscale(cadd(t[e], a[e]), 0.5);
}
__global__ void k_adj_tpost_t(double2* __restrict__ t, double2* __restrict__ a, size_t n) {
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double2 x = cscale(t[e], 0.5);
    t[e] = x;
    a[e] = cadd(a[e], x);
}

// One channel of the right operand at the linearisation point and in the tangent / cotangent
struct AdjR {
    int ch;
    double s1, s2;
    const double2* Pi;    // Π of the channel (device Π layout)
    const double2* dPi;   // its tangent (JVP)
    const double2* F0r;   // [ω, νp, Ω]
    const double2* Gam;   // p: Γpx [ω, νp, Ω]; a, t: Γr transposed [ω, ν, Ω]
    const double2* dGam;
    const double2* L;     // FL.γr.K3, FL.γa.K3 [ν, νp, Ω, P]
    const double2* La;
    const double2* dL;
    const double2* dLa;
    double2* out;         // dR (JVP) [ω, νp, Ω, P]
};
__device__ __forceinline__ double2 adj_ltilde(const double2* L, const double2* La, int t, size_t o) {
    const double2 x = L[o];
    return t ? csub(cscale(x, 2.0), La[o]) : x;
}
// JVP of the right-operand builder (k_k3_R): dR = s1 [dΠ (F0r + Γ~) + Π dΓ~] + s2 (dΠ L~ + Π dL~) 1(ω~ in box)
__global__ void k_adj_dR(K3Geom g, AdjR r) {
    const int NV = g.NV(), NW = g.NW();
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)NV * NV * NW * g.NP) return;
    const int k = (int)(e % NV), j = (int)(e / NV % NV);
    const size_t bb = e / ((size_t)NV * NV);
    const int iW = (int)(bb % NW), p = (int)(bb / NW), m = iW - (g.n3b - 1), w = k - g.n3f;
    const size_t pi = g.pidx(m, w, p);
    const double2 P1 = r.Pi[pi], dP1 = r.dPi[pi];
    const size_t kj = (size_t)k + (size_t)NV * ((size_t)j + (size_t)NV * iW);
    double2 x = cscale(cadd(cmul(dP1, cadd(r.F0r[kj], r.Gam[kj])), cmul(P1, r.dGam[kj])), r.s1);
    const int w_ = crossing(m, w, r.ch);
    if (in_f(w_, g.n3f)) {
        const size_t lo = (size_t)(w_ + g.n3f) + (size_t)NV * ((size_t)j + (size_t)NV * bb);
        const int t = r.ch == TCH;
        x = cadd(x, cscale(cadd(cmul(dP1, adj_ltilde(r.L, r.La, t, lo)), cmul(P1, adj_ltilde(r.dL, r.dLa, t, lo))), r.s2));
    }
    r.out[e] = x;
}

// A_r = cache_Fr at (ν = i, ω = k) of one batch: Fbase (the whole cache when `full`) plus the
// same-channel K2r(Ω, ν, P) + K3r(Ω, ν, ω, P) (t: 2 (t) - (a))
struct AdjAChan {
    int ch, full;
    const double2* Fb;    // Fbase [ν, ω, Ω]
    const double2* K2r;   // vertex K2 [Ω, ν, P] of the channel and of a
    const double2* K2a;
    const double2* K3r;   // vertex K3 [ν, νp, Ω, P]
    const double2* K3a;
};
struct AdjA {
    const double2* Fb; const double2* K2r; const double2* K2a; const double2* K3r; const double2* K3a;
    int full, t, NV, NW2;
    __device__ __forceinline__ double2 operator()(int i, int k) const {
        const size_t o = (size_t)i + (size_t)NV * k;
        const double2 x = Fb[o];
        if (full) return x;
        double2 own = cadd(K2r[(size_t)NW2 * i], K3r[o]);
        if (t) own = csub(cscale(own, 2.0), cadd(K2a[(size_t)NW2 * i], K3a[o]));
        return cadd(x, own);
    }
};
__device__ __forceinline__ AdjA adj_a_bind(const K3Geom& g, const AdjAChan& c, int bb) {
    const int NV = g.NV(), NW = g.NW();
    const int iW = bb % NW, p = bb / NW, m = iW - (g.n3b - 1);
    const int NW2 = 2 * g.n2b - 1, NV2 = 2 * g.n2f;
    const size_t nn = (size_t)NV * NV;
    const size_t k2o = (size_t)(m + g.n2b - 1) + (size_t)NW2 * ((size_t)(g.n2f - g.n3f) + (size_t)NV2 * p);
    AdjA a;
    a.Fb = c.Fb + nn * iW;
    a.full = c.full;
    a.t = c.ch == TCH;
    a.NV = NV;
    a.NW2 = NW2;
    a.K2r = c.full ? nullptr : c.K2r + k2o;
    a.K2a = (c.full || !a.t) ? nullptr : c.K2a + k2o;
    a.K3r = c.full ? nullptr : c.K3r + nn * bb;
    a.K3a = (c.full || !a.t) ? nullptr : c.K3a + nn * bb;
    return a;
}

// JVP GEMM: left operand [dA | A] (inner index kk < Nν3: the tangent), right operand [R; dR]
struct AdjJvpA {
    K3Geom g; AdjAChan d, p;
    struct Bound {
        AdjA d, p; int NV;
        __device__ __forceinline__ double2 operator()(int i, int kk) const { return kk < NV ? d(i, kk) : p(i, kk - NV); }
    };
    __device__ __forceinline__ Bound bind(int b) const { return Bound{adj_a_bind(g, d, b), adj_a_bind(g, p, b), g.NV()}; }
};
struct AdjJvpB {
    K3Geom g; const double2* R; const double2* dR;
    struct Bound {
        const double2* R; const double2* dR; int NV;
        __device__ __forceinline__ double2 operator()(int kk, int j) const {
            return kk < NV ? R[(size_t)kk + (size_t)NV * j] : dR[(size_t)(kk - NV) + (size_t)NV * j];
        }
    };
    __device__ __forceinline__ Bound bind(int b) const {
        const size_t off = (size_t)g.NV() * g.NV() * b;
        return Bound{R + off, dR + off, g.NV()};
    }
};
// d raw = T v + d self (d self = dFL_r, t: 2 dFL_t - dFL_a)
struct AdjJvpE {
    K3Geom g; double T; int t; const double2* dL; const double2* dLa; double2* out;
    struct Bound {
        const double2* dL; const double2* dLa; double2* out; double T; int t, NV;
        __device__ __forceinline__ void operator()(int i, int j, double2 v) const {
            const size_t o = (size_t)i + (size_t)NV * j;
            out[o] = cadd(cscale(v, T), adj_ltilde(dL, dLa, t, o));
        }
    };
    __device__ __forceinline__ Bound bind(int b) const {
        const size_t off = (size_t)g.NV() * g.NV() * b;
        return Bound{dL + off, t ? dLa + off : nullptr, out + off, T, t, g.NV()};
    }
};

// VJP GEMMs.  R̄ = T A^T Ḡ: left operand A transposed; Ā = T Ḡ R^T: right operand R transposed.
struct AdjAT {
    K3Geom g; AdjAChan c;
    struct Bound {
        AdjA a;
        __device__ __forceinline__ double2 operator()(int r, int kk) const { return a(kk, r); }
    };
    __device__ __forceinline__ Bound bind(int b) const { return Bound{adj_a_bind(g, c, b)}; }
};
// dense Nν3 x Nν3 matrix per batch, element (a, b) at a + Nν3 b (TR: at b + Nν3 a)
template <bool TR>
struct AdjDense {
    K3Geom g; const double2* p;
    struct Bound {
        const double2* p; int NV;
        __device__ __forceinline__ double2 operator()(int a, int b) const {
            return TR ? p[(size_t)b + (size_t)NV * a] : p[(size_t)a + (size_t)NV * b];
        }
    };
    __device__ __forceinline__ Bound bind(int b) const { return Bound{p + (size_t)g.NV() * g.NV() * b, g.NV()}; }
};
struct AdjOut {
    K3Geom g; double2* p; double scale;
    struct Bound {
        double2* p; double scale; int NV;
        __device__ __forceinline__ void operator()(int i, int j, double2 v) const { p[(size_t)i + (size_t)NV * j] = cscale(v, scale); }
    };
    __device__ __forceinline__ Bound bind(int b) const { return Bound{p + (size_t)g.NV() * g.NV() * b, scale, g.NV()}; }
};

// The right-operand builder's transpose.  Per channel slot: the linearisation point (Π, F0r, Γ~, FL)
// and R̄ [ω, νp, Ω, P].
struct AdjRT {
    int ch;
    double s1, s2;
    const double2* Pi;
    const double2* F0r;
    const double2* Gam;
    const double2* L;
    const double2* La;
    const double2* Rb;    // R̄
};
__device__ __forceinline__ const AdjRT& adj_pick(const AdjRT& a, const AdjRT& b, const AdjRT& c, int s) {
    return s == 0 ? a : s == 1 ? b : c;
}
// Π̄(Ω, ω, P) = Σ_r Σ_νp R̄_r[ω, νp] (s1 (F0r + Γ~) + s2 L~(ω~) 1(ω~ in box)) over the channels of the
// bubble (Πpp: p; Πph: a and t), one thread per (ω, Ω, P); rp = slot of p, ra / rt = slots of a, t
__global__ void k_adj_pibar(K3Geom g, AdjRT rp, AdjRT ra, AdjRT rt, double2* __restrict__ pipp, double2* __restrict__ piph) {
    const int NV = g.NV(), NW = g.NW();
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)NV * NW * g.NP) return;
    const int k = (int)(e % NV);
    const size_t bb = e / NV;
    const int iW = (int)(bb % NW), p = (int)(bb / NW), m = iW - (g.n3b - 1), w = k - g.n3f;
    double2 acc[2] = {czero(), czero()};
    for (int s = 0; s < 3; ++s) {
        const AdjRT& r = adj_pick(rp, ra, rt, s);
        const int w_ = crossing(m, w, r.ch);
        const bool inb = in_f(w_, g.n3f);
        const int t = r.ch == TCH;
        double2 sum = czero();
        for (int j = 0; j < NV; ++j) {
            const size_t kj = (size_t)k + (size_t)NV * ((size_t)j + (size_t)NV * iW);
            double2 f = cscale(cadd(r.F0r[kj], r.Gam[kj]), r.s1);
            if (inb) {
                const size_t lo = (size_t)(w_ + g.n3f) + (size_t)NV * ((size_t)j + (size_t)NV * bb);
                f = cadd(f, cscale(adj_ltilde(r.L, r.La, t, lo), r.s2));
            }
            sum = cadd(sum, cmul(r.Rb[(size_t)k + (size_t)NV * ((size_t)j + (size_t)NV * bb)], f));
        }
        acc[s == 0 ? 0 : 1] = cadd(acc[s == 0 ? 0 : 1], sum);
    }
    const size_t pi = g.pidx(m, w, p);
    pipp[pi] = acc[0];
    piph[pi] = acc[1];
}
// Γ~̄ = s1 Σ_P Π(ω) R̄[ω, νp] into the cache cotangents: p -> Γpx (slot 0, [ω, νp]); a, t -> Γa, Γt(d)
// (slots 2, 3, [ν, ω]: the transposed copies' cotangents land transposed), one thread per (ω, νp, Ω)
__global__ void k_adj_gambar(K3Geom g, AdjRT rp, AdjRT ra, AdjRT rt, double2* __restrict__ cbar) {
    const int NV = g.NV(), NW = g.NW();
    const size_t n3 = (size_t)NV * NV * NW;
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n3) return;
    const int k = (int)(e % NV), j = (int)(e / NV % NV), iW = (int)(e / ((size_t)NV * NV));
    const int m = iW - (g.n3b - 1), w = k - g.n3f;
    for (int s = 0; s < 3; ++s) {
        const AdjRT& r = adj_pick(rp, ra, rt, s);
        double2 sum = czero();
        for (int p = 0; p < g.NP; ++p) {
            const size_t bb = (size_t)iW + (size_t)NW * p;
            sum = cadd(sum, cmul(r.Pi[g.pidx(m, w, p)], r.Rb[(size_t)k + (size_t)NV * ((size_t)j + (size_t)NV * bb)]));
        }
        sum = cscale(sum, r.s1);
        if (r.ch == PCH) cbar[e] = sum;
        else cbar[(r.ch == ACH ? 2 : 3) * n3 + (size_t)j + (size_t)NV * ((size_t)k + (size_t)NV * iW)] = sum;
    }
}
// FL̄ per element (row i = ω~, νp, Ω, P): the self terms (Ḡ_r; t: 2 Ḡ_t to t, -Ḡ_t to a) and the
// builder's s2 Π(ω) R̄[ω, νp] at the R row ω whose crossing is ω~ (a, t: ω = ω~; p: ω = Ω - ω~ - 1)
__global__ void k_adj_flbar(K3Geom g, AdjRT rp, AdjRT ra, AdjRT rt, const double2* __restrict__ gp,
                            const double2* __restrict__ ga, const double2* __restrict__ gt, double2* __restrict__ lp,
                            double2* __restrict__ la, double2* __restrict__ lt) {
    const int NV = g.NV(), NW = g.NW();
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)NV * NV * NW * g.NP) return;
    const int i = (int)(e % NV), j = (int)(e / NV % NV);
    const size_t bb = e / ((size_t)NV * NV);
    const int iW = (int)(bb % NW), p = (int)(bb / NW), m = iW - (g.n3b - 1), wt = i - g.n3f;
    double2 c[3];
    for (int s = 0; s < 3; ++s) {
        const AdjRT& r = adj_pick(rp, ra, rt, s);
        const int w = r.ch == PCH ? m - wt - 1 : wt;
        c[s] = czero();
        if (in_f(w, g.n3f)) {
            const size_t ro = (size_t)(w + g.n3f) + (size_t)NV * ((size_t)j + (size_t)NV * bb);
            c[s] = cscale(cmul(r.Pi[g.pidx(m, w, p)], r.Rb[ro]), r.s2);
        }
    }
    lp[e] = cadd(gp[e], c[0]);
    la[e] = csub(cadd(ga[e], c[1]), cadd(gt[e], c[2]));
    lt[e] = cscale(cadd(gt[e], c[2]), 2.0);
}
// Fbase̅ = Σ_P Ā into cache slot `slot` [ν, ω, Ω], one thread per (ν, ω, Ω)
__global__ void k_adj_fbase(K3Geom g, const double2* __restrict__ Ab, double2* __restrict__ cbar) {
    const int NV = g.NV(), NW = g.NW();
    const size_t nn = (size_t)NV * NV, n3 = nn * NW;
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n3) return;
    double2 sum = czero();
    for (int p = 0; p < g.NP; ++p) sum = cadd(sum, Ab[e + n3 * p]);
    cbar[e] = sum;
}
// the same-channel terms of A (lattice): K3̄_r += c Ā, K2̄_r(Ω, ν, P) += c Σ_ω Ā[ν, ω], one thread per
// (ν, Ω, P); c = 1 (p, a), 2 (t into t) or -1 (t into a)
__global__ void k_adj_own(K3Geom g, const double2* __restrict__ Ab, double c, double2* __restrict__ K2, double2* __restrict__ K3) {
    const int NV = g.NV(), NW = g.NW();
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)NV * NW * g.NP) return;
    const int i = (int)(e % NV);
    const size_t bb = e / NV;
    const int iW = (int)(bb % NW), p = (int)(bb / NW), m = iW - (g.n3b - 1);
    const size_t nn = (size_t)NV * NV;
    double2 rs = czero();
    for (int k = 0; k < NV; ++k) {
        const size_t o = nn * bb + (size_t)i + (size_t)NV * k;
        const double2 x = cscale(Ab[o], c);
        K3[o] = cadd(K3[o], x);
        rs = cadd(rs, x);
    }
    const int NW2 = 2 * g.n2b - 1, NV2 = 2 * g.n2f;
    const size_t k2 = (size_t)(m + g.n2b - 1) + (size_t)NW2 * ((size_t)(i - g.n3f + g.n2f) + (size_t)NV2 * p);
    K2[k2] = cadd(K2[k2], rs);
}

}  // namespace fdga
