// fdga_forms.h -- the other BSE forms of the K1 and K2 classes and the mfRG caches:
//   * the is_mfRG = Val(true) forms of BSE_K1!/BSE_K2! (src/BSEa/BSEa_K1.jl:29-35, BSEa_K2.jl:70-78;
//     lattice src/nonlocal/BSEa/BSEa_K1.jl:29-37, BSEa_K2.jl:70-78),
//   * BSE_K1_new!/BSE_K2_new! with their lattice is_mfRG forms (src/BSEa/BSEa_K1.jl:57-100,
//     BSEa_K2.jl:108-151; src/nonlocal/BSEa/BSEa_K1.jl:62-108, BSEa_K2.jl:108-155),
//   * build_K3_cache_mfRG! (src/nonlocal_2/build_K3_cache.jl:97-164).
// Each K1/K2 form is a gather-GEMV over the bubble mesh of the shape
//   out(Ω, [ν,] P) = T s Σ_ω [X(ω) Π(Ω, ω, P) + Y(ω) Π0(Ω, ω, P)]
// with X, Y products of point evaluations of F, F0 and FL.  Within the frequency window
// |ω| <= Wc = NK + |Ω| + |ν| + 2 (NK bounds every box of F, F0 and its tables) X and Y are summed
// point by point; beyond it neither depends on ω any more (every converted or crossed argument lies
// outside every box, so only bare vertices and K1/K2 classes at the fixed arguments remain), and the
// far sums are X_far Σ_far Π + Y_far Σ_far Π0 with the tail sums of Π and Π0 (k_tail_sums layout).
#pragma once
#include "fdga_kernels.h"

namespace fdga {

enum { FORM_K1_MFRG = 0, FORM_K2_MFRG = 1, FORM_K1_NEW = 2, FORM_K2_NEW = 3 };

struct FormArgs {
    VtxV F;                    // target vertex
    VtxV FL;                   // left vertex (reference vertex zero)
    F0V f0;                    // reference vertex
    const double2* Pi;         // Π of the channel [ω, Ω, P]
    const double2* Pi0;        // Π0 of the channel
    const double2* Pt;         // their tail sums on the band
    const double2* P0t;
    double2* out;              // Fbuff.γr.K1 or K2
    const double2* outa;       // Fbuff.γa (the t channel's post-processing reads the finished a channel)
    const double2* Lr;         // FL.γr.K2 and FL.γa.K2 (added by the K2 is_mfRG form)
    const double2* La;
    double2 U;                 // bare_vertex(F, Sp) of the _new forms
    double scale;              // T s
    int nl;                    // s-wave lattice solver (its forms differ from the local ones)
    int mfrg;                  // is_mfRG of the _new forms
    int n1, n2b, n2f, nWP, nVP, wb, NP, NK;
    SymCls sym;
};

// X(ω), Y(ω) of form FORM in channel CH at (Ω, ν, ω, P); ν = FDGA_INF for the K1 forms
template <int FORM, int CH>
__device__ __forceinline__ void form_xy(const FormArgs& a, int m, int v, int w, int p, double2& X, double2& Y) {
    constexpr int SP = ChSpin<CH>::SP;
    const int wc = crossing(m, w, CH);
    if (FORM == FORM_K1_MFRG) {
        // F0(Ω, νInf, ω) Π0 FL(Ω, ω~, νInf); the local form in aCh / pSp for every channel
        X = czero();
        if (a.nl) Y = cmul(a.f0.eval(m, FDGA_INF, w, CH, SP), a.FL.eval(m, wc, FDGA_INF, p, CH, SP, FL_ALL));
        else Y = cmul(a.f0.eval(m, FDGA_INF, w, ACH, PSP), a.FL.eval(m, wc, FDGA_INF, p, ACH, PSP, FL_ALL));
        return;
    }
    if (FORM == FORM_K2_MFRG) {
        // (F0(Ω, ν, .) - F0(Ω, νInf, .)) Π0 FL(Ω, ., νInf): crossing on the right vertex (local) or
        // on the left one (lattice)
        X = czero();
        const int wl = a.nl ? wc : w, wr = a.nl ? w : wc;
        const double2 d0 = csub(a.f0.eval(m, v, wl, CH, SP), a.f0.eval(m, FDGA_INF, wl, CH, SP));
        Y = cmul(d0, a.FL.eval(m, wr, FDGA_INF, p, CH, SP, FL_ALL));
        return;
    }
    double2 f, f0;
    if (FORM == FORM_K1_NEW) {
        f = a.F.eval(m, FDGA_INF, w, p, CH, SP, FL_ALL);
        f0 = a.f0.eval(m, FDGA_INF, w, CH, SP);
    } else {
        f = csub(a.F.eval(m, v, w, p, CH, SP, FL_ALL), a.F.eval(m, FDGA_INF, w, p, CH, SP, FL_ALL));
        f0 = csub(a.f0.eval(m, v, w, CH, SP), a.f0.eval(m, FDGA_INF, w, CH, SP));
    }
    if (a.mfrg && a.nl) {      // (F - F0) Π U
        X = cmul(csub(f, f0), a.U);
        Y = czero();
    } else {                   // F Π U - F0 Π0 U
        X = cmul(f, a.U);
        Y = cneg(cmul(f0, a.U));
    }
}

// one wave per output class (K1: classes of (Ω, P); K2: of (Ω, ν, P)); lanes over the window
template <int FORM, int CH>
__device__ __forceinline__ void form_wave(int wid, int lane, const FormArgs& a) {
    constexpr bool K2 = (FORM == FORM_K2_MFRG || FORM == FORM_K2_NEW);
    const int NWo = K2 ? 2 * a.n2b - 1 : 2 * a.n1 - 1, NVo = K2 ? 2 * a.n2f : 1;
    if (wid >= a.sym.count(NWo * NVo * a.NP)) return;
    const int o0 = a.sym.rep(wid);
    const int iW = o0 % NWo, iv = (o0 / NWo) % NVo, p = o0 / (NWo * NVo);
    const int m = iW - (K2 ? a.n2b - 1 : a.n1 - 1);
    const int v = K2 ? iv - a.n2f : FDGA_INF;
    const int Wc = window(a.NK, m, K2 ? v : 0);
    const int nVP = a.nVP;
    const size_t row = (size_t)(m + a.nWP - 1) + (size_t)(2 * a.nWP - 1) * p;
    const double2* __restrict__ pi = a.Pi + (size_t)(2 * nVP) * row;
    const double2* __restrict__ pi0 = a.Pi0 + (size_t)(2 * nVP) * row;
    const int lo = max(0, nVP - Wc), hi = min(2 * nVP - 1, nVP + Wc);
    double sr = 0.0, si = 0.0;
    for (int iw = lo + lane; iw <= hi; iw += WAVE) {
        double2 X, Y;
        form_xy<FORM, CH>(a, m, v, iw - nVP, p, X, Y);
        const double2 t = cadd(cmul(X, pi[iw]), cmul(Y, pi0[iw]));
        sr += t.x;
        si += t.y;
    }
    double2 X, Y, xl, xr, yl, yr;
    form_xy<FORM, CH>(a, m, v, lane == 0 ? -Wc - 1 : Wc + 1, p, X, Y);
    far_pair(X, xl, xr);
    far_pair(Y, yl, yr);
    sr = wave_sum(sr);
    si = wave_sum(si);
    double2 acc = cmk(sr, si);
    if (lo > 0)
        acc = cadd(acc, cadd(cmul(xl, tail_left(a.Pt, row, nVP, a.wb, lo)), cmul(yl, tail_left(a.P0t, row, nVP, a.wb, lo))));
    if (hi < 2 * nVP - 1)
        acc = cadd(acc, cadd(cmul(xr, tail_right(a.Pt, row, nVP, a.wb, hi)), cmul(yr, tail_right(a.P0t, row, nVP, a.wb, hi))));
    const double2 raw = cscale(acc, a.scale);
    int b, e;
    a.sym.range(wid, b, e);
    for (int k = b + lane; k < e; k += WAVE) {
        const int code = a.sym.member(wid, k), o = code >> 2;
        double2 val = sym_apply(code & 3, raw);
        if (FORM == FORM_K2_MFRG) {   // + FL.K2 (t: 2 FL.t - FL.a), src/BSEa/BSEa_K2.jl:95-100
            if (CH == TCH) val = cadd(val, csub(cscale(a.Lr[o], 2.0), a.La[o]));
            else val = cadd(val, a.Lr[o]);
        }
        if (CH == TCH) val = cscale(cadd(val, a.outa[o]), 0.5);   // src/BSE_templates.jl:29-33
        a.out[o] = val;
    }
}
template <int FORM, int CH>
__global__ __launch_bounds__(256) void k_form(FormArgs a) {
    form_wave<FORM, CH>((int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) / WAVE), threadIdx.x & (WAVE - 1), a);
}

// ---------------------------------------------------------------------------------------
// build_K3_cache_mfRG! (src/nonlocal_2/build_K3_cache.jl:97-164), one thread per (x, y, Ω):
//   Γpx  [ω, νp, Ω] = F(Ω, ω, νp, pCh, xSp; γp = false) - F.F0(Ω, ω, νp, pCh, xSp; γp = false)
//   Γr   [ν, ω, Ω]  = F(Ω, ν, ω, r, pSp; γr = false) - F.F0(Ω, ν, ω, r, pSp; γr = false), t: 2 Γt - Γa
//   Fbase_r [ν, ω, Ω] = F0(Ω, ν, ω, r, pSp) - F0(Ω, νInf, ω, r, pSp), t: 2 (t) - (a)   (first call only)
// F(...; γr = false) keeps the reference vertex, which F.F0(...; γr = false) removes again: for a
// RefVertex this cancels except in Γpx, where F reaches its RefVertex through the xSp crossing and
// F.F0 through its own Fp_x table; for a local reference Vertex F.F0 also drops its own γr.  The
// cross channels are P-averaged and the reference vertex is local, so no cache depends on P.  The F0
// caches of loop 1 are not touched (they keep build_K3_cache!'s values).
__device__ __forceinline__ double2 f0_vertex_eval(const F0V& f0, int m, int n, int np, int ch, int sp, int fl) {
    if (f0.kind == 0) return f0.ref.eval(m, n, np, ch, sp);
    VtxV v;
    v.g[0] = f0.g[0]; v.g[1] = f0.g[1]; v.g[2] = f0.g[2];
    v.f0.kind = 0;
    v.f0.ref = f0.ref;
    return v.eval(m, n, np, 0, ch, sp, fl);
}
__global__ __launch_bounds__(256) void k_cache_mfrg(CachePtrs out, VtxV F, int n3b, int n3f, int first) {
    const int NV = 2 * n3f, NW = 2 * n3b - 1;
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)NV * NV * NW) return;
    const int i = (int)(e % NV), j = (int)((e / NV) % NV), iW = (int)(e / ((size_t)NV * NV));
    const int m = iW - (n3b - 1), a = i - n3f, b = j - n3f;
    const int fx = FL_F0 | FL_GT | FL_GA, fa = FL_F0 | FL_GP | FL_GT, ft = FL_F0 | FL_GP | FL_GA;
    out.c[0][e] = csub(F.eval(m, a, b, -1, PCH, XSP, fx), f0_vertex_eval(F.f0, m, a, b, PCH, XSP, fx));
    out.c[4][e] = csub(F.eval(m, a, b, -1, PCH, PSP, fx), f0_vertex_eval(F.f0, m, a, b, PCH, PSP, fx));
    const double2 ga = csub(F.eval(m, a, b, -1, ACH, PSP, fa), f0_vertex_eval(F.f0, m, a, b, ACH, PSP, fa));
    const double2 gt = csub(F.eval(m, a, b, -1, TCH, PSP, ft), f0_vertex_eval(F.f0, m, a, b, TCH, PSP, ft));
    out.c[5][e] = ga;
    out.c[6][e] = csub(cscale(gt, 2.0), ga);
    if (!first) return;
    const double2 dp = csub(F.f0.eval_psp(m, a, b, PCH), F.f0.eval_psp(m, FDGA_INF, b, PCH));
    const double2 da = csub(F.f0.eval_psp(m, a, b, ACH), F.f0.eval_psp(m, FDGA_INF, b, ACH));
    const double2 dt = csub(F.f0.eval_psp(m, a, b, TCH), F.f0.eval_psp(m, FDGA_INF, b, TCH));
    out.c[7][e] = dp;
    out.c[8][e] = da;
    out.c[9][e] = csub(cscale(dt, 2.0), da);
}
// the transposed copies [ω, ν, Ω] of the a and t Γ caches that BSE_K3!'s right operands read
__global__ __launch_bounds__(256) void k_cache_transpose(double2* __restrict__ ta, double2* __restrict__ tt,
                                                         const double2* __restrict__ ga, const double2* __restrict__ gt,
                                                         int NV, int NW) {
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)NV * NV * NW) return;
    const int j = (int)(e % NV), i = (int)((e / NV) % NV);
    const size_t src = (size_t)i + (size_t)NV * ((size_t)j + (size_t)NV * (e / ((size_t)NV * NV)));
    ta[e] = ga[src];
    tt[e] = gt[src];
}

}  // namespace fdga
