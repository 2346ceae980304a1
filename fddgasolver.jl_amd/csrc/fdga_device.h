// fdga_device.h -- device-side types and the vertex point evaluator for gfx950.
//
// Frequencies are integer Matsubara indices: bosonic m (Ω = 2πT m), fermionic n
// (ν = πT(2n+1)).  A box of size N holds m in [-(N-1), N-1] and n in [-N, N-1]
// (MatsubaraFunctions.MatsubaraMesh).  FDGA_INF is νInf (src/types.jl:134-144).
//
// Storage on the device:
//   K1 [Ω, P]           column-major (reference layout)
//   K2 [Ω, ν, P]        column-major (reference layout)
//   K3 [ν, νp, Ω, P]    one column-major ν x νp matrix per (Ω, P): the BSE GEMM operand layout
//   RefVertex tables [Ω, ν, νp] (reference layout)
// Cross-channel lookups of the s-wave solver use the P-averaged copies A1/A2/A3
// (src/nonlocal/swave.jl:32-88); for the local solver (NP == 1) they alias K1/K2/K3.
#pragma once
#include <hip/hip_runtime.h>
#include <climits>

#define FDGA_INF INT_MIN

namespace fdga {

enum { PCH = 0, TCH = 1, ACH = 2 };
enum { PSP = 0, XSP = 1, DSP = 2 };
enum { FL_F0 = 1, FL_GP = 2, FL_GT = 4, FL_GA = 8, FL_ALL = 15 };

// ---------------------------------------------------------------------------------------
// complex FP64 helpers on double2 (x = re, y = im)
__host__ __device__ __forceinline__ double2 cmk(double r, double i) { double2 z; z.x = r; z.y = i; return z; }
__host__ __device__ __forceinline__ double2 cadd(double2 a, double2 b) { return cmk(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ double2 csub(double2 a, double2 b) { return cmk(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ double2 cmul(double2 a, double2 b) { return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ __forceinline__ double2 cscale(double2 a, double s) { return cmk(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ double2 cneg(double2 a) { return cmk(-a.x, -a.y); }
__host__ __device__ __forceinline__ double2 czero() { return cmk(0.0, 0.0); }
__host__ __device__ __forceinline__ double2 cinv(double2 a) {
    double d = a.x * a.x + a.y * a.y;
    return cmk(a.x / d, -a.y / d);
}
// the same on float2 (the FP32 form of the lattice bubble)
__host__ __device__ __forceinline__ float2 cmkf(float r, float i) { float2 z; z.x = r; z.y = i; return z; }
__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return cmkf(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return cmkf(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) { return cmkf(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) { return cmkf(a.x * s, a.y * s); }
// a complex number of type CT (double2 or float2) from FP64 parts
template <class CT>
__host__ __device__ __forceinline__ CT cmk_as(double r, double i) { CT z; z.x = r; z.y = i; return z; }

// ---------------------------------------------------------------------------------------
// Matsubara index arithmetic; νInf absorbs
__host__ __device__ __forceinline__ bool in_b(int m, int N) { return m != FDGA_INF && m >= -(N - 1) && m <= N - 1; }
__host__ __device__ __forceinline__ bool in_f(int n, int N) { return n != FDGA_INF && n >= -N && n <= N - 1; }
__host__ __device__ __forceinline__ int bpf(int m, int n) { return (m == FDGA_INF || n == FDGA_INF) ? FDGA_INF : m + n; }      // Ω + ν
__host__ __device__ __forceinline__ int bmf(int m, int n) { return (m == FDGA_INF || n == FDGA_INF) ? FDGA_INF : m - n - 1; }  // Ω - ν
__host__ __device__ __forceinline__ int fmf(int a, int b) { return (a == FDGA_INF || b == FDGA_INF) ? FDGA_INF : a - b; }      // ν - ν'
__host__ __device__ __forceinline__ int fpf(int a, int b) { return (a == FDGA_INF || b == FDGA_INF) ? FDGA_INF : a + b + 1; }  // ν + ν'
__host__ __device__ __forceinline__ int crossing(int m, int w, int ch) { return ch == PCH ? bmf(m, w) : w; }  // src/BSE_templates.jl:4-6

// _convert_channel (src/convention.jl:4-37)
__host__ __device__ __forceinline__ void convert(int from, int to, int A, int b, int bp, int& o0, int& o1, int& o2) {
    if (from == to) { o0 = A; o1 = b; o2 = bp; return; }
    if (from == PCH) {
        if (to == TCH) { o0 = fmf(bmf(A, b), bp); o1 = bp; o2 = b; return; }
        o0 = fmf(b, bp); o1 = bmf(A, b); o2 = bp; return;                       // p -> a
    }
    if (from == TCH) {
        if (to == PCH) { o0 = fpf(bpf(A, b), bp); o1 = bp; o2 = b; return; }
        o0 = fmf(bp, b); o1 = bpf(A, b); o2 = b; return;                        // t -> a
    }
    if (to == PCH) { o0 = fpf(bpf(A, bp), b); o1 = bpf(A, bp); o2 = bp; return; } // a -> p
    o0 = fmf(b, bp); o1 = bp; o2 = bpf(A, bp);                                  // a -> t
}

// ---------------------------------------------------------------------------------------
// one reducible channel (K1, K2, K3), possibly momentum resolved
struct ChanV {
    const double2* K1;
    const double2* K2;
    const double2* K3;
    const double2* A1;
    const double2* A2;
    const double2* A3;
    int n1, n2b, n2f, n3b, n3f;
    int NP;
    int k3f;   // K3 holds FP32 values (the FP32 lattice iteration's vertex) in the first half of its bytes

    __host__ __device__ __forceinline__ size_t s1() const { return (size_t)(2 * n1 - 1); }
    __host__ __device__ __forceinline__ size_t s2() const { return (size_t)(2 * n2b - 1) * (2 * n2f); }
    __host__ __device__ __forceinline__ size_t s3() const { return (size_t)(2 * n3f) * (2 * n3f) * (2 * n3b - 1); }
    __host__ __device__ __forceinline__ size_t i1(int m) const { return (size_t)(m + n1 - 1); }
    __host__ __device__ __forceinline__ size_t i2(int m, int n) const { return (size_t)(m + n2b - 1) + (size_t)(2 * n2b - 1) * (n + n2f); }
    __host__ __device__ __forceinline__ size_t i3(int m, int n, int np) const {
        return (size_t)(n + n3f) + (size_t)(2 * n3f) * ((size_t)(np + n3f) + (size_t)(2 * n3f) * (m + n3b - 1));
    }
    __device__ __forceinline__ double2 k1(int m, int p) const { return p < 0 ? A1[i1(m)] : K1[i1(m) + s1() * p]; }
    __device__ __forceinline__ double2 k2(int m, int n, int p) const { return p < 0 ? A2[i2(m, n)] : K2[i2(m, n) + s2() * p]; }
    __device__ __forceinline__ double2 k3(int m, int n, int np, int p) const {
        if (k3f && p >= 0) return k3p<true>(m, n, np, p);
        return p < 0 ? A3[i3(m, n, np)] : K3[i3(m, n, np) + s3() * p];   // (one load of a selected address)
    }
    // K3 at momentum p >= 0 in the form K32 (== k3f): loops over K3 elements branch on the form once,
    // outside, and issue one load per element
    template <bool K32>
    __device__ __forceinline__ double2 k3p(int m, int n, int np, int p) const {
        const size_t i = i3(m, n, np) + s3() * p;
        if (K32) { const float2 v = reinterpret_cast<const float2*>(K3)[i]; return cmk(v.x, v.y); }
        return K3[i];
    }

    // γ(Ω, ν, νp[, P]) with all classes (src/channel.jl:220-339; src/nonlocal/channel.jl:90-226)
    __device__ __forceinline__ double2 eval(int m, int n, int np, int p) const {
        double2 v = czero();
        if (!in_b(m, n1)) return v;
        v = k1(m, p);
        if (!in_b(m, n2b)) return v;
        const bool finite_n = n != FDGA_INF, finite_np = np != FDGA_INF;
        const bool i1_ = finite_n && in_f(n, n2f), i2_ = finite_np && in_f(np, n2f);
        if (finite_n && finite_np) {
            if (i1_ && i2_) {
                v = cadd(v, cadd(k2(m, n, p), k2(m, np, p)));
                if (in_b(m, n3b) && in_f(n, n3f) && in_f(np, n3f)) v = cadd(v, k3(m, n, np, p));
            } else if (i1_) {
                v = cadd(v, k2(m, n, p));
            } else if (i2_) {
                v = cadd(v, k2(m, np, p));
            }
        } else if (!finite_n && finite_np) {
            if (i2_) v = cadd(v, k2(m, np, p));
        } else if (finite_n && !finite_np) {
            if (i1_) v = cadd(v, k2(m, n, p));
        }
        return v;
    }
};

// RefVertex U + tabulated Fp_p, Fp_x, Ft_p, Ft_x (src/refvertex.jl:1-216)
struct RefV {
    double2 U;
    int tb, tf;                   // table box (tb == 0: no tables)
    const double2* tab[4];        // Fp_p, Fp_x, Ft_p, Ft_x  [Ω, ν, νp]

    __device__ __forceinline__ double2 t(int k, int m, int n, int np) const {
        if (tb == 0 || !(in_b(m, tb) && in_f(n, tf) && in_f(np, tf))) return czero();
        size_t i = (size_t)(m + tb - 1) + (size_t)(2 * tb - 1) * ((size_t)(n + tf) + (size_t)(2 * tf) * (np + tf));
        return tab[k][i];
    }
    __device__ __forceinline__ double2 bare(int sp) const { return sp == XSP ? cneg(U) : U; }
    __device__ __forceinline__ double2 eval_p(int m, int n, int np, int ch) const {  // pSp, finite
        if (ch == PCH) return cadd(t(0, m, n, np), U);
        if (ch == TCH) return cadd(t(2, m, n, np), U);
        return cadd(cneg(t(3, m, np, n)), U);
    }
    __device__ __forceinline__ double2 eval_x(int m, int n, int np, int ch) const {  // xSp, finite
        if (ch == PCH) return csub(t(1, m, n, np), U);
        if (ch == TCH) return csub(t(3, m, n, np), U);
        return csub(cneg(t(2, m, np, n)), U);
    }
    __device__ __forceinline__ double2 eval(int m, int n, int np, int ch, int sp) const {
        if (n == FDGA_INF || np == FDGA_INF) return bare(sp);
        if (sp == PSP) return eval_p(m, n, np, ch);
        if (sp == XSP) return eval_x(m, n, np, ch);
        return cadd(cscale(eval_p(m, n, np, ch), 2.0), eval_x(m, n, np, ch));
    }
};

// channel selection by a run-time index without indexing the descriptor array dynamically
// (a dynamic index would move the whole kernel-argument descriptor to scratch memory)
__device__ __forceinline__ double2 chan_eval(const ChanV* g, int r, int m, int n, int np, int p) {
    if (r == PCH) return g[0].eval(m, n, np, p);
    if (r == TCH) return g[1].eval(m, n, np, p);
    return g[2].eval(m, n, np, p);
}

// γ part of the parallel-spin Vertex evaluator (src/vertex.jl:209-284; NL s-wave
// src/nonlocal/vertex.jl:216-379): same channel at momentum p, other channels averaged
__device__ __forceinline__ double2 gamma_psp(const ChanV* g, int m, int n, int np, int p, int ch, int fl) {
    double2 v = czero();
    if (n != FDGA_INF && np != FDGA_INF) {
        int a, b, c;
        if (fl & FL_GP) { convert(ch, PCH, m, n, np, a, b, c); v = cadd(v, g[0].eval(a, b, c, ch == PCH ? p : -1)); }
        if (fl & FL_GT) { convert(ch, TCH, m, n, np, a, b, c); v = cadd(v, g[1].eval(a, b, c, ch == TCH ? p : -1)); }
        if (fl & FL_GA) { convert(ch, ACH, m, n, np, a, b, c); v = cadd(v, g[2].eval(a, b, c, ch == ACH ? p : -1)); }
    } else if (fl & (FL_GP << ch)) {
        v = cadd(v, chan_eval(g, ch, m, n, np, p));
    }
    return v;
}

// reference vertex F0: a RefVertex (kind 0) or a local Vertex around the RefVertex (kind 1)
struct F0V {
    int kind;
    RefV ref;
    ChanV g[3];
    __device__ __forceinline__ double2 eval_psp(int m, int n, int np, int ch) const {
        double2 v = ref.eval(m, n, np, ch, PSP);
        if (kind == 1) v = cadd(v, gamma_psp(g, m, n, np, 0, ch, FL_GP | FL_GT | FL_GA));
        return v;
    }
    // full evaluator with spin (src/vertex.jl:288-336 for kind 1, src/refvertex.jl for kind 0)
    __device__ __forceinline__ double2 eval(int m, int n, int np, int ch, int sp) const {
        if (kind == 0) return ref.eval(m, n, np, ch, sp);
        if (sp == PSP) return eval_psp(m, n, np, ch);
        double2 x;
        if (ch == PCH) x = cneg(eval_psp(m, n, bmf(m, np), PCH));
        else if (ch == TCH) x = cneg(eval_psp(m, n, np, ACH));
        else x = cneg(eval_psp(m, n, np, TCH));
        if (sp == XSP) return x;
        return cadd(cscale(eval_psp(m, n, np, ch), 2.0), x);
    }
};

// target-type vertex (F, FL, Fbuff): channels g, reference vertex f0
struct VtxV {
    ChanV g[3];
    F0V f0;
    __device__ __forceinline__ double2 eval_psp(int m, int n, int np, int p, int ch, int fl) const {
        double2 v = czero();
        if (fl & FL_F0) v = f0.eval_psp(m, n, np, ch);
        return cadd(v, gamma_psp(g, m, n, np, p, ch, fl));
    }
    __device__ __forceinline__ static int swap_ta(int fl) {
        return (fl & (FL_F0 | FL_GP)) | ((fl & FL_GT) ? FL_GA : 0) | ((fl & FL_GA) ? FL_GT : 0);
    }
    __device__ __forceinline__ double2 eval_xsp(int m, int n, int np, int p, int ch, int fl) const {
        const int f2 = swap_ta(fl);
        if (ch == PCH) return cneg(eval_psp(m, n, bmf(m, np), p, PCH, f2));
        if (ch == TCH) return cneg(eval_psp(m, n, np, p, ACH, f2));
        return cneg(eval_psp(m, n, np, p, TCH, f2));
    }
    __device__ __forceinline__ double2 eval(int m, int n, int np, int p, int ch, int sp, int fl) const {
        if (sp == PSP) return eval_psp(m, n, np, p, ch, fl);
        if (sp == XSP) return eval_xsp(m, n, np, p, ch, fl);
        double2 v = cscale(eval_psp(m, n, np, p, ch, fl), 2.0);
        return cadd(v, eval_xsp(m, n, np, p, ch, fl));
    }
};

}  // namespace fdga
