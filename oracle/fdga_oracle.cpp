// fdga_oracle.cpp -- CPU FP64 restatement of the fdDGAsolver.jl parquet iteration.
//
// TEST INFRASTRUCTURE ONLY (the parity checker and the CPU baseline).  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library; the HIP
// product in fddgasolver.jl_amd/ never does.
//
// The code follows the reference loop by loop.  Each function cites the reference
// file:line it restates.  Two deliberate differences, both value-preserving:
//  * s-wave momentum averages (src/nonlocal/swave.jl:32-88) are precomputed once per
//    vertex update instead of being re-summed on every lookup;
//  * caches that the reference stores per P but that do not depend on P are still
//    stored per P here (literal), the HIP product stores them once.
// Parity status of every function is listed in the header comment of tests/test_oracle_*.py
// and in the repository design notes.  Parity unpinned: in the symmetry classes (build_group), the
// operation of a member that the exploration reaches along paths with different operations is the
// first path's (MatsubaraFunctions' SymmetryGroup is not vendored); partitions, representatives and
// all path-independent operations are pinned.
#include "fdga_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

#include <omp.h>

namespace {

using cplx = std::complex<double>;
constexpr int INF = ORC_INF;
constexpr int PCH = ORC_PCH, TCH = ORC_TCH, ACH = ORC_ACH;
constexpr int PSP = ORC_PSP, XSP = ORC_XSP, DSP = ORC_DSP;
constexpr int FL_F0 = 1, FL_GP = 2, FL_GT = 4, FL_GA = 8, FL_ALL = 15;

// ---------------------------------------------------------------------------------------
// Matsubara meshes (MatsubaraFunctions.MatsubaraMesh; sizes pinned by test/test_channel.jl:16)
struct MB {  // bosonic, m = -(N-1)..N-1
    int N = 1;
    int size() const { return 2 * N - 1; }
    bool in(int m) const { return m != INF && m >= -(N - 1) && m <= N - 1; }
    int idx(int m) const { return m + N - 1; }
    int val(int i) const { return i - N + 1; }
};
struct MF {  // fermionic, n = -N..N-1, value πT(2n+1)
    int N = 1;
    int size() const { return 2 * N; }
    bool in(int n) const { return n != INF && n >= -N && n <= N - 1; }
    int idx(int n) const { return n + N; }
    int val(int i) const { return i - N; }
};

// MatsubaraFrequency arithmetic; νInf absorbs (src/types.jl:134-144)
inline int bpf(int m, int n) { return (m == INF || n == INF) ? INF : m + n; }      // Ω + ν
inline int bmf(int m, int n) { return (m == INF || n == INF) ? INF : m - n - 1; }  // Ω - ν
inline int fmf(int a, int b) { return (a == INF || b == INF) ? INF : a - b; }      // ν - ν'
inline int fpf(int a, int b) { return (a == INF || b == INF) ? INF : a + b + 1; }  // ν + ν'

// _convert_channel (src/convention.jl:4-37), all arguments finite
inline void convert(int from, int to, int A, int b, int bp, int& o0, int& o1, int& o2) {
    if (from == to) { o0 = A; o1 = b; o2 = bp; return; }
    if (from == PCH && to == TCH) { o0 = fmf(bmf(A, b), bp); o1 = bp; o2 = b; return; }
    if (from == PCH && to == ACH) { o0 = fmf(b, bp); o1 = bmf(A, b); o2 = bp; return; }
    if (from == TCH && to == PCH) { o0 = fpf(bpf(A, b), bp); o1 = bp; o2 = b; return; }
    if (from == TCH && to == ACH) { o0 = fmf(bp, b); o1 = bpf(A, b); o2 = b; return; }
    if (from == ACH && to == PCH) { o0 = fpf(bpf(A, bp), b); o1 = bpf(A, bp); o2 = bp; return; }
    /* ACH -> TCH */              { o0 = fmf(b, bp); o1 = bp; o2 = bpf(A, bp); return; }
}

// _crossing (src/BSE_templates.jl:4-6)
inline int crossing(int m, int w, int ch) { return ch == PCH ? bmf(m, w) : w; }

// ---------------------------------------------------------------------------------------
// Channel / NL_Channel (src/channel.jl:7-48, 220-339; src/nonlocal/channel.jl:3-226)
struct Chan {
    MB b1, b2, b3;
    MF f2, f3;
    int NP = 1;
    std::vector<cplx> K1, K2, K3;  // [b1,NP] [b2,f2,NP] [b3,f3,f3,NP]
    std::vector<cplx> A1, A2, A3;  // P averages (kSW lookups)

    void alloc(int nK1, int nK2b, int nK2f, int nK3b, int nK3f, int np) {
        b1.N = nK1; b2.N = nK2b; f2.N = nK2f; b3.N = nK3b; f3.N = nK3f; NP = np;
        K1.assign((size_t)b1.size() * NP, 0.0);
        K2.assign((size_t)b2.size() * f2.size() * NP, 0.0);
        K3.assign((size_t)b3.size() * f3.size() * f3.size() * NP, 0.0);
        refresh();
    }
    size_t n1() const { return (size_t)b1.size(); }
    size_t n2() const { return (size_t)b2.size() * f2.size(); }
    size_t n3() const { return (size_t)b3.size() * f3.size() * f3.size(); }
    size_t i1(int m) const { return b1.idx(m); }
    size_t i2(int m, int n) const { return b2.idx(m) + (size_t)b2.size() * f2.idx(n); }
    size_t i3(int m, int n, int np) const {
        return b3.idx(m) + (size_t)b3.size() * (f3.idx(n) + (size_t)f3.size() * f3.idx(np));
    }
    // p >= 0: momentum index; p < 0: s-wave average (kSW)
    cplx k1(int m, int p) const { return p < 0 ? A1[i1(m)] : K1[i1(m) + n1() * p]; }
    cplx k2(int m, int n, int p) const { return p < 0 ? A2[i2(m, n)] : K2[i2(m, n) + n2() * p]; }
    cplx k3(int m, int n, int np, int p) const {
        return p < 0 ? A3[i3(m, n, np)] : K3[i3(m, n, np) + n3() * p];
    }
    // s-wave average: sum(view(f, ..., :)) / length(mesh_P) (src/nonlocal/swave.jl:32-72)
    static void avg(const std::vector<cplx>& K, size_t n, int NP, std::vector<cplx>& A) {
        A.assign(n, 0.0);
        for (int p = 0; p < NP; ++p)
            for (size_t i = 0; i < n; ++i) A[i] += K[i + n * p];
        for (size_t i = 0; i < n; ++i) A[i] /= (double)NP;
    }
    void refresh() {
        avg(K1, n1(), NP, A1);
        avg(K2, n2(), NP, A2);
        avg(K3, n3(), NP, A3);
    }

    // evaluator γ(Ω, ν, νp[, P]) with class flags
    cplx eval(int m, int n, int np, int p, bool fK1 = true, bool fK2 = true, bool fK3 = true) const {
        cplx v = 0.0;
        if (n != INF && np != INF) {  // src/channel.jl:220-253
            if (b1.in(m)) {
                if (fK1) v += k1(m, p);
                if (b2.in(m)) {
                    bool in1 = f2.in(n), in2 = f2.in(np);
                    if (in1 && in2) {
                        if (fK2) v += k2(m, n, p) + k2(m, np, p);
                        if (fK3 && b3.in(m) && f3.in(n) && f3.in(np)) v += k3(m, n, np, p);
                    } else if (in1) {
                        if (fK2) v += k2(m, n, p);
                    } else if (in2) {
                        if (fK2) v += k2(m, np, p);
                    }
                }
            }
        } else if (n == INF && np != INF) {  // src/channel.jl:256-285
            if (fK1) {
                if (b1.in(m)) {
                    v += k1(m, p);
                    if (fK2 && b2.in(m) && f2.in(np)) v += k2(m, np, p);
                }
            } else if (fK2 && b2.in(m) && f2.in(np)) {
                v += k2(m, np, p);
            }
        } else if (n != INF && np == INF) {  // src/channel.jl:287-316
            if (fK1) {
                if (b1.in(m)) {
                    v += k1(m, p);
                    if (fK2 && b2.in(m) && f2.in(n)) v += k2(m, n, p);
                }
            } else if (fK2 && b2.in(m) && f2.in(n)) {
                v += k2(m, n, p);
            }
        } else {  // src/channel.jl:318-339
            if (fK1 && b1.in(m)) v += k1(m, p);
        }
        return v;
    }
};

// ---------------------------------------------------------------------------------------
// RefVertex (src/refvertex.jl:1-216)
struct RefV {
    cplx U = 0.0;
    MB tb;
    MF tf;
    std::vector<cplx> tab[4];  // Fp_p, Fp_x, Ft_p, Ft_x  [tb, tf, tf]
    enum { FPP = 0, FPX = 1, FTP = 2, FTX = 3 };

    void alloc(int nb, int nf) {
        tb.N = nb; tf.N = nf;
        for (auto& t : tab) t.assign((size_t)tb.size() * tf.size() * tf.size(), 0.0);
    }
    // MeshFunction call syntax: zero outside the mesh
    cplx t(int k, int m, int n, int np) const {
        if (!(tb.in(m) && tf.in(n) && tf.in(np))) return 0.0;
        return tab[k][tb.idx(m) + (size_t)tb.size() * (tf.idx(n) + (size_t)tf.size() * tf.idx(np))];
    }
    cplx bare(int sp) const { return sp == XSP ? -U : U; }  // :212-216
    cplx eval(int m, int n, int np, int ch, int sp) const {
        if (n == INF || np == INF) return bare(sp);  // :176-210
        if (sp == DSP) return 2.0 * eval(m, n, np, ch, PSP) + eval(m, n, np, ch, XSP);  // :163-174
        if (sp == PSP) {  // :90-124
            if (ch == PCH) return t(FPP, m, n, np) + U;
            if (ch == TCH) return t(FTP, m, n, np) + U;
            return -t(FTX, m, np, n) + U;
        }
        // xSp :127-161
        if (ch == PCH) return t(FPX, m, n, np) - U;
        if (ch == TCH) return t(FTX, m, n, np) - U;
        return -t(FTP, m, np, n) - U;
    }
};

// ---------------------------------------------------------------------------------------
// Vertex / NL_Vertex evaluators (src/vertex.jl:209-336; src/nonlocal/vertex.jl:69-379)
struct Vert {
    const RefV* ref = nullptr;  // non-null: this vertex is a RefVertex
    const Vert* F0 = nullptr;
    Chan g[3];  // γp, γt, γa

    cplx bare() const { return ref ? ref->U : F0->bare(); }
    void refresh() { if (ref) return; for (auto& c : g) c.refresh(); }

    // p: momentum index of the same-channel class (cross channels use the s-wave average)
    cplx eval(int m, int n, int np, int p, int ch, int sp, int fl) const {
        if (ref) return ref->eval(m, n, np, ch, sp);
        if (sp == DSP) {  // src/vertex.jl:316-336
            cplx v = 0.0;
            v += eval(m, n, np, p, ch, PSP, fl) * 2.0;
            v += eval(m, n, np, p, ch, XSP, fl);
            return v;
        }
        if (sp == XSP) {  // src/vertex.jl:288-313 : swap γt <-> γa flags
            int g2 = (fl & (FL_F0 | FL_GP)) | ((fl & FL_GT) ? FL_GA : 0) | ((fl & FL_GA) ? FL_GT : 0);
            if (ch == PCH) return -eval(m, n, bmf(m, np), p, PCH, PSP, g2);
            if (ch == TCH) return -eval(m, n, np, p, ACH, PSP, g2);
            return -eval(m, n, np, p, TCH, PSP, g2);
        }
        cplx v = 0.0;
        if (fl & FL_F0) v += F0->eval(m, n, np, 0, ch, PSP, FL_ALL);  // F0 is local
        const int bits[3] = {FL_GP, FL_GT, FL_GA};
        if (n != INF && np != INF) {  // src/vertex.jl:209-241; NL s-wave :216-379
            for (int r = 0; r < 3; ++r) {
                if (!(fl & bits[r])) continue;
                int a, b, c;
                convert(ch, r, m, n, np, a, b, c);
                v += g[r].eval(a, b, c, r == ch ? p : -1);
            }
        } else {  // src/vertex.jl:246-284 : only the same channel
            for (int r = 0; r < 3; ++r)
                if (r == ch && (fl & bits[r])) v += g[r].eval(m, n, np, p);
        }
        return v;
    }
};

// ---------------------------------------------------------------------------------------
// MatsubaraFunctions.SymmetryGroup: classes grown from the lowest unvisited linear index;
// the representative is the first element; members get Operation(sgn, con) of it.
struct SymGroup {
    bool identity = true;
    std::vector<std::vector<std::pair<int64_t, uint8_t>>> classes;  // op bit0 sgn, bit1 con
};
inline cplx apply_op(uint8_t op, cplx x) {
    if (op & 1) x = -x;
    if (op & 2) x = std::conj(x);
    return x;
}
// Symmetry: map a tuple of frequency indices to another one + operation; returns false if
// the image leaves the mesh.
using SymFn = std::function<bool(const int*, int*, uint8_t&)>;

}  // namespace

// =========================================================================================
struct orc_solver {
    orc_params prm;
    double T = 0.1;
    int NP = 1, NkG = 1;
    MF mG;
    MB mPiB;
    MF mPiF;
    std::vector<cplx> Gbare, G0, Sig0, G, Sig;
    std::vector<cplx> Pi0pp, Pi0ph, Pipp, Piph;
    std::vector<cplx> Lpp, Lph, L0pp, L0ph;
    RefV ref, zero_ref;
    Vert refwrap, zerowrap;
    Vert F0v, F, Fbuff, FL;
    std::vector<cplx> cache[10];
    // symmetry groups (index as in ParquetSolver: [1]=K1 [2]=K2 [3]=K3)
    SymGroup SGS, SGpp[4], SGph[4], SGppL[4], SGphL[4], SG0pp2, SG0ph2;

    size_t piN() const { return (size_t)mPiB.size() * mPiF.size(); }
    size_t pidx(int m, int n, int p) const {
        return mPiB.idx(m) + (size_t)mPiB.size() * (mPiF.idx(n) + (size_t)mPiF.size() * p);
    }
    size_t gidx(int n, int k) const { return mG.idx(n) + (size_t)mG.size() * k; }
    double fval(int n) const { return M_PI * T * (2.0 * n + 1.0); }  // value(ν)
    cplx U() const { return ref.U; }
};

namespace {

// ---------------------------------------------------------------------------------------
// symmetry-group helpers
void build_group(SymGroup& sg, int rank, const int* sizes, const std::function<int(int, int)>& valof,
                 const std::function<bool(int, int)>& inb, const std::function<int(int, int)>& idxof,
                 const std::vector<SymFn>& syms) {
    sg.identity = false;
    sg.classes.clear();
    int64_t total = 1;
    for (int d = 0; d < rank; ++d) total *= sizes[d];
    std::vector<char> checked(total, 0);
    auto lin_to_w = [&](int64_t idx, int* w) {
        for (int d = 0; d < rank; ++d) {
            w[d] = valof(d, (int)(idx % sizes[d]));
            idx /= sizes[d];
        }
    };
    auto w_to_lin = [&](const int* w) {
        int64_t idx = 0, stride = 1;
        for (int d = 0; d < rank; ++d) {
            idx += stride * idxof(d, w[d]);
            stride *= sizes[d];
        }
        return idx;
    };
    for (int64_t idx = 0; idx < total; ++idx) {
        if (checked[idx]) continue;
        std::vector<std::pair<int64_t, uint8_t>> cls;
        cls.push_back({idx, 0});
        checked[idx] = 1;
        // depth-first exploration
        std::vector<std::pair<int64_t, uint8_t>> stack = {{idx, 0}};
        while (!stack.empty()) {
            auto [cur, op] = stack.back();
            stack.pop_back();
            int w[4], wp[4];
            lin_to_w(cur, w);
            for (auto& S : syms) {
                uint8_t o = 0;
                if (!S(w, wp, o)) continue;
                bool ok = true;
                for (int d = 0; d < rank; ++d) ok = ok && inb(d, wp[d]);
                if (!ok) continue;
                int64_t j = w_to_lin(wp);
                if (checked[j]) continue;
                checked[j] = 1;
                uint8_t no = (uint8_t)(op ^ o);
                cls.push_back({j, no});
                stack.push_back({j, no});
            }
        }
        sg.classes.push_back(std::move(cls));
    }
}

// SG(f, InitFunction(diagram)): evaluate class representatives, fill members
void sg_eval(const SymGroup& sg, std::vector<cplx>& out, size_t n, const std::function<cplx(size_t)>& f) {
    if (sg.identity) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t i = 0; i < (int64_t)n; ++i) out[i] = f((size_t)i);
        return;
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t c = 0; c < (int64_t)sg.classes.size(); ++c) {
        const auto& cls = sg.classes[c];
        cplx r = f((size_t)cls[0].first);
        for (auto& [j, op] : cls) out[j] = apply_op(op, r);
    }
}
// SG(f): symmetrize in place (first class element is the reference)
void sg_symmetrize(const SymGroup& sg, std::vector<cplx>& f) {
    if (sg.identity) return;
    for (auto& cls : sg.classes) {
        cplx r = f[cls[0].first];
        for (size_t k = 1; k < cls.size(); ++k) f[cls[k].first] = apply_op(cls[k].second, r);
    }
}

// ---------------------------------------------------------------------------------------
// 2D DFT over the trailing (L, L) dims of data[row + nrows*(j1 + L*j2)]; sign -1 = fft, +1 = bfft
void dft2(std::vector<cplx>& data, size_t nrows, int L, int sign) {
    if (L == 1) return;
    std::vector<cplx> tw(L);
    for (int k = 0; k < L; ++k) tw[k] = std::polar(1.0, sign * 2.0 * M_PI * k / L);
    std::vector<cplx> buf(L);
    // along j1
#pragma omp parallel for collapse(2) firstprivate(buf)
    for (int j2 = 0; j2 < L; ++j2)
        for (size_t r = 0; r < nrows; ++r) {
            for (int k = 0; k < L; ++k) {
                cplx s = 0.0;
                for (int j = 0; j < L; ++j) s += data[r + nrows * (j + (size_t)L * j2)] * tw[(size_t)(j * k) % L];
                buf[k] = s;
            }
            for (int k = 0; k < L; ++k) data[r + nrows * (k + (size_t)L * j2)] = buf[k];
        }
    // along j2
#pragma omp parallel for collapse(2) firstprivate(buf)
    for (int j1 = 0; j1 < L; ++j1)
        for (size_t r = 0; r < nrows; ++r) {
            for (int k = 0; k < L; ++k) {
                cplx s = 0.0;
                for (int j = 0; j < L; ++j) s += data[r + nrows * (j1 + (size_t)L * j)] * tw[(size_t)(j * k) % L];
                buf[k] = s;
            }
            for (int k = 0; k < L; ++k) data[r + nrows * (j1 + (size_t)L * k)] = buf[k];
        }
}
inline int imod(int a, int L) { int r = a % L; return r < 0 ? r + L : r; }

// ---------------------------------------------------------------------------------------
// bubbles! (local, src/bubble.jl:9-36) with use_G_tail = true
void bubbles_local(const orc_solver* s, std::vector<cplx>& pp, std::vector<cplx>& ph, const std::vector<cplx>& G) {
    auto gt = [&](int n) -> cplx { return s->mG.in(n) ? G[s->gidx(n, 0)] : cplx(1.0 / s->fval(n), 0.0); };
    for (int iW = 0; iW < s->mPiB.size(); ++iW)
        for (int iv = 0; iv < s->mPiF.size(); ++iv) {
            int m = s->mPiB.val(iW), n = s->mPiF.val(iv);
            cplx Gv = gt(n), Gmv = gt(bmf(m, n)), Gpv = gt(bpf(m, n));
            pp[s->pidx(m, n, 0)] = Gv * Gmv;
            ph[s->pidx(m, n, 0)] = Gv * Gpv;
        }
}

// bubbles_real_space! (src/nonlocal/bubble.jl:87-158), G1 = G2 = G, use_G_tail = true
void bubbles_nl(const orc_solver* s, std::vector<cplx>& pp, std::vector<cplx>& ph, const std::vector<cplx>& G) {
    const int LG = s->prm.LG, L = s->prm.L, NG = s->mG.size();
    std::vector<cplx> GR = G;
    dft2(GR, NG, LG, -1);
    for (auto& x : GR) x /= (double)LG * LG;
    const size_t nrow = s->piN();
    std::fill(pp.begin(), pp.end(), 0.0);
    std::fill(ph.begin(), ph.end(), 0.0);
    auto call = [&](int n, int k) -> cplx { return s->mG.in(n) ? GR[s->gidx(n, k)] : cplx(0.0); };
    auto tail = [&](int n, int k) -> cplx { return s->mG.in(n) ? GR[s->gidx(n, k)] : cplx(1.0 / s->fval(n), 0.0); };
    const int h = L / 2;
    for (int R2 = -h; R2 <= h; ++R2)
        for (int R1 = -h; R1 <= h; ++R1) {
            double w = 1.0;
            if (LG % 2 == 0) {
                if (std::abs(R1) == LG / 2) w /= 2;
                if (std::abs(R2) == LG / 2) w /= 2;
            }
            int kp = imod(R1, LG) + LG * imod(R2, LG);
            int km = imod(-R1, LG) + LG * imod(-R2, LG);
            int iR = imod(R1, L) + L * imod(R2, L);
            bool origin = (R1 == 0 && R2 == 0);
#pragma omp parallel for
            for (int iv = 0; iv < s->mPiF.size(); ++iv)
                for (int iW = 0; iW < s->mPiB.size(); ++iW) {
                    int m = s->mPiB.val(iW), n = s->mPiF.val(iv);
                    size_t o = iW + (size_t)s->mPiB.size() * iv + nrow * iR;
                    if (origin) {
                        pp[o] += tail(bmf(m, n), kp) * tail(n, kp) * w;
                        ph[o] += tail(bpf(m, n), kp) * tail(n, km) * w;
                    } else {
                        pp[o] += call(bmf(m, n), kp) * call(n, kp) * w;
                        ph[o] += call(bpf(m, n), kp) * call(n, km) * w;
                    }
                }
        }
    dft2(pp, nrow, L, +1);
    dft2(ph, nrow, L, +1);
}

void bubbles_any(const orc_solver* s, std::vector<cplx>& pp, std::vector<cplx>& ph, const std::vector<cplx>& G) {
    if (s->prm.nonlocal) bubbles_nl(s, pp, ph, G);
    else bubbles_local(s, pp, ph, G);
}

// Dyson! (src/dyson.jl:5-31): G = 1 / (1/Gbare + Σ)
void dyson(std::vector<cplx>& G, const std::vector<cplx>& S, const std::vector<cplx>& Gb) {
    for (size_t i = 0; i < G.size(); ++i) G[i] = 1.0 / (1.0 / Gb[i] + S[i]);
}

// compute_occupation (src/dyson.jl:33-41)
double occupation(const orc_solver* s, const std::vector<cplx>& G) {
    cplx sum = 0.0;
    for (auto& x : G) sum += x;
    return 0.5 + sum.imag() * s->T / (double)s->NkG;
}

}  // namespace

// =========================================================================================
extern "C" {

orc_solver* orc_create(const orc_params* p) {
    auto* s = new orc_solver();
    s->prm = *p;
    s->T = p->T;
    s->NP = p->nonlocal ? p->L * p->L : 1;
    s->NkG = p->nonlocal ? p->LG * p->LG : 1;
    if (!p->nonlocal) { s->prm.L = 1; s->prm.LG = 1; }
    s->mG.N = p->nG;
    s->mPiB.N = p->nK1;
    s->mPiF.N = p->nK1 * p->mPinu;
    size_t ng = (size_t)s->mG.size() * s->NkG;
    for (auto* v : {&s->Gbare, &s->G0, &s->Sig0, &s->G, &s->Sig}) v->assign(ng, 0.0);
    size_t npi = s->piN() * s->NP;
    for (auto* v : {&s->Pi0pp, &s->Pi0ph, &s->Pipp, &s->Piph}) v->assign(npi, 0.0);
    size_t nl = (size_t)(2 * p->nK2b - 1) * (2 * p->nK2f) * s->NP;
    for (auto* v : {&s->Lpp, &s->Lph, &s->L0pp, &s->L0ph}) v->assign(nl, 0.0);
    // RefVertex(T, U) with tables (src/refvertex.jl:19-34)
    s->ref.U = p->U;
    s->ref.alloc(p->tab_nb, p->tab_nf);
    s->zero_ref.U = 0.0;
    s->zero_ref.alloc(1, 1);
    s->refwrap.ref = &s->ref;
    s->zerowrap.ref = &s->zero_ref;
    if (p->ref_kind == 0) {
        s->F0v.ref = &s->ref;
    } else {
        s->F0v.F0 = &s->refwrap;
        for (auto& c : s->F0v.g) c.alloc(p->r_nK1, p->r_nK2b, p->r_nK2f, p->r_nK3b, p->r_nK3f, 1);
    }
    s->F.F0 = &s->F0v;
    s->Fbuff.F0 = &s->zerowrap;  // Fbuff, FL: Vertex(RefVertex(T, 0.0), ...) (src/ParquetSolver.jl:118)
    s->FL.F0 = &s->zerowrap;
    for (Vert* v : {&s->F, &s->Fbuff, &s->FL})
        for (auto& c : v->g) c.alloc(p->nK1, p->nK2b, p->nK2f, p->nK3b, p->nK3f, s->NP);
    size_t nc = s->F.g[0].n3() * s->NP;
    for (auto& c : s->cache) c.assign(nc, 0.0);
    return s;
}

void orc_destroy(orc_solver* s) { delete s; }

static std::vector<cplx>* arr(orc_solver* s, int which) {
    switch (which) {
        case ORC_GBARE: return &s->Gbare;
        case ORC_G0: return &s->G0;
        case ORC_SIGMA0: return &s->Sig0;
        case ORC_G: return &s->G;
        case ORC_SIGMA: return &s->Sig;
        case ORC_PI0PP: return &s->Pi0pp;
        case ORC_PI0PH: return &s->Pi0ph;
        case ORC_PIPP: return &s->Pipp;
        case ORC_PIPH: return &s->Piph;
        case ORC_LPP: return &s->Lpp;
        case ORC_LPH: return &s->Lph;
        case ORC_TAB_FPP: case ORC_TAB_FPX: case ORC_TAB_FTP: case ORC_TAB_FTX:
            return &s->ref.tab[which - ORC_TAB_FPP];
        default: break;
    }
    if (which >= ORC_CACHE0 && which < ORC_CACHE0 + 10) return &s->cache[which - ORC_CACHE0];
    if (which >= ORC_VERTEX0 && which < ORC_VERTEX0 + 64) {
        int k = which - ORC_VERTEX0, vtx = k / 16, ch = (k % 16) / 3, cls = (k % 16) % 3 + 1;
        if (ch > 2) return nullptr;
        Vert* v = vtx == 0 ? &s->F : vtx == 1 ? &s->Fbuff : vtx == 2 ? &s->FL : &s->F0v;
        if (v->ref) return nullptr;
        Chan& c = v->g[ch];
        return cls == 1 ? &c.K1 : cls == 2 ? &c.K2 : &c.K3;
    }
    return nullptr;
}

int64_t orc_size(const orc_solver* s, int32_t which) {
    auto* a = arr(const_cast<orc_solver*>(s), which);
    return a ? (int64_t)a->size() : -1;
}
int orc_set(orc_solver* s, int32_t which, const double* data) {
    auto* a = arr(s, which);
    if (!a) return -1;
    std::memcpy(static_cast<void*>(a->data()), data, a->size() * sizeof(cplx));
    for (Vert* v : {&s->F, &s->Fbuff, &s->FL, &s->F0v}) if (!v->ref) v->refresh();
    return 0;
}
int orc_get(const orc_solver* s, int32_t which, double* data) {
    auto* a = arr(const_cast<orc_solver*>(s), which);
    if (!a) return -1;
    std::memcpy(data, a->data(), a->size() * sizeof(cplx));
    return 0;
}
void orc_set_U(orc_solver* s, double re, double im) { s->ref.U = cplx(re, im); }

// siam_bare_Green (src/models/siam.jl:9-29); stores i*G0
void orc_siam_bare_green(orc_solver* s, double e, double Delta, double D) {
    for (int i = 0; i < s->mG.size(); ++i) {
        double v = s->fval(s->mG.val(i));
        double sg = (v > 0) - (v < 0);
        if (D <= 0) s->Gbare[i] = 1.0 / (cplx(v, e) + Delta * sg);
        else s->Gbare[i] = 1.0 / (cplx(v, e) + 2.0 * Delta / M_PI * std::atan(D / v));
    }
}

// hubbard_bare_Green (src/models/hubbard.jl:8-44); k = 2π (j1, j2) / LG, j1 fastest
void orc_hubbard_bare_green(orc_solver* s, double mu, double t1, double t2, double t3) {
    int LG = s->prm.LG;
    for (int j2 = 0; j2 < LG; ++j2)
        for (int j1 = 0; j1 < LG; ++j1) {
            double k1 = 2 * M_PI * j1 / LG, k2 = 2 * M_PI * j2 / LG;
            double ek = -2 * t1 * (std::cos(k1) + std::cos(k2)) - 4 * t2 * std::cos(k1) * std::cos(k2)
                        - 2 * t3 * (std::cos(2 * k1) + std::cos(2 * k2));
            int k = j1 + LG * j2;
            for (int i = 0; i < s->mG.size(); ++i) {
                double v = s->fval(s->mG.val(i));
                s->Gbare[s->gidx(s->mG.val(i), k)] = 1.0 / cplx(mu - ek, v) * cplx(0.0, 1.0);
            }
        }
}

void orc_init(orc_solver* s) {
    bubbles_any(s, s->Pi0pp, s->Pi0ph, s->G0);  // src/ParquetSolver.jl:103
    s->Sig = s->Sig0;
    s->G = s->G0;
    dyson(s->G, s->Sig, s->Gbare);               // :106-109
    bubbles_any(s, s->Pipp, s->Piph, s->G);      // :112-114
    for (Vert* v : {&s->F, &s->Fbuff, &s->FL, &s->F0v}) if (!v->ref) v->refresh();
}

// ---------------------------------------------------------------------------------------
// init_sym_grp! (src/ParquetSolver.jl:203-231), symmetries of src/symmetries.jl:1-94; lattice
// solver: init_sym_grp! of src/nonlocal/ParquetSolver.jl:198-300 with src/nonlocal/symmetries.jl:1-133
int orc_init_sym_grp(orc_solver* s) {
    const orc_params& p = s->prm;
    const bool nl = p.nonlocal != 0;
    // mesh axes: fermionic / bosonic Matsubara meshes, or an L x L momentum mesh whose points
    // BrillouinPoint(j1, j2) are numbered j1 + L j2 (the layout of G[ν, k] and of the P axis)
    enum { FERM = 0, BOS = 1, MOM = 2 };
    struct Dim { int kind; int N; };
    auto build = [&](SymGroup& sg, const std::vector<Dim>& dims, const std::vector<SymFn>& syms) {
        int sizes[5];
        for (size_t d = 0; d < dims.size(); ++d)
            sizes[d] = dims[d].kind == BOS ? 2 * dims[d].N - 1 : dims[d].kind == FERM ? 2 * dims[d].N : dims[d].N * dims[d].N;
        auto valof = [dims](int d, int i) {
            return dims[d].kind == BOS ? i - dims[d].N + 1 : dims[d].kind == FERM ? i - dims[d].N : i;
        };
        auto inb = [dims](int d, int x) {
            if (x == INF) return false;
            if (dims[d].kind == BOS) return x >= -(dims[d].N - 1) && x <= dims[d].N - 1;
            if (dims[d].kind == FERM) return x >= -dims[d].N && x <= dims[d].N - 1;
            return x >= 0 && x < dims[d].N * dims[d].N;
        };
        auto idxof = [dims](int d, int x) {
            return dims[d].kind == BOS ? x + dims[d].N - 1 : dims[d].kind == FERM ? x + dims[d].N : x;
        };
        build_group(sg, (int)dims.size(), sizes, valof, inb, idxof, syms);
    };
    auto nf = [](int n) { return -n - 1; };  // -ν
    const uint8_t SGN = 1, CON = 2;
    SymFn sSig = [&](const int* w, int* o, uint8_t& op) { o[0] = nf(w[0]); op = SGN | CON; return true; };
    SymFn sK1 = [&](const int* w, int* o, uint8_t& op) { o[0] = -w[0]; op = CON; return true; };
    SymFn sK2pp1 = [&](const int* w, int* o, uint8_t& op) { o[0] = -w[0]; o[1] = nf(w[1]); op = CON; return true; };
    SymFn sK2pp2 = [&](const int* w, int* o, uint8_t& op) { o[0] = w[0]; o[1] = bmf(w[0], w[1]); op = 0; return true; };
    SymFn sK2ph2 = [&](const int* w, int* o, uint8_t& op) { o[0] = -w[0]; o[1] = bpf(w[0], w[1]); op = 0; return true; };
    SymFn sK3pp1 = [&](const int* w, int* o, uint8_t& op) { o[0] = -w[0]; o[1] = nf(w[1]); o[2] = nf(w[2]); op = CON; return true; };
    SymFn sK3pp2 = [&](const int* w, int* o, uint8_t& op) { o[0] = w[0]; o[1] = w[2]; o[2] = w[1]; op = 0; return true; };
    SymFn sK3pp3 = [&](const int* w, int* o, uint8_t& op) { o[0] = w[0]; o[1] = bmf(w[0], w[1]); o[2] = bmf(w[0], w[2]); op = 0; return true; };
    SymFn sK3ph3 = [&](const int* w, int* o, uint8_t& op) { o[0] = -w[0]; o[1] = bpf(w[0], w[1]); o[2] = bpf(w[0], w[2]); op = 0; return true; };

    // momentum parts (src/nonlocal/symmetries.jl): -k folded back, _ref (x, y) -> (y, x),
    // _rot (x, y) -> fold_back((y, -x))
    enum { KID = 0, KNEG = 1, KREF = 2, KROT = 3 };
    auto kmove = [](int code, int k, int L) {
        const int x = k % L, y = k / L;
        const int mx = (L - x) % L, my = (L - y) % L;
        if (code == KNEG) return mx + L * my;
        if (code == KREF) return y + L * x;
        if (code == KROT) return y + L * mx;
        return k;
    };
    // one group: the frequency symmetries (each with the momentum part of its lattice form), and for
    // the lattice solver the momentum axis plus sX_ref and sX_rot at the end of the list
    auto group = [&](SymGroup& sg, std::vector<Dim> dims, std::vector<std::pair<SymFn, int>> syms, int L) {
        const int nfreq = (int)dims.size();
        std::vector<SymFn> fns;
        if (!nl) {
            for (auto& f : syms) fns.push_back(f.first);
            build(sg, dims, fns);
            return;
        }
        dims.push_back({MOM, L});
        SymFn keep = [nfreq](const int* w, int* o, uint8_t& op) {
            for (int d = 0; d < nfreq; ++d) o[d] = w[d];
            op = 0;
            return true;
        };
        syms.push_back({keep, KREF});
        syms.push_back({keep, KROT});
        for (auto& f : syms) {
            SymFn fr = f.first;
            const int code = f.second;
            fns.push_back([=](const int* w, int* o, uint8_t& op) {
                if (!fr(w, o, op)) return false;
                o[nfreq] = kmove(code, w[nfreq], L);
                return true;
            });
        }
        build(sg, dims, fns);
    };

    const int L = p.L, LG = p.LG;
    std::vector<Dim> dS = {{FERM, p.nG}};
    std::vector<Dim> d1 = {{BOS, p.nK1}};
    std::vector<Dim> d2 = {{BOS, p.nK2b}, {FERM, p.nK2f}};
    std::vector<Dim> d3 = {{BOS, p.nK3b}, {FERM, p.nK3f}, {FERM, p.nK3f}};
    group(s->SGS, dS, {{sSig, KNEG}}, LG);
    group(s->SGpp[1], d1, {{sK1, KNEG}}, L);
    group(s->SGpp[2], d2, {{sK2pp1, KNEG}, {sK2pp2, KID}}, L);
    group(s->SGpp[3], d3, {{sK3pp1, KNEG}, {sK3pp2, KID}, {sK3pp3, KID}}, L);
    s->SGppL[1] = s->SGpp[1];
    s->SGppL[2] = s->SGpp[2];
    group(s->SGppL[3], d3, {{sK3pp1, KNEG}, {sK3pp3, KID}}, L);
    s->SGph[1] = s->SGpp[1];
    group(s->SGph[2], d2, {{sK2pp1, KNEG}, {sK2ph2, KNEG}}, L);
    group(s->SGph[3], d3, {{sK3pp1, KNEG}, {sK3pp2, KID}, {sK3ph3, KNEG}}, L);
    s->SGphL[1] = s->SGph[1];
    s->SGphL[2] = s->SGph[2];
    group(s->SGphL[3], d3, {{sK3pp1, KNEG}, {sK3ph3, KNEG}}, L);
    group(s->SG0pp2, d2, {{sK2pp1, KNEG}, {sK2pp2, KID}}, L);
    group(s->SG0ph2, d2, {{sK2pp1, KNEG}, {sK2ph2, KNEG}}, L);
    return 0;
}

static const SymGroup* sym_by_id(const orc_solver* s, int32_t id) {
    switch (id) {
        case 0: return &s->SGS;
        case 1: return &s->SGpp[1];
        case 2: return &s->SGpp[2];
        case 3: return &s->SGpp[3];
        case 4: return &s->SGppL[3];
        case 5: return &s->SGph[2];
        case 6: return &s->SGph[3];
        case 7: return &s->SGphL[3];
        case 8: return &s->SG0pp2;
        case 9: return &s->SG0ph2;
        default: return nullptr;
    }
}

int64_t orc_sym_size(const orc_solver* s, int32_t id) {
    const SymGroup* g = sym_by_id(s, id);
    if (!g || g->identity) return -1;
    int64_t n = 0;
    for (auto& cls : g->classes) n += (int64_t)cls.size();
    return n;
}

int orc_sym_classes(const orc_solver* s, int32_t id, int64_t* rep, int32_t* op) {
    const SymGroup* g = sym_by_id(s, id);
    if (!g || g->identity) return -1;
    for (auto& cls : g->classes)
        for (auto& [j, o] : cls) {
            rep[j] = cls[0].first;
            op[j] = o;
        }
    return 0;
}

// ---------------------------------------------------------------------------------------
void orc_dyson(orc_solver* s) { dyson(s->G, s->Sig, s->Gbare); }
void orc_bubbles(orc_solver* s) { bubbles_any(s, s->Pipp, s->Piph, s->G); }

// build_K3_cache! (src/build_K3_cache.jl:20-95; NL src/nonlocal/build_K3_cache.jl:18-94)
void orc_build_k3_cache(orc_solver* s) {
    s->F.refresh();
    const Chan& c = s->F.g[0];
    const int nb = c.b3.size(), nf = c.f3.size();
    const Vert& F = s->F;
    const Vert& F0 = s->F0v;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int P = 0; P < s->NP; ++P)
        for (int i3 = 0; i3 < nf; ++i3)
            for (int i2 = 0; i2 < nf; ++i2)
                for (int i1 = 0; i1 < nb; ++i1) {
                    size_t i = i1 + (size_t)nb * (i2 + (size_t)nf * (i3 + (size_t)nf * P));
                    int m = c.b3.val(i1), a = c.f3.val(i2), b = c.f3.val(i3);
                    // loop 1: (Ω, ω, νp)
                    {
                        int w = a, vp = b;
                        s->cache[0][i] = F.eval(m, w, vp, P, PCH, XSP, FL_GT | FL_GA);
                        s->cache[1][i] = F0.eval(m, w, vp, 0, PCH, XSP, FL_ALL) - F0.eval(m, w, INF, 0, PCH, XSP, FL_ALL);
                        s->cache[2][i] = F0.eval(m, w, vp, 0, ACH, PSP, FL_ALL) - F0.eval(m, w, INF, 0, ACH, PSP, FL_ALL);
                        s->cache[3][i] = F0.eval(m, w, vp, 0, TCH, PSP, FL_ALL) - F0.eval(m, w, INF, 0, TCH, PSP, FL_ALL);
                        s->cache[3][i] = 2.0 * s->cache[3][i] - s->cache[2][i];
                    }
                    // loop 2: (Ω, ν, ω)
                    {
                        int v = a, w = b;
                        cplx Gpp = F.eval(m, v, w, P, PCH, PSP, FL_GT | FL_GA);
                        cplx Ga = F.eval(m, v, w, P, ACH, PSP, FL_GP | FL_GT);
                        cplx Gt = F.eval(m, v, w, P, TCH, PSP, FL_GP | FL_GA);
                        cplx Fp = F.eval(m, v, w, P, PCH, PSP, FL_F0 | FL_GP) - F.eval(m, INF, w, P, PCH, PSP, FL_F0 | FL_GP) + Gpp;
                        cplx Fa = F.eval(m, v, w, P, ACH, PSP, FL_F0 | FL_GA) - F.eval(m, INF, w, P, ACH, PSP, FL_F0 | FL_GA) + Ga;
                        cplx Ft = F.eval(m, v, w, P, TCH, PSP, FL_F0 | FL_GT) - F.eval(m, INF, w, P, TCH, PSP, FL_F0 | FL_GT) + Gt;
                        s->cache[4][i] = Gpp;
                        s->cache[5][i] = Ga;
                        s->cache[6][i] = Gt * 2.0 - Ga;
                        s->cache[7][i] = Fp;
                        s->cache[8][i] = Fa;
                        s->cache[9][i] = Ft * 2.0 - Fa;
                    }
                }
}

// build_K3_cache_mfRG!(S, is_first_iteration) (src/nonlocal_2/build_K3_cache.jl:97-164, the method that
// mfRGLinearMap calls for NL_ParquetSolver): the Γ caches are F(...; γr = false) - F.F0(...; γr = false),
// the F caches (first call only) the reference vertex's F0(ν) - F0(νInf); the F0 caches are left as they
// are.  Each cache is evaluated over the K3 symmetry group of its channel (SGpp[3] / SGph[3]).
void orc_build_k3_cache_mfrg(orc_solver* s, int32_t is_first) {
    s->F.refresh();
    const Chan& c = s->F.g[0];
    const int nb = c.b3.size(), nf = c.f3.size();
    const size_t n3 = c.n3();
    const Vert& F = s->F;
    const Vert& F0 = s->F0v;
    auto idx = [&](size_t ii, int& m, int& a, int& b, int& P) {
        const size_t i = ii % n3;
        P = (int)(ii / n3);
        m = c.b3.val((int)(i % nb));
        a = c.f3.val((int)((i / nb) % nf));
        b = c.f3.val((int)(i / ((size_t)nb * nf)));
    };
    const size_t ntot = n3 * s->NP;
    // γr = false in channel r: F0 and the two other channels
    auto gam = [&](int m, int a, int b, int P, int ch, int sp) {
        const int fl = FL_ALL & ~(ch == PCH ? FL_GP : ch == TCH ? FL_GT : FL_GA);
        return F.eval(m, a, b, P, ch, sp, fl) - F0.eval(m, a, b, 0, ch, sp, fl);
    };
    sg_eval(s->SGpp[3], s->cache[0], ntot, [&](size_t ii) {
        int m, a, b, P; idx(ii, m, a, b, P);
        return gam(m, a, b, P, PCH, XSP);
    });
    sg_eval(s->SGpp[3], s->cache[4], ntot, [&](size_t ii) {
        int m, a, b, P; idx(ii, m, a, b, P);
        return gam(m, a, b, P, PCH, PSP);
    });
    sg_eval(s->SGph[3], s->cache[5], ntot, [&](size_t ii) {
        int m, a, b, P; idx(ii, m, a, b, P);
        return gam(m, a, b, P, ACH, PSP);
    });
    sg_eval(s->SGph[3], s->cache[6], ntot, [&](size_t ii) {
        int m, a, b, P; idx(ii, m, a, b, P);
        return gam(m, a, b, P, TCH, PSP);
    });
    for (size_t i = 0; i < ntot; ++i) s->cache[6][i] = s->cache[6][i] * 2.0 - s->cache[5][i];
    if (!is_first) return;
    auto d0 = [&](int m, int a, int b, int ch) {
        return F0.eval(m, a, b, 0, ch, PSP, FL_ALL) - F0.eval(m, INF, b, 0, ch, PSP, FL_ALL);
    };
    sg_eval(s->SGpp[3], s->cache[7], ntot, [&](size_t ii) {
        int m, a, b, P; idx(ii, m, a, b, P);
        return d0(m, a, b, PCH);
    });
    sg_eval(s->SGph[3], s->cache[8], ntot, [&](size_t ii) {
        int m, a, b, P; idx(ii, m, a, b, P);
        return d0(m, a, b, ACH);
    });
    sg_eval(s->SGph[3], s->cache[9], ntot, [&](size_t ii) {
        int m, a, b, P; idx(ii, m, a, b, P);
        return d0(m, a, b, TCH);
    });
    for (size_t i = 0; i < ntot; ++i) s->cache[9][i] = s->cache[9][i] * 2.0 - s->cache[8][i];
}

static const int kSign[3] = {+1, -1, +1};        // BSE_K1/L_K2/K2 sign: p +1, t -1, a +1
static const int kSp[3] = {PSP, DSP, PSP};        // spin: p pSp, t dSp, a pSp

static SymGroup& sg_for(orc_solver* s, int ch, int cls, bool left) {
    if (ch == PCH) return left ? s->SGppL[cls] : s->SGpp[cls];
    return left ? s->SGphL[cls] : s->SGph[cls];
}

// t channel post-processing: γt = (γt_d + γa) / 2 (src/BSE_templates.jl:29-33 etc.)
static void t_post(std::vector<cplx>& t, const std::vector<cplx>& a) {
    for (size_t i = 0; i < t.size(); ++i) { t[i] += a[i]; t[i] /= 2.0; }
}

// BSE_L_K2! (local src/BSEa/BSEa_K2.jl:1-40; NL src/nonlocal/BSEa/BSEa_K2.jl:1-41)
void orc_bse_l_k2(orc_solver* s, int32_t ch) {
    s->F.refresh();
    Chan& out = s->FL.g[ch];
    const std::vector<cplx>& Pi0 = (ch == PCH) ? s->Pi0pp : s->Pi0ph;
    const int sp = kSp[ch], sg = kSign[ch];
    const int fl = FL_ALL & ~FL_F0 & ~(ch == PCH ? FL_GP : ch == TCH ? FL_GT : FL_GA);
    const size_t n2 = out.n2();
    // one evaluation over every (Ω, ν, P): the lattice groups relate different momenta
    {
        std::vector<cplx> tmp(n2 * s->NP);
        sg_eval(sg_for(s, ch, 2, true), tmp, tmp.size(), [&](size_t ii) {
            const size_t i = ii % n2;
            const int P = (int)(ii / n2);
            int m = out.b2.val((int)(i % out.b2.size())), v = out.f2.val((int)(i / out.b2.size()));
            cplx val = 0.0;
            if (!s->prm.nonlocal) {
                for (int iw = 0; iw < s->mPiF.size(); ++iw) {
                    int w = s->mPiF.val(iw);
                    cplx Gp = s->F.eval(m, v, w, P, ch, sp, fl);
                    cplx F0p = s->F0v.eval(m, crossing(m, w, ch), INF, 0, ch, sp, FL_ALL);
                    val += Gp * Pi0[s->pidx(m, w, P)] * F0p;
                }
            } else {
                for (int iw = 0; iw < out.f2.size(); ++iw) {
                    int w = out.f2.val(iw);
                    cplx Gl = s->F.eval(m, v, crossing(m, w, ch), P, ch, sp, fl);
                    cplx F0r = s->F0v.eval(m, w, INF, 0, ch, sp, FL_ALL);
                    val += Gl * Pi0[s->pidx(m, w, P)] * F0r;
                }
            }
            return s->T * val * (double)sg;
        });
        std::copy(tmp.begin(), tmp.end(), out.K2.begin());
    }
    if (ch == TCH) t_post(s->FL.g[TCH].K2, s->FL.g[ACH].K2);
    s->FL.g[ch].refresh();
}

// BSE_L_K3! (src/BSEa/BSEa_K3.jl:1-41; NL src/nonlocal/BSEa/BSEa_K3.jl:1-41)
void orc_bse_l_k3(orc_solver* s, int32_t ch) {
    Chan& out = s->FL.g[ch];
    const std::vector<cplx>& Pi0 = (ch == PCH) ? s->Pi0pp : s->Pi0ph;
    const std::vector<cplx>& CG = s->cache[ch == PCH ? 4 : ch == ACH ? 5 : 6];
    const std::vector<cplx>& CF0 = s->cache[ch == PCH ? 1 : ch == ACH ? 2 : 3];
    const int sg = (ch == ACH) ? +1 : -1;  // src/BSE_templates.jl:122,130,138
    const int nb = out.b3.size(), nf = out.f3.size();
    const size_t n3 = out.n3();
    {
        std::vector<cplx> tmp(n3 * s->NP);
        sg_eval(sg_for(s, ch, 3, true), tmp, tmp.size(), [&](size_t ii) {
            const size_t i = ii % n3;
            const int P = (int)(ii / n3);
            int i1 = (int)(i % nb), i2 = (int)((i / nb) % nf), i3 = (int)(i / ((size_t)nb * nf));
            int m = out.b3.val(i1);
            cplx val = 0.0;
            for (int iw = 0; iw < nf; ++iw) {
                int w = out.f3.val(iw);
                cplx G = CG[i1 + (size_t)nb * (i2 + (size_t)nf * (iw + (size_t)nf * P))];
                cplx F0 = CF0[i1 + (size_t)nb * (iw + (size_t)nf * (i3 + (size_t)nf * P))];
                val += G * Pi0[s->pidx(m, w, P)] * F0;
            }
            return s->T * val * (double)sg;
        });
        std::copy(tmp.begin(), tmp.end(), out.K3.begin());
    }
    if (ch == TCH) t_post(s->FL.g[TCH].K3, s->FL.g[ACH].K3);
    s->FL.g[ch].refresh();
}

// BSE_K1! (src/BSEa/BSEa_K1.jl:2-54; NL src/nonlocal/BSEa/BSEa_K1.jl:2-57).  is_mfRG = true: the central
// part F0(Ω, νInf, ω) Π0 FL(Ω, ω~, νInf); the local form evaluates both vertices in aCh / pSp whatever the
// channel (src/BSEa/BSEa_K1.jl:29-35), the lattice form in the channel's own (src/nonlocal/BSEa/BSEa_K1.jl:29-37)
void orc_bse_k1(orc_solver* s, int32_t ch, int32_t is_mfrg) {
    s->F.refresh();
    s->FL.refresh();
    Chan& out = s->Fbuff.g[ch];
    const std::vector<cplx>& Pi0 = (ch == PCH) ? s->Pi0pp : s->Pi0ph;
    const std::vector<cplx>& Pi = (ch == PCH) ? s->Pipp : s->Piph;
    const int sp = kSp[ch], sg = kSign[ch];
    const bool nl = s->prm.nonlocal != 0;
    const int mch = nl ? ch : ACH, msp = nl ? sp : PSP;   // vertices' channel / spin of the mfRG form
    const size_t n1 = out.n1();
    {
        std::vector<cplx> tmp(n1 * s->NP);
        sg_eval(sg_for(s, ch, 1, false), tmp, tmp.size(), [&](size_t ii) {
            const size_t i = ii % n1;
            const int P = (int)(ii / n1);
            int m = out.b1.val((int)i);
            cplx val = 0.0;
            for (int iw = 0; iw < s->mPiF.size(); ++iw) {
                int w = s->mPiF.val(iw);
                size_t k = s->pidx(m, w, P);
                if (is_mfrg) {
                    cplx F0l = s->F0v.eval(m, INF, w, 0, mch, msp, FL_ALL);
                    cplx FLr = s->FL.eval(m, crossing(m, w, ch), INF, P, mch, msp, FL_ALL);
                    val += F0l * Pi0[k] * FLr;
                } else {
                    cplx Fl = s->F.eval(m, INF, w, P, ch, sp, FL_ALL);
                    cplx F0r = s->F0v.eval(m, crossing(m, w, ch), INF, 0, ch, sp, FL_ALL);
                    cplx FLr = s->FL.eval(m, crossing(m, w, ch), INF, P, ch, sp, FL_ALL);
                    val += Fl * ((Pi[k] - Pi0[k]) * F0r + Pi[k] * FLr);
                }
            }
            return s->T * val * (double)sg;
        });
        std::copy(tmp.begin(), tmp.end(), out.K1.begin());
    }
    if (ch == TCH) t_post(s->Fbuff.g[TCH].K1, s->Fbuff.g[ACH].K1);
}

// BSE_K1_new! (src/BSEa/BSEa_K1.jl:57-100; NL src/nonlocal/BSEa/BSEa_K1.jl:62-108):
// K1 = T s Σ_ω (F(Ω,νInf,ω) Π - F0(Ω,νInf,ω) Π0) U with U = bare_vertex(F, Sp); the lattice is_mfRG form
// is (F - F0) Π U, the local form has no is_mfRG branch
void orc_bse_k1_new(orc_solver* s, int32_t ch, int32_t is_mfrg) {
    s->F.refresh();
    Chan& out = s->Fbuff.g[ch];
    const std::vector<cplx>& Pi0 = (ch == PCH) ? s->Pi0pp : s->Pi0ph;
    const std::vector<cplx>& Pi = (ch == PCH) ? s->Pipp : s->Piph;
    const int sp = kSp[ch], sg = kSign[ch];
    const bool mf = is_mfrg && s->prm.nonlocal;
    const cplx U = sp == XSP ? -s->F.bare() : s->F.bare();
    const size_t n1 = out.n1();
    {
        std::vector<cplx> tmp(n1 * s->NP);
        sg_eval(sg_for(s, ch, 1, false), tmp, tmp.size(), [&](size_t ii) {
            const size_t i = ii % n1;
            const int P = (int)(ii / n1);
            int m = out.b1.val((int)i);
            cplx val = 0.0;
            for (int iw = 0; iw < s->mPiF.size(); ++iw) {
                int w = s->mPiF.val(iw);
                size_t k = s->pidx(m, w, P);
                cplx Fl = s->F.eval(m, INF, w, P, ch, sp, FL_ALL);
                cplx F0l = s->F0v.eval(m, INF, w, 0, ch, sp, FL_ALL);
                val += mf ? (Fl - F0l) * Pi[k] * U : (Fl * Pi[k] - F0l * Pi0[k]) * U;
            }
            return s->T * val * (double)sg;
        });
        std::copy(tmp.begin(), tmp.end(), out.K1.begin());
    }
    if (ch == TCH) t_post(s->Fbuff.g[TCH].K1, s->Fbuff.g[ACH].K1);
}

// BSE_K2! (src/BSEa/BSEa_K2.jl:43-103; NL src/nonlocal/BSEa/BSEa_K2.jl:45-104).  is_mfRG = true: the central
// part (F0(Ω,ν,.) - F0(Ω,νInf,.)) Π0 FL(Ω, ., νInf) with the crossing on the right vertex (local,
// src/BSEa/BSEa_K2.jl:70-78) or on the left one (lattice, src/nonlocal/BSEa/BSEa_K2.jl:70-78)
void orc_bse_k2(orc_solver* s, int32_t ch, int32_t is_mfrg) {
    s->F.refresh();
    s->FL.refresh();
    Chan& out = s->Fbuff.g[ch];
    const std::vector<cplx>& Pi0 = (ch == PCH) ? s->Pi0pp : s->Pi0ph;
    const std::vector<cplx>& Pi = (ch == PCH) ? s->Pipp : s->Piph;
    const int sp = kSp[ch], sg = kSign[ch];
    const bool nl = s->prm.nonlocal != 0;
    const size_t n2 = out.n2();
    {
        std::vector<cplx> tmp(n2 * s->NP);
        sg_eval(sg_for(s, ch, 2, false), tmp, tmp.size(), [&](size_t ii) {
            const size_t i = ii % n2;
            const int P = (int)(ii / n2);
            int m = out.b2.val((int)(i % out.b2.size())), v = out.f2.val((int)(i / out.b2.size()));
            cplx val = 0.0;
            for (int iw = 0; iw < s->mPiF.size(); ++iw) {
                int w = s->mPiF.val(iw);
                size_t k = s->pidx(m, w, P);
                if (is_mfrg) {
                    const int wl = nl ? crossing(m, w, ch) : w, wr = nl ? w : crossing(m, w, ch);
                    cplx F0l = s->F0v.eval(m, v, wl, 0, ch, sp, FL_ALL) - s->F0v.eval(m, INF, wl, 0, ch, sp, FL_ALL);
                    cplx FLr = s->FL.eval(m, wr, INF, P, ch, sp, FL_ALL);
                    val += F0l * Pi0[k] * FLr;
                } else {
                    cplx Fl = s->F.eval(m, v, w, P, ch, sp, FL_ALL) - s->F.eval(m, INF, w, P, ch, sp, FL_ALL);
                    cplx F0r = s->F0v.eval(m, crossing(m, w, ch), INF, 0, ch, sp, FL_ALL);
                    cplx FLr = s->FL.eval(m, crossing(m, w, ch), INF, P, ch, sp, FL_ALL);
                    val += Fl * ((Pi[k] - Pi0[k]) * F0r + Pi[k] * FLr);
                }
            }
            return s->T * val * (double)sg;
        });
        std::copy(tmp.begin(), tmp.end(), out.K2.begin());
    }
    // add FL.K2 (t: 2 FLt - FLa) (src/BSEa/BSEa_K2.jl:95-100)
    if (ch == TCH) {
        for (size_t i = 0; i < out.K2.size(); ++i) out.K2[i] += s->FL.g[TCH].K2[i] * 2.0;
        for (size_t i = 0; i < out.K2.size(); ++i) out.K2[i] += s->FL.g[ACH].K2[i] * -1.0;
        t_post(s->Fbuff.g[TCH].K2, s->Fbuff.g[ACH].K2);
    } else {
        for (size_t i = 0; i < out.K2.size(); ++i) out.K2[i] += s->FL.g[ch].K2[i];
    }
}

// BSE_K2_new! (src/BSEa/BSEa_K2.jl:108-151; NL src/nonlocal/BSEa/BSEa_K2.jl:108-155):
// K2 = T s Σ_ω ((F(ν) - F(νInf)) Π - (F0(ν) - F0(νInf)) Π0) U; lattice is_mfRG: (ΔF - ΔF0) Π U; no FL term
void orc_bse_k2_new(orc_solver* s, int32_t ch, int32_t is_mfrg) {
    s->F.refresh();
    Chan& out = s->Fbuff.g[ch];
    const std::vector<cplx>& Pi0 = (ch == PCH) ? s->Pi0pp : s->Pi0ph;
    const std::vector<cplx>& Pi = (ch == PCH) ? s->Pipp : s->Piph;
    const int sp = kSp[ch], sg = kSign[ch];
    const bool mf = is_mfrg && s->prm.nonlocal;
    const cplx U = sp == XSP ? -s->F.bare() : s->F.bare();
    const size_t n2 = out.n2();
    {
        std::vector<cplx> tmp(n2 * s->NP);
        sg_eval(sg_for(s, ch, 2, false), tmp, tmp.size(), [&](size_t ii) {
            const size_t i = ii % n2;
            const int P = (int)(ii / n2);
            int m = out.b2.val((int)(i % out.b2.size())), v = out.f2.val((int)(i / out.b2.size()));
            cplx val = 0.0;
            for (int iw = 0; iw < s->mPiF.size(); ++iw) {
                int w = s->mPiF.val(iw);
                size_t k = s->pidx(m, w, P);
                cplx Fl = s->F.eval(m, v, w, P, ch, sp, FL_ALL) - s->F.eval(m, INF, w, P, ch, sp, FL_ALL);
                cplx F0l = s->F0v.eval(m, v, w, 0, ch, sp, FL_ALL) - s->F0v.eval(m, INF, w, 0, ch, sp, FL_ALL);
                val += mf ? (Fl - F0l) * Pi[k] * U : (Fl * Pi[k] - F0l * Pi0[k]) * U;
            }
            return s->T * val * (double)sg;
        });
        std::copy(tmp.begin(), tmp.end(), out.K2.begin());
    }
    if (ch == TCH) t_post(s->Fbuff.g[TCH].K2, s->Fbuff.g[ACH].K2);
}

// BSE_K3! (src/BSEa/BSEa_K3.jl:43-128; NL src/nonlocal/BSEa/BSEa_K3.jl:44-128).  is_mfRG = true
// (src/BSEa/BSEa_K3.jl:83-97): A Π0 (s1 Γ~ + s2 L~), no (Π - Π0) F0 term
void orc_bse_k3(orc_solver* s, int32_t ch, int32_t is_mfrg) {
    Chan& out = s->Fbuff.g[ch];
    const std::vector<cplx>& Pi0 = (ch == PCH) ? s->Pi0pp : s->Pi0ph;
    const std::vector<cplx>& Pi = (ch == PCH) ? s->Pipp : s->Piph;
    const std::vector<cplx>& CG = s->cache[ch == PCH ? 0 : ch == ACH ? 5 : 6];
    const std::vector<cplx>& CF = s->cache[ch == PCH ? 7 : ch == ACH ? 8 : 9];
    const std::vector<cplx>& CF0 = s->cache[ch == PCH ? 1 : ch == ACH ? 2 : 3];
    const int s1 = (ch == ACH) ? +1 : -1;  // src/BSE_templates.jl:156,164,172
    const int s2 = (ch == TCH) ? -1 : +1;
    const Chan& Lt = s->FL.g[TCH];
    const Chan& La = s->FL.g[ACH];
    const Chan& Lr = s->FL.g[ch];
    const int nb = out.b3.size(), nf = out.f3.size();
    const size_t n3 = out.n3();
    auto c3 = [&](const std::vector<cplx>& C, int i1, int i2, int i3, int P) {
        return C[i1 + (size_t)nb * (i2 + (size_t)nf * (i3 + (size_t)nf * P))];
    };
    {
        std::vector<cplx> tmp(n3 * s->NP);
        sg_eval(sg_for(s, ch, 3, false), tmp, tmp.size(), [&](size_t ii) {
            const size_t i = ii % n3;
            const int P = (int)(ii / n3);
            int i1 = (int)(i % nb), i2 = (int)((i / nb) % nf), i3 = (int)(i / ((size_t)nb * nf));
            int m = out.b3.val(i1), v = out.f3.val(i2), vp = out.f3.val(i3);
            cplx val = 0.0;
            for (int iw = 0; iw < nf; ++iw) {
                int w = out.f3.val(iw);
                cplx A = c3(CF, i1, i2, iw, P);
                cplx Gs = (ch == PCH) ? c3(CG, i1, iw, i3, P) : c3(CG, i1, i3, iw, P);
                cplx F0s = c3(CF0, i1, iw, i3, P);
                size_t k = s->pidx(m, w, P);
                if (is_mfrg) val += A * Pi0[k] * Gs * (double)s1;
                else val += A * ((Pi[k] - Pi0[k]) * F0s + Pi[k] * Gs) * (double)s1;
                int w_ = crossing(m, w, ch);
                if (out.f3.in(w_)) {
                    cplx Lx;
                    if (ch == TCH) Lx = 2.0 * Lt.K3[Lt.i3(m, w_, vp) + Lt.n3() * P] - La.K3[La.i3(m, w_, vp) + La.n3() * P];
                    else Lx = Lr.K3[Lr.i3(m, w_, vp) + Lr.n3() * P];
                    val += A * (is_mfrg ? Pi0[k] : Pi[k]) * Lx * (double)s2;
                }
            }
            cplx self;
            if (ch == TCH) self = 2.0 * Lt.K3[Lt.i3(m, v, vp) + Lt.n3() * P] - La.K3[La.i3(m, v, vp) + La.n3() * P];
            else self = Lr.K3[Lr.i3(m, v, vp) + Lr.n3() * P];
            return s->T * val + self;
        });
        std::copy(tmp.begin(), tmp.end(), out.K3.begin());
    }
    if (ch == TCH) t_post(s->Fbuff.g[TCH].K3, s->Fbuff.g[ACH].K3);
}

void orc_update_F(orc_solver* s) {
    for (int r = 0; r < 3; ++r) {
        s->F.g[r].K1 = s->Fbuff.g[r].K1;
        s->F.g[r].K2 = s->Fbuff.g[r].K2;
        s->F.g[r].K3 = s->Fbuff.g[r].K3;
        s->F.g[r].refresh();
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// SDE (src/SDE.jl:3-33, 35-48, 79-283; NL src/nonlocal/SDE.jl:3-330)
namespace {

struct SdeIn {
    const std::vector<cplx>* G;
    const std::vector<cplx>* Pipp;
    const std::vector<cplx>* Piph;
    std::vector<cplx>* Lpp;
    std::vector<cplx>* Lph;
    const SymGroup* SGpp2;
    const SymGroup* SGph2;
};

// SDE_channel_L_pp! / _ph! for a Vertex (src/SDE.jl:79-150; NL src/nonlocal/SDE.jl:3-75)
// and for a RefVertex (src/SDE.jl:154-217; NL src/nonlocal/SDE.jl:79-142)
void sde_channel_L(orc_solver* s, const SdeIn& in, const Vert& Fv) {
    const Chan& shape = s->F.g[0];
    const int nb = shape.b2.size(), nf = shape.f2.size();
    const size_t n2 = (size_t)nb * nf;
    const cplx U = Fv.bare();
    const bool isref = Fv.ref != nullptr;
    {
        std::vector<cplx> tpp(n2 * s->NP), tph(n2 * s->NP);
        sg_eval(*in.SGpp2, tpp, tpp.size(), [&](size_t ii) {
            const size_t i = ii % n2;
            const int P = (int)(ii / n2);
            int m = shape.b2.val((int)(i % nb)), v = shape.f2.val((int)(i / nb));
            cplx val = 0.0;
            for (int iw = 0; iw < s->mPiF.size(); ++iw) {
                int w = s->mPiF.val(iw);
                cplx pi = (*in.Pipp)[s->pidx(m, w, P)];
                if (isref) val += U * pi * (Fv.eval(m, bmf(m, w), v, 0, PCH, PSP, FL_ALL) - Fv.ref->bare(PSP));
                else val += U * pi * Fv.g[PCH].eval(m, bmf(m, w), v, Fv.g[PCH].NP > 1 ? P : 0);
            }
            return s->T * val;
        });
        sg_eval(*in.SGph2, tph, tph.size(), [&](size_t ii) {
            const size_t i = ii % n2;
            const int P = (int)(ii / n2);
            int m = shape.b2.val((int)(i % nb)), v = shape.f2.val((int)(i / nb));
            cplx val = 0.0;
            for (int iw = 0; iw < s->mPiF.size(); ++iw) {
                int w = s->mPiF.val(iw);
                cplx pi = (*in.Piph)[s->pidx(m, w, P)];
                if (isref) {
                    val += U * pi * (Fv.eval(m, v, w, 0, ACH, PSP, FL_ALL) + Fv.eval(m, v, w, 0, TCH, PSP, FL_ALL)
                                     - Fv.ref->bare(PSP) - Fv.ref->bare(PSP));
                } else if (!s->prm.nonlocal) {
                    // local: F(aCh; γa only) + F(tCh; γt only) (src/SDE.jl:134-137)
                    val += U * pi * (Fv.eval(m, v, w, 0, ACH, PSP, FL_GA) + Fv.eval(m, v, w, 0, TCH, PSP, FL_GT));
                } else {
                    int pp = Fv.g[TCH].NP > 1 ? P : 0;
                    val += U * pi * (Fv.g[TCH].eval(m, v, w, pp) + Fv.g[ACH].eval(m, v, w, pp));
                }
            }
            return s->T * val;
        });
        std::copy(tpp.begin(), tpp.end(), in.Lpp->begin());
        std::copy(tph.begin(), tph.end(), in.Lph->begin());
    }
}

// SDE_U2_using_G (local src/SDE.jl:328-364; NL src/nonlocal/SDE.jl:400-447)
std::vector<cplx> sde_u2(orc_solver* s, const std::vector<cplx>& G, cplx U) {
    const int NG = s->mG.size();
    std::vector<cplx> out(G.size(), 0.0);
    const double T = s->T;
    if (!s->prm.nonlocal) {
        std::vector<cplx> tmp(NG);
        sg_eval(s->SGS, tmp, NG, [&](size_t iv) {
            int v = s->mG.val((int)iv);
            cplx val = 0.0;
            for (int i2 = 0; i2 < NG; ++i2)
                for (int i1 = 0; i1 < NG; ++i1) {
                    int w1 = s->mG.val(i1), w2 = s->mG.val(i2);
                    int x = bpf(fmf(w1, w2), v);
                    if (s->mG.in(x)) val += G[i1] * G[i2] * G[s->mG.idx(x)];
                }
            return val * U * U * T * T;
        });
        return tmp;
    }
    const int L = s->prm.LG;
    std::vector<cplx> Gp = G, Gm = G;
    dft2(Gp, NG, L, -1);
    dft2(Gm, NG, L, +1);
    for (auto& x : Gp) x /= (double)L * L;
    for (auto& x : Gm) x /= (double)L * L;
    std::vector<cplx> SR(G.size(), 0.0);
#pragma omp parallel for
    for (int R = 0; R < L * L; ++R)
        for (int i2 = 0; i2 < NG; ++i2)
            for (int i1 = 0; i1 < NG; ++i1) {
                int w1 = s->mG.val(i1), w2 = s->mG.val(i2);
                cplx gg = Gm[i1 + (size_t)NG * R] * Gp[i2 + (size_t)NG * R];
                for (int iv = 0; iv < NG; ++iv) {
                    int x = bpf(fmf(w1, w2), s->mG.val(iv));
                    if (s->mG.in(x)) SR[iv + (size_t)NG * R] += gg * Gp[s->mG.idx(x) + (size_t)NG * R];
                }
            }
    for (auto& x : SR) x *= U * U * T * T;
    dft2(SR, NG, L, +1);
    sg_symmetrize(s->SGS, SR);
    return SR;
}

// SDE_compute! (local src/SDE.jl:219-283; NL src/nonlocal/SDE.jl:146-330 with use_real_space)
std::vector<cplx> sde_compute(orc_solver* s, const SdeIn& in, const Vert& Fv, bool incU2, bool incH) {
    sde_channel_L(s, in, Fv);
    const Chan& shape = s->F.g[0];
    const int nb = shape.b2.size(), nf = shape.f2.size();
    const int NG = s->mG.size();
    const std::vector<cplx>& G = *in.G;
    std::vector<cplx> Sig(G.size(), 0.0);
    if (!s->prm.nonlocal) {
        sg_eval(s->SGS, Sig, NG, [&](size_t iv) {
            int v = s->mG.val((int)iv);
            cplx val = 0.0;
            if (shape.f2.in(v)) {
                for (int i = 0; i < nb; ++i) {
                    int m = shape.b2.val(i);
                    size_t k = i + (size_t)nb * shape.f2.idx(v);
                    int a = bmf(m, v), b = bpf(m, v);
                    val += (s->mG.in(a) ? G[s->mG.idx(a)] : cplx(0.0)) * (*in.Lpp)[k];
                    val += (s->mG.in(b) ? G[s->mG.idx(b)] : cplx(0.0)) * (*in.Lph)[k];
                }
            }
            return s->T * val;
        });
    } else {
        // real-space evaluation (src/nonlocal/SDE.jl:209-280)
        const int LG = s->prm.LG, L = s->prm.L;
        std::vector<cplx> GR = G;
        dft2(GR, NG, LG, -1);
        for (auto& x : GR) x /= (double)LG * LG;
        const size_t n2 = (size_t)nb * nf;
        std::vector<cplx> LppR = *in.Lpp, LphR = *in.Lph;
        dft2(LppR, n2, L, -1);
        dft2(LphR, n2, L, -1);
        for (auto& x : LppR) x /= (double)L * L;
        for (auto& x : LphR) x /= (double)L * L;
        std::vector<cplx> SR(G.size(), 0.0);
        const int h = L / 2;
        for (int R2 = -h; R2 <= h; ++R2)
            for (int R1 = -h; R1 <= h; ++R1) {
                double w = 1.0;
                if (L % 2 == 0) {
                    if (std::abs(R1) == L / 2) w /= 2;
                    if (std::abs(R2) == L / 2) w /= 2;
                }
                int iRL = imod(R1, L) + L * imod(R2, L);
                int imRG = imod(-R1, LG) + LG * imod(-R2, LG);
                int ipRS = imod(R1, LG) + LG * imod(R2, LG);
                int imRS = imRG;
#pragma omp parallel for
                for (int iv = 0; iv < nf; ++iv) {
                    int v = shape.f2.val(iv);
                    if (!s->mG.in(v)) continue;
                    for (int iW = 0; iW < nb; ++iW) {
                        int m = shape.b2.val(iW);
                        size_t k = iW + (size_t)nb * iv + n2 * iRL;
                        int a = bmf(m, v), b = bpf(m, v);
                        if (s->mG.in(a)) SR[s->mG.idx(v) + (size_t)NG * ipRS] += GR[s->mG.idx(a) + (size_t)NG * imRG] * LppR[k] * w;
                        if (s->mG.in(b)) SR[s->mG.idx(v) + (size_t)NG * imRS] += GR[s->mG.idx(b) + (size_t)NG * imRG] * LphR[k] * w;
                    }
                }
            }
        for (auto& x : SR) x *= s->T;
        dft2(SR, NG, LG, +1);
        sg_symmetrize(s->SGS, SR);
        Sig = SR;
    }
    if (Fv.ref) for (auto& x : Sig) x *= 1.0 / 3.0;
    const cplx U = Fv.bare();
    if (incU2) {
        std::vector<cplx> u2 = sde_u2(s, G, U);
        for (size_t i = 0; i < Sig.size(); ++i) Sig[i] += u2[i];
    }
    if (incH) {
        double n = occupation(s, G);
        cplx h = (n - 0.5) * U * cplx(0.0, 1.0);
        for (auto& x : Sig) x += h;
    }
    return Sig;
}

// SDE!(Σ, G, Πpp, Πph, Lpp, Lph, F, ...) recursion over F, F.F0, ... (src/SDE.jl:35-48)
void sde_rec(orc_solver* s, std::vector<cplx>& Sig, const SdeIn& in, const Vert& Fv, bool incU2, bool incH) {
    std::vector<cplx> add = sde_compute(s, in, Fv, incU2, incH);
    for (size_t i = 0; i < Sig.size(); ++i) Sig[i] += add[i];
    if (!Fv.ref) sde_rec(s, Sig, in, *Fv.F0, false, false);
}

}  // namespace

extern "C" {

// SDE!(S; strategy) (src/SDE.jl:3-33)
void orc_sde(orc_solver* s, int32_t strategy, int32_t include_hartree) {
    s->F.refresh();
    s->F0v.refresh();
    std::fill(s->Sig.begin(), s->Sig.end(), 0.0);
    SdeIn in{&s->G, &s->Pipp, &s->Piph, &s->Lpp, &s->Lph, &s->SGpp[2], &s->SGph[2]};
    sde_rec(s, s->Sig, in, s->F, true, include_hartree != 0);
    if (strategy == ORC_FDPA || strategy == ORC_FDPA_NEW) {
        std::vector<cplx> S0(s->Sig.size(), 0.0);
        SdeIn in0{&s->G0, &s->Pi0pp, &s->Pi0ph, &s->L0pp, &s->L0ph, &s->SG0pp2, &s->SG0ph2};
        sde_rec(s, S0, in0, s->F0v, true, include_hartree != 0);
        for (size_t i = 0; i < s->Sig.size(); ++i) s->Sig[i] += S0[i] * -1.0;
        for (size_t i = 0; i < s->Sig.size(); ++i) s->Sig[i] += s->Sig0[i];
        if (include_hartree && s->prm.fdpa_h0) {
            double n0 = occupation(s, s->G0);
            cplx h = (n0 - 0.5) * s->F.bare() * cplx(0.0, 1.0);
            for (auto& x : s->Sig) x -= h;
        }
    }
}

// iterate_solver!(S; strategy, update_Σ, compute_Hartree) (src/solve.jl:4-116), strategies :scPA, :fdPA,
// :scPA_new and :fdPA_new (K3 first, then the alternative K1/K2 forms, src/solve.jl:26-46)
void orc_iterate(orc_solver* s, int32_t strategy, int32_t update_sigma, int32_t compute_hartree) {
    if (update_sigma) {
        orc_dyson(s);
        orc_bubbles(s);
    }
    orc_build_k3_cache(s);
    if (strategy == ORC_SCPA_NEW || strategy == ORC_FDPA_NEW) {
        if (strategy == ORC_FDPA_NEW)
            for (int ch : {PCH, ACH, TCH}) orc_bse_l_k3(s, ch);
        for (int ch : {PCH, ACH, TCH}) orc_bse_k3(s, ch, 0);
        for (int ch : {PCH, ACH, TCH}) orc_bse_k1_new(s, ch, 0);
        for (int ch : {PCH, ACH, TCH}) orc_bse_k2_new(s, ch, 0);
    } else {
        if (strategy == ORC_FDPA) {
            for (int ch : {PCH, ACH, TCH}) orc_bse_l_k2(s, ch);
            for (int ch : {PCH, ACH, TCH}) orc_bse_l_k3(s, ch);
        }
        for (int ch : {PCH, ACH, TCH}) orc_bse_k1(s, ch, 0);
        for (int ch : {PCH, ACH, TCH}) orc_bse_k2(s, ch, 0);
        for (int ch : {PCH, ACH, TCH}) orc_bse_k3(s, ch, 0);
    }
    orc_update_F(s);
    // SDE!(S; strategy): the _new strategies take the Σ update of their base strategy (src/SDE.jl:4-28)
    if (update_sigma) orc_sde(s, (strategy == ORC_FDPA || strategy == ORC_FDPA_NEW) ? ORC_FDPA : ORC_SCPA, compute_hartree);
}

// flatten layout: [γp(K1,K2,K3); γt(...); γa(...); Σ] (src/ParquetSolver.jl:277-306, src/vertex.jl:153-196)
int64_t orc_flat_size(const orc_solver* s) {
    int64_t n = 0;
    for (int r = 0; r < 3; ++r) n += s->F.g[r].K1.size() + s->F.g[r].K2.size() + s->F.g[r].K3.size();
    return n + (int64_t)s->Sig.size();
}
void orc_flatten(const orc_solver* s, double* x) {
    cplx* o = reinterpret_cast<cplx*>(x);
    for (int r = 0; r < 3; ++r)
        for (auto* v : {&s->F.g[r].K1, &s->F.g[r].K2, &s->F.g[r].K3}) o = std::copy(v->begin(), v->end(), o);
    std::copy(s->Sig.begin(), s->Sig.end(), o);
}
void orc_unflatten(orc_solver* s, const double* x) {
    const cplx* in = reinterpret_cast<const cplx*>(x);
    for (int r = 0; r < 3; ++r) {
        for (auto* v : {&s->F.g[r].K1, &s->F.g[r].K2, &s->F.g[r].K3}) {
            std::copy(in, in + v->size(), v->begin());
            in += v->size();
        }
        s->F.g[r].refresh();
    }
    std::copy(in, in + s->Sig.size(), s->Sig.begin());
}

// fixed_point! (src/solve.jl:119-157).  update_Σ = false: x and R hold the vertex only, R = flatten(S.F) - x
// (the vertex-only vector of src/mfRG.jl:111-173: unflatten!(S.F, x) asserts length(x) == length(S.F),
// src/vertex.jl:195)
int64_t orc_flat_size_F(const orc_solver* s) { return orc_flat_size(s) - (int64_t)s->Sig.size(); }

void orc_fixed_point(orc_solver* s, double* R, const double* x, int32_t strategy, int32_t update_sigma,
                     int32_t compute_hartree) {
    if (update_sigma) {
        orc_unflatten(s, x);
    } else {
        std::vector<cplx> keep = s->Sig;
        std::vector<double> full((size_t)orc_flat_size(s) * 2);
        std::memcpy(full.data(), x, (size_t)orc_flat_size_F(s) * 2 * sizeof(double));
        std::memcpy(full.data() + (size_t)orc_flat_size_F(s) * 2, reinterpret_cast<const double*>(keep.data()),
                    keep.size() * 2 * sizeof(double));
        orc_unflatten(s, full.data());
    }
    orc_iterate(s, strategy, update_sigma, compute_hartree);
    int64_t n = (update_sigma ? orc_flat_size(s) : orc_flat_size_F(s)) * 2;
    if (update_sigma) {
        orc_flatten(s, R);
    } else {
        std::vector<double> full((size_t)orc_flat_size(s) * 2);
        orc_flatten(s, full.data());
        std::memcpy(R, full.data(), (size_t)n * sizeof(double));
    }
    for (int64_t i = 0; i < n; ++i) R[i] -= x[i];
}

// mfRGLinearMap * x (src/mfRG.jl:34-89): y = x - G_lin(x / 100) * 100, where G_lin is one pass of the
// linearised BSE around the reference vertex: F = unflatten(x / 100), build_K3_cache_mfRG!, for :fdPA
// BSE_L_K2! and the is_mfRG forms of BSE_K1!/BSE_K2!, for :fdPA_new the is_mfRG forms of BSE_K1_new!/
// BSE_K2_new!, then BSE_L_K3! and the is_mfRG BSE_K3!, F = Fbuff.  x, y: the vertex only.
void orc_mfrg_apply(orc_solver* s, double* y, const double* x, int32_t strategy, int32_t is_first) {
    const double factor = 1e-2;
    const int64_t nF = orc_flat_size_F(s);
    {
        std::vector<double> full((size_t)orc_flat_size(s) * 2);
        for (int64_t i = 0; i < 2 * nF; ++i) full[i] = x[i] * factor;
        std::memcpy(full.data() + 2 * nF, reinterpret_cast<const double*>(s->Sig.data()), s->Sig.size() * 2 * sizeof(double));
        orc_unflatten(s, full.data());
    }
    orc_build_k3_cache_mfrg(s, is_first);
    if (strategy == ORC_FDPA) {
        for (int ch : {PCH, ACH, TCH}) orc_bse_l_k2(s, ch);
        for (int ch : {PCH, ACH, TCH}) orc_bse_k1(s, ch, 1);
        for (int ch : {PCH, ACH, TCH}) orc_bse_k2(s, ch, 1);
    } else {
        for (int ch : {PCH, ACH, TCH}) orc_bse_k1_new(s, ch, 1);
        for (int ch : {PCH, ACH, TCH}) orc_bse_k2_new(s, ch, 1);
    }
    for (int ch : {PCH, ACH, TCH}) orc_bse_l_k3(s, ch);
    for (int ch : {PCH, ACH, TCH}) orc_bse_k3(s, ch, 1);
    orc_update_F(s);
    std::vector<double> full((size_t)orc_flat_size(s) * 2);
    orc_flatten(s, full.data());
    for (int64_t i = 0; i < 2 * nF; ++i) y[i] = x[i] - full[i] / factor;
}

// NLsolve 4.5.1 Anderson (method=:anderson, aa_start=1, droptol=0, xtol=0):
//   g = x + β f(x); converged when max|f| < ftol; history of Δg and Δf of size m;
//   x_{k+1} = g_k - Σ_i γ_i Δg_i with γ = argmin ||f_k - ΔF γ|| (QR least squares).
int orc_solve(orc_solver* s, int32_t strategy, int32_t maxiter, double tol, double beta, int32_t mem,
              int32_t update_sigma, double* final_res) {
    const int64_t n = update_sigma ? orc_flat_size(s) : orc_flat_size_F(s);
    std::vector<cplx> x(n), fx(n), g(n), gold(n), fold(n);
    {
        std::vector<cplx> full(orc_flat_size(s));
        orc_flatten(s, reinterpret_cast<double*>(full.data()));
        std::copy(full.begin(), full.begin() + n, x.begin());
    }
    const int mmax = (int)std::min<int64_t>(mem, n);
    std::vector<std::vector<cplx>> dG, dF;  // oldest first
    int it = 0;
    double res = INFINITY;
    for (it = 1; it <= maxiter; ++it) {
        orc_fixed_point(s, reinterpret_cast<double*>(fx.data()), reinterpret_cast<const double*>(x.data()), strategy, update_sigma, 1);
        for (int64_t i = 0; i < n; ++i) g[i] = x[i] + beta * fx[i];
        res = 0.0;
        for (auto& v : fx) res = std::max(res, std::abs(v));
        if (res < tol) break;
        x = g;
        if (mmax == 0) continue;
        if (it == 1) {
            fold = fx;
            gold = g;
            continue;
        }
        std::vector<cplx> dg(n), df(n);
        for (int64_t i = 0; i < n; ++i) { dg[i] = g[i] - gold[i]; df[i] = fx[i] - fold[i]; }
        gold = g;
        fold = fx;
        dG.push_back(std::move(dg));
        dF.push_back(std::move(df));
        if ((int)dG.size() > mmax) { dG.erase(dG.begin()); dF.erase(dF.begin()); }
        const int k = (int)dF.size();
        // modified Gram-Schmidt QR of ΔF
        std::vector<std::vector<cplx>> Q(k);
        std::vector<cplx> Rm((size_t)k * k, 0.0);
        for (int j = 0; j < k; ++j) {
            std::vector<cplx> v = dF[j];
            for (int i = 0; i < j; ++i) {
                cplx r = 0.0;
                for (int64_t t = 0; t < n; ++t) r += std::conj(Q[i][t]) * v[t];
                Rm[i + (size_t)k * j] = r;
                for (int64_t t = 0; t < n; ++t) v[t] -= r * Q[i][t];
            }
            double nrm = 0.0;
            for (auto& z : v) nrm += std::norm(z);
            nrm = std::sqrt(nrm);
            Rm[j + (size_t)k * j] = nrm;
            for (auto& z : v) z /= nrm;
            Q[j] = std::move(v);
        }
        std::vector<cplx> gam(k, 0.0);
        for (int i = 0; i < k; ++i) {
            cplx r = 0.0;
            for (int64_t t = 0; t < n; ++t) r += std::conj(Q[i][t]) * fx[t];
            gam[i] = r;
        }
        for (int i = k - 1; i >= 0; --i) {
            for (int j = i + 1; j < k; ++j) gam[i] -= Rm[i + (size_t)k * j] * gam[j];
            gam[i] /= Rm[i + (size_t)k * i];
        }
        for (int i = 0; i < k; ++i)
            for (int64_t t = 0; t < n; ++t) x[t] -= gam[i] * dG[i][t];
    }
    if (final_res) *final_res = res;
    return it > maxiter ? maxiter : it;
}

void orc_vertex_eval(const orc_solver* s, int32_t vtx, int32_t m, int32_t n, int32_t np, int32_t iP,
                     int32_t ch, int32_t sp, int32_t flags, double* out) {
    const Vert* v = vtx == 0 ? &s->F : vtx == 1 ? &s->Fbuff : vtx == 2 ? &s->FL : &s->F0v;
    cplx r = v->eval(m, n, np, iP, ch, sp, flags);
    out[0] = r.real();
    out[1] = r.imag();
}

void orc_set_threads(int32_t n) { omp_set_num_threads(n > 0 ? n : 1); }
int32_t orc_max_threads(void) { return omp_get_max_threads(); }

double orc_occupation(const orc_solver* s, int32_t which) {
    return occupation(s, which == ORC_G0 ? s->G0 : which == ORC_GBARE ? s->Gbare : s->G);
}

}  // extern "C"
