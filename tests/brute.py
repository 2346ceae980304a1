"""Brute-force restatement of the reference's BSE formulas in plain Python / numpy, for tiny boxes.

TEST INFRASTRUCTURE ONLY.  It is written from the reference sources, not from the C++ oracle, and
it is the independent side of the oracle's pins (tests/test_oracle_pins.py).  It keeps the target
vertex K1-only (K2 = K3 = 0), where every vertex evaluation has the closed form of the reference's
own test (test/test_nonlocal_vertex.jl:148-154, s-wave form: the cross channels at the momentum
average):

    F(Ω, ν, ω, pCh) = F0 + γp.K1(Ω, P) + γ̄t.K1(Ω - ν - ω) + γ̄a.K1(ν - ω)
    F(Ω, ν, ω, tCh) = F0 + γt.K1(Ω, P) + γ̄p.K1(Ω + ν + ω) + γ̄a.K1(ω - ν)
    F(Ω, ν, ω, aCh) = F0 + γa.K1(Ω, P) + γ̄p.K1(Ω + ν + ω) + γ̄t.K1(ν - ω)

and an infinite fermionic argument keeps only the channel's own class (src/vertex.jl:246-284).  The
reference vertex F0 is a RefVertex with (random, non-symmetric) tables (src/refvertex.jl:90-216);
the left vertex FL of the fdPA has random K2 and K3 classes and no K1 (src/channel.jl:220-339 for
its evaluation).  Π and Π0 are random, non-symmetric arrays.  The BSE stages follow
src/build_K3_cache.jl, src/BSEa/BSEa_K{1,2,3}.jl, src/nonlocal/BSEa/BSEa_K{1,2,3}.jl,
src/BSE_templates.jl and src/nonlocal_2/build_K3_cache.jl:97-164.

Indices: bosonic m (Ω = 2πT m) on a box N: m = -(N-1) .. N-1; fermionic n (ν = πT(2n+1)):
n = -N .. N-1.  In index arithmetic Ω - ν - ω -> m - n - w - 1, ν - ω -> n - w, Ω + ν + ω ->
m + n + w + 1, Ω - ν -> m - n - 1, Ω + ν -> m + n.  INF (None) is νInf.
"""
import numpy as np

PCH, TCH, ACH = 0, 1, 2
PSP, XSP, DSP = 0, 1, 2
F_F0, F_GP, F_GT, F_GA = 1, 2, 4, 8
F_ALL = 15
INF = None
SIGN = {PCH: +1, TCH: -1, ACH: +1}     # BSE_K1!/BSE_L_K2!/BSE_K2! signs (src/BSE_templates.jl:12-111)
SPIN = {PCH: PSP, TCH: DSP, ACH: PSP}


def inb(m, N):
    return m is not None and -(N - 1) <= m <= N - 1


def inf_(n, N):
    return n is not None and -N <= n <= N - 1


def cross(m, w, ch):
    """_crossing (src/BSE_templates.jl:4-6): Ω - ω for the p channel, ω otherwise."""
    if m is None or w is None:
        return None
    return m - w - 1 if ch == PCH else w


class Ref:
    """RefVertex (src/refvertex.jl:90-216): tables Fp_p, Fp_x, Ft_p, Ft_x on (tb, tf) boxes."""

    def __init__(self, U, tb, tf, tabs):
        self.U, self.tb, self.tf = U, tb, tf
        self.t = {k: np.asarray(v).reshape(2 * tb - 1, 2 * tf, 2 * tf, order="F") for k, v in tabs.items()}

    def tab(self, k, m, n, np_):   # MeshFunction call syntax: zero outside the mesh
        if not (inb(m, self.tb) and inf_(n, self.tf) and inf_(np_, self.tf)):
            return 0.0
        return self.t[k][m + self.tb - 1, n + self.tf, np_ + self.tf]

    def bare(self, sp):
        return -self.U if sp == XSP else self.U

    def __call__(self, m, n, np_, ch, sp):
        if n is None or np_ is None:
            return self.bare(sp)
        if sp == DSP:
            return 2 * self(m, n, np_, ch, PSP) + self(m, n, np_, ch, XSP)
        if sp == PSP:
            if ch == PCH:
                return self.tab("Fp_p", m, n, np_) + self.U
            if ch == TCH:
                return self.tab("Ft_p", m, n, np_) + self.U
            return -self.tab("Ft_x", m, np_, n) + self.U
        if ch == PCH:
            return self.tab("Fp_x", m, n, np_) - self.U
        if ch == TCH:
            return self.tab("Ft_x", m, n, np_) - self.U
        return -self.tab("Ft_p", m, np_, n) - self.U


class K1Vertex:
    """Target vertex with only K1 classes: K1[ch] of shape (2 nK1 - 1, NP)."""

    def __init__(self, ref, K1, nK1):
        self.ref, self.nK1 = ref, nK1
        self.K1 = {ch: np.asarray(K1[ch]) for ch in range(3)}
        self.avg = {ch: self.K1[ch].mean(axis=1) for ch in range(3)}

    def k1(self, ch, m, P):
        """same channel at momentum P (P = None: the s-wave average, kSW)"""
        if not inb(m, self.nK1):
            return 0.0
        return self.K1[ch][m + self.nK1 - 1, P] if P is not None else self.avg[ch][m + self.nK1 - 1]

    def psp(self, m, n, w, P, ch, fl):
        v = self.ref(m, n, w, ch, PSP) if fl & F_F0 else 0.0
        own = {PCH: F_GP, TCH: F_GT, ACH: F_GA}[ch]
        if n is None or w is None:
            if fl & own:
                v += self.k1(ch, m, P)
            return v
        if fl & own:
            v += self.k1(ch, m, P)
        if ch == PCH:
            if fl & F_GT:
                v += self.k1(TCH, m - n - w - 1, None)
            if fl & F_GA:
                v += self.k1(ACH, n - w, None)
        elif ch == TCH:
            if fl & F_GP:
                v += self.k1(PCH, m + n + w + 1, None)
            if fl & F_GA:
                v += self.k1(ACH, w - n, None)
        else:
            if fl & F_GP:
                v += self.k1(PCH, m + n + w + 1, None)
            if fl & F_GT:
                v += self.k1(TCH, n - w, None)
        return v

    def __call__(self, m, n, w, P, ch, sp, fl=F_ALL):
        """src/vertex.jl:209-336: xSp by crossing with the γt/γa flags exchanged, dSp = 2 pSp + xSp"""
        if sp == PSP:
            return self.psp(m, n, w, P, ch, fl)
        sw = (fl & (F_F0 | F_GP)) | (F_GA if fl & F_GT else 0) | (F_GT if fl & F_GA else 0)
        if ch == PCH:
            x = -self.psp(m, n, None if w is None else m - w - 1, P, PCH, sw)
        elif ch == TCH:
            x = -self.psp(m, n, w, P, ACH, sw)
        else:
            x = -self.psp(m, n, w, P, TCH, sw)
        if sp == XSP:
            return x
        return 2 * self.psp(m, n, w, P, ch, fl) + x


class LeftVertex:
    """FL: K2 and K3 classes (K1 = 0), evaluated where the BSEs read it, at (Ω, x, νInf): the
    channel evaluator keeps K1 and K2(Ω, x) (src/channel.jl:287-316); the t channel in dSp reads
    2 γt - γa there, since its xSp is -γa (src/vertex.jl:288-336, FL.F0 = RefVertex(T, 0))."""

    def __init__(self, K2, K3, nK2, nK3):
        self.K2 = {ch: np.asarray(K2[ch]) for ch in range(3)}
        self.K3 = {ch: np.asarray(K3[ch]) for ch in range(3)}
        self.nK2, self.nK3 = nK2, nK3

    def k2(self, ch, m, x, P):
        if inb(m, self.nK2[0]) and inf_(x, self.nK2[1]):
            return self.K2[ch][m + self.nK2[0] - 1, x + self.nK2[1], P]
        return 0.0

    def at_inf(self, m, x, P, ch, sp):
        if sp == PSP:
            return self.k2(ch, m, x, P)
        if sp == XSP:
            return -self.k2({PCH: PCH, TCH: ACH, ACH: TCH}[ch], m, None if (ch == PCH and x is None) else x, P)
        return 2 * self.at_inf(m, x, P, ch, PSP) + self.at_inf(m, x, P, ch, XSP)

    def k3(self, ch, m, n, np_, P):
        b, f = self.nK3
        if inb(m, b) and inf_(n, f) and inf_(np_, f):
            return self.K3[ch][m + b - 1, n + f, np_ + f, P]
        return 0.0


class Setup:
    """Meshes and inputs of one tiny instance; `nonlocal_` selects the s-wave lattice forms."""

    def __init__(self, *, nK1, nK2, nK3, nPi, T, NP, nonlocal_, F, F0, FL, Pi, Pi0):
        self.nK1, self.nK2, self.nK3, self.nPi, self.T, self.NP = nK1, nK2, nK3, nPi, T, NP
        self.nonlocal_ = nonlocal_
        self.F, self.F0, self.FL = F, F0, FL
        self.Pi, self.Pi0 = Pi, Pi0     # {ch: array (2 nK1 - 1, 2 nPi, NP)}; pp for p, ph for t and a

    def pi(self, which, ch, m, w, P):
        A = (self.Pi if which == 1 else self.Pi0)[ch]
        return A[m + self.nK1 - 1, w + self.nPi, P]

    def ws(self):
        return range(-self.nPi, self.nPi)

    # ------------------------------------------------------------------ build_K3_cache!
    def caches(self):
        """The ten caches of build_K3_cache! (src/build_K3_cache.jl:20-95; NL
        src/nonlocal/build_K3_cache.jl:18-94), each [Ω, x, y, P]: loop 1 (Γpx, F0p, F0a, F0t) over
        (Ω, ω, νp), loop 2 (Γpp, Γa, Γt, Fp, Fa, Ft) over (Ω, ν, ω)."""
        b, f = self.nK3
        F, F0 = self.F, self.F0.ref
        out = np.zeros((10, 2 * b - 1, 2 * f, 2 * f, self.NP), complex)
        for P in range(self.NP):
            for im, m in enumerate(range(-(b - 1), b)):
                for ia, a in enumerate(range(-f, f)):
                    for ib, c in enumerate(range(-f, f)):
                        out[0, im, ia, ib, P] = F(m, a, c, P, PCH, XSP, F_GT | F_GA)
                        out[1, im, ia, ib, P] = F0(m, a, c, PCH, XSP) - F0(m, a, INF, PCH, XSP)
                        f0a = F0(m, a, c, ACH, PSP) - F0(m, a, INF, ACH, PSP)
                        f0t = F0(m, a, c, TCH, PSP) - F0(m, a, INF, TCH, PSP)
                        out[2, im, ia, ib, P] = f0a
                        out[3, im, ia, ib, P] = 2 * f0t - f0a
                        gpp = F(m, a, c, P, PCH, PSP, F_GT | F_GA)
                        ga = F(m, a, c, P, ACH, PSP, F_GP | F_GT)
                        gt = F(m, a, c, P, TCH, PSP, F_GP | F_GA)
                        fp = F(m, a, c, P, PCH, PSP, F_F0 | F_GP) - F(m, INF, c, P, PCH, PSP, F_F0 | F_GP) + gpp
                        fa = F(m, a, c, P, ACH, PSP, F_F0 | F_GA) - F(m, INF, c, P, ACH, PSP, F_F0 | F_GA) + ga
                        ft = F(m, a, c, P, TCH, PSP, F_F0 | F_GT) - F(m, INF, c, P, TCH, PSP, F_F0 | F_GT) + gt
                        out[4:10, im, ia, ib, P] = [gpp, ga, 2 * gt - ga, fp, fa, 2 * ft - fa]
        return out

    def caches_mfrg(self, old):
        """build_K3_cache_mfRG! (src/nonlocal_2/build_K3_cache.jl:97-164) on top of the caches `old`:
        Γ = F(...; γr = false) - F.F0(...; γr = false) (for a RefVertex F0 the reference vertex cancels),
        F caches = F0(ν) - F0(νInf); the F0 caches are kept."""
        b, f = self.nK3
        F, F0 = self.F, self.F0.ref
        out = old.copy()
        for P in range(self.NP):
            for im, m in enumerate(range(-(b - 1), b)):
                for ia, a in enumerate(range(-f, f)):
                    for ib, c in enumerate(range(-f, f)):
                        out[0, im, ia, ib, P] = F(m, a, c, P, PCH, XSP, F_F0 | F_GT | F_GA) - F0(m, a, c, PCH, XSP)
                        out[4, im, ia, ib, P] = F(m, a, c, P, PCH, PSP, F_F0 | F_GT | F_GA) - F0(m, a, c, PCH, PSP)
                        ga = F(m, a, c, P, ACH, PSP, F_F0 | F_GP | F_GT) - F0(m, a, c, ACH, PSP)
                        gt = F(m, a, c, P, TCH, PSP, F_F0 | F_GP | F_GA) - F0(m, a, c, TCH, PSP)
                        out[5, im, ia, ib, P] = ga
                        out[6, im, ia, ib, P] = 2 * gt - ga
                        d = [F0(m, a, c, ch, PSP) - F0(m, INF, c, ch, PSP) for ch in (PCH, ACH, TCH)]
                        out[7, im, ia, ib, P] = d[0]
                        out[8, im, ia, ib, P] = d[1]
                        out[9, im, ia, ib, P] = 2 * d[2] - d[1]
        return out

    # ------------------------------------------------------------------ BSE_L_K3!, BSE_K3!
    def l_k3(self, C):
        """BSE_L_K3! (src/BSEa/BSEa_K3.jl:1-41) of the three channels with the t post-processing
        (src/BSE_templates.jl:117-145): FL.K3 = T s Σ_ω Γr[Ω, ν, ω] Π0(Ω, ω) F0r[Ω, ω, νp]."""
        b, f = self.nK3
        res = {}
        for ch, ig, i0, s in ((PCH, 4, 1, -1), (ACH, 5, 2, +1), (TCH, 6, 3, -1)):
            pi0 = self.k3pi(0, ch)
            res[ch] = self.T * s * np.einsum("mnwp,mwp,mwvp->mnvp", C[ig], pi0, C[i0])
        res[TCH] = (res[TCH] + res[ACH]) / 2
        return res

    def k3pi(self, which, ch):
        """Π or Π0 on the K3 box: [Ω, ω, P]"""
        b, f = self.nK3
        A = (self.Pi if which == 1 else self.Pi0)[ch]
        return A[self.nK1 - b:self.nK1 + b - 1, self.nPi - f:self.nPi + f, :]

    def k3(self, C, FLK3, is_mfrg=False):
        """BSE_K3! (src/BSEa/BSEa_K3.jl:43-128; NL src/nonlocal/BSEa/BSEa_K3.jl:44-128) of the three
        channels with the t post-processing (src/BSE_templates.jl:151-180).  A = cache_Fr[Ω, ν, ω];
        Γ~ = cache_Γr[Ω, νp, ω] for a and t, cache_Γpx[Ω, ω, νp] for p; L~ = FL.K3(Ω, ω~, νp) with
        ω~ = Ω - ω (p) or ω, zero outside the K3 box; t: 2 FL.t - FL.a."""
        b, f = self.nK3
        T = self.T
        res = {}
        for ch, ig, iF, i0, s1, s2 in ((PCH, 0, 7, 1, -1, +1), (ACH, 5, 8, 2, +1, +1), (TCH, 6, 9, 3, -1, -1)):
            A = C[iF]
            Gt = C[ig] if ch == PCH else np.transpose(C[ig], (0, 2, 1, 3))
            L = FLK3[ch] if ch != TCH else 2 * FLK3[TCH] - FLK3[ACH]
            Lt = np.zeros_like(L)   # [Ω, ω, νp]: L at (Ω, ω~, νp)
            for im, m in enumerate(range(-(b - 1), b)):
                for iw, w in enumerate(range(-f, f)):
                    x = cross(m, w, ch)
                    if inf_(x, f):
                        Lt[im, iw] = L[im, x + f]
            pi, pi0 = self.k3pi(1, ch), self.k3pi(0, ch)
            if is_mfrg:
                R = s1 * pi0[:, :, None, :] * Gt + s2 * pi0[:, :, None, :] * Lt
            else:
                R = s1 * ((pi - pi0)[:, :, None, :] * C[i0] + pi[:, :, None, :] * Gt) + s2 * pi[:, :, None, :] * Lt
            res[ch] = T * np.einsum("mnwp,mwvp->mnvp", A, R) + L
        res[TCH] = (res[TCH] + res[ACH]) / 2
        return res

    # ------------------------------------------------------------------ K1 / K2 classes
    def l_k2(self):
        """BSE_L_K2! (local src/BSEa/BSEa_K2.jl:1-40: Σ over the bubble mesh, Γ(Ω, ν, ω) Π0 F0(Ω, ω~, ∞);
        NL src/nonlocal/BSEa/BSEa_K2.jl:1-41: Σ over the K2 mesh, Γ(Ω, ν, ω~) Π0(ω) F0(Ω, ω, ∞))
        with the t post-processing (src/BSE_templates.jl:40-71)."""
        b2, f2 = self.nK2
        res = {}
        for ch in (PCH, ACH, TCH):
            sp, s = SPIN[ch], SIGN[ch]
            fl = F_ALL & ~F_F0 & ~{PCH: F_GP, TCH: F_GT, ACH: F_GA}[ch]
            out = np.zeros((2 * b2 - 1, 2 * f2, self.NP), complex)
            for P in range(self.NP):
                for im, m in enumerate(range(-(b2 - 1), b2)):
                    for iv, v in enumerate(range(-f2, f2)):
                        val = 0.0
                        if not self.nonlocal_:
                            for w in self.ws():
                                val += (self.F(m, v, w, P, ch, sp, fl) * self.pi(0, ch, m, w, P)
                                        * self.F0.ref(m, cross(m, w, ch), INF, ch, sp))
                        else:
                            for w in range(-f2, f2):
                                val += (self.F(m, v, cross(m, w, ch), P, ch, sp, fl) * self.pi(0, ch, m, w, P)
                                        * self.F0.ref(m, w, INF, ch, sp))
                        out[im, iv, P] = self.T * val * s
            res[ch] = out
        res[TCH] = (res[TCH] + res[ACH]) / 2
        return res

    def k1(self, is_mfrg=False):
        """BSE_K1! (src/BSEa/BSEa_K1.jl:2-54; NL src/nonlocal/BSEa/BSEa_K1.jl:2-57); is_mfRG: the local form
        evaluates in aCh/pSp, the lattice form in the channel's own."""
        N1 = self.nK1
        res = {}
        for ch in (PCH, ACH, TCH):
            sp, s = SPIN[ch], SIGN[ch]
            mch, msp = (ch, sp) if self.nonlocal_ else (ACH, PSP)
            out = np.zeros((2 * N1 - 1, self.NP), complex)
            for P in range(self.NP):
                for im, m in enumerate(range(-(N1 - 1), N1)):
                    val = 0.0
                    for w in self.ws():
                        x = cross(m, w, ch)
                        if is_mfrg:
                            val += (self.F0.ref(m, INF, w, mch, msp) * self.pi(0, ch, m, w, P)
                                    * self.FL.at_inf(m, x, P, mch, msp))
                        else:
                            Fl = self.F(m, INF, w, P, ch, sp)
                            W = ((self.pi(1, ch, m, w, P) - self.pi(0, ch, m, w, P)) * self.F0.ref(m, x, INF, ch, sp)
                                 + self.pi(1, ch, m, w, P) * self.FL.at_inf(m, x, P, ch, sp))
                            val += Fl * W
                    out[im, P] = self.T * val * s
            res[ch] = out
        res[TCH] = (res[TCH] + res[ACH]) / 2
        return res

    def k2(self, is_mfrg=False):
        """BSE_K2! (src/BSEa/BSEa_K2.jl:43-103; NL src/nonlocal/BSEa/BSEa_K2.jl:45-104) plus FL.K2 (t:
        2 FL.t - FL.a) and the t post-processing; is_mfRG: (F0(ν) - F0(∞)) Π0 FL(∞) with the crossing on
        the right vertex (local) or on the left one (lattice)."""
        b2, f2 = self.nK2
        res = {}
        for ch in (PCH, ACH, TCH):
            sp, s = SPIN[ch], SIGN[ch]
            out = np.zeros((2 * b2 - 1, 2 * f2, self.NP), complex)
            for P in range(self.NP):
                for im, m in enumerate(range(-(b2 - 1), b2)):
                    for iv, v in enumerate(range(-f2, f2)):
                        val = 0.0
                        for w in self.ws():
                            x = cross(m, w, ch)
                            if is_mfrg:
                                wl, wr = (x, w) if self.nonlocal_ else (w, x)
                                d0 = self.F0.ref(m, v, wl, ch, sp) - self.F0.ref(m, INF, wl, ch, sp)
                                val += d0 * self.pi(0, ch, m, w, P) * self.FL.at_inf(m, wr, P, ch, sp)
                            else:
                                Fl = self.F(m, v, w, P, ch, sp) - self.F(m, INF, w, P, ch, sp)
                                W = ((self.pi(1, ch, m, w, P) - self.pi(0, ch, m, w, P)) * self.F0.ref(m, x, INF, ch, sp)
                                     + self.pi(1, ch, m, w, P) * self.FL.at_inf(m, x, P, ch, sp))
                                val += Fl * W
                        out[im, iv, P] = self.T * val * s
            fl = self.FL.K2[ch] if ch != TCH else 2 * self.FL.K2[TCH] - self.FL.K2[ACH]
            res[ch] = out + fl
        res[TCH] = (res[TCH] + res[ACH]) / 2
        return res

    def k1_new(self, is_mfrg=False):
        """BSE_K1_new! (src/BSEa/BSEa_K1.jl:57-100; NL src/nonlocal/BSEa/BSEa_K1.jl:62-108):
        T s Σ_ω (F(Ω,∞,ω) Π - F0(Ω,∞,ω) Π0) U; lattice is_mfRG: (F - F0) Π U."""
        N1 = self.nK1
        res = {}
        mf = is_mfrg and self.nonlocal_
        for ch in (PCH, ACH, TCH):
            sp, s = SPIN[ch], SIGN[ch]
            U = self.F0.ref.bare(sp)
            out = np.zeros((2 * N1 - 1, self.NP), complex)
            for P in range(self.NP):
                for im, m in enumerate(range(-(N1 - 1), N1)):
                    val = 0.0
                    for w in self.ws():
                        Fl = self.F(m, INF, w, P, ch, sp)
                        F0l = self.F0.ref(m, INF, w, ch, sp)
                        pi, pi0 = self.pi(1, ch, m, w, P), self.pi(0, ch, m, w, P)
                        val += ((Fl - F0l) * pi if mf else Fl * pi - F0l * pi0) * U
                    out[im, P] = self.T * val * s
            res[ch] = out
        res[TCH] = (res[TCH] + res[ACH]) / 2
        return res

    def k2_new(self, is_mfrg=False):
        """BSE_K2_new! (src/BSEa/BSEa_K2.jl:108-151; NL src/nonlocal/BSEa/BSEa_K2.jl:108-155)."""
        b2, f2 = self.nK2
        res = {}
        mf = is_mfrg and self.nonlocal_
        for ch in (PCH, ACH, TCH):
            sp, s = SPIN[ch], SIGN[ch]
            U = self.F0.ref.bare(sp)
            out = np.zeros((2 * b2 - 1, 2 * f2, self.NP), complex)
            for P in range(self.NP):
                for im, m in enumerate(range(-(b2 - 1), b2)):
                    for iv, v in enumerate(range(-f2, f2)):
                        val = 0.0
                        for w in self.ws():
                            Fl = self.F(m, v, w, P, ch, sp) - self.F(m, INF, w, P, ch, sp)
                            F0l = self.F0.ref(m, v, w, ch, sp) - self.F0.ref(m, INF, w, ch, sp)
                            pi, pi0 = self.pi(1, ch, m, w, P), self.pi(0, ch, m, w, P)
                            val += ((Fl - F0l) * pi if mf else Fl * pi - F0l * pi0) * U
                        out[im, iv, P] = self.T * val * s
            res[ch] = out
        res[TCH] = (res[TCH] + res[ACH]) / 2
        return res


class F0Wrap:
    def __init__(self, ref):
        self.ref = ref


# ---------------------------------------------------------------------------------- lattice SDE
def fold(L, x):
    return x % L


def sigma_momentum_space(G, Lpp, Lph, nG, nK2, T, L):
    """Σ(ν, k) = T / L² Σ_P Σ_Ω [G(Ω - ν, P - k) Lpp(Ω, ν, P) + G(Ω + ν, P + k) Lph(Ω, ν, P)], the momentum-space
    evaluation of SDE_compute_inner! (src/nonlocal/SDE.jl:282-327) for a vertex mesh equal to the G mesh
    (momentum index j1 + L j2, first coordinate fastest).  G [2 nG, L²], Lpp/Lph [2 nK2b - 1, 2 nK2f, L²]."""
    b2, f2 = nK2
    S = np.zeros((2 * nG, L * L), complex)
    for j2 in range(L):
        for j1 in range(L):
            k = j1 + L * j2
            for v in range(-nG, nG):
                if not inf_(v, f2):
                    continue
                val = 0.0
                for p2 in range(L):
                    for p1 in range(L):
                        P = p1 + L * p2
                        pmk = fold(L, p1 - j1) + L * fold(L, p2 - j2)
                        ppk = fold(L, p1 + j1) + L * fold(L, p2 + j2)
                        for m in range(-(b2 - 1), b2):
                            a, c = m - v - 1, m + v
                            if inf_(a, nG):
                                val += G[a + nG, pmk] * Lpp[m + b2 - 1, v + f2, P]
                            if inf_(c, nG):
                                val += G[c + nG, ppk] * Lph[m + b2 - 1, v + f2, P]
                S[v + nG, k] = T * val / (L * L)
    return S


def sigma_u2_momentum_space(G, nG, T, U, L):
    """The second-order term of SDE_U2_using_G (src/nonlocal/SDE.jl:400-447) written in momentum space:
    Σ_U²(ν, k) = U² T² / L⁴ Σ_{ω1, ω2, k1, k2} G(ω1, k1) G(ω2, k2) G(ω1 - ω2 + ν, k + k1 - k2)
    (the real-space product G(ω1, -R) G(ω2, R) G(ω1 - ω2 + ν, R) by the convolution theorem; a local G
    is L = 1)."""
    N = L * L
    S = np.zeros((2 * nG, N), complex)
    ks = [(j1, j2) for j2 in range(L) for j1 in range(L)]
    for ik, (a1, a2) in enumerate(ks):
        for v in range(-nG, nG):
            val = 0.0
            for i1, (b1, b2) in enumerate(ks):
                for i2, (c1, c2) in enumerate(ks):
                    k3 = fold(L, a1 + b1 - c1) + L * fold(L, a2 + b2 - c2)
                    for w1 in range(-nG, nG):
                        g1 = G[w1 + nG, i1]
                        for w2 in range(-nG, nG):
                            x = w1 - w2 + v
                            if inf_(x, nG):
                                val += g1 * G[w2 + nG, i2] * G[x + nG, k3]
            S[v + nG, ik] = val * U * U * T * T / (N * N)
    return S


def channel_eval(K1, K2, K3, nK1, nK2, nK3, m, n, np_):
    """Channel evaluator γ(Ω, ν, νp) for finite arguments (src/channel.jl:220-253); K* of one momentum."""
    v = 0.0
    if not inb(m, nK1):
        return v
    v += K1[m + nK1 - 1]
    if not inb(m, nK2[0]):
        return v
    i1, i2 = inf_(n, nK2[1]), inf_(np_, nK2[1])
    if i1 and i2:
        v += K2[m + nK2[0] - 1, n + nK2[1]] + K2[m + nK2[0] - 1, np_ + nK2[1]]
        if inb(m, nK3[0]) and inf_(n, nK3[1]) and inf_(np_, nK3[1]):
            v += K3[m + nK3[0] - 1, n + nK3[1], np_ + nK3[1]]
    elif i1:
        v += K2[m + nK2[0] - 1, n + nK2[1]]
    elif i2:
        v += K2[m + nK2[0] - 1, np_ + nK2[1]]
    return v


def sde_L(F, U, Pipp, Piph, nK1, nK2, nK3, nPi, T, NP):
    """The L-vertices of a Vertex level of the lattice SDE (SDE_channel_L_pp!/_ph!,
    src/nonlocal/SDE.jl:3-75): Lpp(Ω, ν, P) = T Σ_ω U Π_pp(Ω, ω, P) γp(Ω, Ω - ω, ν, P),
    Lph(Ω, ν, P) = T Σ_ω U Π_ph(Ω, ω, P) (γt + γa)(Ω, ν, ω, P) over the bubble mesh.
    F: {ch: (K1 [NΩ1, NP], K2 [NΩ2, Nν2, NP], K3 [NΩ3, Nν3, Nν3, NP])}."""
    b2, f2 = nK2
    Lpp = np.zeros((2 * b2 - 1, 2 * f2, NP), complex)
    Lph = np.zeros_like(Lpp)
    for P in range(NP):
        g = {ch: (F[ch][0][:, P], F[ch][1][:, :, P], F[ch][2][..., P]) for ch in range(3)}
        for m in range(-(b2 - 1), b2):
            for v in range(-f2, f2):
                app = aph = 0.0
                for w in range(-nPi, nPi):
                    ip = (m + nK1 - 1, w + nPi, P)
                    app += Pipp[ip] * channel_eval(*g[PCH], nK1, nK2, nK3, m, m - w - 1, v)
                    aph += Piph[ip] * (channel_eval(*g[TCH], nK1, nK2, nK3, m, v, w)
                                       + channel_eval(*g[ACH], nK1, nK2, nK3, m, v, w))
                Lpp[m + b2 - 1, v + f2, P] = T * U * app
                Lph[m + b2 - 1, v + f2, P] = T * U * aph
    return Lpp, Lph
