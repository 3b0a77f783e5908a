"""NumPy restatement of the reference's per-pair loop structure, for timing.

TEST INFRASTRUCTURE ONLY: the CPU baseline of the benchmark and its test.
It is never imported by the product package.

The reference evaluates one (term, bra, ket) element per call with small
NumPy operations, lifting every particle-space operator X to the 3n x 3n
kron(X, I3) that acts on the particle-major shift vectors
(python/matel.py:111-115), and keeps the vech(L) gradient algebra in n x n
form (python/gradel.py:114-148, :187-261).  This module does the same: the
same kinds and sizes of NumPy calls per element (matel, python/matel.py:
28-210) and per bra gradient (gradel, python/gradel.py:27-392), the ordered
pair loop of make_grad (python/make_grad.py:37-55) and its GRADA triple loop
(python/make_grad.py:68-79), so that its time per call stands in for the
reference's on a machine where the reference itself is not present.  Its
formulas are the C oracle's (necg_oracle.c), which the golden vectors pin;
``tests/test_kron_ref.py`` checks this module against the oracle.

Indices are 0-based.
"""

from __future__ import annotations

import math
import time

import numpy as np

# python/shared.py:71 (math.pi) and :76-86 (Pade constants)
_PI = math.pi
_C = (0.564691197e-04, 0.758433197e-03, 0.769838037e-02, 0.629344460e-01, 0.213271302e+00,
      0.720266520e-04, 0.955528842e-03, 0.101431553e-01, 0.738522953e-01, 0.338450368e+00,
      0.879937801e+00)


def f0(z):
    """Boys F0, Pade form (python/utils.py:78-90)."""
    if z > 16.3578:
        return math.sqrt(_PI / (4.0 * z))
    c = _C
    p = ((((c[0] * z + c[1]) * z + c[2]) * z + c[3]) * z + c[4]) * z + 1.0
    q = (((((c[5] * z + c[6]) * z + c[7]) * z + c[8]) * z + c[9]) * z + c[10]) * z + 1.0
    return math.sqrt(p / q)


def df0(z):
    """dF0/dz in power form (python/utils.py:93-106)."""
    if z > 16.3578:
        return -math.sqrt(_PI / (4.0 * z)) / (2.0 * z)
    c = _C
    p = c[0] * z ** 5 + c[1] * z ** 4 + c[2] * z ** 3 + c[3] * z ** 2 + c[4] * z + 1.0
    q = (c[5] * z ** 6 + c[6] * z ** 5 + c[7] * z ** 4 + c[8] * z ** 3 + c[9] * z ** 2
         + c[10] * z + 1.0)
    dp = 5 * c[0] * z ** 4 + 4 * c[1] * z ** 3 + 3 * c[2] * z ** 2 + 2 * c[3] * z + c[4]
    dq = (6 * c[5] * z ** 5 + 5 * c[6] * z ** 4 + 4 * c[7] * z ** 3 + 3 * c[8] * z ** 2
          + 2 * c[9] * z + c[10])
    return (q * dp - p * dq) / (2.0 * math.sqrt(p / q) * q * q)


def vech(X):
    """Column-major lower triangle (python/utils.py:42-55)."""
    n = X.shape[0]
    return np.concatenate([X[j:, j] for j in range(n)])


class KronRef:
    """Per-call evaluation of one system in the reference's operator form."""

    def __init__(self, sysd):
        self.s = sysd
        n = sysd.n
        self.n, self.nst = n, sysd.nst
        self.svh, self.skp = n * (n + 1) // 2, 3 * n
        self.blk = self.svh + self.skp
        self.I3 = np.eye(3)
        self.mass = np.asarray(sysd.mass, dtype=np.float64)
        self.massb = np.kron(self.mass, self.I3)
        self.ncv = np.asarray(sysd.ncv, dtype=np.float64)
        self.scvv = np.asarray(sysd.scvv, dtype=np.float64)
        self.sml = np.asarray(sysd.sml, dtype=np.float64)    # (nst, n, n)
        self.sms = np.asarray(sysd.sms, dtype=np.float64)    # (nst, 3n, 3n)
        self.il = np.tril_indices(n)

    def _L(self, iv, k):
        n, b = self.n, k * self.blk
        L = np.zeros((n, n))
        v = iv[b:b + self.svh]
        t = 0
        for j in range(n):            # column-major lower triangle
            for i in range(j, n):
                L[i, j] = v[t]
                t += 1
        return L

    def _setup(self, iv, stf, ii, jj):
        """python/matel.py:56-136 (and gradel.py:46-102)."""
        n, svh, blk = self.n, self.svh, self.blk
        lk = self._L(iv, ii)
        ell = self._L(iv, jj)
        ak = lk @ lk.T
        al = ell @ ell.T
        P = self.sml[stf]
        al = P.T @ (al @ P)
        st1 = iv[ii * blk + svh: ii * blk + blk].copy()
        st2 = self.sms[stf] @ iv[jj * blk + svh: jj * blk + blk]
        akl = ak + al
        akli = np.linalg.inv(akl)
        dett = abs(np.linalg.det(akl))
        detak = math.sqrt(abs(np.prod(np.diag(lk)) ** 2))
        detal = math.sqrt(abs(np.prod(np.diag(ell)) ** 2))
        akb = np.kron(ak, self.I3)
        alb = np.kron(al, self.I3)
        aklib = np.kron(akli, self.I3)
        iv1 = akb @ st1
        i1 = float(st1 @ iv1)
        iv2 = alb @ st2
        i2 = float(st2 @ iv2)
        ee = iv1 + iv2
        st = aklib @ ee
        i3 = float(ee @ st)
        ov = math.exp(-(i1 + i2) + i3) * math.sqrt(8.0 ** n) * math.sqrt(
            (detak * detal / dett) ** 3)
        return dict(lk=lk, ak=ak, al=al, akli=akli, akb=akb, alb=alb, aklib=aklib, st1=st1,
                    st2=st2, ee=ee, st=st, ov=ov)

    def _kinetic(self, w):
        dif1 = w["st"] - w["st1"]
        dif2 = w["st"] - w["st2"]
        im2 = w["akb"] @ (self.massb @ w["alb"])
        ke_i3 = float(dif1 @ (im2 @ dif2))
        im1 = self.massb @ (w["akb"] @ (w["aklib"] @ w["alb"]))
        ke_i2 = float(np.trace(im1))
        return ke_i3, ke_i2, dif1, dif2

    def _jij(self, p, q):
        n = self.n
        J = np.zeros((n, n))
        if p == 0:
            J[q - 1, q - 1] = 1.0
        else:
            J[p - 1, p - 1] = J[q - 1, q - 1] = 1.0
            J[p - 1, q - 1] = J[q - 1, p - 1] = -1.0
        return J

    def matel(self, iv, stf, ii, jj):
        """(ov, tk, ert, zht) of term stf, bra ii, ket jj (python/matel.py:28)."""
        w = self._setup(iv, stf, ii, jj)
        ov = w["ov"]
        ke_i3, ke_i2, _, _ = self._kinetic(w)
        tk = 2.0 * ov * (2.0 * ke_i3 + ke_i2)
        ert = 0.0
        n = self.n
        for p in range(n + 1):
            for q in range(p + 1, n + 1):
                ism1 = self._jij(p, q) @ w["akli"]
                jnb = np.kron(w["akli"] @ ism1, self.I3)
                i3er = float(w["ee"] @ (jnb @ w["ee"]))
                i2er = float(np.trace(ism1))
                if i2er == 0.0:
                    continue
                ert += self.ncv[p] * self.ncv[q] * 2.0 * ov * f0(i3er / i2er) / math.sqrt(
                    i2er * _PI)
        zs = sum(self.ncv[i] * w["st"][3 * i - 1] for i in range(1, n + 1))
        return ov, tk, ert, ov * zs * self.s.efs

    def gradel(self, iv, stf, ii, jj):
        """(tgvl, tgvls): bra derivatives w.r.t. vech(L_ii) and s_ii
        (python/gradel.py:27-392)."""
        n, svh, skp = self.n, self.svh, self.skp
        w = self._setup(iv, stf, ii, jj)
        ov, lk, ak, al, akli = w["ov"], w["lk"], w["ak"], w["al"], w["akli"]
        st, st1, st2, ee = w["st"], w["st1"], w["st2"], w["ee"]
        aki = np.linalg.inv(ak)
        ke_i3, ke_i2, dif1, dif2 = self._kinetic(w)
        tfac = 2.0 * ke_i3 + ke_i2
        # overlap gradients (python/gradel.py:108-148)
        ogs = 2.0 * ov * (w["akb"] @ dif1)
        S1 = st1.reshape(n, 3)
        E = ee.reshape(n, 3)
        St = st.reshape(n, 3)
        S2 = st2.reshape(n, 3)
        ogl = 1.5 * vech(aki @ lk) - 3.0 * vech(akli @ lk)
        ogl -= 2.0 * vech((S1 @ S1.T) @ lk)
        ogl -= 2.0 * vech(akli @ ((E @ E.T) @ akli) @ lk)
        m2 = akli @ (E @ S1.T)
        ogl += 2.0 * vech(m2 @ lk) + 2.0 * vech(m2.T @ lk)
        ogl *= ov
        # kinetic gradients (python/gradel.py:153-261)
        al_mass = self.mass @ al
        aklp = w["akb"] @ (self.massb @ w["alb"])
        aklpt = aklp.T
        akliak = w["aklib"] @ w["akb"]
        iv2k = (aklp @ akliak).T @ dif1
        iv1k = (aklpt @ akliak - aklpt).T @ dif2
        kgs = 2.0 * tfac * ogs + 4.0 * ov * iv2k + 4.0 * ov * iv1k
        kgl = tfac * ogl
        a1 = al_mass @ (akli @ lk)
        kgl -= ov * 6.0 * vech(akli @ (ak @ a1))
        kgl += ov * 6.0 * vech(a1)
        sds, sdsk, skdsk = St @ St.T, St @ S1.T, S1 @ S1.T
        sldsk, slds = S2 @ S1.T, S2 @ St.T
        ap = ak @ al_mass
        apt = ap.T
        p1 = akli @ (ap + apt)
        pp = akli @ ap
        pt = akli @ apt
        kcm = ((al_mass - p1) @ sds + (p1 - al_mass) @ sdsk + pt @ sdsk.T
               + (pp - al_mass) @ slds + (al_mass - pp) @ sldsk - pt @ skdsk)
        kgl += 2.0 * ov * (vech(kcm @ lk) + vech(kcm.T @ lk))
        kgl *= 2.0
        # Coulomb gradients (python/gradel.py:266-337)
        ergst = np.zeros(skp)
        erglt = np.zeros(svh)
        EEt = E @ E.T
        S1Et = S1 @ E.T
        for p in range(n + 1):
            for q in range(p + 1, n + 1):
                J = self._jij(p, q)
                ism1 = J @ akli
                jn = akli @ ism1
                jnb = np.kron(jn, self.I3)
                i3er = float(ee @ (jnb @ ee))
                i2er = float(np.trace(ism1))
                if i2er == 0.0:
                    continue
                z = i3er / i2er
                f, fp = f0(z), df0(z)
                qq = self.ncv[p] * self.ncv[q]
                iv3 = (w["akb"] @ jnb) @ ee
                ergst += (qq * 2.0 * f / math.sqrt(i2er * _PI)) * ogs + (
                    qq * 4.0 * ov * fp / math.sqrt(_PI * i2er ** 3)) * iv3
                jnl = jn @ lk
                ergl = (2.0 * f / math.sqrt(i2er * _PI)) * ogl
                ergl += (2.0 * ov * f / math.sqrt(_PI * i2er ** 3)) * vech(jnl)
                ergl += (4.0 * ov * fp * i3er / math.sqrt(_PI * i2er ** 5)) * vech(jnl)
                m3 = 4.0 * ov * fp / math.sqrt(_PI * i2er ** 3)
                b1 = (akli @ EEt) @ jn
                ergl -= m3 * (vech(b1 @ lk) + vech(b1.T @ lk))
                b2 = S1Et @ jn
                ergl += m3 * (vech(b2 @ lk) + vech(b2.T @ lk))
                erglt += qq * ergl
        # Stark gradients (python/gradel.py:342-370)
        efs = self.s.efs
        akakli = np.kron(ak @ akli, self.I3)
        zgs = np.zeros(skp)
        zgl = np.zeros(svh)
        for i in range(1, n + 1):
            zgs += self.ncv[i] * (ogs * st[3 * i - 1] + ov * akakli[:, 3 * i - 1]) * efs
            Z1 = np.zeros((n, n))
            Z1[:, i - 1] = S1[:, 2] - St[:, 2]
            v1 = vech((Z1 @ akli) @ lk)
            v2 = vech((akli @ Z1.T) @ lk)
            zgl += self.ncv[i] * efs * (st[3 * i - 1] * ogl + ov * (v1 + v2))
        tgvl = np.concatenate([kgl + erglt - zgl, kgs + ergst - zgs])
        tgvls = np.concatenate([ogl, ogs])
        if ii == jj:
            tgvl *= 2.0
            tgvls *= 2.0
        return tgvl, tgvls

    # -- the loops of outprod / make_grad --------------------------------
    def element_rows(self, iv, K, rows):
        """H and S entries of the given bra rows, lower triangle
        (python/outprod.py:49-54): {(i, j): (h, s)}."""
        out = {}
        for i in rows:
            for j in range(i + 1):
                h = s = 0.0
                for stf in range(self.nst):
                    ov, tk, ert, zht = self.matel(iv, stf, i, j)
                    h += self.scvv[stf] * (tk + ert - zht)
                    s += self.scvv[stf] * ov
                out[(i, j)] = (h, s)
        return out

    def gradient_rows(self, x, K, E, norm, rows):
        """GRADA of the given bra rows (python/make_grad.py:37-79): the DH/DS
        row block from gradel over every ket, then the triple loop."""
        blk = self.blk
        iv, C = x[:K * blk], x[K * blk:]
        out = {}
        for l in rows:
            dh = np.zeros((blk, K))
            ds = np.zeros((blk, K))
            for k in range(K):
                for stf in range(self.nst):
                    a, b = self.gradel(iv, stf, l, k)
                    dh[:, k] += self.scvv[stf] * a
                    ds[:, k] += self.scvv[stf] * b
            g = np.zeros(blk)
            for r in range(blk):           # the reference's interpreted loop
                v = 0.0
                for i in range(K):
                    gm = norm * (dh[r, i] - E * ds[r, i])
                    v += (2.0 if i != l else 1.0) * C[i] * C[l] * gm
                g[r] = v
            out[l] = g
        return out


def _work(args):
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1):     # one core per process, as the reference runs
        return _work1(args)


def _work1(args):
    sysd, x, K, E, norm, grows, erows, kets = args
    kr = KronRef(sysd)
    iv = x[:K * kr.blk]
    t0 = time.perf_counter()
    # gradient rows restricted to `kets` kets each (a bounded sample)
    for l in grows:
        for k in range(kets):
            for stf in range(kr.nst):
                kr.gradel(iv, stf, l, k)
    t1 = time.perf_counter()
    for i in erows:
        for j in range(min(i + 1, kets)):
            for stf in range(kr.nst):
                kr.matel(iv, stf, i, j)
    t2 = time.perf_counter()
    return t1 - t0, t2 - t1


def time_sample(sysd, x, K, E, norm, procs=1, rows_per_proc=2, kets=64):
    """Times a bounded sample of the reference-structured loops at K:
    ``rows_per_proc`` gradient rows and element rows per process, each over
    ``kets`` kets and every term, on ``procs`` processes (rows dealt to the
    processes rank-cyclically, as the Fortran code deals its pair rows), plus
    the GRADA triple loop of one row.  Returns the wall time of the sample and
    the extrapolation to one full evaluation (K^2 gradient pair blocks,
    K(K+1)/2 element pairs, K GRADA rows)."""
    import multiprocessing as mp
    kr = KronRef(sysd)
    rows = [(K - 1 - 7 * t) % K for t in range(procs * rows_per_proc)]
    chunks = [rows[r::procs] for r in range(procs)]
    args = [(sysd, x, K, E, norm, c, c, kets) for c in chunks]
    t0 = time.perf_counter()
    if procs == 1:
        res = [_work(args[0])]
    else:
        # fresh interpreters: the workers never touch the GPU the parent holds
        with mp.get_context("spawn").Pool(procs) as pool:
            res = pool.map(_work, args)
    wall = time.perf_counter() - t0
    tg = sum(r[0] for r in res)           # process-seconds
    te = sum(r[1] for r in res)
    ng = len(rows) * kets                 # ordered (bra, ket) blocks sampled
    ne = sum(min(i + 1, kets) for i in rows)
    # GRADA triple loop of one row (interpreted, K * blk iterations)
    dh = np.zeros((kr.blk, K))
    C = x[K * kr.blk:]
    t3 = time.perf_counter()
    for r in range(kr.blk):
        v = 0.0
        for i in range(K):
            v += (2.0 if i != 0 else 1.0) * C[i] * C[0] * (norm * (dh[r, i] - E * dh[r, i]))
    t_grada = time.perf_counter() - t3
    per_g = tg / ng                       # process-seconds per ordered pair (all terms)
    per_e = te / ne
    full_proc_s = per_g * K * K + per_e * K * (K + 1) / 2 + t_grada * K
    return {"wall_s": wall, "process_s": tg + te, "gradient_blocks": ng, "element_pairs": ne,
            "s_per_gradient_block": per_g, "s_per_element_pair": per_e,
            "grada_row_s": t_grada, "est_seconds_per_eval": full_proc_s / procs}
