/*
 * necg_oracle.c -- CPU restatement of the reference energy path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against; it is never linked into, loaded by or called from the
 * product library.  Only tests, the smoke check and the CPU-baseline leg of
 * the benchmark load it.
 *
 * It restates MauricioCafiero/NBO_CorrelatedGaussians python/ (the parity
 * oracle of that project) term by term:
 *   f0 / df0          python/utils.py:78-106
 *   vech              python/utils.py:42-55
 *   matel             python/matel.py:28-210
 *   gradel            python/gradel.py:27-392
 *   outprod           python/outprod.py:20-76
 *   make_grad         python/make_grad.py:18-90
 *   rrms, dipole      python/rrms.py:21-120, python/necg.py:200-232
 * in "n-space": every 3n x 3n operator kron(X, I3) of the reference acts on a
 * particle-major 3n vector exactly as the n x n matrix X acts on the n x 3
 * matrix holding that vector (row i = xyz of pseudo-particle i), so X is used
 * directly.  The shift transform sms is applied as the full 3n x 3n matrix the
 * input provides (python/matel.py:87).  Inverses and determinants come from
 * Gaussian elimination with partial pivoting, the same LU route as
 * numpy.linalg.inv / det (python/matel.py:93, :99).
 *
 * Pinning: the golden vectors written from the reference (outprod, matel,
 * gradel, rrms and the dipole loop), central finite differences of
 * matel against gradel, the hydrogen-atom closed form, the Gaussian product
 * theorem for overlaps, and F0(0) = 1 / erf form of the Boys function; rrms
 * and the dipole by the Gaussian product theorem (one particle: overlap and
 * centre in closed form), a single basis function (the centre is the shift),
 * the overlap of matel for an identity bra term, and the covariance of the
 * centres under one common translation of every shift; see
 * tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NMAX 8
#define SKMAX (3 * NMAX)

typedef struct {
  int n, nst;
  double efs;
  const double *mass; /* n*n row-major */
  const double *ncv;  /* n+1 */
  const double *scvv; /* nst, already divided by nst */
  const double *sml;  /* nst*n*n, sml[p*n*n + i*n + j] = reference sml[i,j,p] */
  const double *sms;  /* nst*(3n)*(3n) */
} osys;

/* python/shared.py:71 (math.pi) and :76-86 (Pade constants). */
static const double PI_ = 3.14159265358979311599796346854;
static const double C1 = 0.564691197e-04, C2 = 0.758433197e-03,
                    C3 = 0.769838037e-02, C4 = 0.629344460e-01,
                    C5 = 0.213271302e+00, C6 = 0.720266520e-04,
                    C7 = 0.955528842e-03, C8 = 0.101431553e-01,
                    C9 = 0.738522953e-01, C10 = 0.338450368e+00,
                    C11 = 0.879937801e+00;

/* python/utils.py:78-90 */
double necg_oracle_f0(double z) {
  if (z <= 16.3578) {
    double num = (((((C1 * z + C2) * z + C3) * z + C4) * z + C5) * z + 1.0);
    double den = ((((((C6 * z + C7) * z + C8) * z + C9) * z + C10) * z + C11) * z + 1.0);
    return sqrt(num / den);
  }
  double r = 2.0 * z;
  return sqrt(PI_ / (2.0 * r));
}

/* python/utils.py:93-106 (quotient rule with power-form polynomials) */
double necg_oracle_df0(double z) {
  if (z <= 16.3578) {
    double num = (((((C1 * z + C2) * z + C3) * z + C4) * z + C5) * z + 1.0);
    double den = ((((((C6 * z + C7) * z + C8) * z + C9) * z + C10) * z + C11) * z + 1.0);
    double f = sqrt(num / den);
    double z2 = z * z, z3 = z2 * z, z4 = z3 * z, z5 = z4 * z, z6 = z5 * z;
    double anum = C1 * z5 + C2 * z4 + C3 * z3 + C4 * z2 + C5 * z + 1.0;
    double aden = C6 * z6 + C7 * z5 + C8 * z4 + C9 * z3 + C10 * z2 + C11 * z + 1.0;
    double danum = 5.0 * C1 * z4 + 4.0 * C2 * z3 + 3.0 * C3 * z2 + 2.0 * C4 * z + C5;
    double daden = 6.0 * C6 * z5 + 5.0 * C7 * z4 + 4.0 * C8 * z3 + 3.0 * C9 * z2 +
                   2.0 * C10 * z + C11;
    return (1.0 / (2.0 * f)) * ((aden * danum - anum * daden) / (aden * aden));
  }
  return -sqrt(PI_ / (4.0 * z)) / (2.0 * z);
}

/* ---------------- small dense helpers (row-major n x m) ---------------- */

static void mm(int n, int k, int m, const double *A, const double *B, double *C) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) {
      double s = 0.0;
      for (int t = 0; t < k; ++t) s += A[i * k + t] * B[t * m + j];
      C[i * m + j] = s;
    }
}
static void tr_(int n, int m, const double *A, double *At) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) At[j * n + i] = A[i * m + j];
}
/* python/utils.py:42-55: column-major lower triangle */
static void vech_(int n, const double *X, double *v) {
  int k = 0;
  for (int j = 0; j < n; ++j)
    for (int i = j; i < n; ++i) v[k++] = X[i * n + j];
}
static void vech_axpy(int n, double a, const double *X, double *v) {
  int k = 0;
  for (int j = 0; j < n; ++j)
    for (int i = j; i < n; ++i) v[k++] += a * X[i * n + j];
}
static double dot(int len, const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < len; ++i) s += a[i] * b[i];
  return s;
}
static double trace(int n, const double *A) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += A[i * n + i];
  return s;
}

/* Gauss-Jordan with partial pivoting; returns 0 if singular. */
static int inv_det(int n, const double *A, double *Ai, double *det) {
  double a[NMAX * NMAX], b[NMAX * NMAX];
  memcpy(a, A, sizeof(double) * n * n);
  for (int i = 0; i < n * n; ++i) b[i] = 0.0;
  for (int i = 0; i < n; ++i) b[i * n + i] = 1.0;
  double d = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (fabs(a[r * n + c]) > fabs(a[p * n + c])) p = r;
    if (a[p * n + c] == 0.0) return 0;
    if (p != c) {
      d = -d;
      for (int j = 0; j < n; ++j) {
        double t = a[c * n + j]; a[c * n + j] = a[p * n + j]; a[p * n + j] = t;
        t = b[c * n + j]; b[c * n + j] = b[p * n + j]; b[p * n + j] = t;
      }
    }
    double piv = a[c * n + c];
    d *= piv;
    for (int j = 0; j < n; ++j) { a[c * n + j] /= piv; b[c * n + j] /= piv; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      double f = a[r * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) { a[r * n + j] -= f * a[c * n + j]; b[r * n + j] -= f * b[c * n + j]; }
    }
  }
  memcpy(Ai, b, sizeof(double) * n * n);
  *det = d;
  return 1;
}

/* Shared set-up of matel/gradel for bra ii, ket jj, term stf (0-based):
 * python/matel.py:56-136 and python/gradel.py:46-102. */
typedef struct {
  int n;
  double lk[NMAX * NMAX], ell[NMAX * NMAX], ak[NMAX * NMAX], al[NMAX * NMAX];
  double akli[NMAX * NMAX], st1[SKMAX], st2[SKMAX], iv1[SKMAX], iv2[SKMAX];
  double ee[SKMAX], st[SKMAX];
  double i1, i2, i3, ov, dett, detak, detal;
} pairwork;

static int pair_setup(const osys *s, const double *iv, int stf, int ii, int jj, pairwork *w) {
  const int n = s->n, svh = n * (n + 1) / 2, skp = 3 * n, blk = svh + skp;
  w->n = n;
  memset(w->lk, 0, sizeof(w->lk));
  memset(w->ell, 0, sizeof(w->ell));
  int kk = ii * blk, ll = jj * blk;
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      w->lk[j * n + i] = iv[kk++];
      w->ell[j * n + i] = iv[ll++];
    }
  double lkt[NMAX * NMAX] = {0}, ellt[NMAX * NMAX] = {0}, tmp[NMAX * NMAX], smlt[NMAX * NMAX];
  tr_(n, n, w->lk, lkt);
  tr_(n, n, w->ell, ellt);
  mm(n, n, n, w->lk, lkt, w->ak);
  double al0[NMAX * NMAX];
  mm(n, n, n, w->ell, ellt, al0);
  const double *sm = s->sml + (size_t)stf * n * n;
  mm(n, n, n, al0, sm, tmp);        /* al @ sml            */
  tr_(n, n, sm, smlt);
  mm(n, n, n, smlt, tmp, w->al);    /* sml^T @ (al @ sml)  */
  memcpy(w->st1, iv + ii * blk + svh, sizeof(double) * skp);
  const double *sv = s->sms + (size_t)stf * skp * skp;
  const double *s2 = iv + jj * blk + svh;
  for (int r = 0; r < skp; ++r) {
    double a = 0.0;
    for (int c = 0; c < skp; ++c) a += sv[r * skp + c] * s2[c];
    w->st2[r] = a;
  }
  double akl[NMAX * NMAX];
  for (int i = 0; i < n * n; ++i) akl[i] = w->ak[i] + w->al[i];
  double d;
  if (!inv_det(n, akl, w->akli, &d)) return 0;
  w->dett = fabs(d);
  double pk = 1.0, pl = 1.0;
  for (int i = 0; i < n; ++i) { pk *= w->lk[i * n + i]; pl *= w->ell[i * n + i]; }
  w->detak = sqrt(fabs(pk * pk));
  w->detal = sqrt(fabs(pl * pl));
  mm(n, n, 3, w->ak, w->st1, w->iv1);
  w->i1 = dot(skp, w->st1, w->iv1);
  mm(n, n, 3, w->al, w->st2, w->iv2);
  w->i2 = dot(skp, w->st2, w->iv2);
  for (int i = 0; i < skp; ++i) w->ee[i] = w->iv1[i] + w->iv2[i];
  mm(n, n, 3, w->akli, w->ee, w->st);
  w->i3 = dot(skp, w->ee, w->st);
  double prefac = sqrt(pow(8.0, n)) * sqrt(pow(w->detak * w->detal / w->dett, 3));
  w->ov = exp(-(w->i1 + w->i2) + w->i3) * prefac;
  return 1;
}

/* Coulomb pair matrix jij (python/matel.py:166-175). */
static void jij_of(int n, int pp, int q, double *J) {
  for (int i = 0; i < n * n; ++i) J[i] = 0.0;
  if (pp == 0) {
    J[(q - 1) * n + (q - 1)] = 1.0;
  } else {
    J[(pp - 1) * n + (pp - 1)] = 1.0;
    J[(q - 1) * n + (q - 1)] = 1.0;
    J[(pp - 1) * n + (q - 1)] = -1.0;
    J[(q - 1) * n + (pp - 1)] = -1.0;
  }
}

/* ke_i3 and ke_i2 (python/matel.py:144-153). */
static void kinetic_parts(const osys *s, const pairwork *w, double *ke_i3, double *ke_i2) {
  const int n = s->n, skp = 3 * n;
  double t1[NMAX * NMAX], t2[NMAX * NMAX], v[SKMAX], dif1[SKMAX], dif2[SKMAX];
  mm(n, n, n, s->mass, w->al, t1);   /* massb @ alb */
  mm(n, n, n, w->ak, t1, t2);        /* akb @ im1   */
  for (int i = 0; i < skp; ++i) { dif1[i] = w->st[i] - w->st1[i]; dif2[i] = w->st[i] - w->st2[i]; }
  mm(n, n, 3, t2, dif2, v);
  *ke_i3 = dot(skp, dif1, v);
  mm(n, n, n, w->akli, w->al, t1);   /* aklib @ alb */
  mm(n, n, n, w->ak, t1, t2);        /* akb @ im1   */
  mm(n, n, n, s->mass, t2, t1);      /* massb @ im2 */
  *ke_i2 = 3.0 * trace(n, t1);       /* trace of kron(X, I3) = 3 tr X */
}

/* python/matel.py:28-210; stf, ii, jj are 0-based here. */
int necg_oracle_matel(const osys *s, const double *iv, int stf, int ii, int jj, double out[4]) {
  pairwork w;
  if (!pair_setup(s, iv, stf, ii, jj, &w)) return 0;
  const int n = s->n;
  double ke_i3, ke_i2;
  kinetic_parts(s, &w, &ke_i3, &ke_i2);
  double tk = 2.0 * w.ov * (2.0 * ke_i3 + ke_i2);
  double ert = 0.0;
  for (int pp = 0; pp <= n; ++pp)
    for (int q = pp + 1; q <= n; ++q) {
      double J[NMAX * NMAX], ism1[NMAX * NMAX], Jn[NMAX * NMAX], v[SKMAX];
      jij_of(n, pp, q, J);
      mm(n, n, n, J, w.akli, ism1);
      mm(n, n, n, w.akli, ism1, Jn);
      mm(n, n, 3, Jn, w.ee, v);
      double i3er = dot(3 * n, w.ee, v);
      double i2er = trace(n, ism1);
      if (i2er == 0.0) continue;
      double i1er = i3er / i2er;
      ert += s->ncv[pp] * s->ncv[q] * 2.0 * w.ov * necg_oracle_f0(i1er) / sqrt(i2er * PI_);
    }
  double zs = 0.0;
  for (int i = 1; i <= n; ++i) zs += s->ncv[i] * w.st[3 * i - 1];
  out[0] = w.ov;
  out[1] = tk;
  out[2] = ert;
  out[3] = w.ov * zs * s->efs;
  return 1;
}

/* python/gradel.py:27-392; stf, ii, jj 0-based; tgvl/tgvls of length blk. */
int necg_oracle_gradel(const osys *s, const double *iv, int stf, int ii, int jj, double *tgvl,
                       double *tgvls) {
  pairwork w;
  if (!pair_setup(s, iv, stf, ii, jj, &w)) return 0;
  const int n = s->n, svh = n * (n + 1) / 2, skp = 3 * n;
  double aki[NMAX * NMAX], dk;
  if (!inv_det(n, w.ak, aki, &dk)) return 0;
  double dif1[SKMAX], dif2[SKMAX];
  for (int i = 0; i < skp; ++i) { dif1[i] = w.st[i] - w.st1[i]; dif2[i] = w.st[i] - w.st2[i]; }
  double A[NMAX * NMAX], Bm[NMAX * NMAX], Cm[NMAX * NMAX], Dm[NMAX * NMAX], T[SKMAX * 3];

  /* ogs (python/gradel.py:108) */
  double ogs[SKMAX];
  mm(n, n, 3, w.ak, dif1, ogs);
  for (int i = 0; i < skp; ++i) ogs[i] *= 2.0 * w.ov;

  /* ogl (python/gradel.py:114-148); E E^T etc. use the n x 3 forms */
  double ogl[NMAX * (NMAX + 1) / 2];
  memset(ogl, 0, sizeof(ogl));
  mm(n, n, n, aki, w.lk, A);            vech_axpy(n, 1.5, A, ogl);
  mm(n, n, n, w.akli, w.lk, A);         vech_axpy(n, -3.0, A, ogl);
  double s1t[3 * NMAX], eet[3 * NMAX];
  tr_(n, 3, w.st1, s1t);
  tr_(n, 3, w.ee, eet);
  mm(n, 3, n, w.st1, s1t, A);  mm(n, n, n, A, w.lk, Bm);  vech_axpy(n, -2.0, Bm, ogl);
  mm(n, 3, n, w.ee, eet, A);   mm(n, n, n, A, w.akli, Bm); mm(n, n, n, w.akli, Bm, A);
  mm(n, n, n, A, w.lk, Bm);    vech_axpy(n, -2.0, Bm, ogl);
  mm(n, 3, n, w.ee, s1t, A);   mm(n, n, n, w.akli, A, Bm); /* ism2 = akli E S1^T */
  mm(n, n, n, Bm, w.lk, Cm);   vech_axpy(n, 2.0, Cm, ogl);
  tr_(n, n, Bm, A);            mm(n, n, n, A, w.lk, Cm);   vech_axpy(n, 2.0, Cm, ogl);
  for (int i = 0; i < svh; ++i) ogl[i] *= w.ov;

  /* kinetic (python/gradel.py:153-164) */
  double ke_i3, ke_i2;
  kinetic_parts(s, &w, &ke_i3, &ke_i2);
  const double tfac = 2.0 * ke_i3 + ke_i2;

  /* kgs (python/gradel.py:167-184) */
  double al_mass[NMAX * NMAX], aklp[NMAX * NMAX], aklpt[NMAX * NMAX], akliak[NMAX * NMAX];
  mm(n, n, n, s->mass, w.al, al_mass);
  mm(n, n, n, w.ak, al_mass, aklp);           /* akb @ (massb @ alb) */
  tr_(n, n, aklp, aklpt);
  mm(n, n, n, w.akli, w.ak, akliak);
  double iv2k[SKMAX], iv1k[SKMAX];
  mm(n, n, n, aklp, akliak, A);  tr_(n, n, A, Bm);  mm(n, n, 3, Bm, dif1, iv2k);
  mm(n, n, n, aklpt, akliak, A);
  for (int i = 0; i < n * n; ++i) A[i] -= aklpt[i];
  tr_(n, n, A, Bm);              mm(n, n, 3, Bm, dif2, iv1k);
  double kgs[SKMAX];
  for (int i = 0; i < skp; ++i)
    kgs[i] = 2.0 * tfac * ogs[i] + 4.0 * w.ov * iv2k[i] + 4.0 * w.ov * iv1k[i];

  /* kgl (python/gradel.py:187-261) */
  double kgl[NMAX * (NMAX + 1) / 2];
  for (int i = 0; i < svh; ++i) kgl[i] = tfac * ogl[i];
  mm(n, n, n, w.akli, w.lk, A); mm(n, n, n, al_mass, A, Bm); mm(n, n, n, w.ak, Bm, Cm);
  mm(n, n, n, w.akli, Cm, Dm);  vech_axpy(n, -w.ov * 6.0, Dm, kgl);
  vech_axpy(n, w.ov * 6.0, Bm, kgl);      /* al_mass @ akli @ lk */
  double stt[3 * NMAX], s2t[3 * NMAX];
  tr_(n, 3, w.st, stt);
  tr_(n, 3, w.st2, s2t);
  double sds[NMAX * NMAX], sdsk[NMAX * NMAX], skdsk[NMAX * NMAX], sldsk[NMAX * NMAX],
      slds[NMAX * NMAX], sdskT[NMAX * NMAX];
  mm(n, 3, n, w.st, stt, sds);
  mm(n, 3, n, w.st, s1t, sdsk);
  mm(n, 3, n, w.st1, s1t, skdsk);
  mm(n, 3, n, w.st2, s1t, sldsk);
  mm(n, 3, n, w.st2, stt, slds);
  tr_(n, n, sdsk, sdskT);
  double kcm[NMAX * NMAX], P1[NMAX * NMAX], Psum[NMAX * NMAX], Pp[NMAX * NMAX], Pt[NMAX * NMAX];
  for (int i = 0; i < n * n; ++i) Psum[i] = aklp[i] + aklpt[i];
  mm(n, n, n, w.akli, Psum, P1);         /* akli (aklp + aklpt) */
  mm(n, n, n, w.akli, aklp, Pp);         /* akli aklp  */
  mm(n, n, n, w.akli, aklpt, Pt);        /* akli aklpt */
  memset(kcm, 0, sizeof(kcm));
  for (int i = 0; i < n * n; ++i) A[i] = al_mass[i] - P1[i];
  mm(n, n, n, A, sds, Bm);   for (int i = 0; i < n * n; ++i) kcm[i] += Bm[i];
  for (int i = 0; i < n * n; ++i) A[i] = P1[i] - al_mass[i];
  mm(n, n, n, A, sdsk, Bm);  for (int i = 0; i < n * n; ++i) kcm[i] += Bm[i];
  mm(n, n, n, Pt, sdskT, Bm); for (int i = 0; i < n * n; ++i) kcm[i] += Bm[i];
  for (int i = 0; i < n * n; ++i) A[i] = Pp[i] - al_mass[i];
  mm(n, n, n, A, slds, Bm);  for (int i = 0; i < n * n; ++i) kcm[i] += Bm[i];
  for (int i = 0; i < n * n; ++i) A[i] = al_mass[i] - Pp[i];
  mm(n, n, n, A, sldsk, Bm); for (int i = 0; i < n * n; ++i) kcm[i] += Bm[i];
  mm(n, n, n, Pt, skdsk, Bm); for (int i = 0; i < n * n; ++i) kcm[i] -= Bm[i];
  mm(n, n, n, kcm, w.lk, A); vech_axpy(n, 2.0 * w.ov, A, kgl);
  tr_(n, n, kcm, Bm); mm(n, n, n, Bm, w.lk, A); vech_axpy(n, 2.0 * w.ov, A, kgl);
  for (int i = 0; i < svh; ++i) kgl[i] *= 2.0;

  /* Coulomb gradients (python/gradel.py:266-337) */
  double ergst[SKMAX], erglt[NMAX * (NMAX + 1) / 2];
  memset(ergst, 0, sizeof(ergst));
  memset(erglt, 0, sizeof(erglt));
  double EEt[NMAX * NMAX], S1Et[NMAX * NMAX];
  mm(n, 3, n, w.ee, eet, EEt);
  mm(n, 3, n, w.st1, eet, S1Et);
  for (int pp = 0; pp <= n; ++pp)
    for (int q = pp + 1; q <= n; ++q) {
      double J[NMAX * NMAX], ism1[NMAX * NMAX], Jn[NMAX * NMAX], v[SKMAX];
      jij_of(n, pp, q, J);
      mm(n, n, n, J, w.akli, ism1);
      mm(n, n, n, w.akli, ism1, Jn);
      mm(n, n, 3, Jn, w.ee, v);
      double i3er = dot(skp, w.ee, v);
      double i2er = trace(n, ism1);
      if (i2er == 0.0) continue;
      double i1er = i3er / i2er;
      double f = necg_oracle_f0(i1er), fp = necg_oracle_df0(i1er);
      double qq = s->ncv[pp] * s->ncv[q];
      /* ergs */
      double akJ[NMAX * NMAX], iv3[SKMAX];
      mm(n, n, n, w.ak, Jn, akJ);
      mm(n, n, 3, akJ, w.ee, iv3);
      double c0 = qq * 2.0 * f / sqrt(i2er * PI_);
      double c1 = qq * 4.0 * w.ov * fp / sqrt(PI_ * pow(i2er, 3));
      for (int i = 0; i < skp; ++i) ergst[i] += c0 * ogs[i] + c1 * iv3[i];
      /* ergl */
      double ergl[NMAX * (NMAX + 1) / 2];
      double JnL[NMAX * NMAX];
      mm(n, n, n, Jn, w.lk, JnL);
      double a0 = 2.0 * f / sqrt(i2er * PI_);
      double a1 = 2.0 * w.ov * f / sqrt(PI_ * pow(i2er, 3));
      for (int i = 0; i < svh; ++i) ergl[i] = a0 * ogl[i];
      vech_axpy(n, a1, JnL, ergl);
      vech_axpy(n, 4.0 * w.ov * fp * i3er / sqrt(PI_ * pow(i2er, 5)), JnL, ergl);
      double m3 = 4.0 * w.ov * fp / sqrt(PI_ * pow(i2er, 3));
      mm(n, n, n, w.akli, EEt, A);  mm(n, n, n, A, Jn, Bm);   /* akli E E^T Jn */
      mm(n, n, n, Bm, w.lk, Cm);    vech_axpy(n, -m3, Cm, ergl);
      tr_(n, n, Bm, A); mm(n, n, n, A, w.lk, Cm); vech_axpy(n, -m3, Cm, ergl);
      mm(n, n, n, S1Et, Jn, Bm);                                /* S1 E^T Jn */
      mm(n, n, n, Bm, w.lk, Cm);    vech_axpy(n, m3, Cm, ergl);
      tr_(n, n, Bm, A); mm(n, n, n, A, w.lk, Cm); vech_axpy(n, m3, Cm, ergl);
      for (int i = 0; i < svh; ++i) erglt[i] += qq * ergl[i];
    }

  /* Stark gradients (python/gradel.py:342-370) */
  double zsum = 0.0;
  for (int i = 1; i <= n; ++i) zsum += s->ncv[i] * w.st[3 * i - 1];
  (void)zsum;
  double zgs[SKMAX];
  memset(zgs, 0, sizeof(zgs));
  double akakli[NMAX * NMAX];
  mm(n, n, n, w.ak, w.akli, akakli);
  for (int i = 1; i <= n; ++i) {
    for (int r = 0; r < skp; ++r) {
      /* column 3i-1 of kron(ak akli, I3): row r = 3a + x is (ak akli)[a, i-1] if x == z */
      double col = (r % 3 == 2) ? akakli[(r / 3) * n + (i - 1)] : 0.0;
      zgs[r] += s->ncv[i] * (ogs[r] * w.st[3 * i - 1] + w.ov * col) * s->efs;
    }
  }
  double zgl[NMAX * (NMAX + 1) / 2];
  memset(zgl, 0, sizeof(zgl));
  for (int kk = 1; kk <= n; ++kk) {
    double Z1[NMAX * NMAX];
    for (int i = 0; i < n * n; ++i) Z1[i] = 0.0;
    for (int a = 0; a < n; ++a) Z1[a * n + (kk - 1)] = w.st1[3 * a + 2] - w.st[3 * a + 2];
    double v1[NMAX * (NMAX + 1) / 2], v2[NMAX * (NMAX + 1) / 2];
    mm(n, n, n, Z1, w.akli, A); mm(n, n, n, A, w.lk, Bm); vech_(n, Bm, v1);
    tr_(n, n, Z1, A); mm(n, n, n, w.akli, A, Bm); mm(n, n, n, Bm, w.lk, Cm); vech_(n, Cm, v2);
    for (int i = 0; i < svh; ++i)
      zgl[i] += s->ncv[kk] * s->efs * (w.st[3 * kk - 1] * ogl[i] + w.ov * (v1[i] + v2[i]));
  }

  /* assembly (python/gradel.py:377-390) */
  for (int i = 0; i < svh; ++i) {
    tgvl[i] = kgl[i] + erglt[i] - zgl[i];
    tgvls[i] = ogl[i];
  }
  for (int i = 0; i < skp; ++i) {
    tgvl[svh + i] = kgs[i] + ergst[i] - zgs[i];
    tgvls[svh + i] = ogs[i];
  }
  if (ii == jj) {
    for (int i = 0; i < svh + skp; ++i) { tgvl[i] *= 2.0; tgvls[i] *= 2.0; }
  }
  (void)T;
  return 1;
}

/* H and S (python/outprod.py:49-60): lower triangle bra i >= ket j, then
 * mirrored.  Returns the number of failed pair set-ups (singular A_k + A_l). */
int necg_oracle_matrices(const osys *s, int K, const double *iv, double *H, double *S,
                         int nthreads) {
  int bad = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
#endif
  for (int i = 0; i < K; ++i) {
    for (int j = 0; j <= i; ++j) {
      double h = 0.0, o = 0.0;
      for (int stf = 0; stf < s->nst; ++stf) {
        double e[4];
        if (!necg_oracle_matel(s, iv, stf, i, j, e)) { bad++; continue; }
        h += s->scvv[stf] * (e[1] + e[2] - e[3]);
        o += s->scvv[stf] * e[0];
      }
      H[(size_t)i * K + j] = h;
      S[(size_t)i * K + j] = o;
    }
  }
  for (int j = 0; j < K; ++j)
    for (int i = 0; i < j; ++i) {
      H[(size_t)i * K + j] = H[(size_t)j * K + i];
      S[(size_t)i * K + j] = S[(size_t)j * K + i];
    }
  return bad;
}

/* DH/DS row block of bra l (python/make_grad.py:37-55):
 * dh[r*K + k] = sum_stf scvv * gradel(stf, l, k)[r]. */
static int grad_rows(const osys *s, int K, const double *iv, int l, double *dh, double *ds) {
  const int n = s->n, blk = n * (n + 1) / 2 + 3 * n;
  double tg[64], tgs[64];
  for (int k = 0; k < K; ++k) {
    double a[64], b[64];
    for (int r = 0; r < blk; ++r) { a[r] = 0.0; b[r] = 0.0; }
    for (int stf = 0; stf < s->nst; ++stf) {
      if (!necg_oracle_gradel(s, iv, stf, l, k, tg, tgs)) return 0;
      for (int r = 0; r < blk; ++r) { a[r] += s->scvv[stf] * tg[r]; b[r] += s->scvv[stf] * tgs[r]; }
    }
    for (int r = 0; r < blk; ++r) { dh[(size_t)r * K + k] = a[r]; ds[(size_t)r * K + k] = b[r]; }
  }
  return 1;
}

/* The gradient rows of bras [row0, row0+nrows) given E and norm = 1/cSc:
 * GRADA (python/make_grad.py:61-79).  Returns failures. */
int necg_oracle_grada_rows(const osys *s, int K, const double *x, double eng, double norm,
                           int row0, int nrows, double *grada, int nthreads) {
  const int n = s->n, blk = n * (n + 1) / 2 + 3 * n;
  const double *iv = x, *C = x + (size_t)K * blk;
  int bad = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel reduction(+ : bad)
#endif
  {
    double *dh = (double *)malloc(sizeof(double) * blk * K);
    double *ds = (double *)malloc(sizeof(double) * blk * K);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int j = row0; j < row0 + nrows; ++j) {
      if (!grad_rows(s, K, iv, j, dh, ds)) { bad++; continue; }
      for (int r = 0; r < blk; ++r) {
        double val = 0.0;
        for (int i = 0; i < K; ++i) {
          double gm = norm * (dh[(size_t)r * K + i] - eng * ds[(size_t)r * K + i]);
          if (i != j) val += 2.0 * C[i] * C[j] * gm;
          else val += C[i] * C[j] * gm;
        }
        grada[(size_t)(j - row0) * blk + r] = val;
      }
    }
    free(dh);
    free(ds);
  }
  return bad;
}

/* outprod (python/outprod.py:20-76 with make_grad python/make_grad.py:18-90).
 * H, S may be NULL.  Returns 0 on success, 1 for a singular pair matrix,
 * 2 for cSc == 0. */
int necg_oracle_outprod(const osys *s, int K, const double *x, double *f, double *g, double *Hout,
                        double *Sout, int nthreads) {
  const int n = s->n, blk = n * (n + 1) / 2 + 3 * n;
  const double *C = x + (size_t)K * blk;
  double *H = (double *)malloc(sizeof(double) * K * K);
  double *S = (double *)malloc(sizeof(double) * K * K);
  int rc = 0;
  if (necg_oracle_matrices(s, K, x, H, S, nthreads)) { rc = 1; goto done; }
  double cHc = 0.0, cSc = 0.0;
  for (int j = 0; j < K; ++j) {  /* (C @ H) @ C */
    double ch = 0.0, cs = 0.0;
    for (int i = 0; i < K; ++i) { ch += C[i] * H[(size_t)i * K + j]; cs += C[i] * S[(size_t)i * K + j]; }
    cHc += ch * C[j];
    cSc += cs * C[j];
  }
  if (cSc == 0.0) { rc = 2; goto done; }
  double eng = cHc / cSc, norm = 1.0 / cSc;
  *f = eng;
  if (necg_oracle_grada_rows(s, K, x, eng, norm, 0, K, g, nthreads)) { rc = 1; goto done; }
  for (int i = 0; i < K; ++i) {  /* GRADC (python/make_grad.py:85) */
    double a = 0.0;
    for (int j = 0; j < K; ++j) a += (H[(size_t)i * K + j] - eng * S[(size_t)i * K + j]) * C[j];
    g[(size_t)K * blk + i] = 2.0 * norm * a;
  }
  if (Hout) memcpy(Hout, H, sizeof(double) * K * K);
  if (Sout) memcpy(Sout, S, sizeof(double) * K * K);
done:
  free(H);
  free(S);
  return rc;
}

/* outprod as above, every sum (the term sums of each element and gradient
 * block, C^T H C, C^T S C, the GRADA row sums and GRADC) carried in long
 * double (x87 80-bit, 64-bit significand).  The per-term values are the same
 * double-precision matel/gradel results.  Returns the double-accumulated
 * (E, g) of necg_oracle_outprod in f[0], g and the long-double-accumulated
 * ones in f[1], gl, so that one pass gives both the reference's summation
 * and a value whose only rounding is that of the terms themselves: the
 * full-size fixture of the K = 2048 evaluation is written from it. */
int necg_oracle_outprod_ld(const osys *s, int K, const double *x, double *f, double *g,
                           double *gl, int nthreads) {
  const int n = s->n, blk = n * (n + 1) / 2 + 3 * n;
  const double *iv = x, *C = x + (size_t)K * blk;
  double *H = (double *)malloc(sizeof(double) * K * K);
  double *S = (double *)malloc(sizeof(double) * K * K);
  long double *HL = (long double *)malloc(sizeof(long double) * K * K);
  long double *SL = (long double *)malloc(sizeof(long double) * K * K);
  int bad = 0, rc = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
#endif
  for (int i = 0; i < K; ++i) {
    for (int j = 0; j <= i; ++j) {
      double h = 0.0, o = 0.0;
      long double hl = 0.0L, ol = 0.0L;
      for (int stf = 0; stf < s->nst; ++stf) {
        double e[4];
        if (!necg_oracle_matel(s, iv, stf, i, j, e)) { bad++; continue; }
        h += s->scvv[stf] * (e[1] + e[2] - e[3]);
        o += s->scvv[stf] * e[0];
        hl += (long double)s->scvv[stf] * ((long double)e[1] + e[2] - e[3]);
        ol += (long double)s->scvv[stf] * e[0];
      }
      H[(size_t)i * K + j] = H[(size_t)j * K + i] = h;
      S[(size_t)i * K + j] = S[(size_t)j * K + i] = o;
      HL[(size_t)i * K + j] = HL[(size_t)j * K + i] = hl;
      SL[(size_t)i * K + j] = SL[(size_t)j * K + i] = ol;
    }
  }
  if (bad) { rc = 1; goto done; }
  double cHc = 0.0, cSc = 0.0;
  long double cHcl = 0.0L, cScl = 0.0L;
  for (int j = 0; j < K; ++j) {
    double ch = 0.0, cs = 0.0;
    long double chl = 0.0L, csl = 0.0L;
    for (int i = 0; i < K; ++i) {
      ch += C[i] * H[(size_t)i * K + j];
      cs += C[i] * S[(size_t)i * K + j];
      chl += (long double)C[i] * HL[(size_t)i * K + j];
      csl += (long double)C[i] * SL[(size_t)i * K + j];
    }
    cHc += ch * C[j];
    cSc += cs * C[j];
    cHcl += chl * C[j];
    cScl += csl * C[j];
  }
  if (cSc == 0.0 || cScl == 0.0L) { rc = 2; goto done; }
  const double eng = cHc / cSc, norm = 1.0 / cSc;
  const long double engl = cHcl / cScl, norml = 1.0L / cScl;
  f[0] = eng;
  f[1] = (double)engl;
#ifdef _OPENMP
#pragma omp parallel reduction(+ : bad)
#endif
  {
    double tg[64], tgs[64];
    long double *dl = (long double *)malloc(sizeof(long double) * 2 * blk * K);
    double *dd = (double *)malloc(sizeof(double) * 2 * blk * K);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int l = 0; l < K; ++l) {
      int ok = 1;
      for (int k = 0; k < K && ok; ++k) {
        double a[64], b[64];
        long double al[64], bl[64];
        for (int r = 0; r < blk; ++r) { a[r] = b[r] = 0.0; al[r] = bl[r] = 0.0L; }
        for (int stf = 0; stf < s->nst; ++stf) {
          if (!necg_oracle_gradel(s, iv, stf, l, k, tg, tgs)) { ok = 0; break; }
          for (int r = 0; r < blk; ++r) {
            a[r] += s->scvv[stf] * tg[r];
            b[r] += s->scvv[stf] * tgs[r];
            al[r] += (long double)s->scvv[stf] * tg[r];
            bl[r] += (long double)s->scvv[stf] * tgs[r];
          }
        }
        for (int r = 0; r < blk; ++r) {
          dd[(size_t)r * K + k] = a[r];
          dd[(size_t)(blk + r) * K + k] = b[r];
          dl[(size_t)r * K + k] = al[r];
          dl[(size_t)(blk + r) * K + k] = bl[r];
        }
      }
      if (!ok) { bad++; continue; }
      for (int r = 0; r < blk; ++r) {
        double val = 0.0;
        long double vall = 0.0L;
        for (int i = 0; i < K; ++i) {
          const double w = (i != l) ? 2.0 : 1.0;
          double gm = norm * (dd[(size_t)r * K + i] - eng * dd[(size_t)(blk + r) * K + i]);
          long double gml = norml * (dl[(size_t)r * K + i] - engl * dl[(size_t)(blk + r) * K + i]);
          val += w * C[i] * C[l] * gm;
          vall += (long double)w * C[i] * C[l] * gml;
        }
        g[(size_t)l * blk + r] = val;
        gl[(size_t)l * blk + r] = (double)vall;
      }
    }
    free(dl);
    free(dd);
  }
  if (bad) { rc = 1; goto done; }
  for (int i = 0; i < K; ++i) {  /* GRADC (python/make_grad.py:85) */
    double a = 0.0;
    long double al = 0.0L;
    for (int j = 0; j < K; ++j) {
      a += (H[(size_t)i * K + j] - eng * S[(size_t)i * K + j]) * C[j];
      al += (HL[(size_t)i * K + j] - engl * SL[(size_t)i * K + j]) * C[j];
    }
    g[(size_t)K * blk + i] = 2.0 * norm * a;
    gl[(size_t)K * blk + i] = (double)(2.0L * norml * al);
  }
done:
  free(H);
  free(S);
  free(HL);
  free(SL);
  return rc;
}

/* Per-pair sums over all terms: H_ij, S_ij of the lower-triangle element
 * (bra max(i,j), ket min(i,j)) and the DH/DS blocks of ordered pair (i, j). */
int necg_oracle_pair_sums(const osys *s, const double *iv, int npairs, const int32_t *pairs,
                          double *Hij, double *Sij, double *dh, double *ds) {
  const int n = s->n, blk = n * (n + 1) / 2 + 3 * n;
  for (int t = 0; t < npairs; ++t) {
    int i = pairs[2 * t], j = pairs[2 * t + 1];
    int a = i > j ? i : j, b = i > j ? j : i;
    double h = 0.0, o = 0.0, tg[64], tgs[64];
    for (int r = 0; r < blk; ++r) { dh[t * blk + r] = 0.0; ds[t * blk + r] = 0.0; }
    for (int stf = 0; stf < s->nst; ++stf) {
      double e[4];
      if (!necg_oracle_matel(s, iv, stf, a, b, e)) return 0;
      h += s->scvv[stf] * (e[1] + e[2] - e[3]);
      o += s->scvv[stf] * e[0];
      if (!necg_oracle_gradel(s, iv, stf, i, j, tg, tgs)) return 0;
      for (int r = 0; r < blk; ++r) {
        dh[t * blk + r] += s->scvv[stf] * tg[r];
        ds[t * blk + r] += s->scvv[stf] * tgs[r];
      }
    }
    Hij[t] = h;
    Sij[t] = o;
  }
  return 1;
}

/* Element rows [row0, row0+nrows) of the lower triangle (bra i >= ket j) of
 * H and S (python/outprod.py:49-54): Hr[(i-row0)*K + j].  Used to time a
 * bounded sample of the matel loop.  Returns failures. */
int necg_oracle_matel_rows(const osys *s, int K, const double *iv, int row0, int nrows,
                           double *Hr, double *Sr, int nthreads) {
  int bad = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
#endif
  for (int i = row0; i < row0 + nrows; ++i) {
    for (int j = 0; j <= i; ++j) {
      double h = 0.0, o = 0.0;
      for (int stf = 0; stf < s->nst; ++stf) {
        double e[4];
        if (!necg_oracle_matel(s, iv, stf, i, j, e)) { bad++; continue; }
        h += s->scvv[stf] * (e[1] + e[2] - e[3]);
        o += s->scvv[stf] * e[0];
      }
      Hr[(size_t)(i - row0) * K + j] = h;
      Sr[(size_t)(i - row0) * K + j] = o;
    }
  }
  return bad;
}

/* rrms (python/rrms.py:21-120, source/rrms.f90:1-250) for bra ii under term
 * stfe and ket jj under term stf (0-based): the overlap ov and ov * st[c] for
 * every component c of st = (A_k' + A_l')^-1 (A_k' s_k' + A_l' s_l') (the
 * reference returns one component, n1; all 3n are returned here).  Both bra
 * and ket are transformed: A' = sml^T A sml, s' = sms s.  Returns 0 if the
 * pair matrix is singular. */
int necg_oracle_rrms(const osys *s, const double *iv, int stfe, int stf, int ii, int jj,
                     double *ov, double *tk) {
  const int n = s->n, svh = n * (n + 1) / 2, skp = 3 * n, blk = svh + skp;
  double lk[NMAX * NMAX] = {0}, ell[NMAX * NMAX] = {0}, lkt[NMAX * NMAX] = {0}, ellt[NMAX * NMAX] = {0};
  int kk = ii * blk, ll = jj * blk;
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      lk[j * n + i] = iv[kk++];
      ell[j * n + i] = iv[ll++];
    }
  double ak0[NMAX * NMAX], al0[NMAX * NMAX], ak[NMAX * NMAX], al[NMAX * NMAX];
  double tmp[NMAX * NMAX], smt[NMAX * NMAX];
  tr_(n, n, lk, lkt);
  tr_(n, n, ell, ellt);
  mm(n, n, n, lk, lkt, ak0);
  mm(n, n, n, ell, ellt, al0);
  const double *pl = s->sml + (size_t)stf * n * n, *pk = s->sml + (size_t)stfe * n * n;
  mm(n, n, n, al0, pl, tmp);       /* ket: sml^T (al sml) */
  tr_(n, n, pl, smt);
  mm(n, n, n, smt, tmp, al);
  mm(n, n, n, ak0, pk, tmp);       /* bra: sml^T (ak sml) */
  tr_(n, n, pk, smt);
  mm(n, n, n, smt, tmp, ak);
  double st1[SKMAX], st2[SKMAX];
  const double *s1 = iv + ii * blk + svh, *s2 = iv + jj * blk + svh;
  const double *vl = s->sms + (size_t)stf * skp * skp, *vk = s->sms + (size_t)stfe * skp * skp;
  for (int r = 0; r < skp; ++r) {
    double a = 0.0, b = 0.0;
    for (int c = 0; c < skp; ++c) {
      a += vl[r * skp + c] * s2[c];
      b += vk[r * skp + c] * s1[c];
    }
    st2[r] = a;
    st1[r] = b;
  }
  double akl[NMAX * NMAX], akli[NMAX * NMAX], d;
  for (int i = 0; i < n * n; ++i) akl[i] = ak[i] + al[i];
  if (!inv_det(n, akl, akli, &d)) return 0;
  const double dett = fabs(d);
  double pk_ = 1.0, pl_ = 1.0;
  for (int i = 0; i < n; ++i) { pk_ *= lk[i * n + i]; pl_ *= ell[i * n + i]; }
  const double detak = sqrt(fabs(pk_ * pk_)), detal = sqrt(fabs(pl_ * pl_));
  double iv1[SKMAX], iv2[SKMAX], ee[SKMAX], st[SKMAX];
  mm(n, n, 3, ak, st1, iv1);
  mm(n, n, 3, al, st2, iv2);
  const double i1 = dot(skp, st1, iv1), i2 = dot(skp, st2, iv2);
  for (int i = 0; i < skp; ++i) ee[i] = iv1[i] + iv2[i];
  mm(n, n, 3, akli, ee, st);
  const double i3 = dot(skp, ee, st);
  const double o = exp(-(i1 + i2) + i3) * sqrt(pow(8.0, n)) * sqrt(pow(detak * detal / dett, 3));
  *ov = o;
  for (int c = 0; c < skp; ++c) tk[c] = st[c] * o;
  return 1;
}

/* The dipole pass of the driver (python/necg.py:200-232, source/necg.f90:387-446):
 * for every component c = 3 i + xyz, h_c and s summed over all term pairs
 * (stfe, stf) with weight scvv[stf] scvv[stfe] on the lower triangle (bra
 * j >= ket i), mirrored, then rbar[c] = C^T h_c C / C^T s C and
 * dip[xyz] = sum_i ncv[i + 1] rbar[3 i + xyz].  *cSc = C^T s C.
 * Returns 0 on success, 1 for a singular pair matrix, 2 for C^T s C == 0. */
int necg_oracle_dipole(const osys *s, int K, const double *x, double *cSc, double *rbar,
                       double *dip, int nthreads) {
  const int n = s->n, skp = 3 * n, blk = n * (n + 1) / 2 + skp, nst = s->nst;
  const double *C = x + (size_t)K * blk;
  double *h = (double *)calloc((size_t)K * K * skp, sizeof(double));
  double *sd = (double *)calloc((size_t)K * K, sizeof(double));
  int bad = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
#endif
  for (int j = 0; j < K; ++j) {
    for (int i = 0; i <= j; ++i) {
      double hs[SKMAX] = {0}, ss = 0.0, tk[SKMAX], ov;
      for (int stfe = 0; stfe < nst; ++stfe)
        for (int stf = 0; stf < nst; ++stf) {
          if (!necg_oracle_rrms(s, x, stfe, stf, j, i, &ov, tk)) { bad++; continue; }
          const double w = s->scvv[stf] * s->scvv[stfe];
          for (int c = 0; c < skp; ++c) hs[c] += w * tk[c];
          ss += w * ov;
        }
      for (int c = 0; c < skp; ++c) {
        h[((size_t)c * K + j) * K + i] = hs[c];
        h[((size_t)c * K + i) * K + j] = hs[c];
      }
      sd[(size_t)j * K + i] = ss;
      sd[(size_t)i * K + j] = ss;
    }
  }
  int rc = 0;
  if (bad) { rc = 1; goto done; }
  double d2 = 0.0;
  for (int i = 0; i < K; ++i)
    for (int j = 0; j < K; ++j) d2 += C[i] * sd[(size_t)i * K + j] * C[j];
  if (d2 == 0.0) { rc = 2; goto done; }
  *cSc = d2;
  for (int a = 0; a < 3; ++a) dip[a] = 0.0;
  for (int c = 0; c < skp; ++c) {
    double d1 = 0.0;
    const double *hc = h + (size_t)c * K * K;
    for (int i = 0; i < K; ++i)
      for (int j = 0; j < K; ++j) d1 += C[i] * hc[(size_t)i * K + j] * C[j];
    rbar[c] = d1 / d2;
    dip[c % 3] += s->ncv[c / 3 + 1] * rbar[c];
  }
done:
  free(h);
  free(sd);
  return rc;
}
