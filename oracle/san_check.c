/*
 * san_check.c -- runs the oracle's entry points on one system read from a
 * binary file, for a build with AddressSanitizer and UndefinedBehaviorSanitizer
 * (make -C oracle san_check).  TEST INFRASTRUCTURE ONLY.
 *
 * Input (little-endian int32 then float64): n, nst, K; efs; mass (n*n);
 * ncv (n+1); scvv (nst); sml (nst*n*n); sms (nst*9*n*n); x (K*(blk+1)).
 * Output (float64): E, g (len x); E and g of the long-double sums; H, S
 * (K*K); per (term p, bra i, ket j) with p < nst, i, j < K in that order,
 * matel (4) and gradel (2*blk); rrms (1 + 3n) of every (p, q, i, j);
 * the dipole pass (1 + 3n + 3).  Exit status 0 if every call succeeded.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

typedef struct {
  int n, nst;
  double efs;
  const double *mass, *ncv, *scvv, *sml, *sms;
} osys;

int necg_oracle_matel(const osys *s, const double *iv, int stf, int ii, int jj, double out[4]);
int necg_oracle_gradel(const osys *s, const double *iv, int stf, int ii, int jj, double *tgvl,
                       double *tgvls);
int necg_oracle_outprod(const osys *s, int K, const double *x, double *f, double *g,
                        double *Hout, double *Sout, int nthreads);
int necg_oracle_outprod_ld(const osys *s, int K, const double *x, double *f, double *g,
                           double *gl, int nthreads);
int necg_oracle_rrms(const osys *s, const double *iv, int stfe, int stf, int ii, int jj,
                     double *ov, double *tk);
int necg_oracle_dipole(const osys *s, int K, const double *x, double *cSc, double *rbar,
                       double *dip, int nthreads);

static double *rd(FILE *f, size_t n) {
  double *p = (double *)malloc(sizeof(double) * (n ? n : 1));
  if (!p || fread(p, sizeof(double), n, f) != n) exit(3);
  return p;
}

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t hdr[3];
  if (fread(hdr, sizeof(int32_t), 3, f) != 3) return 3;
  const int n = hdr[0], nst = hdr[1], K = hdr[2];
  const int blk = n * (n + 1) / 2 + 3 * n, nx = K * (blk + 1);
  osys s;
  s.n = n;
  s.nst = nst;
  double *efs = rd(f, 1);
  s.efs = efs[0];
  s.mass = rd(f, (size_t)n * n);
  s.ncv = rd(f, (size_t)n + 1);
  s.scvv = rd(f, (size_t)nst);
  s.sml = rd(f, (size_t)nst * n * n);
  s.sms = rd(f, (size_t)nst * 9 * n * n);
  double *x = rd(f, (size_t)nx);
  fclose(f);
  FILE *o = fopen(argv[2], "wb");
  if (!o) return 2;
  int bad = 0;
  double E, E2[2];
  double *g = (double *)calloc(nx, sizeof(double)), *gd = (double *)calloc(nx, sizeof(double));
  double *gl = (double *)calloc(nx, sizeof(double));
  double *H = (double *)calloc((size_t)K * K, sizeof(double));
  double *S = (double *)calloc((size_t)K * K, sizeof(double));
  bad |= necg_oracle_outprod(&s, K, x, &E, g, H, S, 2) != 0;
  bad |= necg_oracle_outprod_ld(&s, K, x, E2, gd, gl, 2) != 0;
  fwrite(&E, sizeof(double), 1, o);
  fwrite(g, sizeof(double), nx, o);
  fwrite(&E2[1], sizeof(double), 1, o);
  fwrite(gl, sizeof(double), nx, o);
  fwrite(H, sizeof(double), (size_t)K * K, o);
  fwrite(S, sizeof(double), (size_t)K * K, o);
  double *tg = (double *)malloc(sizeof(double) * blk), *ts = (double *)malloc(sizeof(double) * blk);
  for (int p = 0; p < nst; ++p)
    for (int i = 0; i < K; ++i)
      for (int j = 0; j < K; ++j) {
        double e[4];
        bad |= !necg_oracle_matel(&s, x, p, i, j, e);
        bad |= !necg_oracle_gradel(&s, x, p, i, j, tg, ts);
        fwrite(e, sizeof(double), 4, o);
        fwrite(tg, sizeof(double), blk, o);
        fwrite(ts, sizeof(double), blk, o);
      }
  double *tk = (double *)malloc(sizeof(double) * 3 * n);
  for (int p = 0; p < nst; ++p)
    for (int q = 0; q < nst; ++q)
      for (int i = 0; i < K; ++i)
        for (int j = 0; j < K; ++j) {
          double ov;
          bad |= !necg_oracle_rrms(&s, x, p, q, i, j, &ov, tk);
          fwrite(&ov, sizeof(double), 1, o);
          fwrite(tk, sizeof(double), 3 * n, o);
        }
  double css, rbar[3 * 8], dip[3];
  bad |= necg_oracle_dipole(&s, K, x, &css, rbar, dip, 2) != 0;
  fwrite(&css, sizeof(double), 1, o);
  fwrite(rbar, sizeof(double), 3 * n, o);
  fwrite(dip, sizeof(double), 3, o);
  fclose(o);
  free(g); free(gd); free(gl); free(H); free(S); free(tg); free(ts); free(tk); free(x);
  free(efs); free((void *)s.mass); free((void *)s.ncv); free((void *)s.scvv);
  free((void *)s.sml); free((void *)s.sms);
  return bad ? 1 : 0;
}
