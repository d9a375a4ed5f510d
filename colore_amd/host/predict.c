/*
 * predict.c -- theory predictions written next to the simulation outputs:
 * write_predictions (predictions.c:26-172) with the FFTLog transforms
 * pk2xi / xi2pk (fftlog.c).
 *
 * FFTLog is Hamilton's fast Hankel transform (Hamilton 2000, MNRAS 312,
 * 257, appendix B): on a log-spaced grid the transform of order mu is a
 * convolution whose kernel in Fourier space is
 *   u_m = (2/k0r0)^(-2 pi i m / L) Gamma((mu+1+2 pi i m/L)/2) / Gamma((mu+1-2 pi i m/L)/2)
 * (q = 0), evaluated here with a Stirling series for ln Gamma and a
 * mixed-radix complex FFT of our own; k0r0 is moved to the nearest
 * low-ringing value as the reference does (goodkr).
 */
#include "clr_host.h"
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ln Gamma(z) for Re z > 0: shift to Re z >= 10 by the recurrence, then the
 * Stirling series to z^-15 (error far below 1e-15 there).  The imaginary part
 * is the continuous branch; FFTLog only uses it through exp(i .) and through
 * an integer rounding whose fractional part a shift by 2 pi leaves alone. */
static double complex lngamma_c(double complex z) {
  double complex shift = 0;
  while (creal(z) < 10) {
    shift += clog(z);
    z += 1;
  }
  const double complex iz = 1 / z, iz2 = iz * iz;
  /* B_2n / (2n (2n-1)) for n = 1..8 */
  static const double c[8] = {1.0 / 12, -1.0 / 360, 1.0 / 1260, -1.0 / 1680, 1.0 / 1188,
                              -691.0 / 360360, 1.0 / 156, -3617.0 / 122400};
  double complex s = 0, p = iz;
  for (int n = 0; n < 8; n++) {
    s += c[n] * p;
    p *= iz2;
  }
  return (z - 0.5) * clog(z) - z + 0.5 * log(2 * M_PI) + s - shift;
}

/* ---- complex FFT of any length (radix by smallest prime factor) ---------- */
typedef struct {
  int n;
  double complex *w;     /* exp(sign 2 pi i j / n), j < n */
  double complex *tmp;
} fftplan;

static void fft_rec(const fftplan *P, int n, const double complex *in, int stride, double complex *out) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int p = 2;
  while (n % p) p = p == 2 ? 3 : p + 2;
  const int m = n / p;
  for (int j = 0; j < p; j++) fft_rec(P, m, in + (size_t)j * stride, stride * p, out + (size_t)j * m);
  /* combine: X[k + q m] = sum_j W_n^{j (k + q m)} Y_j[k] */
  const int step = P->n / n;
  double complex *t = P->tmp;
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      double complex acc = 0;
      for (int j = 0; j < p; j++) {
        const long e = ((long)j * (k + q * m)) % n;
        acc += P->w[e * step] * out[(size_t)j * m + k];
      }
      t[q] = acc;
    }
    for (int q = 0; q < p; q++) out[(size_t)q * m + k] = t[q];
  }
}

/* out = sum_j in[j] exp(sign 2 pi i j k / n) */
static void fft(int n, int sign, const double complex *in, double complex *out) {
  fftplan P = {n, malloc(n * sizeof(double complex)), malloc(n * sizeof(double complex))};
  for (int j = 0; j < n; j++) P.w[j] = cexp(sign * 2 * M_PI * I * ((double)j / n));
  fft_rec(&P, n, in, 1, out);
  free(P.w);
  free(P.tmp);
}

/* fht (fftlog.c) with q = 0, kcrc = 1 and the low-ringing adjustment */
static void fht(int N, const double *r, const double complex *a, double *k, double complex *b, double mu) {
  const double L = log(r[N - 1] / r[0]) * N / (N - 1.);
  double kcrc = 1;
  {
    /* goodkr: make ln(2/kr) N/L + arg(Gamma(x+iy)^2)/pi an integer */
    const double x = (mu + 1) / 2, y = M_PI * N / (2 * L);
    const double ph = cimag(lngamma_c(x + I * y));
    const double arg = log(2 / kcrc) * N / L + 2 * ph / M_PI;
    const double iarg = round(arg);
    if (arg != iarg) kcrc *= exp((arg - iarg) * L / N);
  }
  const double y = M_PI / L, k0r0 = kcrc * exp(-L), t = -2 * y * log(k0r0 / 2);
  double complex *u = malloc(N * sizeof(double complex));
  for (int m = 0; m <= N / 2; m++) {
    const double ph = cimag(lngamma_c((mu + 1) / 2 + I * (m * y)));
    u[m] = cexp(I * (m * t + 2 * ph));
  }
  for (int m = N / 2 + 1; m < N; m++) u[m] = conj(u[N - m]);
  if (N % 2 == 0) u[N / 2] = creal(u[N / 2]);
  double complex *f = malloc(N * sizeof(double complex));
  fft(N, -1, a, f);
  for (int m = 0; m < N; m++) f[m] *= u[m] / N;
  fft(N, +1, f, b);
  for (int n = 0; n < N / 2; n++) {
    const double complex s = b[n];
    b[n] = b[N - 1 - n];
    b[N - 1 - n] = s;
  }
  k[0] = k0r0 / r[0];
  for (int n = 1; n < N; n++) k[n] = k[0] * exp(n * L / N);
  free(u);
  free(f);
}

/* fftlog_ComputeXiLM with l = 0, m = 2 */
static void xi_l0(int N, const double *k, const double *pk, double *r, double *xi) {
  double complex *a = malloc(N * sizeof(double complex)), *b = malloc(N * sizeof(double complex));
  for (int i = 0; i < N; i++) a[i] = pow(k[i], 1.5) * pk[i];
  fht(N, k, a, r, b, 0.5);
  for (int i = 0; i < N; i++) xi[i] = creal(pow(2 * M_PI * r[i], -1.5) * b[i]);
  free(a);
  free(b);
}

void clr_pk2xi(int N, const double *k, const double *pk, double *r, double *xi) { xi_l0(N, k, pk, r, xi); }

void clr_xi2pk(int N, const double *r, const double *xi, double *k, double *pk) {
  xi_l0(N, r, xi, k, pk);
  for (int j = 0; j < N; j++) pk[j] *= 8 * M_PI * M_PI * M_PI;
}

/* pk_linear0 (cosmo.c:291-308) on the run's table */
static double pk_lin0(const clr_run *r, double lgk) {
  const int ik = (int)((lgk - r->logkmin) * r->idlogk);
  if (ik < 0) return r->pkarr[0] * pow(10, r->n_scal * (lgk - r->logkmin));
  if (ik < r->numk - 1) return r->pkarr[ik] + (lgk - r->logkarr[ik]) * (r->pkarr[ik + 1] - r->pkarr[ik]) * r->idlogk;
  return r->pkarr[r->numk - 1] * pow(10, -3 * (lgk - r->logkmax));
}

/* write_predictions (predictions.c:26-172) */
int clr_write_predictions(const clr_run *r) {
  if (r->n_tr[0] == 0 && r->n_tr[1] == 0 && r->n_tr[2] == 0) return 0;
  if (!(r->pred_dz > 0)) return 1;
  printf("*** Writing predictions\n");
  const int Nk = 10000;
  const double kmin = 1e-4, kmax = 100, rminout = 0.5, rmaxout = 300.;
  double *ka = malloc(6 * (size_t)Nk * sizeof(double));
  double *pk = ka + Nk, *pklin = ka + 2 * Nk, *ra = ka + 3 * Nk, *xi = ka + 4 * Nk, *xilin = ka + 5 * Nk;
  for (int i = 0; i < Nk; i++) ka[i] = kmin * pow(kmax / kmin, i * 1.0 / (Nk - 1));
  const double r2s = r->r_smooth > 0 ? r->r_smooth * r->r_smooth : r->r_smooth;   /* io.c:453-458 */
  const double dx2 = pow(r->l_box / r->n_grid, 2);
  const double rsm2_gg = r2s + 0.9 * dx2 / 12., rsm2_gm = r2s + 2.0 * dx2 / 6., rsm2_mm = r2s + 1.88 * dx2 / 6.;
  char fname[CLR_PATH + 96];
  snprintf(fname, sizeof(fname), "%s_gbias.txt", r->prefix_out);
  FILE *fg = fopen(fname, "w");
  if (!fg) { free(ka); return 1; }
  fprintf(fg, "#1-z 2-r(z) 3-g(z) ");
  for (int i = 0; i < r->n_tr[0]; i++) fprintf(fg, "%d-bg_%d(z) ", i + 4, i + 1);
  for (int i = 0; i < r->n_tr[1]; i++) fprintf(fg, "%d-bi_%d(z) ", i + 4 + r->n_tr[0], i + 1);
  for (int i = 0; i < r->n_tr[2]; i++) fprintf(fg, "%d-bc_%d(z) ", i + 4 + r->n_tr[0] + r->n_tr[1], i + 1);
  fprintf(fg, "\n");
  static const char *kind_name[3] = {"srcs", "imap", "custom"};
  static const int bz_tag[3] = {CLR_BG_BZ_SRCS, CLR_BG_BZ_IMAP, CLR_BG_BZ_CSTM};
  int rc = 0;
  for (double z = 0; z <= r->z_max; z += r->pred_dz) {
    const double rr = clr_run_r_of_z(r, z), g = clr_run_get_bg(r, rr, CLR_BG_D1, 0);
    fprintf(fg, "%g %g %g ", z, rr, g);
    for (int i = 0; i < Nk; i++) pklin[i] = pk_lin0(r, log10(ka[i])) * g * g;
    clr_pk2xi(Nk, ka, pklin, ra, xilin);
    for (int kind = 0; kind < 3; kind++) {
      for (int ip = 0; ip < r->n_tr[kind]; ip++) {
        const double bias = clr_run_get_bg(r, rr, bz_tag[kind], ip);
        fprintf(fg, "%g ", bias);
        const double rsm = kind == 2 ? rsm2_mm : rsm2_gg;
        for (int i = 0; i < Nk; i++) pk[i] = pklin[i] * bias * bias * exp(-rsm * ka[i] * ka[i]);
        clr_pk2xi(Nk, ka, pk, ra, xi);
        for (int i = 0; i < Nk; i++) xi[i] = exp(xi[i]) - 1;
        clr_xi2pk(Nk, ra, xi, ka, pk);
        char fpk_name[CLR_PATH + 96], fxi_name[CLR_PATH + 96];
        snprintf(fpk_name, sizeof(fpk_name), "%s_pk_%s_pop%i_z%.3lf.txt", r->prefix_out, kind_name[kind], ip, z);
        snprintf(fxi_name, sizeof(fxi_name), "%s_xi_%s_pop%i_z%.3lf.txt", r->prefix_out, kind_name[kind], ip, z);
        FILE *fpk = fopen(fpk_name, "w"), *fxi = fopen(fxi_name, "w");
        if (!fpk || !fxi) {
          if (fpk) fclose(fpk);
          if (fxi) fclose(fxi);
          rc = 1;
          continue;
        }
        fprintf(fpk, "# k[h/Mpc] P_tt P_tl P_ll\n");
        fprintf(fxi, kind == 0 ? "# r[Mpc/h] xi_tt xi_ll*b^2 xi_ll\n" : "# k[Mpc/h] xi_tt xi_ll*b^2 xi_ll\n");
        for (int i = 0; i < Nk; i++) {
          if (ka[i] >= kmin && ka[i] <= kmax)
            fprintf(fpk, "%g %g %g %g\n", ka[i], pk[i], pklin[i] * bias * exp(-rsm2_gm * ka[i] * ka[i]),
                    pklin[i] * exp(-rsm2_mm * ka[i] * ka[i]));
          if (ra[i] >= rminout && ra[i] <= rmaxout) fprintf(fxi, "%g %g %g %g\n", ra[i], xi[i], xilin[i] * bias, xilin[i]);
        }
        fclose(fpk);
        fclose(fxi);
      }
    }
    fprintf(fg, "\n");
  }
  fclose(fg);
  free(ka);
  return rc;
}

/* ---- Limber power spectra (compute_cl, cosmo.c:114-224) ----------------- */
typedef struct {
  const clr_run *r;
  int l, type;
  double c0a, cfa, c0b, cfb, chis;
} limber_par;

static double bg_h(const clr_run *r, const double *f, double f0, double ff, double rr) {
  if (rr <= 0) return f0;
  if (rr >= r->r_arr_r2z[CLR_NA - 1]) return ff;
  const int ir = (int)(rr * r->glob_idr);
  return f[ir] + (f[ir + 1] - f[ir]) * (rr - r->r_arr_r2z[ir]) * r->glob_idr;
}

/* window_kappa_limber / window_isw_limber (cosmo.c:125-160) at chi_l = (l+1/2)/k */
static double limber_window(const limber_par *q, double k, double c0, double cf) {
  const clr_run *r = q->r;
  const double chi_l = (q->l + 0.5) / k;
  if (chi_l <= 0 || chi_l < c0 || chi_l > cf) return 0;
  if (q->type == CLR_CL_KAPPA) {
    const double gf = bg_h(r, r->growth_d_arr, 1, r->growth_d_arr[CLR_NA - 1], chi_l);
    const double aa = 1 / (1 + bg_h(r, r->z_arr_r2z, 0, r->z_arr_r2z[CLR_NA - 1], chi_l));
    return r->prefac_lensing * q->l * (q->l + 1.) * gf * (q->chis - chi_l) / (k * k * q->chis * chi_l * aa);
  }
  const double pd = bg_h(r, r->growth_pd_arr, r->growth_pd_arr[0], r->growth_pd_arr[CLR_NA - 1], chi_l);
  return -2 * r->prefac_lensing * pd / (k * k);
}

/* cl_integrand (cosmo.c:162-185): integrand in log10 k */
static double limber_integrand(double lk, void *vp) {
  const limber_par *q = (const limber_par *)vp;
  const clr_run *r = q->r;
  const double k = pow(10., lk);
  double pk = pk_lin0(r, lk);
  const double wa = limber_window(q, k, q->c0a, q->cfa), wb = limber_window(q, k, q->c0b, q->cfb);
  if (r->r_smooth > 0 && r->smooth_potential) pk *= exp(-k * k * r->r_smooth * r->r_smooth);
  return k * wa * wb * pk;
}

double clr_limber_cl(const clr_run *r, int l, double chi_0_a, double chi_f_a, double chi_0_b, double chi_f_b,
                     double chi_s, int type) {
  limber_par q = {r, l, type, chi_0_a, chi_f_a, chi_0_b, chi_f_b, chi_s};
  double chi_min = 1e16, chi_max = -1;
  const double c[4] = {chi_0_a, chi_0_b, chi_f_a, chi_f_b};
  for (int i = 0; i < 4; i++) {
    if (c[i] < chi_min) chi_min = c[i];
    if (c[i] > chi_max) chi_max = c[i];
  }
  const double lkmax = chi_min <= 0 ? 2. : log10((l + 0.5) / chi_min);
  const double lkmin = chi_max <= 0 ? -5. : log10((l + 0.5) / chi_max);
  /* The integrand vanishes unless chi_l = (l+1/2)/k lies in both windows:
   * integrate that part of [lkmin, lkmax] only (the same integral; an
   * adaptive rule over the whole range can miss a narrow overlap), adaptive
   * Gauss-Kronrod at 1e-9 relative (the reference asks QAG for 1e-4). */
  const double c_lo = fmax(chi_0_a, chi_0_b), c_hi = fmin(chi_f_a, chi_f_b);
  if (!(c_hi > c_lo) || c_hi <= 0) return 0;
  double a = c_hi > 0 ? log10((l + 0.5) / c_hi) : lkmin, b = c_lo > 0 ? log10((l + 0.5) / c_lo) : lkmax;
  if (a < lkmin) a = lkmin;
  if (b > lkmax) b = lkmax;
  if (!(b > a)) return 0;
  const double res = clr_integrate(limber_integrand, &q, a, b, 1e-9);
  return M_LN10 * res / (l + 0.5);
}

