/* cosmo.c -- FP64 background cosmology, r-tables and P(k) normalisation on
 * the host, following cosmo_mad.c:160-520 (background), cosmo.c:516-816
 * (cosmo_set), cosmo.c:444-495 (pk_linear_set) and cosmo.c:370-433 (sigma^2
 * of a smoothed field).  Integrals use adaptive Gauss-Kronrod (7,15) to a
 * relative 1e-11, tighter than the reference's GSL QNG/QAG tolerances
 * (1e-6 / 1e-4), so the tables agree with it to those tolerances. */
#include "clr_host.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CSM_HMPC 2997.92458
#define RTOD_REF 57.2957795
#define TWOPIPIINVLOGTEN 0.1166503235296796

/* ---- quadrature ---------------------------------------------------------- */
static const double gk_x[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                               0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                               0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                               0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
static const double gk_wk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                                0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                                0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                                0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double gk_wg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                                0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

static void gk15(double (*f)(double, void *), void *p, double a, double b, double *res, double *err) {
  const double c = 0.5 * (a + b), h = 0.5 * (b - a);
  const double fc = f(c, p);
  double rk = fc * gk_wk[7], rg = fc * gk_wg[3];
  for (int j = 0; j < 7; j++) {
    const double dx = h * gk_x[j];
    const double f1 = f(c - dx, p), f2 = f(c + dx, p);
    rk += gk_wk[j] * (f1 + f2);
    if (j % 2 == 1) rg += gk_wg[j / 2] * (f1 + f2);
  }
  *res = rk * h;
  *err = fabs((rk - rg) * h);
}

static double adapt(double (*f)(double, void *), void *p, double a, double b, double tol, int depth) {
  double r, e;
  gk15(f, p, a, b, &r, &e);
  if (e <= tol || depth > 40) return r;
  const double m = 0.5 * (a + b);
  return adapt(f, p, a, m, 0.5 * tol, depth + 1) + adapt(f, p, m, b, 0.5 * tol, depth + 1);
}

double clr_integrate(double (*f)(double, void *), void *p, double a, double b, double relerr) {
  if (a == b) return 0;
  double r, e;
  gk15(f, p, a, b, &r, &e);
  double tol = fabs(r) * relerr;
  if (tol == 0) tol = 1e-300;
  return adapt(f, p, a, b, tol, 0);
}

/* ---- background ------------------------------------------------------- */
static double den(const clr_csm *c, double a) {   /* OM + OL a^3 f_DE + OK a */
  if (c->normalDE) return c->OM + c->OL * a * a * a + c->OK * a;
  if (c->constantw) return c->OM + c->OL * pow(a, -3 * c->w0) + c->OK * a;
  return c->OM + c->OL * pow(a, -3 * (c->w0 + c->wa)) * exp(3 * (a - 1) * c->wa) + c->OK * a;
}
static double naHm1(double a, const clr_csm *c) { return sqrt(a / den(c, a)); }
static double f_na2Hm1(double a, void *p) { return 1 / sqrt(a * den((const clr_csm *)p, a)); }
static double f_naHm3(double a, void *p) { const double v = naHm1(a, (const clr_csm *)p); return v * v * v; }

double clr_csm_particle_horizon(const clr_csm *c, double aa) {
  const double alim = 0.01 * c->a_equality;
  if (aa <= alim) return 2 * sqrt(aa / c->OM) * CSM_HMPC;
  const double int0 = 2 * sqrt(alim / c->OM);
  return (int0 + clr_integrate(f_na2Hm1, (void *)c, alim, aa, 1e-11)) * CSM_HMPC;
}

double clr_csm_radial_comoving_distance(const clr_csm *c, double aa) {
  return c->phorizon - clr_csm_particle_horizon(c, aa);
}

double clr_csm_growth_factor(const clr_csm *c, double aa) {
  const double alim = 0.01 * c->a_equality;
  if (aa <= alim) return aa;
  const double int0 = 0.4 * sqrt(alim * alim * alim * alim * alim / (c->OM * c->OM * c->OM));
  const double I = clr_integrate(f_naHm3, (void *)c, alim, aa, 1e-11);
  return (int0 + I) * 2.5 * c->OM / (aa * naHm1(aa, c));
}

double clr_csm_hubble(const clr_csm *c, double aa) { return sqrt(den(c, aa) / (aa * aa * aa)) / CSM_HMPC; }

double clr_csm_omega_m(const clr_csm *c, double aa) { return c->OM / den(c, aa); }

double clr_csm_f_growth(const clr_csm *c, double aa) {
  const double Da = clr_csm_growth_factor(c, aa);
  double coeff, apow;
  if (c->normalDE) { coeff = 0; apow = aa * aa * aa; }
  else if (c->constantw) { coeff = 1 + c->w0; apow = pow(aa, -3 * c->w0); }
  else { coeff = 1 + c->w0 + c->wa * (1 - aa); apow = pow(aa, -3 * (c->w0 + c->wa)) * exp(3 * c->wa * (aa - 1)); }
  return 0.5 * (5 * c->OM * aa / Da - (3 * c->OM + 3 * coeff * c->OL * apow + 2 * c->OK * aa)) /
         (c->OM + c->OL * apow + c->OK * aa);
}

int clr_csm_set(clr_csm *c, double OmegaM, double OmegaL, double OmegaB, double w, double wa, double h,
                char *err, int errlen) {
  memset(c, 0, sizeof(*c));
  c->h = h; c->w0 = w; c->wa = wa; c->OM = OmegaM; c->OL = OmegaL; c->OK = 1 - OmegaM - OmegaL; c->OB = OmegaB;
  if (fabs(wa) < 1E-6) {
    c->constantw = 1;
    c->normalDE = fabs(w + 1) < 1E-6;
  }
  if (fabs(c->OK) < 1E-6) c->ksign = 0;
  else c->ksign = c->OK > 0 ? -1 : 1;
  if (c->OM <= 0) { snprintf(err, errlen, "Wrong matter parameter %.3lf", c->OM); return 1; }
  if (c->OM < c->OB) { snprintf(err, errlen, "Wrong M/B parameter %.3lf > %.3lf", c->OB, c->OM); return 1; }
  if (c->w0 > -0.333333) { snprintf(err, errlen, "DE is too exotic (w=%.3lf", c->w0); return 1; }
  /* aeqmin (cosmo_mad.c:160-177) */
  double aeqK = 1, aeqL = 1;
  if (c->ksign != 0) aeqK = c->OM / fabs(c->OK);
  if (c->OL != 0) aeqL = c->normalDE ? pow(c->OM / c->OL, 0.333) : pow(c->OM / c->OL, -1 / (3 * c->w0));
  c->a_equality = aeqK < aeqL ? aeqK : aeqL;
  c->phorizon = 0;
  c->phorizon = clr_csm_particle_horizon(c, 1);
  c->growth0 = clr_csm_growth_factor(c, 1);
  return 0;
}

/* ---- natural cubic spline ------------------------------------------------ */
void clr_cspline_eval(int n, const double *x, const double *y, int m, const double *xq, double *yq) {
  double *c2 = calloc(n, sizeof(double)), *g = malloc(n * sizeof(double)), *dd = malloc(n * sizeof(double));
  /* second derivatives with c2[0] = c2[n-1] = 0 (tridiagonal, Thomas) */
  for (int i = 1; i < n - 1; i++) {
    const double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
    const double diag = 2 * (h0 + h1);
    const double rhs = 6 * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
    const double sub = h0;
    if (i == 1) { g[i] = h1 / diag; dd[i] = rhs / diag; }
    else {
      const double den = diag - sub * g[i - 1];
      g[i] = h1 / den;
      dd[i] = (rhs - sub * dd[i - 1]) / den;
    }
  }
  for (int i = n - 2; i >= 1; i--) c2[i] = dd[i] - (i < n - 2 ? g[i] * c2[i + 1] : 0);
  for (int j = 0; j < m; j++) {
    const double t = xq[j];
    if (!(t >= x[0] && t <= x[n - 1])) { yq[j] = NAN; continue; }
    int lo = 0, hi = n - 1;
    while (hi > lo + 1) { const int mid = (lo + hi) / 2; if (x[mid] > t) hi = mid; else lo = mid; }
    const double h = x[hi] - x[lo];
    const double A = (x[hi] - t) / h, B = (t - x[lo]) / h;
    yq[j] = A * y[lo] + B * y[hi] + ((A * A * A - A) * c2[lo] + (B * B * B - B) * c2[hi]) * h * h / 6;
  }
  free(c2); free(g); free(dd);
}

/* ---- P(k) and smoothed variances ---------------------------------------- */
typedef struct { int numk; const double *logk, *pk; double logkmin, logkmax, idlogk, ns, R; int w; } s2par;

static double pk0(const s2par *s, double lgk) {   /* pk_linear0, cosmo.c:291-308 */
  const int ik = (int)((lgk - s->logkmin) * s->idlogk);
  if (ik < 0) return s->pk[0] * pow(10, s->ns * (lgk - s->logkmin));
  if (ik < s->numk - 1) return s->pk[ik] + (lgk - s->logk[ik]) * (s->pk[ik + 1] - s->pk[ik]) * s->idlogk;
  return s->pk[s->numk - 1] * pow(10, -3 * (lgk - s->logkmax));
}

static double wind(double x, int w) {   /* cosmo.c:251-289 */
  if (w == 1) return exp(-0.5 * x * x);
  if (x < 0.1)
    return 1. - 0.1 * x * x + 0.003571429 * x * x * x * x - 6.61376E-5 * x * x * x * x * x * x +
           7.51563E-7 * x * x * x * x * x * x * x * x;
  return 3 * (sin(x) - x * cos(x)) / (x * x * x);
}

static double s2_integrand(double logk, void *p) {   /* integxiL_NO at r=0 */
  const s2par *s = p;
  const double kk = pow(10, logk);
  const double wv = wind(kk * s->R, s->w);
  return TWOPIPIINVLOGTEN * pk0(s, logk) * kk * kk * kk * wv * wv;
}

double clr_sigma2_window(int numk, const double *logk, const double *pk, double logkmin, double logkmax,
                         double idlogk, double ns, double R, int w) {
  s2par s = {numk, logk, pk, logkmin, logkmax, idlogk, ns, R, w};
  /* QAGIL on (-inf, logkmax]: the k^(3+ns) tail below logkmin-12 is < 1e-36 of the total */
  double tot = 0;
  const double lo = logkmin - 12;
  const int nseg = 64;
  for (int i = 0; i < nseg; i++) {
    const double a = lo + (logkmax - lo) * i / nseg, b = lo + (logkmax - lo) * (i + 1) / nseg;
    tot += clr_integrate(s2_integrand, &s, a, b, 1e-10);
  }
  return tot;
}

static int read_two_columns(const char *fname, double **x, double **y, int *n, char *err, int errlen) {
  FILE *f = fopen(fname, "r");
  if (!f) { snprintf(err, errlen, "Couldn't open file %s", fname); return 1; }
  int nl = 0;
  char line[1000];
  while (fgets(line, sizeof(line), f)) nl++;
  rewind(f);
  *x = malloc((nl > 0 ? nl : 1) * sizeof(double));
  *y = malloc((nl > 0 ? nl : 1) * sizeof(double));
  for (int i = 0; i < nl; i++) {
    if (fscanf(f, "%lf %lf", &(*x)[i], &(*y)[i]) != 2) {
      snprintf(err, errlen, "Error reading file %s, line %d", fname, i + 1);
      fclose(f);
      return 1;
    }
  }
  fclose(f);
  *n = nl;
  return 0;
}

/* pk_linear_set, cosmo.c:444-495 */
static int pk_linear_set(clr_run *r, char *err, int errlen) {
  double *k, *p;
  int n;
  if (read_two_columns(r->pk_filename, &k, &p, &n, err, errlen)) return 1;
  if (n < 3) { snprintf(err, errlen, "P(k) file %s has %d lines", r->pk_filename, n); return 1; }
  r->numk = n;
  r->logkarr = malloc(n * sizeof(double));
  r->pkarr = malloc(n * sizeof(double));
  for (int i = 0; i < n; i++) { r->logkarr[i] = log10(k[i]); r->pkarr[i] = p[i]; }
  r->logkmin = r->logkarr[0];
  r->logkmax = r->logkarr[n - 1];
  r->idlogk = (n - 1) / (r->logkmax - r->logkmin);
  double *lk = malloc(n * sizeof(double)), *pk = malloc(n * sizeof(double));
  for (int i = 0; i < n - 1; i++) lk[i] = r->logkmin + i / r->idlogk;
  clr_cspline_eval(n, r->logkarr, r->pkarr, n - 1, lk, pk);
  for (int i = 0; i < n - 1; i++) { r->pkarr[i] = pk[i]; r->logkarr[i] = lk[i]; }
  free(lk); free(pk); free(k); free(p);
  const double s8 = clr_sigma2_window(n, r->logkarr, r->pkarr, r->logkmin, r->logkmax, r->idlogk, r->n_scal, 8, 0);
  r->sigma8_original = sqrt(s8);
  const double norm_pk = r->sig8 * r->sig8 / s8;
  for (int i = 0; i < n; i++) r->pkarr[i] *= norm_pk;
  const double reff = sqrt((r->r_smooth > 0 ? r->r_smooth * r->r_smooth : 0) + pow(0.45 * r->l_box / r->n_grid, 2));
  r->sigma2_gauss = clr_sigma2_window(n, r->logkarr, r->pkarr, r->logkmin, r->logkmax, r->idlogk, r->n_scal, reff, 1);
  return 0;
}

static double a_of_r_provisional(const clr_run *r, double rr) {   /* cosmo.c:88-99 */
  if (rr <= 0) return 1.;
  if (rr >= r->r_arr_a2r[0]) return 1E-6;
  int ia = 0;
  while (rr <= r->r_arr_a2r[ia]) ia++;
  return r->a_arr_a2r[ia - 1] + (r->a_arr_a2r[ia] - r->a_arr_a2r[ia - 1]) * (rr - r->r_arr_a2r[ia - 1]) /
                                    (r->r_arr_a2r[ia] - r->r_arr_a2r[ia - 1]);
}

/* cosmo_set, cosmo.c:516-716, and compute_tracer_cosmo, cosmo.c:818-849 */
int clr_cosmo_set(clr_run *r, char *err, int errlen) {
  if (clr_csm_set(&r->csm, r->OmegaM, r->OmegaL, r->OmegaB, r->weos, 0, r->hhub, err, errlen)) return 1;
  const clr_csm *c = &r->csm;
  r->fgrowth_0 = clr_csm_f_growth(c, 1);
  r->hubble_0 = clr_csm_hubble(c, 1);
  r->r_min = clr_csm_radial_comoving_distance(c, 1 / (1 + r->z_min));
  r->r_max = clr_csm_radial_comoving_distance(c, 1 / (1 + r->z_max));
  r->prefac_lensing = 1.5 * r->hubble_0 * r->hubble_0 * r->OmegaM;
  r->l_box = 2 * r->r_max * (1 + 2. / r->n_grid);
  for (int i = 0; i < 3; i++) r->pos_obs[i] = 0.5 * r->l_box;

  /* tracer splines: n(z) in deg^-2 -> comoving density (cosmo.c:566-592) */
  double *zs[3][CLR_NPOP_MAX], *fs[3][CLR_NPOP_MAX], *zb[3][CLR_NPOP_MAX], *fb[3][CLR_NPOP_MAX];
  int ns_[3][CLR_NPOP_MAX], nb_[3][CLR_NPOP_MAX];
  const char *what[3] = {"N(z)", "T(z)", "K(z)"};
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < r->n_tr[k]; i++) {
      if (read_two_columns(r->fname_bz[k][i], &zb[k][i], &fb[k][i], &nb_[k][i], err, errlen)) return 1;
      if (nb_[k][i] < 2 || zb[k][i][0] > r->z_min || zb[k][i][nb_[k][i] - 1] < r->z_max) {
        snprintf(err, errlen, "Bias z-range is too small");
        return 1;
      }
      if (read_two_columns(r->fname_amp[k][i], &zs[k][i], &fs[k][i], &ns_[k][i], err, errlen)) return 1;
      const int n = ns_[k][i];
      if (k == CLR_TRACER_SRCS) {
        for (int j = 0; j < n; j++) {
          const double a = 1. / (1 + zs[k][i][j]);
          const double hz = clr_csm_hubble(c, a), rz = clr_csm_radial_comoving_distance(c, a);
          fs[k][i][j] *= RTOD_REF * RTOD_REF * hz / (rz * rz);
        }
        if (n > 1 && zs[k][i][0] == 0) fs[k][i][0] = fs[k][i][1];
      }
      if (n < 2 || zs[k][i][0] > r->z_min || zs[k][i][n - 1] < r->z_max) {
        snprintf(err, errlen, "%s z-range is too small", what[k]);
        return 1;
      }
    }

  /* r-tables (cosmo.c:669-692) */
  for (int ii = 0; ii < CLR_NA; ii++) {
    const double a = ((double)ii) / (CLR_NA - 1);
    r->a_arr_a2r[ii] = a;
    r->r_arr_a2r[ii] = clr_csm_radial_comoving_distance(c, a);
  }
  const double growth0 = clr_csm_growth_factor(c, 1);
  r->glob_idr = (CLR_NA - 1) / r->r_arr_a2r[0];
  for (int ii = 0; ii < CLR_NA; ii++) {
    const double rr = ii / r->glob_idr;
    const double a = a_of_r_provisional(r, rr);
    const double z = 1. / a - 1;
    const double gz = clr_csm_growth_factor(c, a) / growth0;
    const double om = clr_csm_omega_m(c, a);
    const double fz = clr_csm_f_growth(c, a);
    const double hhz = clr_csm_hubble(c, a);
    r->z_arr_r2z[ii] = z;
    r->r_arr_r2z[ii] = rr;
    r->growth_d_arr[ii] = gz;
    r->growth_d2_arr[ii] = -0.42857142857 * gz * gz * pow(om, -0.00699300699);
    r->growth_v_arr[ii] = (gz * hhz * fz) / (r->fgrowth_0 * r->hubble_0);
    r->growth_pd_arr[ii] = gz * hhz * (fz - 1);
    r->ihub_arr[ii] = 1. / hhz;
  }
  /* tracer tables: spline values at z(r) (NaN outside the file's z-range, as
   * GSL's spline evaluation returns there, cosmo.c:693-716) */
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < r->n_tr[k]; i++) {
      r->amp_arr[k][i] = malloc(CLR_NA * sizeof(double));
      r->bz_arr[k][i] = malloc(CLR_NA * sizeof(double));
      clr_cspline_eval(ns_[k][i], zs[k][i], fs[k][i], CLR_NA, r->z_arr_r2z, r->amp_arr[k][i]);
      clr_cspline_eval(nb_[k][i], zb[k][i], fb[k][i], CLR_NA, r->z_arr_r2z, r->bz_arr[k][i]);
      free(zs[k][i]); free(fs[k][i]); free(zb[k][i]); free(fb[k][i]);
    }
  if (pk_linear_set(r, err, errlen)) return 1;

  for (int i = 0; i < r->n_kappa; i++) {
    if (r->z_kappa_out[i] > r->z_max) { snprintf(err, errlen, "Source plane %d outside redshift range", i + 1); return 1; }
    r->r_kappa[i] = clr_csm_radial_comoving_distance(c, 1. / (1 + r->z_kappa_out[i]));
  }
  for (int i = 0; i < r->n_isw; i++) {
    if (r->z_isw_out[i] > r->z_max) { snprintf(err, errlen, "Source plane %d outside redshift range", i + 1); return 1; }
    r->r_isw[i] = clr_csm_radial_comoving_distance(c, 1. / (1 + r->z_isw_out[i]));
  }
  /* imap shells from the frequency tables (compute_hp_shell_distances_imap,
   * cosmo.c:497-514) */
  for (int i = 0; i < r->n_tr[CLR_TRACER_IMAP]; i++) {
    double *nu0, *nuf;
    int n;
    if (read_two_columns(r->fname_nu[i], &nu0, &nuf, &n, err, errlen)) return 1;
    r->imap_nr[i] = n;
    r->imap_r0[i] = malloc(n * sizeof(double));
    r->imap_rf[i] = malloc(n * sizeof(double));
    for (int j = 0; j < n; j++) {
      r->imap_r0[i][j] = clr_csm_radial_comoving_distance(c, nuf[j] / r->nu0_imap[i]);
      r->imap_rf[i][j] = clr_csm_radial_comoving_distance(c, nu0[j] / r->nu0_imap[i]);
    }
    free(nu0); free(nuf);
  }
  return 0;
}

/* get_bg / f_of_r_linear (cosmo.c:30-86) and r_of_z (cosmo.c:101-112) */
static double f_of_r_h(const clr_run *r, const double *f, double f0, double ff, double rr) {
  if (rr <= 0) return f0;
  if (rr >= r->r_arr_r2z[CLR_NA - 1]) return ff;
  const int ir = (int)(rr * r->glob_idr);
  return f[ir] + (f[ir + 1] - f[ir]) * (rr - r->r_arr_r2z[ir]) * r->glob_idr;
}

double clr_run_get_bg(const clr_run *r, double rr, int tag, int ipop) {
  switch (tag) {
    case CLR_BG_Z: return f_of_r_h(r, r->z_arr_r2z, 0, r->z_arr_r2z[CLR_NA - 1], rr);
    case CLR_BG_D1: return f_of_r_h(r, r->growth_d_arr, 1, r->growth_d_arr[CLR_NA - 1], rr);
    case CLR_BG_V1: return f_of_r_h(r, r->growth_v_arr, r->growth_v_arr[0], r->growth_v_arr[CLR_NA - 1], rr);
    case CLR_BG_PD: return f_of_r_h(r, r->growth_pd_arr, r->growth_pd_arr[0], r->growth_pd_arr[CLR_NA - 1], rr);
    case CLR_BG_IH: return f_of_r_h(r, r->ihub_arr, r->ihub_arr[0], r->ihub_arr[CLR_NA - 1], rr);
    case CLR_BG_BZ_SRCS:
    case CLR_BG_BZ_IMAP:
    case CLR_BG_BZ_CSTM: {
      const int k = tag - CLR_BG_BZ_SRCS;
      if (ipop < 0 || ipop >= r->n_tr[k]) return NAN;
      return f_of_r_h(r, r->bz_arr[k][ipop], r->bz_arr[k][ipop][0], 1, rr);
    }
    default: return NAN;
  }
}

const double *clr_run_a2r(const clr_run *r) { return r->r_arr_a2r; }

double clr_run_r_of_z(const clr_run *r, double z) {
  const double a = 1. / (1 + z);
  if (a >= 1) return 0;
  if (a <= 0) return r->r_arr_a2r[0];
  const int ia = (int)(a * (CLR_NA - 1));
  return r->r_arr_a2r[ia] + (r->r_arr_a2r[ia + 1] - r->r_arr_a2r[ia]) * (a - r->a_arr_a2r[ia]) * (CLR_NA - 1.);
}
