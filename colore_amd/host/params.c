/* params.c -- read_run_params (io.c:246-563): the reference's param-file
 * keys, their mandatory/optional status and its error messages.  One new
 * optional key per tracer section, `bias_model` ("exponential" (1+d)^b,
 * "exp-truncated", "truncated", or 0/1/2), replaces the reference's
 * compile-time bias model; it defaults to "exp-truncated", the model the
 * reference's Makefile selects (Makefile:12). */
#include "clr_host.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static int need(int ok, const char *sec, const char *var, char *err, int errlen) {
  if (!ok) snprintf(err, errlen, "Couldn't read variable %s.%s", sec, var);
  return !ok;
}

#define RD(kind, sec, var, dst)                                                              \
  do {                                                                                       \
    char p_[300];                                                                            \
    snprintf(p_, sizeof(p_), "%s.%s", sec, var);                                             \
    if (need(clr_cfg_##kind(cfg, p_, dst), sec, var, err, errlen)) goto fail;                \
  } while (0)
#define RDS(sec, var, dst)                                                                   \
  do {                                                                                       \
    char p_[300];                                                                            \
    snprintf(p_, sizeof(p_), "%s.%s", sec, var);                                             \
    if (need(clr_cfg_string(cfg, p_, dst, CLR_PATH), sec, var, err, errlen)) goto fail;      \
  } while (0)

static int read_bias_model(const clr_cfg *cfg, const char *sec, int *model, char *err, int errlen) {
  char p[300], s[64];
  int iv;
  snprintf(p, sizeof(p), "%s.bias_model", sec);
  *model = CLR_BIAS_EXP_TRUNCATED;
  if (!clr_cfg_has(cfg, p)) return 0;
  if (clr_cfg_int(cfg, p, &iv)) {
    if (iv < 0 || iv > 2) { snprintf(err, errlen, "Unknown bias model %d in %s", iv, sec); return 1; }
    *model = iv;
    return 0;
  }
  if (!clr_cfg_string(cfg, p, s, sizeof(s))) { snprintf(err, errlen, "Couldn't read variable %s", p); return 1; }
  if (!strcmp(s, "exponential")) *model = CLR_BIAS_EXPONENTIAL;
  else if (!strcmp(s, "exp-truncated")) *model = CLR_BIAS_EXP_TRUNCATED;
  else if (!strcmp(s, "truncated")) *model = CLR_BIAS_TRUNCATED;
  else { snprintf(err, errlen, "Unknown bias model %s in %s", s, sec); return 1; }
  return 0;
}

clr_run *clr_run_read(const char *fname, char *err, int errlen) {
  clr_run *r = calloc(1, sizeof(clr_run));
  clr_cfg *cfg = clr_cfg_read(fname, err, errlen);
  char sec[64], fmt[64];
  if (!cfg) { free(r); return NULL; }
  /* param_colore_new defaults (io.c:24-158) that the keys below override */
  r->OmegaM = 0.3; r->OmegaL = 0.7; r->OmegaB = 0.05; r->hhub = 0.7; r->weos = -1; r->n_scal = 0.96;
  r->sig8 = 0.83; r->n_grid = 512; r->seed = 1234;
  r->nside_base = 2;   /* choose_nside_base on one task (io.c:224-244) */

  RDS("global", "prefix_out", r->prefix_out);
  RDS("global", "pk_filename", r->pk_filename);
  RD(double, "global", "z_min", &r->z_min);
  RD(double, "global", "z_max", &r->z_max);
  RD(bool, "global", "output_density", &r->output_density);
  RD(int, "global", "seed", &r->seed);
  RD(bool, "global", "write_pred", &r->do_pred);
  if (r->do_pred) {
    RD(bool, "global", "just_write_pred", &r->just_do_pred);
    RD(double, "global", "pred_dz", &r->pred_dz);
  }
  RD(double, "field_par", "r_smooth", &r->r_smooth);
  RD(bool, "field_par", "smooth_potential", &r->smooth_potential);
  RD(int, "field_par", "n_grid", &r->n_grid);
  RD(int, "field_par", "dens_type", &r->dens_type);
  RD(double, "field_par", "lpt_buffer_fraction", &r->lpt_buffer_fraction);
  RD(int, "field_par", "lpt_interp_type", &r->lpt_interp_type);
  RD(int, "field_par", "output_lpt", &r->output_lpt);
  {
    char p[300];
    snprintf(p, sizeof(p), "global.output_format");
    if (need(clr_cfg_string(cfg, p, fmt, sizeof(fmt)), "global", "output_format", err, errlen)) goto fail;
    if (!strcmp(fmt, "HDF5")) { snprintf(err, errlen, "HDF5 format not supported"); goto fail; }
    else if (!strcmp(fmt, "FITS")) r->output_format = 1;
    else if (!strcmp(fmt, "ASCII")) r->output_format = 0;
    else { snprintf(err, errlen, "Unrecognized format %s", fmt); goto fail; }
  }
  RD(double, "cosmo_par", "omega_M", &r->OmegaM);
  RD(double, "cosmo_par", "omega_L", &r->OmegaL);
  RD(double, "cosmo_par", "omega_B", &r->OmegaB);
  RD(double, "cosmo_par", "h", &r->hhub);
  RD(double, "cosmo_par", "w", &r->weos);
  RD(double, "cosmo_par", "ns", &r->n_scal);
  RD(double, "cosmo_par", "sigma_8", &r->sig8);

  /* srcsN (io.c:315-345) */
  while (1) {
    snprintf(sec, sizeof(sec), "srcs%d", r->n_tr[0] + 1);
    if (!clr_cfg_has(cfg, sec)) break;
    const int i = r->n_tr[0];
    if (i >= CLR_NPOP_MAX) { snprintf(err, errlen, "You're asking for too many populations %d! Enlarge NPOP_MAX", i + 1); goto fail; }
    RDS(sec, "nz_filename", r->fname_amp[0][i]);
    RDS(sec, "bias_filename", r->fname_bz[0][i]);
    RD(bool, sec, "include_lensing", &r->lensing_srcs[i]);
    RD(bool, sec, "store_skewers", &r->skw_srcs[i]);
    if (r->skw_srcs[i]) {
      char p[300];
      snprintf(p, sizeof(p), "%s.gaussian_skewers", sec);
      clr_cfg_bool(cfg, p, &r->skw_gauss[i]);
    }
    if (read_bias_model(cfg, sec, &r->bias_model[0][i], err, errlen)) goto fail;
    r->n_tr[0]++;
  }
  /* customN (io.c:347-366) */
  while (1) {
    snprintf(sec, sizeof(sec), "custom%d", r->n_tr[2] + 1);
    if (!clr_cfg_has(cfg, sec)) break;
    const int i = r->n_tr[2];
    if (i >= CLR_NPOP_MAX) { snprintf(err, errlen, "You're asking for too many populations %d! Enlarge NPOP_MAX", i + 1); goto fail; }
    RDS(sec, "kz_filename", r->fname_amp[2][i]);
    RDS(sec, "bias_filename", r->fname_bz[2][i]);
    RD(int, sec, "nside", &r->nside_cstm[i]);
    if (read_bias_model(cfg, sec, &r->bias_model[2][i], err, errlen)) goto fail;
    r->n_tr[2]++;
  }
  /* imapN (io.c:368-388) */
  while (1) {
    snprintf(sec, sizeof(sec), "imap%d", r->n_tr[1] + 1);
    if (!clr_cfg_has(cfg, sec)) break;
    const int i = r->n_tr[1];
    if (i >= CLR_NPOP_MAX) { snprintf(err, errlen, "You're asking for too many populations %d! Enlarge NPOP_MAX", i + 1); goto fail; }
    RDS(sec, "tbak_filename", r->fname_amp[1][i]);
    RDS(sec, "bias_filename", r->fname_bz[1][i]);
    RDS(sec, "freq_list", r->fname_nu[i]);
    RD(double, sec, "freq_rest", &r->nu0_imap[i]);
    RD(int, sec, "nside", &r->nside_imap[i]);
    if (read_bias_model(cfg, sec, &r->bias_model[1][i], err, errlen)) goto fail;
    r->n_tr[1]++;
  }
  if (clr_cfg_has(cfg, "kappa")) {
    r->do_kappa = 1;
    r->n_kappa = clr_cfg_double_array(cfg, "kappa.z_out", r->z_kappa_out, CLR_NPLANES_MAX);
    if (r->n_kappa == -2) { snprintf(err, errlen, "Too many elements in kappa.z_out (> %d)", CLR_NPLANES_MAX); goto fail; }
    if (r->n_kappa <= 0) { snprintf(err, errlen, "Couldn't read variable kappa.z_out"); goto fail; }
    RD(int, "kappa", "nside", &r->nside_kappa);
  }
  if (clr_cfg_has(cfg, "isw")) {
    r->do_isw = 1;
    r->n_isw = clr_cfg_double_array(cfg, "isw.z_out", r->z_isw_out, CLR_NPLANES_MAX);
    if (r->n_isw == -2) { snprintf(err, errlen, "Too many elements in isw.z_out (> %d)", CLR_NPLANES_MAX); goto fail; }
    if (r->n_isw <= 0) { snprintf(err, errlen, "Couldn't read variable isw.z_out"); goto fail; }
    RD(int, "isw", "nside", &r->nside_isw);
  }
  clr_cfg_free(cfg);
  cfg = NULL;

  if (r->dens_type < 0 || r->dens_type > 3) {
    snprintf(err, errlen, "Density type %d not supported", r->dens_type);
    goto fail;
  }
  if ((r->dens_type == 1 || r->dens_type == 2) && (r->lpt_interp_type < 0 || r->lpt_interp_type > 2)) {
    snprintf(err, errlen, "Wrong interpolation type");
    goto fail;
  }
  const int nsides[3] = {r->nside_kappa, r->nside_isw, 0};
  for (int i = 0; i < 2; i++)
    if ((i == 0 ? r->do_kappa : r->do_isw) && (nsides[i] < r->nside_base || nsides[i] > 8192)) {
      snprintf(err, errlen, nsides[i] > 8192 ? "Can't go beyond nside=%d" : "Can't go below nside=%d",
               nsides[i] > 8192 ? 8192 : r->nside_base);
      goto fail;
    }
  if (clr_cosmo_set(r, err, errlen)) goto fail;
  return r;
fail:
  clr_cfg_free(cfg);
  clr_run_free(r);
  return NULL;
}

void clr_run_free(clr_run *r) {
  if (!r) return;
  free(r->logkarr);
  free(r->pkarr);
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < CLR_NPOP_MAX; i++) { free(r->amp_arr[k][i]); free(r->bz_arr[k][i]); }
  for (int i = 0; i < CLR_NPOP_MAX; i++) { free(r->imap_r0[i]); free(r->imap_rf[i]); }
  free(r);
}

double clr_run_scalar(const clr_run *r, const char *n) {
#define S(key, val) if (!strcmp(n, key)) return (double)(val)
  S("l_box", r->l_box); S("r_max", r->r_max); S("r_min", r->r_min); S("z_min", r->z_min); S("z_max", r->z_max);
  S("n_grid", r->n_grid); S("seed", r->seed); S("r_smooth", r->r_smooth);
  S("smooth_potential", r->smooth_potential); S("dens_type", r->dens_type); S("sigma2_gauss", r->sigma2_gauss);
  S("n_srcs", r->n_tr[0]); S("n_imap", r->n_tr[1]); S("n_cstm", r->n_tr[2]); S("n_kappa", r->n_kappa);
  S("n_isw", r->n_isw); S("nside_base", r->nside_base); S("nside_kappa", r->nside_kappa);
  S("nside_isw", r->nside_isw); S("fgrowth_0", r->fgrowth_0); S("hubble_0", r->hubble_0);
  S("prefac_lensing", r->prefac_lensing); S("sig8", r->sig8); S("output_format", r->output_format);
  S("output_density", r->output_density); S("sigma8_original", r->sigma8_original);
  S("a_equality", r->csm.a_equality); S("phorizon", r->csm.phorizon); S("growth0", r->csm.growth0);
  S("omega_m", r->OmegaM);
#undef S
  return NAN;
}

int clr_run_imap_shells(const clr_run *r, int ipop, int *nside, int *nr, const double **r0, const double **rf) {
  if (ipop < 0 || ipop >= r->n_tr[1]) return 1;
  *nside = r->nside_imap[ipop];
  *nr = r->imap_nr[ipop];
  *r0 = r->imap_r0[ipop];
  *rf = r->imap_rf[ipop];
  return 0;
}

int clr_run_tables(const clr_run *r, clr_tables *t) {
  memset(t, 0, sizeof(*t));
  t->na = CLR_NA;
  t->glob_idr = r->glob_idr;
  t->r_arr_r2z = r->r_arr_r2z;
  t->z_arr_r2z = r->z_arr_r2z;
  t->growth_d_arr = r->growth_d_arr;
  t->growth_v_arr = r->growth_v_arr;
  t->growth_pd_arr = r->growth_pd_arr;
  t->ihub_arr = r->ihub_arr;
  t->numk = r->numk;
  t->logkmin = r->logkmin;
  t->logkmax = r->logkmax;
  t->idlogk = r->idlogk;
  t->logkarr = r->logkarr;
  t->pkarr = r->pkarr;
  t->n_scal = r->n_scal;
  t->fgrowth_0 = r->fgrowth_0;
  t->hubble_0 = r->hubble_0;
  t->omega_m = r->OmegaM;
  t->prefac_lensing = r->prefac_lensing;
  t->r_max = r->r_max;
  t->growth_d2_arr = r->growth_d2_arr;
  return 0;
}

int clr_run_tracer(const clr_run *r, int kind, int ipop, const double **amp, const double **bz, int *model) {
  if (kind < 0 || kind > 2 || ipop < 0 || ipop >= r->n_tr[kind]) return CLR_ERR_ARG;
  if (amp) *amp = r->amp_arr[kind][ipop];
  if (bz) *bz = r->bz_arr[kind][ipop];
  if (model) *model = r->bias_model[kind][ipop];
  return 0;
}

const char *clr_run_prefix(const clr_run *r) { return r->prefix_out; }

int clr_run_device(clr_run *r, int device, clr_ctx **ctx, char *err, int errlen) {
  clr_tables t;
  int rc = clr_create(ctx, r->n_grid, r->l_box, r->pos_obs, device);
  if (rc) { snprintf(err, errlen, "%s", clr_last_error(*ctx)); return rc; }
  clr_run_tables(r, &t);
  rc = clr_upload_tables(*ctx, &t);
  for (int k = 0; k < 3 && !rc; k++)
    for (int i = 0; i < r->n_tr[k] && !rc; i++)
      rc = clr_set_tracer(*ctx, k, i, r->bias_model[k][i], r->amp_arr[k][i], r->bz_arr[k][i]);
  if (rc) snprintf(err, errlen, "%s", clr_last_error(*ctx));
  return rc;
}
