/* clr_host.h -- run state of the host driver: the fields of the reference's
 * ParamCoLoRe (src/common.h:220-381) that the field -> catalogue path and its
 * lightcone painters use. */
#ifndef CLR_HOST_INTERNAL_H
#define CLR_HOST_INTERNAL_H
#include <stdio.h>
#include "../../include/colore_host.h"

#define CLR_NPLANES_MAX 100
#define CLR_PATH 512

struct clr_run {
  /* global */
  char prefix_out[CLR_PATH];
  char pk_filename[CLR_PATH];
  double z_min, z_max;
  int output_density, seed, do_pred, just_do_pred, output_format;  /* 0 ASCII 1 FITS */
  double pred_dz;
  /* field_par */
  double r_smooth;          /* length as read (the reference squares it) */
  int smooth_potential, n_grid, dens_type, lpt_interp_type, output_lpt;
  double lpt_buffer_fraction;
  /* cosmo_par */
  double OmegaM, OmegaL, OmegaB, hhub, weos, n_scal, sig8;
  /* derived */
  clr_csm csm;
  double fgrowth_0, hubble_0, prefac_lensing, r_min, r_max, l_box, pos_obs[3];
  double sigma2_gauss, sigma8_original;
  double a_arr_a2r[CLR_NA], r_arr_a2r[CLR_NA];
  double z_arr_r2z[CLR_NA], r_arr_r2z[CLR_NA];
  double growth_d_arr[CLR_NA], growth_d2_arr[CLR_NA], growth_v_arr[CLR_NA];
  double growth_pd_arr[CLR_NA], ihub_arr[CLR_NA];
  double glob_idr;
  int numk;
  double logkmin, logkmax, idlogk;
  double *logkarr, *pkarr;
  /* tracers: [kind][ipop] */
  int n_tr[3];
  char fname_amp[3][CLR_NPOP_MAX][CLR_PATH];   /* nz / tbak / kz */
  char fname_bz[3][CLR_NPOP_MAX][CLR_PATH];
  int bias_model[3][CLR_NPOP_MAX];
  double *amp_arr[3][CLR_NPOP_MAX], *bz_arr[3][CLR_NPOP_MAX];
  int lensing_srcs[CLR_NPOP_MAX], skw_srcs[CLR_NPOP_MAX], skw_gauss[CLR_NPOP_MAX];
  char fname_nu[CLR_NPOP_MAX][CLR_PATH];
  double nu0_imap[CLR_NPOP_MAX];
  int nside_imap[CLR_NPOP_MAX], nside_cstm[CLR_NPOP_MAX];
  int imap_nr[CLR_NPOP_MAX];
  double *imap_r0[CLR_NPOP_MAX], *imap_rf[CLR_NPOP_MAX];
  /* kappa / isw */
  int do_kappa, n_kappa, nside_kappa, do_isw, n_isw, nside_isw;
  double z_kappa_out[CLR_NPLANES_MAX], z_isw_out[CLR_NPLANES_MAX];
  double r_kappa[CLR_NPLANES_MAX], r_isw[CLR_NPLANES_MAX];
  int nside_base;
};

/* cosmo.c */
int  clr_cosmo_set(clr_run *r, char *err, int errlen);
/* config.c: libconfig-subset reader */
typedef struct clr_cfg clr_cfg;
clr_cfg *clr_cfg_read(const char *fname, char *err, int errlen);
void clr_cfg_free(clr_cfg *c);
int clr_cfg_has(const clr_cfg *c, const char *path);
int clr_cfg_string(const clr_cfg *c, const char *path, char *out, int outlen);
int clr_cfg_double(const clr_cfg *c, const char *path, double *out);
int clr_cfg_int(const clr_cfg *c, const char *path, int *out);
int clr_cfg_bool(const clr_cfg *c, const char *path, int *out);
int clr_cfg_double_array(const clr_cfg *c, const char *path, double *out, int nmax);  /* -> count or <0 */
/* io.c */
void clr_timer(int i);
void clr_print_info(const char *fmt, ...);

#endif
