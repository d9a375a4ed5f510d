/*
 * colore_host.h -- host side of the drop-in `CoLoRe <param_file>` driver:
 * param-file parsing (the reference's libconfig keys, io.c:246-563), FP64
 * cosmology and r-tables (cosmo.c, cosmo_mad.c), HEALPix helpers and the
 * output writers (io.c:565-1245).  Plain C; calls the device through
 * colore_amd.h.
 */
#ifndef COLORE_HOST_H
#define COLORE_HOST_H
#include <stdint.h>
#include "colore_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- background cosmology (cosmo_mad.c:160-520) ------------------------- */
typedef struct {
  double OM, OB, OL, OK, h, w0, wa;
  int ksign, normalDE, constantw;
  double phorizon, a_equality, growth0;
} clr_csm;
/* csm_background_set; returns 0, or >0 with a message for invalid input */
int    clr_csm_set(clr_csm *c, double OmegaM, double OmegaL, double OmegaB, double w, double wa,
                   double h, char *err, int errlen);
double clr_csm_hubble(const clr_csm *c, double a);              /* H(a) in h/Mpc      */
double clr_csm_omega_m(const clr_csm *c, double a);
double clr_csm_particle_horizon(const clr_csm *c, double a);    /* Mpc/h              */
double clr_csm_radial_comoving_distance(const clr_csm *c, double a);
double clr_csm_growth_factor(const clr_csm *c, double a);       /* unnormalised D(a)  */
double clr_csm_f_growth(const clr_csm *c, double a);

/* natural cubic spline (gsl_interp_cspline): NaN outside [x[0], x[n-1]] */
void   clr_cspline_eval(int n, const double *x, const double *y, int m, const double *xq, double *yq);
/* adaptive Gauss-Kronrod (7/15) integral of f on [a,b] */
double clr_integrate(double (*f)(double, void *), void *p, double a, double b, double relerr);
/* sigma^2 of the linear field smoothed with window w (0 TopHat, 1 Gauss) of
 * radius R from a P(k) table (xi2p_L at r=0, cosmo.c:370-433) */
double clr_sigma2_window(int numk, const double *logk, const double *pk, double logkmin, double logkmax,
                         double idlogk, double ns, double R, int w);

/* ---- run state: the reference's ParamCoLoRe subset used by this path ---- */
typedef struct clr_run clr_run;

/* read_run_params (io.c:246-563): parse, set up cosmology and tables.
 * Returns NULL on error with the message in err. */
clr_run *clr_run_read(const char *param_file, char *err, int errlen);
void     clr_run_free(clr_run *r);
/* scalar by name: "l_box","r_max","r_min","z_min","z_max","n_grid","seed",
 * "r_smooth","smooth_potential","dens_type","sigma2_gauss","n_srcs","n_imap",
 * "n_cstm","n_kappa","n_isw","nside_base","nside_kappa","nside_isw",
 * "fgrowth_0","hubble_0","prefac_lensing","sig8","output_format",
 * "output_density","sigma8_original"; NaN if unknown. */
double   clr_run_scalar(const clr_run *r, const char *name);
/* Points the clr_tables fields at the run's arrays (no copy). */
int      clr_run_tables(const clr_run *r, clr_tables *out);
/* Tracer r-tables of tracer kind/ipop (pointers into the run). */
int      clr_run_tracer(const clr_run *r, int kind, int ipop, const double **amp, const double **bz,
                        int *bias_model);
/* Shells of intensity species ipop (imap_preproc, imap.c:105-122: comoving
 * r0/rf from the frequency table, pointers into the run): nside and nr. */
int      clr_run_imap_shells(const clr_run *r, int ipop, int *nside, int *nr, const double **r0,
                             const double **rf);
/* Output prefix. */
const char *clr_run_prefix(const clr_run *r);
/* get_bg / r_of_z on the run's tables (cosmo.c:40-112). */
double   clr_run_get_bg(const clr_run *r, double rr, int tag, int ipop);
double   clr_run_r_of_z(const clr_run *r, double z);
/* r_arr_a2r: CLR_NA values of r at a = i/(CLR_NA-1) (cosmo.c:669-675). */
const double *clr_run_a2r(const clr_run *r);
#define CLR_BG_Z  0
#define CLR_BG_D1 1
#define CLR_BG_V1 2
#define CLR_BG_PD 3
#define CLR_BG_IH 4
#define CLR_BG_BZ_SRCS 5   /* tracer bias tables: edges b[0] and 1 (cosmo.c:60-79) */
#define CLR_BG_BZ_IMAP 6
#define CLR_BG_BZ_CSTM 7

/* Creates the device context and uploads every table of the run. */
int      clr_run_device(clr_run *r, int device, clr_ctx **ctx, char *err, int errlen);
/* The stage sequence of main.c:61-141 on the device, with the reference's
 * stdout report and output files.  Returns 0 or an error code. */
int      clr_run_execute(clr_run *r, int device);

/* ---- HEALPix (Gorski et al. 2005) ----------------------------------------- */
long long clr_hp_ang2pix_ring(long long nside, double cth, double phi);
long long clr_hp_vec2pix_ring(long long nside, const double *v);
long long clr_hp_ring2nest(long long nside, long long pix);
long long clr_hp_nest2ring(long long nside, long long pix);
void      clr_hp_pix2vec_nest(long long nside, long long pix, double *v);

/* ---- theory predictions (predictions.c, fftlog.c) ------------------------ */
/* FFTLog (Hamilton 2000) of l = 0 on n log-spaced points: xi(r) from P(k) and
 * back (pk2xi / xi2pk, fftlog.c:137-157); the output grid r (or k) is written. */
void clr_pk2xi(int n, const double *k, const double *pk, double *r, double *xi);
void clr_xi2pk(int n, const double *r, const double *xi, double *k, double *pk);
/* write_predictions (predictions.c:26-172): <prefix>_gbias.txt and the
 * lognormal P(k) / xi(r) predictions of every tracer every pred_dz in z. */
int  clr_write_predictions(const clr_run *r);
/* compute_cl (cosmo.c:114-224): the Limber angular power spectrum at
 * multipole l of two radial windows [chi_0_a, chi_f_a] and [chi_0_b, chi_f_b]
 * for a source plane at chi_s, of the convergence (CLR_CL_KAPPA) or of the
 * ISW temperature (CLR_CL_ISW), from the linear P(k) table (smoothed by
 * exp(-k^2 r_smooth^2) when the potential is smoothed). */
#define CLR_CL_KAPPA 0
#define CLR_CL_ISW   2
double clr_limber_cl(const clr_run *r, int l, double chi_0_a, double chi_f_a, double chi_0_b,
                     double chi_f_b, double chi_s, int type);

/* ---- writers ------------------------------------------------------------ */
/* write_density_grid native format (io.c:565-595) from a padded grid */
int clr_write_density_grid(const char *fname, int N, double L, const float *padded);
/* he_write_healpix_map (healpix_extra.c:4-57): one 1E column, RING */
int clr_write_healpix_map(const char *fname, long long nside, const float *map_ring);
/* write_catalog FITS / ASCII (io.c:1019-1236), columns TYPE RA DEC
 * Z_COSMO DZ_RSD from radecz = float[4n] (RA, Dec, z0, dz_rsd), and, when
 * lens is not NULL (has_lensing), E1 E2 KAPPA DRA DDEC from lens = float[5n]. */
int clr_write_catalog_fits(const char *fname, int ipop, long long n, const float *radecz,
                           const float *lens);
int clr_write_catalog_ascii(const char *fname, int ipop, long long n, const float *radecz,
                            const float *lens);
/* write_lpt (io.c:597-695): Gadget-2 snapshot <fname> of n LPT particles
 * (xyz = 3 floats each; header mass 27.7455 Om L^3 / n, zero velocities,
 * IDs 0..n-1). */
int clr_write_lpt_gadget(const char *fname, long long n, double l_box, double omega_m, double omega_l,
                         double hhub, const float *xyz);
/* The skewer HDUs of a FITS catalogue (io.c:1139-1187), appended to fname:
 * an n x nr image of densities (or Gaussian fields), one of velocities, and
 * a table R Z D V (bg4 = float[4 nr]: r, z, D, V1 at each step) whose
 * header holds SIGMA_G = sqrt(sigma2_gauss). */
int clr_append_skewers_fits(const char *fname, int gaussian, long long n, int nr,
                            const float *dens_or_gauss, const float *v, const float *bg4,
                            float sigma_g);

#ifdef __cplusplus
}
#endif
#endif
