/*
 * oracle.h -- CPU FP64 restatement of CoLoRe's field -> lognormal -> sources
 * path and its lightcone painters.  TEST INFRASTRUCTURE ONLY: used by tests/,
 * __graft_entry__.smoke() and bench.py's CPU baseline as the checker, never by
 * the product path.
 *
 * Every function follows the reference source line by line (file:line in the
 * comments of oracle.c).  Random numbers come from an injectable stream:
 * either Philox4x32-10 counters (the production streams of the GPU path) or,
 * for the Fourier modes, a caller-supplied buffer (e.g. the reference's
 * per-thread MT19937 streams, orc_ref_mode_stream).
 *
 * Pins (tests/test_oracle.py, tests/test_host.py): known-answer vectors
 * (Philox, MT19937 raw stream), numpy/scipy routines (FFTs, r2c, splines,
 * quadrature), closed forms (modes, lognormal/clip, bias models, kappa/ISW,
 * source beams, skewers, custom maps, 1LPT/2LPT displacements, imap at
 * constant T), invariants and brute force (shell sums, HEALPix, deposits).
 * Parity unpinned (no library or reference build here to pin against):
 * the BTPE branch of GSL's binomial (n p >= 14) inside orc_ref_poisson's
 * large-mu path (inversion is used there; the multiplication method for
 * mu <= 10, gamma_int and the small-n p binomial are restated and pinned by
 * hand-worked streams and distribution tests), GSL's MT19937 uniform mapping
 * (restated from its documentation; the raw stream is pinned), FFTW's
 * treatment of the non-Hermitian kx = 0 and N/2 planes (numpy's convention
 * only), chealpix's exact code paths (closed forms only).  orc_poisson (the
 * Philox production sampler's checker) uses inversion, pinned statistically.
 */
#ifndef COLORE_ORACLE_H
#define COLORE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random numbers ---------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

typedef struct { uint32_t mt[624]; int mti; } orc_mt19937;
void orc_mt_seed(orc_mt19937 *s, uint32_t seed);     /* gsl_rng_set(mt19937) */
uint32_t orc_mt_next(orc_mt19937 *s);                /* gsl_rng_get          */
/* The reference's mode stream (fourier.c:295-357): thread ithr of nthreads
 * owns a contiguous block of kz planes (static schedule as libgomp splits it),
 * seeds MT19937 with seed+ithr and draws (phase, modulus) per mode with
 * k^2>0 in loop order.  out[2*m], out[2*m+1] for mode m = kx+(N/2+1)(ky+N kz);
 * the k=0 mode gets zeros. */
void orc_ref_mode_stream(int N, uint32_t seed, int nthreads, uint32_t *out);

/* ---- tables ------------------------------------------------------------ */
typedef struct {
  int na; double idr;
  const double *r, *z, *d1, *v1, *pd, *ih;
} orc_bg;
typedef struct {
  int numk; double logkmin, logkmax, idlogk, ns;
  const double *logk, *pk;
} orc_pk;
double orc_f_of_r(const orc_bg *bg, const double *f, double f0, double ff, double r);
double orc_pk_linear0(const orc_pk *pk, double lgk);
double orc_bias_model(double d, double b, int model);

/* ---- fields (fourier.c) ------------------------------------------------ */
/* create_grids_fourier: half spectra as interleaved complex [N][N][N/2+1]. */
void orc_gen_modes(int N, double L, const orc_pk *pk, uint32_t seed, const uint32_t *inj,
                   double r_smooth, int smooth_potential, double prefac_lensing,
                   double *dens_k, double *npot_k);
/* <delta^2> and <phi^2> of the normalised real grids from the modes alone
 * (Parseval with the c2r's handling of the kx = 0, N/2 planes), any N. */
void orc_kspace_variance(int N, double L, const orc_pk *pk, uint32_t seed, const uint32_t *inj,
                         double r_smooth, int smooth_potential, double prefac_lensing, double out[2]);
/* 3-D c2r, FFTW convention (complex IDFT over z and y, then c2r along x using
 * Re of kx=0 and kx=N/2), unnormalised.  out: padded [N][N][2(N/2+1)]. */
void orc_c2r_3d(int N, const double *in, double *out);
/* Same result by direct summation, O(N^5): for N <= 16 only. */
void orc_c2r_3d_brute(int N, const double *in, double *out);
/* 1-D complex inverse DFT (sign +1), in place, n = 2^a 3^b 5^c. */
void orc_cfft_inv(int n, double *data);
/* Normalisation (fourier.c:383-398) of both grids and compute_sigma_dens
 * (fourier.c:24-79); returns sigma^2 (mean and <d^2> in out2 if non-NULL). */
double orc_normalize_sigma2(int N, double L, double *dens, double *npot, double *out2);

/* ---- density (density.c) ----------------------------------------------- */
void orc_density(int N, double L, const double obs[3], const orc_bg *bg, double sigma2,
                 int dens_type, double *grid);
/* collect_density_normalization_from_grid: returns nz; arrays of capacity cap.
 * sums[t*nz + iz]. */
int orc_shell_sums(int N, double L, const double obs[3], const orc_bg *bg, const double *grid,
                   int ntr, const double *const *bz, const double *bz0, const int *model, int cap,
                   double *zmax_out, unsigned long long *cnt, double *zsum, double *sums);
/* compute_density_normalization host part (density.c:1272-1355) for one
 * tracer: norm table on the NA r-grid. */
void orc_norm_table(const orc_bg *bg, int nz, double zmax, const unsigned long long *cnt,
                    const double *zsum, const double *sums_t, double *norm_arr, double *norm0,
                    double *normf);

/* ---- sources (srcs.c) --------------------------------------------------- */
typedef struct {
  int N; double L; double obs[3];
  const orc_bg *bg;
  const double *amp, *bz, *norm;     /* n(r), b(r), norm(r) tables (NA) */
  double bz0, norm0, normf;
  int model, ipop;
  uint32_t seed;
  double fgrowth_0, hubble_0, omega_m;
} orc_pop;
/* lambda of one cell (srcs.c:158-176; 0 outside r < L/2+20 or where n <= 0). */
double orc_cell_lambda(const orc_pop *p, const double *dens, int ix, int iy, int iz);
/* Poisson count of one cell (srcs.c:158-176) with the Philox stream. */
int orc_cell_count(const orc_pop *p, const double *dens, int ix, int iy, int iz);
/* All counts: counts[N^3] in cell order ix + N(iy + N iz); returns total. */
long long orc_counts(const orc_pop *p, const double *dens, int *counts);
/* orc_counts for planes iz0 .. iz0+nz-1 (dens_slab: those planes, padded). */
long long orc_counts_planes(const orc_pop *p, const double *dens_slab, int iz0, int nz, int *counts);
/* Catalogue in z-major cell order: pos4[4n] (x,y,z,dz_rsd), ipix[n]
 * (NEST at nside_base).  Returns n (<= cap), or -1 if cap is too small. */
long long orc_sources(const orc_pop *p, const double *dens, const double *npot, int nside_base,
                      long long cap, float *pos4, int *ipix);
/* gsl_ran_poisson restated on a caller-supplied MT19937 output stream
 * (uniform = x/2^32); *used = uniforms consumed (-1 if the stream ran out). */
int orc_ref_poisson(const uint32_t *stream, long long len, double mu, long long *used);
/* n raw outputs of MT19937 seeded as gsl_rng_set does. */
void orc_mt_fill(uint32_t seed, long long n, uint32_t *out);
/* Sources from the reference's per-thread streams (srcs.c:120-283 on nthr
 * threads with the static z split iz_begin[nthr+1]); streams[t*stream_len+i].
 * Returns n, -1 (cap too small) or -2 (a stream ran out); counts optional. */
long long orc_sources_ref_stream(const orc_pop *p, const double *dens, const double *npot, int nside_base,
                                 int nthr, const int *iz_begin, const uint32_t *streams, long long stream_len,
                                 long long cap, float *pos4, int *ipix, int *counts);
/* srcs_get_local_properties_single: RA, Dec, z0, dz_rsd per source. */
void orc_source_props(const orc_bg *bg, long long n, const float *pos4, float *out4);
/* Poisson inversion used by both paths. */
int orc_poisson(double lam, uint64_t cell, uint32_t stream, uint32_t seed);

/* ---- lightcone painters ------------------------------------------------- */
void orc_kappa_isw(int which, int N, double L, const double obs[3], double r_max,
                   const orc_bg *bg, const double *npot, int nside, int nplanes,
                   const double *r_src, double *out);
/* srcs.c:425-737 (beams without _USE_FAST_LENSING): dz[n] and, when lens is
 * not NULL, lens[5n] = (e1, e2, kappa, dra, ddec); r_a2r = r(a = i/(na-1)). */
void orc_src_beams(int N, double L, const double obs[3], double r_max, const orc_bg *bg,
                   const double *r_a2r, double fgrowth_0, double hubble_0, double omega_m,
                   const double *npot, long long n, const float *pos4, double *dz, double *lens);
/* srcs.c:507-529, 725-733 (skewers): dg[n nr] density or Gaussian field,
 * v[n nr] scaled radial velocity, nr = N/2. */
void orc_src_skewers(int N, double L, const double obs[3], double r_max, const orc_bg *bg,
                     double fgrowth_0, double hubble_0, double omega_m, const double *dens,
                     const double *npot, int gaussian, double sigma2, long long n, const float *pos4,
                     double *dg, double *vs);
/* cstm.c:39-150: one custom map (NEST) of the tracer's K, b and norm tables;
 * returns 1 when the kernel has no support. */
int orc_cstm(int N, double L, const double obs[3], double r_max, const orc_bg *bg, const double *dens,
             const double *kz_arr, const double *bz_arr, const double *norm_arr, double norm0, double normf,
             int model, int nside, double *out);
/* fftw_wrap_r2c: forward unnormalised half spectrum of a padded real grid. */
void orc_r2c_3d(int N, const double *in, double *out);
/* lpt_1 / lpt_2 (density.c:376-1031) with pos_2_ngp/cic/tsc (interp 0/1/2):
 * out = deposited density - 1 (padded); d2 = D2(r) table (order 2); pos
 * (optional) 3 floats per cell. */
int orc_lpt(int order, int N, double L, const double obs[3], const orc_bg *bg, const double *d2, const double *dg,
            int interp, double *out, float *pos);
void orc_imap(int N, double L, const double obs[3], const orc_bg *bg, const double *dens,
              const double *npot, const double *tz, const double *bz, double bz0,
              const double *norm, double norm0, double normf, int model,
              double fgrowth_0, double hubble_0, double omega_m,
              int nside, int nr, const double *r0, const double *rf, double *data, int *nadd);

/* ---- the surrogate network (no reference counterpart) -------------------- */
/* 4 -> 64 -> 64 -> 2 tanh perceptron in FP64: P = W1[64][4], b1, W2[64][64],
 * b2, W3[2][64], b3 (4610 values); x_n = (x - xm)/xs, y = ym + ys out. */
void orc_mlp_forward(const double *P, const double *xm, const double *xs, const double *ym, const double *ys,
                     const double *x, int n, double *y);
/* L = (1/2n) sum |out - (y - ym)/ys|^2 and its gradient (4610 values). */
double orc_mlp_loss_grad(const double *P, const double *xm, const double *xs, const double *ym, const double *ys,
                         const double *x, const double *y, int n, double *grad);

/* ---- HEALPix -------------------------------------------------------------- */
long long orc_ang2pix_ring(long long nside, double cth, double phi);
long long orc_vec2pix_ring(long long nside, const double *v);
long long orc_ring2nest(long long nside, long long pix);
long long orc_nest2ring(long long nside, long long pix);
void orc_pix2vec_nest(long long nside, long long pix, double *v);

#ifdef __cplusplus
}
#endif
#endif
