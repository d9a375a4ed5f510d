/*
 * colore_amd.h -- C ABI of the MI355X (gfx950) implementation of CoLoRe's
 * field -> lognormal -> sources hot path and its lightcone painters.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns an int status: 0 on success, <0 for a HIP/runtime failure, >0 for
 * an input/specification error.  clr_last_error() returns the message of the
 * last failure on a context.  Calls are synchronous at return (stream-ordered
 * inside), with one exception: the y and x passes of the potential phi run on
 * a second stream and may still be running when clr_gaussian_fields returns;
 * every call that reads phi (sources, kappa, ISW, imap, grid copies) waits
 * for them.  One host thread per context.  Host buffers are owned by the
 * caller.
 *
 * In the reference every stage is `void f(ParamCoLoRe *par)` declared in
 * src/common.h:436-543 and sequenced by src/main.c:61-141; each function
 * below names the reference stage it replaces.
 */
#ifndef COLORE_AMD_H
#define COLORE_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clr_ctx clr_ctx;

/* ---- status codes ---------------------------------------------------- */
#define CLR_OK            0
#define CLR_ERR_HIP      (-1)   /* HIP runtime call failed                 */
#define CLR_ERR_OOM      (-2)   /* device allocation failed                */
#define CLR_ERR_ARG       1     /* bad argument / unsupported size         */
#define CLR_ERR_STATE     2     /* stage called out of order               */

/* ---- enumerations ---------------------------------------------------- */
#define CLR_GRID_DENS     0     /* par->grid_dens (delta)                  */
#define CLR_GRID_NPOT     1     /* par->grid_npot (Newtonian potential)    */

#define CLR_TRACER_SRCS   0     /* srcsN sections                          */
#define CLR_TRACER_IMAP   1     /* imapN sections                          */
#define CLR_TRACER_CSTM   2     /* customN sections                        */

/* bias_model(), src/common.h:414-431; the reference picks one at compile
 * time (Makefile:11-14), here it is chosen per tracer. */
#define CLR_BIAS_EXPONENTIAL   0   /* (1+d)^b                 (default branch) */
#define CLR_BIAS_EXP_TRUNCATED 1   /* _BIAS_MODEL_2                            */
#define CLR_BIAS_TRUNCATED     2   /* _BIAS_MODEL_3                            */

/* density types, src/common.h:115-118 (1LPT/2LPT: clr_lpt_density) */
#define CLR_DENS_LOGNORMAL 0
#define CLR_DENS_CLIP      3

/* random streams for the Fourier modes */
#define CLR_RNG_PHILOX    0     /* counter-based Philox4x32-10, key = seed  */
#define CLR_RNG_INJECTED  1     /* two caller-supplied uint32 per mode      */

#define CLR_NA            5001  /* length of the r-tables, src/common.h:131 */
#define CLR_NPOP_MAX      10    /* src/common.h:132                          */

/* ---- context ---------------------------------------------------------- */

/* Allocates the two padded real grids float[N][N][2(N/2+1)] on `device`
 * (fourier.c:211-238 allocate_fftw; the z halo slabs of the reference are
 * the periodic neighbours on one GPU, fourier.c:412-413).  N must be one of
 * 32, 64, 128, 192, 256, 384, 512, 1024, 2048 (the sizes with a tuned FFT
 * tile set); other sizes return CLR_ERR_ARG. */
int  clr_create(clr_ctx **ctx, int n_grid, double l_box, const double pos_obs[3],
                int device);
void clr_destroy(clr_ctx *ctx);
const char *clr_last_error(const clr_ctx *ctx);
int  clr_version(void);
/* Bytes of device memory the context holds now. */
long long clr_device_bytes(const clr_ctx *ctx);

/* FP64 background tables exactly as held in ParamCoLoRe (src/common.h:220-381,
 * filled by cosmo_set, src/cosmo.c:669-692) and the P(k) table
 * (pk_linear_set, src/cosmo.c:444-495). */
typedef struct {
  int           na;            /* CLR_NA                                  */
  double        glob_idr;      /* (NA-1)/r_arr_r2z[NA-1]                  */
  const double *r_arr_r2z;
  const double *z_arr_r2z;
  const double *growth_d_arr;
  const double *growth_v_arr;
  const double *growth_pd_arr;
  const double *ihub_arr;
  int           numk;          /* P(k) table                              */
  double        logkmin, logkmax, idlogk;
  const double *logkarr;
  const double *pkarr;
  double        n_scal;        /* ns: low-k extrapolation slope           */
  double        fgrowth_0;     /* f(a=1)                                  */
  double        hubble_0;      /* H(a=1) in h/Mpc                         */
  double        omega_m;
  double        prefac_lensing;/* 1.5 H0^2 Om                             */
  double        r_max;         /* chi(z_max): sets the LOS step r_max/(N/2)*/
  const double *growth_d2_arr; /* D2(r) of 2LPT (cosmo.c:689); NULL: none */
} clr_tables;
int clr_upload_tables(clr_ctx *ctx, const clr_tables *tab);

/* Per-tracer r-tables (srcs_nz_arr/srcs_bz_arr, imap_tz_arr/imap_bz_arr,
 * cstm_kz_arr/cstm_bz_arr; src/cosmo.c:693-716).  `amp_arr` is n(r) for
 * sources, T(r) for intensity maps, K(r) for custom maps. */
int clr_set_tracer(clr_ctx *ctx, int kind, int ipop, int bias_model,
                   const double *amp_arr, const double *bz_arr);

/* ---- stages ----------------------------------------------------------- */

/* create_cartesian_fields (fourier.c:361-423) = create_grids_fourier
 * (fourier.c:285-359) + 2x c2r (fourier.c:81-102) + normalisation
 * (fourier.c:383-398) + compute_sigma_dens (fourier.c:24-79).
 * r_smooth is the smoothing LENGTH (<=0: none); the reference squares it in
 * io.c:453-458.  With rng_mode CLR_RNG_INJECTED, `injected` (host memory)
 * holds 2 uint32 per half-spectrum mode in layout order
 * kx + (N/2+1)(ky + N kz): (phase draw, modulus draw), as consumed by
 * rng_delta_gauss (common.c:191-201).  Writes the measured sigma^2. */
int clr_gaussian_fields(clr_ctx *ctx, uint32_t seed, double r_smooth,
                        int smooth_potential, int rng_mode,
                        const uint32_t *injected, double *sigma2_out);

/* compute_physical_density_field (density.c:1105-1126): lognormalize
 * (density.c:1070-1103) or densclip (density.c:1034-1067), in place. */
int clr_density(clr_ctx *ctx, int dens_type, double sigma2);

/* The Lagrangian alternatives of clr_density: order 1 = dens_type 1
 * (lpt_1, density.c:376-645), the displacement i k delta_k / (N^3 k^2) of
 * the Gaussian field; order 2 = dens_type 2 (lpt_2, density.c:646-1031), plus
 * D2(r) times the displacement -i k S_k / (N^3 k^2) of the second-order
 * source S (needs clr_tables.growth_d2_arr).  One particle per cell is moved
 * and assigned to the grid with interp CLR_INTERP_NGP, _CIC or _TSC
 * (lpt_interp_type, pos_2_dens, density.c:37-189); the grid then holds
 * rho - 1.  Needs 3 (order 1) or 8 (order 2) grid-sized buffers of device
 * memory for the call.  pos_out (NULL or 3 N^3 floats) receives the particle
 * positions x, y, z in cell order (ix fastest), as write_lpt dumps them
 * (output_lpt, io.c:619-695). */
#define CLR_INTERP_NGP 0
#define CLR_INTERP_CIC 1
#define CLR_INTERP_TSC 2
int clr_lpt_density(clr_ctx *ctx, int order, int interp, float *pos_out);

/* compute_density_normalization (density.c:1227-1393).  Computes the FP64
 * per-shell sums on the device and builds each tracer's NA norm table on the
 * device.  Optional outputs (may be NULL): nz_out, counts[nz], zsum[nz],
 * sums[n_tracers][nz] (tracers in order srcs..., imap..., cstm...).
 * `cap` is the capacity of the arrays (nz <= cap is required when non-NULL).
 * When `lognormal_fused` is non-zero and dens_type is lognormal, the
 * lognormal transform of clr_density is applied in the same pass over the
 * grid (then do not call clr_density first). */
int clr_shell_norms(clr_ctx *ctx, int lognormal_fused, double sigma2, int cap,
                    int *nz_out, unsigned long long *counts, double *zsum,
                    double *sums);

/* clr_shell_norms without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The partial
 * sums, final sums and packed tables live in one static device array per GPU
 * (clr_noalloc_max_bytes() bytes, about 21 MB) instead of context buffers;
 * the host copies use a fixed array.  Limit: nz (2 + ntracers) <= 320 values
 * (e.g. 42 shells with 5 tracers), else CLR_ERR_ARG.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (other host threads,
 * other contexts on the same device) race; results equal clr_shell_norms'. */
int clr_shell_norms_noalloc(clr_ctx *ctx, int lognormal_fused, double sigma2, int cap,
                            int *nz_out, unsigned long long *counts, double *zsum,
                            double *sums);
/* Size in bytes of the static array of clr_shell_norms_noalloc. */
long long clr_noalloc_max_bytes(void);

/* Copies a tracer's NA norm table (and its edge values norm0/normf). */
int clr_get_norm_table(clr_ctx *ctx, int kind, int ipop, double *norm_arr,
                       double *norm0, double *normf);
/* Replaces a tracer's NA norm table (host-built tables, surrogates). */
int clr_set_norm_table(clr_ctx *ctx, int kind, int ipop, const double *norm_arr,
                       double norm0, double normf);

/* srcs_set_cartesian_single (srcs.c:120-283) for population ipop: Poisson
 * counts with lambda = n(r) dx^3 bias_model(delta,b(r)) norm(r) for cells
 * with r < L/2+20 and n(r) > 0, then positions uniform in [x-dx/2,x+dx/2),
 * RSD shift dz = v_r V1(r) from central differences of phi (srcs.c:87-118),
 * and the NEST base pixel at nside_base (srcs.c:264-270).  Sources stay on
 * the device in global z-major cell order; returns their number.  ipop = -1
 * samples every population in one pass over the grid (srcs_set_cartesian,
 * srcs.c:285-294) with the same catalogues as one call per population, and
 * returns the total.  A cell whose lambda is not finite or above 2^20 (a
 * bias model with b < 0 near delta = -1 can overflow) gets no count: the
 * call returns CLR_ERR_ARG naming the population and cell, and no catalogue
 * is written (the reference would pass such a lambda to gsl_ran_poisson). */
int clr_sample_sources(clr_ctx *ctx, int ipop, uint32_t seed, int nside_base,
                       long long *nsrc_out);
/* clr_sample_sources for population ipop fed the reference's own uniform
 * streams instead of Philox (the reference-stream parity mode; grids up to
 * 512^3).  srcs_set_cartesian_single (srcs.c:120-283) runs on nthr OpenMP
 * threads; thread t owns the planes iz_begin[t] .. iz_begin[t+1]-1 (its
 * schedule(static) share; iz_begin[0] = 0, iz_begin[nthr] = N) and draws from
 * its gsl_rng_mt19937 seeded with seed + t + nthr (ipop + n_srcs IThread0)
 * (srcs.c:150).  streams[t * stream_len + i] holds output i of thread t's
 * generator (uniform = x / 2^32).  Pass 1 walks the thread's cells in z, y, x
 * order and calls gsl_ran_poisson (common.c:185-189) for every cell with
 * r < L/2+20 and n(r) > 0; pass 2 re-seeds the generator (srcs.c:231), so the
 * k-th source of thread t takes outputs 3k, 3k+1, 3k+2 for x, y, z.  The
 * catalogue order, RSD and base pixel are those of clr_sample_sources.
 * Returns CLR_ERR_ARG when a stream is too short. */
int clr_sample_sources_stream(clr_ctx *ctx, int ipop, int nthr, const int *iz_begin,
                              const uint32_t *streams, long long stream_len, int nside_base,
                              long long *nsrc_out);
/* Copies population ipop's CatalogCartesian (common.h:162-167): pos[4n]
 * = (x, y, z, dz_rsd) relative to the observer, ipix[n]. */
int clr_fetch_sources(clr_ctx *ctx, int ipop, float *pos4, int32_t *ipix,
                      long long n);
/* Number of sources in population ipop's catalogue (par->nsources_c_this). */
int clr_catalogue_size(clr_ctx *ctx, int ipop, long long *n);
/* srcs_get_local_properties_single (srcs.c:386-416): per source RA, Dec
 * (degrees), z0 = z(r), dz_rsd; written as float[4n]. */
int clr_source_properties(clr_ctx *ctx, int ipop, float *radecz, long long n);

/* kappa_get_beam_properties (kappa.c:78-175) for pixels 0..npix-1 of a
 * HEALPix map at nside in NEST order (the reference's pixel set on one rank,
 * common.c:505-552); out[nplanes][npix] (NEST).  r_src are the source-plane
 * comoving distances (any order; sorted as kappa_beams_preproc does). */
int clr_kappa(clr_ctx *ctx, int nside, int nplanes, const double *r_src,
              float *out);
/* isw_get_beam_properties (isw.c:78-147); same conventions as clr_kappa. */
int clr_isw(clr_ctx *ctx, int nside, int nplanes, const double *r_src,
            float *out);
/* cstm_beams_preproc_single + cstm_get_beam_properties_single (cstm.c:39-150)
 * for custom map ipop (customN, kernel K(r) set with clr_set_tracer kind
 * CLR_TRACER_CSTM; its norm table from clr_shell_norms): per pixel of a
 * HEALPix map at nside (NEST order, out[12 nside^2]), dr = r_max/(N/2) times
 * the sum over the steps rm = (i+1/2) dr where |K/IH| exceeds 1e-4 of its
 * largest value of K(rm)/IH(rm) (bias_model(d, b(rm)) norm(rm) - 1), d the
 * trilinear (CIC) density at rm u.  Needs clr_density (or the fused shell
 * sums) first. */
int clr_custom_map(clr_ctx *ctx, int ipop, int nside, float *out);
/* srcs_beams_preproc_single, srcs_get_beam_properties_single and
 * srcs_beams_postproc_single (srcs.c:425-737, the build without
 * _USE_FAST_LENSING) for source population ipop after clr_sample_sources.
 * The reference runs them whenever it marches lines of sight (need_beaming,
 * io.c:460-461: lensing at sources, kappa, ISW, skewers or custom maps), and
 * they replace the catalogue's dz_rsd.  Per source, in catalogue order:
 *   dz_rsd[i]     0.5 (N/L) (v.u) V1(r_of_z(z0)) (-f0 / (1.5 H0 Om)), v the
 *                 trilinear (CIC) interpolation of the central differences of
 *                 phi at the source (srcs.c:496-504, 655-662);
 *   lens[5i..5i+4] (e1, e2, kappa, dra, ddec) of the nearest-cell march of
 *                 N/2 steps of r_max/(N/2) up to the source (srcs.c:533-614,
 *                 dra and ddec in degrees, srcs.c:718-719), when lens is not
 *                 NULL (include_lensing).  kappa, dra and ddec start from 0
 *                 (the reference adds them onto unset memory, common.c:373).
 * r_arr_a2r: CLR_NA values of r at a = i/(CLR_NA-1), the table of r_of_z
 * (cosmo.c:101-112).  n must be the population's source count. */
int clr_source_beams(clr_ctx *ctx, int ipop, const double *r_arr_a2r,
                     float *dz_rsd, float *lens, long long n);
/* The skewers of the same pass (store_skewers, srcs.c:507-529) with their
 * postproc scaling (srcs.c:725-733), for population ipop after
 * clr_sample_sources and clr_density (or the fused shell sums).  Rows of
 * nr = N/2 radial steps per source, in catalogue order (the layout of
 * d_skw/g_skw and v_skw, common.c:381-387): at step j, rm = (j+1/2) r_max/nr,
 *   dens_or_gauss[i nr + j]  trilinear (CIC) delta, or with gaussian != 0 the
 *                            Gaussian field ln(1+delta)/D + D sigma2/2, D at
 *                            the grid cell's radius (beaming.c:46-57);
 *   v[i nr + j]              CIC radial velocity 0.5 (N/L)(v.u) times
 *                            V1(rm) (-f0/(1.5 H0 Om));
 * steps past min((int)(r/dr + 1/2), nr-1) hold 0.  Each row is scaled once
 * (the reference's MAX loop bound, srcs.c:725, runs into the next row for
 * sources beyond r_max - dr/2).  sigma2: the Gaussian variance
 * (sigma2_gauss) that clr_gaussian_fields returned. */
int clr_source_skewers(clr_ctx *ctx, int ipop, int gaussian, double sigma2,
                       float *dens_or_gauss, float *v, long long n);
/* imap_set_cartesian_single (imap.c:135-245) for intensity species ipop.
 * r0/rf: the nr shell edges (any order, sorted as imap_preproc does).
 * Outputs full-sky RING maps data[nr][12 nside^2] (sum of T) and
 * nadd[nr][12 nside^2] (number of sub-samples), as the reference stores them
 * before io.c:738-741 divides. */
int clr_imap(clr_ctx *ctx, int ipop, int nside, int nr, const double *r0,
             const double *rf, float *data, int32_t *nadd);

/* Grid copies (density dumps io.c:565-595, parity tests).  Padded layout
 * float[N][N][2(N/2+1)]. */
int clr_fetch_grid(clr_ctx *ctx, int which, float *host);
int clr_upload_grid(clr_ctx *ctx, int which, const float *host);
/* The z-slab of planes iz0 .. iz0+nz-1 of a grid, in the same padded layout
 * float[nz][N][2(N/2+1)] -- what one rank's density dump holds (nz_here
 * planes from iz0_here, io.c:565-595). */
int clr_fetch_grid_slab(clr_ctx *ctx, int which, int iz0, int nz, float *host);
/* Per-cell source counts of population ipop from the last clr_sample_sources
 * call that sampled it (the nsources array of srcs_set_cartesian_single,
 * srcs.c:125-183, without its row padding), for the planes iz0 .. iz0+nz-1:
 * counts[(iz-iz0) N^2 + iy N + ix]. */
int clr_fetch_cell_counts(clr_ctx *ctx, int ipop, int iz0, int nz, int32_t *counts);
/* Declares uploaded grids as the output of the field stage (sigma2 is the
 * measured variance to use) and, if density_transformed, of the density
 * stage too, so that later stages can run on caller-supplied grids. */
int clr_declare_grids(clr_ctx *ctx, double sigma2, int density_transformed);

/* ---- learned surrogate of the per-shell normalisation ------------------
 * No counterpart in the reference: a 4 -> 64 -> 64 -> 2 perceptron (tanh
 * hidden layers) that predicts, from a shell's x = (b, sigma_c, skewness,
 * excess kurtosis) (clr_surrogate_inputs), y0 = the part of ln norm (count/
 * sum of compute_density_normalization, density.c:1272-1355) that the
 * cumulant closure (clr_bias_moment of the shell's measured Gaussian
 * cumulants) misses, and y1 = -y0 (the same for ln <bias_model>, the
 * expected count per cell of the un-normalised sampler in units of
 * n dx^3).  FP32 weights; trained
 * on the device by its own forward/backward kernels (mean squared error of
 * the standardised outputs, full-batch Adam); inference in FP32 or with BF16
 * MFMA (FP32 accumulate).  Parameters: W1[64][4], b1[64], W2[64][64], b2[64],
 * W3[2][64], b3[2] (W[out][in]). */
typedef struct clr_mlp clr_mlp;
#define CLR_MLP_IN      4
#define CLR_MLP_HID     64
#define CLR_MLP_OUT     2
#define CLR_MLP_NPARAM  4610
#define CLR_MLP_FP32      0
#define CLR_MLP_BF16_MFMA 1
int  clr_mlp_create(clr_mlp **net, int device, uint32_t seed);   /* Glorot-uniform init */
void clr_mlp_destroy(clr_mlp *net);
const char *clr_mlp_last_error(const clr_mlp *net);
/* x_n = (x - x_mean)/x_scale on input, y = y_mean + y_scale y_n on output */
int clr_mlp_set_scaling(clr_mlp *net, const float *x_mean, const float *x_scale,
                        const float *y_mean, const float *y_scale);
int clr_mlp_set_params(clr_mlp *net, const float *params);   /* resets Adam */
int clr_mlp_get_params(clr_mlp *net, float *params);
/* Gradient of L = (1/2n) sum |y_n - target_n|^2 over the batch (x: n x 4,
 * y: n x 2 raw targets) at the current parameters; loss optional. */
int clr_mlp_grad(clr_mlp *net, const float *x, const float *y, int n, float *grad, double *loss);
/* `steps` full-batch Adam steps, learning rate decaying geometrically from lr
 * to lr_final; loss_out (NULL or steps values) gets L before each step. */
int clr_mlp_train(clr_mlp *net, const float *x, const float *y, int n, int steps, float lr,
                  float lr_final, double *loss_out);
/* y (n x 2, raw units) from x (n x 4) with CLR_MLP_FP32 or CLR_MLP_BF16_MFMA. */
int clr_mlp_predict(clr_mlp *net, const float *x, int n, float *y, int precision);
/* Mean of bias_model (common.h:414-431) of delta = exp(G - sigma_c^2/2) - 1
 * (the lognormal transform of one cell) over a variate G with mean mu,
 * variance var and third cumulant k3: exp(b mu + b^2 var/2 + b^3 k3/6 -
 * b sigma_c^2/2) for the exponential model; for the two truncated models a
 * 16-point Gauss-Legendre quadrature over the Edgeworth density (normal
 * density times 1 + k3/(6 var^1.5) He_3), split at the model's kink.  With
 * mu = 0, var = sigma_c^2, k3 = 0 it is the one-cell expectation
 * <bias_model> of the lognormal field.  Host only (no device call). */
double clr_bias_moment(int model, double b, double sigma_c, double mu, double var, double k3);

/* Inputs of the surrogate for source population ipop, one row per
 * normalisation shell with cells (bins 1 .. nz-1 of density.c:1233-1244),
 * from a read-only pass over the Gaussian grid that measures, per shell, the
 * sample cumulants of G = D(r) delta_G (the variate the lognormal transform
 * exponentiates): x[i] = (b, sigma_c, skewness, excess kurtosis) with b(z)
 * at the shell's mean redshift and sigma_c^2 = <D^2> sigma^2 (sigma^2 the
 * measured variance of the Gaussian grid); base[i] = -ln clr_bias_moment(
 * model, b, sigma_c, mean, variance, third cumulant), the closure's ln norm;
 * shell[i] is the bin.  Needs clr_gaussian_fields (density not transformed).
 * *n_out = number of rows (x, base, shell may be NULL to query it). */
int clr_surrogate_inputs(clr_ctx *ctx, int ipop, int cap, float *x, double *base, int *shell,
                         int *n_out);
/* Surrogate-fused sampling: clr_sample_sources (same catalogues, streams
 * and order) for population ipop (-1: all) right after clr_gaussian_fields,
 * without clr_shell_norms.  Each shell's ln norm is the closure's, plus the
 * network's y0 for populations with the exponential bias model (BF16 MFMA
 * or FP32 inference; net = NULL: the closure alone);
 * the norm tables are built from them as compute_density_normalization
 * builds them from the shell sums (density.c:1272-1355); the count pass
 * reads the Gaussian field and applies the lognormal transform per cell (the
 * same FP32 value the shell pass of clr_shell_norms(lognormal_fused)
 * stores), so given equal norm tables the counts equal the two-pass path's.
 * The grid keeps the Gaussian field. */
int clr_sample_sources_surrogate(clr_ctx *ctx, clr_mlp *net, int ipop, uint32_t seed,
                                 int nside_base, int precision, long long *nsrc_out);

/* ---- ensembles in one captured graph ----------------------------------
 * No counterpart in the reference (one realisation per run there).  n
 * independent realisations, each an existing context (its own N, box,
 * tables, populations; all on one device), run as one HIP graph: per
 * realisation clr_gaussian_fields (Philox), clr_shell_norms with the fused
 * lognormal, clr_sample_sources(-1) and a kappa map of one source plane at
 * r_kappa[i] (nside_kappa; 0: no map), spread over nstreams captured streams.
 * Buffers are set up by clr_ensemble_create (catalogue capacity src_cap[i]
 * per population).  clr_ensemble_run captures and instantiates the graph on
 * the first call and whenever the seeds or smoothing change, then replays
 * it; afterwards every context is in the state the single calls leave
 * (clr_fetch_sources, clr_fetch_grid, clr_get_norm_table work) and its
 * results equal those calls' bit for bit.  A realisation whose counts hit a
 * case the graph does not handle (counts of 254 or more in a cell, an
 * overflowing FP64 queue, a catalogue past src_cap) is reported with a
 * non-zero flag: re-run it with the single calls. */
typedef struct clr_ensemble clr_ensemble;
int  clr_ensemble_create(clr_ensemble **ens, int n, clr_ctx *const *ctxs, int nside_base,
                         int nside_kappa, const double *r_kappa, const long long *src_cap,
                         int nstreams);
void clr_ensemble_destroy(clr_ensemble *ens);
const char *clr_ensemble_last_error(const clr_ensemble *ens);
/* seeds[n], r_smooth[n]; n_invalid: realisations with a flag; ms: device
 * time of the graph launch. */
int  clr_ensemble_run(clr_ensemble *ens, const uint32_t *seeds, const double *r_smooth,
                      int smooth_potential, int *n_invalid, float *ms);
/* Realisation i: sigma^2, its flag, and (kappa != NULL) its map, NEST
 * float[12 nside_kappa^2]. */
int  clr_ensemble_result(clr_ensemble *ens, int i, double *sigma2, int *flag, float *kappa);
/* FFT tile autotuner: every tile-shape candidate for this context's N runs
 * the field stage `reps` times on its grids (which it overwrites); the
 * first candidate whose median time is within 1 % of the smallest is used
 * from then on by every context of that N in the process (so that timing
 * noise between near-equal shapes does not flip the choice).  ms[cap]:
 * median per candidate. */
int  clr_fft_autotune(clr_ctx *ctx, int reps, int cap, float *ms, int *ncand, int *best);
/* The FFT tile-shape candidates of grid size n_grid: their number (0 when
 * the size has no FFT configuration), and the selection of candidate k for
 * every context of that size in the process -- what clr_fft_autotune
 * chooses by timing, here chosen by the caller (to repeat a run with the
 * shapes of an earlier one, or to test every shape).  No counterpart in the
 * reference (FFTW plans its own).  Returns CLR_ERR_ARG for a k out of range. */
int  clr_fft_candidates(int n_grid);
int  clr_fft_select(int n_grid, int k);

/* Device time in milliseconds of the last call of each stage (0 if never),
 * in this order: mode generation + z pass, y pass, x pass, density
 * transform, shell sums (whole call), sources (whole call), kappa, ISW, imap,
 * the counting pass of the sources call, its catalogue writers (descriptors
 * and per-source positions), the y and x passes of phi (these run on a second
 * stream, overlapping the density's passes and the stages after them; every
 * call that reads phi waits for them), the shell-sum kernel alone, the
 * source-beam kernel, the custom-map kernel, the skewer kernels (summed over
 * their chunks), then phi's y pass alone and phi's x pass alone.  Fills at
 * most n values; returns the number of stages. */
int clr_stage_times(const clr_ctx *ctx, double *ms, int n);
/* Waits for every device operation of the context, the potential's passes
 * on the second stream included (no counterpart in the reference, whose
 * stages return when done). */
int clr_synchronize(clr_ctx *ctx);

/* Page-locked host memory for the outputs a caller fills again and again
 * (maps, catalogues): device-to-host copies into it run at the link's rate
 * instead of through the runtime's pageable staging (a 37.7 MB set of three
 * nside-512 maps: 0.7 ms against ~4 ms).  No counterpart in the reference
 * (its maps and catalogues are malloc'd once per run).  Returns NULL when the
 * allocation fails; clr_host_free returns CLR_ERR_HIP for a pointer the
 * runtime does not know. */
void *clr_host_alloc(size_t bytes);
int clr_host_free(void *p);

#ifdef __cplusplus
}
#endif
#endif /* COLORE_AMD_H */
