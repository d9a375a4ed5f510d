// clr_internal.h -- the context behind the C ABI and the launchers each
// stage file provides.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "../../include/colore_amd.h"
#include "clr_device.h"

struct TracerDev {
  int set = 0;
  int model = 0;
  double *amp = nullptr;    // n(r) / T(r) / K(r), NA
  double *bz = nullptr;     // b(r), NA
  double *norm = nullptr;   // norm(r), NA
  float *bzf = nullptr;     // FP32 copy of bz (feeds FP32 estimates only)
  void *ptab = nullptr;     // float4 per bin: (n dx^3, b, norm, flags) for sources
  double bz0 = 0;           // f0 of the bias table (its first entry)
  double norm0 = 1, normf = 1;
  int norm_set = 0;
  std::vector<double> h_bz; // host copy of bz (surrogate inputs)
  double *edge = nullptr;   // device copy of (norm0, normf) (graph-captured stages)
};

struct CatalogDev {
  float *pos = nullptr;     // 4 floats per source
  int32_t *ipix = nullptr;
  long long cap = 0;
  long long n = 0;
};

enum StageId {
  ST_GEN = 0, ST_FFT, ST_XPASS, ST_DENS, ST_NORM, ST_SRCS, ST_KAPPA, ST_ISW, ST_IMAP,
  ST_SRC_PASS,    // k_count of a sample_sources call
  ST_SRC_WRITE,   // its k_fill and k_write launches
  ST_FFT_PHI,     // y and x passes of phi (second stream, overlapping later stages)
  ST_SHELLS,      // the k_shells launch of a normalisation call
  ST_BEAMS,       // k_src_beams of a clr_source_beams call
  ST_CSTM,        // k_cstm of a clr_custom_map call
  ST_SKW,         // the k_src_skewers launches of a clr_source_skewers call (kernels only)
  ST_PHI_Y,       // phi's y pass alone (second stream)
  ST_PHI_X,       // phi's x pass alone (second stream)
  ST_COUNT
};

struct clr_ctx {
  int device = 0;
  int N = 0;
  int NH = 0;               // N/2+1
  int ngx = 0;              // floats per device row: grid_pitch(N) >= 2*(N/2+1)
  double L = 0, dx = 0;
  double obs[3] = {0, 0, 0};
  hipStream_t stream = nullptr;
  // phi's y and x passes run on stream2 while stream goes on with the density;
  // ev_phi marks their end, every reader of npot waits on it (clr_phi_ready)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_gen = nullptr, ev_phi = nullptr, ev4 = nullptr, ev5 = nullptr, ev6 = nullptr;
  int phi_pending = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  double stage_ms[ST_COUNT] = {0};

  float *dens = nullptr, *npot = nullptr;   // padded real grids
  size_t grid_floats = 0;

  // FP64 tables
  int tables_set = 0;
  double *bg = nullptr;     // r, z, d1, v1, pd, ih (NA each)
  float *d1f = nullptr;     // FP32 copy of the growth table
  int na = 0;
  double glob_idr = 0;
  double r_last = 0;
  double h_bg_first[6] = {0}, h_bg_last[6] = {0};   // first/last entries of each bg table
  double zmax_half_box = 0;                           // z(L/2), get_bg(L/2, BG_Z)
  int32_t *r2n = nullptr;   // RING -> NEST table at r2n_nside
  int r2n_nside = 0;
  double *pk = nullptr;     // logkarr[numk], pkarr[numk]
  int numk = 0;
  double logkmin = 0, logkmax = 0, idlogk = 0, n_scal = 0;
  double fgrowth_0 = 0, hubble_0 = 0, omega_m = 0, prefac_lensing = 0, r_max = 0;

  TracerDev tr[3][CLR_NPOP_MAX];
  int n_tr[3] = {0, 0, 0};
  CatalogDev cat[CLR_NPOP_MAX];

  // scratch
  float2 *tw = nullptr;     // exp(+2 pi i n/N), n < N
  float2 *amp = nullptr;    // per integer |k|^2 index: (delta amp, phi amp)
  size_t amp_n = 0;
  double *red = nullptr;    // reduction partials
  size_t red_n = 0;
  double *red_out = nullptr;
  size_t red_out_n = 0;
  int32_t *rowcnt = nullptr;   // per (z,y) row counts / offsets (+1)
  long long *rowoff = nullptr;
  size_t rows_n = 0;
  uint32_t *inj = nullptr;  // injected uniforms (parity mode)
  size_t inj_n = 0;
  double *hp_vec = nullptr; // pixel centres scratch
  size_t hp_n = 0;
  void *esc = nullptr;      // sources: escape list of saturated count bytes (+ counter)
  size_t esc_bytes = 0;
  double *d2 = nullptr;     // D2(r) table of 2LPT (NA), when the tables carry it
  double d2_first = 0, d2_last = 0;
  void *misc = nullptr;     // generic scratch
  size_t misc_n = 0;
  // the count bytes of the last clr_sample_sources call: populations
  // cnt_p0 .. cnt_p0+cnt_np-1 at rowcnt + (ipop-cnt_p0)*cnt_per_pop, cnt_nesc
  // sorted escape entries at esc+16
  int cnt_p0 = -1, cnt_np = 0;
  size_t cnt_per_pop = 0;
  unsigned int cnt_nesc = 0;
  // per-cell counts of the last reference-stream call (population ref_ipop)
  void *ref_cnt = nullptr;
  size_t ref_cnt_n = 0;
  int ref_ipop = -1;

  // host copies of r, z, D (surrogate inputs) and the shell geometry of the
  // normalisation shells (cells and sum of z per shell; geo_nz = 0: not yet)
  std::vector<double> h_r, h_z, h_d1, h_pd;
  int geo_nz = 0;
  double geo_zmax = 0;
  std::vector<double> geo_cnt, geo_zsum;
  double r_smooth = 0;      // smoothing length of the last clr_gaussian_fields
  // graph-captured kappa: kernel arguments, radial weights + map buffer
  void *tidal = nullptr;     // tidal records, 6 floats per cell (clr_los.hip)
  size_t tidal_n = 0;
  void *velrec = nullptr;    // velocity-difference + density records, 4 floats per cell (clr_los.hip)
  size_t velrec_n = 0;
  std::string los_plan;
  void *los_buf = nullptr;
  size_t los_buf_n = 0;
  int los_nside = 0, los_nplanes = 0;

  int have_fields = 0;
  int density_done = 0;
  int norms_done = 0;
  double sigma2 = 0;
  std::string err;
};

// status helpers
int clr_fail(clr_ctx *c, int code, const char *fmt, ...);
#define CLR_HIP(ctx, call)                                                        \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return clr_fail((ctx), CLR_ERR_HIP, "%s failed: %s (%s:%d)", #call,         \
                      hipGetErrorString(e_), __FILE__, __LINE__);                 \
  } while (0)

int clr_ensure(clr_ctx *c, void **p, size_t *have, size_t bytes);
int clr_ring2nest_table(clr_ctx *c, int nside);

// stage launchers (one per file)
int clr_launch_fields(clr_ctx *c, uint32_t seed, double r_smooth, int smooth_potential,
                      int rng_mode, const uint32_t *injected, double *sigma2_out);
int clr_launch_density(clr_ctx *c, int dens_type, double sigma2);
// scratch_base: NULL (buffers from the context, grown on demand) or the
// device's static scratch (clr_shell_norms_noalloc)
int clr_launch_norms(clr_ctx *c, int lognormal_fused, double sigma2, int cap, int *nz_out,
                     unsigned long long *counts, double *zsum, double *sums, void *scratch_base);
// The no-allocation scratch: its size, and its device address on the current
// device (0, or -1 when hipGetSymbolAddress fails).
size_t clr_noalloc_scratch_bytes();
int clr_noalloc_scratch_lookup(void **ptr);
// Makes npot safe to use: host = 0 orders c->stream after phi's passes,
// host = 1 waits for them on the host.
int clr_phi_ready(clr_ctx *c, int host);
// fused: the grid holds the Gaussian field; counts use its lognormal
// transform per cell (surrogate-fused sampling, norm tables already set)
int clr_launch_sources(clr_ctx *c, int ipop, uint32_t seed, int nside_base, long long *nsrc, int fused);
int clr_launch_source_props(clr_ctx *c, int ipop, float *out, long long n);
int clr_launch_fetch_counts(clr_ctx *c, int ipop, int iz0, int nz, int32_t *out);
int clr_launch_sources_stream(clr_ctx *c, int ipop, int nthr, const int *iz_begin, const uint32_t *streams,
                              long long stream_len, int nside_base, long long *nsrc);
int clr_launch_los(clr_ctx *c, int which, int nside, int nplanes, const double *r_src, float *out);
int clr_launch_src_beams(clr_ctx *c, int ipop, const double *r_arr_a2r, float *dz, float *lens, long long n);
int clr_launch_cstm(clr_ctx *c, int ipop, int nside, float *out);
int clr_launch_lpt(clr_ctx *c, int order, int interp, float *pos_out);
int clr_launch_src_skewers(clr_ctx *c, int ipop, int gaussian, double sigma2, float *dg, float *vs, long long n);
int clr_launch_imap(clr_ctx *c, int ipop, int nside, int nr, const double *r0, const double *rf,
                    float *data, int32_t *nadd);

// Shell geometry of the normalisation (cells and z sums per shell), cached in
// the context; and norm tables built from given per-shell norms of tracers
// (kind, ipop) as compute_density_normalization builds them from sums.
int clr_launch_shell_geometry(clr_ctx *c);
int clr_launch_shell_moments(clr_ctx *c, std::vector<double> &out);
bool clr_cell_records(clr_ctx *c, bool tidal);
int clr_launch_norm_tables_given(clr_ctx *c, int ntr, const int *kinds, const int *ipops, const double *norm_s);

// Graph-capturable stages (clr_ensemble.hip): buffers first (prepare, may
// allocate), then kernel launches only on the given stream (enqueue: no
// allocation, no memset, no host wait; sigma^2 and the norm edges stay on the
// device).  Same kernels and results as the launchers above.
int clr_prepare_fields(clr_ctx *c);
int clr_autotune_fields(clr_ctx *c, int reps, int cap, float *ms, int *ncand, int *best);
int clr_fft_candidate_count(int N);
int clr_fft3d(clr_ctx *c, float *g, int forward);
int clr_fft_select_candidate(int N, int k);
int clr_enqueue_fields(clr_ctx *c, hipStream_t s, uint32_t seed, double r_smooth, int smooth_potential,
                       double *sigma2_dev);
int clr_prepare_norms(clr_ctx *c);
int clr_enqueue_norms(clr_ctx *c, hipStream_t s, const double *sigma2_dev);
int clr_prepare_sources(clr_ctx *c, int nside_base, const long long *cap);
int clr_enqueue_sources(clr_ctx *c, hipStream_t s, uint32_t seed, int nside_base, int *flag_dev);
int clr_finish_sources(clr_ctx *c);
int clr_prepare_kappa(clr_ctx *c, int nside, int nplanes, const double *r_src);
int clr_enqueue_kappa(clr_ctx *c, hipStream_t s);
int clr_fetch_kappa(clr_ctx *c, float *out);

// ln norm of the lognormal moment for surrogate inputs x (D, b, sigma_G,
// r_smooth) of a tracer with bias model `model`; the network predicts the
// difference from it (clr_surrogate.hip).

// Prediction of the surrogate network on device buffers (clr_mlp.hip).
int clr_mlp_predict_device(clr_mlp *net, const float *dx, int n, float *dy, int precision, hipStream_t stream);
