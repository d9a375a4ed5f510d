// clr_api.hip -- context management and the C ABI entry points.
#include "clr_internal.h"
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <mutex>

bool clr_fft_size_supported(int N);

int clr_fail(clr_ctx *c, int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

int clr_ensure(clr_ctx *c, void **p, size_t *have, size_t bytes) {
  if (*have >= bytes && *p) return CLR_OK;
  if (*p) {
    CLR_HIP(c, hipStreamSynchronize(c->stream));
    hipFree(*p);
  }
  *p = nullptr;
  *have = 0;
  if (hipMalloc(p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return clr_fail(c, CLR_ERR_OOM, "device allocation of %zu bytes failed", bytes);
  }
  *have = bytes;
  return CLR_OK;
}

int clr_phi_ready(clr_ctx *c, int host) {
  if (!c->phi_pending) return CLR_OK;
  if (!host) {
    CLR_HIP(c, hipStreamWaitEvent(c->stream, c->ev_phi, 0));
    return CLR_OK;
  }
  CLR_HIP(c, hipStreamSynchronize(c->stream2));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, c->ev4, c->ev5);
  c->stage_ms[ST_FFT_PHI] = ms;
  (void)hipEventElapsedTime(&ms, c->ev4, c->ev6);
  c->stage_ms[ST_PHI_Y] = ms;
  (void)hipEventElapsedTime(&ms, c->ev6, c->ev5);
  c->stage_ms[ST_PHI_X] = ms;
  c->phi_pending = 0;
  return CLR_OK;
}

int clr_ring2nest_table(clr_ctx *c, int nside) {
  if (c->r2n && c->r2n_nside == nside) return CLR_OK;
  const int64_t npix = 12LL * nside * nside;
  std::vector<int32_t> t(npix);
  for (int64_t p = 0; p < npix; p++) t[p] = (int32_t)clr::ring2nest(nside, p);
  if (c->r2n) { CLR_HIP(c, hipStreamSynchronize(c->stream)); hipFree(c->r2n); c->r2n = nullptr; }
  CLR_HIP(c, hipMalloc((void **)&c->r2n, npix * sizeof(int32_t)));
  CLR_HIP(c, hipMemcpy(c->r2n, t.data(), npix * sizeof(int32_t), hipMemcpyHostToDevice));
  c->r2n_nside = nside;
  return CLR_OK;
}

// Device address of the no-allocation scratch (clr_shell_norms_noalloc), one
// per device: looked up once per device by clr_create with that device
// current.  Each GPU holds its own copy of the array at its own address.
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];
static int g_scratch_state[64];   // 0 not looked up, 1 found, -1 failed

static int noalloc_scratch_setup(clr_ctx *c) {
  if (c->device < 0 || c->device >= 64) return clr_fail(c, CLR_ERR_ARG, "device %d out of range", c->device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_state[c->device] == 0) {
    void *p = nullptr;
    if (clr_noalloc_scratch_lookup(&p) == 0 && p) {
      g_scratch_ptr[c->device] = p;
      g_scratch_state[c->device] = 1;
    } else {
      (void)hipGetLastError();
      g_scratch_state[c->device] = -1;
    }
  }
  return CLR_OK;
}

static void free_all(clr_ctx *c) {
  auto f = [](void *p) { if (p) hipFree(p); };
  f(c->dens); f(c->npot); f(c->bg); f(c->pk); f(c->tw); f(c->amp); f(c->red); f(c->red_out);
  f(c->rowcnt); f(c->ref_cnt); f(c->los_buf); f(c->esc); f(c->inj); f(c->hp_vec); f(c->misc); f(c->d2); f(c->d1f); f(c->r2n); f(c->tidal); f(c->velrec);
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < CLR_NPOP_MAX; i++) {
      f(c->tr[k][i].amp); f(c->tr[k][i].bz); f(c->tr[k][i].norm); f(c->tr[k][i].bzf); f(c->tr[k][i].ptab);
      f(c->tr[k][i].edge);
    }
  for (int i = 0; i < CLR_NPOP_MAX; i++) { f(c->cat[i].pos); f(c->cat[i].ipix); }
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->ev2) hipEventDestroy(c->ev2);
  if (c->ev_gen) hipEventDestroy(c->ev_gen);
  if (c->ev_phi) hipEventDestroy(c->ev_phi);
  if (c->ev4) hipEventDestroy(c->ev4);
  if (c->ev5) hipEventDestroy(c->ev5);
  if (c->ev6) hipEventDestroy(c->ev6);
  if (c->ev3) hipEventDestroy(c->ev3);
  if (c->stream2) { hipStreamSynchronize(c->stream2); hipStreamDestroy(c->stream2); }
  if (c->stream) hipStreamDestroy(c->stream);
}

extern "C" {

int clr_version(void) { return 100; }

const char *clr_last_error(const clr_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int clr_create(clr_ctx **out, int n_grid, double l_box, const double pos_obs[3], int device) {
  if (!out) return CLR_ERR_ARG;
  *out = nullptr;
  clr_ctx *c = new clr_ctx();
  if (!clr_fft_size_supported(n_grid) || !(l_box > 0)) {
    clr_fail(c, CLR_ERR_ARG, "unsupported grid n_grid=%d l_box=%g", n_grid, l_box);
    *out = c;
    return CLR_ERR_ARG;
  }
  c->device = device;
  c->N = n_grid;
  c->NH = n_grid / 2 + 1;
  c->ngx = clr::grid_pitch(n_grid);
  c->L = l_box;
  c->dx = l_box / n_grid;
  for (int i = 0; i < 3; i++) c->obs[i] = pos_obs ? pos_obs[i] : 0.5 * l_box;
  *out = c;
  CLR_HIP(c, hipSetDevice(device));
  CLR_HIP(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  CLR_HIP(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
  CLR_HIP(c, hipEventCreateWithFlags(&c->ev_gen, hipEventDisableTiming));
  CLR_HIP(c, hipEventCreateWithFlags(&c->ev_phi, hipEventDisableTiming));
  CLR_HIP(c, hipEventCreate(&c->ev4));
  CLR_HIP(c, hipEventCreate(&c->ev5));
  CLR_HIP(c, hipEventCreate(&c->ev6));
  CLR_HIP(c, hipEventCreate(&c->ev0));
  CLR_HIP(c, hipEventCreate(&c->ev1));
  CLR_HIP(c, hipEventCreate(&c->ev2));
  CLR_HIP(c, hipEventCreate(&c->ev3));
  int rs = noalloc_scratch_setup(c);
  if (rs) return rs;
  c->grid_floats = (size_t)n_grid * n_grid * c->ngx;
  if (hipMalloc((void **)&c->dens, c->grid_floats * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&c->npot, c->grid_floats * sizeof(float)) != hipSuccess) {
    (void)hipGetLastError();
    return clr_fail(c, CLR_ERR_OOM, "cannot allocate 2 grids of %zu floats", c->grid_floats);
  }
  // twiddles exp(+2 pi i n / N), FP64-rounded
  std::vector<float2> tw(n_grid);
  for (int n = 0; n < n_grid; n++) {
    const double a = 2 * M_PI * (double)n / n_grid;
    tw[n] = make_float2((float)cos(a), (float)sin(a));
  }
  CLR_HIP(c, hipMalloc((void **)&c->tw, n_grid * sizeof(float2)));
  CLR_HIP(c, hipMemcpy(c->tw, tw.data(), n_grid * sizeof(float2), hipMemcpyHostToDevice));
  return CLR_OK;
}

void clr_destroy(clr_ctx *ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  if (ctx->stream2) hipStreamSynchronize(ctx->stream2);
  free_all(ctx);
  delete ctx;
}

long long clr_device_bytes(const clr_ctx *c) {
  if (!c) return 0;
  long long b = 2LL * (long long)c->grid_floats * 4 + c->amp_n + c->red_n + c->red_out_n + c->rows_n + c->inj_n +
                c->hp_n + c->misc_n + c->tidal_n + c->velrec_n;
  for (int i = 0; i < CLR_NPOP_MAX; i++) b += c->cat[i].cap * 20;
  return b;
}

int clr_upload_tables(clr_ctx *c, const clr_tables *t) {
  if (!c || !t) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  if (t->na < 2 || !t->r_arr_r2z || !t->z_arr_r2z || !t->growth_d_arr || !t->growth_v_arr ||
      !t->growth_pd_arr || !t->ihub_arr)
    return clr_fail(c, CLR_ERR_ARG, "incomplete background tables");
  if (t->numk < 2 || !t->logkarr || !t->pkarr) return clr_fail(c, CLR_ERR_ARG, "incomplete P(k) table");
  const int na = t->na;
  if (c->bg && c->na != na) { hipFree(c->bg); c->bg = nullptr; }
  if (!c->bg) CLR_HIP(c, hipMalloc((void **)&c->bg, 6 * (size_t)na * sizeof(double)));
  const double *src[6] = {t->r_arr_r2z, t->z_arr_r2z, t->growth_d_arr, t->growth_v_arr, t->growth_pd_arr,
                          t->ihub_arr};
  for (int i = 0; i < 6; i++)
    CLR_HIP(c, hipMemcpy(c->bg + (size_t)i * na, src[i], na * sizeof(double), hipMemcpyHostToDevice));
  c->na = na;
  c->glob_idr = t->glob_idr;
  c->r_last = t->r_arr_r2z[na - 1];
  c->h_r.assign(t->r_arr_r2z, t->r_arr_r2z + na);
  c->h_z.assign(t->z_arr_r2z, t->z_arr_r2z + na);
  c->h_d1.assign(t->growth_d_arr, t->growth_d_arr + na);
  c->h_pd.assign(t->growth_pd_arr, t->growth_pd_arr + na);
  c->geo_nz = 0;   // shell geometry depends on the tables
  for (int i = 0; i < 6; i++) {
    c->h_bg_first[i] = src[i][0];
    c->h_bg_last[i] = src[i][na - 1];
  }
  if (c->d2) hipFree(c->d2);
  c->d2 = nullptr;
  if (t->growth_d2_arr) {
    CLR_HIP(c, hipMalloc((void **)&c->d2, na * sizeof(double)));
    CLR_HIP(c, hipMemcpy(c->d2, t->growth_d2_arr, na * sizeof(double), hipMemcpyHostToDevice));
    c->d2_first = t->growth_d2_arr[0];
    c->d2_last = t->growth_d2_arr[na - 1];
  }
  // get_bg(L/2, BG_Z): the top of the normalisation shells (density.c:1233)
  c->zmax_half_box = clr::f_of_r(t->r_arr_r2z, t->z_arr_r2z, t->glob_idr, c->r_last, 0.5 * c->L, 0.0,
                                 t->z_arr_r2z[na - 1]);
  std::vector<float> d1f(na);
  for (int i = 0; i < na; i++) d1f[i] = (float)t->growth_d_arr[i];
  if (c->d1f) hipFree(c->d1f);
  c->d1f = nullptr;
  CLR_HIP(c, hipMalloc((void **)&c->d1f, na * sizeof(float)));
  CLR_HIP(c, hipMemcpy(c->d1f, d1f.data(), na * sizeof(float), hipMemcpyHostToDevice));
  if (c->pk && c->numk != t->numk) { hipFree(c->pk); c->pk = nullptr; }
  if (!c->pk) CLR_HIP(c, hipMalloc((void **)&c->pk, 2 * (size_t)t->numk * sizeof(double)));
  CLR_HIP(c, hipMemcpy(c->pk, t->logkarr, t->numk * sizeof(double), hipMemcpyHostToDevice));
  CLR_HIP(c, hipMemcpy(c->pk + t->numk, t->pkarr, t->numk * sizeof(double), hipMemcpyHostToDevice));
  c->numk = t->numk;
  c->logkmin = t->logkmin;
  c->logkmax = t->logkmax;
  c->idlogk = t->idlogk;
  c->n_scal = t->n_scal;
  c->fgrowth_0 = t->fgrowth_0;
  c->hubble_0 = t->hubble_0;
  c->omega_m = t->omega_m;
  c->prefac_lensing = t->prefac_lensing;
  c->r_max = t->r_max;
  c->tables_set = 1;
  return CLR_OK;
}

int clr_set_tracer(clr_ctx *c, int kind, int ipop, int bias_model, const double *amp_arr, const double *bz_arr) {
  if (!c) return CLR_ERR_ARG;
  if (kind < 0 || kind > 2 || ipop < 0 || ipop >= CLR_NPOP_MAX)
    return clr_fail(c, CLR_ERR_ARG, "bad tracer kind %d / index %d", kind, ipop);
  if (bias_model < 0 || bias_model > 2) return clr_fail(c, CLR_ERR_ARG, "bad bias model %d", bias_model);
  if (!c->tables_set) return clr_fail(c, CLR_ERR_STATE, "clr_upload_tables must be called first");
  if (!amp_arr || !bz_arr) return clr_fail(c, CLR_ERR_ARG, "missing tracer tables");
  if (ipop > c->n_tr[kind]) return clr_fail(c, CLR_ERR_ARG, "tracers must be set in order (next is %d)", c->n_tr[kind]);
  CLR_HIP(c, hipSetDevice(c->device));
  TracerDev &T = c->tr[kind][ipop];
  const size_t b = (size_t)c->na * sizeof(double);
  if (!T.amp) CLR_HIP(c, hipMalloc((void **)&T.amp, b));
  if (!T.bz) CLR_HIP(c, hipMalloc((void **)&T.bz, b));
  if (!T.norm) CLR_HIP(c, hipMalloc((void **)&T.norm, b));
  if (!T.bzf) CLR_HIP(c, hipMalloc((void **)&T.bzf, (size_t)c->na * sizeof(float)));
  CLR_HIP(c, hipMemcpy(T.amp, amp_arr, b, hipMemcpyHostToDevice));
  CLR_HIP(c, hipMemcpy(T.bz, bz_arr, b, hipMemcpyHostToDevice));
  T.h_bz.assign(bz_arr, bz_arr + c->na);
  std::vector<float> bzf(c->na);
  for (int i = 0; i < c->na; i++) bzf[i] = (float)bz_arr[i];
  CLR_HIP(c, hipMemcpy(T.bzf, bzf.data(), (size_t)c->na * sizeof(float), hipMemcpyHostToDevice));
  T.bz0 = bz_arr[0];
  T.model = bias_model;
  T.set = 1;
  T.norm_set = 0;
  if (ipop == c->n_tr[kind]) c->n_tr[kind]++;
  return CLR_OK;
}

int clr_gaussian_fields(clr_ctx *c, uint32_t seed, double r_smooth, int smooth_potential, int rng_mode,
                        const uint32_t *injected, double *sigma2_out) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_launch_fields(c, seed, r_smooth, smooth_potential, rng_mode, injected, sigma2_out);
}

int clr_density(clr_ctx *c, int dens_type, double sigma2) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_launch_density(c, dens_type, sigma2);
}

int clr_shell_norms(clr_ctx *c, int lognormal_fused, double sigma2, int cap, int *nz_out,
                    unsigned long long *counts, double *zsum, double *sums) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_launch_norms(c, lognormal_fused, sigma2, cap, nz_out, counts, zsum, sums, nullptr);
}

int clr_shell_norms_noalloc(clr_ctx *c, int lognormal_fused, double sigma2, int cap, int *nz_out,
                            unsigned long long *counts, double *zsum, double *sums) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  void *p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c->device >= 0 && c->device < 64 && g_scratch_state[c->device] == 1) p = g_scratch_ptr[c->device];
  }
  if (!p) return clr_fail(c, CLR_ERR_HIP, "no static scratch on device %d (hipGetSymbolAddress failed)", c->device);
  return clr_launch_norms(c, lognormal_fused, sigma2, cap, nz_out, counts, zsum, sums, p);
}

long long clr_noalloc_max_bytes(void) { return (long long)clr_noalloc_scratch_bytes(); }

int clr_get_norm_table(clr_ctx *c, int kind, int ipop, double *norm_arr, double *norm0, double *normf) {
  if (!c) return CLR_ERR_ARG;
  if (kind < 0 || kind > 2 || ipop < 0 || ipop >= c->n_tr[kind] || !c->tr[kind][ipop].norm_set)
    return clr_fail(c, CLR_ERR_STATE, "no norm table for tracer %d/%d", kind, ipop);
  CLR_HIP(c, hipSetDevice(c->device));
  TracerDev &T = c->tr[kind][ipop];
  if (norm_arr) CLR_HIP(c, hipMemcpy(norm_arr, T.norm, (size_t)c->na * sizeof(double), hipMemcpyDeviceToHost));
  if (norm0) *norm0 = T.norm0;
  if (normf) *normf = T.normf;
  return CLR_OK;
}

int clr_set_norm_table(clr_ctx *c, int kind, int ipop, const double *norm_arr, double norm0, double normf) {
  if (!c || !norm_arr) return CLR_ERR_ARG;
  if (kind < 0 || kind > 2 || ipop < 0 || ipop >= c->n_tr[kind] || !c->tr[kind][ipop].set)
    return clr_fail(c, CLR_ERR_STATE, "tracer %d/%d not set", kind, ipop);
  CLR_HIP(c, hipSetDevice(c->device));
  TracerDev &T = c->tr[kind][ipop];
  CLR_HIP(c, hipMemcpy(T.norm, norm_arr, (size_t)c->na * sizeof(double), hipMemcpyHostToDevice));
  T.norm0 = norm0;
  T.normf = normf;
  T.norm_set = 1;
  return CLR_OK;
}

int clr_sample_sources(clr_ctx *c, int ipop, uint32_t seed, int nside_base, long long *nsrc_out) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_launch_sources(c, ipop, seed, nside_base, nsrc_out, 0);
}

int clr_sample_sources_stream(clr_ctx *c, int ipop, int nthr, const int *iz_begin, const uint32_t *streams,
                              long long stream_len, int nside_base, long long *nsrc_out) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_launch_sources_stream(c, ipop, nthr, iz_begin, streams, stream_len, nside_base, nsrc_out);
}

int clr_fetch_sources(clr_ctx *c, int ipop, float *pos4, int32_t *ipix, long long n) {
  if (!c) return CLR_ERR_ARG;
  if (ipop < 0 || ipop >= CLR_NPOP_MAX) return clr_fail(c, CLR_ERR_ARG, "bad population %d", ipop);
  CatalogDev &C = c->cat[ipop];
  if (n != C.n) return clr_fail(c, CLR_ERR_ARG, "population %d has %lld sources, not %lld", ipop, C.n, n);
  if (n == 0) return CLR_OK;
  CLR_HIP(c, hipSetDevice(c->device));
  if (pos4) CLR_HIP(c, hipMemcpy(pos4, C.pos, (size_t)n * 4 * sizeof(float), hipMemcpyDeviceToHost));
  if (ipix) CLR_HIP(c, hipMemcpy(ipix, C.ipix, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
  return CLR_OK;
}

int clr_catalogue_size(clr_ctx *c, int ipop, long long *n) {
  if (!c || !n) return CLR_ERR_ARG;
  if (ipop < 0 || ipop >= CLR_NPOP_MAX) return clr_fail(c, CLR_ERR_ARG, "bad population %d", ipop);
  *n = c->cat[ipop].n;
  return CLR_OK;
}

int clr_source_properties(clr_ctx *c, int ipop, float *radecz, long long n) {
  if (!c || !radecz) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_launch_source_props(c, ipop, radecz, n);
}

int clr_kappa(clr_ctx *c, int nside, int nplanes, const double *r_src, float *out) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_launch_los(c, 0, nside, nplanes, r_src, out);
}

int clr_isw(clr_ctx *c, int nside, int nplanes, const double *r_src, float *out) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_launch_los(c, 1, nside, nplanes, r_src, out);
}

int clr_source_beams(clr_ctx *c, int ipop, const double *r_arr_a2r, float *dz_rsd, float *lens, long long n) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_launch_src_beams(c, ipop, r_arr_a2r, dz_rsd, lens, n);
}

int clr_lpt_density(clr_ctx *c, int order, int interp, float *pos_out) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_launch_lpt(c, order, interp, pos_out);
}

int clr_custom_map(clr_ctx *c, int ipop, int nside, float *out) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_launch_cstm(c, ipop, nside, out);
}

int clr_source_skewers(clr_ctx *c, int ipop, int gaussian, double sigma2, float *dens_or_gauss, float *v,
                       long long n) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_launch_src_skewers(c, ipop, gaussian, sigma2, dens_or_gauss, v, n);
}

int clr_imap(clr_ctx *c, int ipop, int nside, int nr, const double *r0, const double *rf, float *data,
             int32_t *nadd) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_launch_imap(c, ipop, nside, nr, r0, rf, data, nadd);
}

int clr_fetch_grid(clr_ctx *c, int which, float *host) {
  if (!c || !host) return CLR_ERR_ARG;
  if (which != CLR_GRID_DENS && which != CLR_GRID_NPOT) return clr_fail(c, CLR_ERR_ARG, "bad grid %d", which);
  CLR_HIP(c, hipSetDevice(c->device));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  int rc = clr_phi_ready(c, 1);
  if (rc) return rc;
  // device rows are grid_pitch(N) floats, host rows 2(N/2+1)
  const size_t hrow = (size_t)2 * c->NH * sizeof(float), drow = (size_t)c->ngx * sizeof(float);
  CLR_HIP(c, hipMemcpy2D(host, hrow, which == CLR_GRID_DENS ? c->dens : c->npot, drow, hrow, (size_t)c->N * c->N,
                         hipMemcpyDeviceToHost));
  return CLR_OK;
}

int clr_fetch_grid_slab(clr_ctx *c, int which, int iz0, int nz, float *host) {
  if (!c || !host) return CLR_ERR_ARG;
  if (which != CLR_GRID_DENS && which != CLR_GRID_NPOT) return clr_fail(c, CLR_ERR_ARG, "bad grid %d", which);
  if (iz0 < 0 || nz < 0 || iz0 + nz > c->N) return clr_fail(c, CLR_ERR_ARG, "planes %d..%d outside the grid", iz0, iz0 + nz);
  CLR_HIP(c, hipSetDevice(c->device));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  int rc = clr_phi_ready(c, 1);
  if (rc) return rc;
  if (nz == 0) return CLR_OK;
  const size_t hrow = (size_t)2 * c->NH * sizeof(float), drow = (size_t)c->ngx * sizeof(float);
  const float *src = (which == CLR_GRID_DENS ? c->dens : c->npot) + (size_t)iz0 * c->N * c->ngx;
  CLR_HIP(c, hipMemcpy2D(host, hrow, src, drow, hrow, (size_t)nz * c->N, hipMemcpyDeviceToHost));
  return CLR_OK;
}

int clr_fetch_cell_counts(clr_ctx *c, int ipop, int iz0, int nz, int32_t *counts) {
  if (!c || !counts) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_launch_fetch_counts(c, ipop, iz0, nz, counts);
}

int clr_upload_grid(clr_ctx *c, int which, const float *host) {
  if (!c || !host) return CLR_ERR_ARG;
  if (which != CLR_GRID_DENS && which != CLR_GRID_NPOT) return clr_fail(c, CLR_ERR_ARG, "bad grid %d", which);
  CLR_HIP(c, hipSetDevice(c->device));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  int rc = clr_phi_ready(c, 1);
  if (rc) return rc;
  const size_t hrow = (size_t)2 * c->NH * sizeof(float), drow = (size_t)c->ngx * sizeof(float);
  CLR_HIP(c, hipMemcpy2D(which == CLR_GRID_DENS ? c->dens : c->npot, drow, host, hrow, hrow, (size_t)c->N * c->N,
                         hipMemcpyHostToDevice));
  return CLR_OK;
}

int clr_declare_grids(clr_ctx *c, double sigma2, int density_transformed) {
  if (!c) return CLR_ERR_ARG;
  int rc = clr_phi_ready(c, 1);
  if (rc) return rc;
  c->have_fields = 1;
  c->sigma2 = sigma2;
  c->density_done = density_transformed ? 1 : 0;
  c->norms_done = 0;
  return CLR_OK;
}

int clr_synchronize(clr_ctx *c) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  int rc = clr_phi_ready(c, 1);
  if (rc) return rc;
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  return CLR_OK;
}

void *clr_host_alloc(size_t bytes) {
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}

int clr_host_free(void *p) {
  if (!p) return CLR_OK;
  return hipHostFree(p) == hipSuccess ? CLR_OK : CLR_ERR_HIP;
}

int clr_stage_times(const clr_ctx *c, double *ms, int n) {
  if (!c || !ms) return CLR_ERR_ARG;
  for (int i = 0; i < n && i < ST_COUNT; i++) ms[i] = c->stage_ms[i];
  // phi's passes: read without waiting (the last finished value otherwise)
  if (c->phi_pending && ST_FFT_PHI < n && hipEventQuery(c->ev5) == hipSuccess) {
    float f = 0;
    if (hipEventElapsedTime(&f, c->ev4, c->ev5) == hipSuccess) ms[ST_FFT_PHI] = f;
    if (ST_PHI_X < n && hipEventElapsedTime(&f, c->ev4, c->ev6) == hipSuccess) ms[ST_PHI_Y] = f;
    if (ST_PHI_X < n && hipEventElapsedTime(&f, c->ev6, c->ev5) == hipSuccess) ms[ST_PHI_X] = f;
  }
  return ST_COUNT;
}

}  // extern "C"
