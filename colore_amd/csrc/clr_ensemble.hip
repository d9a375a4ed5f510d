// clr_ensemble.hip -- ensembles of independent realisations (each its own
// context: its own grid size, box, tables and seed) run as ONE captured HIP
// graph: for every realisation the field stage (fourier.c:361-423), the
// lognormal transform fused with the shell sums and norm tables
// (density.c:1070-1103, 1128-1393), the sampling of every source population
// (srcs.c:120-283) and a kappa map (kappa.c:78-175), as kernel launches only
// (no allocation, no memset node, no host wait inside the graph).  The graph
// is captured once per seed set, instantiated, and replayed by every later
// clr_ensemble_run with the same seeds; realisations are spread round-robin
// over `nstreams` captured streams, so independent realisations overlap on
// the device.  Each realisation's results equal a single-context run of the
// same stages bit for bit (same kernels, same reductions).
#include "clr_internal.h"
#include <algorithm>
#include <string>
#include <vector>

struct clr_ensemble {
  int device = 0;
  int n = 0;
  std::vector<clr_ctx *> ctx;
  int nside_base = 2, nside_kappa = 0;
  std::vector<double> r_kappa;
  hipStream_t origin = nullptr;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> ev_join;
  hipEvent_t ev_fork = nullptr, ev_a = nullptr, ev_b = nullptr;
  double *sigma2_dev = nullptr;
  int *flags_dev = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint32_t> seeds;
  std::vector<double> r_smooth;
  int smooth_potential = 1;
  std::vector<int> flags;
  std::vector<double> sigma2;
  float ms = 0;
  std::string err;
};

namespace {

int ens_fail(clr_ensemble *e, int code, const std::string &msg) {
  if (e) e->err = msg;
  return code;
}
#define ENS_HIP(e, call)                                                                         \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) return ens_fail((e), CLR_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

__global__ void k_zero_ints(int *__restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

void drop_graph(clr_ensemble *e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
  e->seeds.clear();
}

int capture(clr_ensemble *e) {
  const int ns = (int)e->streams.size();
  ENS_HIP(e, hipStreamBeginCapture(e->origin, hipStreamCaptureModeThreadLocal));
  int rc = CLR_OK;
  hipLaunchKernelGGL(k_zero_ints, dim3((e->n + 255) / 256), dim3(256), 0, e->origin, e->flags_dev, e->n);
  hipError_t he = hipEventRecord(e->ev_fork, e->origin);
  for (int k = 0; k < ns && he == hipSuccess; k++) he = hipStreamWaitEvent(e->streams[k], e->ev_fork, 0);
  for (int i = 0; i < e->n && he == hipSuccess && rc == CLR_OK; i++) {
    clr_ctx *c = e->ctx[i];
    hipStream_t s = e->streams[i % ns];
    rc = clr_enqueue_fields(c, s, e->seeds[i], e->r_smooth[i], e->smooth_potential, e->sigma2_dev + i);
    if (!rc) rc = clr_enqueue_norms(c, s, e->sigma2_dev + i);
    if (!rc) rc = clr_enqueue_sources(c, s, e->seeds[i], e->nside_base, e->flags_dev + i);
    if (!rc && e->nside_kappa > 0) rc = clr_enqueue_kappa(c, s);
    if (rc) e->err = std::string("realisation ") + std::to_string(i) + ": " + c->err;
  }
  for (int k = 0; k < ns && he == hipSuccess; k++) {
    he = hipEventRecord(e->ev_join[k], e->streams[k]);
    if (he == hipSuccess) he = hipStreamWaitEvent(e->origin, e->ev_join[k], 0);
  }
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(e->origin, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (he != hipSuccess || ee != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return ens_fail(e, CLR_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(he != hipSuccess ? he : ee));
  }
  e->graph = g;
  ENS_HIP(e, hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
  return CLR_OK;
}

}  // namespace

extern "C" {

int clr_ensemble_create(clr_ensemble **out, int n, clr_ctx *const *ctxs, int nside_base, int nside_kappa,
                        const double *r_kappa, const long long *src_cap, int nstreams) {
  if (!out) return CLR_ERR_ARG;
  clr_ensemble *e = new clr_ensemble();
  *out = e;
  if (n < 1 || !ctxs || !src_cap || nstreams < 1 || nstreams > 64 || (nside_kappa > 0 && !r_kappa))
    return ens_fail(e, CLR_ERR_ARG, "bad ensemble arguments");
  e->n = n;
  e->device = ctxs[0]->device;
  e->nside_base = nside_base;
  e->nside_kappa = nside_kappa;
  ENS_HIP(e, hipSetDevice(e->device));
  for (int i = 0; i < n; i++) {
    clr_ctx *c = ctxs[i];
    if (!c || c->device != e->device) return ens_fail(e, CLR_ERR_ARG, "every context must be on the same device");
    if (!c->tables_set || c->n_tr[CLR_TRACER_SRCS] < 1)
      return ens_fail(e, CLR_ERR_STATE, "realisation " + std::to_string(i) + " has no tables or populations");
    int rc = clr_phi_ready(c, 1);
    if (!rc) rc = clr_prepare_fields(c);
    if (!rc) rc = clr_prepare_norms(c);
    if (!rc) {
      std::vector<long long> cap(c->n_tr[CLR_TRACER_SRCS], src_cap[i]);
      rc = clr_prepare_sources(c, nside_base, cap.data());
    }
    if (!rc && nside_kappa > 0) rc = clr_prepare_kappa(c, nside_kappa, 1, &r_kappa[i]);
    if (rc) return ens_fail(e, rc, "realisation " + std::to_string(i) + ": " + c->err);
    e->ctx.push_back(c);
    e->r_kappa.push_back(nside_kappa > 0 ? r_kappa[i] : 0.0);
  }
  ENS_HIP(e, hipStreamCreateWithFlags(&e->origin, hipStreamNonBlocking));
  e->streams.resize(nstreams);
  e->ev_join.resize(nstreams);
  for (int k = 0; k < nstreams; k++) {
    ENS_HIP(e, hipStreamCreateWithFlags(&e->streams[k], hipStreamNonBlocking));
    ENS_HIP(e, hipEventCreateWithFlags(&e->ev_join[k], hipEventDisableTiming));
  }
  ENS_HIP(e, hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming));
  ENS_HIP(e, hipEventCreate(&e->ev_a));
  ENS_HIP(e, hipEventCreate(&e->ev_b));
  ENS_HIP(e, hipMalloc((void **)&e->sigma2_dev, n * sizeof(double)));
  ENS_HIP(e, hipMalloc((void **)&e->flags_dev, n * sizeof(int)));
  e->flags.assign(n, 0);
  e->sigma2.assign(n, 0.0);
  return CLR_OK;
}

void clr_ensemble_destroy(clr_ensemble *e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->origin) (void)hipStreamSynchronize(e->origin);
  drop_graph(e);
  for (hipStream_t s : e->streams)
    if (s) (void)hipStreamDestroy(s);
  for (hipEvent_t v : e->ev_join)
    if (v) (void)hipEventDestroy(v);
  for (hipEvent_t v : {e->ev_fork, e->ev_a, e->ev_b})
    if (v) (void)hipEventDestroy(v);
  if (e->origin) (void)hipStreamDestroy(e->origin);
  if (e->sigma2_dev) (void)hipFree(e->sigma2_dev);
  if (e->flags_dev) (void)hipFree(e->flags_dev);
  delete e;
}

const char *clr_ensemble_last_error(const clr_ensemble *e) { return e ? e->err.c_str() : "null ensemble"; }

int clr_ensemble_run(clr_ensemble *e, const uint32_t *seeds, const double *r_smooth, int smooth_potential,
                     int *n_invalid, float *ms) {
  if (!e || !seeds || !r_smooth) return CLR_ERR_ARG;
  ENS_HIP(e, hipSetDevice(e->device));
  const bool same = e->exec && e->smooth_potential == (smooth_potential ? 1 : 0) &&
                    std::equal(seeds, seeds + e->n, e->seeds.begin()) &&
                    std::equal(r_smooth, r_smooth + e->n, e->r_smooth.begin());
  if (!same) {
    drop_graph(e);
    e->seeds.assign(seeds, seeds + e->n);
    e->r_smooth.assign(r_smooth, r_smooth + e->n);
    e->smooth_potential = smooth_potential ? 1 : 0;
    int rc = capture(e);
    if (rc) {
      drop_graph(e);
      return rc;
    }
  }
  ENS_HIP(e, hipEventRecord(e->ev_a, e->origin));
  ENS_HIP(e, hipGraphLaunch(e->exec, e->origin));
  ENS_HIP(e, hipEventRecord(e->ev_b, e->origin));
  ENS_HIP(e, hipStreamSynchronize(e->origin));
  (void)hipEventElapsedTime(&e->ms, e->ev_a, e->ev_b);
  if (ms) *ms = e->ms;
  ENS_HIP(e, hipMemcpy(e->flags.data(), e->flags_dev, e->n * sizeof(int), hipMemcpyDeviceToHost));
  ENS_HIP(e, hipMemcpy(e->sigma2.data(), e->sigma2_dev, e->n * sizeof(double), hipMemcpyDeviceToHost));
  int bad = 0;
  for (int i = 0; i < e->n; i++) {
    clr_ctx *c = e->ctx[i];
    // the host-side state the stage calls leave
    c->sigma2 = e->sigma2[i];
    c->r_smooth = e->r_smooth[i];
    c->have_fields = 1;
    c->density_done = 1;
    c->norms_done = 1;
    for (int k = 0; k < 3; k++)
      for (int t = 0; t < c->n_tr[k]; t++) {
        TracerDev &T = c->tr[k][t];
        double ed[2];
        ENS_HIP(e, hipMemcpy(ed, T.edge, sizeof(ed), hipMemcpyDeviceToHost));
        T.norm0 = ed[0];
        T.normf = ed[1];
        T.norm_set = 1;
      }
    const int rc = clr_finish_sources(c);
    if (rc) return ens_fail(e, rc, c->err);
    if (e->flags[i]) bad++;
  }
  if (n_invalid) *n_invalid = bad;
  return CLR_OK;
}

int clr_ensemble_result(clr_ensemble *e, int i, double *sigma2, int *flag, float *kappa) {
  if (!e || i < 0 || i >= e->n) return CLR_ERR_ARG;
  if (sigma2) *sigma2 = e->sigma2[i];
  if (flag) *flag = e->flags[i];
  if (kappa) {
    ENS_HIP(e, hipSetDevice(e->device));
    const int rc = clr_fetch_kappa(e->ctx[i], kappa);
    if (rc) return ens_fail(e, rc, e->ctx[i]->err);
  }
  return CLR_OK;
}

int clr_fft_autotune(clr_ctx *c, int reps, int cap, float *ms, int *ncand, int *best) {
  if (!c) return CLR_ERR_ARG;
  CLR_HIP(c, hipSetDevice(c->device));
  return clr_autotune_fields(c, reps, cap, ms, ncand, best);
}

int clr_fft_candidates(int n_grid) { return clr_fft_candidate_count(n_grid); }

int clr_fft_select(int n_grid, int k) { return clr_fft_select_candidate(n_grid, k) ? CLR_ERR_ARG : CLR_OK; }

}  // extern "C"
