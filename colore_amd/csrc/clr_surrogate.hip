// clr_surrogate.hip -- surrogate-fused source sampling: the per-shell bias
// normalisations of compute_density_normalization (density.c:1128-1393)
// without the shell-sum pass over the transformed grid, so that
// srcs_set_cartesian_single (srcs.c:120-283) runs right after the Gaussian
// field: the lognormal transform of lognormalize (density.c:1070-1103) is
// applied per cell inside the count pass.
//
// A shell's norm is count / sum_cells bias_model(delta).  With G = D(r)
// delta_G the Gaussian variate a cell's lognormal transform exponentiates,
// delta = exp(G - sigma_c^2/2) - 1, the sum is the shell's count times the
// mean of f(G) = bias_model(exp(G - sigma_c^2/2) - 1) over the shell's
// cells.  A read-only pass over the Gaussian grid (clr_launch_shell_moments)
// measures each shell's sample cumulants of G (mean, variance, third and
// fourth); the closure evaluates the mean of f for a distribution with the
// first three (closed form for the exponential model, Edgeworth-weighted
// quadrature for the truncated ones); the learned network (clr_mlp.hip)
// predicts the rest of ln norm from (b, sigma_c, skewness, excess kurtosis)
// for the exponential model, whose closure (the cumulant series cut after
// the third term) leaves the fourth and higher cumulants' part.
// The exact two-pass reduction (clr_shell_norms) stays the fallback and the
// judge of the number-count error.
#include "clr_internal.h"
#include <math.h>
#include <algorithm>
#include <vector>

namespace {

constexpr int kMomQ = 6;   // per shell: count, sums of G, G^2, G^3, G^4, D^2

// r of redshift z on the FP64 tables (z(r) increases with r): the inverse of
// the linear interpolation f_of_r applies.
double r_of_z_tab(const clr_ctx *c, double z) {
  const std::vector<double> &zz = c->h_z, &rr = c->h_r;
  const int na = (int)zz.size();
  if (z <= zz[0]) return rr[0];
  if (z >= zz[na - 1]) return rr[na - 1];
  int lo = 0, hi = na - 1;
  while (hi > lo + 1) {
    const int m = (lo + hi) / 2;
    if (zz[m] > z) hi = m; else lo = m;
  }
  const double t = (z - zz[lo]) / (zz[hi] - zz[lo]);
  return rr[lo] + t * (rr[hi] - rr[lo]);
}

// Network inputs and closure of population ipop for every shell with cells
// (bins 1 .. nz-1), from the moments pass mom[nz][kMomQ]: b at the shell's
// mean redshift (from the cached shell geometry); sigma_c^2 = <D^2> sigma^2
// (the variance the lognormal transform subtracts, sigma^2 the measured
// variance of the Gaussian grid); the sample mean, variance, third and fourth
// cumulants of G.  x = (b, sigma_c,
// skewness, excess kurtosis); base = -ln of the closure's mean of f.
int shell_inputs(clr_ctx *c, int ipop, const std::vector<double> &mom, std::vector<float> &x,
                 std::vector<double> &base, std::vector<int> &idx) {
  const TracerDev &T = c->tr[CLR_TRACER_SRCS][ipop];
  if (!T.set || T.h_bz.size() != (size_t)c->na) return clr_fail(c, CLR_ERR_STATE, "population %d has no tables", ipop);
  const int nz = c->geo_nz;
  if (mom.size() != (size_t)nz * kMomQ) return clr_fail(c, CLR_ERR_STATE, "shell moments missing");
  x.clear();
  base.clear();
  idx.clear();
  for (int s = 1; s < nz; s++) {
    const double *m = &mom[(size_t)s * kMomQ];
    const double n = m[0];
    if (!(n > 0) || !(c->geo_cnt[s] > 0)) continue;
    const double r = r_of_z_tab(c, c->geo_zsum[s] / c->geo_cnt[s]);
    const double b = clr::f_of_r(c->h_r.data(), T.h_bz.data(), c->glob_idr, c->r_last, r, T.bz0, 1.0);
    const double mu = m[1] / n, e2 = m[2] / n, e3 = m[3] / n, e4 = m[4] / n;
    const double var = std::max(e2 - mu * mu, 0.0);
    const double k3 = e3 - 3 * mu * e2 + 2 * mu * mu * mu;
    const double k4 = e4 - 4 * e3 * mu - 3 * e2 * e2 + 12 * e2 * mu * mu - 6 * mu * mu * mu * mu;
    const double sc = sqrt(std::max(m[5] / n, 0.0) * c->sigma2);
    const double mean_f = clr_bias_moment(T.model, b, sc, mu, var, k3);
    if (!(mean_f > 0) || !std::isfinite(mean_f))
      return clr_fail(c, CLR_ERR_STATE, "shell %d of population %d: closure mean %g", s, ipop, mean_f);
    x.push_back((float)b);
    x.push_back((float)sc);
    x.push_back(var > 0 ? (float)(k3 / (var * sqrt(var))) : 0.f);
    x.push_back(var > 0 ? (float)(k4 / (var * var)) : 0.f);
    base.push_back(-log(mean_f));
    idx.push_back(s);
  }
  return CLR_OK;
}

// The moments pass on the cached shell geometry.
int shell_moments(clr_ctx *c, std::vector<double> &mom) {
  int rc = clr_launch_shell_geometry(c);
  if (rc) return rc;
  rc = clr_launch_shell_moments(c, mom);
  if (rc) return rc;
  if (mom.size() != (size_t)c->geo_nz * kMomQ) return clr_fail(c, CLR_ERR_STATE, "shell moments: bad size");
  return CLR_OK;
}

// 16-point Gauss-Legendre nodes and weights on [-1, 1]: roots of P_16 by
// Newton's iteration from x = cos(pi (i - 1/4) / (n + 1/2)), weights
// 2 / ((1 - x^2) P_16'(x)^2).
struct GaussLegendre16 {
  double x[16], w[16];
  GaussLegendre16() {
    const int n = 16;
    for (int i = 0; i < n / 2; i++) {
      double z = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0;
      for (int it = 0; it < 100; it++) {
        double p0 = 1, p1 = 0;
        for (int j = 1; j <= n; j++) {
          const double p2 = p1;
          p1 = p0;
          p0 = ((2 * j - 1) * z * p1 - (j - 1) * p2) / j;
        }
        dp = n * (z * p0 - p1) / (z * z - 1);
        const double dz = p0 / dp;
        z -= dz;
        if (fabs(dz) < 1e-16) break;
      }
      x[i] = -z;
      x[n - 1 - i] = z;
      w[i] = w[n - 1 - i] = 2 / ((1 - z * z) * dp * dp);
    }
  }
};

}  // namespace

extern "C" {

double clr_bias_moment(int model, double b, double sigma_c, double mu, double var, double k3) {
  const double sc2 = sigma_c * sigma_c;
  if (model != CLR_BIAS_EXP_TRUNCATED && model != CLR_BIAS_TRUNCATED)
    return exp(b * mu + 0.5 * b * b * var + b * b * b * k3 / 6 - 0.5 * b * sc2);
  const double sd = var > 0 ? sqrt(var) : 0.0;
  if (!(sd > 0)) return clr::bias_model_f64(expm1(mu - 0.5 * sc2), b, model);
  const double g1 = k3 / (sd * sd * sd);
  static const GaussLegendre16 gl;
  // t = (G - mu) / sd over [-12, 12], split where bias_model has a kink: the
  // truncation 1 + b delta = 0, or delta = 0 for the exp-truncated model
  double cut[3] = {-12.0, 12.0, 0};
  int nc = 2;
  if (model == CLR_BIAS_TRUNCATED && (b > 1 || b < 0)) cut[nc++] = (log(1 - 1 / b) + 0.5 * sc2 - mu) / sd;
  if (model == CLR_BIAS_EXP_TRUNCATED) cut[nc++] = (0.5 * sc2 - mu) / sd;
  std::sort(cut, cut + nc);
  const double inv_sqrt2pi = 0.3989422804014327;
  double acc = 0;
  for (int c = 0; c + 1 < nc; c++) {
    const double a0 = std::max(cut[c], -12.0), a1 = std::min(cut[c + 1], 12.0);
    if (!(a1 > a0)) continue;
    const int m = 24;
    const double h = (a1 - a0) / m;
    for (int j = 0; j < m; j++) {
      const double mid = a0 + (j + 0.5) * h;
      for (int q = 0; q < 16; q++) {
        const double t = mid + 0.5 * h * gl.x[q];
        // Edgeworth weight: the normal density times 1 + g1/6 He_3(t)
        const double w = inv_sqrt2pi * exp(-0.5 * t * t) * (1 + g1 / 6 * (t * t * t - 3 * t));
        acc += 0.5 * h * gl.w[q] * w * clr::bias_model_f64(expm1(mu + sd * t - 0.5 * sc2), b, model);
      }
    }
  }
  return acc;
}

int clr_surrogate_inputs(clr_ctx *c, int ipop, int cap, float *x, double *base, int *shell, int *n_out) {
  if (!c) return CLR_ERR_ARG;
  if (ipop < 0 || ipop >= c->n_tr[CLR_TRACER_SRCS]) return clr_fail(c, CLR_ERR_ARG, "no source population %d", ipop);
  if (!c->have_fields || c->density_done)
    return clr_fail(c, CLR_ERR_STATE,
                    "the surrogate's inputs need the Gaussian field (fields done, density not transformed)");
  CLR_HIP(c, hipSetDevice(c->device));
  std::vector<double> mom;
  int rc = shell_moments(c, mom);
  if (rc) return rc;
  std::vector<float> xs;
  std::vector<double> bs;
  std::vector<int> idx;
  rc = shell_inputs(c, ipop, mom, xs, bs, idx);
  if (rc) return rc;
  const int n = (int)idx.size();
  if (n_out) *n_out = n;
  if ((x || base || shell) && cap < n) return clr_fail(c, CLR_ERR_ARG, "capacity %d < %d shells", cap, n);
  for (int i = 0; i < n; i++) {
    if (x)
      for (int k = 0; k < CLR_MLP_IN; k++) x[(size_t)i * CLR_MLP_IN + k] = xs[(size_t)i * CLR_MLP_IN + k];
    if (base) base[i] = bs[i];
    if (shell) shell[i] = idx[i];
  }
  return CLR_OK;
}

int clr_sample_sources_surrogate(clr_ctx *c, clr_mlp *net, int ipop, uint32_t seed, int nside_base, int precision,
                                 long long *nsrc_out) {
  if (!c) return CLR_ERR_ARG;
  const int nall = c->n_tr[CLR_TRACER_SRCS];
  if (ipop < -1 || ipop >= nall) return clr_fail(c, CLR_ERR_ARG, "no source population %d", ipop);
  if (!c->have_fields || c->density_done)
    return clr_fail(c, CLR_ERR_STATE, "surrogate sampling needs the Gaussian field (fields done, density not transformed)");
  CLR_HIP(c, hipSetDevice(c->device));
  int rc = clr_launch_shell_geometry(c);
  if (rc) return rc;
  const int p0 = ipop < 0 ? 0 : ipop, p1 = ipop < 0 ? nall : ipop + 1, np = p1 - p0, nz = c->geo_nz;
  CLR_HIP(c, hipEventRecord(c->ev0, c->stream));
  std::vector<double> mom;
  rc = shell_moments(c, mom);
  if (rc) return rc;
  // every population's shells in one inference batch
  std::vector<float> x, xs;
  std::vector<double> base, bs;
  std::vector<int> idx, owner;
  for (int p = p0; p < p1; p++) {
    std::vector<int> ip;
    rc = shell_inputs(c, p, mom, xs, bs, ip);
    if (rc) return rc;
    x.insert(x.end(), xs.begin(), xs.end());
    base.insert(base.end(), bs.begin(), bs.end());
    idx.insert(idx.end(), ip.begin(), ip.end());
    owner.insert(owner.end(), ip.size(), p - p0);
  }
  const int n = (int)idx.size();
  std::vector<float> y((size_t)n * CLR_MLP_OUT, 0.f);
  if (n > 0 && net) {
    rc = clr_ensure(c, &c->misc, &c->misc_n, (size_t)n * (CLR_MLP_IN + CLR_MLP_OUT) * sizeof(float));
    if (rc) return rc;
    float *dx = (float *)c->misc, *dy = dx + (size_t)n * CLR_MLP_IN;
    CLR_HIP(c, hipMemcpyAsync(dx, x.data(), x.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (clr_mlp_predict_device(net, dx, n, dy, precision, c->stream))
      return clr_fail(c, CLR_ERR_HIP, "surrogate inference failed: %s", clr_mlp_last_error(net));
    CLR_HIP(c, hipMemcpyAsync(y.data(), dy, y.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    CLR_HIP(c, hipStreamSynchronize(c->stream));
  }
  // per-shell norms: ln norm = closure + the network's correction y0 for
  // populations with the exponential model (the truncated models' closures
  // take none); the closure alone without a network; shells without cells
  // keep 1 (the edge fill of the norm table replaces bin 0 and the last bin
  // anyway)
  std::vector<double> norm_s((size_t)np * nz, 1.0);
  for (int i = 0; i < n; i++) {
    const bool corr = c->tr[CLR_TRACER_SRCS][p0 + owner[i]].model == CLR_BIAS_EXPONENTIAL;
    const double v = exp((corr ? (double)y[(size_t)i * CLR_MLP_OUT] : 0.0) + base[i]);
    if (!(v > 0) || !std::isfinite(v)) return clr_fail(c, CLR_ERR_STATE, "surrogate predicted norm %g", v);
    norm_s[(size_t)owner[i] * nz + idx[i]] = v;
  }
  std::vector<int> kinds(np, CLR_TRACER_SRCS), ipops(np);
  for (int p = 0; p < np; p++) ipops[p] = p0 + p;
  rc = clr_launch_norm_tables_given(c, np, kinds.data(), ipops.data(), norm_s.data());
  if (rc) return rc;
  CLR_HIP(c, hipEventRecord(c->ev1, c->stream));
  CLR_HIP(c, hipEventSynchronize(c->ev1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
  const double t_norm = ms;
  rc = clr_launch_sources(c, ipop, seed, nside_base, nsrc_out, 1);
  if (rc) return rc;
  c->stage_ms[ST_NORM] = t_norm;
  c->norms_done = 1;
  return CLR_OK;
}

}  // extern "C"
