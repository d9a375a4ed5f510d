// clr_los.hip -- kappa_get_beam_properties (kappa.c:78-175) and
// isw_get_beam_properties (isw.c:78-147) on gfx950.
//
// One thread per HEALPix pixel (NEST order, so neighbouring lanes march
// neighbouring rays and their stencil reads share cache lines).  The march
// uses nr = N/2 radial steps of dr = r_max/nr (get_radial_params,
// common.c:333-337); at each step the NGP cell of (rm u + x_obs)/dx is read
// (interpolate_from_grid, beaming.c:120-268, periodic in x, y and z on one
// GPU).  kappa: transverse trace of the 19-point tidal tensor of phi
// (get_element, beaming.c:31-117) weighted by rm D (1+z) dr and rm^2 D (1+z)
// dr, cumulative over the source planes sorted by distance; ISW: phi weighted
// by 2 PD(rm) dr.  Accumulation in FP64.
#include "clr_internal.h"
#include <algorithm>
#include <math.h>

using namespace clr;

namespace {

constexpr int kMaxPlanes = 100;   // NPLANES_MAX, common.h:133

struct LosArgs {
  const float *phi;
  int N, ngx;
  double idx, dr, ox, oy, oz;
  int nr, nplanes, which;
  int irmin[kMaxPlanes], irmax[kMaxPlanes];
  double inv_rmax[kMaxPlanes];
};

__device__ inline int wrapi(long v, int N) {
  if (v >= N) v -= N;
  else if (v < 0) v += N;
  return (int)v;
}

__global__ __launch_bounds__(256) void k_los(LosArgs a, const double *__restrict__ fac1,
                                             const double *__restrict__ fac2, int64_t nside, int64_t npix,
                                             float *__restrict__ out) {
#pragma clang fp contract(off)
  const int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= npix) return;
  double u[3];
  pix2vec_nest(nside, ip, u);
  const int N = a.N;
  const size_t ngx = (size_t)a.ngx, plane = (size_t)N * ngx;
  const float *__restrict__ phi = a.phi;
  double rot[6];
  if (a.which == 0) {
    double cth = u[2];
    if (cth >= 1) cth = 1;
    if (cth <= -1) cth = -1;
    const double sth = sqrt((1 - cth) * (1 + cth));
    double cph = 1, sph = 0;
    if (sth != 0) { cph = u[0] / sth; sph = u[1] / sth; }
    const double pre = a.idx * a.idx;
    rot[0] = (cth * cth * cph * cph + sph * sph) * pre;
    rot[1] = (2 * cph * sph * (cth * cth - 1)) * pre;
    rot[2] = (-2 * cth * sth * cph) * pre;
    rot[3] = (cth * cth * sph * sph + cph * cph) * pre;
    rot[4] = (-2 * cth * sth * sph) * pre;
    rot[5] = (sth * sth) * pre;
  }
  double k1 = 0, k2 = 0;
  const double obs[3] = {a.ox, a.oy, a.oz};
  for (int pl = 0; pl < a.nplanes; pl++) {
    for (int irr = a.irmin[pl]; irr <= a.irmax[pl]; irr++) {
      const double rm = (irr + 0.5) * a.dr;
      int ix[3];
      for (int ax = 0; ax < 3; ax++) {
        const double xn = (rm * u[ax] + obs[ax]) * a.idx;
        ix[ax] = wrapi((long)(xn + 0.5), N);
      }
      const size_t o = (size_t)ix[2] * plane + (size_t)ix[1] * ngx + ix[0];
      if (a.which == 1) {
        k1 += (double)phi[o] * fac1[irr];
        continue;
      }
      const int xh = ix[0] == N - 1 ? 0 : ix[0] + 1, xl = ix[0] == 0 ? N - 1 : ix[0] - 1;
      const int yh = ix[1] == N - 1 ? 0 : ix[1] + 1, yl = ix[1] == 0 ? N - 1 : ix[1] - 1;
      const int zh = ix[2] == N - 1 ? 0 : ix[2] + 1, zl = ix[2] == 0 ? N - 1 : ix[2] - 1;
      const size_t Z0 = (size_t)ix[2] * plane, ZH = (size_t)zh * plane, ZL = (size_t)zl * plane;
      const size_t Y0 = (size_t)ix[1] * ngx, YH = (size_t)yh * ngx, YL = (size_t)yl * ngx;
      const double c0 = phi[ix[0] + Y0 + Z0];
      double t[6];
      t[0] = ((double)phi[xh + Y0 + Z0] + (double)phi[xl + Y0 + Z0] - 2 * c0);
      t[3] = ((double)phi[ix[0] + YH + Z0] + (double)phi[ix[0] + YL + Z0] - 2 * c0);
      t[1] = 0.25 * ((double)phi[xh + YH + Z0] + (double)phi[xl + YL + Z0] - (double)phi[xh + YL + Z0] -
                     (double)phi[xl + YH + Z0]);
      t[5] = ((double)phi[ix[0] + Y0 + ZH] + (double)phi[ix[0] + Y0 + ZL] - 2 * c0);
      t[2] = 0.25 * ((double)phi[xh + Y0 + ZH] + (double)phi[xl + Y0 + ZL] - (double)phi[xh + Y0 + ZL] -
                     (double)phi[xl + Y0 + ZH]);
      t[4] = 0.25 * ((double)phi[ix[0] + YH + ZH] + (double)phi[ix[0] + YL + ZL] -
                     (double)phi[ix[0] + YH + ZL] - (double)phi[ix[0] + YL + ZH]);
      double dotp = 0;
      for (int q = 0; q < 6; q++) dotp += rot[q] * t[q];
      k1 += dotp * fac1[irr];
      k2 += dotp * fac2[irr];
    }
    const double v = a.which == 0 ? (k1 - a.inv_rmax[pl] * k2) : k1;
    out[(size_t)pl * npix + ip] = (float)v;
  }
}

}  // namespace

int clr_launch_los(clr_ctx *c, int which, int nside, int nplanes, const double *r_src, float *out) {
  if (!c->have_fields) return clr_fail(c, CLR_ERR_STATE, "clr_gaussian_fields must run first");
  if (nside < 1 || nside > 8192 || (nside & (nside - 1))) return clr_fail(c, CLR_ERR_ARG, "bad nside %d", nside);
  if (nplanes < 1 || nplanes > kMaxPlanes || !r_src || !out)
    return clr_fail(c, CLR_ERR_ARG, "need 1..%d source planes", kMaxPlanes);
  const int N = c->N, na = c->na;
  const double r_max = c->r_max;
  if (!(r_max > 0)) return clr_fail(c, CLR_ERR_STATE, "tables carry no r_max");
  const int nr = N / 2;
  const double dr = r_max / nr, idr = 1. / dr;
  std::vector<double> rs(r_src, r_src + nplanes);
  std::vector<float> rsf(nplanes);
  for (int i = 0; i < nplanes; i++) rsf[i] = (float)rs[i];   // ind_sort sorts flouble values
  std::vector<int> order(nplanes);
  for (int i = 0; i < nplanes; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rsf[a] < rsf[b]; });
  LosArgs a{};
  {
    const int prc = clr_phi_ready(c, 0);   // phi's passes may still run on the second stream
    if (prc) return prc;
  }
  a.phi = c->npot;
  a.N = N;
  a.ngx = c->ngx;
  a.idx = N / c->L;
  a.dr = dr;
  a.ox = c->obs[0];
  a.oy = c->obs[1];
  a.oz = c->obs[2];
  a.nr = nr;
  a.nplanes = nplanes;
  a.which = which;
  for (int i = 0; i < nplanes; i++) {
    const int ih = (int)(rs[order[i]] * idr + 0.5);
    a.inv_rmax[i] = 1. / (ih * dr);
    a.irmax[i] = std::min(ih, nr - 1);
  }
  a.irmin[0] = 0;
  for (int i = 1; i < nplanes; i++) a.irmin[i] = a.irmax[i - 1] + 1;

  // radial kernels (FP64, host): need D, z and PD at rm -> from host copies
  std::vector<double> rr(na), zz(na), dd(na), pd(na);
  CLR_HIP(c, hipMemcpy(rr.data(), c->bg, na * sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipMemcpy(zz.data(), c->bg + na, na * sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipMemcpy(dd.data(), c->bg + 2 * (size_t)na, na * sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipMemcpy(pd.data(), c->bg + 4 * (size_t)na, na * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<double> f1(nr), f2(nr);
  for (int i = 0; i < nr; i++) {
    const double rm = (i + 0.5) * dr;
    if (which == 0) {
      const double pg = f_of_r(rr.data(), dd.data(), c->glob_idr, c->r_last, rm, 1.0, dd[na - 1]) *
                        (1 + f_of_r(rr.data(), zz.data(), c->glob_idr, c->r_last, rm, 0.0, zz[na - 1]));
      f1[i] = rm * pg * dr;
      f2[i] = rm * rm * pg * dr;
    } else {
      const double pdg = f_of_r(rr.data(), pd.data(), c->glob_idr, c->r_last, rm, pd[0], pd[na - 1]);
      f1[i] = 2 * pdg * dr;
      f2[i] = 0;
    }
  }
  const int64_t npix = 12LL * nside * nside;
  const size_t need = 2 * (size_t)nr * sizeof(double) + (size_t)npix * nplanes * sizeof(float);
  int rc = clr_ensure(c, &c->misc, &c->misc_n, need);
  if (rc) return rc;
  double *df = (double *)c->misc;
  float *dout = (float *)(df + 2 * nr);
  CLR_HIP(c, hipMemcpyAsync(df, f1.data(), nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
  CLR_HIP(c, hipMemcpyAsync(df + nr, f2.data(), nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
  CLR_HIP(c, hipEventRecord(c->ev0, c->stream));
  hipLaunchKernelGGL(k_los, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, c->stream, a, df, df + nr,
                     (int64_t)nside, npix, dout);
  CLR_HIP(c, hipGetLastError());
  CLR_HIP(c, hipEventRecord(c->ev1, c->stream));
  CLR_HIP(c, hipMemcpyAsync(out, dout, (size_t)npix * nplanes * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->stage_ms[which == 0 ? ST_KAPPA : ST_ISW] = ms;
  return CLR_OK;
}
