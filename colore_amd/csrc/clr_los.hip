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
#include <string.h>

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
      // the second differences in FP32, as get_element evaluates them on
      // flouble values (beaming.c:79-116)
      const float c0 = phi[ix[0] + Y0 + Z0];
      float t[6];
      t[0] = phi[xh + Y0 + Z0] + phi[xl + Y0 + Z0] - 2 * c0;
      t[3] = phi[ix[0] + YH + Z0] + phi[ix[0] + YL + Z0] - 2 * c0;
      t[1] = 0.25f * (phi[xh + YH + Z0] + phi[xl + YL + Z0] - phi[xh + YL + Z0] - phi[xl + YH + Z0]);
      t[5] = phi[ix[0] + Y0 + ZH] + phi[ix[0] + Y0 + ZL] - 2 * c0;
      t[2] = 0.25f * (phi[xh + Y0 + ZH] + phi[xl + Y0 + ZL] - phi[xh + Y0 + ZL] - phi[xl + Y0 + ZH]);
      t[4] = 0.25f * (phi[ix[0] + YH + ZH] + phi[ix[0] + YL + ZL] - phi[ix[0] + YH + ZL] - phi[ix[0] + YL + ZH]);
      double dotp = 0;
      for (int q = 0; q < 6; q++) dotp += rot[q] * (double)t[q];
      k1 += dotp * fac1[irr];
      k2 += dotp * fac2[irr];
    }
    const double v = a.which == 0 ? (k1 - a.inv_rmax[pl] * k2) : k1;
    out[(size_t)pl * npix + ip] = (float)v;
  }
}

// The transverse-Hessian inputs of get_element (beaming.c:79-116) for every
// cell a march can reach (centre within sqrt(r2lim) of the observer) once:
// t6[cell][6] = (xx, xy, xz, yy, yz, zz) second differences of phi in FP32,
// the same expressions k_los evaluates per sample, so a march over t6 reads
// one 24-byte record per step instead of 19 stencil values.
__global__ __launch_bounds__(256) void k_tidal(const float *__restrict__ phi, int N, int ngx, float dxf, float ox,
                                               float oy, float oz, float r2lim, float2 *__restrict__ t6) {
#pragma clang fp contract(off)
  const size_t row = blockIdx.x;
  const int iz = (int)(row / N), iy = (int)(row % N);
  const float yc = iy * dxf - oy, zc = iz * dxf - oz;
  const float rho2 = yc * yc + zc * zc;
  if (rho2 > r2lim) return;   // no march reaches this row
  const size_t plane = (size_t)N * ngx;
  const int yh = iy == N - 1 ? 0 : iy + 1, yl = iy == 0 ? N - 1 : iy - 1;
  const int zh = iz == N - 1 ? 0 : iz + 1, zl = iz == 0 ? N - 1 : iz - 1;
  const size_t Z0 = (size_t)iz * plane, ZH = (size_t)zh * plane, ZL = (size_t)zl * plane;
  const size_t Y0 = (size_t)iy * ngx, YH = (size_t)yh * ngx, YL = (size_t)yl * ngx;
  // 256 cells at a time: the records go through LDS so that consecutive
  // lanes store consecutive float2 of the row (a lane's own 24-byte record
  // would spread each store over 1.5 KB)
  __shared__ float2 st[256 * 3];
  __shared__ unsigned char ok[256];
  for (int x0 = 0; x0 < N; x0 += 256) {
    const int ix = x0 + threadIdx.x;
    const float xc = ix * dxf - ox;
    const bool in = ix < N && !(xc * xc + rho2 > r2lim);
    ok[threadIdx.x] = in;
    if (in) {
      const int xh = ix == N - 1 ? 0 : ix + 1, xl = ix == 0 ? N - 1 : ix - 1;
      const float c0 = phi[ix + Y0 + Z0];
      float t[6];
      t[0] = phi[xh + Y0 + Z0] + phi[xl + Y0 + Z0] - 2 * c0;
      t[3] = phi[ix + YH + Z0] + phi[ix + YL + Z0] - 2 * c0;
      t[1] = 0.25f * (phi[xh + YH + Z0] + phi[xl + YL + Z0] - phi[xh + YL + Z0] - phi[xl + YH + Z0]);
      t[5] = phi[ix + Y0 + ZH] + phi[ix + Y0 + ZL] - 2 * c0;
      t[2] = 0.25f * (phi[xh + Y0 + ZH] + phi[xl + Y0 + ZL] - phi[xh + Y0 + ZL] - phi[xl + Y0 + ZH]);
      t[4] = 0.25f * (phi[ix + YH + ZH] + phi[ix + YL + ZL] - phi[ix + YH + ZL] - phi[ix + YL + ZH]);
      st[3 * threadIdx.x] = make_float2(t[0], t[1]);
      st[3 * threadIdx.x + 1] = make_float2(t[2], t[3]);
      st[3 * threadIdx.x + 2] = make_float2(t[4], t[5]);
    }
    __syncthreads();
    float2 *o = t6 + (row * N + x0) * 3;
    const int nk = 3 * min(256, N - x0);
    for (int k = threadIdx.x; k < nk; k += 256)
      if (ok[k / 3]) o[k] = st[k];
    __syncthreads();
  }
}

// The kappa march of k_los over the precomputed t6 records.
__global__ __launch_bounds__(256) void k_los_t6(LosArgs a, const float2 *__restrict__ t6,
                                                const double *__restrict__ fac1, const double *__restrict__ fac2,
                                                int64_t nside, int64_t npix, float *__restrict__ out) {
#pragma clang fp contract(off)
  const int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= npix) return;
  double u[3];
  pix2vec_nest(nside, ip, u);
  const int N = a.N;
  double rot[6];
  {
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
      const float2 *r = t6 + (((size_t)ix[2] * N + ix[1]) * N + ix[0]) * 3;
      const float2 p0 = r[0], p1 = r[1], p2 = r[2];
      const float t[6] = {p0.x, p0.y, p1.x, p1.y, p2.x, p2.y};
      double dotp = 0;
      for (int q = 0; q < 6; q++) dotp += rot[q] * (double)t[q];
      k1 += dotp * fac1[irr];
      k2 += dotp * fac2[irr];
    }
    out[(size_t)pl * npix + ip] = (float)(k1 - a.inv_rmax[pl] * k2);
  }
}

}  // namespace

// Host part of a kappa / ISW march: the kernel arguments (planes sorted as
// kappa_beams_preproc sorts them, their radial ranges) and the radial
// weights f1, f2 (FP64, from the host copies of the tables).
static int los_setup(clr_ctx *c, int which, int nside, int nplanes, const double *r_src, LosArgs &a,
                     std::vector<double> &f12) {
  if (nside < 1 || nside > 8192 || (nside & (nside - 1))) return clr_fail(c, CLR_ERR_ARG, "bad nside %d", nside);
  if (nplanes < 1 || nplanes > kMaxPlanes || !r_src)
    return clr_fail(c, CLR_ERR_ARG, "need 1..%d source planes", kMaxPlanes);
  const int N = c->N, na = c->na;
  const double r_max = c->r_max;
  if (!(r_max > 0)) return clr_fail(c, CLR_ERR_STATE, "tables carry no r_max");
  if ((int)c->h_pd.size() != na) return clr_fail(c, CLR_ERR_STATE, "background tables missing");
  const int nr = N / 2;
  const double dr = r_max / nr, idr = 1. / dr;
  std::vector<double> rs(r_src, r_src + nplanes);
  std::vector<float> rsf(nplanes);
  for (int i = 0; i < nplanes; i++) rsf[i] = (float)rs[i];   // ind_sort sorts flouble values
  std::vector<int> order(nplanes);
  for (int i = 0; i < nplanes; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return rsf[x] < rsf[y]; });
  a = LosArgs{};
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
  const double *rr = c->h_r.data(), *zz = c->h_z.data(), *dd = c->h_d1.data(), *pd = c->h_pd.data();
  f12.assign(2 * (size_t)nr, 0.0);
  for (int i = 0; i < nr; i++) {
    const double rm = (i + 0.5) * dr;
    if (which == 0) {
      const double pg = f_of_r(rr, dd, c->glob_idr, c->r_last, rm, 1.0, dd[na - 1]) *
                        (1 + f_of_r(rr, zz, c->glob_idr, c->r_last, rm, 0.0, zz[na - 1]));
      f12[i] = rm * pg * dr;
      f12[nr + i] = rm * rm * pg * dr;
    } else {
      const double pdg = f_of_r(rr, pd, c->glob_idr, c->r_last, rm, pd[0], pd[na - 1]);
      f12[i] = 2 * pdg * dr;
      f12[nr + i] = 0;
    }
  }
  return CLR_OK;
}

int clr_launch_los(clr_ctx *c, int which, int nside, int nplanes, const double *r_src, float *out) {
  if (!c->have_fields) return clr_fail(c, CLR_ERR_STATE, "clr_gaussian_fields must run first");
  if (!out) return clr_fail(c, CLR_ERR_ARG, "no output map");
  LosArgs a{};
  std::vector<double> f12;
  int rc = los_setup(c, which, nside, nplanes, r_src, a, f12);
  if (rc) return rc;
  {
    const int prc = clr_phi_ready(c, 0);   // phi's passes may still run on the second stream
    if (prc) return prc;
  }
  const int nr = a.nr;
  const int64_t npix = 12LL * nside * nside;
  const size_t need = 2 * (size_t)nr * sizeof(double) + (size_t)npix * nplanes * sizeof(float);
  rc = clr_ensure(c, &c->misc, &c->misc_n, need);
  if (rc) return rc;
  double *df = (double *)c->misc;
  float *dout = (float *)(df + 2 * nr);
  CLR_HIP(c, hipMemcpyAsync(df, f12.data(), 2 * nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
  // kappa over a precomputed tidal grid (24 bytes per cell) when it fits in
  // a quarter of the free device memory; the direct 19-point march otherwise
  const size_t t6_bytes = (size_t)c->N * c->N * c->N * 6 * sizeof(float);
  bool use_t6 = false;
  if (which == 0) {
    size_t fr = 0, tot = 0;
    if (c->tidal_n >= t6_bytes) use_t6 = true;
    else if (hipMemGetInfo(&fr, &tot) == hipSuccess && t6_bytes <= fr / 4) use_t6 = true;
    else (void)hipGetLastError();
    if (use_t6) {
      rc = clr_ensure(c, &c->tidal, &c->tidal_n, t6_bytes);
      if (rc) return rc;
    }
  }
  CLR_HIP(c, hipEventRecord(c->ev0, c->stream));
  if (use_t6) {
    hipLaunchKernelGGL(k_tidal, dim3((unsigned)((size_t)c->N * c->N)), dim3(256), 0, c->stream, c->npot, c->N,
                       c->ngx, (float)c->dx, (float)c->obs[0], (float)c->obs[1], (float)c->obs[2],
                       (float)((c->r_max + 2 * c->dx) * (c->r_max + 2 * c->dx)), (float2 *)c->tidal);
    hipLaunchKernelGGL(k_los_t6, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, c->stream, a,
                       (const float2 *)c->tidal, df, df + nr, (int64_t)nside, npix, dout);
  } else {
    hipLaunchKernelGGL(k_los, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, c->stream, a, df, df + nr,
                       (int64_t)nside, npix, dout);
  }
  CLR_HIP(c, hipGetLastError());
  CLR_HIP(c, hipEventRecord(c->ev1, c->stream));
  CLR_HIP(c, hipMemcpyAsync(out, dout, (size_t)npix * nplanes * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->stage_ms[which == 0 ? ST_KAPPA : ST_ISW] = ms;
  return CLR_OK;
}

// Graph-captured kappa (clr_ensemble.hip): the plan and its buffers first,
// then the march as one kernel launch into the context's map buffer.
int clr_prepare_kappa(clr_ctx *c, int nside, int nplanes, const double *r_src) {
  LosArgs a{};
  std::vector<double> f12;
  int rc = los_setup(c, 0, nside, nplanes, r_src, a, f12);
  if (rc) return rc;
  const int64_t npix = 12LL * nside * nside;
  const size_t need = 2 * (size_t)a.nr * sizeof(double) + (size_t)npix * nplanes * sizeof(float);
  if (c->los_buf_n < need) {
    if (c->los_buf) CLR_HIP(c, hipFree(c->los_buf));
    c->los_buf = nullptr;
    c->los_buf_n = 0;
    CLR_HIP(c, hipMalloc(&c->los_buf, need));
    c->los_buf_n = need;
  }
  CLR_HIP(c, hipMemcpy(c->los_buf, f12.data(), f12.size() * sizeof(double), hipMemcpyHostToDevice));
  c->los_plan.assign((const char *)&a, (const char *)&a + sizeof(LosArgs));
  c->los_nside = nside;
  c->los_nplanes = nplanes;
  return CLR_OK;
}

int clr_enqueue_kappa(clr_ctx *c, hipStream_t s) {
  if (c->los_plan.size() != sizeof(LosArgs) || !c->los_buf) return clr_fail(c, CLR_ERR_STATE, "kappa not prepared");
  LosArgs a;
  memcpy(&a, c->los_plan.data(), sizeof(LosArgs));
  a.phi = c->npot;
  const int64_t npix = 12LL * c->los_nside * c->los_nside;
  double *df = (double *)c->los_buf;
  hipLaunchKernelGGL(k_los, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s, a, df, df + a.nr,
                     (int64_t)c->los_nside, npix, (float *)(df + 2 * a.nr));
  CLR_HIP(c, hipGetLastError());
  return CLR_OK;
}

// The last graph-captured kappa map (NEST), out[nplanes][npix].
int clr_fetch_kappa(clr_ctx *c, float *out) {
  if (c->los_plan.size() != sizeof(LosArgs) || !c->los_buf) return clr_fail(c, CLR_ERR_STATE, "kappa not prepared");
  LosArgs a;
  memcpy(&a, c->los_plan.data(), sizeof(LosArgs));
  const int64_t npix = 12LL * c->los_nside * c->los_nside;
  CLR_HIP(c, hipMemcpy(out, (double *)c->los_buf + 2 * a.nr, (size_t)npix * c->los_nplanes * sizeof(float),
                       hipMemcpyDeviceToHost));
  return CLR_OK;
}

// ---------------------------------------------------------------------------
// Sources' line-of-sight properties: srcs_beams_preproc_single
// (srcs.c:425-445), srcs_get_beam_properties_single (srcs.c:452-625, the
// build without _USE_FAST_LENSING) and srcs_beams_postproc_single
// (srcs.c:634-737) for one population.  One thread per source in catalogue
// order: sources of one wavefront come from neighbouring cells, so their
// rays run side by side and share the cache lines of the stencil reads.
// ---------------------------------------------------------------------------
namespace {

struct BeamArgs {
  const float *phi;
  int N, ngx, nr, na;
  double idx, dr, idr, ox, oy, oz, factor_vel;
  double glob_idr, r_last, z_last, v1_0, v1_last;
  const double *r_arr, *z_arr, *v1_arr;   // background r-tables
  const double *a2r;                      // r(a) at a = i/(na-1) (r_of_z, cosmo.c:101-112)
  const double *fac;                      // fac_r_0, fac_r_1, fac_r_2 (nr each)
  const float4 *vrec;                     // per cell (k_cell_records): phi's central differences, delta
  const float2 *t6;                       // per cell (k_tidal): the tidal tensor's second differences
};

// The reference wraps once (beaming.c:129-133), which covers every cell the
// march and the source reads reach when r_max is below the box size; the
// full modulo keeps other inputs inside the grid.
__device__ inline int wrap1(int v, int N) {
  if (v >= N || v < 0) {
    v %= N;
    if (v < 0) v += N;
  }
  return v;
}

// get_element (beaming.c:31-117), RETURN_VEL: unscaled central differences,
// periodic in x, y and z on one GPU (fourier.c:412-413).
__device__ inline void vel_at(const float *__restrict__ phi, int N, size_t ngx, size_t plane, int ix, int iy,
                              int iz, double v[3]) {
  const int xh = ix == N - 1 ? 0 : ix + 1, xl = ix == 0 ? N - 1 : ix - 1;
  const int yh = iy == N - 1 ? 0 : iy + 1, yl = iy == 0 ? N - 1 : iy - 1;
  const int zh = iz == N - 1 ? 0 : iz + 1, zl = iz == 0 ? N - 1 : iz - 1;
  const size_t Z0 = (size_t)iz * plane, Y0 = (size_t)iy * ngx;
  v[0] = phi[xh + Y0 + Z0] - phi[xl + Y0 + Z0];   // FP32 differences, as get_element
  v[1] = phi[ix + (size_t)yh * ngx + Z0] - phi[ix + (size_t)yl * ngx + Z0];
  v[2] = phi[ix + Y0 + (size_t)zh * plane] - phi[ix + Y0 + (size_t)zl * plane];
}

// get_element, RETURN_VEL | RETURN_TID: velocity and the 19-point tidal
// tensor (xx, xy, xz, yy, yz, zz; mixed terms with 1/4).
__device__ inline void vel_tid_at(const float *__restrict__ phi, int N, size_t ngx, size_t plane, int ix, int iy,
                                  int iz, double v[3], double t[6]) {
#pragma clang fp contract(off)
  const int xh = ix == N - 1 ? 0 : ix + 1, xl = ix == 0 ? N - 1 : ix - 1;
  const int yh = iy == N - 1 ? 0 : iy + 1, yl = iy == 0 ? N - 1 : iy - 1;
  const int zh = iz == N - 1 ? 0 : iz + 1, zl = iz == 0 ? N - 1 : iz - 1;
  const size_t Z0 = (size_t)iz * plane, ZH = (size_t)zh * plane, ZL = (size_t)zl * plane;
  const size_t Y0 = (size_t)iy * ngx, YH = (size_t)yh * ngx, YL = (size_t)yl * ngx;
  // FP32 differences, as get_element evaluates them on flouble values
  const float c0 = phi[ix + Y0 + Z0];
  const float pxh = phi[xh + Y0 + Z0], pxl = phi[xl + Y0 + Z0];
  const float pyh = phi[ix + YH + Z0], pyl = phi[ix + YL + Z0];
  const float pzh = phi[ix + Y0 + ZH], pzl = phi[ix + Y0 + ZL];
  v[0] = pxh - pxl;
  v[1] = pyh - pyl;
  v[2] = pzh - pzl;
  t[0] = pxh + pxl - 2 * c0;
  t[3] = pyh + pyl - 2 * c0;
  t[5] = pzh + pzl - 2 * c0;
  t[1] = 0.25f * (phi[xh + YH + Z0] + phi[xl + YL + Z0] - phi[xh + YL + Z0] - phi[xl + YH + Z0]);
  t[2] = 0.25f * (phi[xh + Y0 + ZH] + phi[xl + Y0 + ZL] - phi[xh + Y0 + ZL] - phi[xl + Y0 + ZH]);
  t[4] = 0.25f * (phi[ix + YH + ZH] + phi[ix + YL + ZL] - phi[ix + YH + ZL] - phi[ix + YL + ZH]);
}

// Per cell: phi's unscaled central differences in FP32 (as vel_at takes
// them) and the density, one 16-byte record, so that the CIC and NGP reads
// of the source passes take one load per cell instead of six (plus one).
__global__ __launch_bounds__(256) void k_cell_records(const float *__restrict__ phi, const float *__restrict__ dens,
                                                      int N, int ngx, float4 *__restrict__ out) {
#pragma clang fp contract(off)
  const size_t row = blockIdx.x;
  const int iz = (int)(row / N), iy = (int)(row % N);
  const size_t plane = (size_t)N * ngx;
  for (int ix = threadIdx.x; ix < N; ix += blockDim.x) {
    double v[3];
    vel_at(phi, N, (size_t)ngx, plane, ix, iy, iz, v);
    out[row * N + ix] = make_float4((float)v[0], (float)v[1], (float)v[2], dens[row * ngx + ix]);
  }
}

// A cell's velocity differences (and density) from its record or from phi.
template <bool REC>
__device__ inline void vel_cell(const float4 *__restrict__ vrec, const float *__restrict__ phi, int N, size_t ngx,
                                size_t plane, int ix, int iy, int iz, double v[3]) {
  if (REC) {
    const float4 q = vrec[((size_t)iz * N + iy) * N + ix];
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
  } else {
    vel_at(phi, N, ngx, plane, ix, iy, iz, v);
  }
}

template <bool LENS, bool REC>
__global__ __launch_bounds__(256) void k_src_beams(BeamArgs a, const float4 *__restrict__ pos, long long n,
                                                   float *__restrict__ dz, float *__restrict__ lens) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = pos[i];
  const int N = a.N;
  const size_t ngx = (size_t)a.ngx, plane = (size_t)N * ngx;
  const float *__restrict__ phi = a.phi;
  const double px = p.x, py = p.y, pz = p.z;
  const double r = sqrt(px * px + py * py + pz * pz);
  const double ir = 1. / (r > 0.001 ? r : 0.001);
  const double u[3] = {px * ir, py * ir, pz * ir};
  const double xn[3] = {(px + a.ox) * a.idx, (py + a.oy) * a.idx, (pz + a.oz) * a.idx};

  // CIC velocity at the source (interpolate_from_grid with INTERP_TYPE_SKW,
  // beaming.c:193-266)
  int i0[3], i1[3];
  double h0[3], h1[3];
  for (int ax = 0; ax < 3; ax++) {
    const int c = (int)xn[ax];   // (long)x: truncation toward zero
    h0[ax] = xn[ax] - c;
    h1[ax] = 1 - h0[ax];
    i0[ax] = wrap1(c, N);
    i1[ax] = wrap1(c + 1, N);
  }
  double v[3] = {0, 0, 0};
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int bx = c & 1, by = (c >> 1) & 1, bz = c >> 2;
    const double w = (bz ? h0[2] : h1[2]) * (by ? h0[1] : h1[1]) * (bx ? h0[0] : h1[0]);
    double vc[3];
    vel_cell<REC>(a.vrec, phi, N, ngx, plane, bx ? i1[0] : i0[0], by ? i1[1] : i0[1], bz ? i1[2] : i0[2], vc);
    v[0] += vc[0] * w;
    v[1] += vc[1] * w;
    v[2] += vc[2] * w;
  }
  const float vr = (float)(0.5 * a.idx * (v[0] * u[0] + v[1] * u[1] + v[2] * u[2]));
  // postproc: dz_rsd *= V1(r_of_z(z0)) factor_vel, z0 the catalogue's float z(r)
  const float z0 = (float)f_of_r(a.r_arr, a.z_arr, a.glob_idr, a.r_last, r, 0.0, a.z_last);
  double rz;
  {
    const double aa = 1. / (1 + (double)z0);
    if (aa >= 1) rz = 0;
    else if (aa <= 0) rz = a.a2r[0];
    else {
      const int ia = (int)(aa * (a.na - 1));
      rz = a.a2r[ia] + (a.a2r[ia + 1] - a.a2r[ia]) * (aa - (double)ia / (a.na - 1)) * (a.na - 1.);
    }
  }
  const double vg = f_of_r(a.r_arr, a.v1_arr, a.glob_idr, a.r_last, rz, a.v1_0, a.v1_last);
  dz[i] = (float)((double)vr * (vg * a.factor_vel));
  if (!LENS) return;

  double cth = u[2];
  if (cth >= 1) cth = 1;
  if (cth <= -1) cth = -1;
  const double sth = sqrt((1 - cth) * (1 + cth));
  double cph = 1, sph = 0;
  if (sth != 0) { cph = u[0] / sth; sph = u[1] / sth; }
  const double pre = a.idx * a.idx * ir, prem = 0.5 * a.idx * ir;
  const double ux[3] = {cth * cph * prem, cth * sph * prem, -sth * prem};
  const double uy[2] = {-sph * prem, cph * prem};
  const double rk[6] = {(cth * cth * cph * cph + sph * sph) * pre, (2 * cph * sph * (cth * cth - 1)) * pre,
                        (-2 * cth * sth * cph) * pre,                (cth * cth * sph * sph + cph * cph) * pre,
                        (-2 * cth * sth * sph) * pre,                (sth * sth) * pre};
  const double re1[6] = {(cth * cth * cph * cph - sph * sph) * pre, (2 * cph * sph * (cth * cth + 1)) * pre,
                         (-2 * cth * sth * cph) * pre,                (cth * cth * sph * sph - cph * cph) * pre,
                         (-2 * cth * sth * sph) * pre,                (sth * sth) * pre};
  const double re2[5] = {(-2 * cth * cph * sph) * pre, (2 * cth * (cph * cph - sph * sph)) * pre,
                         (2 * sth * sph) * pre,        (2 * cth * sph * cph) * pre,
                         (-2 * sth * cph) * pre};
  const double *__restrict__ f0 = a.fac, *__restrict__ f1 = a.fac + a.nr, *__restrict__ f2 = a.fac + 2 * a.nr;
  double dtx = 0, dty = 0, kp = 0, e1 = 0, e2 = 0;
  const int irmax = min((int)(r * a.idr + 0.5), a.nr - 1);
  for (int i_r = 0; i_r <= irmax; i_r++) {
    const double rm = (i_r + 0.5) * a.dr;
    int ix[3];
    ix[0] = wrap1((int)((rm * u[0] + a.ox) * a.idx + 0.5), N);
    ix[1] = wrap1((int)((rm * u[1] + a.oy) * a.idx + 0.5), N);
    ix[2] = wrap1((int)((rm * u[2] + a.oz) * a.idx + 0.5), N);
    double vp[3], tp[6];
    if (REC) {
      const size_t o = ((size_t)ix[2] * N + ix[1]) * N + ix[0];
      const float4 q = a.vrec[o];
      const float2 t0 = a.t6[3 * o], t1 = a.t6[3 * o + 1], t2 = a.t6[3 * o + 2];
      vp[0] = q.x;
      vp[1] = q.y;
      vp[2] = q.z;
      tp[0] = t0.x;
      tp[1] = t0.y;
      tp[2] = t1.x;
      tp[3] = t1.y;
      tp[4] = t2.x;
      tp[5] = t2.y;
    } else {
      vel_tid_at(phi, N, ngx, plane, ix[0], ix[1], ix[2], vp, tp);
    }
    const double fr = f1[i_r] * r - f2[i_r];
    const double frm = 2 * (f0[i_r] * r - f1[i_r]);
    double de1 = 0, de2 = 0, dk = 0;
#pragma unroll
    for (int q = 0; q < 6; q++) {
      de1 += re1[q] * tp[q];
      de2 += (q < 5 ? re2[q] : 0.0) * tp[q];
      dk += rk[q] * tp[q];
    }
    const double dvx = ux[0] * vp[0] + ux[1] * vp[1] + ux[2] * vp[2];
    const double dvy = uy[0] * vp[0] + uy[1] * vp[1] + 0.0 * vp[2];
    e1 += de1 * fr;
    e2 += de2 * fr;
    kp += dk * fr;
    dtx += dvx * frm;
    dty += dvy * frm;
  }
  const double RTOD = 57.2957795;
  float *o = lens + 5 * i;
  o[0] = (float)e1;
  o[1] = (float)e2;
  o[2] = (float)kp;
  o[3] = (float)(dty * RTOD);
  o[4] = (float)(dtx * RTOD);
}

}  // namespace

// Builds the per-cell records on the context's stream: velocity differences
// and density (16 bytes per cell), with `tidal` also the tidal tensor (24
// bytes per cell, k_tidal), when they fit in a quarter of the free device
// memory (buffers kept for later calls); false: the callers read phi directly.
bool clr_cell_records(clr_ctx *c, bool tidal) {
  const size_t cells = (size_t)c->N * c->N * c->N;
  const size_t vb = cells * sizeof(float4), tb = tidal ? cells * 6 * sizeof(float) : 0;
  size_t need = (c->velrec_n >= vb ? 0 : vb) + (c->tidal_n >= tb ? 0 : tb);
  if (need > 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    if (need > fr / 4) return false;
  }
  if (clr_ensure(c, &c->velrec, &c->velrec_n, vb)) return false;
  if (tidal && clr_ensure(c, &c->tidal, &c->tidal_n, tb)) return false;
  hipLaunchKernelGGL(k_cell_records, dim3((unsigned)((size_t)c->N * c->N)), dim3(256), 0, c->stream, c->npot, c->dens,
                     c->N, c->ngx, (float4 *)c->velrec);
  if (tidal)
    hipLaunchKernelGGL(k_tidal, dim3((unsigned)((size_t)c->N * c->N)), dim3(256), 0, c->stream, c->npot, c->N,
                       c->ngx, (float)c->dx, (float)c->obs[0], (float)c->obs[1], (float)c->obs[2],
                       (float)((c->r_max + 2 * c->dx) * (c->r_max + 2 * c->dx)), (float2 *)c->tidal);
  return hipGetLastError() == hipSuccess;
}

int clr_launch_src_beams(clr_ctx *c, int ipop, const double *r_arr_a2r, float *dz, float *lens, long long n) {
  if (ipop < 0 || ipop >= c->n_tr[CLR_TRACER_SRCS]) return clr_fail(c, CLR_ERR_ARG, "no source population %d", ipop);
  if (!c->have_fields) return clr_fail(c, CLR_ERR_STATE, "clr_gaussian_fields must run first");
  const CatalogDev &C = c->cat[ipop];
  if (n != C.n) return clr_fail(c, CLR_ERR_ARG, "population %d holds %lld sources, not %lld", ipop, C.n, n);
  if (!r_arr_a2r || (n > 0 && !dz)) return clr_fail(c, CLR_ERR_ARG, "null r(a) table or output");
  const int N = c->N, na = c->na;
  const double r_max = c->r_max;
  if (!(r_max > 0)) return clr_fail(c, CLR_ERR_STATE, "tables carry no r_max");
  if (!(c->hubble_0 > 0) || !(c->omega_m > 0)) return clr_fail(c, CLR_ERR_STATE, "tables carry no H0 or Omega_M");
  if (n == 0) return CLR_OK;
  const int nr = N / 2;
  const double dr = r_max / nr;
  // fac_r_0/1/2 (srcs.c:470-480), FP64 on the host from the uploaded tables
  std::vector<double> rr(na), zz(na), dd(na);
  CLR_HIP(c, hipMemcpy(rr.data(), c->bg, na * sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipMemcpy(zz.data(), c->bg + na, na * sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipMemcpy(dd.data(), c->bg + 2 * (size_t)na, na * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<double> fac(3 * (size_t)nr);
  for (int i = 0; i < nr; i++) {
    const double rm = (i + 0.5) * dr;
    const double pg = f_of_r(rr.data(), dd.data(), c->glob_idr, c->r_last, rm, 1.0, dd[na - 1]) *
                      (1 + f_of_r(rr.data(), zz.data(), c->glob_idr, c->r_last, rm, 0.0, zz[na - 1]));
    fac[i] = pg * dr;
    fac[nr + i] = rm * pg * dr;
    fac[2 * nr + i] = rm * rm * pg * dr;
  }
  const size_t tab_bytes = (3 * (size_t)nr + na) * sizeof(double);
  int rc = clr_ensure(c, &c->misc, &c->misc_n, tab_bytes);
  if (rc) return rc;
  double *dfac = (double *)c->misc, *da2r = dfac + 3 * nr;
  CLR_HIP(c, hipMemcpyAsync(dfac, fac.data(), 3 * (size_t)nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
  CLR_HIP(c, hipMemcpyAsync(da2r, r_arr_a2r, na * sizeof(double), hipMemcpyHostToDevice, c->stream));
  BeamArgs a{};
  {
    const int prc = clr_phi_ready(c, 0);
    if (prc) return prc;
  }
  a.phi = c->npot;
  a.N = N;
  a.ngx = c->ngx;
  a.nr = nr;
  a.na = na;
  a.idx = N / c->L;
  a.dr = dr;
  a.idr = 1. / dr;
  a.ox = c->obs[0];
  a.oy = c->obs[1];
  a.oz = c->obs[2];
  a.factor_vel = -c->fgrowth_0 / (1.5 * c->hubble_0 * c->omega_m);
  a.glob_idr = c->glob_idr;
  a.r_last = c->r_last;
  a.z_last = c->h_bg_last[1];
  a.v1_0 = c->h_bg_first[3];
  a.v1_last = c->h_bg_last[3];
  a.r_arr = c->bg;
  a.z_arr = c->bg + na;
  a.v1_arr = c->bg + 3 * (size_t)na;
  a.a2r = da2r;
  a.fac = dfac;
  const int per = lens ? 6 : 1;
  float *tmp = nullptr;
  CLR_HIP(c, hipMalloc((void **)&tmp, (size_t)n * per * sizeof(float)));
  CLR_HIP(c, hipEventRecord(c->ev0, c->stream));
  const dim3 grid((unsigned)((n + 255) / 256));
  // per-cell records (velocity differences; with lensing also the tidal
  // tensor) when they fit in a quarter of the free device memory
  const bool rec = clr_cell_records(c, lens != nullptr);
  a.vrec = (const float4 *)c->velrec;
  a.t6 = (const float2 *)c->tidal;
  if (lens && rec)
    hipLaunchKernelGGL((k_src_beams<true, true>), grid, dim3(256), 0, c->stream, a, (const float4 *)C.pos, n, tmp,
                       tmp + n);
  else if (lens)
    hipLaunchKernelGGL((k_src_beams<true, false>), grid, dim3(256), 0, c->stream, a, (const float4 *)C.pos, n, tmp,
                       tmp + n);
  else if (rec)
    hipLaunchKernelGGL((k_src_beams<false, true>), grid, dim3(256), 0, c->stream, a, (const float4 *)C.pos, n, tmp,
                       (float *)nullptr);
  else
    hipLaunchKernelGGL((k_src_beams<false, false>), grid, dim3(256), 0, c->stream, a, (const float4 *)C.pos, n, tmp,
                       (float *)nullptr);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dz, tmp, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && lens)
    e = hipMemcpyAsync(lens, tmp + n, (size_t)n * 5 * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return clr_fail(c, CLR_ERR_HIP, "source beams: %s", hipGetErrorString(e));
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->stage_ms[ST_BEAMS] = ms;
  return CLR_OK;
}

// ---------------------------------------------------------------------------
// Skewers (srcs.c:507-529) with their postproc scaling (srcs.c:725-733): one
// thread per (source, radial step), steps of a source in consecutive lanes,
// so a wavefront covers one or a few rays and writes whole rows.
// ---------------------------------------------------------------------------
namespace {

struct SkwArgs {
  const float *phi, *dens;
  const float4 *vrec;   // per cell (k_cell_records), or null
  int N, ngx, nr, na, gaussian;
  double idx, dr, idr, ox, oy, oz, factor_vel, sigma2;
  double glob_idr, r_last, v1_0, v1_last, d1_last;
  float dxf;
  const double *r_arr, *d1_arr, *v1_arr;
};

// A block computes 16 steps of 16 consecutive sources (lane = step + 16 x
// source): neighbouring sources lie in neighbouring cells, so their lines of
// sight cross nearly the same cells at the same step and share the CIC
// corners' cache lines, where one thread per (source, step) in row order
// fetched ~170 bytes of HBM per value.  The blocks of one group of sources
// follow each other on one XCD (workgroups go to the 8 XCDs round-robin by
// linear id, each XCD with its own L2): logical block
// (b % 8) (gridDim.x / 8) + b / 8, gridDim.x a multiple of 8.  nr is a
// multiple of 16.
template <bool REC>
__global__ __launch_bounds__(256) void k_src_skewers(SkwArgs a, const float4 *__restrict__ pos, long long n,
                                                     float *__restrict__ dg, float *__restrict__ vs) {
#pragma clang fp contract(off)
  const long long blk = (long long)(blockIdx.x % 8u) * (gridDim.x / 8u) + blockIdx.x / 8u;
  const int ntile = a.nr / 16;
  const long long i = (blk / ntile) * 16 + threadIdx.x / 16;
  const int i_r = (int)(blk % ntile) * 16 + (int)(threadIdx.x % 16);
  if (i >= n) return;
  const long long t = i * a.nr + i_r;
  const float4 p = pos[i];
  const int N = a.N;
  const size_t ngx = (size_t)a.ngx, plane = (size_t)N * ngx;
  const double px = p.x, py = p.y, pz = p.z;
  const double r = sqrt(px * px + py * py + pz * pz);
  const double ir = 1. / (r > 0.001 ? r : 0.001);
  const double u[3] = {px * ir, py * ir, pz * ir};
  const int irmax = min((int)(r * a.idr + 0.5), a.nr - 1);
  if (i_r > irmax) {   // calloc'd rows (common.c:381-387)
    dg[t] = 0;
    vs[t] = 0;
    return;
  }
  const double rm = (i_r + 0.5) * a.dr;
  const double xn[3] = {(rm * u[0] + a.ox) * a.idx, (rm * u[1] + a.oy) * a.idx, (rm * u[2] + a.oz) * a.idx};
  int i0[3], i1[3];
  double h0[3], h1[3];
  for (int ax = 0; ax < 3; ax++) {
    const int c = (int)xn[ax];
    h0[ax] = xn[ax] - c;
    h1[ax] = 1 - h0[ax];
    i0[ax] = wrap1(c, N);
    i1[ax] = wrap1(c + 1, N);
  }
  double v[3] = {0, 0, 0}, d = 0;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int bx = c & 1, by = (c >> 1) & 1, bz = c >> 2;
    const double w = (bz ? h0[2] : h1[2]) * (by ? h0[1] : h1[1]) * (bx ? h0[0] : h1[0]);
    const int cx = bx ? i1[0] : i0[0], cy = by ? i1[1] : i0[1], cz = bz ? i1[2] : i0[2];
    double vc[3];
    float dc;
    if (REC) {
      const float4 q = a.vrec[((size_t)cz * N + cy) * N + cx];
      vc[0] = q.x;
      vc[1] = q.y;
      vc[2] = q.z;
      dc = q.w;
    } else {
      vel_at(a.phi, N, ngx, plane, cx, cy, cz, vc);
      dc = a.dens[(size_t)cz * plane + (size_t)cy * ngx + cx];
    }
    v[0] += vc[0] * w;
    v[1] += vc[1] * w;
    v[2] += vc[2] * w;
    if (a.gaussian) {
      // get_element RETURN_GAUSS (beaming.c:46-57): the cell's offsets and
      // radius in float, D(r) from the FP64 table
      const float x0 = (float)((cx + 0.0) * a.dxf - a.ox), y0 = (float)((cy + 0.0) * a.dxf - a.oy);
      const float z0 = (float)((cz + 0.0) * a.dxf - a.oz);
      const double rc = sqrt((double)(x0 * x0 + y0 * y0 + z0 * z0));
      const double D = f_of_r(a.r_arr, a.d1_arr, a.glob_idr, a.r_last, rc, 1.0, a.d1_last);
      d += (log((double)(1.0f + dc)) / D + D * a.sigma2 * 0.5) * w;
    } else {
      d += (double)dc * w;
    }
  }
  const float vr = (float)(0.5 * a.idx * (v[0] * u[0] + v[1] * u[1] + v[2] * u[2]));
  const double vg = f_of_r(a.r_arr, a.v1_arr, a.glob_idr, a.r_last, rm, a.v1_0, a.v1_last);
  dg[t] = (float)d;
  vs[t] = (float)((double)vr * (vg * a.factor_vel));
}

}  // namespace

int clr_launch_src_skewers(clr_ctx *c, int ipop, int gaussian, double sigma2, float *dg, float *vs, long long n) {
  if (ipop < 0 || ipop >= c->n_tr[CLR_TRACER_SRCS]) return clr_fail(c, CLR_ERR_ARG, "no source population %d", ipop);
  if (!c->have_fields) return clr_fail(c, CLR_ERR_STATE, "clr_gaussian_fields must run first");
  if (!c->density_done) return clr_fail(c, CLR_ERR_STATE, "density field not computed");
  const CatalogDev &C = c->cat[ipop];
  if (n != C.n) return clr_fail(c, CLR_ERR_ARG, "population %d holds %lld sources, not %lld", ipop, C.n, n);
  if (n > 0 && (!dg || !vs)) return clr_fail(c, CLR_ERR_ARG, "null skewer output");
  if (gaussian && !(sigma2 >= 0)) return clr_fail(c, CLR_ERR_ARG, "bad sigma2 %g", sigma2);
  const int N = c->N, na = c->na;
  if (!(c->r_max > 0)) return clr_fail(c, CLR_ERR_STATE, "tables carry no r_max");
  if (!(c->hubble_0 > 0) || !(c->omega_m > 0)) return clr_fail(c, CLR_ERR_STATE, "tables carry no H0 or Omega_M");
  if (n == 0) return CLR_OK;
  {
    const int prc = clr_phi_ready(c, 0);
    if (prc) return prc;
  }
  SkwArgs a{};
  a.phi = c->npot;
  a.dens = c->dens;
  a.N = N;
  a.ngx = c->ngx;
  a.nr = N / 2;
  if (a.nr % 16) return clr_fail(c, CLR_ERR_ARG, "source skewers need N/2 a multiple of 16 (N = %d)", N);
  a.na = na;
  a.gaussian = gaussian ? 1 : 0;
  a.idx = N / c->L;
  a.dr = c->r_max / a.nr;
  a.idr = 1. / a.dr;
  a.ox = c->obs[0];
  a.oy = c->obs[1];
  a.oz = c->obs[2];
  a.factor_vel = -c->fgrowth_0 / (1.5 * c->hubble_0 * c->omega_m);
  a.sigma2 = sigma2;
  a.glob_idr = c->glob_idr;
  a.r_last = c->r_last;
  a.v1_0 = c->h_bg_first[3];
  a.v1_last = c->h_bg_last[3];
  a.d1_last = c->h_bg_last[2];
  a.dxf = (float)(c->L / N);
  a.r_arr = c->bg;
  a.d1_arr = c->bg + 2 * (size_t)na;
  a.v1_arr = c->bg + 3 * (size_t)na;
  a.vrec = clr_cell_records(c, false) ? (const float4 *)c->velrec : nullptr;
  // sources per launch: at most 2^26 skewer values (512 MiB of staging)
  const long long chunk = std::max(1LL, (1LL << 26) / a.nr);
  const long long cap = std::min(n, chunk);
  float *tmp = nullptr;
  CLR_HIP(c, hipMalloc((void **)&tmp, (size_t)cap * a.nr * 2 * sizeof(float)));
  hipError_t e = hipEventRecord(c->ev0, c->stream);
  float kms = 0;
  for (long long s0 = 0; s0 < n && e == hipSuccess; s0 += cap) {
    const long long m = std::min(cap, n - s0);
    const long long tot = m * a.nr;
    // 16 sources x 16 steps per block, the count rounded up to the 8 XCDs
    // (the extra blocks' sources are >= m and return)
    const unsigned nblk = (unsigned)(((m + 15) / 16 * (a.nr / 16) + 7) / 8 * 8);
    e = hipEventRecord(c->ev2, c->stream);
    if (e == hipSuccess) {
      if (a.vrec)
        hipLaunchKernelGGL(k_src_skewers<true>, dim3(nblk), dim3(256), 0, c->stream, a,
                           (const float4 *)C.pos + s0, m, tmp, tmp + (size_t)cap * a.nr);
      else
        hipLaunchKernelGGL(k_src_skewers<false>, dim3(nblk), dim3(256), 0, c->stream, a,
                           (const float4 *)C.pos + s0, m, tmp, tmp + (size_t)cap * a.nr);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(c->ev3, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dg + (size_t)s0 * a.nr, tmp, (size_t)tot * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(vs + (size_t)s0 * a.nr, tmp + (size_t)cap * a.nr, (size_t)tot * sizeof(float),
                         hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    float cms = 0;
    if (e == hipSuccess && hipEventElapsedTime(&cms, c->ev2, c->ev3) == hipSuccess) kms += cms;
  }
  c->stage_ms[ST_SKW] = kms;
  if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return clr_fail(c, CLR_ERR_HIP, "source skewers: %s", hipGetErrorString(e));
  return CLR_OK;
}

// ---------------------------------------------------------------------------
// Custom projected maps: cstm_beams_preproc_single +
// cstm_get_beam_properties_single (cstm.c:39-150).  One thread per pixel
// centre (NEST order), the same radial march as kappa with a trilinear
// density read per step.
// ---------------------------------------------------------------------------
namespace {

struct CstmArgs {
  const float *dens;
  int N, ngx, ir_min, ir_max, model;
  double idx, dr, ox, oy, oz;
};

__global__ __launch_bounds__(256) void k_cstm(CstmArgs a, const double *__restrict__ kz,
                                              const double *__restrict__ bz, const double *__restrict__ nz,
                                              int64_t nside, int64_t npix, float *__restrict__ out) {
#pragma clang fp contract(off)
  const int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= npix) return;
  double u[3];
  pix2vec_nest(nside, ip, u);
  const int N = a.N;
  const size_t ngx = (size_t)a.ngx, plane = (size_t)N * ngx;
  double cval = 0;
  for (int irr = a.ir_min; irr <= a.ir_max; irr++) {
    const double rm = (irr + 0.5) * a.dr;
    const double xn[3] = {(rm * u[0] + a.ox) * a.idx, (rm * u[1] + a.oy) * a.idx, (rm * u[2] + a.oz) * a.idx};
    int i0[3], i1[3];
    double h0[3], h1[3];
    for (int ax = 0; ax < 3; ax++) {
      const int c = (int)xn[ax];
      h0[ax] = xn[ax] - c;
      h1[ax] = 1 - h0[ax];
      i0[ax] = wrap1(c, N);
      i1[ax] = wrap1(c + 1, N);
    }
    double d = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int bx = c & 1, by = (c >> 1) & 1, bzc = c >> 2;
      const double w = (bzc ? h0[2] : h1[2]) * (by ? h0[1] : h1[1]) * (bx ? h0[0] : h1[0]);
      d += (double)a.dens[(size_t)(bzc ? i1[2] : i0[2]) * plane + (size_t)(by ? i1[1] : i0[1]) * ngx +
                          (bx ? i1[0] : i0[0])] * w;
    }
    // the reference holds d as flouble before bias_model
    cval += kz[irr] * (bias_model_f64((double)(float)d, bz[irr], a.model) * nz[irr] - 1);
  }
  out[ip] = (float)(cval * a.dr);
}

}  // namespace

int clr_launch_cstm(clr_ctx *c, int ipop, int nside, float *out) {
  if (ipop < 0 || ipop >= c->n_tr[CLR_TRACER_CSTM]) return clr_fail(c, CLR_ERR_ARG, "no custom map %d", ipop);
  if (!c->density_done) return clr_fail(c, CLR_ERR_STATE, "density field not computed");
  if (nside < 1 || nside > 8192 || (nside & (nside - 1))) return clr_fail(c, CLR_ERR_ARG, "bad nside %d", nside);
  if (!out) return clr_fail(c, CLR_ERR_ARG, "null output");
  const TracerDev &T = c->tr[CLR_TRACER_CSTM][ipop];
  if (!T.set || !T.norm_set) return clr_fail(c, CLR_ERR_STATE, "custom map %d has no norm table", ipop);
  const int N = c->N, na = c->na;
  if (!(c->r_max > 0)) return clr_fail(c, CLR_ERR_STATE, "tables carry no r_max");
  const int nr = N / 2;
  const double dr = c->r_max / nr;
  // kernel K/IH, bias and norm at each step (cstm.c:87-99), FP64 on the host
  std::vector<double> rr(na), ih(na), kt(na), bt(na), nt(na);
  CLR_HIP(c, hipMemcpy(rr.data(), c->bg, na * sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipMemcpy(ih.data(), c->bg + 5 * (size_t)na, na * sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipMemcpy(kt.data(), T.amp, na * sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipMemcpy(bt.data(), T.bz, na * sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipMemcpy(nt.data(), T.norm, na * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<double> tab(3 * (size_t)nr);
  double kmax = -1e100;
  for (int ir = 0; ir < nr; ir++) {
    const double rm = (ir + 0.5) * dr;
    tab[ir] = f_of_r(rr.data(), kt.data(), c->glob_idr, c->r_last, rm, 0.0, 0.0) /
              f_of_r(rr.data(), ih.data(), c->glob_idr, c->r_last, rm, ih[0], ih[na - 1]);
    tab[nr + ir] = f_of_r(rr.data(), bt.data(), c->glob_idr, c->r_last, rm, bt[0], 1.0);
    tab[2 * nr + ir] = f_of_r(rr.data(), nt.data(), c->glob_idr, c->r_last, rm, T.norm0, T.normf);
    kmax = std::max(kmax, fabs(tab[ir]));
  }
  int ir_min = 0, ir_max = nr - 1;
  for (int ir = 0; ir < nr; ir++)
    if (fabs(tab[ir]) > 1e-4 * kmax) { ir_min = ir; break; }
  for (int ir = nr - 1; ir >= 0; ir--)
    if (fabs(tab[ir]) > 1e-4 * kmax) { ir_max = ir; break; }
  if (ir_max < ir_min) return clr_fail(c, CLR_ERR_ARG, "Custom kernel has no suppport");
  const int64_t npix = 12LL * nside * nside;
  const size_t need = 3 * (size_t)nr * sizeof(double) + (size_t)npix * sizeof(float);
  int rc = clr_ensure(c, &c->misc, &c->misc_n, need);
  if (rc) return rc;
  double *dt = (double *)c->misc;
  float *dout = (float *)(dt + 3 * nr);
  CLR_HIP(c, hipMemcpyAsync(dt, tab.data(), 3 * (size_t)nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
  CstmArgs a{};
  a.dens = c->dens;
  a.N = N;
  a.ngx = c->ngx;
  a.ir_min = ir_min;
  a.ir_max = ir_max;
  a.model = T.model;
  a.idx = N / c->L;
  a.dr = dr;
  a.ox = c->obs[0];
  a.oy = c->obs[1];
  a.oz = c->obs[2];
  CLR_HIP(c, hipEventRecord(c->ev0, c->stream));
  hipLaunchKernelGGL(k_cstm, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, c->stream, a, dt, dt + nr,
                     dt + 2 * nr, (int64_t)nside, npix, dout);
  CLR_HIP(c, hipGetLastError());
  CLR_HIP(c, hipEventRecord(c->ev1, c->stream));
  CLR_HIP(c, hipMemcpyAsync(out, dout, (size_t)npix * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->stage_ms[ST_CSTM] = ms;
  return CLR_OK;
}
