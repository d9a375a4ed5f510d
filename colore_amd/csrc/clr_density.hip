// clr_density.hip -- physical density (density.c:1034-1126) and the per-shell
// bias normalisation (density.c:1128-1393) on gfx950.
//
// k_transform : lognormalize / densclip in place (FP64 geometry and table
//               lookups, FP32 exp).
// k_shells    : one pass over the grid; per cell the FP64 radius, z(r) and
//               shell index s = (int)(z*idz)+1, then count, z and
//               bias_model(delta, b_t(r)) for every tracer t, accumulated in
//               FP64.  Lanes of a wavefront that share a shell are reduced
//               with shuffles and added by one lane into that wavefront's own
//               LDS slots, so the order of every addition is fixed: the sums
//               are bitwise repeatable.  With `fused`, the lognormal
//               transform is applied in the same pass (one read + one write
//               of the grid instead of two of each).
// k_shell_final / k_norm_table : per-block partials summed in a fixed order,
//               norm = count/sum, edge fill, and the linear interpolation in z
//               onto the NA r-grid (density.c:1272-1355), all on the device.
#include "clr_internal.h"
#include <math.h>

using namespace clr;

namespace {

constexpr int kShellBlocks = 1024;
constexpr int kShellThreads = 256;
constexpr int kWaves = kShellThreads / 64;

struct TracerArgs {
  const double *bz[3 * CLR_NPOP_MAX];
  double bz0[3 * CLR_NPOP_MAX];
  int model[3 * CLR_NPOP_MAX];
  int n;
};

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__global__ __launch_bounds__(256) void k_transform(float *__restrict__ g, int N, int ngx, double dx,
                                                   double ox, double oy, double oz,
                                                   const double *__restrict__ r_arr,
                                                   const double *__restrict__ d1, double idr,
                                                   double r_last, double d1_last, double sigma2,
                                                   int dens_type) {
#pragma clang fp contract(off)
  const double obs[3] = {ox, oy, oz};
  const size_t row = blockIdx.x;
  const int iz = (int)(row / N), iy = (int)(row % N);
  float *p = g + row * (size_t)ngx;
  for (int ix = threadIdx.x; ix < N; ix += blockDim.x) {
    double x0, y0, z0;
    const double r = cell_radius(ix, iy, iz, dx, obs, &x0, &y0, &z0);
    const double dg = f_of_r(r_arr, d1, idr, r_last, r, 1.0, d1_last);
    const double delta = p[ix];
    if (dens_type == CLR_DENS_CLIP) {
      p[ix] = (float)(fmax(1 + dg * delta, 0.0) - 1);
    } else {
      const float a = (float)(dg * (delta - 0.5 * dg * sigma2));
      p[ix] = expf(a) - 1.f;
    }
  }
}

// nq = 2 + ntr values per shell: [0] count, [1] z sum, [2+t] tracer t.
__global__ __launch_bounds__(kShellThreads) void k_shells(
    float *__restrict__ g, int N, int ngx, double dx, double ox, double oy, double oz,
    const double *__restrict__ r_arr, const double *__restrict__ zt, const double *__restrict__ d1,
    double idr, double r_last, double z_last, double d1_last, double sigma2, int fused, int nz,
    double idz, TracerArgs ta, double *__restrict__ partial) {
#pragma clang fp contract(off)
  extern __shared__ double acc[];   // [kWaves][nz][nq]
  const int nq = 2 + ta.n;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  double *myacc = acc + (size_t)wave * nz * nq;
  for (int i = threadIdx.x; i < kWaves * nz * nq; i += blockDim.x) acc[i] = 0;
  __syncthreads();
  const double obs[3] = {ox, oy, oz};
  const size_t nrows = (size_t)N * N;
  for (size_t row = blockIdx.x; row < nrows; row += gridDim.x) {
    const int iz = (int)(row / N), iy = (int)(row % N);
    float *p = g + row * (size_t)ngx;
    for (int x0i = 0; x0i < N; x0i += blockDim.x) {
      const int ix = x0i + threadIdx.x;
      int s = -1;
      double zz = 0, r = 0;
      float dl = 0.f;
      if (ix < N) {
        double x0, y0, z0;
        r = cell_radius(ix, iy, iz, dx, obs, &x0, &y0, &z0);
        zz = f_of_r(r_arr, zt, idr, r_last, r, 0.0, z_last);
        dl = p[ix];
        if (fused) {
          const double dg = f_of_r(r_arr, d1, idr, r_last, r, 1.0, d1_last);
          const float a = (float)(dg * ((double)dl - 0.5 * dg * sigma2));
          dl = expf(a) - 1.f;
          p[ix] = dl;
        }
        const int ind = (int)(zz * idz) + 1;
        if (ind >= 0 && ind < nz) s = ind;
      }
      unsigned long long active = __ballot(s >= 0);
      while (active) {
        const int leader = __ffsll((long long)active) - 1;
        const int sl = __shfl(s, leader);
        const bool mine = (s == sl);
        const unsigned long long mmask = __ballot(mine);
        double v = wave_sum(mine ? 1.0 : 0.0);
        if (lane == 0) myacc[sl * nq + 0] += v;
        v = wave_sum(mine ? zz : 0.0);
        if (lane == 0) myacc[sl * nq + 1] += v;
        for (int t = 0; t < ta.n; t++) {
          double term = 0;
          if (mine) {
            const double b = f_of_r(r_arr, ta.bz[t], idr, r_last, r, ta.bz0[t], 1.0);
            term = (double)bias_model_f32(dl, (float)b, ta.model[t]);
          }
          v = wave_sum(term);
          if (lane == 0) myacc[sl * nq + 2 + t] += v;
        }
        active &= ~mmask;
      }
    }
  }
  __syncthreads();
  double *out = partial + (size_t)blockIdx.x * nz * nq;
  for (int i = threadIdx.x; i < nz * nq; i += blockDim.x) {
    double v = 0;
    for (int w = 0; w < kWaves; w++) v += acc[(size_t)w * nz * nq + i];
    out[i] = v;
  }
}

__global__ void k_shell_final(const double *__restrict__ partial, int nblocks, int nvals,
                              double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvals) return;
  double v = 0;
  for (int b = 0; b < nblocks; b++) v += partial[(size_t)b * nvals + i];
  out[i] = v;
}

// Per tracer t: norm table on the NA r-grid.  sums = [nz][nq] from
// k_shell_final; work = scratch of 2*nz doubles per tracer.
__global__ void k_norm_table(const double *__restrict__ sums, int nz, int nq, int t, double zmax,
                             const double *__restrict__ r_arr, const double *__restrict__ zt,
                             double idr, double r_last, double z_last, int na, double *__restrict__ work,
                             double *__restrict__ norm_out, double *__restrict__ edges) {
#pragma clang fp contract(off)
  double *zarr = work, *nrm = work + nz;
  if (threadIdx.x == 0) {
    for (int iz = 0; iz < nz; iz++) {
      const double n = sums[iz * nq + 0];
      zarr[iz] = sums[iz * nq + 1];
      nrm[iz] = sums[iz * nq + 2 + t];
      if (n > 0) {
        zarr[iz] /= n;
        nrm[iz] = n / nrm[iz];
      }
    }
    zarr[0] = 0;
    zarr[nz - 1] = zmax;
    nrm[0] = nrm[1];
    nrm[nz - 1] = nrm[nz - 2];
    edges[0] = nrm[0];
    edges[1] = nrm[nz - 1];
  }
  __syncthreads();
  const double z0 = zarr[0], zf = zarr[nz - 1];
  const double n0 = nrm[0], nf = nrm[nz - 1];
  for (int ii = threadIdx.x; ii < na; ii += blockDim.x) {
    const double z = f_of_r(r_arr, zt, idr, r_last, r_arr[ii], 0.0, z_last);
    double nm;
    if (z < z0) nm = n0;
    else if (z >= zf) nm = nf;
    else {
      int lo = 0, hi = nz - 1;
      while (hi > lo + 1) {
        const int i = (hi + lo) / 2;
        if (zarr[i] > z) hi = i; else lo = i;
      }
      const double xl = zarr[lo], xh = zarr[lo + 1];
      const double dxz = xh - xl;
      nm = dxz > 0.0 ? nrm[lo] + (z - xl) / dxz * (nrm[lo + 1] - nrm[lo]) : 0.0;
    }
    norm_out[ii] = nm;
  }
}

}  // namespace

int clr_launch_density(clr_ctx *c, int dens_type, double sigma2) {
  if (!c->have_fields) return clr_fail(c, CLR_ERR_STATE, "clr_gaussian_fields must run first");
  if (dens_type != CLR_DENS_LOGNORMAL && dens_type != CLR_DENS_CLIP)
    return clr_fail(c, CLR_ERR_ARG, "Density type %d not supported", dens_type);
  const int N = c->N;
  const double *r = c->bg, *d1 = c->bg + 2 * c->na;
  double d1_last;
  CLR_HIP(c, hipMemcpy(&d1_last, d1 + c->na - 1, sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipEventRecord(c->ev0, c->stream));
  hipLaunchKernelGGL(k_transform, dim3((unsigned)((size_t)N * N)), dim3(N >= 256 ? 256 : N), 0, c->stream,
                     c->dens, N, c->ngx, c->dx, c->obs[0], c->obs[1], c->obs[2], r, d1, c->glob_idr,
                     c->r_last, d1_last, sigma2, dens_type);
  CLR_HIP(c, hipGetLastError());
  CLR_HIP(c, hipEventRecord(c->ev1, c->stream));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->stage_ms[ST_DENS] = ms;
  c->density_done = 1;
  c->norms_done = 0;
  return CLR_OK;
}

int clr_launch_norms(clr_ctx *c, int lognormal_fused, double sigma2, int cap, int *nz_out,
                     unsigned long long *counts, double *zsum, double *sums) {
  if (!c->have_fields) return clr_fail(c, CLR_ERR_STATE, "clr_gaussian_fields must run first");
  if (lognormal_fused && c->density_done)
    return clr_fail(c, CLR_ERR_STATE, "density already transformed; call with lognormal_fused=0");
  if (!lognormal_fused && !c->density_done)
    return clr_fail(c, CLR_ERR_STATE, "clr_density must run first (or use lognormal_fused)");
  const int N = c->N, na = c->na;
  const double *r = c->bg, *zt = c->bg + na, *d1 = c->bg + 2 * na;
  double lastv[2];
  CLR_HIP(c, hipMemcpy(&lastv[0], zt + na - 1, sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipMemcpy(&lastv[1], d1 + na - 1, sizeof(double), hipMemcpyDeviceToHost));
  const double z_last = lastv[0], d1_last = lastv[1];

  // shells: zmax = z(L/2), nz = (int)(zmax/DZ_SIGMA)+2, idz = (nz-2)/zmax
  double zmax;
  {
    double hv[2];
    const double rr = 0.5 * c->L;
    const int ir = (int)(rr * c->glob_idr);
    if (rr <= 0) zmax = 0;
    else if (rr >= c->r_last) zmax = z_last;
    else {
#pragma clang fp contract(off)
      CLR_HIP(c, hipMemcpy(hv, zt + ir, 2 * sizeof(double), hipMemcpyDeviceToHost));
      double rv;
      CLR_HIP(c, hipMemcpy(&rv, r + ir, sizeof(double), hipMemcpyDeviceToHost));
      zmax = hv[0] + (hv[1] - hv[0]) * (rr - rv) * c->glob_idr;
    }
  }
  const int nz = (int)(zmax / 0.05) + 2;
  const double idz = (nz - 2) / zmax;

  TracerArgs ta{};
  int tk[3 * CLR_NPOP_MAX], tp[3 * CLR_NPOP_MAX];
  ta.n = 0;
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < c->n_tr[k]; i++) {
      if (!c->tr[k][i].set) return clr_fail(c, CLR_ERR_STATE, "tracer %d/%d has no tables", k, i);
      ta.bz[ta.n] = c->tr[k][i].bz;
      ta.bz0[ta.n] = c->tr[k][i].bz0;
      ta.model[ta.n] = c->tr[k][i].model;
      tk[ta.n] = k;
      tp[ta.n] = i;
      ta.n++;
    }
  const int nq = 2 + ta.n;
  const size_t lds = (size_t)kWaves * nz * nq * sizeof(double);
  if (lds > 160 * 1024) return clr_fail(c, CLR_ERR_ARG, "too many shells x tracers (%d x %d)", nz, ta.n);
  if (lds > 64 * 1024) {
    CLR_HIP(c, hipFuncSetAttribute((const void *)k_shells, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
  }
  const size_t nvals = (size_t)nz * nq;
  size_t need = (size_t)kShellBlocks * nvals + nvals + 2 * nz + 2;
  int rc = clr_ensure(c, (void **)&c->red, &c->red_n, need * sizeof(double));
  if (rc) return rc;
  double *partial = c->red, *fin = c->red + (size_t)kShellBlocks * nvals, *work = fin + nvals;

  CLR_HIP(c, hipEventRecord(c->ev0, c->stream));
  const int fused = lognormal_fused ? 1 : 0;
  hipLaunchKernelGGL(k_shells, dim3(kShellBlocks), dim3(kShellThreads), lds, c->stream, c->dens, N, c->ngx,
                     c->dx, c->obs[0], c->obs[1], c->obs[2], r, zt, d1, c->glob_idr, c->r_last, z_last,
                     d1_last, sigma2, fused, nz, idz, ta, partial);
  CLR_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(k_shell_final, dim3((unsigned)((nvals + 255) / 256)), dim3(256), 0, c->stream, partial,
                     kShellBlocks, (int)nvals, fin);
  CLR_HIP(c, hipGetLastError());
  for (int t = 0; t < ta.n; t++) {
    TracerDev &T = c->tr[tk[t]][tp[t]];
    hipLaunchKernelGGL(k_norm_table, dim3(1), dim3(256), 0, c->stream, fin, nz, nq, t, zmax, r, zt,
                       c->glob_idr, c->r_last, z_last, na, work, T.norm, work + 2 * nz);
    CLR_HIP(c, hipGetLastError());
    double edges[2];
    CLR_HIP(c, hipMemcpyAsync(edges, work + 2 * nz, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CLR_HIP(c, hipStreamSynchronize(c->stream));
    T.norm0 = edges[0];
    T.normf = edges[1];
    T.norm_set = 1;
  }
  CLR_HIP(c, hipEventRecord(c->ev1, c->stream));
  std::vector<double> h(nvals);
  CLR_HIP(c, hipMemcpyAsync(h.data(), fin, nvals * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->stage_ms[ST_NORM] = ms;
  if (nz_out) *nz_out = nz;
  if ((counts || zsum || sums) && cap < nz)
    return clr_fail(c, CLR_ERR_ARG, "output capacity %d < nz=%d", cap, nz);
  for (int iz = 0; iz < nz; iz++) {
    if (counts) counts[iz] = (unsigned long long)h[(size_t)iz * nq + 0];
    if (zsum) zsum[iz] = h[(size_t)iz * nq + 1];
    if (sums)
      for (int t = 0; t < ta.n; t++) sums[(size_t)t * nz + iz] = h[(size_t)iz * nq + 2 + t];
  }
  if (fused) c->density_done = 1;
  c->norms_done = 1;
  return CLR_OK;
}
