// clr_sources.hip -- srcs_set_cartesian_single (srcs.c:120-283) and
// srcs_get_local_properties_single (srcs.c:386-416) on gfx950.
//
//   k_count : one block per (z,y) row; per cell lambda and the Poisson count
//             (clr_poisson.h), block total -> rowcnt[row].
//   k_scan_*: exclusive prefix of the row totals (64-bit) -> row offsets.
//   k_fill  : recomputes each cell's count (counter-based draws make this
//             exact), scans the row in x order and writes the sources of
//             every occupied cell: uniform position in [x-dx/2, x+dx/2)
//             (srcs.c:261-263), the RSD shift from central differences of
//             phi (get_rvel, srcs.c:87-118) and the NEST base pixel.
// The catalogue is in global z-major cell order, sources of a cell in draw
// order, as the reference's thread-ordered offsets produce on one thread.
#include "clr_internal.h"
#include "clr_poisson.h"
#include <math.h>

using namespace clr;

namespace {

struct PopArgs {
  const double *r_arr, *amp, *bz, *norm, *v1;
  double idr, r_last, bz0, norm0, normf, v1_0, v1_last;
  double dx, cell_vol, half_box_20, factor_vel, idx;
  double ox, oy, oz;
  int N, ngx, model, ipop;
  uint32_t seed;
};

__device__ inline int cell_count(const PopArgs &a, int ix, int iy, int iz, float dl, double *rr_out,
                                 double *xyz) {
#pragma clang fp contract(off)
  const double obs[3] = {a.ox, a.oy, a.oz};
  double x0, y0, z0;
  const double r = cell_radius(ix, iy, iz, a.dx, obs, &x0, &y0, &z0);
  xyz[0] = x0; xyz[1] = y0; xyz[2] = z0;
  *rr_out = r;
  if (!(r < a.half_box_20)) return 0;
  const double nd = f_of_r(a.r_arr, a.amp, a.idr, a.r_last, r, 0.0, 0.0);
  if (!(nd > 0)) return 0;
  const double bias = f_of_r(a.r_arr, a.bz, a.idr, a.r_last, r, a.bz0, 1.0);
  const double dnorm = f_of_r(a.r_arr, a.norm, a.idr, a.r_last, r, a.norm0, a.normf);
  const uint64_t cell = (uint64_t)ix + (uint64_t)a.N * ((uint64_t)iy + (uint64_t)a.N * (uint64_t)iz);
  const uint32_t stream = stream_poisson(a.ipop);
  const u32x4 rr = draw(cell, stream, 0u, a.seed);
  const float lam32 = (float)(nd * a.cell_vol * dnorm) * bias_model_f32(dl, (float)bias, a.model);
  const int k = poisson_fast_f32(lam32, rr.x);
  if (k >= 0) return k;
  const double lam = nd * a.cell_vol * bias_model_f64((double)dl, bias, a.model) * dnorm;
  return poisson_exact(lam, cell, stream, a.seed, rr.x);
}

__device__ inline int64_t block_sum_i64(int64_t v, int64_t *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  int64_t t = 0;
  for (int i = 0; i < (int)(blockDim.x / 64); i++) t += sh[i];
  return t;
}

__global__ __launch_bounds__(256) void k_count(const float *__restrict__ g, PopArgs a,
                                               int32_t *__restrict__ rowcnt) {
  __shared__ int64_t sh[4];
  const size_t row = blockIdx.x;
  const int iz = (int)(row / a.N), iy = (int)(row % a.N);
  const float *p = g + row * (size_t)a.ngx;
  int64_t tot = 0;
  for (int ix = threadIdx.x; ix < a.N; ix += blockDim.x) {
    double rr, xyz[3];
    tot += cell_count(a, ix, iy, iz, p[ix], &rr, xyz);
  }
  tot = block_sum_i64(tot, sh);
  if (threadIdx.x == 0) rowcnt[row] = (int32_t)tot;
}

// exclusive scan of n int32 -> int64 offsets; out[n] = total.  Three kernels:
// per-block sums, scan of block sums (one block), then the block-local scan.
constexpr int kScanB = 1024;
__global__ __launch_bounds__(kScanB) void k_scan_blocks(const int32_t *__restrict__ in, size_t n,
                                                        long long *__restrict__ bsum) {
  __shared__ long long s[kScanB];
  const size_t i = (size_t)blockIdx.x * kScanB + threadIdx.x;
  s[threadIdx.x] = i < n ? in[i] : 0;
  __syncthreads();
  for (int w = kScanB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = s[0];
}
__global__ __launch_bounds__(kScanB) void k_scan_top(long long *__restrict__ bsum, size_t nb) {
  __shared__ long long s[kScanB];
  __shared__ long long carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (size_t base = 0; base < nb; base += kScanB) {
    const size_t i = base + threadIdx.x;
    const long long v = i < nb ? bsum[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanB; off <<= 1) {
      long long t = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == kScanB - 1) carry += s[kScanB - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}
__global__ __launch_bounds__(kScanB) void k_scan_apply(const int32_t *__restrict__ in, size_t n,
                                                       const long long *__restrict__ bsum,
                                                       long long *__restrict__ out) {
  __shared__ long long s[kScanB];
  const size_t i = (size_t)blockIdx.x * kScanB + threadIdx.x;
  const long long v = i < n ? in[i] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kScanB; off <<= 1) {
    long long t = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = bsum[blockIdx.x] + s[threadIdx.x] - v;
}

__device__ inline double get_rvel(const float *__restrict__ phi, const PopArgs &a, int ix, int iy, int iz,
                                  double x0, double y0, double z0, double rr) {
#pragma clang fp contract(off)
  const int N = a.N;
  const size_t ngx = (size_t)a.ngx;
  const int ixh = ix == N - 1 ? 0 : ix + 1, ixl = ix == 0 ? N - 1 : ix - 1;
  const int iyh = iy == N - 1 ? 0 : iy + 1, iyl = iy == 0 ? N - 1 : iy - 1;
  const int izh = iz == N - 1 ? 0 : iz + 1, izl = iz == 0 ? N - 1 : iz - 1;
  const size_t z0o = (size_t)iz * N * ngx, y0o = (size_t)iy * ngx;
  const double u0 = x0 / rr, u1 = y0 / rr, u2 = z0 / rr;
  const double v0 = (double)phi[ixh + y0o + z0o] - (double)phi[ixl + y0o + z0o];
  const double v1 = (double)phi[ix + (size_t)iyh * ngx + z0o] - (double)phi[ix + (size_t)iyl * ngx + z0o];
  const double v2 = (double)phi[ix + y0o + (size_t)izh * N * ngx] - (double)phi[ix + y0o + (size_t)izl * N * ngx];
  return 0.5 * a.idx * (v0 * u0 + v1 * u1 + v2 * u2);
}

__global__ __launch_bounds__(256) void k_fill(const float *__restrict__ g, const float *__restrict__ phi,
                                              PopArgs a, const long long *__restrict__ rowoff, int nside_base,
                                              float4 *__restrict__ pos, int32_t *__restrict__ ipix) {
#pragma clang fp contract(off)
  __shared__ int32_t s[256];
  __shared__ long long chunk_base;
  const size_t row = blockIdx.x;
  const int iz = (int)(row / a.N), iy = (int)(row % a.N);
  const float *p = g + row * (size_t)a.ngx;
  if (threadIdx.x == 0) chunk_base = rowoff[row];
  for (int x0i = 0; x0i < a.N; x0i += blockDim.x) {
    const int ix = x0i + threadIdx.x;
    int k = 0;
    double rr = 0, xyz[3] = {0, 0, 0};
    if (ix < a.N) k = cell_count(a, ix, iy, iz, p[ix], &rr, xyz);
    __syncthreads();
    s[threadIdx.x] = k;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
      const int t = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    const long long base = chunk_base + s[threadIdx.x] - k;
    if (k > 0) {
      const uint64_t cell = (uint64_t)ix + (uint64_t)a.N * ((uint64_t)iy + (uint64_t)a.N * (uint64_t)iz);
      const double rvel = a.factor_vel * get_rvel(phi, a, ix, iy, iz, xyz[0], xyz[1], xyz[2], rr);
      const double v1 = f_of_r(a.r_arr, a.v1, a.idr, a.r_last, rr, a.v1_0, a.v1_last);
      const double dz_rsd = rvel * v1;
      const uint32_t stream = stream_position(a.ipop);
      u32x4 rnd{};
      for (int ip = 0; ip < k; ip++) {
        float q[3];
        for (int ax = 0; ax < 3; ax++) {
          const uint32_t j = 3u * (uint32_t)ip + (uint32_t)ax;
          if ((j & 3u) == 0u) rnd = draw(cell, stream, j >> 2, a.seed);
          q[ax] = (float)(xyz[ax] + a.dx * (u01(word(rnd, (int)(j & 3u))) - 0.5));
        }
        const long long pid = base + ip;
        pos[pid] = make_float4(q[0], q[1], q[2], (float)dz_rsd);
        const int64_t pr = vec2pix_ring(nside_base, (double)q[0], (double)q[1], (double)q[2]);
        ipix[pid] = (int32_t)ring2nest(nside_base, pr);
      }
    }
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) chunk_base += s[threadIdx.x];
    __syncthreads();
  }
}

// Src fields of srcs_get_local_properties_single: RA, Dec, z0, dz_rsd.
__global__ void k_src_props(const float4 *__restrict__ pos, long long n, const double *__restrict__ r_arr,
                            const double *__restrict__ zt, double idr, double r_last, double z_last,
                            float4 *__restrict__ out) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = pos[i];
  const double x = p.x, y = p.y, z = p.z;
  const double r = sqrt(x * x + y * y + z * z);
  double cth = 1, phi = 0;
  if (r != 0) {
    const double xn = x / r, yn = y / r;
    cth = z / r;
    phi = atan2(yn, xn);
    if (phi < 0) phi += 2 * M_PI;
  }
  const double RTOD = 57.2957795;
  out[i] = make_float4((float)(RTOD * phi), (float)(90 - RTOD * acos(cth)),
                       (float)f_of_r(r_arr, zt, idr, r_last, r, 0.0, z_last), p.w);
}

}  // namespace

int clr_launch_sources(clr_ctx *c, int ipop, uint32_t seed, int nside_base, long long *nsrc) {
  if (ipop < 0 || ipop >= c->n_tr[CLR_TRACER_SRCS] || !c->tr[CLR_TRACER_SRCS][ipop].set)
    return clr_fail(c, CLR_ERR_ARG, "no source population %d", ipop);
  TracerDev &T = c->tr[CLR_TRACER_SRCS][ipop];
  if (!T.norm_set) return clr_fail(c, CLR_ERR_STATE, "population %d has no norm table", ipop);
  if (!c->density_done) return clr_fail(c, CLR_ERR_STATE, "density field not computed");
  if (nside_base < 1 || (nside_base & (nside_base - 1)))
    return clr_fail(c, CLR_ERR_ARG, "nside_base=%d must be a power of two", nside_base);
  const int N = c->N, na = c->na;
  PopArgs a{};
  a.r_arr = c->bg;
  a.v1 = c->bg + 3 * na;
  a.amp = T.amp;
  a.bz = T.bz;
  a.norm = T.norm;
  a.idr = c->glob_idr;
  a.r_last = c->r_last;
  a.bz0 = T.bz0;
  a.norm0 = T.norm0;
  a.normf = T.normf;
  double v1e[2];
  CLR_HIP(c, hipMemcpy(&v1e[0], a.v1, sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipMemcpy(&v1e[1], a.v1 + na - 1, sizeof(double), hipMemcpyDeviceToHost));
  a.v1_0 = v1e[0];
  a.v1_last = v1e[1];
  a.dx = c->L / N;
  a.cell_vol = a.dx * a.dx * a.dx;
  a.half_box_20 = c->L / 2 + 20.;
  a.factor_vel = -c->fgrowth_0 / (1.5 * c->hubble_0 * c->omega_m);
  a.idx = N / c->L;
  a.ox = c->obs[0];
  a.oy = c->obs[1];
  a.oz = c->obs[2];
  a.N = N;
  a.ngx = c->ngx;
  a.model = T.model;
  a.ipop = ipop;
  a.seed = seed;

  const size_t nrows = (size_t)N * N;
  const size_t nb = (nrows + kScanB - 1) / kScanB;
  int rc = clr_ensure(c, (void **)&c->rowcnt, &c->rows_n, nrows * sizeof(int32_t) + 8 +
                      (nrows + 1) * sizeof(long long) + (nb + 1) * sizeof(long long));
  if (rc) return rc;
  int32_t *rowcnt = c->rowcnt;
  long long *rowoff = (long long *)(rowcnt + nrows + (nrows & 1));
  long long *bsum = rowoff + nrows + 1;

  CLR_HIP(c, hipEventRecord(c->ev0, c->stream));
  const int bt = N >= 256 ? 256 : ((N + 63) / 64) * 64;
  hipLaunchKernelGGL(k_count, dim3((unsigned)nrows), dim3(bt), 0, c->stream, c->dens, a, rowcnt);
  CLR_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(kScanB), 0, c->stream, rowcnt, nrows, bsum);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanB), 0, c->stream, bsum, nb);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kScanB), 0, c->stream, rowcnt, nrows, bsum,
                     rowoff);
  CLR_HIP(c, hipGetLastError());
  long long total = 0;
  CLR_HIP(c, hipMemcpyAsync(&total, bsum + nb, sizeof(long long), hipMemcpyDeviceToHost, c->stream));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  CatalogDev &C = c->cat[ipop];
  if (total > C.cap) {
    if (C.pos) hipFree(C.pos);
    if (C.ipix) hipFree(C.ipix);
    C.pos = nullptr;
    C.ipix = nullptr;
    C.cap = 0;
    const long long cap = total + total / 16 + 1024;
    if (hipMalloc((void **)&C.pos, (size_t)cap * 4 * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&C.ipix, (size_t)cap * sizeof(int32_t)) != hipSuccess) {
      (void)hipGetLastError();
      return clr_fail(c, CLR_ERR_OOM, "cannot allocate catalogue of %lld sources", total);
    }
    C.cap = cap;
  }
  C.n = total;
  if (total > 0) {
    hipLaunchKernelGGL(k_fill, dim3((unsigned)nrows), dim3(bt), 0, c->stream, c->dens, c->npot, a, rowoff,
                       nside_base, (float4 *)C.pos, C.ipix);
    CLR_HIP(c, hipGetLastError());
  }
  CLR_HIP(c, hipEventRecord(c->ev1, c->stream));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->stage_ms[ST_SRCS] = ms;
  if (nsrc) *nsrc = total;
  return CLR_OK;
}

int clr_launch_source_props(clr_ctx *c, int ipop, float *out, long long n) {
  if (ipop < 0 || ipop >= CLR_NPOP_MAX) return clr_fail(c, CLR_ERR_ARG, "bad population %d", ipop);
  CatalogDev &C = c->cat[ipop];
  if (n != C.n) return clr_fail(c, CLR_ERR_ARG, "population %d has %lld sources, not %lld", ipop, C.n, n);
  if (n == 0) return CLR_OK;
  const int na = c->na;
  double z_last;
  CLR_HIP(c, hipMemcpy(&z_last, c->bg + na + na - 1, sizeof(double), hipMemcpyDeviceToHost));
  float4 *tmp = nullptr;
  CLR_HIP(c, hipMalloc((void **)&tmp, (size_t)n * sizeof(float4)));
  hipLaunchKernelGGL(k_src_props, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     (const float4 *)C.pos, n, c->bg, c->bg + na, c->glob_idr, c->r_last, z_last, tmp);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, (size_t)n * sizeof(float4), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(tmp);
  if (e != hipSuccess) return clr_fail(c, CLR_ERR_HIP, "source properties: %s", hipGetErrorString(e));
  return CLR_OK;
}
