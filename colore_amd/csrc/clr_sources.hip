// clr_sources.hip -- srcs_set_cartesian_single (srcs.c:120-283) for one or
// all populations, and srcs_get_local_properties_single (srcs.c:386-416).
//
// A segment is up to 1024 consecutive cells of one (z,y) row; each lane owns
// 4 consecutive cells (two 8-byte loads, one Philox draw per population for
// the 4 Poisson uniforms).  Population geometry (radius, table index) is
// computed once per cell for every population of the pass.
//   k_count : per cell and population the Poisson count (clr_poisson.h);
//             per segment and population the total -> segcnt.
//   k_scan_*: exclusive prefix of the segment totals (64-bit), per population.
//   k_fill  : recomputes the counts (counter-based draws make this exact),
//             scans the segment in x order and writes the sources of every
//             occupied cell: uniform position in [x-dx/2, x+dx/2)
//             (srcs.c:261-263), RSD shift from central differences of phi
//             (get_rvel, srcs.c:87-118), NEST base pixel (vec2pix_ring then a
//             RING->NEST table, srcs.c:264-270).
// Each catalogue is in global z-major cell order, sources of a cell in draw
// order -- the order the reference produces on one thread.
#include "clr_internal.h"
#include "clr_poisson.h"
#include <math.h>

using namespace clr;

namespace {

struct PopTab {
  const double *amp, *bz, *norm;
  const float *bzf, *normf32;
  double bz0, norm0, normf;
  int model, ipop;
  float4 *pos;
  int32_t *ipix;
};

struct SrcArgs {
  const float *g, *phi;
  const double *r_arr, *v1;
  double idr, r_last, v1_0, v1_last;
  double dx, cell_vol, half_box_20, factor_vel, idx, ox, oy, oz;
  float idr_f;
  int N, ngx, na, nseg, seglen, npop, nside_base;
  uint32_t seed;
  const int32_t *r2n;
  PopTab pop[CLR_NPOP_MAX];
};

// FP32 table lookup for values that only feed the FP32 lambda estimate
__device__ inline float f_of_r32(const float *__restrict__ f, float idr, int nm1, float r, float f0, float ff) {
  if (r <= 0.f) return f0;
  const float x = r * idr;
  if (x >= (float)nm1) return ff;
  const int ir = (int)x;
  const float t = x - (float)ir;
  return f[ir] + (f[ir + 1] - f[ir]) * t;
}

// Count of cell `cell` for population P given its exact radius r.
__device__ inline int cell_count(const SrcArgs &a, const PopTab &P, double r, float dl, uint64_t cell,
                                 uint32_t word0) {
#pragma clang fp contract(off)
  if (!(r < a.half_box_20)) return 0;
  const double nd = f_of_r(a.r_arr, P.amp, a.idr, a.r_last, r, 0.0, 0.0);
  if (!(nd > 0)) return 0;
  const float rf = (float)r;
  const float b32 = f_of_r32(P.bzf, a.idr_f, a.na - 1, rf, (float)P.bz0, 1.f);
  const float n32 = f_of_r32(P.normf32, a.idr_f, a.na - 1, rf, (float)P.norm0, (float)P.normf);
  const float lam32 = (float)(nd * a.cell_vol) * n32 * bias_model_f32(dl, b32, P.model);
  const int k = poisson_fast_f32(lam32, word0);
  if (k >= 0) return k;
  const double bias = f_of_r(a.r_arr, P.bz, a.idr, a.r_last, r, P.bz0, 1.0);
  const double dnorm = f_of_r(a.r_arr, P.norm, a.idr, a.r_last, r, P.norm0, P.normf);
  const double lam = nd * a.cell_vol * bias_model_f64((double)dl, bias, P.model) * dnorm;
  return poisson_exact(lam, cell, stream_poisson_extra(P.ipop), a.seed, word0);
}

// The lane's 4 cells: delta values, radii and offsets.
struct Quad {
  float d[4];
  double r[4], x0[4];
  double y0, z0;
  int ix0, iy, iz;
  bool ok[4];
  uint64_t cell0;
};

__device__ inline void load_quad(const SrcArgs &a, size_t seg, int lane, Quad &q) {
#pragma clang fp contract(off)
  const size_t row = seg / a.nseg;
  const int s = (int)(seg % a.nseg);
  q.iz = (int)(row / a.N);
  q.iy = (int)(row % a.N);
  q.ix0 = s * a.seglen + 4 * lane;
  const bool in = q.ix0 < a.N && 4 * lane < a.seglen;
  const float *p = a.g + row * (size_t)a.ngx + (in ? q.ix0 : 0);
  float2 v0 = make_float2(0.f, 0.f), v1 = v0;
  if (in) {
    v0 = *(const float2 *)p;
    v1 = *(const float2 *)(p + 2);
  }
  q.d[0] = v0.x; q.d[1] = v0.y; q.d[2] = v1.x; q.d[3] = v1.y;
  q.z0 = (q.iz + 0.0) * a.dx - a.oz;
  q.y0 = (q.iy + 0.0) * a.dx - a.oy;
  q.cell0 = (uint64_t)q.ix0 + (uint64_t)a.N * ((uint64_t)q.iy + (uint64_t)a.N * (uint64_t)q.iz);
#pragma unroll
  for (int c = 0; c < 4; c++) {
    q.ok[c] = in;
    const double xx = (q.ix0 + c + 0.0) * a.dx - a.ox;
    q.x0[c] = xx;
    q.r[c] = sqrt(xx * xx + q.y0 * q.y0 + q.z0 * q.z0);
  }
}

__device__ inline void quad_counts(const SrcArgs &a, const PopTab &P, const Quad &q, int k[4]) {
  u32x4 rr{0u, 0u, 0u, 0u};
  bool any = false;
#pragma unroll
  for (int c = 0; c < 4; c++) any |= q.ok[c] && q.r[c] < a.half_box_20;
  if (any) rr = draw(q.cell0 >> 2, stream_poisson(P.ipop), 0u, a.seed);
#pragma unroll
  for (int c = 0; c < 4; c++) k[c] = q.ok[c] ? cell_count(a, P, q.r[c], q.d[c], q.cell0 + c, word(rr, c)) : 0;
}

template <class T>
__device__ inline T block_excl_scan(T v, T *sh, T *total) {
  // inclusive scan within the wavefront, then across wavefronts through LDS
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64, nw = blockDim.x / 64;
  T x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const T y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  __syncthreads();
  if (lane == 63) sh[w] = x;
  __syncthreads();
  T base = 0, tot = 0;
  for (int i = 0; i < nw; i++) {
    if (i < w) base += sh[i];
    tot += sh[i];
  }
  *total = tot;
  return base + x - v;
}

__global__ __launch_bounds__(256) void k_count(SrcArgs a, int32_t *__restrict__ segcnt, size_t nsegs) {
  __shared__ int32_t sh[CLR_NPOP_MAX][4];
  const size_t seg = blockIdx.x;
  const int lane = threadIdx.x;
  Quad q;
  load_quad(a, seg, lane, q);
  for (int ip = 0; ip < a.npop; ip++) {
    int k[4];
    quad_counts(a, a.pop[ip], q, k);
    int t = k[0] + k[1] + k[2] + k[3];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
    if ((threadIdx.x & 63) == 0) sh[ip][threadIdx.x >> 6] = t;
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)a.npop) {
    int t = 0;
    for (int w = 0; w < (int)(blockDim.x / 64); w++) t += sh[threadIdx.x][w];
    segcnt[(size_t)threadIdx.x * nsegs + seg] = t;
  }
}

// exclusive scan of n int32 -> int64 offsets, bsum[nb] = total.  Three
// kernels: per-block sums, scan of the block sums (one block), block-local scan.
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

__device__ inline double get_rvel(const SrcArgs &a, int ix, int iy, int iz, double x0, double y0, double z0,
                                  double rr) {
#pragma clang fp contract(off)
  const float *__restrict__ phi = a.phi;
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

// RING index at a power-of-two nside of (x,y,z): the arithmetic of
// ang2pix_ring in 32-bit integers (nside <= 8192 keeps every term < 2^31).
__device__ inline int vec2pix_ring32(int nside, double x, double y, double z) {
#pragma clang fp contract(off)
  const double vlen = sqrt(x * x + y * y + z * z);
  if (!(vlen > 0)) return 0;   // a source exactly at the observer
  const double cth = z / vlen;
  double ph = atan2(y, x);
  const double twopi = 6.283185307179586476925286766559;
  if (ph < 0 || ph >= twopi) {
    ph = fmod(ph, twopi);
    if (ph < 0) { ph += twopi; if (ph == twopi) ph = 0; }
  }
  const double tt = ph * 0.6366197723675813430755350534900574;
  const double za = fabs(cth);
  if (za <= 2.0 / 3.0) {
    const double t1 = nside * (0.5 + tt), t2 = nside * cth * 0.75;
    const int jp = (int)(t1 - t2), jm = (int)(t1 + t2);
    const int ir = nside + 1 + jp - jm;
    const int kshift = 1 - (ir & 1);
    const int ip = ((jp + jm - nside + kshift + 1) / 2) & (4 * nside - 1);
    return nside * (nside - 1) * 2 + (ir - 1) * 4 * nside + ip;
  }
  const double tp = tt - (int)tt;
  const double tmp = nside * sqrt(3 * (1 - za));
  const int jp = (int)(tp * tmp), jm = (int)((1.0 - tp) * tmp);
  const int ir = jp + jm + 1;
  int ip = (int)(tt * ir);
  if (ip >= 4 * ir) ip -= 4 * ir;
  if (cth > 0) return 2 * ir * (ir - 1) + ip;
  return 12 * nside * nside - 2 * ir * (ir + 1) + ip;
}

__global__ __launch_bounds__(256) void k_fill(SrcArgs a, const long long *__restrict__ segoff, size_t nsegs) {
#pragma clang fp contract(off)
  __shared__ long long sh[4];
  const size_t seg = blockIdx.x;
  const int lane = threadIdx.x;
  Quad q;
  load_quad(a, seg, lane, q);
  for (int ip = 0; ip < a.npop; ip++) {
    const PopTab &P = a.pop[ip];
    int k[4];
    quad_counts(a, P, q, k);
    const long long mine = k[0] + k[1] + k[2] + k[3];
    long long tot;
    long long base = block_excl_scan<long long>(mine, sh, &tot) + segoff[(size_t)ip * (nsegs + 1) + seg];
    if (mine == 0) continue;
    const uint32_t stream = stream_position(P.ipop);
    for (int c = 0; c < 4; c++) {
      if (k[c] == 0) continue;
      const int ix = q.ix0 + c;
      const uint64_t cell = q.cell0 + c;
      const double rvel = a.factor_vel * get_rvel(a, ix, q.iy, q.iz, q.x0[c], q.y0, q.z0, q.r[c]);
      const double dz_rsd = rvel * f_of_r(a.r_arr, a.v1, a.idr, a.r_last, q.r[c], a.v1_0, a.v1_last);
      const double ctr[3] = {q.x0[c], q.y0, q.z0};
      u32x4 rnd{};
      for (int is = 0; is < k[c]; is++) {
        float pp[3];
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
          const uint32_t j = 3u * (uint32_t)is + (uint32_t)ax;
          if ((j & 3u) == 0u) rnd = draw(cell, stream, j >> 2, a.seed);
          pp[ax] = (float)(ctr[ax] + a.dx * (u01(word(rnd, (int)(j & 3u))) - 0.5));
        }
        P.pos[base] = make_float4(pp[0], pp[1], pp[2], (float)dz_rsd);
        P.ipix[base] = a.r2n[vec2pix_ring32(a.nside_base, (double)pp[0], (double)pp[1], (double)pp[2])];
        base++;
      }
    }
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
  const int nall = c->n_tr[CLR_TRACER_SRCS];
  if (ipop < -1 || ipop >= nall) return clr_fail(c, CLR_ERR_ARG, "no source population %d", ipop);
  if (!c->density_done) return clr_fail(c, CLR_ERR_STATE, "density field not computed");
  if (nside_base < 1 || nside_base > 64 || (nside_base & (nside_base - 1)))
    return clr_fail(c, CLR_ERR_ARG, "nside_base=%d must be a power of two <= 64", nside_base);
  const int p0 = ipop < 0 ? 0 : ipop, p1 = ipop < 0 ? nall : ipop + 1;
  const int N = c->N, na = c->na;
  for (int i = p0; i < p1; i++) {
    const TracerDev &T = c->tr[CLR_TRACER_SRCS][i];
    if (!T.set || !T.norm_set) return clr_fail(c, CLR_ERR_STATE, "population %d has no norm table", i);
  }
  int rc = clr_ring2nest_table(c, nside_base);
  if (rc) return rc;
  // FP32 copies of the norm tables (allocated on first use): before the
  // kernel arguments take their addresses
  for (int i = p0; i < p1; i++) {
    rc = clr_norm_f32(c, c->tr[CLR_TRACER_SRCS][i]);
    if (rc) return rc;
  }
  SrcArgs a{};
  a.g = c->dens;
  a.phi = c->npot;
  a.r_arr = c->bg;
  a.v1 = c->bg + 3 * (size_t)na;
  a.idr = c->glob_idr;
  a.r_last = c->r_last;
  a.v1_0 = c->h_bg_first[3];
  a.v1_last = c->h_bg_last[3];
  a.dx = c->L / N;
  a.cell_vol = a.dx * a.dx * a.dx;
  a.half_box_20 = c->L / 2 + 20.;
  a.factor_vel = -c->fgrowth_0 / (1.5 * c->hubble_0 * c->omega_m);
  a.idx = N / c->L;
  a.ox = c->obs[0];
  a.oy = c->obs[1];
  a.oz = c->obs[2];
  a.idr_f = (float)c->glob_idr;
  a.N = N;
  a.ngx = c->ngx;
  a.na = na;
  a.seglen = N < 1024 ? N : 1024;
  a.nseg = N / a.seglen;
  a.npop = p1 - p0;
  a.nside_base = nside_base;
  a.seed = seed;
  a.r2n = c->r2n;
  for (int i = p0; i < p1; i++) {
    const TracerDev &T = c->tr[CLR_TRACER_SRCS][i];
    PopTab &P = a.pop[i - p0];
    P.amp = T.amp;
    P.bz = T.bz;
    P.norm = T.norm;
    P.bzf = T.bzf;
    P.normf32 = T.normf32;
    P.bz0 = T.bz0;
    P.norm0 = T.norm0;
    P.normf = T.normf;
    P.model = T.model;
    P.ipop = i;
  }
  const size_t nsegs = (size_t)N * N * a.nseg;
  const size_t nb = (nsegs + kScanB - 1) / kScanB;
  const int np = a.npop;
  rc = clr_ensure(c, (void **)&c->rowcnt, &c->rows_n,
                  np * (nsegs * sizeof(int32_t) + 8 + (nsegs + 1) * sizeof(long long) + (nb + 1) * sizeof(long long)));
  if (rc) return rc;
  int32_t *segcnt = c->rowcnt;
  long long *segoff = (long long *)(((uintptr_t)(segcnt + np * nsegs) + 7) & ~(uintptr_t)7);
  long long *bsum = segoff + np * (nsegs + 1);

  const int bt = ((a.seglen / 4) + 63) / 64 * 64;
  for (int i = p0; i < p1; i++)
    if (!a.pop[i - p0].normf32 || !a.pop[i - p0].bzf || !a.pop[i - p0].norm || !a.pop[i - p0].amp)
      return clr_fail(c, CLR_ERR_STATE, "population %d tables are not on the device", i);
  if (!a.r2n) return clr_fail(c, CLR_ERR_STATE, "no RING->NEST table");
  CLR_HIP(c, hipEventRecord(c->ev0, c->stream));
  hipLaunchKernelGGL(k_count, dim3((unsigned)nsegs), dim3(bt), 0, c->stream, a, segcnt, nsegs);
  CLR_HIP(c, hipGetLastError());
  for (int i = 0; i < np; i++) {
    const int32_t *in = segcnt + (size_t)i * nsegs;
    long long *bs = bsum + (size_t)i * (nb + 1);
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(kScanB), 0, c->stream, in, nsegs, bs);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanB), 0, c->stream, bs, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kScanB), 0, c->stream, in, nsegs, bs,
                       segoff + (size_t)i * (nsegs + 1));
    CLR_HIP(c, hipGetLastError());
  }
  std::vector<long long> tot(np);
  for (int i = 0; i < np; i++)
    CLR_HIP(c, hipMemcpyAsync(&tot[i], bsum + (size_t)i * (nb + 1) + nb, sizeof(long long), hipMemcpyDeviceToHost,
                              c->stream));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  long long total = 0;
  for (int i = 0; i < np; i++) {
    CatalogDev &C = c->cat[p0 + i];
    if (tot[i] > C.cap) {
      if (C.pos) hipFree(C.pos);
      if (C.ipix) hipFree(C.ipix);
      C.pos = nullptr;
      C.ipix = nullptr;
      C.cap = 0;
      const long long cap = tot[i] + tot[i] / 16 + 1024;
      if (hipMalloc((void **)&C.pos, (size_t)cap * 4 * sizeof(float)) != hipSuccess ||
          hipMalloc((void **)&C.ipix, (size_t)cap * sizeof(int32_t)) != hipSuccess) {
        (void)hipGetLastError();
        return clr_fail(c, CLR_ERR_OOM, "cannot allocate catalogue of %lld sources", tot[i]);
      }
      C.cap = cap;
    }
    C.n = tot[i];
    a.pop[i].pos = (float4 *)C.pos;
    a.pop[i].ipix = C.ipix;
    total += tot[i];
  }
  if (total > 0) {
    hipLaunchKernelGGL(k_fill, dim3((unsigned)nsegs), dim3(bt), 0, c->stream, a, segoff, nsegs);
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
  float4 *tmp = nullptr;
  CLR_HIP(c, hipMalloc((void **)&tmp, (size_t)n * sizeof(float4)));
  hipLaunchKernelGGL(k_src_props, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     (const float4 *)C.pos, n, c->bg, c->bg + na, c->glob_idr, c->r_last, c->h_bg_last[1], tmp);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, (size_t)n * sizeof(float4), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(tmp);
  if (e != hipSuccess) return clr_fail(c, CLR_ERR_HIP, "source properties: %s", hipGetErrorString(e));
  return CLR_OK;
}
