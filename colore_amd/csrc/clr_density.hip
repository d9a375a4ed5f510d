// clr_density.hip -- physical density (density.c:1034-1126) and the per-shell
// bias normalisation (density.c:1128-1393) on gfx950.
//
// k_transform : lognormalize / densclip in place.
// k_shells    : one pass over the grid.  Each wavefront walks tiles of
//               16 (x) x 16 (y) x 16 (z) cells, 64 rows of 16 x 4 cells (one
//               cell per lane: a compact footprint keeps the table bins the
//               lanes gather close together); per cell the
//               radius, then from one packed FP32 table D(r), z(r) and every
//               b_t(r); the shell index s = (int)(z*idz)+1 is taken from the
//               FP32 z where that is more than 5e-4 of a shell away from a
//               shell edge, otherwise from the reference's FP64 expressions
//               (so the counts match bit for bit); then count, z and
//               bias_model(delta, b_t(r)) for every tracer.  A lane's radius
//               changes slowly along its 64 cells, so it adds into private
//               FP64 accumulators of its current shell and logs finished
//               shells (at most kSlots) in its own LDS slots; the lanes are
//               combined once per tile with DPP wave sums, and lane 0 adds the
//               result into the wavefront's own LDS accumulators.  Every
//               addition happens in a fixed order: the sums are bitwise
//               repeatable.  With `fused` the lognormal transform is applied
//               in the same pass (one read and one write of the grid for both
//               stages).
// k_shell_final / k_norm_table : fixed-order reduction of the per-block
//               partials, norm = count/sum, edge fill and linear
//               interpolation in z onto the NA r-grid (density.c:1272-1355),
//               all on the device.
#include "clr_internal.h"
#include <math.h>
#include <vector>

using namespace clr;

namespace {

constexpr int kShellBlocks = 8192;
constexpr int kShellThreads = 256;
constexpr int kWaves = kShellThreads / 64;
// A wavefront row is kTX consecutive x cells of kSub = 64/kTX adjacent y
// rows (lane = x + kTX * sub-row): a short x extent keeps the radii, and so
// the table bins the lanes gather, close together.  A tile is kTY x kTZ such
// rows: kTX x (kTY kSub) x kTZ cells.
#ifndef CLR_SHELL_TX
#define CLR_SHELL_TX 16
#endif
constexpr int kTX = CLR_SHELL_TX, kSub = 64 / kTX;
constexpr int kTY = 4, kTZ = 16;
constexpr int kSlots = 1;
#ifndef CLR_SHELL_PREFETCH
#define CLR_SHELL_PREFETCH 2
#endif
constexpr int kPrefetch = CLR_SHELL_PREFETCH;   // rows of density in flight per wavefront
static_assert((kTY * kTZ) % kPrefetch == 0, "prefetch distance must divide the tile rows");
constexpr int kMaxTr = 8;

struct ShellArgs {
  float *g;
  int N, ngx, nz, ntr, fused, na;
  double dx, ox, oy, oz, idr, r_last, z_last, sigma2, idz;
  const double *r_arr, *zt;
  const double *sigma2_dev;   // when set, sigma^2 is read here instead of `sigma2`
  const float4 *tab;   // per bin: (value, slope) of D, z, b_0 .. b_{NT-1} (k_pack_shell_table)
  float d1_last_f, z_last_f, idr_f, idz_f;
  float bz0[kMaxTr];
  int model[kMaxTr];
};

// FP32 tables of one shell pass packed per bin i as (f_i, f_{i+1} - f_i) for
// quantity 0 = D(r), 1 = z(r), 2 + t = b_t(r): one gather per cell, and
// f_i + slope * t is the same FP32 value as f_i + (f_{i+1} - f_i) * t.
// shell_tab_w(ntr) float4s per bin.
struct ShellTabSrc {
  const float *d1f;
  const double *zt;
  const float *bzf[kMaxTr];
};
__host__ __device__ constexpr int shell_tab_w(int ntr) { return (2 * (2 + ntr) + 3) / 4; }
__global__ void k_pack_shell_table(ShellTabSrc src, int ntr, int na, float4 *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na) return;
  const int w = shell_tab_w(ntr), j = i + 1 < na ? i + 1 : i;
  float v[4 * shell_tab_w(kMaxTr)];
  for (int k = 0; k < 4 * w; k++) v[k] = 0.f;
  const float d0 = src.d1f[i], d1 = src.d1f[j], z0 = (float)src.zt[i], z1 = (float)src.zt[j];
  v[0] = d0;
  v[1] = d1 - d0;
  v[2] = z0;
  v[3] = z1 - z0;
  for (int t = 0; t < ntr; t++) {
    const float b0 = src.bzf[t][i], b1 = src.bzf[t][j];
    v[4 + 2 * t] = b0;
    v[5 + 2 * t] = b1 - b0;
  }
  for (int k = 0; k < w; k++) out[(size_t)i * w + k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
}

template <int CTRL, int RM>
__device__ inline double dpp_get(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int l2 = __builtin_amdgcn_update_dpp(0, lo, CTRL, RM, 0xF, true);
  const int h2 = __builtin_amdgcn_update_dpp(0, hi, CTRL, RM, 0xF, true);
  return __hiloint2double(h2, l2);
}

// Sum over the 64 lanes of a wavefront (all lanes active), broadcast.
__device__ inline double wave_sum(double x) {
  x += dpp_get<0xB1, 0xF>(x);    // quad_perm [1,0,3,2]
  x += dpp_get<0x4E, 0xF>(x);    // quad_perm [2,3,0,1]
  x += dpp_get<0x141, 0xF>(x);   // row_half_mirror
  x += dpp_get<0x140, 0xF>(x);   // row_mirror
  x += dpp_get<0x142, 0xA>(x);   // row_bcast:15 into rows 1,3
  x += dpp_get<0x143, 0xC>(x);   // row_bcast:31 into rows 2,3
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), 63);
  return __hiloint2double(hi, lo);
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

// Combine the slot-k accumulators of all lanes (slot values in LDS,
// vals[q * 64 + lane]): lanes are grouped by shell id (in order of the lowest
// lane) and each group is summed with DPP.
template <int NQ>
__device__ inline void flush_slot(unsigned long long mask, int key, const double *vals, double *myacc,
                                  int lane) {
  while (mask) {
    const int leader = __ffsll((long long)mask) - 1;
    const int sl = __builtin_amdgcn_readlane(key, leader);
    const bool mine = ((mask >> lane) & 1ull) && key == sl;
    const unsigned long long mm = __ballot(mine);
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const double v = wave_sum(mine ? vals[q * 64 + lane] : 0.0);
      if (lane == 0) myacc[sl * NQ + q] += v;
    }
    mask &= ~mm;
  }
}

// bias_model (common.h:414-431) for the shell sums, on the raw hardware
// exp2/log2 and a fast reciprocal (a few ulp: the sums are of FP32 terms),
// without branches (the model is uniform, the sign of d is not):
//   0: (1+d)^b = exp2(b l2),            l2 = log2(1+d)
//   1: d < 0: exp(b d/(1+d)) = exp2(b q), q = log2(e) d/(1+d); else 1 + b d
//   2: max(1 + b d, 0)
// l2 and q are shared by the tracers; d <= -1 (or NaN) gives 0.
__device__ inline float bias_sum_term(float d, float l2, float q, float b, int model) {
  const float e = __builtin_amdgcn_exp2f(b * (model == 0 ? l2 : q));
  const float lin = fmaxf(__builtin_fmaf(b, d, 1.f), 0.f);
  const bool use_exp = model == 0 || (model == 1 && d < 0.f);
  return d > -1.f ? (use_exp ? e : lin) : 0.f;
}

// LDS per block: kWaves x nz x NQ accumulators, then per wavefront kSlots x
// NQ x 64 logged shell sums and kSlots x 64 shell ids.
template <int NT>
constexpr size_t shell_lds_bytes(int nz) {
  return (size_t)kWaves * nz * (2 + NT) * sizeof(double) +
         (size_t)kWaves * kSlots * 64 * ((2 + NT) * sizeof(double) + sizeof(int));
}

// z(r) of cell (ix,iy,iz) by the reference's FP64 expressions (density.c:
// 1084-1094, f_of_r_linear).
__device__ inline double cell_z_f64(const ShellArgs &a, int ix, int iy, int iz) {
#pragma clang fp contract(off)
  const double obs[3] = {a.ox, a.oy, a.oz};
  double x0, y0, z0;
  const double r = cell_radius(ix, iy, iz, a.dx, obs, &x0, &y0, &z0);
  return f_of_r(a.r_arr, a.zt, a.idr, a.r_last, r, 0.0, a.z_last);
}

// M0: every tracer uses bias model 0 (no other model's code in the loop).
template <int NT, bool M0>
__global__ __launch_bounds__(kShellThreads) __attribute__((amdgpu_waves_per_eu(6))) void k_shells(ShellArgs a,
                                                                                      double *__restrict__ partial) {
  constexpr int NQ = 2 + NT;
  constexpr int W = shell_tab_w(NT);
  extern __shared__ double acc[];
  const int nz = a.nz, N = a.N;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  double *myacc = acc + (size_t)wave * nz * NQ;
  double *logv = acc + (size_t)kWaves * nz * NQ + (size_t)wave * kSlots * NQ * 64;   // [slot][q][lane]
  int *logid = (int *)(acc + (size_t)kWaves * nz * NQ + (size_t)kWaves * kSlots * NQ * 64) + wave * kSlots * 64;
  for (int i = threadIdx.x; i < kWaves * nz * NQ; i += blockDim.x) acc[i] = 0;
  __syncthreads();
  const float ox = (float)a.ox, oy = (float)a.oy, oz = (float)a.oz, dxf = (float)a.dx;
  const float s2f = (float)(a.sigma2_dev ? *a.sigma2_dev : a.sigma2);
  const int ntx = (N + kTX - 1) / kTX, nty = N / (kTY * kSub), ntz = N / kTZ;
  const long ntiles = (long)ntx * nty * ntz;
  const long nw = (long)gridDim.x * kWaves;
  // all lanes' logged shells into the wavefront's accumulators
  auto flush_all = [&](int nslot) {
#pragma unroll
    for (int k = 0; k < kSlots; k++)
      flush_slot<NQ>(__ballot(k < nslot), logid[k * 64 + lane], logv + (size_t)k * NQ * 64, myacc, lane);
  };
  for (long tile = (long)blockIdx.x * kWaves + wave; tile < ntiles; tile += nw) {
    const int tx = (int)(tile % ntx);
    const long rest = tile / ntx;
    const int ty = (int)(rest % nty), tz = (int)(rest / nty);
    const int ix = tx * kTX + lane % kTX, ly = lane / kTX;
    const int iy0 = ty * kTY * kSub + ly;   // the lane's y in the tile's first row
    const bool valid = ix < N;
    int nslot = 0, cur = -1;
    double cacc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) cacc[q] = 0;
    const float x32 = coord_f32(ix, dxf, ox);
    // The next row's density and table bins are loaded while the current row
    // is processed.  Bin lookup: FP32 radius (absolute error ~1e-3 cell), one
    // index for D, z and every b_t (f_of_r32 semantics).
    struct Bin {
      float4 C[W];
      float tq;
      bool e0, e1;
    };
    auto fetch = [&](int rr2, Bin &bn) {
      const int iy2 = iy0 + (rr2 % kTY) * kSub, iz2 = tz * kTZ + (rr2 / kTY);
      const float y32 = coord_f32(iy2, dxf, oy), z32 = coord_f32(iz2, dxf, oz);
      const TabPos tp = tab_pos_f32(radius_f32(x32, y32, z32), a.idr_f, a.na);
      bn.e0 = tp.e0;
      bn.e1 = tp.e1;
      bn.tq = tp.tq;
#pragma unroll
      for (int k = 0; k < W; k++) bn.C[k] = a.tab[(size_t)tp.i0 * W + k];
    };
    // the lane's cell in the tile's first row; rows at uniform offsets from it
    float *const tb = a.g + ((size_t)(tz * kTZ) * N + iy0) * a.ngx + (valid ? ix : 0);
    const auto roff = [&](int r) { return ((size_t)(r / kTY) * N + (r % kTY) * kSub) * a.ngx; };
    // densities kPrefetch rows ahead (the loads that wait on HBM), table bins
    // one row ahead (cache hits), issued bins first so that waiting for them
    // never waits for a later density
    float dq[kPrefetch];
#pragma unroll
    for (int j = 0; j < kPrefetch; j++) dq[j] = valid ? tb[roff(j)] : 0.f;
    Bin bnext;
    fetch(0, bnext);
    for (int r0 = 0; r0 < kTY * kTZ; r0 += kPrefetch)
#pragma unroll
    for (int j = 0; j < kPrefetch; j++) {
      const int rr = r0 + j;
      const int iy = iy0 + (rr % kTY) * kSub, iz = tz * kTZ + (rr / kTY);
      int s = -1;
      double v[NQ];
      float dl = dq[j];
      const Bin bn = bnext;
      if (rr + 1 < kTY * kTZ) fetch(rr + 1, bnext);
      if (rr + kPrefetch < kTY * kTZ && valid) dq[j] = tb[roff(rr + kPrefetch)];
      if (valid) {
        float comp[NQ];
        const bool e0 = bn.e0, e1 = bn.e1;
        {
          const float tq = bn.tq;
          float f[4 * W];
#pragma unroll
          for (int k = 0; k < W; k++) {
            f[4 * k + 0] = bn.C[k].x;
            f[4 * k + 1] = bn.C[k].y;
            f[4 * k + 2] = bn.C[k].z;
            f[4 * k + 3] = bn.C[k].w;
          }
#pragma unroll
          for (int q = 0; q < NQ; q++) comp[q] = f[2 * q] + f[2 * q + 1] * tq;
          if (e0 || e1) {
            comp[0] = e0 ? 1.f : a.d1_last_f;
            comp[1] = e0 ? 0.f : a.z_last_f;
#pragma unroll
            for (int t = 0; t < NT; t++) comp[2 + t] = e0 ? a.bz0[t] : 1.f;
          }
        }
        if (a.fused) {
          // the same float as lognormal_f32 of the fused count pass
          const TabPos tp{0, bn.tq, e0, e1};
          const float D = growth_at(tp, bn.C[0].x, bn.C[0].y, a.d1_last_f);
          dl = lognormal_f32(dl, D, s2f);
          tb[roff(rr)] = dl;
        }
        // shell index: FP32 z (error < 5e-5 shell) unless within 5e-4 shell of
        // an edge (or at the table ends), where the reference's FP64 radius
        // and z decide
        double zz = (double)comp[1];
        const float sf = comp[1] * a.idz_f;
        int ind = (int)sf + 1;
        const float fr = sf - floorf(sf);
        if (e0 || e1 || !(fr > 5e-4f && fr < 1.f - 5e-4f)) {
          zz = cell_z_f64(a, ix, iy, iz);
          ind = (int)(zz * a.idz) + 1;
        }
        if (ind >= 0 && ind < nz) {
          s = ind;
          v[0] = 1.0;
          v[1] = zz;
          const float l2 = __builtin_amdgcn_logf(1.f + dl);
          const float q = M0 ? 0.f : 1.44269504f * dl * __builtin_amdgcn_rcpf(1.f + dl);
#pragma unroll
          for (int t = 0; t < NT; t++)
            v[2 + t] = (double)(M0 ? (dl > -1.f ? __builtin_amdgcn_exp2f(comp[2 + t] * l2) : 0.f)
                                   : bias_sum_term(dl, l2, q, comp[2 + t], a.model[t]));
        }
      }
      const bool change = s >= 0 && s != cur;
      // a lane whose slots are full flushes them first (wave-collective, rare)
      if (__ballot(change && cur >= 0 && nslot == kSlots)) {
        flush_all(nslot);
        nslot = 0;
      }
      if (change) {
        if (cur >= 0) {
#pragma unroll
          for (int q = 0; q < NQ; q++) logv[((size_t)nslot * NQ + q) * 64 + lane] = cacc[q];
          logid[nslot * 64 + lane] = cur;
          nslot++;
        }
        cur = s;
#pragma unroll
        for (int q = 0; q < NQ; q++) cacc[q] = 0;
      }
      if (s >= 0) {
#pragma unroll
        for (int q = 0; q < NQ; q++) cacc[q] += v[q];
      }
    }
    // end of tile: the current shell goes into a slot (flushing first if full)
    if (__ballot(cur >= 0 && nslot == kSlots)) {
      flush_all(nslot);
      nslot = 0;
    }
    if (cur >= 0) {
#pragma unroll
      for (int q = 0; q < NQ; q++) logv[((size_t)nslot * NQ + q) * 64 + lane] = cacc[q];
      logid[nslot * 64 + lane] = cur;
      nslot++;
    }
    flush_all(nslot);
  }
  __syncthreads();
  double *out = partial + (size_t)blockIdx.x * nz * NQ;
  for (int i = threadIdx.x; i < nz * NQ; i += blockDim.x) {
    double v = 0;
    for (int w = 0; w < kWaves; w++) v += acc[(size_t)w * nz * NQ + i];
    out[i] = v;
  }
}

// ---------------------------------------------------------------------------
// k_shell_moments: the surrogate's read-only pass over the Gaussian grid.
// A work item is 256 consecutive x cells of one y row (4 per lane, one float4
// load) walked over zrun consecutive z planes: the wavefront reads 1 KB
// contiguous per plane (kMomPre planes in flight; 8 cells per lane, 2 KB,
// measured the same) and a lane's radius changes slowly,
// so it sums its current shell in FP32 registers and, when the shell
// changes, adds those sums to the wavefront's FP64 accumulators in LDS with
// atomics (lanes cross shell edges at different planes, so few collide).
// A lane's D(r) over its item is the quadratic through the FP64 table's
// values at its nearest, middle and farthest radius; the item's z length is
// chosen so that a lane spans at most ~160 Mpc/h, over which the quadratic
// is within ~1e-6 of the table (these are inputs of an approximation).  Per
// block the wavefronts' sums go to partial[block][nz][kMomQ], reduced in
// fixed order by k_shell_final.
// ---------------------------------------------------------------------------
constexpr int kMomQ = 6;            // count, G, G^2, G^3, G^4, D^2
constexpr int kMomBlocks = 4096;
constexpr int kMomMaxShells = 256;
constexpr int kMomZMax = 32;        // z planes per item, at most
constexpr int kMomPre = 8;          // planes loaded ahead (the item length is a multiple)
constexpr int kMomCPL = 4;          // x cells per lane (float4 loads)
constexpr int kMomXW = 64 * kMomCPL;  // x cells per item

struct MomArgs {
  const float *g;
  const double *r_arr, *d1;   // FP64 tables (f_of_r)
  const float *edge;          // nz FP32 edge radii
  int N, ngx, nz, zrun;
  double idr, r_last, d_last;
  float dxf, ox, oy, oz;
};

// largest s in [1, nz] with edge[s-1] <= r (s = nz: past the last shell)
__device__ inline int mom_shell(const float *edge, int nz, float r) {
  int lo = 1, hi = nz;
  while (hi > lo) {
    const int m = (lo + hi + 1) / 2;
    if (edge[m - 1] <= r) lo = m; else hi = m - 1;
  }
  return lo;
}

template <bool FULL>
__global__ __launch_bounds__(kShellThreads) void k_shell_moments(MomArgs a, double *__restrict__ partial) {
  extern __shared__ double macc[];   // [wave][nz][kMomQ] FP64, then nz FP32 edges
  const int nz = a.nz, N = a.N;
  float *ledge = (float *)(macc + (size_t)kWaves * nz * kMomQ);
  for (int i = threadIdx.x; i < kWaves * nz * kMomQ; i += blockDim.x) macc[i] = 0;
  for (int i = threadIdx.x; i < nz; i += blockDim.x) ledge[i] = a.edge[i];
  __syncthreads();
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  double *myacc = macc + (size_t)wave * nz * kMomQ;
  const int zrun = a.zrun;
  const int nxc = (N + kMomXW - 1) / kMomXW, nzc = (N + zrun - 1) / zrun;
  const long nwork = (long)nxc * N * nzc;
  const long nw = (long)gridDim.x * kWaves;
  const size_t plane = (size_t)N * a.ngx;
  for (long w = (long)blockIdx.x * kWaves + wave; w < nwork; w += nw) {
    const int xc = (int)(w % nxc);
    const long rest = w / nxc;
    const int iy = (int)(rest % N), zc = (int)(rest / N);
    const int ix = xc * kMomXW + lane * kMomCPL, iz0 = zc * zrun;   // the lane's cells: ix .. ix+kMomCPL-1
    const int nzi = FULL ? zrun : min(zrun, N - iz0);
    const bool valid = FULL || ix < N;   // N is a multiple of kMomCPL
    // D(r) on the lane's cells: quadratic through the table at its radius
    // range
    double c0, c1, c2;
    {
      const double xa = (valid ? ix : 0) * (double)a.dxf - a.ox, xb = xa + (kMomCPL - 1) * (double)a.dxf;
      const double yy = iy * (double)a.dxf - a.oy;
      const double za = iz0 * (double)a.dxf - a.oz, zb = (iz0 + nzi - 1) * (double)a.dxf - a.oz;
      const double dxn = xa > 0 ? xa : (xb < 0 ? -xb : 0), dzn = za > 0 ? za : (zb < 0 ? -zb : 0);
      const double r0 = sqrt(dxn * dxn + yy * yy + dzn * dzn);
      const double r2 = sqrt(fmax(xa * xa, xb * xb) + yy * yy + fmax(za * za, zb * zb));
      const double r1 = 0.5 * (r0 + r2);
      const double f0 = f_of_r(a.r_arr, a.d1, a.idr, a.r_last, r0, 1.0, a.d_last);
      const double f1 = f_of_r(a.r_arr, a.d1, a.idr, a.r_last, r1, 1.0, a.d_last);
      const double f2 = f_of_r(a.r_arr, a.d1, a.idr, a.r_last, r2, 1.0, a.d_last);
      const double h = r1 - r0;
      if (h > 1e-3) {
        // D(r) = f0 + (r - r0) (d01 + (r - r1) d012)
        const double d01 = (f1 - f0) / h, d012 = ((f2 - f1) / h - d01) / (2 * h);
        c2 = d012;
        c1 = d01 - d012 * (r0 + r1);
        c0 = f0 - r0 * d01 + d012 * r0 * r1;
      } else {
        c0 = f0; c1 = 0; c2 = 0;
      }
    }
    const float q0 = (float)c0, q1 = (float)c1, q2 = (float)c2;
    const float y = coord_f32(iy, a.dxf, a.oy);
    float rho2[kMomCPL];
#pragma unroll
    for (int k = 0; k < kMomCPL; k++) {
      const float x = coord_f32(valid ? ix + k : 0, a.dxf, a.ox);
      rho2[k] = __builtin_fmaf(x, x, y * y);
    }
    constexpr int V = kMomCPL / 4;   // float4 per lane and plane
    const float4 *col = (const float4 *)(a.g + ((size_t)iz0 * N + iy) * a.ngx + (valid ? ix : 0));
    const size_t plane4 = plane / 4;
    int key = -1;
    float e_lo = 0.f, e_hi = 0.f;
    float acc[kMomQ];
#pragma unroll
    for (int q = 0; q < kMomQ; q++) acc[q] = 0.f;
    float4 vq[kMomPre][V];
#pragma unroll
    for (int j = 0; j < kMomPre; j++)
#pragma unroll
      for (int v = 0; v < V; v++)
        vq[j][v] = (valid && (FULL || j < nzi)) ? col[(size_t)j * plane4 + v] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (int j0 = 0; j0 < nzi; j0 += kMomPre) {
#pragma unroll
      for (int u = 0; u < kMomPre; u++) {
        const int j = j0 + u;
        float dls[kMomCPL];
#pragma unroll
        for (int v = 0; v < V; v++) {
          dls[4 * v] = vq[u][v].x;
          dls[4 * v + 1] = vq[u][v].y;
          dls[4 * v + 2] = vq[u][v].z;
          dls[4 * v + 3] = vq[u][v].w;
        }
        if (valid && (FULL ? j0 + kMomPre < zrun : j + kMomPre < nzi)) {
#pragma unroll
          for (int v = 0; v < V; v++) vq[u][v] = col[(size_t)(j + kMomPre) * plane4 + v];
        }
        if (FULL || (j < nzi && valid)) {
          const float zc2 = coord_f32(iz0 + j, a.dxf, a.oz);
#pragma unroll
          for (int k = 0; k < kMomCPL; k++) {
            const float r = __builtin_amdgcn_sqrtf(__builtin_fmaf(zc2, zc2, rho2[k]));
            if (!(r >= e_lo && r < e_hi)) {
              if (key >= 1 && key < nz) {
#pragma unroll
                for (int q = 0; q < kMomQ; q++) atomicAdd(&myacc[key * kMomQ + q], (double)acc[q]);
              }
#pragma unroll
              for (int q = 0; q < kMomQ; q++) acc[q] = 0.f;
              key = mom_shell(ledge, nz, r);
              e_lo = ledge[key - 1];
              e_hi = key < nz ? ledge[key] : 3.0e38f;
            }
            if (key < nz) {
              const float D = __builtin_fmaf(r, __builtin_fmaf(r, q2, q1), q0);
              const float G = D * dls[k], G2 = G * G;
              acc[0] += 1.f;
              acc[1] += G;
              acc[2] += G2;
              acc[3] = __builtin_fmaf(G2, G, acc[3]);
              acc[4] = __builtin_fmaf(G2, G2, acc[4]);
              acc[5] = __builtin_fmaf(D, D, acc[5]);
            }
          }
        }
      }
    }
    if (key >= 1 && key < nz) {
#pragma unroll
      for (int q = 0; q < kMomQ; q++) atomicAdd(&myacc[key * kMomQ + q], (double)acc[q]);
    }
  }
  __syncthreads();
  double *out = partial + (size_t)blockIdx.x * nz * kMomQ;
  for (int i = threadIdx.x; i < nz * kMomQ; i += blockDim.x) {
    double t = 0;
    for (int w2 = 0; w2 < kWaves; w2++) t += macc[(size_t)w2 * nz * kMomQ + i];
    out[i] = t;
  }
}

// out[i] = sum over blocks of partial[b][i], one 256-thread block per value,
// fixed-order tree: deterministic.
__global__ __launch_bounds__(256) void k_shell_final(const double *__restrict__ partial, int nblocks, int nvals,
                                                     double *__restrict__ out) {
  __shared__ double s[256];
  const int i = blockIdx.x;
  double v = 0;
  for (int b = threadIdx.x; b < nblocks; b += 256) v += partial[(size_t)b * nvals + i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = s[0];
}

// Per tracer t: norm table on the NA r-grid.  sums = [nz][nq] from
// k_shell_final; work = scratch of 2*nz doubles.
constexpr int kNormMaxShells = 1024;
struct NormOuts {
  double *p[3 * CLR_NPOP_MAX];
  double *e[3 * CLR_NPOP_MAX];   // optional per-tracer copies of (norm0, normf)
};

// Every tracer's norm table in one launch: blockIdx.y is the tracer, the x
// blocks split the NA entries; each block rebuilds the tracer's (z, norm)
// nodes in LDS from the FP64 shell sums.
__global__ __launch_bounds__(256) void k_norm_table(const double *__restrict__ sums, int nz, int nq, double zmax,
                                                   const double *__restrict__ r_arr, const double *__restrict__ zt,
                                                   double idr, double r_last, double z_last, int na, NormOuts outs,
                                                   double *__restrict__ edges) {
#pragma clang fp contract(off)
  __shared__ double zarr[kNormMaxShells], nrm[kNormMaxShells];
  const int t = blockIdx.y;
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
    if (blockIdx.x == 0) {
      edges[2 * t] = nrm[0];
      edges[2 * t + 1] = nrm[nz - 1];
      if (outs.e[t]) {
        outs.e[t][0] = nrm[0];
        outs.e[t][1] = nrm[nz - 1];
      }
    }
  }
  __syncthreads();
  const double z0 = zarr[0], zf = zarr[nz - 1];
  const double n0 = nrm[0], nf = nrm[nz - 1];
  double *__restrict__ norm_out = outs.p[t];
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  if (ii >= na) return;
  {
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

template <int NT, bool M0>
void launch_shells(dim3 g, dim3 b, size_t lds, hipStream_t s, const ShellArgs &a, double *partial) {
  if (lds > 64 * 1024)
    (void)hipFuncSetAttribute((const void *)k_shells<NT, M0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((k_shells<NT, M0>), g, b, lds, s, a, partial);
}
using ShellFn = void (*)(dim3, dim3, size_t, hipStream_t, const ShellArgs &, double *);
#define CLR_SHELL_FNS(M0)                                                                                  \
  {launch_shells<0, M0>, launch_shells<1, M0>, launch_shells<2, M0>, launch_shells<3, M0>, launch_shells<4, M0>, \
   launch_shells<5, M0>, launch_shells<6, M0>, launch_shells<7, M0>, launch_shells<8, M0>}
const ShellFn kShellFns[2][kMaxTr + 1] = {CLR_SHELL_FNS(false), CLR_SHELL_FNS(true)};
#undef CLR_SHELL_FNS

}  // namespace

int clr_launch_density(clr_ctx *c, int dens_type, double sigma2) {
  if (!c->have_fields) return clr_fail(c, CLR_ERR_STATE, "clr_gaussian_fields must run first");
  if (dens_type != CLR_DENS_LOGNORMAL && dens_type != CLR_DENS_CLIP)
    return clr_fail(c, CLR_ERR_ARG, "Density type %d not supported", dens_type);
  const int N = c->N;
  const double *r = c->bg, *d1 = c->bg + 2 * c->na;
  CLR_HIP(c, hipEventRecord(c->ev0, c->stream));
  hipLaunchKernelGGL(k_transform, dim3((unsigned)((size_t)N * N)), dim3(N >= 256 ? 256 : N), 0, c->stream,
                     c->dens, N, c->ngx, c->dx, c->obs[0], c->obs[1], c->obs[2], r, d1, c->glob_idr,
                     c->r_last, c->h_bg_last[2], sigma2, dens_type);
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

// ---------------------------------------------------------------------------
// Static scratch of the no-allocation variant (clr_shell_norms_noalloc):
// the per-block FP64 partials of k_shells, the final sums, k_norm_table's
// edge values and the packed FP32 table, for nz (2 + ntr) <= kNoAllocVals.
// One copy per GPU for the life of the process; host code reaches it only
// through the address hipGetSymbolAddress returns on that device.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocVals = 320;
constexpr size_t kNoAllocTabFloat4 = (size_t)CLR_NA * shell_tab_w(kMaxTr);
constexpr size_t kNoAllocDoubles = (size_t)kShellBlocks * kNoAllocVals + kNoAllocVals + 2 * (size_t)kNormMaxShells +
                                   2 * (size_t)kMaxTr;
constexpr size_t kNoAllocBytes = (kNoAllocDoubles * sizeof(double) + 255) / 256 * 256 + kNoAllocTabFloat4 * 16;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

size_t clr_noalloc_scratch_bytes() { return kNoAllocBytes; }

int clr_noalloc_scratch_lookup(void **ptr) {
  return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch)) == hipSuccess ? 0 : -1;
}

int clr_launch_norms(clr_ctx *c, int lognormal_fused, double sigma2, int cap, int *nz_out,
                     unsigned long long *counts, double *zsum, double *sums, void *scratch_base) {
  if (!c->have_fields) return clr_fail(c, CLR_ERR_STATE, "clr_gaussian_fields must run first");
  if (lognormal_fused && c->density_done)
    return clr_fail(c, CLR_ERR_STATE, "density already transformed; call with lognormal_fused=0");
  if (!lognormal_fused && !c->density_done)
    return clr_fail(c, CLR_ERR_STATE, "clr_density must run first (or use lognormal_fused)");
  const int N = c->N, na = c->na;
  const double *r = c->bg, *zt = c->bg + na;
  const double z_last = c->h_bg_last[1];

  // shells: zmax = z(L/2), nz = (int)(zmax/DZ_SIGMA)+2, idz = (nz-2)/zmax
  const double zmax = c->zmax_half_box;
  const int nz = (int)(zmax / 0.05) + 2;
  if (nz > kNormMaxShells) return clr_fail(c, CLR_ERR_ARG, "%d redshift shells (at most %d)", nz, kNormMaxShells);
  const double idz = (nz - 2) / zmax;

  ShellArgs a{};
  ShellTabSrc tsrc{};
  int tk[3 * CLR_NPOP_MAX], tp[3 * CLR_NPOP_MAX];
  int ntr = 0;
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < c->n_tr[k]; i++) {
      if (!c->tr[k][i].set) return clr_fail(c, CLR_ERR_STATE, "tracer %d/%d has no tables", k, i);
      if (ntr >= kMaxTr) return clr_fail(c, CLR_ERR_ARG, "at most %d tracers per normalisation pass", kMaxTr);
      tsrc.bzf[ntr] = c->tr[k][i].bzf;
      a.bz0[ntr] = (float)c->tr[k][i].bz0;
      a.model[ntr] = c->tr[k][i].model;
      tk[ntr] = k;
      tp[ntr] = i;
      ntr++;
    }
  for (int t = 0; t < ntr; t++)
    if (!tsrc.bzf[t] || !c->tr[tk[t]][tp[t]].norm) return clr_fail(c, CLR_ERR_STATE, "tracer tables missing");
  if (!c->d1f || !c->bg) return clr_fail(c, CLR_ERR_STATE, "background tables missing");
  a.g = c->dens;
  a.N = N;
  a.ngx = c->ngx;
  a.nz = nz;
  a.ntr = ntr;
  a.fused = lognormal_fused ? 1 : 0;
  a.na = na;
  a.dx = c->dx;
  a.ox = c->obs[0];
  a.oy = c->obs[1];
  a.oz = c->obs[2];
  a.idr = c->glob_idr;
  a.r_last = c->r_last;
  a.z_last = z_last;
  a.sigma2 = sigma2;
  a.idz = idz;
  a.r_arr = r;
  a.zt = zt;
  tsrc.d1f = c->d1f;
  tsrc.zt = zt;
  a.z_last_f = (float)z_last;
  a.idz_f = (float)idz;
  a.d1_last_f = (float)c->h_bg_last[2];
  a.idr_f = (float)c->glob_idr;
  const int nq = 2 + ntr;
  static size_t (*const kLds[kMaxTr + 1])(int) = {shell_lds_bytes<0>, shell_lds_bytes<1>, shell_lds_bytes<2>,
                                                   shell_lds_bytes<3>, shell_lds_bytes<4>, shell_lds_bytes<5>,
                                                   shell_lds_bytes<6>, shell_lds_bytes<7>, shell_lds_bytes<8>};
  const size_t lds = kLds[ntr](nz);
  if (lds > 160 * 1024) return clr_fail(c, CLR_ERR_ARG, "too many shells x tracers (%d x %d)", nz, ntr);
  bool all0 = true;
  for (int t = 0; t < ntr; t++) all0 = all0 && a.model[t] == CLR_BIAS_EXPONENTIAL;
  const size_t nvals = (size_t)nz * nq;
  const int w = shell_tab_w(ntr);
  double *partial, *fin, *work;
  float4 *tab;
  if (scratch_base) {
    // no allocation: everything in the device's static scratch
    if (nvals > kNoAllocVals)
      return clr_fail(c, CLR_ERR_ARG, "%d shells x (2 + %d tracers) = %zu values exceed the %zu of the static scratch",
                      nz, ntr, nvals, kNoAllocVals);
    if ((size_t)na * w > kNoAllocTabFloat4)
      return clr_fail(c, CLR_ERR_ARG, "r-tables of %d entries exceed the static scratch", na);
    partial = (double *)scratch_base;
    fin = partial + (size_t)kShellBlocks * nvals;
    work = fin + nvals;
    tab = (float4 *)((char *)scratch_base + (kNoAllocDoubles * sizeof(double) + 255) / 256 * 256);
  } else {
    const size_t need = (size_t)kShellBlocks * nvals + nvals + 2 * nz + 2 * ntr;
    int rc = clr_ensure(c, (void **)&c->red, &c->red_n, need * sizeof(double));
    if (rc) return rc;
    partial = c->red;
    fin = c->red + (size_t)kShellBlocks * nvals;
    work = fin + nvals;
    rc = clr_ensure(c, &c->misc, &c->misc_n, (size_t)na * w * sizeof(float4));
    if (rc) return rc;
    tab = (float4 *)c->misc;
  }
  a.tab = tab;

  CLR_HIP(c, hipEventRecord(c->ev0, c->stream));
  hipLaunchKernelGGL(k_pack_shell_table, dim3((na + 255) / 256), dim3(256), 0, c->stream, tsrc, ntr, na, tab);
  CLR_HIP(c, hipGetLastError());
  CLR_HIP(c, hipEventRecord(c->ev2, c->stream));
  kShellFns[all0 ? 1 : 0][ntr](dim3(kShellBlocks), dim3(kShellThreads), lds, c->stream, a, partial);
  CLR_HIP(c, hipGetLastError());
  CLR_HIP(c, hipEventRecord(c->ev3, c->stream));
  hipLaunchKernelGGL(k_shell_final, dim3((unsigned)nvals), dim3(256), 0, c->stream, partial, kShellBlocks,
                     (int)nvals, fin);
  CLR_HIP(c, hipGetLastError());
  if (ntr > 0) {
    NormOuts outs{};
    for (int t = 0; t < ntr; t++) {
      outs.p[t] = c->tr[tk[t]][tp[t]].norm;
      outs.e[t] = c->tr[tk[t]][tp[t]].edge;
    }
    hipLaunchKernelGGL(k_norm_table, dim3((unsigned)((na + 255) / 256), (unsigned)ntr), dim3(256), 0, c->stream,
                       fin, nz, nq, zmax, r, zt, c->glob_idr, c->r_last, z_last, na, outs, work + 2 * nz);
    CLR_HIP(c, hipGetLastError());
  }
  CLR_HIP(c, hipEventRecord(c->ev1, c->stream));
  // host copies of the sums and edge values (a fixed array when they fit, so
  // that the no-allocation variant allocates nothing on the host either)
  double hfix[kNoAllocVals + 2 * kMaxTr];
  std::vector<double> hv;
  double *h = hfix;
  if (nvals > kNoAllocVals) {
    hv.resize(nvals + 2 * ntr);
    h = hv.data();
  }
  CLR_HIP(c, hipMemcpyAsync(h, fin, nvals * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CLR_HIP(c, hipMemcpyAsync(h + nvals, work + 2 * nz, 2 * ntr * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  for (int t = 0; t < ntr; t++) {
    TracerDev &T = c->tr[tk[t]][tp[t]];
    T.norm0 = h[nvals + 2 * t];
    T.normf = h[nvals + 2 * t + 1];
    T.norm_set = 1;
  }
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->stage_ms[ST_NORM] = ms;
  hipEventElapsedTime(&ms, c->ev2, c->ev3);
  c->stage_ms[ST_SHELLS] = ms;
  if (nz_out) *nz_out = nz;
  if ((counts || zsum || sums) && cap < nz)
    return clr_fail(c, CLR_ERR_ARG, "output capacity %d < nz=%d", cap, nz);
  for (int iz = 0; iz < nz; iz++) {
    if (counts) counts[iz] = (unsigned long long)h[(size_t)iz * nq + 0];
    if (zsum) zsum[iz] = h[(size_t)iz * nq + 1];
    if (sums)
      for (int t = 0; t < ntr; t++) sums[(size_t)t * nz + iz] = h[(size_t)iz * nq + 2 + t];
  }
  if (lognormal_fused) c->density_done = 1;
  c->norms_done = 1;
  return CLR_OK;
}

namespace {

// Arguments of a shell pass over the geometry only (no tracer columns): the
// shell binning of compute_density_normalization (density.c:1128-1211).
int shell_geo_args(clr_ctx *c, ShellArgs &a, ShellTabSrc &tsrc, int &nz) {
  if (!c->tables_set || !c->d1f) return clr_fail(c, CLR_ERR_STATE, "clr_upload_tables must be called first");
  const int na = c->na;
  const double zmax = c->zmax_half_box;
  nz = (int)(zmax / 0.05) + 2;
  if (nz > kNormMaxShells) return clr_fail(c, CLR_ERR_ARG, "%d redshift shells (at most %d)", nz, kNormMaxShells);
  a = ShellArgs{};
  tsrc = ShellTabSrc{};
  a.g = c->dens;
  a.N = c->N;
  a.ngx = c->ngx;
  a.nz = nz;
  a.ntr = 0;
  a.fused = 0;
  a.na = na;
  a.dx = c->dx;
  a.ox = c->obs[0];
  a.oy = c->obs[1];
  a.oz = c->obs[2];
  a.idr = c->glob_idr;
  a.r_last = c->r_last;
  a.z_last = c->h_bg_last[1];
  a.idz = (nz - 2) / zmax;
  a.r_arr = c->bg;
  a.zt = c->bg + na;
  tsrc.d1f = c->d1f;
  tsrc.zt = c->bg + na;
  a.z_last_f = (float)c->h_bg_last[1];
  a.idz_f = (float)a.idz;
  a.d1_last_f = (float)c->h_bg_last[2];
  a.idr_f = (float)c->glob_idr;
  return CLR_OK;
}

// One geometry-type shell pass with ntr table columns, reduced to
// fin[nz][2 + ntr] on the host.
template <int NT>
int shell_pass_to_host(clr_ctx *c, ShellArgs &a, const ShellTabSrc &tsrc, int nz, std::vector<double> &h) {
  const int na = c->na, nq = 2 + NT;
  const size_t nvals = (size_t)nz * nq;
  const int w = shell_tab_w(NT);
  int rc = clr_ensure(c, (void **)&c->red, &c->red_n, ((size_t)kShellBlocks * nvals + nvals) * sizeof(double));
  if (rc) return rc;
  double *partial = c->red, *fin = c->red + (size_t)kShellBlocks * nvals;
  rc = clr_ensure(c, &c->misc, &c->misc_n, (size_t)na * w * sizeof(float4));
  if (rc) return rc;
  a.tab = (const float4 *)c->misc;
  hipLaunchKernelGGL(k_pack_shell_table, dim3((na + 255) / 256), dim3(256), 0, c->stream, tsrc, NT, na,
                     (float4 *)c->misc);
  launch_shells<NT, true>(dim3(kShellBlocks), dim3(kShellThreads), shell_lds_bytes<NT>(nz), c->stream, a, partial);
  hipLaunchKernelGGL(k_shell_final, dim3((unsigned)nvals), dim3(256), 0, c->stream, partial, kShellBlocks,
                     (int)nvals, fin);
  CLR_HIP(c, hipGetLastError());
  h.resize(nvals);
  CLR_HIP(c, hipMemcpyAsync(h.data(), fin, nvals * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  return CLR_OK;
}

}  // namespace

// Cells and z sums per normalisation shell (the count and zsum columns of
// collect_density_normalization_from_grid, density.c:1128-1211): the shell
// pass without tracers; they depend only on the grid geometry and tables.
int clr_launch_shell_geometry(clr_ctx *c) {
  if (c->geo_nz > 0) return CLR_OK;
  ShellArgs a;
  ShellTabSrc tsrc;
  int nz = 0;
  int rc = shell_geo_args(c, a, tsrc, nz);
  if (rc) return rc;
  std::vector<double> h;
  rc = shell_pass_to_host<0>(c, a, tsrc, nz, h);
  if (rc) return rc;
  c->geo_cnt.resize(nz);
  c->geo_zsum.resize(nz);
  for (int iz = 0; iz < nz; iz++) {
    c->geo_cnt[iz] = h[(size_t)iz * 2];
    c->geo_zsum[iz] = h[(size_t)iz * 2 + 1];
  }
  c->geo_zmax = c->zmax_half_box;
  c->geo_nz = nz;
  return CLR_OK;
}

// Gaussian moments per normalisation shell of the current (untransformed)
// grid, for the surrogate: with G = D(r) delta_G per cell, out[nz][6] = the
// cell count and the sums of G, G^2, G^3, G^4 and D^2 (k_shell_moments).
// The shell of a cell is the one its FP32 radius falls in between the FP32
// radii of the z edges (nz shells of compute_density_normalization,
// density.c:1233-1244); D(r) from the FP32 growth table.
int clr_launch_shell_moments(clr_ctx *c, std::vector<double> &out) {
  if (!c->have_fields || c->density_done)
    return clr_fail(c, CLR_ERR_STATE, "the shell moments need the Gaussian field (fields done, density not transformed)");
  if (!c->tables_set || !c->d1f) return clr_fail(c, CLR_ERR_STATE, "clr_upload_tables must be called first");
  const int na = c->na;
  const double zmax = c->zmax_half_box;
  const int nz = (int)(zmax / 0.05) + 2;
  if (nz > kMomMaxShells) return clr_fail(c, CLR_ERR_ARG, "%d redshift shells (the moments pass takes %d)", nz, kMomMaxShells);
  const double idz = (nz - 2) / zmax;
  // E[k] = r(z = k / idz), k = 0 .. nz-1: shell s holds E[s-1] <= r < E[s]
  std::vector<float> hedge(nz);
  const std::vector<double> &zz = c->h_z, &rr = c->h_r;
  for (int k = 0; k < nz; k++) {
    const double z = k / idz;
    double r;
    if (z <= zz[0]) r = rr[0];
    else if (z >= zz[na - 1]) r = rr[na - 1] * 4;   // past the table: no cell reaches it
    else {
      int lo = 0, hi = na - 1;
      while (hi > lo + 1) {
        const int m = (lo + hi) / 2;
        if (zz[m] > z) hi = m; else lo = m;
      }
      r = rr[lo] + (z - zz[lo]) / (zz[hi] - zz[lo]) * (rr[hi] - rr[lo]);
    }
    hedge[k] = (float)r;
  }
  hedge[0] = 0.f;
  const size_t nvals = (size_t)nz * kMomQ;
  int rc = clr_ensure(c, (void **)&c->red, &c->red_n, ((size_t)kMomBlocks * nvals + nvals) * sizeof(double));
  if (rc) return rc;
  double *partial = c->red, *fin = c->red + (size_t)kMomBlocks * nvals;
  rc = clr_ensure(c, &c->misc, &c->misc_n, (size_t)nz * sizeof(float));
  if (rc) return rc;
  float *dedge = (float *)c->misc;
  CLR_HIP(c, hipMemcpyAsync(dedge, hedge.data(), nz * sizeof(float), hipMemcpyHostToDevice, c->stream));
  MomArgs a{};
  a.g = c->dens;
  a.r_arr = c->bg;
  a.d1 = c->bg + 2 * (size_t)na;
  a.edge = dedge;
  a.N = c->N;
  a.ngx = c->ngx;
  a.nz = nz;
  a.idr = c->glob_idr;
  a.r_last = c->r_last;
  a.d_last = c->h_bg_last[2];
  a.dxf = (float)c->dx;
  a.ox = (float)c->obs[0];
  a.oy = (float)c->obs[1];
  a.oz = (float)c->obs[2];
  // items of zrun planes: a multiple of kMomPre, at most kMomZMax, spanning
  // at most ~160 Mpc/h
  a.zrun = std::max(kMomPre, std::min(kMomZMax, (int)(160.0 / c->dx) / kMomPre * kMomPre));
  const size_t lds = (size_t)kWaves * nz * kMomQ * sizeof(double) + (size_t)nz * sizeof(float);
  if (c->N % kMomCPL) return clr_fail(c, CLR_ERR_ARG, "the shell moments need N divisible by %d", kMomCPL);
  if (c->N % kMomXW == 0 && c->N % a.zrun == 0)
    hipLaunchKernelGGL(k_shell_moments<true>, dim3(kMomBlocks), dim3(kShellThreads), lds, c->stream, a, partial);
  else
    hipLaunchKernelGGL(k_shell_moments<false>, dim3(kMomBlocks), dim3(kShellThreads), lds, c->stream, a, partial);
  hipLaunchKernelGGL(k_shell_final, dim3((unsigned)nvals), dim3(256), 0, c->stream, partial, kMomBlocks, (int)nvals,
                     fin);
  CLR_HIP(c, hipGetLastError());
  out.resize(nvals);
  CLR_HIP(c, hipMemcpyAsync(out.data(), fin, nvals * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  return CLR_OK;
}

// Norm tables of the given tracers from per-shell norms norm_s[t][nz] on the
// cached geometry: the sums are count/norm, and k_norm_table then applies the
// reference's edge fill and interpolation in z (density.c:1272-1355).
int clr_launch_norm_tables_given(clr_ctx *c, int ntr, const int *kinds, const int *ipops, const double *norm_s) {
  if (c->geo_nz <= 0) return clr_fail(c, CLR_ERR_STATE, "no shell geometry");
  if (ntr < 1 || ntr > 3 * CLR_NPOP_MAX) return clr_fail(c, CLR_ERR_ARG, "bad tracer count %d", ntr);
  const int nz = c->geo_nz, na = c->na, nq = 2 + ntr;
  std::vector<double> fin((size_t)nz * nq + 2 * (size_t)ntr, 0.0);
  for (int iz = 0; iz < nz; iz++) {
    const double n = c->geo_cnt[iz];
    fin[(size_t)iz * nq] = n;
    fin[(size_t)iz * nq + 1] = c->geo_zsum[iz];
    for (int t = 0; t < ntr; t++) {
      const double nm = norm_s[(size_t)t * nz + iz];
      fin[(size_t)iz * nq + 2 + t] = n > 0 ? n / nm : 0.0;
    }
  }
  const size_t bytes = fin.size() * sizeof(double);
  int rc = clr_ensure(c, (void **)&c->red, &c->red_n, bytes);
  if (rc) return rc;
  double *dfin = c->red, *edges = c->red + (size_t)nz * nq;
  CLR_HIP(c, hipMemcpyAsync(dfin, fin.data(), (size_t)nz * nq * sizeof(double), hipMemcpyHostToDevice, c->stream));
  NormOuts outs{};
  for (int t = 0; t < ntr; t++) {
    TracerDev &T = c->tr[kinds[t]][ipops[t]];
    if (!T.set || !T.norm) return clr_fail(c, CLR_ERR_STATE, "tracer %d/%d has no tables", kinds[t], ipops[t]);
    outs.p[t] = T.norm;
    outs.e[t] = T.edge;
  }
  hipLaunchKernelGGL(k_norm_table, dim3((unsigned)((na + 255) / 256), (unsigned)ntr), dim3(256), 0, c->stream,
                     dfin, nz, nq, c->geo_zmax, c->bg, c->bg + na, c->glob_idr, c->r_last, c->h_bg_last[1], na,
                     outs, edges);
  CLR_HIP(c, hipGetLastError());
  std::vector<double> e(2 * (size_t)ntr);
  CLR_HIP(c, hipMemcpyAsync(e.data(), edges, 2 * ntr * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  for (int t = 0; t < ntr; t++) {
    TracerDev &T = c->tr[kinds[t]][ipops[t]];
    T.norm0 = e[2 * t];
    T.normf = e[2 * t + 1];
    T.norm_set = 1;
  }
  return CLR_OK;
}

// ---------------------------------------------------------------------------
// The shell-sum stage for graph capture: buffers first (clr_prepare_norms),
// then kernel launches only (clr_enqueue_norms): the packed table, the shell
// pass with the fused lognormal (sigma^2 read from the device), the final
// sums and every tracer's norm table, its edge values left in each tracer's
// device copy.  The same kernels and order as clr_launch_norms.
// ---------------------------------------------------------------------------
static int norms_tracers(clr_ctx *c, int *tk, int *tp) {
  int ntr = 0;
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < c->n_tr[k]; i++) {
      tk[ntr] = k;
      tp[ntr] = i;
      ntr++;
    }
  return ntr;
}

int clr_prepare_norms(clr_ctx *c) {
  int tk[3 * CLR_NPOP_MAX], tp[3 * CLR_NPOP_MAX];
  const int ntr = norms_tracers(c, tk, tp);
  if (ntr < 1 || ntr > kMaxTr) return clr_fail(c, CLR_ERR_ARG, "%d tracers (1 .. %d)", ntr, kMaxTr);
  const int nz = (int)(c->zmax_half_box / 0.05) + 2;
  const size_t nvals = (size_t)nz * (2 + ntr);
  int rc = clr_ensure(c, (void **)&c->red, &c->red_n,
                      ((size_t)kShellBlocks * nvals + nvals + 2 * nz + 2 * ntr) * sizeof(double));
  if (rc) return rc;
  rc = clr_ensure(c, &c->misc, &c->misc_n, (size_t)c->na * shell_tab_w(ntr) * sizeof(float4));
  if (rc) return rc;
  for (int t = 0; t < ntr; t++) {
    TracerDev &T = c->tr[tk[t]][tp[t]];
    if (!T.edge) CLR_HIP(c, hipMalloc((void **)&T.edge, 2 * sizeof(double)));
  }
  return CLR_OK;
}

int clr_enqueue_norms(clr_ctx *c, hipStream_t st, const double *sigma2_dev) {
  int tk[3 * CLR_NPOP_MAX], tp[3 * CLR_NPOP_MAX];
  const int ntr = norms_tracers(c, tk, tp);
  const int N = c->N, na = c->na;
  const double zmax = c->zmax_half_box;
  const int nz = (int)(zmax / 0.05) + 2;
  const int nq = 2 + ntr;
  const size_t nvals = (size_t)nz * nq;
  if (ntr < 1 || ntr > kMaxTr || c->red_n < ((size_t)kShellBlocks * nvals + nvals + 2 * nz + 2 * ntr) * sizeof(double))
    return clr_fail(c, CLR_ERR_STATE, "normalisation buffers not prepared");
  ShellArgs a{};
  ShellTabSrc tsrc{};
  bool all0 = true;
  NormOuts outs{};
  for (int t = 0; t < ntr; t++) {
    TracerDev &T = c->tr[tk[t]][tp[t]];
    if (!T.set || !T.bzf || !T.norm || !T.edge) return clr_fail(c, CLR_ERR_STATE, "tracer tables missing");
    tsrc.bzf[t] = T.bzf;
    a.bz0[t] = (float)T.bz0;
    a.model[t] = T.model;
    all0 = all0 && T.model == CLR_BIAS_EXPONENTIAL;
    outs.p[t] = T.norm;
    outs.e[t] = T.edge;
  }
  a.g = c->dens;
  a.N = N;
  a.ngx = c->ngx;
  a.nz = nz;
  a.ntr = ntr;
  a.fused = 1;
  a.na = na;
  a.dx = c->dx;
  a.ox = c->obs[0];
  a.oy = c->obs[1];
  a.oz = c->obs[2];
  a.idr = c->glob_idr;
  a.r_last = c->r_last;
  a.z_last = c->h_bg_last[1];
  a.sigma2 = 0;
  a.sigma2_dev = sigma2_dev;
  a.idz = (nz - 2) / zmax;
  a.r_arr = c->bg;
  a.zt = c->bg + na;
  tsrc.d1f = c->d1f;
  tsrc.zt = c->bg + na;
  a.z_last_f = (float)c->h_bg_last[1];
  a.idz_f = (float)a.idz;
  a.d1_last_f = (float)c->h_bg_last[2];
  a.idr_f = (float)c->glob_idr;
  a.tab = (const float4 *)c->misc;
  static size_t (*const kLds[kMaxTr + 1])(int) = {shell_lds_bytes<0>, shell_lds_bytes<1>, shell_lds_bytes<2>,
                                                   shell_lds_bytes<3>, shell_lds_bytes<4>, shell_lds_bytes<5>,
                                                   shell_lds_bytes<6>, shell_lds_bytes<7>, shell_lds_bytes<8>};
  double *partial = c->red, *fin = c->red + (size_t)kShellBlocks * nvals, *work = fin + nvals;
  hipLaunchKernelGGL(k_pack_shell_table, dim3((na + 255) / 256), dim3(256), 0, st, tsrc, ntr, na, (float4 *)c->misc);
  kShellFns[all0 ? 1 : 0][ntr](dim3(kShellBlocks), dim3(kShellThreads), kLds[ntr](nz), st, a, partial);
  hipLaunchKernelGGL(k_shell_final, dim3((unsigned)nvals), dim3(256), 0, st, partial, kShellBlocks, (int)nvals, fin);
  hipLaunchKernelGGL(k_norm_table, dim3((unsigned)((na + 255) / 256), (unsigned)ntr), dim3(256), 0, st, fin, nz, nq,
                     zmax, c->bg, c->bg + na, c->glob_idr, c->r_last, c->h_bg_last[1], na, outs, work + 2 * nz);
  CLR_HIP(c, hipGetLastError());
  return CLR_OK;
}

