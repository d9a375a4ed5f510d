// clr_sources.hip -- srcs_set_cartesian_single (srcs.c:120-283) for one or
// all populations, and srcs_get_local_properties_single (srcs.c:386-416).
//
// Blocks are one wavefront over 512 consecutive cells (z-major flat order; a
// lane owns 8 cells of one row); k_count's wavefronts take two such blocks.
// k_count : per lane and population, one upper bound on lambda over the
//           lane's 8 cells (largest delta, coarse bounds of the bins their
//           radii can fall in) settles k = 0 wherever the 16 high uniform
//           bits lie below exp(-bound); the few open cells are compacted
//           (ballot) into an LDS queue that all lanes then work: the cell's
//           own FP32 lambda with a bounded error (clr_poisson.h), the full
//           32-bit uniform and the FP32 search, and where that is too close
//           to call the reference's FP64 expressions (k_exact, count_exact)
//           -- every count equals the FP64 one.  Writes one byte per cell
//           and population (255 = "255 or more", the count is in the escape
//           list) and the block's total.
// k_scan_*: exclusive prefix of the block totals (64-bit), per population.
// k_fill  : reads the bytes, scans its cells and writes one descriptor
//           (cell, index within the cell) per source at its final place in
//           the catalogue: global z-major cell order, draw order within a
//           cell -- the order of the reference on one thread.
// k_write : one thread per source turns its descriptor into the position
//           in [x-dx/2, x+dx/2) (srcs.c:261-263), the RSD shift from central
//           differences of phi (get_rvel, srcs.c:87-118) and the NEST base
//           pixel (vec2pix_ring then a RING->NEST table, srcs.c:264-270).
#include "clr_internal.h"
#include "clr_poisson.h"
#include <math.h>
#include <limits.h>
#include <algorithm>
#include <vector>

using namespace clr;

namespace {

constexpr int kLanes = 64, kCellsPerLane = 8, kBlockCells = kLanes * kCellsPerLane;
constexpr int kByteEscape = 255;    // stored for counts >= kBytePending: the count is in the escape list
constexpr int kBytePending = 254;   // k_count left the cell to k_exact (never seen by k_fill)
// escape-list count of a cell whose Poisson mean has no count (see kPoisMax)
constexpr unsigned long long kCountInvalid = ~0ull;

struct PopTab {
  const double *amp, *bz, *norm;
  const float4 *tab;         // k_pop_table's bins, bound and coarse bounds
  double bz0, norm0, normf;
  int model, ipop;
  uint8_t *cnt8;             // one byte per cell
  uint32_t *btot;            // one total per block
  const long long *boff;     // exclusive prefix of btot
  uint4 *desc;               // the catalogue's position buffer, holding descriptors
  const double *edges;       // device (norm0, normf) when set (graph-captured sampling)
  long long cap;             // catalogue capacity (descriptors beyond it are dropped)
};

struct SrcArgs {
  const float *g;
  const double *r_arr;
  double idr, r_last;
  double dx, cell_vol, half_box_20, ox, oy, oz;
  int N, ngx, na, npop;
  uint32_t seed;
  // exact counts of cells whose byte saturated: key = pop << 56 | cell,
  // value = count; appended by k_count, sorted by key on the host for k_fill
  ulonglong2 *esc;
  unsigned int *esc_n;
  unsigned int esc_cap, esc_count;
  // cells whose count needs the FP64 path: (cell lo, cell hi, pass population,
  // 32-bit uniform), appended by k_count and counted by k_exact
  uint4 *exq;
  unsigned int *exq_n;
  unsigned int exq_cap;
  PopTab pop[CLR_NPOP_MAX];
  // surrogate-fused sampling: the grid holds the Gaussian field and the
  // lognormal transform is applied per cell as the shell pass would
  int fused;
  const float *d1f;
  float s2f, d1_last_f, idr_f, dxf, oxf, oyf, ozf;
};

// The density of cell (ix,iy,iz) that the counts use: the stored value, or
// with a.fused its FP32 lognormal transform (lognormal_f32), the same float
// the shell pass would have stored.
__device__ inline float cell_density(const SrcArgs &a, float g, int ix, int iy, int iz) {
  if (!a.fused) return g;
  const TabPos p = tab_pos_f32(radius_f32(coord_f32(ix, a.dxf, a.oxf), coord_f32(iy, a.dxf, a.oyf),
                                          coord_f32(iz, a.dxf, a.ozf)), a.idr_f, a.na);
  return lognormal_f32(g, growth_f32(a.d1f, p, a.d1_last_f), a.s2f);
}

// Per bin i (the interval [r_i, r_i+1] of the FP64 tables) one float4 that
// the sampler gathers once per cell:
//   Pos   (x >= 0): n, norm > 0 and finite at both ends, P = n dx^3 norm in
//         [1e-30, 1e30] at both ends within a factor 2, |b| < 1e6:
//         x = P(r_i), y = P(r_i+1) - P(r_i), z = b(r_i), w = b(r_i+1) - b(r_i).
//         The sampler interpolates P linearly; the reference interpolates n
//         and norm separately, whose product differs by at most
//         |dn dnorm| dx^3 / 4, a relative bound kept (maximum over bins) in
//         the x of the extra element na.
//   Zero  (x = -1): n is NaN at an end, or n <= 0 at both ends (the FP64
//         interpolation is then <= 0 or NaN): no sources.
//   Tiny  (x = -2): n >= 0 and norm > 0 at both ends with P <= 1e-20: the
//         reference's lambda is at most 1e-20 bias, far below 2^-33, where
//         exp(-lambda) exceeds every uniform: no sources unless the bias is
//         huge (z, w as for Pos).
//   Other (x = -3): the FP64 path decides.
constexpr float kBinZero = -1.f, kBinTiny = -2.f, kBinOther = -3.f;

// After the bins and the bound (element na) come the coarse bounds, written
// by k_pop_coarse: entry j covers the bins [16j, 16j+16), i.e. the nodes
// 16j..16j+16, as (Pmax, bmin, bmax, 0):
//   Pmax >= n(r) dx^3 norm(r) for the reference's separate FP64 interpolants
//   of n and norm anywhere in those bins (the largest n times the largest
//   norm at the nodes, rounded up; NaN nodes of n give no sources there),
//   bmin, bmax: the extreme b at the nodes.
//   (+inf, -inf, +inf): no bound -- a node with n = +-inf, norm < 0 or not
//   finite, b < 0, |b| >= 1e6 or NaN, or Pmax > 1e30.
constexpr int kCoarse = 16;
inline int coarse_count(int na) { return (na + kCoarse - 1) / kCoarse; }
inline size_t pop_table_len(int na) { return (size_t)na + 1 + coarse_count(na); }

__global__ void k_pop_table(const double *__restrict__ amp, const double *__restrict__ bz,
                            const double *__restrict__ norm, double cv, int na, float4 *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na) return;
  const int j = i + 1 < na ? i + 1 : i;
  const double f0 = amp[i], f1 = amp[j], n0 = norm[i], n1 = norm[j], b0 = bz[i], b1 = bz[j];
  const bool bfin = fabs(b0) < 1e6 && fabs(b1) < 1e6;
  const bool npos = n0 > 0 && n1 > 0 && n0 < 1e300 && n1 < 1e300;
  const float bz0 = (float)b0, dbz = (float)(b1 - b0);
  if (isnan(f0) || isnan(f1) || (f0 <= 0 && f1 <= 0)) {
    out[i] = make_float4(kBinZero, 0.f, 0.f, 0.f);
    return;
  }
  if (f0 >= 0 && f1 >= 0 && npos && bfin) {
    const double P0 = f0 * cv * n0, P1 = f1 * cv * n1;
    if (P0 <= 1e-20 && P1 <= 1e-20) {
      out[i] = make_float4(kBinTiny, 0.f, bz0, dbz);
      return;
    }
    if (P0 >= 1e-30 && P1 >= 1e-30 && P0 <= 1e30 && P1 <= 1e30 && P0 <= 2 * P1 && P1 <= 2 * P0) {
      const double rel = fabs((f1 - f0) * (n1 - n0)) * cv / 4 / fmin(P0, P1);
      atomicMax((unsigned int *)&out[na].x, __float_as_uint((float)rel * 1.0001f));
      out[i] = make_float4((float)P0, (float)(P1 - P0), bz0, dbz);
      return;
    }
  }
  out[i] = make_float4(kBinOther, 0.f, 0.f, 0.f);
}

__global__ void k_pop_coarse(const double *__restrict__ amp, const double *__restrict__ bz,
                             const double *__restrict__ norm, double cv, int na, float4 *__restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= (na + kCoarse - 1) / kCoarse) return;
  const int i0 = j * kCoarse, i1 = min(i0 + kCoarse, na - 1);
  double nmax = 0, mmax = 0, bmin = 1e300, bmax = 0;
  bool ok = true;
  for (int i = i0; i <= i1; i++) {
    const double n = amp[i], m = norm[i], b = bz[i];
    if (isinf(n)) ok = false;
    else if (n > nmax) nmax = n;
    if (m >= 0 && m < 1e300) mmax = fmax(mmax, m);
    else ok = false;
    if (b >= 0 && b < 1e6) {
      bmin = fmin(bmin, b);
      bmax = fmax(bmax, b);
    } else {
      ok = false;
    }
  }
  const double P = nmax * cv * mmax;
  out[na + 1 + j] = ok && P <= 1e30 ? make_float4((float)(P * (1 + 1e-6)), (float)bmin, (float)bmax, 0.f)
                                    : make_float4(INFINITY, -INFINITY, INFINITY, 0.f);
}

// The reference's FP64 count of cell (ix,iy,iz) (srcs.c:136-187) from the
// full 32-bit uniform x.
__device__ __noinline__ int count_exact(int ix, int iy, int iz, float dl, uint64_t cell, uint32_t x,
                                        const double *r_arr, const double *amp, const double *bz,
                                        const double *norm, double idr, double r_last, double bz0, double norm0,
                                        double normf, double dx, double ox, double oy, double oz, double cell_vol,
                                        double half_box_20, int model, int ipop, uint32_t seed) {
#pragma clang fp contract(off)
  const double obs[3] = {ox, oy, oz};
  double x0, y0, z0;
  const double r = cell_radius(ix, iy, iz, dx, obs, &x0, &y0, &z0);
  if (!(r < half_box_20)) return 0;
  const double nd = f_of_r(r_arr, amp, idr, r_last, r, 0.0, 0.0);
  if (!(nd > 0)) return 0;
  const double bias = f_of_r(r_arr, bz, idr, r_last, r, bz0, 1.0);
  const double dnorm = f_of_r(r_arr, norm, idr, r_last, r, norm0, normf);
  const double lam = nd * cell_vol * bias_model_f64((double)dl, bias, model) * dnorm;
  return poisson_exact(lam, cell, stream_poisson_extra(ipop), seed, x);
}

// FP32 bias_model (common.h:414-431) for a bias model known at compile time,
// branch-free, with an absolute bound E on its error against the FP64 value
// at the same (delta, b): b carries the interpolation error of the table
// (bmax = larger |b| at the bin ends), the raw hardware exp2/log2 (about 1
// ulp; results below FP32's normal range flush to 0, which only lowers lambda
// where k = 0 is certain anyway) and __fdividef theirs.
template <int MODEL>
__device__ inline float bias_fast_t(float b, float bmax, float dl, float &E) {
  float v, e;
  if (MODEL == 0) {
    const float l2 = __builtin_amdgcn_logf(1.f + dl);
    v = __builtin_amdgcn_exp2f(b * l2);
    e = v * (2e-6f + 1e-6f * bmax + 2e-6f * bmax * fabsf(l2) * 0.6931472f);
  } else if (MODEL == 1) {
    const float qd = __fdividef(dl, 1.f + dl);
    const float ve = __expf(b * qd);
    const bool neg = dl < 0.f;
    v = neg ? ve : 1.f + b * dl;
    e = neg ? ve * (2e-6f + 1e-6f * bmax + 2e-6f * bmax * fabsf(qd)) : 3e-7f + 1e-6f * bmax * fabsf(dl);
  } else {
    v = fmaxf(1.f + b * dl, 0.f);
    e = 3e-7f + 1e-6f * bmax * fabsf(dl);
  }
  const bool dead = !(dl > -1.f);
  E = dead ? 0.f : e;
  return dead ? 0.f : v;
}

template <int MODEL>
__device__ inline float bias_upper(float dmax, float blo, float bhi) {
  float E1, E2;
  const float v1 = bias_fast_t<MODEL>(blo, bhi, dmax, E1);
  const float v2 = bias_fast_t<MODEL>(bhi, bhi, dmax, E2);
  return fmaxf(v1 + E1, v2 + E2) * (1.f + 1e-6f);
}

__device__ inline float bias_fast_rt(int model, float b, float bmax, float dl, float &E) {
  if (model == 0) return bias_fast_t<0>(b, bmax, dl, E);
  if (model == 1) return bias_fast_t<1>(b, bmax, dl, E);
  return bias_fast_t<2>(b, bmax, dl, E);
}

// Appends a cell to the FP64 queue of k_exact with its full uniform.
__device__ inline void to_exact(const SrcArgs &a, size_t cell, int ip, uint32_t x) {
  const unsigned j = atomicAdd(a.exq_n, 1u);
  if (j < a.exq_cap) a.exq[j] = make_uint4((uint32_t)cell, (uint32_t)(cell >> 32), (uint32_t)ip, x);
}

// Count of one cell that the lane bound left open (srcs.c:136-187): the
// cell's own geometry and bin, then
//   outside r < L/2+20, or an empty bin: 0;
//   a Pos bin: the FP32 lambda and its error bound decide k = 0 from the 16
//   high bits, else the FP32 search with the full uniform decides;
//   a Tiny bin: 0 while the bias stays below 1e9;
// and everything else -- the radius within 1e-9 of a bin or box edge, the
// last bins, Other bins, a search too close to call -- goes to k_exact (-1).
__device__ inline int settle_cell(const SrcArgs &a, const float4 *tab, int model, int pipop, int ip, size_t cell,
                                  int ix, int iy, int iz, float dl, uint32_t hi) {
  const double xx = ix * a.dx - a.ox, y0 = iy * a.dx - a.oy, z0 = iz * a.dx - a.oz;
  const double r2 = xx * xx + (y0 * y0 + z0 * z0);
  // r = r2 / sqrt(r2): FP32 rsqrt refined once in FP64 (rel. error ~2e-14)
  const float y = __builtin_amdgcn_rsqf((float)r2);
  const double yd = (double)y * (1.5 - 0.5 * r2 * (double)y * (double)y);
  const double r = r2 * yd;
  const double xr = r * a.idr;
  const int i = (int)xr;
  const double fr = xr - i;
  const double hb = a.half_box_20;
  const bool fast = r2 > 0 && fr > 1e-9 && fr < 1 - 1e-9 && i < a.na - 2 && fabs(r - hb) > 1e-9 * hb;
  if (fast && !(r < hb)) return 0;
  if (fast) {
    const float4 T = tab[i];
    const float t = (float)fr;
    if (T.x == kBinZero) return 0;
    if (T.x == kBinTiny) {
      float E;
      const float bias = bias_fast_rt(model, T.z + T.w * t, fmaxf(fabsf(T.z), fabsf(T.z + T.w)), dl, E);
      if (bias + E < 1e9f) return 0;
    } else if (T.x >= 0.f) {
      float E;
      const float bias = bias_fast_rt(model, T.z + T.w * t, fmaxf(fabsf(T.z), fabsf(T.z + T.w)), dl, E);
      const float nn = T.x + T.y * t;
      const float l = nn * bias;
      const float err = l * (2e-6f + tab[a.na].x) + nn * E;
      const float ex = __expf(-l);
      if (poisson_zero_from_hi(l, ex, err, hi)) return 0;
      const uint32_t x = (hi << 16) | poisson_lo16(cell, pipop, a.seed);
      const int k = poisson_fast_f32(l, ex, err, x);
      if (k >= 0) return k;
      to_exact(a, cell, ip, x);
      return -1;
    }
  }
  to_exact(a, cell, ip, (hi << 16) | poisson_lo16(cell, pipop, a.seed));
  return -1;
}

// k_count works on 1024 cells per wavefront (16 per lane, two of the 512-cell
// blocks whose totals k_scan and k_fill use), so that its fixed costs per
// wavefront -- the density loads' latency, the coarse-bound gathers, the queue
// pass -- are shared by twice the cells.
constexpr int kCountLaneCells = 16;
constexpr int kCountCells = kLanes * kCountLaneCells;
constexpr int kCountSubs = kCountCells / kBlockCells;
// queue entry: cell within the wavefront's 1024 (bits 0-9) | population
// (bits 10-13) | the uniform's high 16 bits << 16.  The queue is worked when
// another population's 1024 entries might not fit, so usually once.
constexpr int kQueueCap = kCountCells + kCountCells / 4;

template <bool FUSED>
__global__ __launch_bounds__(kLanes) __attribute__((amdgpu_waves_per_eu(6))) void k_count(SrcArgs a) {
  __shared__ uint32_t queue[kQueueCap];
  extern __shared__ uint4 sbytes[];   // [npop][kLanes]: one byte per cell and population
  __shared__ uint2 sgeo[kLanes];      // per lane: first column | row y << 16, row z
  __shared__ uint32_t stot[CLR_NPOP_MAX][kCountSubs];
  __shared__ const float4 *stab[CLR_NPOP_MAX];
  __shared__ int smodel[CLR_NPOP_MAX];
  const int lane = threadIdx.x;
  const size_t blk = blockIdx.x;
  const size_t base = blk * kCountCells;
  const size_t cell0 = base + (size_t)kCountLaneCells * lane;
  const int N = a.N;
  // row and column of the lane's first cell: the wavefront part in scalar
  // 64-bit arithmetic, the lane part in 32 bits (rows < 2^22)
  const size_t row0 = base / (size_t)N;
  const int obase = (int)(base - row0 * N);
  const int o = obase + kCountLaneCells * lane;
  const int drow = o / N;
  const int ix0 = o - drow * N;
  const unsigned row = (unsigned)row0 + (unsigned)drow;
  const int iy = (int)(row % (unsigned)N), iz = (int)(row / (unsigned)N);
  const unsigned long long below = (1ull << lane) - 1ull;
  sgeo[lane] = make_uint2((uint32_t)ix0 | ((uint32_t)iy << 16), (uint32_t)iz);
  for (int ip = 0; ip < a.npop; ip++) {
    sbytes[ip * kLanes + lane] = make_uint4(0u, 0u, 0u, 0u);
    if (lane < kCountSubs) stot[ip][lane] = 0;
    if (lane == 0) {
      stab[ip] = a.pop[ip].tab;
      smodel[ip] = a.pop[ip].model;
    }
  }

  // --- the lane's 16 cells: largest delta and the range of their radii ------
  // Every count of the lane is 0 where u < exp(-lambda_max) for an upper
  // bound lambda_max of the reference's lambda over the 16 cells:
  //   lambda = n(r) dx^3 norm(r) bias_model(delta, b(r)) <= Pmax * bias_max
  // with Pmax and the range of b from the coarse bounds of the bins that the
  // radii can fall in, and bias_max the larger corner (delta_max, b_min or
  // b_max): with b >= 0 every bias model is non-decreasing in delta and
  // monotone in b.  Cells outside r < L/2+20 have lambda = 0 and need no
  // special case.  The radii come from the two end cells (the nearest point
  // of the row when it lies between them), in FP32 with a 1e-5 margin.
  float dmax;
  {
    const float4 *p = (const float4 *)(a.g + (size_t)row * a.ngx + ix0);   // rows start on 128 bytes
    float4 v[kCountLaneCells / 4];
#pragma unroll
    for (int h = 0; h < kCountLaneCells / 4; h++) v[h] = p[h];
    dmax = fmaxf(fmaxf(v[0].x, v[0].y), fmaxf(v[0].z, v[0].w));
#pragma unroll
    for (int h = 1; h < kCountLaneCells / 4; h++)
      dmax = fmaxf(dmax, fmaxf(fmaxf(v[h].x, v[h].y), fmaxf(v[h].z, v[h].w)));
  }
  bool out;
  int ilo, ihi;
  {
    const double y0 = iy * a.dx - a.oy, z0 = iz * a.dx - a.oz;
    const double yz2 = y0 * y0 + z0 * z0;
    const double xa = ix0 * a.dx - a.ox, xb = xa + (kCountLaneCells - 1) * a.dx;
    const double ra2 = xa * xa + yz2, rb2 = xb * xb + yz2;
    const double lo2 = (xa <= 0 && xb >= 0) ? yz2 : fmin(ra2, rb2);
    const float rlo = sqrtf((float)lo2) * (1.f - 1e-5f);
    const float rhi = sqrtf((float)fmax(ra2, rb2)) * (1.f + 1e-5f);
    const float hb = (float)a.half_box_20 * (1.f + 1e-5f);
    const float idr = (float)a.idr;
    out = rlo > hb;
    ilo = (int)(rlo * idr * (1.f - 1e-5f));
    ihi = (int)(fminf(rhi, hb) * idr * (1.f + 1e-5f));
    if constexpr (FUSED) {
      // dmax holds the largest Gaussian value g; every cell's density is
      // exp(D (g_c - D s2/2)) - 1 with D in [D(rhi), D(rlo)] (D decreases with
      // r), and D g - D^2 s2/2 is concave in D: its largest value on that
      // range, widened for the FP32 roundings of the cell's own evaluation.
      const TabPos plo = tab_pos_f32(rlo, a.idr_f, a.na), phi = tab_pos_f32(rhi, a.idr_f, a.na);
      const float dhi = growth_f32(a.d1f, plo, a.d1_last_f) * (1.f + 1e-6f);
      const float dlo = growth_f32(a.d1f, phi, a.d1_last_f) * (1.f - 1e-6f);
      const float g = dmax, s2 = a.s2f;
      const float dstar = s2 > 0.f ? fminf(fmaxf(g / s2, dlo), dhi) : (g > 0.f ? dhi : dlo);
      const float fm = dstar * (g - 0.5f * dstar * s2);
      dmax = expf(fm + 1e-5f * fabsf(fm) + 1e-6f) * (1.f + 1e-5f) - 1.f;
    }
  }

  int qn = 0;   // uniform
  for (int ip = 0; ip < a.npop; ip++) {
    const PopTab &P = a.pop[ip];
    const int model = P.model;
    // --- A: the lane bound, then the cells it leaves open to the queue ------
    int hcrit = 65536;   // hi < hcrit: u < (hi+1)/2^16 <= exp(-lambda_max), k = 0
    if (!out) {
      hcrit = 0;
      if (ihi < a.na - 2) {
        const float4 *coarse = P.tab + a.na + 1;
        float pm = 0.f, blo = 3.4e38f, bhi = 0.f;
        for (int j = ilo / kCoarse; j <= ihi / kCoarse; j++) {
          const float4 C = coarse[j];
          pm = fmaxf(pm, C.x);
          blo = fminf(blo, C.y);
          bhi = fmaxf(bhi, C.z);
        }
        if (pm == 0.f) {
          hcrit = 65536;
        } else if (pm < 1e30f && blo >= 0.f) {
          float bub;
          if (model == 0) bub = bias_upper<0>(dmax, blo, bhi);
          else if (model == 1) bub = bias_upper<1>(dmax, blo, bhi);
          else bub = bias_upper<2>(dmax, blo, bhi);
          const float lub = pm * bub * (1.f + 1e-5f);
          // __expf: relative error below 1e-6 + 2e-7 lub
          if (lub < 16.f) hcrit = (int)(65536.f * __expf(-lub) * (1.f - 2e-6f - 4e-7f * lub));
        }
      }
    }
    if (__ballot(hcrit < 65536) != 0ull) {
#pragma unroll
      for (int g = 0; g < kCountLaneCells / 8; g++) {
        // the high 16 bits of 8 cells come from one draw at site cell/8
        const u32x4 rr = draw((cell0 >> 3) + g, stream_poisson(P.ipop), 0u, a.seed);
#pragma unroll
        for (int c = 0; c < 8; c++) {
          const uint32_t hi = poisson_hi16(rr, (uint64_t)c);
          const bool open = (int)hi >= hcrit;
          const unsigned long long m = __ballot(open);
          if (m != 0ull) {
            if (open)
              queue[qn + __popcll(m & below)] =
                  (uint32_t)(kCountLaneCells * lane + 8 * g + c) | ((uint32_t)ip << 10) | (hi << 16);
            qn += __popcll(m);
          }
        }
      }
    }
    // --- B: the queue, all lanes busy; run when another population might
    // not fit, and after the last -------------------------------------------
    if (qn > kQueueCap - kCountCells || ip == a.npop - 1) {
      __syncthreads();   // one wavefront: orders the LDS writes above before the reads below
      for (int i = lane; i < qn; i += kLanes) {
        const uint32_t e = queue[i];
        const int cb = (int)(e & (kCountCells - 1));
        const int q = (int)((e >> 10) & 15u);
        const uint32_t hi = e >> 16;
        const size_t cell = base + cb;
        const uint2 gl = sgeo[cb / kCountLaneCells];
        const int ix = (int)(gl.x & 0xFFFFu) + cb % kCountLaneCells, iy = (int)(gl.x >> 16), iz = (int)gl.y;
        const float dl = cell_density(a, a.g[((size_t)iz * N + iy) * a.ngx + ix], ix, iy, iz);
        const int k = settle_cell(a, stab[q], smodel[q], a.pop[0].ipop + q, q, cell, ix, iy, iz, dl, hi);
        ((uint8_t *)(sbytes + q * kLanes))[cb] =
            (uint8_t)(k < 0 ? kBytePending : (k < kBytePending ? k : kByteEscape));
        if (k > 0) atomicAdd(&stot[q][cb / kBlockCells], (uint32_t)k);
        if (k >= kBytePending) {
          // the exact count of a saturated byte goes to the escape list
          const unsigned x = atomicAdd(a.esc_n, 1u);
          if (x < a.esc_cap)
            a.esc[x] = make_ulonglong2(((unsigned long long)q << 56) | (unsigned long long)cell, (unsigned long long)k);
        }
      }
      __syncthreads();
      qn = 0;
    }
  }
  // --- the bytes of the lane's 16 cells and the two block totals -------------
  for (int ip = 0; ip < a.npop; ip++) {
    const PopTab &P = a.pop[ip];
    *(uint4 *)(P.cnt8 + cell0) = sbytes[ip * kLanes + lane];
    if (lane < kCountSubs) P.btot[blk * kCountSubs + lane] = stot[ip][lane];
  }
}

// The FP64 counts of the cells k_count left pending: block totals, bytes and
// escape entries completed in place (one thread per entry).
__global__ __launch_bounds__(256) void k_exact(SrcArgs a) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned n = *a.exq_n;
  if (j >= n || j >= a.exq_cap) return;
  const uint4 e = a.exq[j];
  const uint64_t cell = (uint64_t)e.x | ((uint64_t)e.y << 32);
  const int ip = (int)e.z;
  const PopTab &P = a.pop[ip];
  const int N = a.N;
  const uint64_t rw = cell / (uint64_t)N;
  const int cx = (int)(cell - rw * N);
  const double n0 = P.edges ? P.edges[0] : P.norm0, nf = P.edges ? P.edges[1] : P.normf;
  const int k = count_exact(cx, (int)(rw % N), (int)(rw / N),
                            cell_density(a, a.g[rw * (size_t)a.ngx + cx], cx, (int)(rw % N), (int)(rw / N)), cell,
                            e.w, a.r_arr,
                            P.amp, P.bz, P.norm, a.idr, a.r_last, P.bz0, n0, nf, a.dx, a.ox, a.oy, a.oz,
                            a.cell_vol, a.half_box_20, P.model, P.ipop, a.seed);
  if (k < 0) {
    // no count for this mean: an escape entry the host turns into an error
    P.cnt8[cell] = kByteEscape;
    const unsigned q = atomicAdd(a.esc_n, 1u);
    if (q < a.esc_cap)
      a.esc[q] = make_ulonglong2(((unsigned long long)ip << 56) | (unsigned long long)cell, kCountInvalid);
    return;
  }
  if (k > 0) atomicAdd(&P.btot[cell / kBlockCells], (uint32_t)k);
  P.cnt8[cell] = (uint8_t)(k < kBytePending ? k : kByteEscape);
  if (k >= kBytePending) {
    const unsigned q = atomicAdd(a.esc_n, 1u);
    if (q < a.esc_cap)
      a.esc[q] = make_ulonglong2(((unsigned long long)ip << 56) | (unsigned long long)cell, (unsigned long long)k);
  }
}

// exclusive scan of n uint32 -> int64 offsets (out[n] = total), for every
// population at once (blockIdx.y).  Three kernels: per-block sums, scan of
// the block sums (one block per population), block-local scan.
constexpr int kScanB = 1024;
struct ScanPtrs {
  const uint32_t *in[CLR_NPOP_MAX];
  long long *bsum[CLR_NPOP_MAX];
  long long *out[CLR_NPOP_MAX];
};
__global__ __launch_bounds__(kScanB) void k_scan_blocks(ScanPtrs p, size_t n) {
  __shared__ long long s[kScanB];
  const uint32_t *__restrict__ in = p.in[blockIdx.y];
  const size_t i = (size_t)blockIdx.x * kScanB + threadIdx.x;
  s[threadIdx.x] = i < n ? in[i] : 0;
  __syncthreads();
  for (int w = kScanB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) p.bsum[blockIdx.y][blockIdx.x] = s[0];
}
__global__ __launch_bounds__(kScanB) void k_scan_top(ScanPtrs p, size_t nb) {
  __shared__ long long s[kScanB];
  __shared__ long long carry;
  long long *__restrict__ bsum = p.bsum[blockIdx.y];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (size_t base = 0; base < nb; base += kScanB) {
    const size_t i = base + threadIdx.x;
    const long long v = i < nb ? bsum[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanB; off <<= 1) {
      const long long u = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += u;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == kScanB - 1) carry += s[kScanB - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}
__global__ __launch_bounds__(kScanB) void k_scan_apply(ScanPtrs p, size_t n) {
  __shared__ long long s[kScanB];
  const uint32_t *__restrict__ in = p.in[blockIdx.y];
  const long long *__restrict__ bsum = p.bsum[blockIdx.y];
  long long *__restrict__ out = p.out[blockIdx.y];
  const size_t i = (size_t)blockIdx.x * kScanB + threadIdx.x;
  const long long v = i < n ? in[i] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kScanB; off <<= 1) {
    const long long u = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += u;
    __syncthreads();
  }
  if (i < n) out[i] = bsum[blockIdx.x] + s[threadIdx.x] - v;
  if (i == n - 1) out[n] = bsum[gridDim.x];
}

// count of an escaped cell from the sorted escape list
__device__ inline int escaped_count(const ulonglong2 *esc, unsigned n, unsigned long long key) {
  unsigned lo = 0, hi = n;
  while (lo < hi) {
    const unsigned mid = (lo + hi) / 2;
    if (esc[mid].x < key) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && esc[lo].x == key) ? (int)esc[lo].y : 0;
}

// One wavefront per kCountCells cells (two 512-cell blocks of the scan): a lane
// reads the 16 count bytes of its cells per population.
__global__ __launch_bounds__(kLanes) void k_fill(SrcArgs a) {
  constexpr int G = 2;   // populations whose loads are in flight together
  const int lane = threadIdx.x;
  const size_t w0 = (size_t)blockIdx.x * kCountSubs;   // first 512-cell block
  const uint64_t cell0 = (uint64_t)blockIdx.x * kCountCells + (uint64_t)kCountLaneCells * lane;
  for (int g0 = 0; g0 < a.npop; g0 += G) {
    long long b0[G], b1[G];
    uint4 w[G];
#pragma unroll
    for (int q = 0; q < G; q++) {
      b0[q] = b1[q] = 0;
      w[q] = make_uint4(0u, 0u, 0u, 0u);
      if (g0 + q < a.npop) {
        const PopTab &P = a.pop[g0 + q];
        b0[q] = P.boff[w0];
        b1[q] = P.boff[w0 + kCountSubs];
        w[q] = *(const uint4 *)(P.cnt8 + cell0);
      }
    }
#pragma unroll
    for (int q = 0; q < G; q++) {
      if (g0 + q >= a.npop || b1[q] == b0[q]) continue;   // uniform: no sources in these cells
      const uint32_t wd[4] = {w[q].x, w[q].y, w[q].z, w[q].w};
      int k[kCountLaneCells];
      long long mine = 0;
#pragma unroll
      for (int c = 0; c < kCountLaneCells; c++) {
        k[c] = (int)((wd[c / 4] >> (8 * (c & 3))) & 0xFFu);
        if (k[c] == kByteEscape)
          k[c] = escaped_count(a.esc, a.esc_count, ((unsigned long long)(g0 + q) << 56) | (cell0 + c));
        mine += k[c];
      }
      long long incl = mine;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const long long y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
      }
      long long pid = b0[q] + incl - mine;
      uint4 *desc = a.pop[g0 + q].desc;
#pragma unroll
      for (int c = 0; c < kCountLaneCells; c++) {
        const uint64_t cell = cell0 + c;
        const long long lim = b1[q] < a.pop[g0 + q].cap ? b1[q] : a.pop[g0 + q].cap;
        for (int is = 0; is < k[c]; is++, pid++)
          if (pid < lim) desc[pid] = make_uint4((uint32_t)cell, (uint32_t)(cell >> 32), (uint32_t)is, 0u);
      }
    }
  }
}

struct WriteArgs {
  const float *phi;
  const double *r_arr, *v1;
  double idr, r_last, v1_0, v1_last;
  double dx, factor_vel, idx, ox, oy, oz;
  int N, ngx, nside_base, ipop;
  uint32_t seed;
  const int32_t *r2n;
  float4 *pos;
  int32_t *ipix;
  long long n;
  const long long *n_dev;   // when set: the source count on the device (n is the capacity)
};

__device__ inline double get_rvel(const WriteArgs &a, int ix, int iy, int iz, double x0, double y0, double z0,
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
  // FP32 differences of the stored potential, as get_rvel takes them
  const double v0 = phi[ixh + y0o + z0o] - phi[ixl + y0o + z0o];
  const double v1 = phi[ix + (size_t)iyh * ngx + z0o] - phi[ix + (size_t)iyl * ngx + z0o];
  const double v2 = phi[ix + y0o + (size_t)izh * N * ngx] - phi[ix + y0o + (size_t)izl * N * ngx];
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

__global__ __launch_bounds__(256) void k_write(WriteArgs a) {
#pragma clang fp contract(off)
  const long long pid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pid >= a.n) return;
  if (a.n_dev && pid >= *a.n_dev) return;
  const uint4 ds = ((const uint4 *)a.pos)[pid];
  const int N = a.N;
  uint64_t cell = (uint64_t)ds.x | ((uint64_t)ds.y << 32);
  if (cell >= (uint64_t)N * N * N) cell = 0;   // never: every slot holds a descriptor
  const int is = (int)ds.z;
  const uint64_t rw = cell / (uint64_t)N;
  const int ix = (int)(cell - rw * N), iy = (int)(rw % N), iz = (int)(rw / N);
  const double x0 = (ix + 0.0) * a.dx - a.ox, y0 = (iy + 0.0) * a.dx - a.oy, z0 = (iz + 0.0) * a.dx - a.oz;
  const double rr = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
  const double rvel = a.factor_vel * get_rvel(a, ix, iy, iz, x0, y0, z0, rr);
  const double dz_rsd = rvel * f_of_r(a.r_arr, a.v1, a.idr, a.r_last, rr, a.v1_0, a.v1_last);
  const uint32_t stream = stream_position(a.ipop);
  const double ctr[3] = {x0, y0, z0};
  float pp[3];
  u32x4 rnd{};
  uint32_t last = 0xFFFFFFFFu;
#pragma unroll
  for (int ax = 0; ax < 3; ax++) {
    const uint32_t j = 3u * (uint32_t)is + (uint32_t)ax;
    if ((j >> 2) != last) {
      rnd = draw(cell, stream, j >> 2, a.seed);
      last = j >> 2;
    }
    pp[ax] = (float)(ctr[ax] + a.dx * (u01(word(rnd, (int)(j & 3u))) - 0.5));
  }
  a.pos[pid] = make_float4(pp[0], pp[1], pp[2], (float)dz_rsd);
  a.ipix[pid] = a.r2n[vec2pix_ring32(a.nside_base, (double)pp[0], (double)pp[1], (double)pp[2])];
}

// Zero n 32-bit words (a kernel rather than a memset node: graph-safe).
__global__ void k_zero_words(uint32_t *__restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Status of a graph-captured sampling pass: 1 saturated counts (escape list
// not sorted in graph mode), 2 FP64 queue overflow, 4 a catalogue beyond
// its capacity.
struct GraphTotals {
  const long long *tot[CLR_NPOP_MAX];
  long long cap[CLR_NPOP_MAX];
};
__global__ void k_graph_flags(const unsigned *__restrict__ esc_n, const unsigned *__restrict__ exq_n,
                              unsigned exq_cap, int npop, GraphTotals g, int *__restrict__ flag) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int f = 0;
  if (*esc_n > 0) f |= 1;
  if (*exq_n > exq_cap) f |= 2;
  for (int i = 0; i < npop; i++)
    if (*g.tot[i] > g.cap[i]) f |= 4;
  *flag |= f;
}

// Per-cell counts (the reference's nsources array, srcs.c:125, without the
// row padding) of cells [cell0, cell0+n) of population slot q: the count
// byte, or the escape-list entry of a saturated byte.
__global__ void k_expand_counts(const uint8_t *__restrict__ cnt8, const ulonglong2 *__restrict__ esc,
                                unsigned nesc, unsigned long long q, uint64_t cell0, uint64_t n,
                                int32_t *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t cell = cell0 + i;
  int k = cnt8[cell];
  if (k == kByteEscape) k = escaped_count(esc, nesc, (q << 56) | cell);
  out[i] = k;
}

// ---------------------------------------------------------------------------
// Reference-stream mode: srcs_set_cartesian_single (srcs.c:120-283) fed the
// reference's own per-thread uniform streams.  Thread t of the reference owns
// the planes izb[t] .. izb[t+1]-1 (schedule(static)) and draws from its
// MT19937 stream (seed + t + nthr (ipop + n_srcs IThread0), srcs.c:150);
// the caller passes the streams' raw 32-bit outputs.
//   k_ref_lambda: per cell the reference's FP64 lambda, or -1 where it draws
//                 nothing (r >= L/2+20 or n(r) <= 0);
//   k_ref_walk  : one lane per reference thread walks its cells in z, y, x
//                 order through gsl_ran_poisson as GSL documents it (below),
//                 consuming its stream exactly as the reference does;
//   k_ref_write : pass 2 re-seeds (srcs.c:231): the k-th source of thread t
//                 takes uniforms 3k, 3k+1, 3k+2 of stream t for x, y, z.
// gsl_ran_poisson: while mu > 10, m = (unsigned)(7 mu/8), X = gamma_int(m)
// (a < 12: -log of a product of uniform_pos; else Knuth's tan-based
// rejection); X >= mu returns k + binomial(mu/X, m-1), else k += m,
// mu -= X; then the multiplication method.  binomial: p > 1/2 mirrored,
// inversion by sequential search (at most 111 terms per uniform) from
// f(0) = q^n by binary powering.  GSL's BTPE branch (n p >= 14) is not
// restated: inversion there too.
// ---------------------------------------------------------------------------
struct RefStream {
  const uint32_t *s;
  long long len, pos;
  int overflow;
};

__device__ inline double rs_uniform(RefStream &r) {
  if (r.pos >= r.len) { r.overflow = 1; return 0.5; }
  return u01(r.s[r.pos++]);
}
__device__ inline double rs_uniform_pos(RefStream &r) {
  double x;
  do { x = rs_uniform(r); } while (x == 0 && !r.overflow);
  return x;
}
__device__ inline double pow_uint(double x, unsigned n) {
#pragma clang fp contract(off)
  double v = 1.0;
  do {
    if (n & 1u) v *= x;
    n >>= 1;
    x *= x;
  } while (n);
  return v;
}
__device__ double ref_gamma_int(RefStream &r, unsigned a) {
#pragma clang fp contract(off)
  if (a < 12) {
    double prod = 1;
    for (unsigned i = 0; i < a; i++) prod *= rs_uniform_pos(r);
    return -log(prod);
  }
  const double sqa = sqrt(2.0 * a - 1), am1 = (double)a - 1;
  double x, y, v;
  do {
    do {
      y = tan(M_PI * rs_uniform(r));
      x = sqa * y + am1;
    } while (x <= 0 && !r.overflow);
    v = rs_uniform(r);
  } while (v > (1 + y * y) * exp(am1 * log(x / am1) - sqa * y) && !r.overflow);
  return x;
}
__device__ unsigned ref_binomial(RefStream &r, double p, unsigned n) {
#pragma clang fp contract(off)
  if (n == 0) return 0;
  bool flipped = false;
  if (p > 0.5) { p = 1.0 - p; flipped = true; }
  const double q = 1 - p, s = p / q;
  const double f0 = pow_uint(q, n);
  unsigned ix = 0;
  for (;;) {
    double f = f0, u = rs_uniform(r);
    bool done = false;
    for (ix = 0; ix <= 110; ix++) {
      if (u < f) { done = true; break; }
      u -= f;
      f *= s * (double)(n - ix) / (double)(ix + 1);
    }
    if (done || r.overflow) break;
  }
  return flipped ? n - ix : ix;
}
__device__ int ref_poisson(RefStream &r, double mu) {
#pragma clang fp contract(off)
  unsigned k = 0;
  while (mu > 10 && !r.overflow) {
    const unsigned m = (unsigned)(mu * (7.0 / 8.0));
    const double X = ref_gamma_int(r, m);
    if (X >= mu) return (int)(k + ref_binomial(r, mu / X, m - 1));
    k += m;
    mu -= X;
  }
  const double emu = exp(-mu);
  double prod = 1.0;
  do {
    prod *= rs_uniform(r);
    k++;
  } while (prod > emu && !r.overflow);
  return (int)(k - 1);
}

struct RefArgs {
  const float *g;
  const double *r_arr, *amp, *bz, *norm;
  double idr, r_last, bz0, norm0, normf;
  double dx, cell_vol, half_box_20, ox, oy, oz;
  int N, ngx, model, nthr;
  const int *izb;                 // nthr + 1 plane bounds
  const uint32_t *streams;        // nthr x len
  long long len;
  double *lam;                    // N^3: lambda, or -1 (no draw)
  uint32_t *cnt;                  // N^3 counts
  int *overflow;                  // set when a stream runs out
};

__global__ void k_ref_lambda(RefArgs a) {
#pragma clang fp contract(off)
  const uint64_t cell = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.N;
  if (cell >= (uint64_t)N * N * N) return;
  const uint64_t rw = cell / (uint64_t)N;
  const int ix = (int)(cell - rw * N), iy = (int)(rw % N), iz = (int)(rw / N);
  const double obs[3] = {a.ox, a.oy, a.oz};
  double x0, y0, z0;
  const double r = cell_radius(ix, iy, iz, a.dx, obs, &x0, &y0, &z0);
  double lam = -1.0;
  if (r < a.half_box_20) {
    const double nd = f_of_r(a.r_arr, a.amp, a.idr, a.r_last, r, 0.0, 0.0);
    if (nd > 0) {
      const double bias = f_of_r(a.r_arr, a.bz, a.idr, a.r_last, r, a.bz0, 1.0);
      const double dnorm = f_of_r(a.r_arr, a.norm, a.idr, a.r_last, r, a.norm0, a.normf);
      const double dl = a.g[(size_t)ix + (size_t)a.ngx * ((size_t)iy + (size_t)N * iz)];
      lam = nd * a.cell_vol * bias_model_f64(dl, bias, a.model) * dnorm;
    }
  }
  a.lam[cell] = lam;
}

// One lane per reference thread; the lanes of a wavefront walk different
// threads' planes side by side.
__global__ __launch_bounds__(64) void k_ref_walk(RefArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nthr) return;
  const size_t plane = (size_t)a.N * a.N;
  RefStream rs{a.streams + (size_t)t * a.len, a.len, 0, 0};
  const size_t c0 = (size_t)a.izb[t] * plane, c1 = (size_t)a.izb[t + 1] * plane;
  for (size_t c = c0; c < c1; c++) {
    const double lam = a.lam[c];
    a.cnt[c] = lam == -1.0 ? 0u : (uint32_t)ref_poisson(rs, lam);
  }
  if (rs.overflow) atomicOr(a.overflow, 1);
}

struct RefWriteExtra {
  const uint32_t *cnt;
  const long long *off;           // exclusive prefix of cnt (N^3 + 1)
  const int *izb;
  const uint32_t *streams;
  long long len;
  int nthr;
  int *overflow;
};

__global__ __launch_bounds__(256) void k_ref_write(WriteArgs a, RefWriteExtra e) {
#pragma clang fp contract(off)
  const uint64_t cell = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.N;
  if (cell >= (uint64_t)N * N * N) return;
  const int k = (int)e.cnt[cell];
  if (k == 0) return;
  const uint64_t rw = cell / (uint64_t)N;
  const int ix = (int)(cell - rw * N), iy = (int)(rw % N), iz = (int)(rw / N);
  int t = 0;
  while (t + 1 < e.nthr && iz >= e.izb[t + 1]) t++;
  const long long first = e.off[(size_t)e.izb[t] * N * N];   // thread t's first source
  const uint32_t *st = e.streams + (size_t)t * e.len;
  const double x0 = (ix + 0.0) * a.dx - a.ox, y0 = (iy + 0.0) * a.dx - a.oy, z0 = (iz + 0.0) * a.dx - a.oz;
  const double rr = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
  const double rvel = a.factor_vel * get_rvel(a, ix, iy, iz, x0, y0, z0, rr);
  const double dz_rsd = rvel * f_of_r(a.r_arr, a.v1, a.idr, a.r_last, rr, a.v1_0, a.v1_last);
  const double ctr[3] = {x0, y0, z0};
  for (int is = 0; is < k; is++) {
    const long long pid = e.off[cell] + is;
    const long long j = 3 * (pid - first);
    if (j + 2 >= e.len) { atomicOr(e.overflow, 2); return; }
    float pp[3];
    for (int ax = 0; ax < 3; ax++) pp[ax] = (float)(ctr[ax] + a.dx * (u01(st[j + ax]) - 0.5));
    a.pos[pid] = make_float4(pp[0], pp[1], pp[2], (float)dz_rsd);
    a.ipix[pid] = a.r2n[vec2pix_ring32(a.nside_base, (double)pp[0], (double)pp[1], (double)pp[2])];
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

int clr_launch_sources(clr_ctx *c, int ipop, uint32_t seed, int nside_base, long long *nsrc, int fused) {
  const int nall = c->n_tr[CLR_TRACER_SRCS];
  if (ipop < -1 || ipop >= nall) return clr_fail(c, CLR_ERR_ARG, "no source population %d", ipop);
  if (!fused && !c->density_done) return clr_fail(c, CLR_ERR_STATE, "density field not computed");
  if (fused && (!c->have_fields || c->density_done))
    return clr_fail(c, CLR_ERR_STATE, "fused sampling needs the Gaussian field (fields done, density not transformed)");
  if (fused && !c->d1f) return clr_fail(c, CLR_ERR_STATE, "background tables missing");
  if (nside_base < 1 || nside_base > 64 || (nside_base & (nside_base - 1)))
    return clr_fail(c, CLR_ERR_ARG, "nside_base=%d must be a power of two <= 64", nside_base);
  const int p0 = ipop < 0 ? 0 : ipop, p1 = ipop < 0 ? nall : ipop + 1;
  const int N = c->N, na = c->na;
  const size_t ncells = (size_t)N * N * N;
  if (N % kCountLaneCells || ncells % kCountCells)
    return clr_fail(c, CLR_ERR_ARG, "grid of %d^3 cells is not a whole number of %d-cell blocks", N, kCountCells);
  for (int i = p0; i < p1; i++) {
    const TracerDev &T = c->tr[CLR_TRACER_SRCS][i];
    if (!T.set || !T.norm_set) return clr_fail(c, CLR_ERR_STATE, "population %d has no norm table", i);
  }
  int rc = clr_ring2nest_table(c, nside_base);
  if (rc) return rc;
  const double dx = c->L / N, cell_vol = dx * dx * dx;
  // packed FP32 tables, rebuilt from the current norm tables
  for (int i = p0; i < p1; i++) {
    TracerDev &T = c->tr[CLR_TRACER_SRCS][i];
    if (!T.ptab) CLR_HIP(c, hipMalloc((void **)&T.ptab, pop_table_len(na) * sizeof(float4)));
    CLR_HIP(c, hipMemsetAsync((float4 *)T.ptab + na, 0, sizeof(float4), c->stream));
    hipLaunchKernelGGL(k_pop_table, dim3((na + 255) / 256), dim3(256), 0, c->stream, T.amp, T.bz, T.norm, cell_vol,
                       na, (float4 *)T.ptab);
    hipLaunchKernelGGL(k_pop_coarse, dim3((coarse_count(na) + 255) / 256), dim3(256), 0, c->stream, T.amp, T.bz,
                       T.norm, cell_vol, na, (float4 *)T.ptab);
    CLR_HIP(c, hipGetLastError());
  }
  const int np = p1 - p0;
  const size_t nblocks = ncells / kBlockCells;
  const size_t nsb = (nblocks + kScanB - 1) / kScanB;
  // scratch per population: bytes (ncells), block totals, offsets, scan partials
  // (a multiple of 256 bytes: every population's bytes start aligned)
  const size_t per_pop = (ncells + nblocks * 4 + (nblocks + 1) * 8 + (nsb + 1) * 8 + 64 + 255) / 256 * 256;
  rc = clr_ensure(c, (void **)&c->rowcnt, &c->rows_n, (size_t)np * per_pop);
  if (rc) return rc;

  SrcArgs a{};
  a.g = c->dens;
  a.r_arr = c->bg;
  a.idr = c->glob_idr;
  a.r_last = c->r_last;
  a.dx = dx;
  a.cell_vol = cell_vol;
  a.half_box_20 = c->L / 2 + 20.;
  a.ox = c->obs[0];
  a.oy = c->obs[1];
  a.oz = c->obs[2];
  a.N = N;
  a.ngx = c->ngx;
  a.na = na;
  a.npop = np;
  a.seed = seed;
  a.fused = fused ? 1 : 0;
  a.d1f = c->d1f;
  a.s2f = (float)c->sigma2;
  a.d1_last_f = (float)c->h_bg_last[2];
  a.idr_f = (float)c->glob_idr;
  a.dxf = (float)dx;
  a.oxf = (float)c->obs[0];
  a.oyf = (float)c->obs[1];
  a.ozf = (float)c->obs[2];
  std::vector<long long *> bsum(np);
  for (int i = p0; i < p1; i++) {
    const TracerDev &T = c->tr[CLR_TRACER_SRCS][i];
    PopTab &P = a.pop[i - p0];
    uint8_t *base = (uint8_t *)c->rowcnt + (size_t)(i - p0) * per_pop;
    P.cnt8 = base;
    P.btot = (uint32_t *)(base + ncells);
    long long *off = (long long *)(((uintptr_t)(P.btot + nblocks) + 7) & ~(uintptr_t)7);
    P.boff = off;
    bsum[i - p0] = off + nblocks + 1;
    P.amp = T.amp;
    P.bz = T.bz;
    P.norm = T.norm;
    P.tab = (const float4 *)T.ptab;
    P.bz0 = T.bz0;
    P.norm0 = T.norm0;
    P.normf = T.normf;
    P.model = T.model;
    P.ipop = i;
    P.edges = nullptr;
    P.cap = LLONG_MAX;
    if (!P.amp || !P.bz || !P.norm || !P.tab) return clr_fail(c, CLR_ERR_STATE, "population %d tables missing", i);
  }
  if (!c->r2n) return clr_fail(c, CLR_ERR_STATE, "no RING->NEST table");
  if (!a.g || !a.r_arr) return clr_fail(c, CLR_ERR_STATE, "density grid or background tables missing");

  CLR_HIP(c, hipEventRecord(c->ev0, c->stream));
  std::vector<long long> tot(np);
  unsigned int nesc = 0;
  for (int attempt = 0; attempt < 2; attempt++) {
    // escape list: a counter, then entries (grown and recounted if it overflows)
    const size_t want = c->esc_bytes > 0 ? c->esc_bytes : (size_t)16 + 16 * 65536;
    rc = clr_ensure(c, &c->esc, &c->esc_bytes, want);
    if (rc) return rc;
    a.esc_n = (unsigned int *)c->esc;
    a.esc = (ulonglong2 *)((char *)c->esc + 16);
    a.esc_cap = (unsigned int)((c->esc_bytes - 16) / 16);
    // FP64 queue: a counter, then entries
    const size_t xwant = c->misc_n >= 16 + 16 * 4096 ? c->misc_n : (size_t)16 + 16 * 65536;
    rc = clr_ensure(c, &c->misc, &c->misc_n, xwant);
    if (rc) return rc;
    a.exq_n = (unsigned int *)c->misc;
    a.exq = (uint4 *)((char *)c->misc + 16);
    a.exq_cap = (unsigned int)((c->misc_n - 16) / 16);
    CLR_HIP(c, hipMemsetAsync(a.esc_n, 0, sizeof(unsigned int), c->stream));
    CLR_HIP(c, hipMemsetAsync(a.exq_n, 0, sizeof(unsigned int), c->stream));
    CLR_HIP(c, hipEventRecord(c->ev2, c->stream));
    if (fused)
      hipLaunchKernelGGL(k_count<true>, dim3((unsigned)(ncells / kCountCells)), dim3(kLanes),
                         (size_t)np * kLanes * sizeof(uint4), c->stream, a);
    else
      hipLaunchKernelGGL(k_count<false>, dim3((unsigned)(ncells / kCountCells)), dim3(kLanes),
                         (size_t)np * kLanes * sizeof(uint4), c->stream, a);
    CLR_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_exact, dim3((a.exq_cap + 255) / 256), dim3(256), 0, c->stream, a);
    CLR_HIP(c, hipGetLastError());
    CLR_HIP(c, hipEventRecord(c->ev3, c->stream));
    ScanPtrs sp{};
    for (int i = 0; i < np; i++) {
      sp.in[i] = a.pop[i].btot;
      sp.bsum[i] = bsum[i];
      sp.out[i] = (long long *)a.pop[i].boff;
    }
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nsb, (unsigned)np), dim3(kScanB), 0, c->stream, sp, nblocks);
    hipLaunchKernelGGL(k_scan_top, dim3(1, (unsigned)np), dim3(kScanB), 0, c->stream, sp, nsb);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nsb, (unsigned)np), dim3(kScanB), 0, c->stream, sp, nblocks);
    CLR_HIP(c, hipGetLastError());
    for (int i = 0; i < np; i++)
      CLR_HIP(c, hipMemcpyAsync(&tot[i], a.pop[i].boff + nblocks, sizeof(long long), hipMemcpyDeviceToHost,
                                c->stream));
    unsigned int nexq = 0;
    CLR_HIP(c, hipMemcpyAsync(&nesc, a.esc_n, sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream));
    CLR_HIP(c, hipMemcpyAsync(&nexq, a.exq_n, sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream));
    CLR_HIP(c, hipStreamSynchronize(c->stream));
    if (nesc <= a.esc_cap && nexq <= a.exq_cap) break;
    if (attempt == 1) return clr_fail(c, CLR_ERR_STATE, "source lists overflow (%u, %u entries)", nesc, nexq);
    if (nesc > a.esc_cap) {
      rc = clr_ensure(c, &c->esc, &c->esc_bytes, (size_t)16 + 16 * ((size_t)nesc + nesc / 4 + 1024));
      if (rc) return rc;
    }
    if (nexq > a.exq_cap) {
      rc = clr_ensure(c, &c->misc, &c->misc_n, (size_t)16 + 16 * ((size_t)nexq + nexq / 4 + 1024));
      if (rc) return rc;
    }
  }
  a.esc_count = nesc;
  c->cnt_p0 = p0;
  c->cnt_np = np;
  c->ref_ipop = -1;
  c->cnt_per_pop = per_pop;
  c->cnt_nesc = nesc;
  if (nesc > 0) {
    // sorted by key for k_fill's binary search
    std::vector<ulonglong2> h(nesc);
    CLR_HIP(c, hipMemcpy(h.data(), a.esc, (size_t)nesc * sizeof(ulonglong2), hipMemcpyDeviceToHost));
    std::sort(h.begin(), h.end(), [](const ulonglong2 &u, const ulonglong2 &v) { return u.x < v.x; });
    for (const ulonglong2 &u : h)
      if (u.y == kCountInvalid) {
        c->cnt_p0 = -1;   // no counts to fetch
        c->cnt_np = 0;
        return clr_fail(c, CLR_ERR_ARG,
                        "population %d, cell %llu: the Poisson mean is not finite or above 2^20 (a bias model "
                        "with b < 0 near delta = -1?)",
                        p0 + (int)(u.x >> 56), (unsigned long long)(u.x & ((1ull << 56) - 1)));
      }
    CLR_HIP(c, hipMemcpy(a.esc, h.data(), (size_t)nesc * sizeof(ulonglong2), hipMemcpyHostToDevice));
  }
  float ms = 0;
  (void)hipEventElapsedTime(&ms, c->ev2, c->ev3);
  c->stage_ms[ST_SRC_PASS] = ms;
  long long total = 0;
  for (int i = 0; i < np; i++) {
    CatalogDev &C = c->cat[p0 + i];
    const long long n = tot[i];
    if (n > C.cap) {
      if (C.pos) (void)hipFree(C.pos);
      if (C.ipix) (void)hipFree(C.ipix);
      C.pos = nullptr;
      C.ipix = nullptr;
      C.cap = 0;
      const long long cap = n + n / 16 + 1024;
      if (hipMalloc((void **)&C.pos, (size_t)cap * 4 * sizeof(float)) != hipSuccess ||
          hipMalloc((void **)&C.ipix, (size_t)cap * sizeof(int32_t)) != hipSuccess) {
        (void)hipGetLastError();
        return clr_fail(c, CLR_ERR_OOM, "cannot allocate catalogue of %lld sources", n);
      }
      C.cap = cap;
    }
    C.n = n;
    a.pop[i].desc = (uint4 *)C.pos;
    total += n;
    if (n > 0 && !C.pos) return clr_fail(c, CLR_ERR_STATE, "no catalogue buffer");
  }
  CLR_HIP(c, hipEventRecord(c->ev2, c->stream));
  if (total > 0) {
    hipLaunchKernelGGL(k_fill, dim3((unsigned)(ncells / kCountCells)), dim3(kLanes), 0, c->stream, a);
    CLR_HIP(c, hipGetLastError());
  }
  rc = clr_phi_ready(c, 0);   // k_write reads phi
  if (rc) return rc;
  for (int i = 0; i < np; i++) {
    CatalogDev &C = c->cat[p0 + i];
    if (C.n == 0) continue;
    WriteArgs w{};
    w.phi = c->npot;
    w.r_arr = c->bg;
    w.v1 = c->bg + 3 * (size_t)na;
    w.idr = c->glob_idr;
    w.r_last = c->r_last;
    w.v1_0 = c->h_bg_first[3];
    w.v1_last = c->h_bg_last[3];
    w.dx = dx;
    w.factor_vel = -c->fgrowth_0 / (1.5 * c->hubble_0 * c->omega_m);
    w.idx = N / c->L;
    w.ox = c->obs[0];
    w.oy = c->obs[1];
    w.oz = c->obs[2];
    w.N = N;
    w.ngx = c->ngx;
    w.nside_base = nside_base;
    w.ipop = p0 + i;
    w.seed = seed;
    w.r2n = c->r2n;
    w.pos = (float4 *)C.pos;
    w.ipix = C.ipix;
    w.n = C.n;
    if (!w.phi || !w.pos || !w.ipix) return clr_fail(c, CLR_ERR_STATE, "no potential or catalogue buffer");
    hipLaunchKernelGGL(k_write, dim3((unsigned)((C.n + 255) / 256)), dim3(256), 0, c->stream, w);
    CLR_HIP(c, hipGetLastError());
  }
  CLR_HIP(c, hipEventRecord(c->ev3, c->stream));
  CLR_HIP(c, hipEventRecord(c->ev1, c->stream));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  (void)hipEventElapsedTime(&ms, c->ev2, c->ev3);
  c->stage_ms[ST_SRC_WRITE] = ms;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
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
  (void)hipFree(tmp);
  if (e != hipSuccess) return clr_fail(c, CLR_ERR_HIP, "source properties: %s", hipGetErrorString(e));
  return CLR_OK;
}

int clr_launch_fetch_counts(clr_ctx *c, int ipop, int iz0, int nz, int32_t *out) {
  const int N = c->N;
  if (iz0 < 0 || nz < 0 || iz0 + nz > N) return clr_fail(c, CLR_ERR_ARG, "planes %d..%d outside the grid", iz0, iz0 + nz);
  if (c->ref_ipop >= 0 && ipop == c->ref_ipop) {
    // counts of the last reference-stream call
    const size_t n = (size_t)N * N * nz, c0 = (size_t)N * N * iz0;
    if (nz == 0) return CLR_OK;
    CLR_HIP(c, hipMemcpy(out, (const uint32_t *)c->ref_cnt + c0, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    return CLR_OK;
  }
  if (c->cnt_p0 < 0 || ipop < c->cnt_p0 || ipop >= c->cnt_p0 + c->cnt_np)
    return clr_fail(c, CLR_ERR_STATE, "population %d was not sampled by the last clr_sample_sources call", ipop);
  if (nz == 0) return CLR_OK;
  const int q = ipop - c->cnt_p0;
  const uint8_t *cnt8 = (const uint8_t *)c->rowcnt + (size_t)q * c->cnt_per_pop;
  const uint64_t n = (uint64_t)N * N * nz, cell0 = (uint64_t)N * N * iz0;
  int32_t *tmp = nullptr;
  CLR_HIP(c, hipMalloc((void **)&tmp, n * sizeof(int32_t)));
  hipLaunchKernelGGL(k_expand_counts, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, cnt8,
                     (const ulonglong2 *)((char *)c->esc + 16), c->cnt_nesc, (unsigned long long)q, cell0, n, tmp);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return clr_fail(c, CLR_ERR_HIP, "cell counts: %s", hipGetErrorString(e));
  return CLR_OK;
}

int clr_launch_sources_stream(clr_ctx *c, int ipop, int nthr, const int *iz_begin, const uint32_t *streams,
                              long long stream_len, int nside_base, long long *nsrc) {
  const int N = c->N, na = c->na;
  if (ipop < 0 || ipop >= c->n_tr[CLR_TRACER_SRCS]) return clr_fail(c, CLR_ERR_ARG, "no source population %d", ipop);
  if (!c->density_done) return clr_fail(c, CLR_ERR_STATE, "density field not computed");
  if (N > 512) return clr_fail(c, CLR_ERR_ARG, "the reference-stream mode is for grids up to 512^3 (N=%d)", N);
  if (nthr < 1 || nthr > 4096 || !iz_begin || !streams || stream_len < 3)
    return clr_fail(c, CLR_ERR_ARG, "bad thread streams (nthr=%d, len=%lld)", nthr, stream_len);
  if (iz_begin[0] != 0 || iz_begin[nthr] != N)
    return clr_fail(c, CLR_ERR_ARG, "thread planes must cover 0..%d", N);
  for (int t = 0; t < nthr; t++)
    if (iz_begin[t + 1] < iz_begin[t]) return clr_fail(c, CLR_ERR_ARG, "thread planes must not decrease");
  if (nside_base < 1 || nside_base > 64 || (nside_base & (nside_base - 1)))
    return clr_fail(c, CLR_ERR_ARG, "nside_base=%d must be a power of two <= 64", nside_base);
  const TracerDev &T = c->tr[CLR_TRACER_SRCS][ipop];
  if (!T.set || !T.norm_set) return clr_fail(c, CLR_ERR_STATE, "population %d has no norm table", ipop);
  int rc = clr_ring2nest_table(c, nside_base);
  if (rc) return rc;
  rc = clr_phi_ready(c, 0);   // k_ref_write reads phi
  if (rc) return rc;
  const size_t ncells = (size_t)N * N * N;
  const size_t nsb = (ncells + kScanB - 1) / kScanB;
  rc = clr_ensure(c, (void **)&c->ref_cnt, &c->ref_cnt_n, ncells * sizeof(uint32_t));
  if (rc) return rc;
  c->cnt_p0 = -1;
  c->ref_ipop = -1;
  double *lam = nullptr;
  long long *off = nullptr, *bsum = nullptr;
  uint32_t *dst = nullptr;
  int *dizb = nullptr, *dflag = nullptr;
  const double dx = c->L / N;
  hipError_t e = hipMalloc((void **)&lam, ncells * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&off, (ncells + 1) * sizeof(long long));
  if (e == hipSuccess) e = hipMalloc((void **)&bsum, (nsb + 1) * sizeof(long long));
  if (e == hipSuccess) e = hipMalloc((void **)&dst, (size_t)nthr * stream_len * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc((void **)&dizb, (size_t)(nthr + 1) * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void **)&dflag, sizeof(int));
  auto release = [&]() {
    (void)hipStreamSynchronize(c->stream);
    for (void *q : {(void *)lam, (void *)off, (void *)bsum, (void *)dst, (void *)dizb, (void *)dflag})
      if (q) (void)hipFree(q);
  };
  if (e != hipSuccess) {
    (void)hipGetLastError();
    release();
    return clr_fail(c, CLR_ERR_OOM, "reference-stream buffers: %s", hipGetErrorString(e));
  }
  int flag = 0;
  long long total = 0;
  e = hipMemcpyAsync(dst, streams, (size_t)nthr * stream_len * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dizb, iz_begin, (size_t)(nthr + 1) * sizeof(int), hipMemcpyHostToDevice,
                                          c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(dflag, 0, sizeof(int), c->stream);
  RefArgs a{};
  a.g = c->dens;
  a.r_arr = c->bg;
  a.amp = T.amp;
  a.bz = T.bz;
  a.norm = T.norm;
  a.idr = c->glob_idr;
  a.r_last = c->r_last;
  a.bz0 = T.bz0;
  a.norm0 = T.norm0;
  a.normf = T.normf;
  a.dx = dx;
  a.cell_vol = dx * dx * dx;
  a.half_box_20 = c->L / 2 + 20.;
  a.ox = c->obs[0];
  a.oy = c->obs[1];
  a.oz = c->obs[2];
  a.N = N;
  a.ngx = c->ngx;
  a.model = T.model;
  a.nthr = nthr;
  a.izb = dizb;
  a.streams = dst;
  a.len = stream_len;
  a.lam = lam;
  a.cnt = (uint32_t *)c->ref_cnt;
  a.overflow = dflag;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_ref_lambda, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, c->stream, a);
    hipLaunchKernelGGL(k_ref_walk, dim3((unsigned)((nthr + 63) / 64)), dim3(64), 0, c->stream, a);
    ScanPtrs sp{};
    sp.in[0] = a.cnt;
    sp.bsum[0] = bsum;
    sp.out[0] = off;
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nsb, 1), dim3(kScanB), 0, c->stream, sp, ncells);
    hipLaunchKernelGGL(k_scan_top, dim3(1, 1), dim3(kScanB), 0, c->stream, sp, nsb);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nsb, 1), dim3(kScanB), 0, c->stream, sp, ncells);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&total, off + ncells, sizeof(long long), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    release();
    return clr_fail(c, CLR_ERR_HIP, "reference-stream counts: %s", hipGetErrorString(e));
  }
  if (flag) {
    release();
    return clr_fail(c, CLR_ERR_ARG, "a thread stream of %lld uniforms ran out in the counting pass", stream_len);
  }
  CatalogDev &C = c->cat[ipop];
  if (total > C.cap) {
    if (C.pos) (void)hipFree(C.pos);
    if (C.ipix) (void)hipFree(C.ipix);
    C.pos = nullptr;
    C.ipix = nullptr;
    C.cap = 0;
    const long long cap = total + total / 16 + 1024;
    if (hipMalloc((void **)&C.pos, (size_t)cap * 4 * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&C.ipix, (size_t)cap * sizeof(int32_t)) != hipSuccess) {
      (void)hipGetLastError();
      release();
      return clr_fail(c, CLR_ERR_OOM, "cannot allocate catalogue of %lld sources", total);
    }
    C.cap = cap;
  }
  C.n = total;
  if (total > 0) {
    WriteArgs w{};
    w.phi = c->npot;
    w.r_arr = c->bg;
    w.v1 = c->bg + 3 * (size_t)na;
    w.idr = c->glob_idr;
    w.r_last = c->r_last;
    w.v1_0 = c->h_bg_first[3];
    w.v1_last = c->h_bg_last[3];
    w.dx = dx;
    w.factor_vel = -c->fgrowth_0 / (1.5 * c->hubble_0 * c->omega_m);
    w.idx = N / c->L;
    w.ox = c->obs[0];
    w.oy = c->obs[1];
    w.oz = c->obs[2];
    w.N = N;
    w.ngx = c->ngx;
    w.nside_base = nside_base;
    w.ipop = ipop;
    w.r2n = c->r2n;
    w.pos = (float4 *)C.pos;
    w.ipix = C.ipix;
    w.n = total;
    RefWriteExtra x{a.cnt, off, dizb, dst, stream_len, nthr, dflag};
    hipLaunchKernelGGL(k_ref_write, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, c->stream, w, x);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  release();
  if (e != hipSuccess) return clr_fail(c, CLR_ERR_HIP, "reference-stream catalogue: %s", hipGetErrorString(e));
  if (flag) return clr_fail(c, CLR_ERR_ARG, "a thread stream of %lld uniforms ran out in the position pass", stream_len);
  c->ref_ipop = ipop;
  if (nsrc) *nsrc = total;
  return CLR_OK;
}

// ---------------------------------------------------------------------------
// The sampling stage for graph capture: buffers first (clr_prepare_sources:
// RING->NEST table, packed tables, count scratch, lists, catalogues of the
// given capacities), then kernel launches only (clr_enqueue_sources): every
// population in one pass, the norm edges read from the device, counters
// cleared by kernels, the source count left on the device (k_write reads it),
// and a status word: saturated counts (whose escape list this path does not
// sort), an overflowing FP64 queue or a catalogue past its capacity make the
// result invalid (the caller then re-runs the realisation with
// clr_sample_sources).  Same kernels, streams and order as clr_launch_sources.
// ---------------------------------------------------------------------------
static size_t src_per_pop(size_t ncells) {
  const size_t nblocks = ncells / kBlockCells, nsb = (nblocks + kScanB - 1) / kScanB;
  return (ncells + nblocks * 4 + (nblocks + 1) * 8 + (nsb + 1) * 8 + 64 + 255) / 256 * 256;
}

int clr_prepare_sources(clr_ctx *c, int nside_base, const long long *cap) {
  const int np = c->n_tr[CLR_TRACER_SRCS], N = c->N, na = c->na;
  const size_t ncells = (size_t)N * N * N;
  if (np < 1) return clr_fail(c, CLR_ERR_STATE, "no source population");
  if (N % kCountLaneCells || ncells % kCountCells)
    return clr_fail(c, CLR_ERR_ARG, "grid of %d^3 cells is not a whole number of %d-cell blocks", N, kCountCells);
  int rc = clr_ring2nest_table(c, nside_base);
  if (rc) return rc;
  for (int i = 0; i < np; i++) {
    TracerDev &T = c->tr[CLR_TRACER_SRCS][i];
    if (!T.set) return clr_fail(c, CLR_ERR_STATE, "population %d has no tables", i);
    if (!T.ptab) CLR_HIP(c, hipMalloc((void **)&T.ptab, pop_table_len(na) * sizeof(float4)));
    if (!T.edge) CLR_HIP(c, hipMalloc((void **)&T.edge, 2 * sizeof(double)));
    CatalogDev &C = c->cat[i];
    if (C.cap < cap[i]) {
      if (C.pos) (void)hipFree(C.pos);
      if (C.ipix) (void)hipFree(C.ipix);
      C.pos = nullptr;
      C.ipix = nullptr;
      C.cap = 0;
      if (hipMalloc((void **)&C.pos, (size_t)cap[i] * 4 * sizeof(float)) != hipSuccess ||
          hipMalloc((void **)&C.ipix, (size_t)cap[i] * sizeof(int32_t)) != hipSuccess) {
        (void)hipGetLastError();
        return clr_fail(c, CLR_ERR_OOM, "cannot allocate a catalogue of %lld sources", cap[i]);
      }
      C.cap = cap[i];
    }
  }
  rc = clr_ensure(c, (void **)&c->rowcnt, &c->rows_n, (size_t)np * src_per_pop(ncells));
  if (rc) return rc;
  rc = clr_ensure(c, &c->esc, &c->esc_bytes, c->esc_bytes > 0 ? c->esc_bytes : (size_t)16 + 16 * 65536);
  if (rc) return rc;
  return clr_ensure(c, &c->misc, &c->misc_n, c->misc_n >= 16 + 16 * 65536 ? c->misc_n : (size_t)16 + 16 * 65536);
}

int clr_enqueue_sources(clr_ctx *c, hipStream_t st, uint32_t seed, int nside_base, int *flag_dev) {
  const int np = c->n_tr[CLR_TRACER_SRCS], N = c->N, na = c->na;
  const size_t ncells = (size_t)N * N * N, per_pop = src_per_pop(ncells);
  const size_t nblocks = ncells / kBlockCells, nsb = (nblocks + kScanB - 1) / kScanB;
  if (c->rows_n < (size_t)np * per_pop || !c->esc || c->misc_n < 16 + 16 * 4096 || c->r2n_nside != nside_base)
    return clr_fail(c, CLR_ERR_STATE, "sampling buffers not prepared");
  const double dx = c->L / N, cell_vol = dx * dx * dx;
  SrcArgs a{};
  a.g = c->dens;
  a.r_arr = c->bg;
  a.idr = c->glob_idr;
  a.r_last = c->r_last;
  a.dx = dx;
  a.cell_vol = cell_vol;
  a.half_box_20 = c->L / 2 + 20.;
  a.ox = c->obs[0];
  a.oy = c->obs[1];
  a.oz = c->obs[2];
  a.N = N;
  a.ngx = c->ngx;
  a.na = na;
  a.npop = np;
  a.seed = seed;
  a.fused = 0;
  a.esc_n = (unsigned int *)c->esc;
  a.esc = (ulonglong2 *)((char *)c->esc + 16);
  a.esc_cap = (unsigned int)((c->esc_bytes - 16) / 16);
  a.esc_count = 0;
  a.exq_n = (unsigned int *)c->misc;
  a.exq = (uint4 *)((char *)c->misc + 16);
  a.exq_cap = (unsigned int)((c->misc_n - 16) / 16);
  std::vector<long long *> bsum(np);
  GraphTotals gt{};
  for (int i = 0; i < np; i++) {
    TracerDev &T = c->tr[CLR_TRACER_SRCS][i];
    CatalogDev &C = c->cat[i];
    if (!T.ptab || !T.edge || !C.pos || !C.ipix) return clr_fail(c, CLR_ERR_STATE, "sampling buffers not prepared");
    PopTab &P = a.pop[i];
    uint8_t *base = (uint8_t *)c->rowcnt + (size_t)i * per_pop;
    P.cnt8 = base;
    P.btot = (uint32_t *)(base + ncells);
    long long *off = (long long *)(((uintptr_t)(P.btot + nblocks) + 7) & ~(uintptr_t)7);
    P.boff = off;
    bsum[i] = off + nblocks + 1;
    P.amp = T.amp;
    P.bz = T.bz;
    P.norm = T.norm;
    P.tab = (const float4 *)T.ptab;
    P.bz0 = T.bz0;
    P.norm0 = 1;
    P.normf = 1;
    P.edges = T.edge;
    P.model = T.model;
    P.ipop = i;
    P.cap = C.cap;
    P.desc = (uint4 *)C.pos;
    gt.tot[i] = off + nblocks;
    gt.cap[i] = C.cap;
    // packed tables from the current norm tables
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, (uint32_t *)((float4 *)T.ptab + na), 4);
    hipLaunchKernelGGL(k_pop_table, dim3((na + 255) / 256), dim3(256), 0, st, T.amp, T.bz, T.norm, cell_vol, na,
                       (float4 *)T.ptab);
    hipLaunchKernelGGL(k_pop_coarse, dim3((coarse_count(na) + 255) / 256), dim3(256), 0, st, T.amp, T.bz, T.norm,
                       cell_vol, na, (float4 *)T.ptab);
  }
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, (uint32_t *)a.esc_n, 1);
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, (uint32_t *)a.exq_n, 1);
  hipLaunchKernelGGL(k_count<false>, dim3((unsigned)(ncells / kCountCells)), dim3(kLanes),
                     (size_t)np * kLanes * sizeof(uint4), st, a);
  hipLaunchKernelGGL(k_exact, dim3((a.exq_cap + 255) / 256), dim3(256), 0, st, a);
  ScanPtrs sp{};
  for (int i = 0; i < np; i++) {
    sp.in[i] = a.pop[i].btot;
    sp.bsum[i] = bsum[i];
    sp.out[i] = (long long *)a.pop[i].boff;
  }
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nsb, (unsigned)np), dim3(kScanB), 0, st, sp, nblocks);
  hipLaunchKernelGGL(k_scan_top, dim3(1, (unsigned)np), dim3(kScanB), 0, st, sp, nsb);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nsb, (unsigned)np), dim3(kScanB), 0, st, sp, nblocks);
  hipLaunchKernelGGL(k_graph_flags, dim3(1), dim3(64), 0, st, (const unsigned *)a.esc_n,
                     (const unsigned *)a.exq_n, a.exq_cap, np, gt, flag_dev);
  hipLaunchKernelGGL(k_fill, dim3((unsigned)(ncells / kCountCells)), dim3(kLanes), 0, st, a);
  for (int i = 0; i < np; i++) {
    CatalogDev &C = c->cat[i];
    WriteArgs w{};
    w.phi = c->npot;
    w.r_arr = c->bg;
    w.v1 = c->bg + 3 * (size_t)na;
    w.idr = c->glob_idr;
    w.r_last = c->r_last;
    w.v1_0 = c->h_bg_first[3];
    w.v1_last = c->h_bg_last[3];
    w.dx = dx;
    w.factor_vel = -c->fgrowth_0 / (1.5 * c->hubble_0 * c->omega_m);
    w.idx = N / c->L;
    w.ox = c->obs[0];
    w.oy = c->obs[1];
    w.oz = c->obs[2];
    w.N = N;
    w.ngx = c->ngx;
    w.nside_base = nside_base;
    w.ipop = i;
    w.seed = seed;
    w.r2n = c->r2n;
    w.pos = (float4 *)C.pos;
    w.ipix = C.ipix;
    w.n = C.cap;
    w.n_dev = gt.tot[i];
    hipLaunchKernelGGL(k_write, dim3((unsigned)((C.cap + 255) / 256)), dim3(256), 0, st, w);
  }
  CLR_HIP(c, hipGetLastError());
  return CLR_OK;
}

// After a graph-captured pass: the host-side state the normal path keeps
// (source counts, count layout for clr_fetch_cell_counts).
int clr_finish_sources(clr_ctx *c) {
  const int np = c->n_tr[CLR_TRACER_SRCS], N = c->N;
  const size_t ncells = (size_t)N * N * N, per_pop = src_per_pop(ncells), nblocks = ncells / kBlockCells;
  for (int i = 0; i < np; i++) {
    uint8_t *base = (uint8_t *)c->rowcnt + (size_t)i * per_pop;
    const uint32_t *btot = (const uint32_t *)(base + ncells);
    const long long *off = (const long long *)(((uintptr_t)(btot + nblocks) + 7) & ~(uintptr_t)7);
    long long n = 0;
    CLR_HIP(c, hipMemcpy(&n, off + nblocks, sizeof(long long), hipMemcpyDeviceToHost));
    c->cat[i].n = n;
  }
  c->cnt_p0 = 0;
  c->cnt_np = np;
  c->cnt_per_pop = per_pop;
  c->cnt_nesc = 0;
  c->ref_ipop = -1;
  return CLR_OK;
}

