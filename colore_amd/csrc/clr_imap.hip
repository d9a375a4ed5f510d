// clr_imap.hip -- imap_set_cartesian_single (imap.c:135-245) on gfx950.
//
// Every grid cell inside [r0_min - 20, rf_max + 20] with T(r) > 0:
// temperature T(r) bias_model(delta, b(r)) norm(r), RSD shift
// v_r V1(r)/H(r) (imap.c:199-200), then nsub^3 sub-points at
// x + (i+1/2) dx/nsub (imap.c:212-218); each sub-point whose shifted radius
// falls in a shell adds T and 1 to its RING pixel (he_ang2pix,
// healpix_extra.c:141-172).  k_imap paints the cells of shells with small
// nsub (a thread per cell, blocks on bricks of cells), k_imap_wave those of
// the inner shells (a wavefront per cell).  Both sum their sub-points per
// (shell, pixel) bin in LDS tables and add each bin's n T and n to the maps
// with device atomics (the reference adds T and 1 per sub-point with an omp
// atomic; the counts are the same, the float sums differ only by rounding,
// whose order neither fixes).  The shell of a radius is the reference's
// (get_r_index_imap, imap.c:76-103, for contiguous shells): found by binary
// search over the sorted shells, or by one comparison inside the two-shell
// window of the cell.
#include "clr_internal.h"
#include <algorithm>
#include <math.h>

using namespace clr;

namespace {

constexpr int kMaxShells = 1024;
constexpr int kThreadSub = 6;   // nsub above which a cell is painted by a whole wavefront

struct ImapArgs {
  const float *dens, *phi;
  const double *r_arr, *tz, *bz, *norm, *v1, *ih;
  double idr, r_last, bz0, norm0, normf, v1_0, v1_last, ih_0, ih_last;
  double dx, ox, oy, oz, rmin, rmax, factor_vel, idx;
  int N, ngx, model, nr;
  int64_t nside, npix;
  double rf_last;   // rf of the last shell
};

// Sub-point arithmetic.  The reference's chain for a sub-point (cart2sph,
// imap.c:24-41, then he_ang2pix) costs three FP64 divisions, an FP64 atan2
// and the pixel's floors.  pixel_fast evaluates the same chain with one
// division (cos theta = z/r, exact as the reference takes it) and an
// approximate azimuth -- atan of min/max by the Abramowitz & Stegun 4.4.49
// polynomial (Hastings; |error| <= 2e-8 rad on [0, 1]) -- and accepts its
// result only when every floor it takes is certain: the value lies further
// than the azimuth's error bound (carried through each product) from an
// integer.  Otherwise, and on the z axis or within the bound of phi = 0
// (where the reference's wrap to [0, 2pi) could differ), the exact chain
// runs.  The accepted pixel is therefore the exact chain's pixel.
constexpr double kAtanErr = 2e-8;   // A&S 4.4.49 bound; the polynomial's measured maximum is 1.36e-8

__device__ inline double atan01(double t) {
#pragma clang fp contract(off)
  const double t2 = t * t;
  double p = 0.0028662257;
  p = p * t2 - 0.0161657367;
  p = p * t2 + 0.0429096138;
  p = p * t2 - 0.0752896400;
  p = p * t2 + 0.1065626393;
  p = p * t2 - 0.1420889944;
  p = p * t2 + 0.1999355085;
  p = p * t2 - 0.3333314528;
  p = p * t2 + 1.0;
  return t * p;
}

// floor(v) when [v - e, v + e] holds no integer and v - e >= 0, else -1
__device__ inline int certain_floor(double v, double e) {
  const double lo = floor(v - e);
  return (lo >= 0 && lo == floor(v + e)) ? (int)lo : -1;
}

// RING pixel of the sub-point (x, y, z) with cth = z / r (r > 0) as the
// reference's chain gives it, or -1 when a floor is too close to call.
// nside is a power of two.
__device__ inline int pixel_fast(int nside, double x, double y, double cth) {
#pragma clang fp contract(off)
  const double ax = fabs(x), ay = fabs(y);
  const double mx = fmax(ax, ay), mn = fmin(ax, ay);
  if (!(mx > 0)) return -1;
  // the ratio through the hardware reciprocal and two Newton steps (a few
  // ulp; the exact quotient is not needed: the error stays far inside the
  // margin the floors are certified with)
  double rm = __builtin_amdgcn_rcp(mx);
  rm = rm * (2.0 - mx * rm);
  rm = rm * (2.0 - mx * rm);
  double a = atan01(mn * rm);
  if (ay > ax) a = 1.5707963267948966 - a;
  if (x < 0) a = 3.141592653589793 - a;
  if (y < 0) a = -a;
  // |a - atan2(y/r, x/r)| <= kAtanErr + 1e-14 (the reference's own roundings
  // and the constants' included)
  if (fabs(a) <= 4 * kAtanErr) return -1;
  const double ph = a < 0 ? a + 6.283185307179586 : a;
  const double tt = ph * 0.6366197723675814;
  const double ett = 0.6366197723675814 * (kAtanErr + 1e-14) + 1e-13;   // bound on |tt - tt_ref|
  const double dn = (double)nside;
  if (fabs(cth) <= 2.0 / 3.0) {
    const double t1 = dn * (0.5 + tt), t2 = dn * cth * 0.75;
    const double e = dn * ett + 1e-10;
    const int jp = certain_floor(t1 - t2, e), jm = certain_floor(t1 + t2, e);
    if (jp < 0 || jm < 0) return -1;
    const int ir = nside + 1 + jp - jm;
    const int kshift = 1 - (ir & 1);
    const int ip = ((jp + jm - nside + kshift + 1) / 2) & (4 * nside - 1);
    return nside * (nside - 1) * 2 + (ir - 1) * 4 * nside + ip;
  }
  const int k = certain_floor(tt, ett);
  if (k < 0) return -1;
  const double tp = tt - k;
  const double tmp = dn * sqrt(3 * (1 - fabs(cth)));
  const int jp = certain_floor(tp * tmp, tmp * ett + 1e-10);
  const int jm = certain_floor((1.0 - tp) * tmp, tmp * (ett + 1e-15) + 1e-10);
  if (jp < 0 || jm < 0) return -1;
  const int ir = jp + jm + 1;
  int ip = certain_floor(tt * ir, ir * ett + 1e-10);
  if (ip < 0) return -1;
  if (ip >= 4 * ir) ip -= 4 * ir;
  if (cth > 0) return 2 * ir * (ir - 1) + ip;
  return 12 * nside * nside - 2 * ir * (ir + 1) + ip;
}

// The reference's chain (cart2sph, he_ang2pix), out of line: it runs for
// the few sub-points pixel_fast leaves open.
__device__ __noinline__ int64_t pixel_exact(int64_t nside, double x, double y, double z, double r) {
#pragma clang fp contract(off)
  double cth = 1, phi = 0;
  if (r != 0) {
    const double xn = x / r, yn = y / r;
    cth = z / r;
    phi = atan2(yn, xn);
    if (phi < 0) phi += 2 * M_PI;
  }
  return ang2pix_ring(nside, cth, phi);
}

// shell index: -1 below the first shell, nr above the last, -2 in a gap
__device__ inline int shell_of(const double *__restrict__ r0, const double *__restrict__ rf, int nr, double r) {
  if (r < r0[0]) return -1;
  if (r >= rf[nr - 1]) return nr;
  int lo = 0, hi = nr - 1;   // largest i with r0[i] <= r
  while (hi > lo) {
    const int m = (lo + hi + 1) / 2;
    if (r0[m] <= r) lo = m; else hi = m - 1;
  }
  if (r < rf[lo]) return lo;
  return -2;
}

__device__ inline double rvel_cell(const ImapArgs &a, int ix, int iy, int iz, double x0, double y0, double z0,
                                   double rr) {
#pragma clang fp contract(off)
  const int N = a.N;
  const size_t ngx = (size_t)a.ngx;
  const float *phi = a.phi;
  const int ixh = ix == N - 1 ? 0 : ix + 1, ixl = ix == 0 ? N - 1 : ix - 1;
  const int iyh = iy == N - 1 ? 0 : iy + 1, iyl = iy == 0 ? N - 1 : iy - 1;
  const int izh = iz == N - 1 ? 0 : iz + 1, izl = iz == 0 ? N - 1 : iz - 1;
  const size_t z0o = (size_t)iz * N * ngx, y0o = (size_t)iy * ngx;
  const double u0 = x0 / rr, u1 = y0 / rr, u2 = z0 / rr;
  // FP32 differences of the stored potential, as the reference takes them
  const double v0 = phi[ixh + y0o + z0o] - phi[ixl + y0o + z0o];
  const double v1 = phi[ix + (size_t)iyh * ngx + z0o] - phi[ix + (size_t)iyl * ngx + z0o];
  const double v2 = phi[ix + y0o + (size_t)izh * N * ngx] - phi[ix + y0o + (size_t)izl * N * ngx];
  return 0.5 * a.idx * (v0 * u0 + v1 * u1 + v2 * u2);
}

// One cell's painting inputs: temperature T(r) bias_model(delta, b(r)) norm(r)
// and the RSD shift v_r V1(r)/H(r) (imap.c:185-200), and the sub-sampling
// nsub of its shell (imap.c:201-209).  false when the cell is outside the
// shells' range or has T(r) = 0.
// The shells its sub-points can reach: their shifted radii lie in [rlo, rhi]
// (the cube's centre radius +- half its diagonal, plus the cell's shift),
// and when no shell edge but r0[wi0 + 1] lies inside (wok), a sub-point's
// shell is wi0 or wi0 + 1 by one comparison (wf0, wf1: their rf).
struct CellPaint {
  double x0, y0, z0, dr_rsd, dxs;
  double rlo, rhi, wb, wf0, wf1;
  float temp;
  int ns, wi0;
  bool wok;
};
__device__ inline bool cell_paint(const ImapArgs &a, const double *__restrict__ r0s, const double *__restrict__ rfs,
                                  const int *__restrict__ nsub, int ix, int iy, int iz, CellPaint &cp) {
#pragma clang fp contract(off)
  const double x0 = ix * a.dx - a.ox, y0 = iy * a.dx - a.oy, z0 = iz * a.dx - a.oz;
  const double r0 = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
  const double tmean = f_of_r(a.r_arr, a.tz, a.idr, a.r_last, r0, 0.0, 0.0);
  if (!(r0 <= a.rmax && r0 >= a.rmin)) return false;
  if (!(tmean > 0)) return false;
  const double bias = f_of_r(a.r_arr, a.bz, a.idr, a.r_last, r0, a.bz0, 1.0);
  const double dnorm = f_of_r(a.r_arr, a.norm, a.idr, a.r_last, r0, a.norm0, a.normf);
  double rvel = 0;
  if (r0 > 0) rvel = a.factor_vel * rvel_cell(a, ix, iy, iz, x0, y0, z0, r0);
  cp.dr_rsd = rvel * f_of_r(a.r_arr, a.v1, a.idr, a.r_last, r0, a.v1_0, a.v1_last) *
              f_of_r(a.r_arr, a.ih, a.idr, a.r_last, r0, a.ih_0, a.ih_last);
  // The temperature only enters the FP32 map sums: bias_model (common.h:
  // 414-431) in FP32 on the hardware exp2/log2 (a few 1e-7 relative; the
  // FP64 pow cost as much as a cell's whole sub-point walk at nsub = 2).
  const float dl = a.dens[((size_t)iz * a.N + iy) * (size_t)a.ngx + ix];
  float bm;
  {
    const float b = (float)bias;
    if (!(dl > -1.f)) {
      bm = 0.f;
    } else if (a.model == 1) {
      bm = dl < 0.f ? __builtin_amdgcn_exp2f(1.44269504f * b * dl * __builtin_amdgcn_rcpf(1.f + dl)) : 1.f + b * dl;
    } else if (a.model == 2) {
      bm = fmaxf(1.f + b * dl, 0.f);
    } else {
      bm = __builtin_amdgcn_exp2f(b * __builtin_amdgcn_logf(1.f + dl));
    }
  }
  cp.temp = (float)(tmean * dnorm) * bm;
  const int irad = shell_of(r0s, rfs, a.nr, r0);
  cp.ns = irad < 0 ? nsub[0] : (irad >= a.nr ? nsub[a.nr - 1] : nsub[irad]);
  cp.dxs = a.dx / cp.ns;
  cp.x0 = x0;
  cp.y0 = y0;
  cp.z0 = z0;
  {
    const double xc = x0 + 0.5 * a.dx, yc = y0 + 0.5 * a.dx, zc = z0 + 0.5 * a.dx;
    const double rc = sqrt(xc * xc + yc * yc + zc * zc), hd = 0.8660254038 * a.dx;
    const double m = 1e-9 * (rc + hd + fabs(cp.dr_rsd)) + 1e-9;
    cp.rlo = rc - hd + cp.dr_rsd - m;
    cp.rhi = rc + hd + cp.dr_rsd + m;
    int lo = -1, hi = a.nr - 1;   // largest i with r0[i] <= rlo, or -1
    while (hi > lo) {
      const int mid = (lo + hi + 1) / 2;
      if (r0s[mid] <= cp.rlo) lo = mid; else hi = mid - 1;
    }
    cp.wi0 = lo;
    cp.wb = lo + 1 < a.nr ? r0s[lo + 1] : INFINITY;
    cp.wf0 = lo >= 0 ? rfs[lo] : 0.0;
    cp.wf1 = lo + 1 < a.nr ? rfs[lo + 1] : 0.0;
    cp.wok = lo + 2 >= a.nr || cp.rhi < r0s[lo + 2];
  }
  return true;
}

// shell_of for a sub-point of the cell (the same index)
__device__ inline int shell_near(const ImapArgs &a, const double *__restrict__ r0s, const double *__restrict__ rfs,
                                 const CellPaint &cp, double rs) {
  if (!(cp.wok && rs >= cp.rlo && rs <= cp.rhi)) return shell_of(r0s, rfs, a.nr, rs);
  const int i = rs >= cp.wb ? cp.wi0 + 1 : cp.wi0;
  if (i < 0) return -1;
  if (rs >= a.rf_last) return a.nr;
  return rs < (i == cp.wi0 ? cp.wf0 : cp.wf1) ? i : -2;
}

// The shell (or < 0 outside the shells) and RING pixel of sub-point
// (izz, iyy, ixx) of a cell (imap.c:210-240).
__device__ inline int subpoint_shell_pixel(const ImapArgs &a, const double *__restrict__ r0s,
                                           const double *__restrict__ rfs, const CellPaint &cp, int izz, int iyy,
                                           int ixx, int64_t &pix) {
#pragma clang fp contract(off)
  const double z = cp.z0 + (izz + 0.5) * cp.dxs;
  const double y = cp.y0 + (iyy + 0.5) * cp.dxs;
  const double x = cp.x0 + (ixx + 0.5) * cp.dxs;
  const double r = sqrt(x * x + y * y + z * z);
  const int irad = shell_near(a, r0s, rfs, cp, r + cp.dr_rsd);
  if (irad < 0 || irad >= a.nr) return -1;
  pix = -1;
  if (r != 0) pix = pixel_fast((int)a.nside, x, y, z / r);
  if (pix < 0) pix = pixel_exact(a.nside, x, y, z, r);
  return irad;
}

// The (shell, RING pixel) bin of a sub-point, or -1 outside the shells.
__device__ inline int64_t subpoint_bin(const ImapArgs &a, const double *__restrict__ r0s,
                                       const double *__restrict__ rfs, const CellPaint &cp, int izz, int iyy,
                                       int ixx) {
  int64_t pix;
  const int irad = subpoint_shell_pixel(a, r0s, rfs, cp, izz, iyy, ixx, pix);
  return irad < 0 ? -1 : (int64_t)irad * a.npix + pix;
}

// Cells whose shell has nsub <= kThreadSub: one thread per cell, each
// sub-point adding T and 1 to its bin in the block's LDS table.
//
// A block works on bricks of kBrick^3 cells, in Morton order (consecutive
// workgroups on neighbouring bricks).  A brick's sub-points fall in the
// bins of its footprint, a few hundred at most, each many times: the block
// sums them in its LDS table -- open addressing on the bin's map offset,
// slots claimed with LDS compare-and-swap and listed -- and adds each bin to
// the maps with one pair of device atomics when the brick is done.  A bin
// that finds no slot within kProbe probes goes to the maps directly.
// Measured on 1024^3, nside 512, 64 shells, cells with nsub <= 4: row order
// with device atomics per cache entry 219 ms; bricks 177 ms; bricks with the
// LDS table 106 ms.  The split at nsub 6 (k_imap_wave's per-cell costs
// outweigh its width for smaller cells) measured 164 ms for the whole pass
// against 180 at 4 and 174 at 8.
// (Giving each XCD a contiguous eighth of the Morton order measured the
// same for this kernel and much slower for k_imap_wave.)
constexpr int kBrick = 8;
constexpr int kImapBlocks = 8192;
constexpr int kProbe = 8;
constexpr int kBrickSlots = 2048;

// bits 0, 3, 6, ... of m, packed
__device__ inline unsigned morton_compact(unsigned long long m) {
  m &= 0x1249249249249249ull;
  m = (m ^ (m >> 2)) & 0x10C30C30C30C30C3ull;
  m = (m ^ (m >> 4)) & 0x100F00F00F00F00Full;
  m = (m ^ (m >> 8)) & 0x1F0000FF0000FFull;
  m = (m ^ (m >> 16)) & 0x1F00000000FFFFull;
  m = (m ^ (m >> 32)) & 0x1FFFFFull;
  return (unsigned)m;
}

struct BrickBins {
  unsigned long long key[kBrickSlots];   // map offset, ~0: empty
  float temp[kBrickSlots];
  int hits[kBrickSlots];
  short used[kBrickSlots];
  int nused;
};

// the table's slot of map offset `key`, claimed if new, or -1
__device__ inline int brick_slot(BrickBins &b, unsigned long long key) {
  unsigned h = (((unsigned)key * 2654435761u) ^ ((unsigned)(key >> 32) * 0x85EBCA6Bu)) >> 21;   // 11 bits
  for (int p = 0; p < kProbe; p++) {
    const unsigned long long k = __hip_atomic_load(&b.key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == key) return (int)h;
    if (k == ~0ull) {
      const unsigned long long old = atomicCAS(&b.key[h], ~0ull, key);
      if (old == ~0ull) {
        b.used[atomicAdd(&b.nused, 1)] = (short)h;
        return (int)h;
      }
      if (old == key) return (int)h;
    }
    h = (h + 1) & (kBrickSlots - 1);
  }
  return -1;
}

// At least 4 waves per SIMD (128 VGPRs at most): 141 -> 118 ms against 3
// waves when the kernel needed 146.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_imap(ImapArgs a, const double *__restrict__ r0s,
                                              const double *__restrict__ rfs, const int *__restrict__ nsub,
                                              int nbs, float *__restrict__ data, int *__restrict__ nadd) {
  __shared__ BrickBins bb;
  for (int i = threadIdx.x; i < kBrickSlots; i += blockDim.x) {
    bb.key[i] = ~0ull;
    bb.temp[i] = 0.f;
    bb.hits[i] = 0;
  }
  if (threadIdx.x == 0) bb.nused = 0;
  __syncthreads();
  const int nb1 = (a.N + kBrick - 1) / kBrick;   // bricks per side; nbs: the power of two >= nb1
  const unsigned long long nbt = (unsigned long long)nbs * nbs * nbs;
  for (unsigned long long m = blockIdx.x; m < nbt; m += gridDim.x) {
    const int bx = (int)morton_compact(m), by = (int)morton_compact(m >> 1), bz = (int)morton_compact(m >> 2);
    if (bx >= nb1 || by >= nb1 || bz >= nb1) continue;   // uniform over the block
    for (int c = threadIdx.x; c < kBrick * kBrick * kBrick; c += blockDim.x) {
      const int ix = bx * kBrick + c % kBrick, iy = by * kBrick + (c / kBrick) % kBrick;
      const int iz = bz * kBrick + c / (kBrick * kBrick);
      if (ix >= a.N || iy >= a.N || iz >= a.N) continue;
      CellPaint cp;
      if (!cell_paint(a, r0s, rfs, nsub, ix, iy, iz, cp)) continue;
      if (cp.ns > kThreadSub) continue;   // k_imap_wave's cell
      const int ns = cp.ns;
      // each sub-point adds T and 1 to its bin's slot (a per-cell register
      // cache of 4 bins in front of the table measured 123 against 98 ms)
      for (int izz = 0; izz < ns; izz++)
        for (int iyy = 0; iyy < ns; iyy++)
          for (int ixx = 0; ixx < ns; ixx++) {
            const int64_t o = subpoint_bin(a, r0s, rfs, cp, izz, iyy, ixx);
            if (o < 0) continue;
            const int sl = brick_slot(bb, (unsigned long long)o);
            if (sl >= 0) {
              atomicAdd(&bb.temp[sl], cp.temp);
              atomicAdd(&bb.hits[sl], 1);
            } else {
              atomicAdd(&data[o], cp.temp);
              atomicAdd(&nadd[o], 1);
            }
          }
    }
    // the brick's bins into the maps
    __syncthreads();
    const int nu = bb.nused;
    for (int i = threadIdx.x; i < nu; i += blockDim.x) {
      const int h = bb.used[i];
      const unsigned long long o = bb.key[h];
      atomicAdd(&data[o], bb.temp[h]);
      atomicAdd(&nadd[o], bb.hits[h]);
      bb.key[h] = ~0ull;
      bb.temp[h] = 0.f;
      bb.hits[h] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) bb.nused = 0;
    __syncthreads();
  }
}

// Cells whose shell has nsub > kThreadSub (the shells nearest the observer,
// up to 22^3 sub-points per cell at 1024^3): one wavefront per cell, one
// sub-point per lane.  A cell's sub-points fall in the few hundred bins of
// its footprint (and at most two shells when its window holds), each many
// times: the wavefront sums them in its own LDS table of bins -- open
// addressing, key (shell - wi0) npix + pixel, slots claimed with LDS
// compare-and-swap and listed for the flush -- and adds each bin's sums to
// the maps with one pair of device atomics when the cell is done (or the
// table is three quarters full).  Sub-points whose bin finds no slot within
// kProbe probes, and cells whose window does not hold, add to the maps
// directly.  Cells of the cube [lo, lo + side)^3 around the observer, taken
// by the wavefronts in turn.  Measured on 1024^3, nside 512, 64 shells,
// cells with nsub > 4: device atomics per sub-point 257 ms, with the LDS
// tables 133 ms.
constexpr int kWaveSlots = 512;                // bins per wavefront table (1024: 61 ms, the LDS then limits occupancy)
constexpr int kWaveFill = 3 * kWaveSlots / 4;  // emptied past this many bins
constexpr int kImapWaves = 2;                  // wavefronts per block of k_imap_wave

struct WaveBins {
  int key[kWaveSlots];   // -1: empty
  float temp[kWaveSlots];
  int hits[kWaveSlots];
  short used[kWaveSlots];
  int nused;
};

// the table's slot of `key`, claimed if new, or -1 after kProbe probes
__device__ inline int wave_slot(WaveBins &b, int key) {
  unsigned h = ((unsigned)key * 2654435761u) >> 23;   // top 9 bits: kWaveSlots = 512
  for (int p = 0; p < kProbe; p++) {
    const int k = __hip_atomic_load(&b.key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    if (k == key) return (int)h;
    if (k == -1) {
      const int old = atomicCAS(&b.key[h], -1, key);
      if (old == -1) {
        b.used[atomicAdd(&b.nused, 1)] = (short)h;
        return (int)h;
      }
      if (old == key) return (int)h;
    }
    h = (h + 1) & (kWaveSlots - 1);
  }
  return -1;
}

// adds the table's bins to the maps and empties it (wavefront-collective)
__device__ inline void wave_flush(WaveBins &b, int lane, int wi0, int64_t npix, float *__restrict__ data,
                                  int *__restrict__ nadd) {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const int nu = __hip_atomic_load(&b.nused, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  for (int i = lane; i < nu; i += 64) {
    const int h = b.used[i];
    const int key = b.key[h];
    const int ds = key >= npix ? 1 : 0;
    const int64_t o = (int64_t)(wi0 + ds) * npix + (key - ds * npix);
    atomicAdd(&data[o], b.temp[h]);
    atomicAdd(&nadd[o], b.hits[h]);
    b.key[h] = -1;
    b.temp[h] = 0.f;
    b.hits[h] = 0;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) b.nused = 0;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(64 * kImapWaves) void k_imap_wave(ImapArgs a, const double *__restrict__ r0s,
                                                             const double *__restrict__ rfs,
                                                             const int *__restrict__ nsub, int lo_x, int lo_y,
                                                             int lo_z, int side, float *__restrict__ data,
                                                             int *__restrict__ nadd) {
  __shared__ WaveBins tabs[kImapWaves];
  const int lane = threadIdx.x % 64;
  WaveBins &b = tabs[threadIdx.x / 64];
  for (int i = lane; i < kWaveSlots; i += 64) {
    b.key[i] = -1;
    b.temp[i] = 0.f;
    b.hits[i] = 0;
  }
  if (lane == 0) b.nused = 0;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  // cells in row order, consecutive wavefronts on neighbouring cells of a
  // row (a compact order -- neighbouring cells along the line of sight share
  // their bins -- measured 2.5x slower: the flushes of concurrent wavefronts
  // then contend for the same map lines)
  const long nw = (long)gridDim.x * kImapWaves;
  const long ncell = (long)side * side * side;
  for (long w = (long)blockIdx.x * kImapWaves + threadIdx.x / 64; w < ncell; w += nw) {
    const int ix = lo_x + (int)(w % side), iy = lo_y + (int)((w / side) % side), iz = lo_z + (int)(w / ((long)side * side));
    CellPaint cp;
    if (!cell_paint(a, r0s, rfs, nsub, ix, iy, iz, cp)) continue;
    if (cp.ns <= kThreadSub) continue;   // k_imap's cell
    const int ns = cp.ns, nq = ns * ns * ns;
    for (int q0 = 0; q0 < nq; q0 += 64) {
      const int q = q0 + lane;
      if (q < nq) {
        const int ixx = q % ns, iyy = (q / ns) % ns, izz = q / (ns * ns);
        int64_t pix;
        const int s = subpoint_shell_pixel(a, r0s, rfs, cp, izz, iyy, ixx, pix);
        if (s >= 0) {
          const int ds = s - cp.wi0;   // 0 or 1 when the window holds
          const int slot = (cp.wok && (ds == 0 || ds == 1)) ? wave_slot(b, (int)(ds * a.npix + pix)) : -1;
          if (slot >= 0) {
            atomicAdd(&b.temp[slot], cp.temp);
            atomicAdd(&b.hits[slot], 1);
          } else {
            const int64_t o = (int64_t)s * a.npix + pix;
            atomicAdd(&data[o], cp.temp);
            atomicAdd(&nadd[o], 1);
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (__hip_atomic_load(&b.nused, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) > kWaveFill)
        wave_flush(b, lane, cp.wi0, a.npix, data, nadd);
    }
    wave_flush(b, lane, cp.wi0, a.npix, data, nadd);
  }
}

}  // namespace

int clr_launch_imap(clr_ctx *c, int ipop, int nside, int nr, const double *r0, const double *rf, float *data,
                    int32_t *nadd) {
  if (ipop < 0 || ipop >= c->n_tr[CLR_TRACER_IMAP] || !c->tr[CLR_TRACER_IMAP][ipop].set)
    return clr_fail(c, CLR_ERR_ARG, "no intensity species %d", ipop);
  TracerDev &T = c->tr[CLR_TRACER_IMAP][ipop];
  if (!T.norm_set) return clr_fail(c, CLR_ERR_STATE, "species %d has no norm table", ipop);
  if (!c->density_done) return clr_fail(c, CLR_ERR_STATE, "density field not computed");
  if (nside < 1 || nside > 8192 || (nside & (nside - 1))) return clr_fail(c, CLR_ERR_ARG, "bad nside %d", nside);
  if (nr < 1 || nr > kMaxShells || !r0 || !rf || !data || !nadd)
    return clr_fail(c, CLR_ERR_ARG, "need 1..%d shells and output maps", kMaxShells);
  const int N = c->N, na = c->na;
  // imap_preproc (imap.c:105-133): sort shells by r0 (flouble compare)
  std::vector<int> order(nr);
  for (int i = 0; i < nr; i++) order[i] = i;
  std::vector<float> r0f(nr);
  for (int i = 0; i < nr; i++) r0f[i] = (float)r0[i];
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return r0f[a] < r0f[b]; });
  std::vector<double> sr0(nr), srf(nr);
  for (int i = 0; i < nr; i++) { sr0[i] = r0[order[i]]; srf[i] = rf[order[i]]; }
  const int64_t npix = 12LL * nside * nside;
  const double pixel_size = sqrt(4 * M_PI / (double)npix);
  const double dx = c->L / N;
  std::vector<int> nsub(nr);
  for (int i = 0; i < nr; i++) {
    const double a0 = sr0[i], af = srf[i];
    const double vol = (af * af * af - a0 * a0 * a0) * pixel_size * pixel_size / 3;
    const double sct = a0 * pixel_size, scr = af - a0, scv = pow(vol, 0.333333);
    const double sc0 = fmin(scv, fmin(sct, scr));
    nsub[i] = (int)(dx / sc0 + 0.5) + 1;
    if (nsub[i] < 1 || nsub[i] > 512)
      return clr_fail(c, CLR_ERR_ARG, "shell %d gives %d sub-samples per cell side", i, nsub[i]);
  }
  ImapArgs a{};
  a.dens = c->dens;
  {
    const int prc = clr_phi_ready(c, 0);   // phi's passes may still run on the second stream
    if (prc) return prc;
  }
  a.phi = c->npot;
  a.r_arr = c->bg;
  a.tz = T.amp;
  a.bz = T.bz;
  a.norm = T.norm;
  a.v1 = c->bg + 3 * (size_t)na;
  a.ih = c->bg + 5 * (size_t)na;
  double e[4];
  CLR_HIP(c, hipMemcpy(&e[0], a.v1, sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipMemcpy(&e[1], a.v1 + na - 1, sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipMemcpy(&e[2], a.ih, sizeof(double), hipMemcpyDeviceToHost));
  CLR_HIP(c, hipMemcpy(&e[3], a.ih + na - 1, sizeof(double), hipMemcpyDeviceToHost));
  a.v1_0 = e[0]; a.v1_last = e[1]; a.ih_0 = e[2]; a.ih_last = e[3];
  a.idr = c->glob_idr;
  a.r_last = c->r_last;
  a.bz0 = T.bz0;
  a.norm0 = T.norm0;
  a.normf = T.normf;
  a.dx = dx;
  a.ox = c->obs[0];
  a.oy = c->obs[1];
  a.oz = c->obs[2];
  a.rmin = sr0[0] - 20.;
  a.rmax = srf[nr - 1] + 20.;
  a.factor_vel = -c->fgrowth_0 / (1.5 * c->hubble_0 * c->omega_m);
  a.idx = N / c->L;
  a.N = N;
  a.ngx = c->ngx;
  a.model = T.model;
  a.nr = nr;
  a.nside = nside;
  a.npix = npix;
  a.rf_last = srf[nr - 1];

  const size_t mapn = (size_t)npix * nr;
  const size_t need = 2 * (size_t)nr * sizeof(double) + (size_t)nr * sizeof(int) + 64 + mapn * (sizeof(float) + sizeof(int));
  int rc = clr_ensure(c, &c->misc, &c->misc_n, need);
  if (rc) return rc;
  double *dr0 = (double *)c->misc, *drf = dr0 + nr;
  int *dns = (int *)(drf + nr);
  float *dmap = (float *)(((uintptr_t)(dns + nr) + 63) & ~(uintptr_t)63);
  int *dnadd = (int *)(dmap + mapn);
  CLR_HIP(c, hipMemcpyAsync(dr0, sr0.data(), nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
  CLR_HIP(c, hipMemcpyAsync(drf, srf.data(), nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
  CLR_HIP(c, hipMemcpyAsync(dns, nsub.data(), nr * sizeof(int), hipMemcpyHostToDevice, c->stream));
  CLR_HIP(c, hipMemsetAsync(dmap, 0, mapn * (sizeof(float) + sizeof(int)), c->stream));
  CLR_HIP(c, hipEventRecord(c->ev0, c->stream));
  int nbs = 1;
  while (nbs < (N + kBrick - 1) / kBrick) nbs *= 2;
  hipLaunchKernelGGL(k_imap, dim3(kImapBlocks), dim3(256), 0, c->stream, a, dr0, drf, dns, nbs, dmap, dnadd);
  CLR_HIP(c, hipGetLastError());
  // the cells of shells with nsub > kThreadSub lie within r_wave of the
  // observer (cells below the first shell and in gaps take the first
  // shell's nsub, cells beyond the last take the last one's)
  double r_wave = -1;
  for (int i = 0; i < nr; i++)
    if (nsub[i] > kThreadSub) r_wave = std::max(r_wave, srf[i]);
  bool gaps = false;   // cells between two shells also take nsub[0]
  for (int i = 0; i + 1 < nr; i++) gaps = gaps || srf[i] < sr0[i + 1];
  if (nsub[0] > kThreadSub) r_wave = std::max(r_wave, gaps ? a.rmax : sr0[0]);
  if (nsub[nr - 1] > kThreadSub) r_wave = a.rmax;
  if (r_wave > 0) {
    r_wave = std::min(r_wave, a.rmax) + dx;
    int lo[3], hi[3];
    for (int ax = 0; ax < 3; ax++) {
      lo[ax] = std::max(0, (int)floor((c->obs[ax] - r_wave) / dx));
      hi[ax] = std::min(N, (int)ceil((c->obs[ax] + r_wave) / dx) + 1);
    }
    const int side = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    // a cube of `side` cells from lo, clamped to the grid
    for (int ax = 0; ax < 3; ax++) lo[ax] = std::max(0, std::min(lo[ax], N - side));
    if (side > 0 && side <= N)
      hipLaunchKernelGGL(k_imap_wave, dim3(kImapBlocks), dim3(64 * kImapWaves), 0, c->stream, a, dr0, drf, dns,
                         lo[0], lo[1], lo[2], side, dmap, dnadd);
    CLR_HIP(c, hipGetLastError());
  }
  CLR_HIP(c, hipEventRecord(c->ev1, c->stream));
  CLR_HIP(c, hipMemcpyAsync(data, dmap, mapn * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  CLR_HIP(c, hipMemcpyAsync(nadd, dnadd, mapn * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  CLR_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->stage_ms[ST_IMAP] = ms;
  return CLR_OK;
}
