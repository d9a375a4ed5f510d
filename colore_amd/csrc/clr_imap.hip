// clr_imap.hip -- imap_set_cartesian_single (imap.c:135-245) on gfx950.
//
// One thread per grid cell inside [r0_min - 20, rf_max + 20] with T(r) > 0:
// temperature T(r) bias_model(delta, b(r)) norm(r), RSD shift
// v_r V1(r)/H(r) (imap.c:199-200), then nsub^3 sub-points at
// x + (i+1/2) dx/nsub (imap.c:212-218); each sub-point whose shifted radius
// falls in a shell adds T and 1 to its RING pixel (he_ang2pix,
// healpix_extra.c:141-172).  A cell's sub-points land in a few (shell,
// pixel) bins: the thread counts them in a small register cache and adds
// n T and n per bin with device atomics (the reference adds T and 1 per
// sub-point with an omp atomic; the counts are the same, the float sums
// differ only by rounding, whose order neither fixes).  The shell of a radius is found by binary search over the
// sorted shells; for contiguous shells this is the shell the reference's walk
// (get_r_index_imap, imap.c:76-103) returns.
#include "clr_internal.h"
#include <algorithm>
#include <math.h>

using namespace clr;

namespace {

constexpr int kMaxShells = 1024;
constexpr int kThreadSub = 4;   // nsub above which a cell is painted by a whole wavefront

struct ImapArgs {
  const float *dens, *phi;
  const double *r_arr, *tz, *bz, *norm, *v1, *ih;
  double idr, r_last, bz0, norm0, normf, v1_0, v1_last, ih_0, ih_last;
  double dx, ox, oy, oz, rmin, rmax, factor_vel, idx;
  int N, ngx, model, nr;
  int64_t nside, npix;
};

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
struct CellPaint {
  double x0, y0, z0, dr_rsd, dxs;
  float temp;
  int ns;
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
  const double dl = a.dens[((size_t)iz * a.N + iy) * (size_t)a.ngx + ix];
  cp.temp = (float)(tmean * bias_model_f64(dl, bias, a.model) * dnorm);
  const int irad = shell_of(r0s, rfs, a.nr, r0);
  cp.ns = irad < 0 ? nsub[0] : (irad >= a.nr ? nsub[a.nr - 1] : nsub[irad]);
  cp.dxs = a.dx / cp.ns;
  cp.x0 = x0;
  cp.y0 = y0;
  cp.z0 = z0;
  return true;
}

// The (shell, RING pixel) bin of sub-point (izz, iyy, ixx) of a cell
// (imap.c:210-240), or -1 outside the shells.
__device__ inline int64_t subpoint_bin(const ImapArgs &a, const double *__restrict__ r0s,
                                       const double *__restrict__ rfs, const CellPaint &cp, int izz, int iyy,
                                       int ixx) {
#pragma clang fp contract(off)
  const double z = cp.z0 + (izz + 0.5) * cp.dxs;
  const double y = cp.y0 + (iyy + 0.5) * cp.dxs;
  const double x = cp.x0 + (ixx + 0.5) * cp.dxs;
  const double r = sqrt(x * x + y * y + z * z);
  double cth = 1, phi = 0;
  if (r != 0) {
    const double xn = x / r, yn = y / r;
    cth = z / r;
    phi = atan2(yn, xn);
    if (phi < 0) phi += 2 * M_PI;
  }
  const int irad = shell_of(r0s, rfs, a.nr, r + cp.dr_rsd);
  if (irad < 0 || irad >= a.nr) return -1;
  return (int64_t)irad * a.npix + ang2pix_ring(a.nside, cth, phi);
}

// Cells whose shell has nsub <= kThreadSub: one thread per cell.  Its
// sub-points land in a few bins: the thread counts them in a small register
// cache and adds n T and n per bin with device atomics.
__global__ __launch_bounds__(256) void k_imap(ImapArgs a, const double *__restrict__ r0s,
                                              const double *__restrict__ rfs, const int *__restrict__ nsub,
                                              float *__restrict__ data, int *__restrict__ nadd) {
  const size_t row = blockIdx.x;
  const int iz = (int)(row / a.N), iy = (int)(row % a.N);
  for (int ix = threadIdx.x; ix < a.N; ix += blockDim.x) {
    CellPaint cp;
    if (!cell_paint(a, r0s, rfs, nsub, ix, iy, iz, cp)) continue;
    if (cp.ns > kThreadSub) continue;   // k_imap_wave's cell
    const int ns = cp.ns;
    const float temp = cp.temp;
    // (bin, hits) of the cell's sub-points; a full cache flushes its oldest
    // entry
    constexpr int kCache = 4;   // measured: 2 / 4 / 8 / 16 / 32 entries -> 610 / 512 / 526 / 550 / 651 ms
    int64_t key[kCache];
    int hits[kCache];
#pragma unroll
    for (int k = 0; k < kCache; k++) {
      key[k] = -1;
      hits[k] = 0;
    }
    int next = 0;
    const auto flush = [&](int64_t o, int n) {
      atomicAdd(&data[o], temp * (float)n);
      atomicAdd(&nadd[o], n);
    };
    for (int izz = 0; izz < ns; izz++)
      for (int iyy = 0; iyy < ns; iyy++)
        for (int ixx = 0; ixx < ns; ixx++) {
          const int64_t o = subpoint_bin(a, r0s, rfs, cp, izz, iyy, ixx);
          if (o < 0) continue;
          bool found = false;
#pragma unroll
          for (int k = 0; k < kCache; k++)
            if (key[k] == o) {
              hits[k]++;
              found = true;
            }
          if (!found) {
#pragma unroll
            for (int k = 0; k < kCache; k++)
              if (k == next) {
                if (key[k] >= 0) flush(key[k], hits[k]);
                key[k] = o;
                hits[k] = 1;
              }
            next = next == kCache - 1 ? 0 : next + 1;
          }
        }
#pragma unroll
    for (int k = 0; k < kCache; k++)
      if (key[k] >= 0) flush(key[k], hits[k]);
  }
}

// Cells whose shell has nsub > kThreadSub (the shells nearest the observer,
// up to 22^3 sub-points per cell at 1024^3): one wavefront per cell, one
// sub-point per lane, so the lanes of one atomic instruction mostly add into
// the same few bins.  Cells of the cube [lo, lo + side)^3 around the
// observer, taken by the wavefronts in turn.
__global__ __launch_bounds__(256) void k_imap_wave(ImapArgs a, const double *__restrict__ r0s,
                                                   const double *__restrict__ rfs, const int *__restrict__ nsub,
                                                   int lo_x, int lo_y, int lo_z, int side, float *__restrict__ data,
                                                   int *__restrict__ nadd) {
  const int lane = threadIdx.x % 64;
  const long nw = (long)gridDim.x * (blockDim.x / 64);
  const long ncell = (long)side * side * side;
  for (long w = (long)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; w < ncell; w += nw) {
    const int ix = lo_x + (int)(w % side), iy = lo_y + (int)((w / side) % side), iz = lo_z + (int)(w / ((long)side * side));
    CellPaint cp;
    if (!cell_paint(a, r0s, rfs, nsub, ix, iy, iz, cp)) continue;
    if (cp.ns <= kThreadSub) continue;   // k_imap's cell
    const int ns = cp.ns, nq = ns * ns * ns;
    for (int q = lane; q < nq; q += 64) {
      const int ixx = q % ns, iyy = (q / ns) % ns, izz = q / (ns * ns);
      const int64_t o = subpoint_bin(a, r0s, rfs, cp, izz, iyy, ixx);
      if (o >= 0) {
        atomicAdd(&data[o], cp.temp);
        atomicAdd(&nadd[o], 1);
      }
    }
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
  hipLaunchKernelGGL(k_imap, dim3((unsigned)((size_t)N * N)), dim3(N >= 256 ? 256 : N), 0, c->stream, a, dr0,
                     drf, dns, dmap, dnadd);
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
      hipLaunchKernelGGL(k_imap_wave, dim3(4096), dim3(256), 0, c->stream, a, dr0, drf, dns, lo[0], lo[1], lo[2],
                         side, dmap, dnadd);
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
