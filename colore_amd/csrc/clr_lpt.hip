// clr_lpt.hip -- Lagrangian perturbation theory densities (lpt_1,
// density.c:376-645, and lpt_2, density.c:646-1031, with pos_2_ngp /
// pos_2_cic / pos_2_tsc, density.c:37-175) on gfx950.
//
// The Gaussian field in the density grid is transformed forward (r2c, in
// place on the padded rows), each displacement component i k_a delta_k /
// (N^3 k^2) is written to its own half spectrum and transformed back (c2r, in
// place); 2LPT adds the six digradients, the second-order source and its
// displacement.  The transforms are the project's own (clr_fft3d: the
// Stockham engine of the field stage, run forward as the conjugate of the
// inverse transform of the conjugate).  One thread per grid cell then
// computes the cell's particle position (FP32 as the reference stores it,
// wrapped into the box), one thread per particle adds its NGP, CIC or TSC
// weights to the zeroed density grid with FP32 atomics, and a last pass
// subtracts 1.
#include "clr_internal.h"
#include <math.h>
#include <vector>

using namespace clr;

namespace {

// one thread per half-spectrum mode: psi_a = i k_a delta_k / (N^3 k^2)
__global__ __launch_bounds__(256) void k_lpt_disp(const float2 *__restrict__ dk, float2 *__restrict__ p0,
                                                  float2 *__restrict__ p1, float2 *__restrict__ p2, int N,
                                                  int nhp, double dkf, double fftnorm) {
#pragma clang fp contract(off)
  const int NH = N / 2 + 1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N * N * NH) return;
  const int kk = (int)(t % NH);
  const long long rw = t / NH;
  const int jj = (int)(rw % N), ii = (int)(rw / N);
  const size_t o = (size_t)kk + (size_t)nhp * ((size_t)jj + (size_t)N * ii);
  const double kz = 2 * ii <= N ? ii * dkf : -(N - ii) * dkf;
  const double ky = 2 * jj <= N ? jj * dkf : -(N - jj) * dkf;
  const double kx = kk * dkf;
  const double km2 = fftnorm * (kx * kx + ky * ky + kz * kz);
  const float2 d = dk[o];
  if (km2 <= 0) {
    p0[o] = p1[o] = p2[o] = make_float2(0.f, 0.f);
    return;
  }
  // i kv (re + i im) / km2 = (-kv im, kv re) / km2
  p0[o] = make_float2((float)(-kx * d.y / km2), (float)(kx * d.x / km2));
  p1[o] = make_float2((float)(-ky * d.y / km2), (float)(ky * d.x / km2));
  p2[o] = make_float2((float)(-kz * d.y / km2), (float)(kz * d.x / km2));
}

// 2LPT (density.c:683-722): the displacement spectra and the six
// "digradient" spectra i k_b psi_a of the stored (FP32) psi_a, for
// (a,b) = xx, xy, xz, yy, yz, zz; psi_z is written over delta_k itself
__global__ __launch_bounds__(256) void k_lpt2_grad(float2 *__restrict__ dk, float2 *__restrict__ p0,
                                                   float2 *__restrict__ p1, float2 *__restrict__ g0,
                                                   float2 *__restrict__ g1, float2 *__restrict__ g2,
                                                   float2 *__restrict__ g3, float2 *__restrict__ g4,
                                                   float2 *__restrict__ g5, int N, int nhp, double dkf,
                                                   double fftnorm) {
#pragma clang fp contract(off)
  const int NH = N / 2 + 1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N * N * NH) return;
  const int kk = (int)(t % NH);
  const long long rw = t / NH;
  const int jj = (int)(rw % N), ii = (int)(rw / N);
  const size_t o = (size_t)kk + (size_t)nhp * ((size_t)jj + (size_t)N * ii);
  const double kv[3] = {kk * dkf, 2 * jj <= N ? jj * dkf : -(N - jj) * dkf, 2 * ii <= N ? ii * dkf : -(N - ii) * dkf};
  const double km2 = fftnorm * (kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
  const float2 d = dk[o];
  float2 ps[3];
#pragma unroll
  for (int ax = 0; ax < 3; ax++)
    ps[ax] = km2 <= 0 ? make_float2(0.f, 0.f)
                      : make_float2((float)(-kv[ax] * d.y / km2), (float)(kv[ax] * d.x / km2));
  // i kb (re + i im) = (-kb im, kb re)
  auto ik = [](double kb, float2 v) { return make_float2((float)(-kb * v.y), (float)(kb * v.x)); };
  p0[o] = ps[0];
  p1[o] = ps[1];
  dk[o] = ps[2];
  g0[o] = ik(kv[0], ps[0]);
  g1[o] = ik(kv[1], ps[0]);
  g2[o] = ik(kv[2], ps[0]);
  g3[o] = ik(kv[1], ps[1]);
  g4[o] = ik(kv[2], ps[1]);
  g5[o] = ik(kv[2], ps[2]);
}

// density.c:732-752: the second-order source xx yy + xx zz + yy zz - xy^2 -
// xz^2 - yz^2 (FP32, the reference's order), into g5 on the real cells
__global__ __launch_bounds__(256) void k_lpt2_source(const float *__restrict__ g0, const float *__restrict__ g1,
                                                     const float *__restrict__ g2, const float *__restrict__ g3,
                                                     const float *__restrict__ g4, float *__restrict__ g5, int N,
                                                     int ngx) {
#pragma clang fp contract(off)
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N * N * N) return;
  const int ix = (int)(t % N);
  const size_t idx = (size_t)ix + (size_t)ngx * (size_t)(t / N);
  const float xx = g0[idx], xy = g1[idx], xz = g2[idx], yy = g3[idx], yz = g4[idx], zz = g5[idx];
  g5[idx] = xx * yy + xx * zz + yy * zz - xy * xy - xz * xz - yz * yz;
}

// density.c:778-805: psi2_a = -i k_a S_k / (N^3 k^2)
__global__ __launch_bounds__(256) void k_lpt2_disp2(const float2 *__restrict__ sk, float2 *__restrict__ q0,
                                                    float2 *__restrict__ q1, float2 *__restrict__ q2, int N,
                                                    int nhp, double dkf, double fftnorm) {
#pragma clang fp contract(off)
  const int NH = N / 2 + 1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N * N * NH) return;
  const int kk = (int)(t % NH);
  const long long rw = t / NH;
  const int jj = (int)(rw % N), ii = (int)(rw / N);
  const size_t o = (size_t)kk + (size_t)nhp * ((size_t)jj + (size_t)N * ii);
  const double kv[3] = {kk * dkf, 2 * jj <= N ? jj * dkf : -(N - jj) * dkf, 2 * ii <= N ? ii * dkf : -(N - ii) * dkf};
  const double km2 = fftnorm * (kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
  const float2 s = sk[o];
  float2 *q[3] = {q0, q1, q2};
#pragma unroll
  for (int ax = 0; ax < 3; ax++)
    // -i kv (re + i im) / km2 = (kv im, -kv re) / km2
    q[ax][o] = km2 <= 0 ? make_float2(0.f, 0.f)
                        : make_float2((float)(kv[ax] * s.y / km2), (float)(-kv[ax] * s.x / km2));
}

struct PosArgs {
  const float *d0, *d1, *d2;      // first-order displacement
  const float *s0, *s1, *s2;      // second-order displacement (2LPT) or null
  float *p0, *p1, *p2;            // positions out (may alias d*)
  const double *r_arr, *dd1, *dd2;
  double idr, r_last, d1_last, d2_first, d2_last, ox, oy, oz, L;
  float dxf;
  int N, ngx;
};

// The reference wraps once (density.c:47-50, 80-85, 125-130), enough for
// positions inside [0, L); the modulo keeps any other value on the grid.
__device__ inline int wrapn(int v, int N) {
  if (v < 0 || v >= N) {
    v %= N;
    if (v < 0) v += N;
  }
  return v;
}

// one thread per cell: the particle's position x + D1 psi1 (+ D2 psi2) +
// x_obs, FP32, wrapped into [0, L) (density.c:481-505, 852-875)
__global__ __launch_bounds__(256) void k_lpt_pos(PosArgs a) {
#pragma clang fp contract(off)
  const int N = a.N;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N * N * N) return;
  const int ix = (int)(t % N);
  const long long rw = t / N;
  const int iy = (int)(rw % N), iz = (int)(rw / N);
  const size_t idx = (size_t)ix + (size_t)a.ngx * (size_t)rw;
  float xv[3];
  xv[0] = (float)((ix + 0.0) * (double)a.dxf - a.ox);
  xv[1] = (float)((iy + 0.0) * (double)a.dxf - a.oy);
  xv[2] = (float)((iz + 0.0) * (double)a.dxf - a.oz);
  const float r2 = xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2];
  const double r = sqrt((double)r2);
  const double dg = f_of_r(a.r_arr, a.dd1, a.idr, a.r_last, r, 1.0, a.d1_last);
  const double obs[3] = {a.ox, a.oy, a.oz};
  const float d[3] = {a.d0[idx], a.d1[idx], a.d2[idx]};
  float p[3];
  if (a.s0) {
    const double d2g = f_of_r(a.r_arr, a.dd2, a.idr, a.r_last, r, a.d2_first, a.d2_last);
    const float s[3] = {a.s0[idx], a.s1[idx], a.s2[idx]};
#pragma unroll
    for (int ax = 0; ax < 3; ax++) p[ax] = (float)(xv[ax] + dg * d[ax] + d2g * s[ax] + obs[ax]);
  } else {
#pragma unroll
    for (int ax = 0; ax < 3; ax++) p[ax] = (float)(xv[ax] + dg * d[ax] + obs[ax]);
  }
#pragma unroll
  for (int ax = 0; ax < 3; ax++) {
    if (p[ax] < 0) p[ax] = (float)(p[ax] + a.L);
    if (p[ax] >= a.L) p[ax] = (float)(p[ax] - a.L);
  }
  a.p0[idx] = p[0];
  a.p1[idx] = p[1];
  a.p2[idx] = p[2];
}

// one thread per particle: NGP, CIC or TSC weights (density.c:37-175) added
// to the zeroed grid with FP32 atomics
__global__ __launch_bounds__(256) void k_lpt_deposit(const float *__restrict__ px, const float *__restrict__ py,
                                                     const float *__restrict__ pz, float *__restrict__ g, int N,
                                                     int ngxi, float iag, int interp) {
#pragma clang fp contract(off)
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N * N * N) return;
  const size_t ngx = (size_t)ngxi;
  const size_t idx = (size_t)(t % N) + ngx * (size_t)(t / N);
  const float p[3] = {px[idx], py[idx], pz[idx]};
  const size_t pl = (size_t)N * ngx;
  if (interp == 0) {
    int i0[3];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) i0[ax] = wrapn((int)((double)(p[ax] * iag) + 0.5), N);
    atomicAdd(g + i0[0] + ngx * i0[1] + pl * i0[2], 1.0f);
  } else if (interp == 1) {
    int i0[3], i1[3];
    float a0[3], a1[3];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
      const float xs = p[ax] * iag;
      const int c = (int)xs;
      a1[ax] = xs - c;
      a0[ax] = 1 - a1[ax];
      i0[ax] = wrapn(c, N);
      i1[ax] = wrapn(c + 1, N);
    }
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int bx = c & 1, by = (c >> 1) & 1, bz = c >> 2;
      const float w = (bx ? a1[0] : a0[0]) * (by ? a1[1] : a0[1]) * (bz ? a1[2] : a0[2]);
      atomicAdd(g + (bx ? i1[0] : i0[0]) + ngx * (by ? i1[1] : i0[1]) + pl * (bz ? i1[2] : i0[2]), w);
    }
  } else {
    int ic[3][3];
    float wt[3][3];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
      const float xs = p[ax] * iag;
      const int i0 = (int)floorf((float)((double)xs + 0.5));
      const float d = xs - i0;
      wt[ax][0] = (float)(0.5 * (0.5 - d) * (0.5 - d));
      wt[ax][2] = (float)(0.5 * (0.5 + d) * (0.5 + d));
      wt[ax][1] = (float)(0.75 - d * d);
      ic[ax][0] = wrapn(i0 - 1, N);
      ic[ax][1] = wrapn(i0, N);
      ic[ax][2] = wrapn(i0 + 1, N);
    }
    for (int sz = 0; sz < 3; sz++)
      for (int sy = 0; sy < 3; sy++)
#pragma unroll
        for (int sx = 0; sx < 3; sx++)
          atomicAdd(g + ic[0][sx] + ngx * ic[1][sy] + pl * ic[2][sz], wt[0][sx] * wt[1][sy] * wt[2][sz]);
  }
}

// The same CIC / TSC assignment with the weights first summed in LDS: a
// block takes a 16^3 tile of particles (16 per thread) and holds the grid
// region of the tile grown by 2 cells on every side (20^3 floats, 32 KB, so
// that several blocks share a CU); weights landing in the region go to LDS,
// the ones further out straight to the grid, and the region is then added to
// the grid once (about 2 global atomics per particle instead of 8 for CIC, 27
// for TSC).  Tiles of 32 x 16 x 16 with 4-cell margins (90 KB, two blocks
// per CU) took 8 ms more for a 1024^3 CIC deposit; tiles from 8^3 to
// 32 x 8 x 8 with 1- to 3-cell margins came within 1 ms of this one.  Needs N >= 64 and
// N a multiple of 16, so that the region never wraps onto itself.
constexpr int kDTX = 16, kDTY = 16, kDTZ = 16, kDHalo = 2;
constexpr int kDRX = kDTX + 2 * kDHalo, kDRY = kDTY + 2 * kDHalo, kDRZ = kDTZ + 2 * kDHalo;
constexpr size_t kDepositLds = (size_t)kDRX * kDRY * kDRZ * sizeof(float);

__global__ __launch_bounds__(256) void k_lpt_deposit_tiled(const float *__restrict__ px, const float *__restrict__ py,
                                                           const float *__restrict__ pz, float *__restrict__ g,
                                                           int N, int ngxi, float iag, int interp) {
#pragma clang fp contract(off)
  extern __shared__ float reg[];
  const size_t ngx = (size_t)ngxi, pl = (size_t)N * ngx;
  const int ntx = N / kDTX, nty = N / kDTY;
  const int b = blockIdx.x;
  const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
  const int ox = tx * kDTX - kDHalo, oy = ty * kDTY - kDHalo, oz = tz * kDTZ - kDHalo;
  for (int i = threadIdx.x; i < kDRX * kDRY * kDRZ; i += blockDim.x) reg[i] = 0.f;
  __syncthreads();
  // relative coordinate of cell c in the region along one axis (c in [0, N))
  auto rel = [N](int c, int o) {
    int d = c - o;
    if (d < 0) d += N;
    if (d >= N) d -= N;
    return d;
  };
  auto add = [&](int cx, int cy, int cz, float w) {
    const int rx = rel(cx, ox), ry = rel(cy, oy), rz = rel(cz, oz);
    if (rx < kDRX && ry < kDRY && rz < kDRZ) atomicAdd(&reg[rx + kDRX * (ry + kDRY * rz)], w);
    else atomicAdd(g + cx + ngx * cy + pl * cz, w);
  };
  for (int q = threadIdx.x; q < kDTX * kDTY * kDTZ; q += blockDim.x) {
    const int lx = q % kDTX, ly = (q / kDTX) % kDTY, lz = q / (kDTX * kDTY);
    const size_t idx = (size_t)(tx * kDTX + lx) + ngx * ((size_t)(ty * kDTY + ly) + (size_t)N * (tz * kDTZ + lz));
    const float p[3] = {px[idx], py[idx], pz[idx]};
    if (interp == 1) {
      int i0[3], i1[3];
      float a0[3], a1[3];
#pragma unroll
      for (int ax = 0; ax < 3; ax++) {
        const float xs = p[ax] * iag;
        const int c = (int)xs;
        a1[ax] = xs - c;
        a0[ax] = 1 - a1[ax];
        i0[ax] = wrapn(c, N);
        i1[ax] = wrapn(c + 1, N);
      }
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const int bx = c & 1, by = (c >> 1) & 1, bz = c >> 2;
        add(bx ? i1[0] : i0[0], by ? i1[1] : i0[1], bz ? i1[2] : i0[2],
            (bx ? a1[0] : a0[0]) * (by ? a1[1] : a0[1]) * (bz ? a1[2] : a0[2]));
      }
    } else {
      int ic[3][3];
      float wt[3][3];
#pragma unroll
      for (int ax = 0; ax < 3; ax++) {
        const float xs = p[ax] * iag;
        const int i0 = (int)floorf((float)((double)xs + 0.5));
        const float d = xs - i0;
        wt[ax][0] = (float)(0.5 * (0.5 - d) * (0.5 - d));
        wt[ax][2] = (float)(0.5 * (0.5 + d) * (0.5 + d));
        wt[ax][1] = (float)(0.75 - d * d);
        ic[ax][0] = wrapn(i0 - 1, N);
        ic[ax][1] = wrapn(i0, N);
        ic[ax][2] = wrapn(i0 + 1, N);
      }
      for (int sz = 0; sz < 3; sz++)
        for (int sy = 0; sy < 3; sy++)
#pragma unroll
          for (int sx = 0; sx < 3; sx++)
            add(ic[0][sx], ic[1][sy], ic[2][sz], wt[0][sx] * wt[1][sy] * wt[2][sz]);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kDRX * kDRY * kDRZ; i += blockDim.x) {
    const float v = reg[i];
    if (v == 0.f) continue;
    const int rx = i % kDRX, ry = (i / kDRX) % kDRY, rz = i / (kDRX * kDRY);
    const int cx = wrapn(ox + rx, N), cy = wrapn(oy + ry, N), cz = wrapn(oz + rz, N);
    atomicAdd(g + cx + ngx * cy + pl * cz, v);
  }
}

// density.c:609-622: delta = rho * 1 - 1 on the real cells
__global__ __launch_bounds__(256) void k_lpt_norm(float *__restrict__ g, int N, int ngx) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N * N * N) return;
  const int ix = (int)(t % N);
  const long long rw = t / N;
  const size_t idx = (size_t)ix + (size_t)ngx * (size_t)rw;
  g[idx] = (float)(g[idx] * 1.0f - 1.);
}


}  // namespace

int clr_launch_lpt(clr_ctx *c, int order, int interp, float *pos_out) {
  if (!c->have_fields) return clr_fail(c, CLR_ERR_STATE, "clr_gaussian_fields must run first");
  if (c->density_done) return clr_fail(c, CLR_ERR_STATE, "the density grid is already transformed");
  if (order != 1 && order != 2) return clr_fail(c, CLR_ERR_ARG, "Density type %d not supported", order);
  if (order == 2 && !c->d2) return clr_fail(c, CLR_ERR_STATE, "2LPT needs the D2 growth table (clr_tables.growth_d2_arr)");
  if (interp < 0 || interp > 2) return clr_fail(c, CLR_ERR_ARG, "Wrong interpolation type");
  const int N = c->N, ngx = c->ngx, nhp = ngx / 2, NH = N / 2 + 1;
  const size_t gbytes = (size_t)ngx * N * N * sizeof(float);
  // order 1: psi_x, psi_y, psi_z; order 2: psi_x, psi_y (psi_z over delta) and
  // the six digradients, whose first three become psi2 and last three hold
  // the positions
  const int nbuf = order == 1 ? 3 : 8;
  float *ps[8] = {nullptr};
  for (int i = 0; i < nbuf; i++) {
    if (hipMalloc((void **)&ps[i], gbytes) != hipSuccess) {
      for (int j = 0; j < i; j++) (void)hipFree(ps[j]);
      return clr_fail(c, CLR_ERR_HIP, "out of device memory for the LPT grids");
    }
  }
  int rc = CLR_OK, fr = CLR_OK;   // fr: the first transform error
  hipError_t e = hipSuccess;
  const long long modes = (long long)N * N * NH, cells = (long long)N * N * N;
  const dim3 gm((unsigned)((modes + 255) / 256)), gc((unsigned)((cells + 255) / 256)), blk(256);
  const double dkf = 2 * M_PI / c->L, fftnorm = (double)N * ((double)N * N);
  auto ok = [&]() { return fr == CLR_OK && e == hipSuccess; };
  // in place on the padded rows: real rows of ngx floats, complex rows of
  // ngx/2 values (>= N/2+1), unnormalised like FFTW's plans
  auto r2c = [&](float *g) { if (ok()) fr = clr_fft3d(c, g, 1); };
  auto c2r = [&](float *g) { if (ok()) fr = clr_fft3d(c, g, 0); };
  if (ok()) e = hipEventRecord(c->ev0, c->stream);
  r2c(c->dens);
  const float *dsp[3], *sec[3] = {nullptr, nullptr, nullptr};
  float *pos[3];
  if (order == 1) {
    if (ok()) {
      hipLaunchKernelGGL(k_lpt_disp, gm, blk, 0, c->stream, (const float2 *)c->dens, (float2 *)ps[0],
                         (float2 *)ps[1], (float2 *)ps[2], N, nhp, dkf, fftnorm);
      e = hipGetLastError();
    }
    for (int i = 0; i < 3; i++) c2r(ps[i]);
    for (int i = 0; i < 3; i++) dsp[i] = pos[i] = ps[i];
  } else {
    float **g = ps + 2;   // g[0..5]: xx, xy, xz, yy, yz, zz
    if (ok()) {
      hipLaunchKernelGGL(k_lpt2_grad, gm, blk, 0, c->stream, (float2 *)c->dens, (float2 *)ps[0], (float2 *)ps[1],
                         (float2 *)g[0], (float2 *)g[1], (float2 *)g[2], (float2 *)g[3], (float2 *)g[4],
                         (float2 *)g[5], N, nhp, dkf, fftnorm);
      e = hipGetLastError();
    }
    for (int i = 0; i < 6; i++) c2r(g[i]);
    if (ok()) {
      hipLaunchKernelGGL(k_lpt2_source, gc, blk, 0, c->stream, g[0], g[1], g[2], g[3], g[4], g[5], N, ngx);
      e = hipGetLastError();
    }
    r2c(g[5]);
    if (ok()) {
      hipLaunchKernelGGL(k_lpt2_disp2, gm, blk, 0, c->stream, (const float2 *)g[5], (float2 *)g[0], (float2 *)g[1],
                         (float2 *)g[2], N, nhp, dkf, fftnorm);
      e = hipGetLastError();
    }
    c2r(ps[0]);
    c2r(ps[1]);
    c2r(c->dens);
    for (int i = 0; i < 3; i++) c2r(g[i]);
    dsp[0] = ps[0];
    dsp[1] = ps[1];
    dsp[2] = c->dens;
    for (int i = 0; i < 3; i++) {
      sec[i] = g[i];
      pos[i] = g[3 + i];
    }
  }
  if (ok()) {
    PosArgs a{};
    a.d0 = dsp[0];
    a.d1 = dsp[1];
    a.d2 = dsp[2];
    a.s0 = sec[0];
    a.s1 = sec[1];
    a.s2 = sec[2];
    a.p0 = pos[0];
    a.p1 = pos[1];
    a.p2 = pos[2];
    a.r_arr = c->bg;
    a.dd1 = c->bg + 2 * (size_t)c->na;
    a.dd2 = c->d2;
    a.idr = c->glob_idr;
    a.r_last = c->r_last;
    a.d1_last = c->h_bg_last[2];
    a.d2_first = c->d2_first;
    a.d2_last = c->d2_last;
    a.ox = c->obs[0];
    a.oy = c->obs[1];
    a.oz = c->obs[2];
    a.L = c->L;
    a.dxf = (float)(c->L / N);
    a.N = N;
    a.ngx = ngx;
    hipLaunchKernelGGL(k_lpt_pos, gc, blk, 0, c->stream, a);
    e = hipGetLastError();
  }
  if (ok() && pos_out) {
    // the particle positions (write_lpt, io.c:619-695): x, y, z per particle in cell order
    std::vector<float> plane((size_t)N * N * N);
    for (int ax = 0; ax < 3 && ok(); ax++) {
      e = hipMemcpy2DAsync(plane.data(), (size_t)N * sizeof(float), pos[ax], (size_t)ngx * sizeof(float),
                           (size_t)N * sizeof(float), (size_t)N * N, hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e == hipSuccess)
        for (size_t i = 0; i < plane.size(); i++) pos_out[3 * i + ax] = plane[i];
    }
  }
  if (ok()) e = hipMemsetAsync(c->dens, 0, gbytes, c->stream);
  if (ok()) {
    if (interp != 0 && N >= 64 && N % kDTX == 0) {
      e = hipFuncSetAttribute((const void *)k_lpt_deposit_tiled, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kDepositLds);
      const unsigned tiles = (unsigned)((N / kDTX) * (N / kDTY) * (N / kDTZ));
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_lpt_deposit_tiled, dim3(tiles), blk, kDepositLds, c->stream, pos[0], pos[1], pos[2],
                           c->dens, N, ngx, (float)(N / c->L), interp);
    } else {
      hipLaunchKernelGGL(k_lpt_deposit, gc, blk, 0, c->stream, pos[0], pos[1], pos[2], c->dens, N, ngx,
                         (float)(N / c->L), interp);
    }
    if (e == hipSuccess) e = hipGetLastError();
  }
  if (ok()) {
    hipLaunchKernelGGL(k_lpt_norm, gc, blk, 0, c->stream, c->dens, N, ngx);
    e = hipGetLastError();
  }
  if (ok()) e = hipEventRecord(c->ev1, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (fr != CLR_OK) rc = fr;   // clr_fft3d has set the message
  else if (e != hipSuccess) rc = clr_fail(c, CLR_ERR_HIP, "LPT: %s", hipGetErrorString(e));
  for (int i = 0; i < nbuf; i++) (void)hipFree(ps[i]);
  if (rc) return rc;
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->stage_ms[ST_DENS] = ms;
  c->density_done = 1;
  c->norms_done = 0;
  return CLR_OK;
}
