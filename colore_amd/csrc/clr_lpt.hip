// clr_lpt.hip -- first-order Lagrangian perturbation theory density
// (lpt_1, density.c:376-645, with pos_2_ngp / pos_2_cic / pos_2_tsc,
// density.c:37-175) on gfx950.
//
// The Gaussian field in the density grid is transformed forward (r2c, in
// place on the padded rows), each displacement component i k_a delta_k /
// (N^3 k^2) is written to its own half spectrum and transformed back (c2r, in
// place); the three FFT kinds are rocFFT's (through hipFFT): this is a
// secondary path, not the per-step hot path.  One thread per grid cell then
// moves the cell's particle by D(r) psi (positions in FP32 as the reference
// stores them, wrapped into the box) and adds its NGP, CIC or TSC weights to
// the zeroed density grid with FP32 atomics; a last pass subtracts 1.
#include "clr_internal.h"
#include <hipfft/hipfft.h>
#include <math.h>

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

struct MoveArgs {
  const float *ps0, *ps1, *ps2;
  float *dens;
  const double *r_arr, *d1;
  double idr, r_last, d1_last, ox, oy, oz, L;
  float dxf, iag;
  int N, ngx, interp;
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

// one thread per cell: the particle's position (density.c:481-505) and its
// mass assignment (density.c:37-175)
__global__ __launch_bounds__(256) void k_lpt_move(MoveArgs a) {
#pragma clang fp contract(off)
  const int N = a.N;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N * N * N) return;
  const int ix = (int)(t % N);
  const long long rw = t / N;
  const int iy = (int)(rw % N), iz = (int)(rw / N);
  const size_t ngx = (size_t)a.ngx;
  const size_t idx = (size_t)ix + ngx * ((size_t)iy + (size_t)N * iz);
  float xv[3];
  xv[0] = (float)((ix + 0.0) * (double)a.dxf - a.ox);
  xv[1] = (float)((iy + 0.0) * (double)a.dxf - a.oy);
  xv[2] = (float)((iz + 0.0) * (double)a.dxf - a.oz);
  const float r2 = xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2];
  const double dg = f_of_r(a.r_arr, a.d1, a.idr, a.r_last, sqrt((double)r2), 1.0, a.d1_last);
  const double obs[3] = {a.ox, a.oy, a.oz};
  const float disp[3] = {a.ps0[idx], a.ps1[idx], a.ps2[idx]};
  float p[3];
#pragma unroll
  for (int ax = 0; ax < 3; ax++) {
    float q = (float)(xv[ax] + dg * disp[ax] + obs[ax]);
    if (q < 0) q = (float)(q + a.L);
    if (q >= a.L) q = (float)(q - a.L);
    p[ax] = q;
  }
  float *__restrict__ g = a.dens;
  const size_t pl = (size_t)N * ngx;
  if (a.interp == 0) {
    int i0[3];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
      i0[ax] = wrapn((int)((double)(p[ax] * a.iag) + 0.5), N);
    }
    atomicAdd(g + i0[0] + ngx * i0[1] + pl * i0[2], 1.0f);
  } else if (a.interp == 1) {
    int i0[3], i1[3];
    float a0[3], a1[3];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
      const float xs = p[ax] * a.iag;
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
      const float xs = p[ax] * a.iag;
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

// density.c:609-622: delta = rho * 1 - 1 on the real cells
__global__ __launch_bounds__(256) void k_lpt_norm(float *__restrict__ g, int N, int ngx) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N * N * N) return;
  const int ix = (int)(t % N);
  const long long rw = t / N;
  const size_t idx = (size_t)ix + (size_t)ngx * (size_t)rw;
  g[idx] = (float)(g[idx] * 1.0f - 1.);
}

int fft_fail(clr_ctx *c, hipfftResult r, const char *what) {
  return clr_fail(c, CLR_ERR_HIP, "hipFFT %s failed (%d)", what, (int)r);
}

}  // namespace

int clr_launch_lpt(clr_ctx *c, int order, int interp) {
  if (!c->have_fields) return clr_fail(c, CLR_ERR_STATE, "clr_gaussian_fields must run first");
  if (c->density_done) return clr_fail(c, CLR_ERR_STATE, "the density grid is already transformed");
  if (order != 1) return clr_fail(c, CLR_ERR_ARG, "Density type %d not supported", order == 2 ? 2 : order);
  if (interp < 0 || interp > 2) return clr_fail(c, CLR_ERR_ARG, "Wrong interpolation type");
  const int N = c->N, ngx = c->ngx, nhp = ngx / 2, NH = N / 2 + 1;
  const size_t gbytes = (size_t)ngx * N * N * sizeof(float);
  float *ps[3] = {nullptr, nullptr, nullptr};
  for (int i = 0; i < 3; i++) {
    if (hipMalloc((void **)&ps[i], gbytes) != hipSuccess) {
      for (int j = 0; j < i; j++) (void)hipFree(ps[j]);
      return clr_fail(c, CLR_ERR_HIP, "out of device memory for the displacement grids");
    }
  }
  hipfftHandle fwd = 0, inv = 0;
  int rc = CLR_OK;
  // in place on the padded rows: real rows of ngx floats, complex rows of
  // ngx/2 values (>= N/2+1)
  long long n[3] = {N, N, N}, re[3] = {N, N, ngx}, ce[3] = {N, N, nhp};
  const long long rdist = (long long)N * N * ngx, cdist = (long long)N * N * nhp;
  size_t ws = 0;
  hipfftResult fr = hipfftCreate(&fwd);
  if (fr == HIPFFT_SUCCESS) fr = hipfftCreate(&inv);
  if (fr == HIPFFT_SUCCESS) fr = hipfftMakePlanMany64(fwd, 3, n, re, 1, rdist, ce, 1, cdist, HIPFFT_R2C, 1, &ws);
  if (fr == HIPFFT_SUCCESS) fr = hipfftMakePlanMany64(inv, 3, n, ce, 1, cdist, re, 1, rdist, HIPFFT_C2R, 1, &ws);
  if (fr == HIPFFT_SUCCESS) fr = hipfftSetStream(fwd, c->stream);
  if (fr == HIPFFT_SUCCESS) fr = hipfftSetStream(inv, c->stream);
  hipError_t e = hipSuccess;
  if (fr == HIPFFT_SUCCESS) e = hipEventRecord(c->ev0, c->stream);
  if (fr == HIPFFT_SUCCESS && e == hipSuccess)
    fr = hipfftExecR2C(fwd, (hipfftReal *)c->dens, (hipfftComplex *)c->dens);
  if (fr == HIPFFT_SUCCESS && e == hipSuccess) {
    const long long modes = (long long)N * N * NH;
    hipLaunchKernelGGL(k_lpt_disp, dim3((unsigned)((modes + 255) / 256)), dim3(256), 0, c->stream,
                       (const float2 *)c->dens, (float2 *)ps[0], (float2 *)ps[1], (float2 *)ps[2], N, nhp,
                       2 * M_PI / c->L, (double)N * ((double)N * N));
    e = hipGetLastError();
  }
  for (int i = 0; i < 3 && fr == HIPFFT_SUCCESS && e == hipSuccess; i++)
    fr = hipfftExecC2R(inv, (hipfftComplex *)ps[i], (hipfftReal *)ps[i]);
  if (fr == HIPFFT_SUCCESS && e == hipSuccess) e = hipMemsetAsync(c->dens, 0, gbytes, c->stream);
  if (fr == HIPFFT_SUCCESS && e == hipSuccess) {
    MoveArgs a{};
    a.ps0 = ps[0];
    a.ps1 = ps[1];
    a.ps2 = ps[2];
    a.dens = c->dens;
    a.r_arr = c->bg;
    a.d1 = c->bg + 2 * (size_t)c->na;
    a.idr = c->glob_idr;
    a.r_last = c->r_last;
    a.d1_last = c->h_bg_last[2];
    a.ox = c->obs[0];
    a.oy = c->obs[1];
    a.oz = c->obs[2];
    a.L = c->L;
    a.dxf = (float)(c->L / N);
    a.iag = (float)(N / c->L);
    a.N = N;
    a.ngx = ngx;
    a.interp = interp;
    const long long cells = (long long)N * N * N;
    hipLaunchKernelGGL(k_lpt_move, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, c->stream, a);
    e = hipGetLastError();
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_lpt_norm, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, c->stream, c->dens, N,
                         ngx);
      e = hipGetLastError();
    }
  }
  if (fr == HIPFFT_SUCCESS && e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (fr != HIPFFT_SUCCESS) rc = fft_fail(c, fr, "plan or transform");
  else if (e != hipSuccess) rc = clr_fail(c, CLR_ERR_HIP, "1LPT: %s", hipGetErrorString(e));
  if (fwd) hipfftDestroy(fwd);
  if (inv) hipfftDestroy(inv);
  for (int i = 0; i < 3; i++) (void)hipFree(ps[i]);
  if (rc) return rc;
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->stage_ms[ST_DENS] = ms;
  c->density_done = 1;
  c->norms_done = 0;
  return CLR_OK;
}
