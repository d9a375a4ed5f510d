// clr_poisson.h -- per-cell Poisson count of srcs_set_cartesian_single
// (srcs.c:136-187) with counter-based uniforms.
//
// Sampling algorithm (both the device and the CPU oracle): inversion by
// sequential search, k = min{k : u <= F(k; lambda)} with u = x/2^32 the
// cell's word of Philox(cell/4, stream_poisson(ipop), 0); for lambda > 500
// the count is the sum of m = ceil(lambda/500) inversions at lambda/m
// (Poisson is closed under sums), the i-th extra one (i >= 1) using word
// (i-1)%4 of Philox(cell, stream_poisson_extra(ipop), (i-1)/4).
//
// The device evaluates lambda and the search in FP32 first and accepts the
// result only when u lies further than a stated error bound from both CDF
// values that bracket it; otherwise it redoes the cell in FP64 with the
// reference's expression order.  The bound covers the FP32 error of lambda
// and of the running CDF, so the accepted count is the FP64 count.
#pragma once
#include "clr_device.h"

namespace clr {

constexpr double kPoisSplit = 500.0;

__host__ __device__ inline int poisson_inv_f64(double lam, double u) {
#pragma clang fp contract(off)
  double p = exp(-lam);
  double F = p;
  int k = 0;
  const int kmax = (int)(10 * lam) + 1000;
  while (u > F && k < kmax) {
    k++;
    p *= lam / k;
    F += p;
  }
  return k;
}

__host__ __device__ inline int poisson_exact(double lam, uint64_t cell, uint32_t stream_extra, uint32_t seed,
                                             uint32_t first_word) {
#pragma clang fp contract(off)
  if (!(lam > 0)) return 0;
  if (lam <= kPoisSplit) return poisson_inv_f64(lam, u01(first_word));
  const int m = (int)ceil(lam / kPoisSplit);
  const double lp = lam / m;
  int k = poisson_inv_f64(lp, u01(first_word));
  u32x4 r{};
  for (int i = 1; i < m; i++) {
    const int j = i - 1;
    if ((j & 3) == 0) r = draw(cell, stream_extra, (uint32_t)(j >> 2), seed);
    k += poisson_inv_f64(lp, u01(word(r, j & 3)));
  }
  return k;
}

#if defined(__HIPCC__)
// FP32 attempt; returns -1 when the decision is too close to call.
__device__ inline int poisson_fast_f32(float lam, uint32_t x) {
  if (!(lam > 0.f)) return -2;          // caller decides with FP64
  if (lam >= 64.f) return -1;
  const float uf = (float)x * 2.3283064365386963e-10f;
  float p = expf(-lam);
  float F = p, Fprev = 0.f;
  int k = 0;
  while (uf > F && k < 256) {
    k++;
    p *= lam / (float)k;
    Fprev = F;
    F += p;
  }
  const float eps = 2e-6f * (lam + 2.f) + 1e-6f;
  if (F - uf <= eps) return -1;
  if (k > 0 && uf - Fprev <= eps) return -1;
  return k;
}
#endif

}  // namespace clr
