// clr_poisson.h -- per-cell Poisson count of srcs_set_cartesian_single
// (srcs.c:136-187) with counter-based uniforms.
//
// The cell's uniform is u = (hi * 2^16 + lo) / 2^32 where
//   hi = 16-bit half (cell % 8) of Philox(cell/8, stream_poisson(ipop), 0)
//   lo = low 16 bits of word 0 of Philox(cell, stream_poisson_lo(ipop), 0).
// Sampling (device and CPU oracle): inversion by sequential search,
// k = min{k : u <= F(k; lambda)}; for lambda > 500 the count is the sum of
// m = ceil(lambda/500) inversions at lambda/m (Poisson is closed under sums),
// the i-th extra one (i >= 1) using word (i-1)%4 of
// Philox(cell, stream_poisson_extra(ipop), (i-1)/4).
//
// The device decides most cells from hi alone: u < (hi+1)/2^16, so
// (hi+1)/2^16 <= exp(-lambda) already means k = 0, and lo is drawn only when
// it is not.  Lambda and the search run in FP32 first; a result is accepted
// only when u lies further than a stated error bound from both CDF values
// that bracket it, otherwise the cell is redone in FP64 in the reference's
// expression order.  The bound covers the FP32 error of lambda, of exp and of
// the running CDF, so every accepted count is the FP64 count.
#pragma once
#include "clr_device.h"

namespace clr {

constexpr double kPoisSplit = 500.0;
// Largest Poisson mean sampled: a mean that is not finite or above 2^20 (a
// bias model with b < 0 near delta = -1 gives such means) has no count here;
// poisson_exact returns -1 and the library reports the cell as an error.
constexpr double kPoisMax = 1048576.0;

__host__ __device__ inline uint32_t stream_poisson_lo(int ipop) { return 0x40000100u + (uint32_t)ipop; }

__host__ __device__ inline uint32_t poisson_hi16(const u32x4 &r, uint64_t cell) {
  const int h = (int)(cell & 7);
  return (word(r, h >> 1) >> ((h & 1) * 16)) & 0xFFFFu;
}
__host__ __device__ inline uint32_t poisson_lo16(uint64_t cell, int ipop, uint32_t seed) {
  return draw(cell, stream_poisson_lo(ipop), 0u, seed).x & 0xFFFFu;
}

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

// FP64 count from the full 32-bit uniform x = hi << 16 | lo; -1 for a mean
// that is not finite or above kPoisMax.
__host__ __device__ inline int poisson_exact(double lam, uint64_t cell, uint32_t stream_extra, uint32_t seed,
                                             uint32_t x) {
#pragma clang fp contract(off)
  if (!(lam > 0)) return 0;
  if (!(lam <= kPoisMax)) return -1;
  if (lam <= kPoisSplit) return poisson_inv_f64(lam, u01(x));
  const int m = (int)ceil(lam / kPoisSplit);
  const double lp = lam / m;
  int k = poisson_inv_f64(lp, u01(x));
  u32x4 r{};
  for (int i = 1; i < m; i++) {
    const int j = i - 1;
    if ((j & 3) == 0) r = draw(cell, stream_extra, (uint32_t)(j >> 2), seed);
    k += poisson_inv_f64(lp, u01(word(r, j & 3)));
  }
  return k;
}

#if defined(__HIPCC__)
// The FP32 fast path.  `lam` is an FP32 estimate of the reference's FP64
// lambda with |lam - lambda| <= err (the caller's bound); e = __expf(-lam).
// __expf is within (2e-7 + 1e-7 lam) relative of exp(-lam), and
// |exp(-lam) - exp(-lambda)| <= err exp(-min); every margin below adds those
// terms and the FP32 rounding of the running sums with room to spare.

// k = 0 decided from the high 16 bits: u < (hi+1)/2^16 <= exp(-lambda).
__device__ inline bool poisson_zero_from_hi(float lam, float e, float err, uint32_t hi) {
  return (float)(hi + 1) * 1.52587890625e-05f <= e * (1.f - 1e-6f - 2e-7f * lam - 1.1f * err);
}

// Inversion with the full uniform; -1 when u lies within the error bound of
// a CDF value that brackets it (or lambda is outside (0, 64)).
__device__ inline int poisson_fast_f32(float lam, float e, float err, uint32_t x) {
  if (!(lam > 0.f) || lam >= 64.f) return -1;
  const float uf = (float)x * 2.3283064365386963e-10f;
  float p = e;
  float F = p, Fprev = 0.f;
  int k = 0;
  while (uf > F && k < 256) {
    k++;
    p *= lam / (float)k;
    Fprev = F;
    F += p;
  }
  const float eps = 2e-6f + 3e-7f * lam + 4e-7f * (float)k + 1.1f * err;
  if (F - uf <= eps) return -1;
  if (k > 0 && uf - Fprev <= eps) return -1;
  return k;
}
#endif

}  // namespace clr
