// clr_device.h -- device helpers shared by the gfx950 kernels: Philox4x32-10,
// FP64 r-table interpolation, bias models and HEALPix pixel functions.
//
// The FP64 helpers are compiled with floating-point contraction off, so that
// a*b+c is rounded twice exactly as a plain C build on x86-64 rounds it; the
// integer decisions built on them (shell index, Poisson counts, pixel index)
// then agree bit for bit with a CPU evaluation of the same expressions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clr {

// Floats per x row of a real grid in device memory: the 2(N/2+1) of the
// in-place c2r layout rounded up to 32, so that every row starts on a
// 128-byte cache line (the host-side layout keeps 2(N/2+1); the grid copies
// convert).  Half of it is the row pitch of the complex half spectrum.
__host__ __device__ constexpr int grid_pitch(int N) { return (2 * (N / 2 + 1) + 31) / 32 * 32; }

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11).  Counter-based: one call
// maps (counter[4], key[2]) to four independent 32-bit words.
// ---------------------------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

__host__ __device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int i = 0; i < 10; i++) {
    if (i) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    // one 32x32->64 product per multiplier (a single v_mad_u64_u32 on gfx950)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
    c = u32x4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0};
  }
  return c;
}

// Stream identifiers (counter word z).  Counter word x/y = 64-bit site index,
// word w = draw index within the site.
//   modes    : site = kx + (N/2+1)(ky + N (kz mod N/2)); words (0,1) are the
//              (phase, modulus) draws of kz < N/2, words (2,3) those of
//              kz + N/2.
//   Poisson  : (cell = ix + N(iy + N iz)) the high 16 bits of the cell's
//              uniform are half cell%8 of the draw at site cell/8, the low 16
//              bits word 0 of site cell in stream 0x40000100+pop; extra
//              uniforms of the lambda > 500 split come from site = cell in
//              stream_poisson_extra (clr_poisson.h).
//   positions: site = cell; uniform j of the cell's sources is word j%4 of
//              draw j/4 (three per source, x y z).
constexpr uint32_t kKey1        = 0x436F4C52u;   // second key word
constexpr uint32_t kStreamModes = 0x00000000u;
__host__ __device__ inline uint32_t stream_poisson(int ipop) { return 0x100u + (uint32_t)ipop; }
__host__ __device__ inline uint32_t stream_poisson_extra(int ipop) { return 0x80000100u + (uint32_t)ipop; }
__host__ __device__ inline uint32_t stream_position(int ipop) { return 0x200u + (uint32_t)ipop; }

__host__ __device__ inline u32x4 draw(uint64_t site, uint32_t stream, uint32_t idx, uint32_t seed) {
  return philox4x32_10(u32x4{(uint32_t)site, (uint32_t)(site >> 32), stream, idx}, seed, kKey1);
}
__host__ __device__ inline uint32_t word(const u32x4 &r, int i) {
  return i == 0 ? r.x : (i == 1 ? r.y : (i == 2 ? r.z : r.w));
}
// u = x / 2^32 in [0,1): the mapping of gsl_rng_uniform for 32-bit generators.
__host__ __device__ inline double u01(uint32_t x) { return (double)x * 2.3283064365386963e-10; }

// ---------------------------------------------------------------------------
// r-tables: f_of_r_linear (cosmo.c:30-38).
// ---------------------------------------------------------------------------
struct BgTables {
  const double *r, *z, *d1, *v1, *pd, *ih;   // NA each
  double idr;
  double r_last;
  int na;
};

__host__ __device__ inline double f_of_r(const double *__restrict__ r_arr, const double *__restrict__ f,
                                         double idr, double r_last, double r, double f0, double ff) {
#pragma clang fp contract(off)
  if (r <= 0) return f0;
  if (r >= r_last) return ff;
  const int ir = (int)(r * idr);
  return f[ir] + (f[ir + 1] - f[ir]) * (r - r_arr[ir]) * idr;
}

// Geometry of cell (ix,iy,iz): offsets from the observer and radius, in the
// reference's FP64 expression order (density.c:1084-1094, srcs.c:158-167).
__host__ __device__ inline double cell_radius(int ix, int iy, int iz, double dx, const double *obs,
                                              double *x0, double *y0, double *z0) {
#pragma clang fp contract(off)
  const double zz = (iz + 0.0) * dx - obs[2];
  const double yy = (iy + 0.0) * dx - obs[1];
  const double xx = (ix + 0.0) * dx - obs[0];
  *x0 = xx; *y0 = yy; *z0 = zz;
  return sqrt(xx * xx + yy * yy + zz * zz);
}

// ---------------------------------------------------------------------------
// bias_model (common.h:414-431), selected per tracer.
// ---------------------------------------------------------------------------
__host__ __device__ inline double bias_model_f64(double d, double b, int model) {
#pragma clang fp contract(off)
  if (d <= -1) return 0;
  if (model == 1) {
    if (d < 0) return exp(b * d / (1 + d));
    return 1 + b * d;
  }
  if (model == 2) {
    if (1 + b * d > 0) return 1 + b * d;
    return 0;
  }
  return pow(1 + d, b);
}

// ---------------------------------------------------------------------------
// The FP32 lognormal transform of one cell (lognormalize, density.c:1070-1103)
// as the shell pass applies it in place and the surrogate-fused count pass
// applies it on the fly: cell offsets and radius in FP32, D(r) by linear
// interpolation of the FP32 growth table (D(0) = 1, D beyond the table = its
// last value, as get_bg), delta = exp(D (dG - D sigma^2 / 2)) - 1.  Every
// rounding is spelled out (explicit fma, contraction off), so both passes
// produce the same float for the same cell.
// ---------------------------------------------------------------------------
__device__ inline float coord_f32(int i, float dxf, float o) { return __builtin_fmaf((float)i, dxf, -o); }
__device__ inline float radius_f32(float x, float y, float z) {
  return __builtin_amdgcn_sqrtf(__builtin_fmaf(x, x, __builtin_fmaf(y, y, z * z)));
}
// Table position of radius rf: bin i0 and fraction tq, with the two edge flags.
struct TabPos {
  int i0;
  float tq;
  bool e0, e1;
};
__device__ inline TabPos tab_pos_f32(float rf, float idr_f, int na) {
  TabPos p;
  const float xq = rf * idr_f;
  p.e0 = !(rf > 0.f);
  p.e1 = xq >= (float)(na - 1);
  p.i0 = (p.e0 || p.e1) ? 0 : (int)xq;
  p.tq = xq - (float)p.i0;
  return p;
}
// D from the bin's (value, slope) pair.
__device__ inline float growth_at(const TabPos &p, float d0, float slope, float d1_last_f) {
  const float v = __builtin_fmaf(slope, p.tq, d0);
  return p.e0 ? 1.f : (p.e1 ? d1_last_f : v);
}
__device__ inline float growth_f32(const float *__restrict__ d1f, const TabPos &p, float d1_last_f) {
  const int j = p.i0 + 1;
  const float d0 = d1f[p.i0], d1 = d1f[j];
  return growth_at(p, d0, d1 - d0, d1_last_f);
}
// e^x to about 2 ulp: the hardware exp2 of x log2(e) rounded, times the
// first-order correction for that rounding (log2(e) = 1.44269502 + 1.925963e-8).
__device__ inline float exp_lognormal(float x) {
  const float t = x * 1.44269502f;
  const float e = __builtin_fmaf(x, 1.44269502f, -t) + x * 1.925963e-8f;
  const float r = __builtin_amdgcn_exp2f(t);
  return __builtin_fmaf(r, e * 0.693147181f, r);
}
__device__ inline float lognormal_f32(float dg, float D, float s2f) {
#pragma clang fp contract(off)
  const float hh = (0.5f * D) * s2f;
  const float arg = D * (dg - hh);
  return exp_lognormal(arg) - 1.f;
}

__device__ inline float bias_model_f32(float d, float b, int model) {
  if (d <= -1.f) return 0.f;
  if (model == 1) {
    if (d < 0.f) return __expf(b * d / (1.f + d));
    return 1.f + b * d;
  }
  if (model == 2) return fmaxf(1.f + b * d, 0.f);
  return exp2f(b * __log2f(1.f + d));
}

// ---------------------------------------------------------------------------
// HEALPix (Gorski et al. 2005, ApJ 622, 759): RING/NEST indices of a
// direction and pixel centres.
// ---------------------------------------------------------------------------
__host__ __device__ inline int64_t imod64(int64_t a, int64_t b) { int64_t v = a % b; return v >= 0 ? v : v + b; }

// RING index from (cos(theta), phi); he_ang2pix (healpix_extra.c:141-172).
__host__ __device__ inline int64_t ang2pix_ring(int64_t nside, double cth, double phi) {
#pragma clang fp contract(off)
  const double twopi = 6.283185307179586476925286766559;
  const double inv_halfpi = 0.6366197723675813430755350534900574;
  const double za = fabs(cth);
  double ph = phi;
  if (ph < 0 || ph >= twopi) {
    ph = fmod(ph, twopi);
    if (ph < 0) { ph += twopi; if (ph == twopi) ph = 0; }
  }
  const double tt = ph * inv_halfpi;
  if (za <= 2.0 / 3.0) {
    const double t1 = nside * (0.5 + tt), t2 = nside * cth * 0.75;
    const int64_t jp = (int64_t)(t1 - t2), jm = (int64_t)(t1 + t2);
    const int64_t ir = nside + 1 + jp - jm;
    const int64_t kshift = 1 - (ir & 1);
    int64_t ip = (jp + jm - nside + kshift + 1) / 2;
    ip = imod64(ip, 4 * nside);
    return nside * (nside - 1) * 2 + (ir - 1) * 4 * nside + ip;
  }
  const double tp = tt - (int)tt;
  const double tmp = nside * sqrt(3 * (1 - za));
  const int64_t jp = (int64_t)(tp * tmp), jm = (int64_t)((1.0 - tp) * tmp);
  const int64_t ir = jp + jm + 1;
  int64_t ip = (int64_t)(tt * ir);
  ip = imod64(ip, 4 * ir);
  if (cth > 0) return 2 * ir * (ir - 1) + ip;
  return 12 * nside * nside - 2 * ir * (ir + 1) + ip;
}

__host__ __device__ inline int64_t vec2pix_ring(int64_t nside, double x, double y, double z) {
#pragma clang fp contract(off)
  const double vlen = sqrt(x * x + y * y + z * z);
  return ang2pix_ring(nside, z / vlen, atan2(y, x));
}

__host__ __device__ inline int64_t isqrt64(int64_t v) {
  int64_t r = (int64_t)sqrt((double)v + 0.5);
  while (r * r > v) r--;
  while ((r + 1) * (r + 1) <= v) r++;
  return r;
}

__host__ __device__ inline int jrll(int f) { return f < 4 ? 2 : (f < 8 ? 3 : 4); }
__host__ __device__ inline int jpll(int f) {
  const int t[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};
  return t[f];
}

__host__ __device__ inline int64_t spread_bits(int64_t v) {   // bit i -> bit 2i
  uint64_t x = (uint64_t)v & 0xFFFFFFFFull;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return (int64_t)x;
}
__host__ __device__ inline int64_t compress_bits(int64_t v) { // bit 2i -> bit i
  uint64_t x = (uint64_t)v & 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
  x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
  return (int64_t)x;
}

__host__ __device__ inline void ring2xyf(int64_t nside, int64_t pix, int64_t *ix, int64_t *iy, int *face) {
  const int64_t ncap = 2 * nside * (nside - 1), npix = 12 * nside * nside, nl2 = 2 * nside;
  int64_t iring, iphi, kshift, nr;
  if (pix < ncap) {
    iring = (1 + isqrt64(1 + 2 * pix)) >> 1;
    iphi = (pix + 1) - 2 * iring * (iring - 1);
    kshift = 0; nr = iring;
    *face = (int)((iphi - 1) / nr);
  } else if (pix < npix - ncap) {
    const int64_t ip = pix - ncap;
    const int64_t tmp = ip / (4 * nside);
    iring = tmp + nside;
    iphi = ip - tmp * 4 * nside + 1;
    kshift = (iring + nside) & 1;
    nr = nside;
    const int64_t ire = tmp + 1, irm = nl2 + 1 - tmp;
    const int64_t ifm = (iphi - ire / 2 + nside - 1) / nside;
    const int64_t ifp = (iphi - irm / 2 + nside - 1) / nside;
    *face = (ifp == ifm) ? (int)(ifp | 4) : ((ifp < ifm) ? (int)ifp : (int)(ifm + 8));
  } else {
    const int64_t ip = npix - pix;
    iring = (1 + isqrt64(2 * ip - 1)) >> 1;
    iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
    kshift = 0; nr = iring;
    iring = 2 * nl2 - iring;
    *face = (int)(8 + (iphi - 1) / nr);
  }
  const int64_t irt = iring - ((2 + (*face >> 2)) * nside) + 1;
  int64_t ipt = 2 * iphi - jpll(*face) * nr - kshift - 1;
  if (ipt >= nl2) ipt -= 8 * nside;
  *ix = (ipt - irt) >> 1;
  *iy = (-ipt - irt) >> 1;
}

__host__ __device__ inline int64_t xyf2nest(int64_t nside, int64_t ix, int64_t iy, int face) {
  return (int64_t)face * nside * nside + spread_bits(ix) + (spread_bits(iy) << 1);
}

__host__ __device__ inline void nest2xyf(int64_t nside, int64_t pix, int64_t *ix, int64_t *iy, int *face) {
  const int64_t np = nside * nside;
  *face = (int)(pix / np);
  const int64_t ipf = pix - (int64_t)(*face) * np;
  *ix = compress_bits(ipf);
  *iy = compress_bits(ipf >> 1);
}

__host__ __device__ inline int64_t xyf2ring(int64_t nside, int64_t ix, int64_t iy, int face) {
  const int64_t nl4 = 4 * nside, npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
  const int64_t jr = (int64_t)jrll(face) * nside - ix - iy - 1;
  int64_t nr, kshift, nb;
  if (jr < nside) { nr = jr; nb = 2 * nr * (nr - 1); kshift = 0; }
  else if (jr > 3 * nside) { nr = nl4 - jr; nb = npix - 2 * (nr + 1) * nr; kshift = 0; }
  else { nr = nside; nb = ncap + (jr - nside) * nl4; kshift = (jr - nside) & 1; }
  int64_t jp = ((int64_t)jpll(face) * nr + ix - iy + 1 + kshift) / 2;
  if (jp > nl4) jp -= nl4; else if (jp < 1) jp += nl4;
  return nb + jp - 1;
}

__host__ __device__ inline int64_t ring2nest(int64_t nside, int64_t pix) {
  int64_t ix, iy; int f;
  ring2xyf(nside, pix, &ix, &iy, &f);
  return xyf2nest(nside, ix, iy, f);
}
__host__ __device__ inline int64_t nest2ring(int64_t nside, int64_t pix) {
  int64_t ix, iy; int f;
  nest2xyf(nside, pix, &ix, &iy, &f);
  return xyf2ring(nside, ix, iy, f);
}

// Pixel centre unit vector of NEST pixel `pix`.
__host__ __device__ inline void pix2vec_nest(int64_t nside, int64_t pix, double *v) {
#pragma clang fp contract(off)
  int64_t ix, iy; int face;
  nest2xyf(nside, pix, &ix, &iy, &face);
  const int64_t nl4 = 4 * nside;
  const double fact2 = 4.0 / (12.0 * (double)nside * (double)nside);
  const int64_t jr = (int64_t)jrll(face) * nside - ix - iy - 1;
  int64_t nr, kshift;
  double z;
  if (jr < nside) { nr = jr; z = 1 - (double)(nr * nr) * fact2; kshift = 0; }
  else if (jr > 3 * nside) { nr = nl4 - jr; z = (double)(nr * nr) * fact2 - 1; kshift = 0; }
  else { const double fact1 = (double)(nside << 1) * fact2; nr = nside; z = (double)(2 * nside - jr) * fact1; kshift = (jr - nside) & 1; }
  int64_t jp = ((int64_t)jpll(face) * nr + ix - iy + 1 + kshift) / 2;
  if (jp > nl4) jp -= nl4;
  if (jp < 1) jp += nl4;
  const double phi = ((double)jp - (double)(kshift + 1) * 0.5) * (1.5707963267948966 / (double)nr);
  const double st = sqrt((1.0 - z) * (1.0 + z));
  v[0] = st * cos(phi);
  v[1] = st * sin(phi);
  v[2] = z;
}

}  // namespace clr
