// clr_fft.h -- register/LDS Stockham engine for the batched inverse DFTs of the
// 3-D complex-to-real transform (sign +1, unnormalised, as fftw_plan_dft_c2r_3d
// in fourier.c:81-102).
//
// A line of length L is held by P threads, E = L/P complex values each.  Each
// pass of radix R does E/R butterflies per thread; the first pass reads
// element t + b*P + m*(L/R) straight from HBM and the last pass writes element
// t + b*P + q*(L/R) straight back, so only the passes in between go through
// LDS.
#pragma once
#include <hip/hip_runtime.h>

namespace clr {

__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ inline float2 cmuli(float2 a) { return make_float2(-a.y, a.x); }   // i*a

// Compile-time cos/sin of 2*pi*k/R (Taylor series on the reduced argument).
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double ctaylor_cos(double x) {
  double s = 1, t = 1;
  for (int n = 1; n < 40; n++) { t *= -x * x / ((2.0 * n - 1) * (2.0 * n)); s += t; }
  return s;
}
constexpr double ctaylor_sin(double x) {
  double s = x, t = x;
  for (int n = 1; n < 40; n++) { t *= -x * x / ((2.0 * n) * (2.0 * n + 1)); s += t; }
  return s;
}
template <int R> struct TwTab {
  float c[R], s[R];
  constexpr TwTab() : c(), s() {
    for (int k = 0; k < R; k++) {
      double a = 2 * kPi * k / R;
      if (a > kPi) a -= 2 * kPi;
      c[k] = (float)ctaylor_cos(a);
      s[k] = (float)ctaylor_sin(a);
    }
  }
};
template <int R> __device__ inline float2 wconst(int k) {
  constexpr TwTab<R> t{};
  return make_float2(t.c[k % R], t.s[k % R]);
}

// In-place inverse DFT of size R: a[q] <- sum_m a[m] exp(+2 pi i m q / R).
template <int R> struct Dft;
template <> struct Dft<1> { __device__ static void run(float2 *) {} };
template <> struct Dft<2> {
  __device__ static void run(float2 *a) {
    const float2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};
template <> struct Dft<3> {
  __device__ static void run(float2 *a) {
    const float h = 0.86602540378443864676f;
    const float2 s = cadd(a[1], a[2]), d = csub(a[1], a[2]);
    const float2 m = make_float2(a[0].x - 0.5f * s.x, a[0].y - 0.5f * s.y);
    const float2 r = make_float2(-h * d.y, h * d.x);   // i*h*d
    a[0] = cadd(a[0], s);
    a[1] = cadd(m, r);
    a[2] = csub(m, r);
  }
};
template <> struct Dft<4> {
  __device__ static void run(float2 *a) {
    const float2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    const float2 t2 = cadd(a[1], a[3]), t3 = cmuli(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};
// R = R1*R2 with m = m2 + R2*m1, q = q1 + R1*q2.
template <int R1, int R2> struct DftSplit {
  __device__ static void run(float2 *a) {
    constexpr int R = R1 * R2;
    float2 B[R2][R1];
#pragma unroll
    for (int m2 = 0; m2 < R2; m2++) {
      float2 t[R1];
#pragma unroll
      for (int m1 = 0; m1 < R1; m1++) t[m1] = a[m2 + R2 * m1];
      Dft<R1>::run(t);
#pragma unroll
      for (int q1 = 0; q1 < R1; q1++) B[m2][q1] = (m2 * q1 == 0) ? t[q1] : cmul(t[q1], wconst<R>(m2 * q1));
    }
#pragma unroll
    for (int q1 = 0; q1 < R1; q1++) {
      float2 t[R2];
#pragma unroll
      for (int m2 = 0; m2 < R2; m2++) t[m2] = B[m2][q1];
      Dft<R2>::run(t);
#pragma unroll
      for (int q2 = 0; q2 < R2; q2++) a[q1 + R1 * q2] = t[q2];
    }
  }
};
template <> struct Dft<6> : DftSplit<2, 3> {};
template <> struct Dft<8> : DftSplit<2, 4> {};
template <> struct Dft<12> : DftSplit<3, 4> {};
template <> struct Dft<16> : DftSplit<4, 4> {};

// LDS views.  Column-interleaved: element e of column c at s[e*T + c] (lanes
// of a wavefront run over consecutive columns).  Row: element e at s[e].
template <int T> struct LdsCol {
  float2 *s; int c;
  __device__ float2 ld(int e) const { return s[e * T + c]; }
  __device__ void st(int e, float2 v) const { s[e * T + c] = v; }
};
struct LdsRow {
  float2 *s;
  __device__ float2 ld(int e) const { return s[e + (e >> 5)]; }
  __device__ void st(int e, float2 v) const { s[e + (e >> 5)] = v; }
};
template <int L> constexpr int lds_row_len() { return L + (L >> 5); }

// Stockham passes.  Ns = product of the radices already applied; tw holds
// exp(+2 pi i n / NT) for n < NT, NT a multiple of L.
template <int L, int P, int NT, int Ns, int R, int... Rest> struct Stages {
  template <class IO>
  __device__ static void run(float2 *v, int t, const IO &io, const float2 *__restrict__ tw) {
    constexpr int E = L / P;
    constexpr int NB = E / R;
    static_assert(E % R == 0, "radix must divide elements per thread");
    if constexpr (Ns > 1) {
#pragma unroll
      for (int b = 0; b < NB; b++) {
        const int j = t + b * P;
#pragma unroll
        for (int m = 0; m < R; m++) v[b * R + m] = io.ld(j + m * (L / R));
      }
    }
#pragma unroll
    for (int b = 0; b < NB; b++) {
      if constexpr (Ns > 1) {
        const int j = t + b * P;
        const int k = j % Ns;
#pragma unroll
        for (int m = 1; m < R; m++) v[b * R + m] = cmul(v[b * R + m], tw[(m * k) * (NT / (Ns * R))]);
      }
      Dft<R>::run(&v[b * R]);
    }
    if constexpr (sizeof...(Rest) > 0) {
      __syncthreads();
#pragma unroll
      for (int b = 0; b < NB; b++) {
        const int j = t + b * P;
        const int k = j % Ns;
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int q = 0; q < R; q++) io.st(base + q * Ns, v[b * R + q]);
      }
      __syncthreads();
      Stages<L, P, NT, Ns * R, Rest...>::run(v, t, io, tw);
    }
  }
};

template <int... Rs> struct RadixList;
template <int R0, int... Rs> struct RadixList<R0, Rs...> {
  static constexpr int first = R0;
  static constexpr int last = RadixList<Rs...>::last;
  static constexpr int product = R0 * RadixList<Rs...>::product;
};
template <int R0> struct RadixList<R0> {
  static constexpr int first = R0;
  static constexpr int last = R0;
  static constexpr int product = R0;
};

}  // namespace clr
