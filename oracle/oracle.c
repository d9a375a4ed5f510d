/*
 * oracle.c -- CPU FP64 restatement of CoLoRe's hot path.  TEST INFRASTRUCTURE
 * ONLY (see oracle.h).  Compiled with -ffp-contract=off: every a*b+c is
 * rounded twice, as the reference's plain C build rounds it on x86-64.
 */
#include "oracle.h"
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ========================================================================
 * Random numbers
 * ====================================================================== */

/* Philox4x32-10, Salmon et al. (SC'11), Random123 reference definition. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* counter layout shared with the GPU path: (site lo, site hi, stream, draw),
 * key (seed, 0x436F4C52) */
static void orc_draw(uint64_t site, uint32_t stream, uint32_t idx, uint32_t seed, uint32_t out[4]) {
  const uint32_t ctr[4] = {(uint32_t)site, (uint32_t)(site >> 32), stream, idx};
  const uint32_t key[2] = {seed, 0x436F4C52u};
  orc_philox4x32_10(ctr, key, out);
}
static double orc_u01(uint32_t x) { return (double)x / 4294967296.0; }

/* MT19937 (Matsumoto & Nishimura 1998) with init_genrand seeding, as GSL's
 * mt19937_set (seed 0 -> 4357) and mt19937_get. */
void orc_mt_seed(orc_mt19937 *s, uint32_t seed) {
  if (seed == 0) seed = 4357;
  s->mt[0] = seed;
  for (int i = 1; i < 624; i++)
    s->mt[i] = (uint32_t)(1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i);
  s->mti = 624;
}
uint32_t orc_mt_next(orc_mt19937 *s) {
  static const uint32_t mag01[2] = {0x0u, 0x9908b0dfu};
  uint32_t y;
  if (s->mti >= 624) {
    int kk;
    for (kk = 0; kk < 624 - 397; kk++) {
      y = (s->mt[kk] & 0x80000000u) | (s->mt[kk + 1] & 0x7fffffffu);
      s->mt[kk] = s->mt[kk + 397] ^ (y >> 1) ^ mag01[y & 1];
    }
    for (; kk < 623; kk++) {
      y = (s->mt[kk] & 0x80000000u) | (s->mt[kk + 1] & 0x7fffffffu);
      s->mt[kk] = s->mt[kk + (397 - 624)] ^ (y >> 1) ^ mag01[y & 1];
    }
    y = (s->mt[623] & 0x80000000u) | (s->mt[0] & 0x7fffffffu);
    s->mt[623] = s->mt[396] ^ (y >> 1) ^ mag01[y & 1];
    s->mti = 0;
  }
  y = s->mt[s->mti++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

/* fourier.c:295-357 with an OpenMP static schedule over kz as libgomp splits
 * n iterations over t threads: the first n%t threads get one extra. */
void orc_ref_mode_stream(int N, uint32_t seed, int nthreads, uint32_t *out) {
  const int NH = N / 2 + 1;
  const int q = N / nthreads, rem = N % nthreads;
  for (int ithr = 0; ithr < nthreads; ithr++) {
    orc_mt19937 mt;
    orc_mt_seed(&mt, seed + (uint32_t)ithr);
    const int i0 = ithr * q + (ithr < rem ? ithr : rem);
    const int i1 = i0 + q + (ithr < rem ? 1 : 0);
    for (int ii = i0; ii < i1; ii++)
      for (int jj = 0; jj < N; jj++)
        for (int kk = 0; kk <= N / 2; kk++) {
          const size_t m = (size_t)kk + (size_t)NH * ((size_t)jj + (size_t)N * ii);
          if (ii == 0 && jj == 0 && kk == 0) { out[2 * m] = 0; out[2 * m + 1] = 0; continue; }
          out[2 * m] = orc_mt_next(&mt);       /* phase   */
          out[2 * m + 1] = orc_mt_next(&mt);   /* modulus */
        }
  }
}

/* ========================================================================
 * Tables
 * ====================================================================== */

/* f_of_r_linear, cosmo.c:30-38 */
double orc_f_of_r(const orc_bg *bg, const double *f, double f0, double ff, double r) {
  if (r <= 0) return f0;
  if (r >= bg->r[bg->na - 1]) return ff;
  const int ir = (int)(r * bg->idr);
  return f[ir] + (f[ir + 1] - f[ir]) * (r - bg->r[ir]) * bg->idr;
}

/* pk_linear0, cosmo.c:291-308 (the last table bin uses the power-law tail) */
double orc_pk_linear0(const orc_pk *p, double lgk) {
  const int ik = (int)((lgk - p->logkmin) * p->idlogk);
  if (ik < 0) return p->pk[0] * pow(10, p->ns * (lgk - p->logkmin));
  if (ik < p->numk - 1) return p->pk[ik] + (lgk - p->logk[ik]) * (p->pk[ik + 1] - p->pk[ik]) * p->idlogk;
  return p->pk[p->numk - 1] * pow(10, -3 * (lgk - p->logkmax));
}

/* bias_model, common.h:414-431: 0 = (1+d)^b, 1 = _BIAS_MODEL_2, 2 = _BIAS_MODEL_3 */
double orc_bias_model(double d, double b, int model) {
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

/* ========================================================================
 * Fields
 * ====================================================================== */

/* create_grids_fourier, fourier.c:285-359 */
/* One mode of create_grids_fourier (fourier.c:314-353) at half-spectrum
 * index (kk, jj, ii): frequencies i <= N/2 -> +i dk, else -(N-i) dk; k = 0
 * is zero; (phase, modulus) from rng_delta_gauss (common.c:191-201). */
static void mode_at(int N, double L, const orc_pk *pk, uint32_t seed, const uint32_t *inj, double r2,
                    int do_smoothing, int smooth_potential, double prefac_lensing, int kk, int jj, int ii,
                    double d[2], double p[2]) {
  const int NH = N / 2 + 1;
  const double dk = 2 * M_PI / L;
  const double idk3 = 1. / (dk * dk * dk);
  const double kz = (2 * ii <= N) ? ii * dk : -(N - ii) * dk;
  const double ky = (2 * jj <= N) ? jj * dk : -(N - jj) * dk;
  const double kx = kk * dk;
  const size_t index = (size_t)kk + (size_t)NH * ((size_t)jj + (size_t)N * ii);
  const double k_mod2 = kx * kx + ky * ky + kz * kz;
  if (k_mod2 <= 0) { d[0] = d[1] = p[0] = p[1] = 0; return; }
  const double lgk = 0.5 * log10(k_mod2);
  const double sigma2 = orc_pk_linear0(pk, lgk) * idk3;
  uint32_t x0, x1;
  if (inj) { x0 = inj[2 * index]; x1 = inj[2 * index + 1]; }
  else {
    /* one Philox draw per (kz, kz+N/2) pair of modes */
    uint32_t r[4];
    const size_t site = (size_t)kk + (size_t)NH * ((size_t)jj + (size_t)N * (ii % (N / 2)));
    orc_draw(site, 0u, 0u, seed, r);
    const int hi = ii >= N / 2;
    x0 = r[2 * hi]; x1 = r[2 * hi + 1];
  }
  const double phase = 2 * M_PI * orc_u01(x0);
  const double u = orc_u01(x1);
  const double mod = sqrt(-sigma2 * log(1 - u));
  double dre = mod * cos(phase), dim = mod * sin(phase);
  double pre = -prefac_lensing * dre / k_mod2, pim = -prefac_lensing * dim / k_mod2;
  if (do_smoothing) {
    const double sm = exp(-0.5 * r2 * k_mod2);
    dre *= sm; dim *= sm;
    if (smooth_potential) { pre *= sm; pim *= sm; }
  }
  d[0] = dre; d[1] = dim; p[0] = pre; p[1] = pim;
}

void orc_gen_modes(int N, double L, const orc_pk *pk, uint32_t seed, const uint32_t *inj,
                   double r_smooth, int smooth_potential, double prefac_lensing,
                   double *dens_k, double *npot_k) {
  const int NH = N / 2 + 1;
  const int do_smoothing = r_smooth > 0;
  const double r2 = do_smoothing ? r_smooth * r_smooth : 0;
#pragma omp parallel for schedule(static)
  for (int ii = 0; ii < N; ii++)
    for (int jj = 0; jj < N; jj++)
      for (int kk = 0; kk <= N / 2; kk++) {
        const size_t index = (size_t)kk + (size_t)NH * ((size_t)jj + (size_t)N * ii);
        mode_at(N, L, pk, seed, inj, r2, do_smoothing, smooth_potential, prefac_lensing, kk, jj, ii,
                dens_k + 2 * index, npot_k + 2 * index);
      }
}

/* Parseval for the c2r of fourier.c:81-102 followed by the normalisation
 * (sqrt(2 pi)/L)^3 of fourier.c:383-398: the real grid is the inverse DFT of
 * the Hermitian spectrum F with F = X on 0 < kx < N/2 and, on the kx = 0 and
 * kx = N/2 planes (whose imaginary parts the c2r drops after the z and y
 * transforms), F(ky,kz) = (X(ky,kz) + conj X(-ky,-kz))/2.  So
 *   <grid^2> = norm^2 [ 2 sum_{0<kx<N/2} |X|^2 + sum_{kx=0,N/2} |F|^2 ].
 * Modes are regenerated one at a time (no spectrum is stored), so this runs
 * at any N.  out[0] = <delta^2>, out[1] = <phi^2> over the N^3 cells. */
void orc_kspace_variance(int N, double L, const orc_pk *pk, uint32_t seed, const uint32_t *inj,
                         double r_smooth, int smooth_potential, double prefac_lensing, double out[2]) {
  const int do_smoothing = r_smooth > 0;
  const double r2 = do_smoothing ? r_smooth * r_smooth : 0;
  double sd = 0, sp = 0;
#pragma omp parallel for schedule(static) reduction(+ : sd, sp)
  for (int ii = 0; ii < N; ii++)
    for (int jj = 0; jj < N; jj++)
      for (int kk = 0; kk <= N / 2; kk++) {
        double d[2], p[2];
        mode_at(N, L, pk, seed, inj, r2, do_smoothing, smooth_potential, prefac_lensing, kk, jj, ii, d, p);
        if (kk > 0 && 2 * kk < N) {
          sd += 2 * (d[0] * d[0] + d[1] * d[1]);
          sp += 2 * (p[0] * p[0] + p[1] * p[1]);
        } else {
          double dm[2], pm[2];
          mode_at(N, L, pk, seed, inj, r2, do_smoothing, smooth_potential, prefac_lensing, kk, (N - jj) % N,
                  (N - ii) % N, dm, pm);
          const double fr = 0.5 * (d[0] + dm[0]), fi = 0.5 * (d[1] - dm[1]);
          const double gr = 0.5 * (p[0] + pm[0]), gi = 0.5 * (p[1] - pm[1]);
          sd += fr * fr + fi * fi;
          sp += gr * gr + gi * gi;
        }
      }
  const double nrm = pow(sqrt(2 * M_PI) / L, 3);
  out[0] = sd * nrm * nrm;
  out[1] = sp * nrm * nrm;
}

/* Stockham autosort complex inverse FFT (decimation in frequency), radices
 * 4/2/3/5 and a direct DFT for any other prime factor.  Stage with current
 * length n_c = p*mp and stride s:
 *   y[s(p q + k) + j] = w_{n_c}^{q k} sum_r x[s(q + mp r) + j] w_p^{r k}. */
static void cfft_stockham(int n, double *x, double *y) {
  double *a = x, *b = y;
  int s = 1, nc = n;
  while (nc > 1) {
    int p;
    if (nc % 4 == 0) p = 4;
    else if (nc % 2 == 0) p = 2;
    else if (nc % 3 == 0) p = 3;
    else if (nc % 5 == 0) p = 5;
    else { p = 7; while (nc % p) p += 2; }
    const int mp = nc / p;
    double *wr = malloc(2 * (size_t)p * sizeof(double)), *wi = wr + p;
    for (int k = 0; k < p; k++) { wr[k] = cos(2 * M_PI * k / p); wi[k] = sin(2 * M_PI * k / p); }
    double *vr = malloc(6 * (size_t)p * sizeof(double)), *vi = vr + p, *tr = vr + 2 * p, *ti = vr + 3 * p;
    double *qc = vr + 4 * p, *qs = vr + 5 * p;
    for (int q = 0; q < mp; q++) {
      for (int k = 0; k < p; k++) {
        const double ang = 2 * M_PI * (double)q * k / nc;
        qc[k] = cos(ang);
        qs[k] = sin(ang);
      }
      for (int sidx = 0; sidx < s; sidx++) {
        for (int r = 0; r < p; r++) {
          const double *src = a + 2 * ((size_t)sidx + (size_t)s * (q + (size_t)mp * r));
          vr[r] = src[0];
          vi[r] = src[1];
        }
        for (int k = 0; k < p; k++) {
          double sr = 0, si = 0;
          for (int r = 0; r < p; r++) {
            const int e = (r * k) % p;
            sr += vr[r] * wr[e] - vi[r] * wi[e];
            si += vr[r] * wi[e] + vi[r] * wr[e];
          }
          tr[k] = sr * qc[k] - si * qs[k];
          ti[k] = sr * qs[k] + si * qc[k];
        }
        for (int k = 0; k < p; k++) {
          double *dst = b + 2 * ((size_t)sidx + (size_t)s * (p * (size_t)q + k));
          dst[0] = tr[k];
          dst[1] = ti[k];
        }
      }
    }
    free(vr);
    free(wr);
    double *t = a; a = b; b = t;
    s *= p;
    nc = mp;
  }
  if (a != x) memcpy(x, a, 2 * (size_t)n * sizeof(double));
}

void orc_cfft_inv(int n, double *data) {
  double *tmp = malloc(2 * (size_t)n * sizeof(double));
  cfft_stockham(n, data, tmp);
  free(tmp);
}

/* c2r along one row: X[0..N/2] (interleaved) -> out[0..N-1] */
static void c2r_row(int N, const double *X, double *out, double *work) {
  /* full Hermitian completion with Re of DC and Nyquist only */
  for (int k = 0; k <= N / 2; k++) { work[2 * k] = X[2 * k]; work[2 * k + 1] = X[2 * k + 1]; }
  work[1] = 0;
  work[2 * (N / 2) + 1] = 0;
  for (int k = 1; k < N / 2; k++) { work[2 * (N - k)] = X[2 * k]; work[2 * (N - k) + 1] = -X[2 * k + 1]; }
  orc_cfft_inv(N, work);
  for (int x = 0; x < N; x++) out[x] = work[2 * x];
}

void orc_c2r_3d(int N, const double *in, double *out) {
  const int NH = N / 2 + 1, ngx = 2 * NH;
  double *h = malloc(2 * (size_t)N * N * NH * sizeof(double));
  memcpy(h, in, 2 * (size_t)N * N * NH * sizeof(double));
  /* z then y complex transforms for every kx */
#pragma omp parallel
  {
    double *line = malloc(2 * (size_t)N * sizeof(double));
#pragma omp for schedule(static)
    for (int jj = 0; jj < N; jj++)
      for (int kk = 0; kk < NH; kk++) {
        for (int ii = 0; ii < N; ii++) {
          const size_t o = 2 * ((size_t)kk + (size_t)NH * ((size_t)jj + (size_t)N * ii));
          line[2 * ii] = h[o]; line[2 * ii + 1] = h[o + 1];
        }
        orc_cfft_inv(N, line);
        for (int ii = 0; ii < N; ii++) {
          const size_t o = 2 * ((size_t)kk + (size_t)NH * ((size_t)jj + (size_t)N * ii));
          h[o] = line[2 * ii]; h[o + 1] = line[2 * ii + 1];
        }
      }
#pragma omp for schedule(static)
    for (int z = 0; z < N; z++)
      for (int kk = 0; kk < NH; kk++) {
        for (int jj = 0; jj < N; jj++) {
          const size_t o = 2 * ((size_t)kk + (size_t)NH * ((size_t)jj + (size_t)N * z));
          line[2 * jj] = h[o]; line[2 * jj + 1] = h[o + 1];
        }
        orc_cfft_inv(N, line);
        for (int jj = 0; jj < N; jj++) {
          const size_t o = 2 * ((size_t)kk + (size_t)NH * ((size_t)jj + (size_t)N * z));
          h[o] = line[2 * jj]; h[o + 1] = line[2 * jj + 1];
        }
      }
    double *work = malloc(2 * (size_t)N * sizeof(double));
#pragma omp for schedule(static)
    for (long row = 0; row < (long)N * N; row++) {
      c2r_row(N, h + 2 * (size_t)NH * row, out + (size_t)ngx * row, work);
      out[(size_t)ngx * row + N] = 0;
      out[(size_t)ngx * row + N + 1] = 0;
    }
    free(work);
    free(line);
  }
  free(h);
}

void orc_c2r_3d_brute(int N, const double *in, double *out) {
  const int NH = N / 2 + 1, ngx = 2 * NH;
  double *G = calloc(2 * (size_t)NH * N * N, sizeof(double));   /* [kx][z][y] */
  for (int kx = 0; kx < NH; kx++)
    for (int z = 0; z < N; z++)
      for (int y = 0; y < N; y++) {
        double sr = 0, si = 0;
        for (int kz = 0; kz < N; kz++)
          for (int ky = 0; ky < N; ky++) {
            const double *h = in + 2 * ((size_t)kx + (size_t)NH * ((size_t)ky + (size_t)N * kz));
            const double ang = 2 * M_PI * (double)(((long)kz * z + (long)ky * y) % N) / N;
            sr += h[0] * cos(ang) - h[1] * sin(ang);
            si += h[0] * sin(ang) + h[1] * cos(ang);
          }
        double *g = G + 2 * (((size_t)kx * N + z) * N + y);
        g[0] = sr; g[1] = si;
      }
  for (int z = 0; z < N; z++)
    for (int y = 0; y < N; y++)
      for (int x = 0; x < N; x++) {
        double s = G[2 * (((size_t)0 * N + z) * N + y)];
        s += ((x & 1) ? -1.0 : 1.0) * G[2 * (((size_t)(N / 2) * N + z) * N + y)];
        for (int kx = 1; kx < N / 2; kx++) {
          const double *g = G + 2 * (((size_t)kx * N + z) * N + y);
          const double ang = 2 * M_PI * (double)(((long)kx * x) % N) / N;
          s += 2 * (g[0] * cos(ang) - g[1] * sin(ang));
        }
        out[(size_t)ngx * ((size_t)y + (size_t)N * z) + x] = s;
      }
  for (long row = 0; row < (long)N * N; row++) { out[(size_t)ngx * row + N] = 0; out[(size_t)ngx * row + N + 1] = 0; }
  free(G);
}

/* fourier.c:383-398 and compute_sigma_dens, fourier.c:24-79 */
double orc_normalize_sigma2(int N, double L, double *dens, double *npot, double *out2) {
  const int ngx = 2 * (N / 2 + 1);
  const size_t ntot = (size_t)ngx * N * N;
  const double norm = pow(sqrt(2 * M_PI) / L, 3);
  for (size_t i = 0; i < ntot; i++) { dens[i] *= norm; if (npot) npot[i] *= norm; }
  double s1 = 0, s2 = 0;
  const double ng = (double)N * N * N;
#pragma omp parallel for reduction(+ : s1, s2) schedule(static)
  for (long row = 0; row < (long)N * N; row++)
    for (int x = 0; x < N; x++) {
      const double d = dens[(size_t)ngx * row + x];
      s1 += d; s2 += d * d;
    }
  const double mean = s1 / ng;
  const double sig2 = s2 / ng - mean * mean;
  if (out2) { out2[0] = mean; out2[1] = s2 / ng; }
  return sig2;
}

/* ========================================================================
 * Density
 * ====================================================================== */

/* lognormalize (density.c:1070-1103) / densclip (density.c:1034-1067) */
void orc_density(int N, double L, const double obs[3], const orc_bg *bg, double sigma2,
                 int dens_type, double *grid) {
  const int ngx = 2 * (N / 2 + 1);
  const double dx = L / N;
  const double d1f = bg->d1[bg->na - 1];
#pragma omp parallel for schedule(static)
  for (int iz = 0; iz < N; iz++) {
    const double z0 = (iz + 0.0) * dx - obs[2];
    for (int iy = 0; iy < N; iy++) {
      const double y0 = (iy + 0.0) * dx - obs[1];
      for (int ix = 0; ix < N; ix++) {
        const double x0 = (ix + 0.0) * dx - obs[0];
        const size_t index = (size_t)ix + (size_t)ngx * ((size_t)iy + (size_t)N * iz);
        const double r = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
        const double dg = orc_f_of_r(bg, bg->d1, 1, d1f, r);
        const double delta = grid[index];
        if (dens_type == 3) grid[index] = fmax(1 + dg * delta, 0) - 1;
        else grid[index] = exp(dg * (delta - 0.5 * dg * sigma2)) - 1;
      }
    }
  }
}

/* collect_density_normalization_from_grid (density.c:1128-1211) */
int orc_shell_sums(int N, double L, const double obs[3], const orc_bg *bg, const double *grid,
                   int ntr, const double *const *bz, const double *bz0, const int *model, int cap,
                   double *zmax_out, unsigned long long *cnt, double *zsum, double *sums) {
  const int ngx = 2 * (N / 2 + 1);
  const double dx = L / N;
  const double zlast = bg->z[bg->na - 1];
  const double zmax = orc_f_of_r(bg, bg->z, 0, zlast, L * 0.5);
  const int nz = (int)(zmax / 0.05) + 2;
  const double idz = (nz - 2) / zmax;
  if (zmax_out) *zmax_out = zmax;
  if (nz > cap) return nz;
  memset(cnt, 0, nz * sizeof(unsigned long long));
  memset(zsum, 0, nz * sizeof(double));
  memset(sums, 0, (size_t)ntr * nz * sizeof(double));
#pragma omp parallel
  {
    unsigned long long *c_t = calloc(nz, sizeof(unsigned long long));
    double *z_t = calloc(nz, sizeof(double));
    double *s_t = calloc((size_t)ntr * nz, sizeof(double));
#pragma omp for schedule(static)
    for (int iz = 0; iz < N; iz++) {
      const double z0 = (iz + 0.0) * dx - obs[2];
      for (int iy = 0; iy < N; iy++) {
        const double y0 = (iy + 0.0) * dx - obs[1];
        for (int ix = 0; ix < N; ix++) {
          const double x0 = (ix + 0.0) * dx - obs[0];
          const size_t index = (size_t)ix + (size_t)ngx * ((size_t)iy + (size_t)N * iz);
          const double r = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
          const double redshift = orc_f_of_r(bg, bg->z, 0, zlast, r);
          const int ind_z = (int)(redshift * idz) + 1;
          if (ind_z >= 0 && ind_z < nz) {
            const double d = grid[index];
            c_t[ind_z]++;
            z_t[ind_z] += redshift;
            for (int t = 0; t < ntr; t++)
              s_t[(size_t)t * nz + ind_z] += orc_bias_model(d, orc_f_of_r(bg, bz[t], bz0[t], 1, r), model[t]);
          }
        }
      }
    }
#pragma omp critical
    {
      for (int i = 0; i < nz; i++) { cnt[i] += c_t[i]; zsum[i] += z_t[i]; }
      for (size_t i = 0; i < (size_t)ntr * nz; i++) sums[i] += s_t[i];
    }
    free(c_t); free(z_t); free(s_t);
  }
  return nz;
}

/* density.c:1272-1355 for one tracer; gsl_interp_linear between shell means */
void orc_norm_table(const orc_bg *bg, int nz, double zmax, const unsigned long long *cnt,
                    const double *zsum, const double *sums_t, double *norm_arr, double *norm0,
                    double *normf) {
  double *zarr = malloc(nz * sizeof(double)), *nrm = malloc(nz * sizeof(double));
  for (int iz = 0; iz < nz; iz++) {
    zarr[iz] = zsum[iz];
    nrm[iz] = sums_t[iz];
    if (cnt[iz] > 0) {
      zarr[iz] /= cnt[iz];
      nrm[iz] = cnt[iz] / nrm[iz];
    }
  }
  zarr[0] = 0;
  zarr[nz - 1] = zmax;
  nrm[0] = nrm[1];
  nrm[nz - 1] = nrm[nz - 2];
  const double z0 = zarr[0], zf = zarr[nz - 1];
  *norm0 = nrm[0];
  *normf = nrm[nz - 1];
  const double zlast = bg->z[bg->na - 1];
  for (int ii = 0; ii < bg->na; ii++) {
    const double z = orc_f_of_r(bg, bg->z, 0, zlast, bg->r[ii]);
    double nm;
    if (z < z0) nm = *norm0;
    else if (z >= zf) nm = *normf;
    else {
      int lo = 0, hi = nz - 1;
      while (hi > lo + 1) {
        const int i = (hi + lo) / 2;
        if (zarr[i] > z) hi = i; else lo = i;
      }
      const double dxz = zarr[lo + 1] - zarr[lo];
      nm = dxz > 0.0 ? nrm[lo] + (z - zarr[lo]) / dxz * (nrm[lo + 1] - nrm[lo]) : 0.0;
    }
    norm_arr[ii] = nm;
  }
  free(zarr); free(nrm);
}

/* ========================================================================
 * Sources
 * ====================================================================== */

static int poisson_inv(double lam, double u) {
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

/* inversion by sequential search with u = (hi 2^16 + lo)/2^32: hi the 16-bit
 * half (cell % 8) of Philox(cell/8, stream, 0), lo the low 16 bits of word 0
 * of Philox(cell, stream | 0x40000000, 0); lambda > 500 split into
 * ceil(lambda/500) equal parts, the i-th extra part using word (i-1)%4 of
 * Philox(cell, stream | 0x80000000, (i-1)/4) */
int orc_poisson(double lam, uint64_t cell, uint32_t stream, uint32_t seed) {
  if (!(lam > 0)) return 0;
  uint32_t r[4], q[4];
  orc_draw(cell >> 3, stream, 0u, seed, r);
  orc_draw(cell, stream | 0x40000000u, 0u, seed, q);
  const uint32_t hi = (r[(cell & 7) >> 1] >> (((cell & 7) & 1) * 16)) & 0xFFFFu;
  const uint32_t x = (hi << 16) | (q[0] & 0xFFFFu);
  const double u0 = orc_u01(x);
  if (lam <= 500.0) return poisson_inv(lam, u0);
  const int m = (int)ceil(lam / 500.0);
  const double lp = lam / m;
  int k = poisson_inv(lp, u0);
  for (int i = 1; i < m; i++) {
    const int j = i - 1;
    if ((j & 3) == 0) orc_draw(cell, stream | 0x80000000u, (uint32_t)(j >> 2), seed, r);
    k += poisson_inv(lp, orc_u01(r[j & 3]));
  }
  return k;
}

/* lambda of srcs.c:158-176: n(r) dx^3 bias_model(delta, b(r)) norm(r) for a
 * cell with r < L/2 + 20 and n(r) > 0, else 0 (no draw). */
double orc_cell_lambda(const orc_pop *p, const double *dens, int ix, int iy, int iz) {
  const int N = p->N, ngx = 2 * (N / 2 + 1);
  const double dx = p->L / N;
  const double cell_vol = dx * dx * dx;
  const double z0 = (iz + 0.0) * dx - p->obs[2];
  const double y0 = (iy + 0.0) * dx - p->obs[1];
  const double x0 = (ix + 0.0) * dx - p->obs[0];
  const double r = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
  if (!(r < p->L / 2 + 20.)) return 0;
  const double ndens = orc_f_of_r(p->bg, p->amp, 0, 0, r);
  if (!(ndens > 0)) return 0;
  const double bias = orc_f_of_r(p->bg, p->bz, p->bz0, 1, r);
  const double dnorm = orc_f_of_r(p->bg, p->norm, p->norm0, p->normf, r);
  const size_t index = (size_t)ix + (size_t)ngx * ((size_t)iy + (size_t)N * iz);
  return ndens * cell_vol * orc_bias_model(dens[index], bias, p->model) * dnorm;
}

int orc_cell_count(const orc_pop *p, const double *dens, int ix, int iy, int iz) {
  const int N = p->N;
  const double lambda = orc_cell_lambda(p, dens, ix, iy, iz);
  const uint64_t cell = (uint64_t)ix + (uint64_t)N * ((uint64_t)iy + (uint64_t)N * iz);
  return orc_poisson(lambda, cell, 0x100u + (uint32_t)p->ipop, p->seed);
}

long long orc_counts(const orc_pop *p, const double *dens, int *counts) {
  const int N = p->N;
  long long tot = 0;
#pragma omp parallel for reduction(+ : tot) schedule(static)
  for (int iz = 0; iz < N; iz++)
    for (int iy = 0; iy < N; iy++)
      for (int ix = 0; ix < N; ix++) {
        const int c = orc_cell_count(p, dens, ix, iy, iz);
        if (counts) counts[(size_t)ix + (size_t)N * ((size_t)iy + (size_t)N * iz)] = c;
        tot += c;
      }
  return tot;
}

/* orc_counts for the planes iz0 .. iz0+nz-1 only; dens_slab holds those
 * planes in the padded layout, counts[(iz-iz0) N^2 + iy N + ix]. */
long long orc_counts_planes(const orc_pop *p, const double *dens_slab, int iz0, int nz, int *counts) {
  const int N = p->N, ngx = 2 * (N / 2 + 1);
  const double *dens = dens_slab - (ptrdiff_t)iz0 * N * ngx;   /* indexed by the global iz */
  long long tot = 0;
#pragma omp parallel for reduction(+ : tot) schedule(static)
  for (long row = 0; row < (long)nz * N; row++) {
    const int iz = iz0 + (int)(row / N), iy = (int)(row % N);
    for (int ix = 0; ix < N; ix++) {
      const int c = orc_cell_count(p, dens, ix, iy, iz);
      counts[(size_t)row * N + ix] = c;
      tot += c;
    }
  }
  return tot;
}

/* get_rvel, srcs.c:87-118 (periodic z on one rank) */
static double rvel(int N, double L, const double *phi, int ix, int iy, int iz, double x0, double y0,
                   double z0, double rr) {
  const long ngx = 2 * (N / 2 + 1);
  const double idx = N / L;
  long iz_hi = iz + 1, iz_lo = iz - 1, iz_0 = iz;
  long iy_hi = iy + 1, iy_lo = iy - 1, iy_0 = iy;
  long ix_hi = ix + 1, ix_lo = ix - 1, ix_0 = ix;
  if (iy == 0) iy_lo = N - 1;
  if (iy == N - 1) iy_hi = 0;
  if (ix == 0) ix_lo = N - 1;
  if (ix == N - 1) ix_hi = 0;
  if (iz == 0) iz_lo = N - 1;
  if (iz == N - 1) iz_hi = 0;
  iz_0 *= ngx * N; iz_lo *= ngx * N; iz_hi *= ngx * N;
  iy_0 *= ngx; iy_lo *= ngx; iy_hi *= ngx;
  double u[3] = {x0 / rr, y0 / rr, z0 / rr}, v[3];
  /* flouble differences, rounded to float as the reference computes them */
  v[0] = (float)phi[ix_hi + iy_0 + iz_0] - (float)phi[ix_lo + iy_0 + iz_0];
  v[1] = (float)phi[ix_0 + iy_hi + iz_0] - (float)phi[ix_0 + iy_lo + iz_0];
  v[2] = (float)phi[ix_0 + iy_0 + iz_hi] - (float)phi[ix_0 + iy_0 + iz_lo];
  return 0.5 * idx * (v[0] * u[0] + v[1] * u[1] + v[2] * u[2]);
}

/* srcs.c:205-278 on one thread: cells in z-major order, sources of a cell in
 * draw order; positions x0 + dx (u - 1/2) from the Philox position stream. */
long long orc_sources(const orc_pop *p, const double *dens, const double *npot, int nside_base,
                      long long cap, float *pos4, int *ipix) {
  const int N = p->N;
  const double dx = p->L / N;
  const double factor_vel = -p->fgrowth_0 / (1.5 * p->hubble_0 * p->omega_m);
  const double v1f0 = p->bg->v1[0], v1ff = p->bg->v1[p->bg->na - 1];
  /* pass 1: counts per (z,y) row; pass 2: every row writes from its offset,
   * which reproduces the sequential z-major order */
  long long *off = malloc(((size_t)N * N + 1) * sizeof(long long));
#pragma omp parallel for schedule(static)
  for (long row = 0; row < (long)N * N; row++) {
    long long c = 0;
    for (int ix = 0; ix < N; ix++) c += orc_cell_count(p, dens, ix, (int)(row % N), (int)(row / N));
    off[row + 1] = c;
  }
  off[0] = 0;
  for (long row = 0; row < (long)N * N; row++) off[row + 1] += off[row];
  const long long total = off[(size_t)N * N];
  if (total > cap) { free(off); return -1; }
#pragma omp parallel for schedule(static)
  for (long row = 0; row < (long)N * N; row++) {
    const int iz = (int)(row / N), iy = (int)(row % N);
    const double z0 = (iz + 0.0) * dx - p->obs[2];
    const double y0 = (iy + 0.0) * dx - p->obs[1];
    long long pid = off[row];
    for (int ix = 0; ix < N; ix++) {
      const double x0 = (ix + 0.0) * dx - p->obs[0];
      const int npp = orc_cell_count(p, dens, ix, iy, iz);
      if (npp <= 0) continue;
      const double rr = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
      const double rv = factor_vel * rvel(N, p->L, npot, ix, iy, iz, x0, y0, z0, rr);
      const double dz_rsd = rv * orc_f_of_r(p->bg, p->bg->v1, v1f0, v1ff, rr);
      const uint64_t cell = (uint64_t)ix + (uint64_t)N * ((uint64_t)iy + (uint64_t)N * iz);
      uint32_t rnd[4] = {0, 0, 0, 0};
      for (int ip = 0; ip < npp; ip++) {
        const double c0[3] = {x0, y0, z0};
        float q[3];
        for (int ax = 0; ax < 3; ax++) {
          const uint32_t j = 3u * (uint32_t)ip + (uint32_t)ax;
          if ((j & 3u) == 0) orc_draw(cell, 0x200u + (uint32_t)p->ipop, j >> 2, p->seed, rnd);
          q[ax] = (float)(c0[ax] + dx * (orc_u01(rnd[j & 3u]) - 0.5));
        }
        pos4[4 * pid + 0] = q[0];
        pos4[4 * pid + 1] = q[1];
        pos4[4 * pid + 2] = q[2];
        pos4[4 * pid + 3] = (float)dz_rsd;
        const double v[3] = {q[0], q[1], q[2]};
        ipix[pid] = (int)orc_ring2nest(nside_base, orc_vec2pix_ring(nside_base, v));
        pid++;
      }
    }
  }
  free(off);
  return total;
}

/* ---- the reference's own streams (srcs.c:136-278 on OpenMP threads) ------
 * gsl_rng_mt19937 output x maps to gsl_rng_uniform = x / 2^32 and
 * gsl_rng_uniform_pos repeats while that is 0 (GSL's documented mapping).
 * gsl_ran_poisson (called through rng_poisson, common.c:185-189), as GSL
 * documents it:
 *   while mu > 10: m = (unsigned)(mu 7/8); X = gamma_int(m);
 *     X >= mu: return k + binomial(mu/X, m-1); else k += m, mu -= X;
 *   then the multiplication method: prod of uniforms until prod <= e^-mu,
 *     returning (number of uniforms) - 1.
 * gamma_int(a): a < 12 -> -log(prod of a uniform_pos); else the rejection
 * method of Knuth (TAOCP vol. 2, 3.4.1, algorithm A, tan-based).
 * binomial(p, n): p > 1/2 uses 1-p and returns n - k; inversion by sequential
 * search with f(x+1) = f(x) (n-x)/(x+1) p/q from f(0) = q^n (q^n by binary
 * powering), at most 111 terms per uniform and a fresh uniform when they are
 * exhausted.  GSL switches to the BTPE algorithm when n p >= 14; that branch
 * is not restated (inversion is used there too): parity unpinned for cells
 * that reach it. */
typedef struct { const uint32_t *s; long long len, pos; int overflow; } ref_stream;

static double rs_uniform(ref_stream *r) {
  if (r->pos >= r->len) { r->overflow = 1; return 0.5; }
  return r->s[r->pos++] / 4294967296.0;
}
static double rs_uniform_pos(ref_stream *r) {
  double x;
  do { x = rs_uniform(r); } while (x == 0 && !r->overflow);
  return x;
}
static double pow_int_binary(double x, unsigned n) {
  double v = 1.0;
  do {
    if (n & 1u) v *= x;
    n >>= 1;
    x *= x;
  } while (n);
  return v;
}
static double ref_gamma_int(ref_stream *r, unsigned a) {
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
    } while (x <= 0 && !r->overflow);
    v = rs_uniform(r);
  } while (v > (1 + y * y) * exp(am1 * log(x / am1) - sqa * y) && !r->overflow);
  return x;
}
static unsigned ref_binomial(ref_stream *r, double p, unsigned n) {
  if (n == 0) return 0;
  int flipped = 0;
  if (p > 0.5) { p = 1.0 - p; flipped = 1; }
  const double q = 1 - p, s = p / q;
  const double f0 = pow_int_binary(q, n);
  unsigned ix = 0;
  for (;;) {
    double f = f0, u = rs_uniform(r);
    int done = 0;
    for (ix = 0; ix <= 110; ix++) {
      if (u < f) { done = 1; break; }
      u -= f;
      f *= s * (double)(n - ix) / (double)(ix + 1);
    }
    if (done || r->overflow) break;
  }
  return flipped ? n - ix : ix;
}
static int ref_poisson(ref_stream *r, double mu) {
  unsigned k = 0;
  while (mu > 10 && !r->overflow) {
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
  } while (prod > emu && !r->overflow);
  return (int)(k - 1);
}

int orc_ref_poisson(const uint32_t *stream, long long len, double mu, long long *used) {
  ref_stream r = {stream, len, 0, 0};
  const int k = ref_poisson(&r, mu);
  if (used) *used = r.overflow ? -1 : r.pos;
  return k;
}

void orc_mt_fill(uint32_t seed, long long n, uint32_t *out) {
  orc_mt19937 mt;
  orc_mt_seed(&mt, seed);
  for (long long i = 0; i < n; i++) out[i] = orc_mt_next(&mt);
}

/* srcs_set_cartesian_single (srcs.c:120-283) on nthr OpenMP threads, thread t
 * owning the planes iz_begin[t] .. iz_begin[t+1]-1 (schedule(static)) and
 * drawing from streams[t*stream_len ...] (the MT19937 stream seeded with
 * seed + t + nthr (ipop + n_srcs IThread0), srcs.c:150).  Pass 1: every cell
 * with r < L/2+20 and n(r) > 0 draws its count with gsl_ran_poisson (even
 * when lambda is 0, which takes one uniform).  Pass 2 re-seeds (srcs.c:231):
 * the k-th source of thread t takes uniforms 3k, 3k+1, 3k+2 of the stream
 * for x, y, z.  Returns the number of sources (<= cap), -1 when cap is too
 * small, -2 when a stream runs out.  counts (N^3, cell order) may be NULL. */
long long orc_sources_ref_stream(const orc_pop *p, const double *dens, const double *npot, int nside_base,
                                 int nthr, const int *iz_begin, const uint32_t *streams, long long stream_len,
                                 long long cap, float *pos4, int *ipix, int *counts) {
  const int N = p->N;
  const double dx = p->L / N;
  const double factor_vel = -p->fgrowth_0 / (1.5 * p->hubble_0 * p->omega_m);
  const double v1f0 = p->bg->v1[0], v1ff = p->bg->v1[p->bg->na - 1];
  int *cnt = malloc((size_t)N * N * N * sizeof(int));
  long long *tot = calloc((size_t)nthr + 1, sizeof(long long));
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
  for (int t = 0; t < nthr; t++) {
    ref_stream rs = {streams + (size_t)t * stream_len, stream_len, 0, 0};
    long long nt = 0;
    for (int iz = iz_begin[t]; iz < iz_begin[t + 1]; iz++)
      for (int iy = 0; iy < N; iy++)
        for (int ix = 0; ix < N; ix++) {
          const size_t c = (size_t)ix + (size_t)N * ((size_t)iy + (size_t)N * iz);
          const double z0 = (iz + 0.0) * dx - p->obs[2], y0 = (iy + 0.0) * dx - p->obs[1];
          const double x0 = (ix + 0.0) * dx - p->obs[0];
          const double r = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
          int npp = 0;
          if (r < p->L / 2 + 20. && orc_f_of_r(p->bg, p->amp, 0, 0, r) > 0)
            npp = ref_poisson(&rs, orc_cell_lambda(p, dens, ix, iy, iz));
          cnt[c] = npp;
          nt += npp;
        }
    tot[t + 1] = nt;
    if (rs.overflow) bad |= 1;
  }
  long long total = 0;
  for (int t = 0; t < nthr; t++) { tot[t + 1] += tot[t]; }
  total = tot[nthr];
  if (counts) memcpy(counts, cnt, (size_t)N * N * N * sizeof(int));
  if (bad) { free(cnt); free(tot); return -2; }
  if (total > cap) { free(cnt); free(tot); return -1; }
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
  for (int t = 0; t < nthr; t++) {
    const uint32_t *st = streams + (size_t)t * stream_len;
    long long pid = tot[t], k = 0;
    for (int iz = iz_begin[t]; iz < iz_begin[t + 1]; iz++)
      for (int iy = 0; iy < N; iy++)
        for (int ix = 0; ix < N; ix++) {
          const size_t c = (size_t)ix + (size_t)N * ((size_t)iy + (size_t)N * iz);
          const int npp = cnt[c];
          if (npp <= 0) continue;
          const double z0 = (iz + 0.0) * dx - p->obs[2], y0 = (iy + 0.0) * dx - p->obs[1];
          const double x0 = (ix + 0.0) * dx - p->obs[0];
          const double rr = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
          const double rv = factor_vel * rvel(N, p->L, npot, ix, iy, iz, x0, y0, z0, rr);
          const double dz_rsd = rv * orc_f_of_r(p->bg, p->bg->v1, v1f0, v1ff, rr);
          const double c0[3] = {x0, y0, z0};
          for (int ip = 0; ip < npp; ip++, k++, pid++) {
            float q[3];
            for (int ax = 0; ax < 3; ax++) {
              const long long j = 3 * k + ax;
              if (j >= stream_len) { bad |= 1; q[ax] = 0; continue; }
              q[ax] = (float)(c0[ax] + dx * (st[j] / 4294967296.0 - 0.5));
            }
            pos4[4 * pid + 0] = q[0];
            pos4[4 * pid + 1] = q[1];
            pos4[4 * pid + 2] = q[2];
            pos4[4 * pid + 3] = (float)dz_rsd;
            const double v[3] = {q[0], q[1], q[2]};
            ipix[pid] = (int)orc_ring2nest(nside_base, orc_vec2pix_ring(nside_base, v));
          }
        }
  }
  free(cnt); free(tot);
  return bad ? -2 : total;
}

/* srcs_get_local_properties_single, srcs.c:386-416, cart2sph srcs.c:68-85 */
void orc_source_props(const orc_bg *bg, long long n, const float *pos4, float *out4) {
  const double RTOD = 57.2957795;
  const double zl = bg->z[bg->na - 1];
  for (long long i = 0; i < n; i++) {
    const double x = pos4[4 * i], y = pos4[4 * i + 1], z = pos4[4 * i + 2];
    const double r = sqrt(x * x + y * y + z * z);
    double cth = 1, phi = 0;
    if (r != 0) {
      cth = z / r;
      phi = atan2(y / r, x / r);
      if (phi < 0) phi += 2 * M_PI;
    }
    out4[4 * i] = (float)(RTOD * phi);
    out4[4 * i + 1] = (float)(90 - RTOD * acos(cth));
    out4[4 * i + 2] = (float)orc_f_of_r(bg, bg->z, 0, zl, r);
    out4[4 * i + 3] = pos4[4 * i + 3];
  }
}

/* ========================================================================
 * Lightcone painters
 * ====================================================================== */

static int cmp_dbl_idx(const void *a, const void *b, void *ctx) {
  const float *v = ctx;
  const int ia = *(const int *)a, ib = *(const int *)b;
  return v[ia] < v[ib] ? -1 : (v[ia] > v[ib] ? 1 : (ia < ib ? -1 : 1));
}

static void sort_idx(int n, const double *vals, int *idx) {
  float *v = malloc(n * sizeof(float));
  for (int i = 0; i < n; i++) { v[i] = (float)vals[i]; idx[i] = i; }
  qsort_r(idx, n, sizeof(int), cmp_dbl_idx, v);
  free(v);
}

/* kappa_get_beam_properties (kappa.c:78-175) / isw_get_beam_properties
 * (isw.c:78-147), pixels in NEST order, out[nplanes][12 nside^2] */
void orc_kappa_isw(int which, int N, double L, const double obs[3], double r_max,
                   const orc_bg *bg, const double *npot, int nside, int nplanes,
                   const double *r_src, double *out) {
  const long ngx = 2 * (N / 2 + 1);
  const double idx = N / L;
  const int nr = N / 2;
  const double dr = r_max / nr, idr = 1. / dr;
  int *ord = malloc(nplanes * sizeof(int));
  sort_idx(nplanes, r_src, ord);
  int *irmax = malloc(nplanes * sizeof(int)), *irmin = malloc(nplanes * sizeof(int));
  double *inv = malloc(nplanes * sizeof(double));
  for (int i = 0; i < nplanes; i++) {
    const int ih = (int)(r_src[ord[i]] * idr + 0.5);
    inv[i] = 1. / (ih * dr);
    irmax[i] = ih < nr - 1 ? ih : nr - 1;
  }
  irmin[0] = 0;
  for (int i = 1; i < nplanes; i++) irmin[i] = irmax[i - 1] + 1;
  double *f1 = malloc(nr * sizeof(double)), *f2 = malloc(nr * sizeof(double));
  const int na = bg->na;
  for (int i = 0; i < nr; i++) {
    const double rm = (i + 0.5) * dr;
    if (which == 0) {
      const double pg = orc_f_of_r(bg, bg->d1, 1, bg->d1[na - 1], rm) * (1 + orc_f_of_r(bg, bg->z, 0, bg->z[na - 1], rm));
      f1[i] = rm * pg * dr;
      f2[i] = rm * rm * pg * dr;
    } else {
      f1[i] = 2 * orc_f_of_r(bg, bg->pd, bg->pd[0], bg->pd[na - 1], rm) * dr;
      f2[i] = 0;
    }
  }
  const long long npix = 12LL * nside * nside;
#pragma omp parallel for schedule(dynamic, 256)
  for (long long ip = 0; ip < npix; ip++) {
    double u[3], rot[6] = {0, 0, 0, 0, 0, 0};
    orc_pix2vec_nest(nside, ip, u);
    if (which == 0) {
      double cth = u[2];
      if (cth >= 1) cth = 1;
      if (cth <= -1) cth = -1;
      const double sth = sqrt((1 - cth) * (1 + cth));
      double cph = 1, sph = 0;
      if (sth != 0) { cph = u[0] / sth; sph = u[1] / sth; }
      const double pre = idx * idx;
      rot[0] = (cth * cth * cph * cph + sph * sph) * pre;
      rot[1] = (2 * cph * sph * (cth * cth - 1)) * pre;
      rot[2] = (-2 * cth * sth * cph) * pre;
      rot[3] = (cth * cth * sph * sph + cph * cph) * pre;
      rot[4] = (-2 * cth * sth * sph) * pre;
      rot[5] = (sth * sth) * pre;
    }
    double k1 = 0, k2 = 0;
    for (int pl = 0; pl < nplanes; pl++) {
      for (int irr = irmin[pl]; irr <= irmax[pl]; irr++) {
        const double rm = (irr + 0.5) * dr;
        long ix[3];
        for (int ax = 0; ax < 3; ax++) {
          const double xn = (rm * u[ax] + obs[ax]) * idx;
          ix[ax] = (long)(xn + 0.5);
          if (ix[ax] >= N) ix[ax] -= N;
          else if (ix[ax] < 0) ix[ax] += N;
        }
        const long xh = ix[0] == N - 1 ? 0 : ix[0] + 1, xl = ix[0] == 0 ? N - 1 : ix[0] - 1;
        const long yh = ix[1] == N - 1 ? 0 : ix[1] + 1, yl = ix[1] == 0 ? N - 1 : ix[1] - 1;
        const long zh = ix[2] == N - 1 ? 0 : ix[2] + 1, zl = ix[2] == 0 ? N - 1 : ix[2] - 1;
        const long Z0 = ix[2] * ngx * N, ZH = zh * ngx * N, ZL = zl * ngx * N;
        const long Y0 = ix[1] * ngx, YH = yh * ngx, YL = yl * ngx;
        const long X0 = ix[0];
        if (which == 1) { k1 += npot[X0 + Y0 + Z0] * f1[irr]; continue; }
        double t[6];
        t[0] = (npot[xh + Y0 + Z0] + npot[xl + Y0 + Z0] - 2 * npot[X0 + Y0 + Z0]);
        t[3] = (npot[X0 + YH + Z0] + npot[X0 + YL + Z0] - 2 * npot[X0 + Y0 + Z0]);
        t[1] = 0.25 * (npot[xh + YH + Z0] + npot[xl + YL + Z0] - npot[xh + YL + Z0] - npot[xl + YH + Z0]);
        t[5] = (npot[X0 + Y0 + ZH] + npot[X0 + Y0 + ZL] - 2 * npot[X0 + Y0 + Z0]);
        t[2] = 0.25 * (npot[xh + Y0 + ZH] + npot[xl + Y0 + ZL] - npot[xh + Y0 + ZL] - npot[xl + Y0 + ZH]);
        t[4] = 0.25 * (npot[X0 + YH + ZH] + npot[X0 + YL + ZL] - npot[X0 + YH + ZL] - npot[X0 + YL + ZH]);
        double dotp = 0;
        for (int q = 0; q < 6; q++) dotp += rot[q] * t[q];
        k1 += dotp * f1[irr];
        k2 += dotp * f2[irr];
      }
      out[(size_t)pl * npix + ip] = which == 0 ? (k1 - inv[pl] * k2) : k1;
    }
  }
  free(ord); free(irmax); free(irmin); free(inv); free(f1); free(f2);
}

/* get_element (beaming.c:31-117) on one rank: the unscaled central
 * differences of phi (v) and its 19-point second differences (t: xx, xy, xz,
 * yy, yz, zz), periodic in x, y and z (fourier.c:412-413 sets the z halo
 * slices to the periodic neighbours). */
static void grid_element(const double *phi, int N, long ix, long iy, long iz, double v[3], double t[6]) {
  const long ngx = 2 * (N / 2 + 1);
  const long xh = ix == N - 1 ? 0 : ix + 1, xl = ix == 0 ? N - 1 : ix - 1;
  const long yh = iy == N - 1 ? 0 : iy + 1, yl = iy == 0 ? N - 1 : iy - 1;
  const long zh = iz == N - 1 ? 0 : iz + 1, zl = iz == 0 ? N - 1 : iz - 1;
  const long Z0 = iz * ngx * N, ZH = zh * ngx * N, ZL = zl * ngx * N;
  const long Y0 = iy * ngx, YH = yh * ngx, YL = yl * ngx;
  if (v) {
    v[0] = phi[xh + Y0 + Z0] - phi[xl + Y0 + Z0];
    v[1] = phi[ix + YH + Z0] - phi[ix + YL + Z0];
    v[2] = phi[ix + Y0 + ZH] - phi[ix + Y0 + ZL];
  }
  if (t) {
    const double c0 = phi[ix + Y0 + Z0];
    t[0] = phi[xh + Y0 + Z0] + phi[xl + Y0 + Z0] - 2 * c0;
    t[3] = phi[ix + YH + Z0] + phi[ix + YL + Z0] - 2 * c0;
    t[5] = phi[ix + Y0 + ZH] + phi[ix + Y0 + ZL] - 2 * c0;
    t[1] = 0.25 * (phi[xh + YH + Z0] + phi[xl + YL + Z0] - phi[xh + YL + Z0] - phi[xl + YH + Z0]);
    t[2] = 0.25 * (phi[xh + Y0 + ZH] + phi[xl + Y0 + ZL] - phi[xh + Y0 + ZL] - phi[xl + Y0 + ZH]);
    t[4] = 0.25 * (phi[ix + YH + ZH] + phi[ix + YL + ZL] - phi[ix + YH + ZL] - phi[ix + YL + ZH]);
  }
}

static long wrap_cell(long i, int N) { return i >= N ? i - N : (i < 0 ? i + N : i); }

/* r_of_z (cosmo.c:101-112): r(a) tabulated at a = i/(NA-1). */
static double r_of_z_tab(int na, const double *r_a2r, double z) {
  const double a = 1. / (1 + z);
  if (a >= 1) return 0;
  if (a <= 0) return r_a2r[0];
  const int ia = (int)(a * (na - 1));
  return r_a2r[ia] + (r_a2r[ia + 1] - r_a2r[ia]) * (a - (double)ia / (na - 1)) * (na - 1.);
}

/* srcs_beams_preproc_single (srcs.c:425-445), srcs_get_beam_properties_single
 * (srcs.c:452-625, the build without _USE_FAST_LENSING) and
 * srcs_beams_postproc_single (srcs.c:634-737) on one rank.  pos4[4i..4i+2]
 * are the positions relative to the observer (CatalogCartesian).  dz[i] is
 * the CIC radial velocity at the source (interpolate_from_grid with
 * INTERP_TYPE_SKW = CIC, beaming.c:193-268) scaled by V1(r_of_z(z0)) times
 * -f0/(1.5 H0 Om); with lens != NULL, lens[5i..5i+4] = (e1, e2, kappa, dra,
 * ddec) of the NGP march of nr = N/2 steps of r_max/nr up to the source
 * (kappa, dra and ddec start at 0, dra/ddec in degrees). */
void orc_src_beams(int N, double L, const double obs[3], double r_max, const orc_bg *bg,
                   const double *r_a2r, double fgrowth_0, double hubble_0, double omega_m,
                   const double *npot, long long n, const float *pos4, double *dz, double *lens) {
  const double idx = N / L;
  const int nr = N / 2, na = bg->na;
  const double dr = r_max / nr, idr = 1. / dr;
  const double factor_vel = -fgrowth_0 / (1.5 * hubble_0 * omega_m);
  const double RTOD = 57.2957795;
  double *f0 = malloc(3 * (size_t)nr * sizeof(double)), *f1 = f0 + nr, *f2 = f0 + 2 * nr;
  for (int i = 0; i < nr; i++) {
    const double rm = (i + 0.5) * dr;
    const double pg = orc_f_of_r(bg, bg->d1, 1, bg->d1[na - 1], rm) * (1 + orc_f_of_r(bg, bg->z, 0, bg->z[na - 1], rm));
    f0[i] = pg * dr;
    f1[i] = rm * pg * dr;
    f2[i] = rm * rm * pg * dr;
  }
#pragma omp parallel for schedule(dynamic, 256)
  for (long long ip = 0; ip < n; ip++) {
    const float *pos = pos4 + 4 * ip;
    const double r = sqrt((double)pos[0] * pos[0] + (double)pos[1] * pos[1] + (double)pos[2] * pos[2]);
    const double ir = 1. / (r > 0.001 ? r : 0.001);
    double xn[3], u[3];
    for (int ax = 0; ax < 3; ax++) {
      xn[ax] = (pos[ax] + obs[ax]) * idx;
      u[ax] = pos[ax] * ir;
    }
    /* trilinear (CIC) velocity at the source, beaming.c:193-266 */
    long i0[3], i1[3];
    double h0[3], h1[3];
    for (int ax = 0; ax < 3; ax++) {
      i0[ax] = (long)xn[ax];
      h0[ax] = xn[ax] - i0[ax];
      h1[ax] = 1 - h0[ax];
      i1[ax] = wrap_cell(i0[ax] + 1, N);
      i0[ax] = wrap_cell(i0[ax], N);
    }
    double v[3] = {0, 0, 0};
    for (int c = 0; c < 8; c++) {
      const int bx = c & 1, by = (c >> 1) & 1, bz = c >> 2;
      const double w = (bz ? h0[2] : h1[2]) * (by ? h0[1] : h1[1]) * (bx ? h0[0] : h1[0]);
      double vc[3];
      grid_element(npot, N, bx ? i1[0] : i0[0], by ? i1[1] : i0[1], bz ? i1[2] : i0[2], vc, NULL);
      for (int ax = 0; ax < 3; ax++) v[ax] += vc[ax] * w;
    }
    const double vr = 0.5 * idx * (v[0] * u[0] + v[1] * u[1] + v[2] * u[2]);
    /* postproc: z0 is the catalogue's float z(r) (srcs.c:409) */
    const float z0 = (float)orc_f_of_r(bg, bg->z, 0, bg->z[na - 1], r);
    const double vg = orc_f_of_r(bg, bg->v1, bg->v1[0], bg->v1[na - 1], r_of_z_tab(na, r_a2r, z0));
    dz[ip] = (double)(float)vr * (vg * factor_vel);
    if (!lens) continue;
    double cth_h = u[2], sth_h, cph_h = 1, sph_h = 0;
    if (cth_h >= 1) cth_h = 1;
    if (cth_h <= -1) cth_h = -1;
    sth_h = sqrt((1 - cth_h) * (1 + cth_h));
    if (sth_h != 0) { cph_h = u[0] / sth_h; sph_h = u[1] / sth_h; }
    const double pre = idx * idx * ir, prem = 0.5 * idx * ir;
    const double ux[3] = {cth_h * cph_h * prem, cth_h * sph_h * prem, -sth_h * prem};
    const double uy[3] = {-sph_h * prem, cph_h * prem, 0};
    const double rk[6] = {(cth_h * cth_h * cph_h * cph_h + sph_h * sph_h) * pre,
                          (2 * cph_h * sph_h * (cth_h * cth_h - 1)) * pre,
                          (-2 * cth_h * sth_h * cph_h) * pre,
                          (cth_h * cth_h * sph_h * sph_h + cph_h * cph_h) * pre,
                          (-2 * cth_h * sth_h * sph_h) * pre,
                          (sth_h * sth_h) * pre};
    const double re1[6] = {(cth_h * cth_h * cph_h * cph_h - sph_h * sph_h) * pre,
                           (2 * cph_h * sph_h * (cth_h * cth_h + 1)) * pre,
                           (-2 * cth_h * sth_h * cph_h) * pre,
                           (cth_h * cth_h * sph_h * sph_h - cph_h * cph_h) * pre,
                           (-2 * cth_h * sth_h * sph_h) * pre,
                           (sth_h * sth_h) * pre};
    const double re2[6] = {(-2 * cth_h * cph_h * sph_h) * pre,
                           (2 * cth_h * (cph_h * cph_h - sph_h * sph_h)) * pre,
                           (2 * sth_h * sph_h) * pre,
                           (2 * cth_h * sph_h * cph_h) * pre,
                           (-2 * sth_h * cph_h) * pre,
                           0};
    double dtx = 0, dty = 0, kp = 0, e1 = 0, e2 = 0;
    int irmax = (int)(r * idr + 0.5);
    if (irmax > nr - 1) irmax = nr - 1;
    for (int i_r = 0; i_r <= irmax; i_r++) {
      const double rm = (i_r + 0.5) * dr;
      long ix[3];
      for (int ax = 0; ax < 3; ax++) ix[ax] = wrap_cell((long)((rm * u[ax] + obs[ax]) * idx + 0.5), N);
      double vp[3], tp[6];
      grid_element(npot, N, ix[0], ix[1], ix[2], vp, tp);
      const double fr = f1[i_r] * r - f2[i_r];
      const double frm = 2 * (f0[i_r] * r - f1[i_r]);
      double de1 = 0, de2 = 0, dk = 0, dvx = 0, dvy = 0;
      for (int q = 0; q < 6; q++) {
        de1 += re1[q] * tp[q];
        de2 += re2[q] * tp[q];
        dk += rk[q] * tp[q];
      }
      for (int ax = 0; ax < 3; ax++) {
        dvx += ux[ax] * vp[ax];
        dvy += uy[ax] * vp[ax];
      }
      e1 += de1 * fr;
      e2 += de2 * fr;
      kp += dk * fr;
      dtx += dvx * frm;
      dty += dvy * frm;
    }
    double *o = lens + 5 * ip;
    o[0] = e1;
    o[1] = e2;
    o[2] = kp;
    o[3] = dty * RTOD;
    o[4] = dtx * RTOD;
  }
  free(f0);
}

/* The skewers of srcs_get_beam_properties_single (srcs.c:507-529) and their
 * postproc scaling (srcs.c:725-733) on one rank.  Per source i and step
 * j < nr = N/2 (rows of nr): dg[i nr + j] is the trilinear (CIC) density
 * (RETURN_DENS) or, with gaussian != 0, the Gaussian field ln(1+delta)/D +
 * D sigma2/2 with D at the grid cell's radius (RETURN_GAUSS,
 * beaming.c:46-57); v[i nr + j] is the CIC radial velocity 0.5 (N/L)(v.u)
 * times V1(rm) (-f0/(1.5 H0 Om)).  Steps past min((int)(r/dr+0.5), nr-1)
 * stay 0 (calloc, common.c:381-387).  Each row's nr values are scaled once
 * (the reference's loop bound MAX(..., nr-1), srcs.c:725, runs past the row
 * for sources beyond r_max - dr/2). */
void orc_src_skewers(int N, double L, const double obs[3], double r_max, const orc_bg *bg,
                     double fgrowth_0, double hubble_0, double omega_m, const double *dens,
                     const double *npot, int gaussian, double sigma2, long long n, const float *pos4,
                     double *dg, double *vs) {
  const long ngx = 2 * (N / 2 + 1);
  const double idx = N / L;
  const int nr = N / 2, na = bg->na;
  const double dr = r_max / nr, idr = 1. / dr;
  const double factor_vel = -fgrowth_0 / (1.5 * hubble_0 * omega_m);
  const float dxf = (float)(L / N);
#pragma omp parallel for schedule(dynamic, 64)
  for (long long ip = 0; ip < n; ip++) {
    const float *pos = pos4 + 4 * ip;
    const double r = sqrt((double)pos[0] * pos[0] + (double)pos[1] * pos[1] + (double)pos[2] * pos[2]);
    const double ir = 1. / (r > 0.001 ? r : 0.001);
    const double u[3] = {pos[0] * ir, pos[1] * ir, pos[2] * ir};
    int irmax = (int)(r * idr + 0.5);
    if (irmax > nr - 1) irmax = nr - 1;
    double *drow = dg + (size_t)ip * nr, *vrow = vs + (size_t)ip * nr;
    for (int j = 0; j < nr; j++) drow[j] = vrow[j] = 0;
    for (int i_r = 0; i_r <= irmax; i_r++) {
      const double rm = (i_r + 0.5) * dr;
      long i0[3], i1[3];
      double h0[3], h1[3];
      for (int ax = 0; ax < 3; ax++) {
        const double xn = (rm * u[ax] + obs[ax]) * idx;
        i0[ax] = (long)xn;
        h0[ax] = xn - i0[ax];
        h1[ax] = 1 - h0[ax];
        i1[ax] = wrap_cell(i0[ax] + 1, N);
        i0[ax] = wrap_cell(i0[ax], N);
      }
      double v[3] = {0, 0, 0}, d = 0;
      for (int c = 0; c < 8; c++) {
        const int bx = c & 1, by = (c >> 1) & 1, bz = c >> 2;
        const double w = (bz ? h0[2] : h1[2]) * (by ? h0[1] : h1[1]) * (bx ? h0[0] : h1[0]);
        const long cx = bx ? i1[0] : i0[0], cy = by ? i1[1] : i0[1], cz = bz ? i1[2] : i0[2];
        double vc[3];
        grid_element(npot, N, cx, cy, cz, vc, NULL);
        for (int ax = 0; ax < 3; ax++) v[ax] += vc[ax] * w;
        const double dc = dens[cx + cy * ngx + cz * ngx * N];
        if (gaussian) {
          /* beaming.c:51-56: the cell's offsets in float */
          const float x0 = (float)((cx + 0.0) * dxf - obs[0]), y0 = (float)((cy + 0.0) * dxf - obs[1]);
          const float z0 = (float)((cz + 0.0) * dxf - obs[2]);
          const double rc = sqrt((double)(x0 * x0 + y0 * y0 + z0 * z0));
          const double D = orc_f_of_r(bg, bg->d1, 1, bg->d1[na - 1], rc);
          d += (log((double)(1.0f + (float)dc)) / D + D * sigma2 * 0.5) * w;
        } else {
          d += dc * w;
        }
      }
      const double vr = 0.5 * idx * (v[0] * u[0] + v[1] * u[1] + v[2] * u[2]);
      const double vg = orc_f_of_r(bg, bg->v1, bg->v1[0], bg->v1[na - 1], rm);
      drow[i_r] = d;
      vrow[i_r] = (double)(float)vr * (vg * factor_vel);
    }
  }
}

/* cstm_beams_preproc_single + cstm_get_beam_properties_single (cstm.c:39-150)
 * on one rank: per HEALPix pixel centre (NEST order), the march over
 * rm = (ir+1/2) dr, dr = r_max/(N/2), between the first and last step where
 * |K(rm)/IH(rm)| exceeds 1e-4 of its maximum, of K/IH (bias_model(d, b)
 * norm - 1) with d the trilinear (CIC) density (as float), times dr.  kz, bz,
 * norm are the tracer's r-tables (get_bg edges: K 0/0, b b[0]/1, norm
 * norm0/normf).  Returns 1 when the kernel has no support. */
int orc_cstm(int N, double L, const double obs[3], double r_max, const orc_bg *bg, const double *dens,
             const double *kz_arr, const double *bz_arr, const double *norm_arr, double norm0, double normf,
             int model, int nside, double *out) {
  const long ngx = 2 * (N / 2 + 1);
  const double idx = N / L;
  const int nr = N / 2, na = bg->na;
  const double dr = r_max / nr;
  double *kz = malloc(3 * (size_t)nr * sizeof(double)), *bz = kz + nr, *nz = kz + 2 * nr;
  double kmax = -1e100;
  for (int ir = 0; ir < nr; ir++) {
    const double rm = (ir + 0.5) * dr;
    kz[ir] = orc_f_of_r(bg, kz_arr, 0, 0, rm) / orc_f_of_r(bg, bg->ih, bg->ih[0], bg->ih[na - 1], rm);
    bz[ir] = orc_f_of_r(bg, bz_arr, bz_arr[0], 1, rm);
    nz[ir] = orc_f_of_r(bg, norm_arr, norm0, normf, rm);
    if (fabs(kz[ir]) > kmax) kmax = fabs(kz[ir]);
  }
  int irmin = 0, irmax = nr - 1;
  for (int ir = 0; ir < nr; ir++)
    if (fabs(kz[ir]) > 1e-4 * kmax) { irmin = ir; break; }
  for (int ir = nr - 1; ir >= 0; ir--)
    if (fabs(kz[ir]) > 1e-4 * kmax) { irmax = ir; break; }
  if (irmax < irmin) { free(kz); return 1; }
  const long long npix = 12LL * nside * nside;
#pragma omp parallel for schedule(dynamic, 256)
  for (long long ip = 0; ip < npix; ip++) {
    double u[3];
    orc_pix2vec_nest(nside, ip, u);
    double cval = 0;
    for (int irr = irmin; irr <= irmax; irr++) {
      const double rm = (irr + 0.5) * dr;
      long i0[3], i1[3];
      double h0[3], h1[3];
      for (int ax = 0; ax < 3; ax++) {
        const double xn = (rm * u[ax] + obs[ax]) * idx;
        i0[ax] = (long)xn;
        h0[ax] = xn - i0[ax];
        h1[ax] = 1 - h0[ax];
        i1[ax] = wrap_cell(i0[ax] + 1, N);
        i0[ax] = wrap_cell(i0[ax], N);
      }
      double d = 0;
      for (int c = 0; c < 8; c++) {
        const int bx = c & 1, by = (c >> 1) & 1, bzc = c >> 2;
        const double w = (bzc ? h0[2] : h1[2]) * (by ? h0[1] : h1[1]) * (bx ? h0[0] : h1[0]);
        d += dens[(bx ? i1[0] : i0[0]) + (by ? i1[1] : i0[1]) * ngx + (bzc ? i1[2] : i0[2]) * ngx * N] * w;
      }
      cval += kz[irr] * (orc_bias_model((float)d, bz[irr], model) * nz[irr] - 1);
    }
    out[ip] = cval * dr;
  }
  free(kz);
  return 0;
}

/* ========================================================================
 * 1LPT density (density.c:37-175, 376-645)
 * ====================================================================== */

/* 1-D forward DFT (sign -1) in place by conjugation around the inverse */
static void cfft_fwd(int n, double *d) {
  for (int i = 0; i < n; i++) d[2 * i + 1] = -d[2 * i + 1];
  orc_cfft_inv(n, d);
  for (int i = 0; i < n; i++) d[2 * i + 1] = -d[2 * i + 1];
}

/* fftw_wrap_r2c: unnormalised forward transform of the padded real grid,
 * half spectrum out[kz][ky][kx <= N/2] (interleaved complex). */
void orc_r2c_3d(int N, const double *in, double *out) {
  const int NH = N / 2 + 1, ngx = 2 * NH;
  double *line = malloc(2 * (size_t)N * sizeof(double));
  for (int z = 0; z < N; z++)
    for (int y = 0; y < N; y++) {
      const double *row = in + ((size_t)z * N + y) * ngx;
      for (int x = 0; x < N; x++) { line[2 * x] = row[x]; line[2 * x + 1] = 0; }
      cfft_fwd(N, line);
      for (int kx = 0; kx < NH; kx++) {
        const size_t o = 2 * ((size_t)kx + (size_t)NH * ((size_t)y + (size_t)N * z));
        out[o] = line[2 * kx];
        out[o + 1] = line[2 * kx + 1];
      }
    }
  for (int pass = 0; pass < 2; pass++)   /* y, then z */
    for (int a = 0; a < N; a++)
      for (int kx = 0; kx < NH; kx++) {
        for (int b = 0; b < N; b++) {
          const size_t o = pass == 0 ? 2 * ((size_t)kx + (size_t)NH * ((size_t)b + (size_t)N * a))
                                     : 2 * ((size_t)kx + (size_t)NH * ((size_t)a + (size_t)N * b));
          line[2 * b] = out[o];
          line[2 * b + 1] = out[o + 1];
        }
        cfft_fwd(N, line);
        for (int b = 0; b < N; b++) {
          const size_t o = pass == 0 ? 2 * ((size_t)kx + (size_t)NH * ((size_t)b + (size_t)N * a))
                                     : 2 * ((size_t)kx + (size_t)NH * ((size_t)a + (size_t)N * b));
          out[o] = line[2 * b];
          out[o + 1] = line[2 * b + 1];
        }
      }
  free(line);
}

/* k vector of half-spectrum mode (kk, jj, ii) (density.c:405-426) */
static void lpt_kvec(int N, double dk, int kk, int jj, int ii, double kv[3]) {
  kv[0] = kk * dk;
  kv[1] = 2 * jj <= N ? jj * dk : -(N - jj) * dk;
  kv[2] = 2 * ii <= N ? ii * dk : -(N - ii) * dk;
}

/* lpt_1 / lpt_2 on one rank: the displacement i k delta_k / (N^3 k^2) of
 * the Gaussian field dg (padded real) and, for order 2, the second-order
 * displacement -i k S_k / (N^3 k^2) of S = xx yy + xx zz + yy zz - xy^2 -
 * xz^2 - yz^2 (the digradients i k_b psi_a, density.c:714-752); particles at
 * the cell corners moved by D1 psi (+ D2 psi2, d2 = the D2(r) table), positions
 * in float wrapped into [0, L), deposited with NGP, CIC or TSC
 * (pos_2_ngp/cic/tsc, float weights) and out = rho - 1 (padded real, padding
 * 0).  Grids pass through float as the reference stores them.  pos (3 N^3
 * floats, optional) receives the positions. */
int orc_lpt(int order, int N, double L, const double obs[3], const orc_bg *bg, const double *d2, const double *dg,
            int interp, double *out, float *pos) {
  if (interp < 0 || interp > 2 || order < 1 || order > 2 || (order == 2 && !d2)) return 1;
  const int NH = N / 2 + 1, ngx = 2 * NH, na = bg->na;
  const size_t nh = (size_t)N * N * NH, ng = (size_t)N * N * ngx;
  double *H = malloc(2 * nh * sizeof(double)), *S = malloc(2 * nh * sizeof(double));
  double *disp = malloc(3 * ng * sizeof(double)), *sec = order == 2 ? calloc(3 * ng, sizeof(double)) : NULL;
  orc_r2c_3d(N, dg, H);
  const double dk = 2 * M_PI / L, fftnorm = (double)N * ((long)N * N);
  if (order == 2) {
    /* the six digradients from the FP32-stored first-order spectra */
    double *P = malloc(6 * nh * sizeof(double)), *G = malloc(6 * ng * sizeof(double));
    for (int ii = 0; ii < N; ii++)
      for (int jj = 0; jj < N; jj++)
        for (int kk = 0; kk < NH; kk++) {
          double kv[3];
          lpt_kvec(N, dk, kk, jj, ii, kv);
          const size_t o = 2 * ((size_t)kk + (size_t)NH * ((size_t)jj + (size_t)N * ii));
          const double km2 = fftnorm * (kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
          for (int ax = 0; ax < 3; ax++) {
            P[2 * ax * nh + o] = km2 <= 0 ? 0 : (float)(-kv[ax] * H[o + 1] / km2);
            P[2 * ax * nh + o + 1] = km2 <= 0 ? 0 : (float)(kv[ax] * H[o] / km2);
          }
        }
    static const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < 6; q++) {
      for (int ii = 0; ii < N; ii++)
        for (int jj = 0; jj < N; jj++)
          for (int kk = 0; kk < NH; kk++) {
            double kv[3];
            lpt_kvec(N, dk, kk, jj, ii, kv);
            const size_t o = 2 * ((size_t)kk + (size_t)NH * ((size_t)jj + (size_t)N * ii));
            const double *pp = P + 2 * pa[q] * nh + o;
            S[o] = (float)(-kv[pb[q]] * pp[1]);
            S[o + 1] = (float)(kv[pb[q]] * pp[0]);
          }
      orc_c2r_3d(N, S, G + q * ng);
      for (size_t i = 0; i < ng; i++) G[q * ng + i] = (float)G[q * ng + i];
    }
    double *src = malloc(ng * sizeof(double));
    for (size_t i = 0; i < ng; i++) src[i] = 0;
    for (int iz = 0; iz < N; iz++)
      for (int iy = 0; iy < N; iy++)
        for (int ix = 0; ix < N; ix++) {
          const size_t i = (size_t)ix + (size_t)ngx * ((size_t)iy + (size_t)N * iz);
          const float xx = G[i], xy = G[ng + i], xz = G[2 * ng + i], yy = G[3 * ng + i], yz = G[4 * ng + i],
                      zz = G[5 * ng + i];
          src[i] = xx * yy + xx * zz + yy * zz - xy * xy - xz * xz - yz * yz;
        }
    orc_r2c_3d(N, src, H);   /* H now holds S_k; the first-order spectra are in P */
    for (int ax = 0; ax < 3; ax++) {
      for (int ii = 0; ii < N; ii++)
        for (int jj = 0; jj < N; jj++)
          for (int kk = 0; kk < NH; kk++) {
            double kv[3];
            lpt_kvec(N, dk, kk, jj, ii, kv);
            const size_t o = 2 * ((size_t)kk + (size_t)NH * ((size_t)jj + (size_t)N * ii));
            const double km2 = fftnorm * (kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
            const double hr = (float)H[o], hi = (float)H[o + 1];
            S[o] = km2 <= 0 ? 0 : (float)(kv[ax] * hi / km2);
            S[o + 1] = km2 <= 0 ? 0 : (float)(-kv[ax] * hr / km2);
          }
      orc_c2r_3d(N, S, sec + ax * ng);
      orc_c2r_3d(N, P + 2 * ax * nh, disp + ax * ng);
    }
    free(P); free(G); free(src);
  }
  for (int ax = 0; ax < 3 && order == 1; ax++) {
    for (int ii = 0; ii < N; ii++) {
      const double kz = 2 * ii <= N ? ii * dk : -(N - ii) * dk;
      for (int jj = 0; jj < N; jj++) {
        const double ky = 2 * jj <= N ? jj * dk : -(N - jj) * dk;
        for (int kk = 0; kk < NH; kk++) {
          const double kx = kk * dk;
          const size_t o = 2 * ((size_t)kk + (size_t)NH * ((size_t)jj + (size_t)N * ii));
          const double km2 = fftnorm * (kx * kx + ky * ky + kz * kz);
          const double kv = ax == 0 ? kx : (ax == 1 ? ky : kz);
          if (km2 <= 0) { S[o] = S[o + 1] = 0; continue; }
          /* i kv (re + i im) / km2 */
          S[o] = -kv * H[o + 1] / km2;
          S[o + 1] = kv * H[o] / km2;
        }
      }
    }
    orc_c2r_3d(N, S, disp + ax * ng);
  }
  for (size_t i = 0; i < ng; i++) out[i] = 0;
  const float dxf = (float)(L / N), iag = (float)(N / L);
  for (int iz = 0; iz < N; iz++)
    for (int iy = 0; iy < N; iy++)
      for (int ix = 0; ix < N; ix++) {
        const size_t idx = (size_t)ix + (size_t)ngx * ((size_t)iy + (size_t)N * iz);
        float xv[3];
        xv[0] = (float)((ix + 0.0) * dxf - obs[0]);
        xv[1] = (float)((iy + 0.0) * dxf - obs[1]);
        xv[2] = (float)((iz + 0.0) * dxf - obs[2]);
        const double r = sqrt((double)(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]));
        const double d1 = orc_f_of_r(bg, bg->d1, 1, bg->d1[na - 1], r);
        const double d2g = order == 2 ? orc_f_of_r(bg, d2, d2[0], d2[na - 1], r) : 0;
        float p[3];
        for (int ax = 0; ax < 3; ax++) {
          float q = order == 1 ? (float)(xv[ax] + d1 * (float)disp[ax * ng + idx] + obs[ax])
                               : (float)(xv[ax] + d1 * (float)disp[ax * ng + idx] +
                                         d2g * (float)sec[ax * ng + idx] + obs[ax]);
          if (q < 0) q = (float)(q + L);
          if (q >= L) q = (float)(q - L);
          p[ax] = q;
        }
        if (pos) {
          float *pp = pos + 3 * ((size_t)ix + (size_t)N * ((size_t)iy + (size_t)N * iz));
          pp[0] = p[0]; pp[1] = p[1]; pp[2] = p[2];
        }
        if (interp == 0) {
          int i0[3];
          for (int ax = 0; ax < 3; ax++) {
            i0[ax] = (int)((double)(p[ax] * iag) + 0.5);
            if (i0[ax] >= N) i0[ax] -= N;
            if (i0[ax] < 0) i0[ax] += N;
          }
          out[i0[0] + ngx * ((size_t)i0[1] + (size_t)N * i0[2])] += 1.;
        } else if (interp == 1) {
          int i0[3], i1[3];
          float a0[3], a1[3];
          for (int ax = 0; ax < 3; ax++) {
            i0[ax] = (int)(p[ax] * iag);
            a1[ax] = p[ax] * iag - i0[ax];
            a0[ax] = 1 - a1[ax];
            i1[ax] = i0[ax] + 1;
            if (i0[ax] < 0) i0[ax] += N;
            if (i1[ax] < 0) i1[ax] += N;
            if (i0[ax] >= N) i0[ax] -= N;
            if (i1[ax] >= N) i1[ax] -= N;
          }
          for (int c = 0; c < 8; c++) {
            const int bx = c & 1, by = (c >> 1) & 1, bz = c >> 2;
            const float w = (bx ? a1[0] : a0[0]) * (by ? a1[1] : a0[1]) * (bz ? a1[2] : a0[2]);
            out[(bx ? i1[0] : i0[0]) + ngx * ((size_t)(by ? i1[1] : i0[1]) + (size_t)N * (bz ? i1[2] : i0[2]))] += w;
          }
        } else {
          int ic[3][3];
          float wt[3][3];
          for (int ax = 0; ax < 3; ax++) {
            const int i0 = (int)floorf((float)((double)(p[ax] * iag) + 0.5));
            const float a = p[ax] * iag - i0;
            wt[ax][0] = (float)(0.5 * (0.5 - a) * (0.5 - a));
            wt[ax][2] = (float)(0.5 * (0.5 + a) * (0.5 + a));
            wt[ax][1] = (float)(0.75 - a * a);
            for (int s = 0; s < 3; s++) {
              int v = i0 + s - 1;
              if (v < 0) v += N;
              if (v >= N) v -= N;
              ic[ax][s] = v;
            }
          }
          for (int sz = 0; sz < 3; sz++)
            for (int sy = 0; sy < 3; sy++)
              for (int sx = 0; sx < 3; sx++)
                out[ic[0][sx] + ngx * ((size_t)ic[1][sy] + (size_t)N * ic[2][sz])] +=
                    wt[0][sx] * wt[1][sy] * wt[2][sz];
        }
      }
  for (int iz = 0; iz < N; iz++)
    for (int iy = 0; iy < N; iy++)
      for (int ix = 0; ix < N; ix++) {
        const size_t idx = (size_t)ix + (size_t)ngx * ((size_t)iy + (size_t)N * iz);
        out[idx] = out[idx] - 1.;
      }
  free(H); free(S); free(disp); free(sec);
  return 0;
}

/* get_r_index_imap (imap.c:76-103), walk bounded to 2 nr steps */
static int r_index_imap(int nr, const double *r0, const double *rf, double r, int start) {
  int ir0 = start < 0 ? 0 : (start >= nr ? nr - 1 : start);
  for (int it = 0; it < 2 * nr + 4; it++) {
    if (ir0 == -1 || ir0 == nr) return ir0;
    if (r < r0[ir0]) ir0--;
    else if (r >= rf[ir0]) ir0++;
    else return ir0;
  }
  return -2;
}

/* imap_set_cartesian_single (imap.c:135-245) */
void orc_imap(int N, double L, const double obs[3], const orc_bg *bg, const double *dens,
              const double *npot, const double *tz, const double *bz, double bz0,
              const double *norm, double norm0, double normf, int model,
              double fgrowth_0, double hubble_0, double omega_m,
              int nside, int nr, const double *r0u, const double *rfu, double *data, int *nadd) {
  const long ngx = 2 * (N / 2 + 1);
  int *ord = malloc(nr * sizeof(int));
  sort_idx(nr, r0u, ord);
  double *r0 = malloc(nr * sizeof(double)), *rf = malloc(nr * sizeof(double));
  for (int i = 0; i < nr; i++) { r0[i] = r0u[ord[i]]; rf[i] = rfu[ord[i]]; }
  const long long npix = 12LL * nside * nside;
  memset(data, 0, (size_t)npix * nr * sizeof(double));
  memset(nadd, 0, (size_t)npix * nr * sizeof(int));
  const double dx = L / N;
  const double factor_vel = -fgrowth_0 / (1.5 * hubble_0 * omega_m);
  const double pixel_size = sqrt(4 * M_PI / npix);
  const double rmin_here = r0[0] - 20., rmax_here = rf[nr - 1] + 20.;
  int *nsub = malloc(nr * sizeof(int));
  for (int i = 0; i < nr; i++) {
    const double vol = (rf[i] * rf[i] * rf[i] - r0[i] * r0[i] * r0[i]) * pixel_size * pixel_size / 3;
    const double sct = r0[i] * pixel_size, scr = rf[i] - r0[i], scv = pow(vol, 0.333333);
    const double sc0 = fmin(scv, fmin(sct, scr));
    nsub[i] = (int)(dx / sc0 + 0.5) + 1;
  }
  const int na = bg->na;
  for (int iz = 0; iz < N; iz++) {
    int irad = 0;
    const double z0 = iz * dx - obs[2];
    for (int iy = 0; iy < N; iy++) {
      const double y0 = iy * dx - obs[1];
      for (int ix = 0; ix < N; ix++) {
        const double x0 = ix * dx - obs[0];
        const long index = ix + ngx * (iy + (long)N * iz);
        const double rr = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
        const double tmean = orc_f_of_r(bg, tz, 0, 0, rr);
        if (!(rr <= rmax_here && rr >= rmin_here)) continue;
        if (!(tmean > 0)) continue;
        const double bias = orc_f_of_r(bg, bz, bz0, 1, rr);
        const double dnorm = orc_f_of_r(bg, norm, norm0, normf, rr);
        const double rv = rr > 0 ? factor_vel * rvel(N, L, npot, ix, iy, iz, x0, y0, z0, rr) : 0;
        const double dr_rsd = rv * orc_f_of_r(bg, bg->v1, bg->v1[0], bg->v1[na - 1], rr) *
                              orc_f_of_r(bg, bg->ih, bg->ih[0], bg->ih[na - 1], rr);
        const double temp = tmean * orc_bias_model(dens[index], bias, model) * dnorm;
        irad = r_index_imap(nr, r0, rf, rr, irad);
        const int ns = irad < 0 ? nsub[0] : (irad >= nr ? nsub[nr - 1] : nsub[irad]);
        const double dxs = dx / ns;
        for (int izz = 0; izz < ns; izz++) {
          const double z = z0 + (izz + 0.5) * dxs;
          for (int iyy = 0; iyy < ns; iyy++) {
            const double y = y0 + (iyy + 0.5) * dxs;
            for (int ixx = 0; ixx < ns; ixx++) {
              const double x = x0 + (ixx + 0.5) * dxs;
              const double r = sqrt(x * x + y * y + z * z);
              double cth = 1, phi = 0;
              if (r != 0) {
                cth = z / r;
                phi = atan2(y / r, x / r);
                if (phi < 0) phi += 2 * M_PI;
              }
              irad = r_index_imap(nr, r0, rf, r + dr_rsd, irad);
              if (irad >= 0 && irad < nr) {
                const long long pix = orc_ang2pix_ring(nside, cth, phi);
                data[(size_t)irad * npix + pix] += temp;
                nadd[(size_t)irad * npix + pix]++;
              }
            }
          }
        }
      }
    }
  }
  free(ord); free(r0); free(rf); free(nsub);
}

/* ========================================================================
 * HEALPix (Gorski et al. 2005)
 * ====================================================================== */
static const int JRLL[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
static const int JPLL[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};

long long orc_ang2pix_ring(long long nside, double cth, double phi) {
  const double twopi = 2 * M_PI;
  const double za = fabs(cth);
  double ph = phi;
  if (ph < 0 || ph >= twopi) {
    ph = fmod(ph, twopi);
    if (ph < 0) { ph += twopi; if (ph == twopi) ph = 0; }
  }
  const double tt = ph / (0.5 * M_PI);
  if (za <= 2. / 3.) {
    const double t1 = nside * (0.5 + tt), t2 = nside * cth * 0.75;
    const long long jp = (long long)(t1 - t2), jm = (long long)(t1 + t2);
    const long long ir = nside + 1 + jp - jm;
    const long long kshift = 1 - (ir & 1);
    long long ip = (jp + jm - nside + kshift + 1) / 2;
    ip %= 4 * nside; if (ip < 0) ip += 4 * nside;
    return nside * (nside - 1) * 2 + (ir - 1) * 4 * nside + ip;
  }
  const double tp = tt - (int)tt;
  const double tmp = nside * sqrt(3 * (1 - za));
  const long long jp = (long long)(tp * tmp), jm = (long long)((1.0 - tp) * tmp);
  const long long ir = jp + jm + 1;
  long long ip = (long long)(tt * ir);
  ip %= 4 * ir; if (ip < 0) ip += 4 * ir;
  if (cth > 0) return 2 * ir * (ir - 1) + ip;
  return 12 * nside * nside - 2 * ir * (ir + 1) + ip;
}

long long orc_vec2pix_ring(long long nside, const double *v) {
  const double vlen = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  return orc_ang2pix_ring(nside, v[2] / vlen, atan2(v[1], v[0]));
}

static long long isqrt_ll(long long v) {
  long long r = (long long)sqrt((double)v + 0.5);
  while (r * r > v) r--;
  while ((r + 1) * (r + 1) <= v) r++;
  return r;
}

static void nest2xyf(long long nside, long long pix, long long *ix, long long *iy, int *face) {
  const long long np = nside * nside;
  *face = (int)(pix / np);
  long long ipf = pix % np, x = 0, y = 0;
  for (int b = 0; ipf; b++, ipf >>= 2) {
    x |= (ipf & 1) << b;
    y |= ((ipf >> 1) & 1) << b;
  }
  *ix = x; *iy = y;
}
static long long xyf2nest(long long nside, long long ix, long long iy, int face) {
  long long r = 0;
  for (int b = 0; (ix >> b) || (iy >> b); b++) r |= (((ix >> b) & 1) << (2 * b)) | (((iy >> b) & 1) << (2 * b + 1));
  return (long long)face * nside * nside + r;
}
static void ring2xyf(long long nside, long long pix, long long *ix, long long *iy, int *face) {
  const long long ncap = 2 * nside * (nside - 1), npix = 12 * nside * nside, nl2 = 2 * nside;
  long long iring, iphi, kshift, nr;
  if (pix < ncap) {
    iring = (1 + isqrt_ll(1 + 2 * pix)) >> 1;
    iphi = (pix + 1) - 2 * iring * (iring - 1);
    kshift = 0; nr = iring;
    *face = (int)((iphi - 1) / nr);
  } else if (pix < npix - ncap) {
    const long long ip = pix - ncap, tmp = ip / (4 * nside);
    iring = tmp + nside;
    iphi = ip - tmp * 4 * nside + 1;
    kshift = (iring + nside) & 1;
    nr = nside;
    const long long ire = tmp + 1, irm = nl2 + 1 - tmp;
    const long long ifm = (iphi - ire / 2 + nside - 1) / nside, ifp = (iphi - irm / 2 + nside - 1) / nside;
    *face = (ifp == ifm) ? (int)(ifp | 4) : ((ifp < ifm) ? (int)ifp : (int)(ifm + 8));
  } else {
    const long long ip = npix - pix;
    iring = (1 + isqrt_ll(2 * ip - 1)) >> 1;
    iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
    kshift = 0; nr = iring;
    iring = 2 * nl2 - iring;
    *face = (int)(8 + (iphi - 1) / nr);
  }
  const long long irt = iring - ((2 + (*face >> 2)) * nside) + 1;
  long long ipt = 2 * iphi - JPLL[*face] * nr - kshift - 1;
  if (ipt >= nl2) ipt -= 8 * nside;
  *ix = (ipt - irt) >> 1;
  *iy = (-ipt - irt) >> 1;
}
static long long xyf2ring(long long nside, long long ix, long long iy, int face) {
  const long long nl4 = 4 * nside, npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
  const long long jr = JRLL[face] * nside - ix - iy - 1;
  long long nr, kshift, nb;
  if (jr < nside) { nr = jr; nb = 2 * nr * (nr - 1); kshift = 0; }
  else if (jr > 3 * nside) { nr = nl4 - jr; nb = npix - 2 * (nr + 1) * nr; kshift = 0; }
  else { nr = nside; nb = ncap + (jr - nside) * nl4; kshift = (jr - nside) & 1; }
  long long jp = (JPLL[face] * nr + ix - iy + 1 + kshift) / 2;
  if (jp > nl4) jp -= nl4; else if (jp < 1) jp += nl4;
  return nb + jp - 1;
}
long long orc_ring2nest(long long nside, long long pix) {
  long long ix, iy; int f;
  ring2xyf(nside, pix, &ix, &iy, &f);
  return xyf2nest(nside, ix, iy, f);
}
long long orc_nest2ring(long long nside, long long pix) {
  long long ix, iy; int f;
  nest2xyf(nside, pix, &ix, &iy, &f);
  return xyf2ring(nside, ix, iy, f);
}
void orc_pix2vec_nest(long long nside, long long pix, double *v) {
  long long ix, iy; int face;
  nest2xyf(nside, pix, &ix, &iy, &face);
  const long long nl4 = 4 * nside;
  const double fact2 = 4.0 / (12.0 * (double)nside * (double)nside);
  const long long jr = JRLL[face] * nside - ix - iy - 1;
  long long nr, kshift;
  double z;
  if (jr < nside) { nr = jr; z = 1 - (double)(nr * nr) * fact2; kshift = 0; }
  else if (jr > 3 * nside) { nr = nl4 - jr; z = (double)(nr * nr) * fact2 - 1; kshift = 0; }
  else { const double fact1 = (double)(nside << 1) * fact2; nr = nside; z = (double)(2 * nside - jr) * fact1; kshift = (jr - nside) & 1; }
  long long jp = (JPLL[face] * nr + ix - iy + 1 + kshift) / 2;
  if (jp > nl4) jp -= nl4;
  if (jp < 1) jp += nl4;
  const double phi = ((double)jp - (double)(kshift + 1) * 0.5) * (0.5 * M_PI / (double)nr);
  const double st = sqrt((1.0 - z) * (1.0 + z));
  v[0] = st * cos(phi);
  v[1] = st * sin(phi);
  v[2] = z;
}

/* ========================================================================
 * The surrogate network (no reference counterpart): 4 -> 64 -> 64 -> 2,
 * tanh hidden layers, FP64.  P = W1[64][4], b1[64], W2[64][64], b2[64],
 * W3[2][64], b3[2]; x_n = (x - xm)/xs, y = ym + ys out.
 * ====================================================================== */
enum { MI = 4, MH = 64, MO = 2, oW1 = 0, oB1 = MH * MI, oW2 = oB1 + MH, oB2 = oW2 + MH * MH, oW3 = oB2 + MH,
       oB3 = oW3 + MO * MH, MNP = oB3 + MO };

static void mlp_fwd1(const double *P, const double *xn, double *h1, double *h2, double *out) {
  for (int o = 0; o < MH; o++) {
    double a = P[oB1 + o];
    for (int i = 0; i < MI; i++) a += P[oW1 + o * MI + i] * xn[i];
    h1[o] = tanh(a);
  }
  for (int o = 0; o < MH; o++) {
    double a = P[oB2 + o];
    for (int i = 0; i < MH; i++) a += P[oW2 + o * MH + i] * h1[i];
    h2[o] = tanh(a);
  }
  for (int k = 0; k < MO; k++) {
    double a = P[oB3 + k];
    for (int i = 0; i < MH; i++) a += P[oW3 + k * MH + i] * h2[i];
    out[k] = a;
  }
}

void orc_mlp_forward(const double *P, const double *xm, const double *xs, const double *ym, const double *ys,
                     const double *x, int n, double *y) {
  for (int s = 0; s < n; s++) {
    double xn[MI], h1[MH], h2[MH], out[MO];
    for (int i = 0; i < MI; i++) xn[i] = (x[s * MI + i] - xm[i]) / xs[i];
    mlp_fwd1(P, xn, h1, h2, out);
    for (int k = 0; k < MO; k++) y[s * MO + k] = ym[k] + ys[k] * out[k];
  }
}

/* L = (1/2n) sum_s sum_k (out_k - t_k)^2 with t = (y - ym)/ys; grad[MNP]. */
double orc_mlp_loss_grad(const double *P, const double *xm, const double *xs, const double *ym, const double *ys,
                         const double *x, const double *y, int n, double *grad) {
  memset(grad, 0, MNP * sizeof(double));
  double loss = 0;
  for (int s = 0; s < n; s++) {
    double xn[MI], h1[MH], h2[MH], out[MO], g3[MO], g2[MH], g1[MH];
    for (int i = 0; i < MI; i++) xn[i] = (x[s * MI + i] - xm[i]) / xs[i];
    mlp_fwd1(P, xn, h1, h2, out);
    for (int k = 0; k < MO; k++) {
      const double e = out[k] - (y[s * MO + k] - ym[k]) / ys[k];
      loss += 0.5 * e * e / n;
      g3[k] = e / n;
      grad[oB3 + k] += g3[k];
      for (int i = 0; i < MH; i++) grad[oW3 + k * MH + i] += g3[k] * h2[i];
    }
    for (int i = 0; i < MH; i++) {
      double a = 0;
      for (int k = 0; k < MO; k++) a += P[oW3 + k * MH + i] * g3[k];
      g2[i] = a * (1 - h2[i] * h2[i]);
    }
    for (int o = 0; o < MH; o++) {
      grad[oB2 + o] += g2[o];
      for (int i = 0; i < MH; i++) grad[oW2 + o * MH + i] += g2[o] * h1[i];
    }
    for (int i = 0; i < MH; i++) {
      double a = 0;
      for (int o = 0; o < MH; o++) a += P[oW2 + o * MH + i] * g2[o];
      g1[i] = a * (1 - h1[i] * h1[i]);
    }
    for (int o = 0; o < MH; o++) {
      grad[oB1 + o] += g1[o];
      for (int i = 0; i < MI; i++) grad[oW1 + o * MI + i] += g1[o] * xn[i];
    }
  }
  return loss;
}

