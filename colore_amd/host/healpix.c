/* healpix.c -- HEALPix pixel functions on the host (Gorski et al. 2005,
 * ApJ 622, 759): the RING index of a direction (he_ang2pix,
 * healpix_extra.c:141-172), RING<->NEST through (face, ix, iy), and NEST
 * pixel centres (the chealpix calls of common.c:530-540 and
 * healpix_extra.c:292-317). */
#include "clr_host.h"
#include <math.h>

static const int jrll_t[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
static const int jpll_t[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};

long long clr_hp_ang2pix_ring(long long nside, double cth, double phi) {
  const double twopi = 6.283185307179586476925286766559;
  const double za = fabs(cth);
  double ph = phi;
  if (ph < 0 || ph >= twopi) {
    ph = fmod(ph, twopi);
    if (ph < 0) { ph += twopi; if (ph == twopi) ph = 0; }
  }
  const double tt = ph * 0.6366197723675813430755350534900574;
  if (za <= 2.0 / 3.0) {
    const double t1 = nside * (0.5 + tt), t2 = nside * cth * 0.75;
    const long long jp = (long long)(t1 - t2), jm = (long long)(t1 + t2);
    const long long ir = nside + 1 + jp - jm, kshift = 1 - (ir & 1);
    long long ip = (jp + jm - nside + kshift + 1) / 2;
    ip %= 4 * nside; if (ip < 0) ip += 4 * nside;
    return nside * (nside - 1) * 2 + (ir - 1) * 4 * nside + ip;
  }
  const double tp = tt - (int)tt, tmp = nside * sqrt(3 * (1 - za));
  const long long jp = (long long)(tp * tmp), jm = (long long)((1.0 - tp) * tmp);
  const long long ir = jp + jm + 1;
  long long ip = (long long)(tt * ir);
  ip %= 4 * ir; if (ip < 0) ip += 4 * ir;
  return cth > 0 ? 2 * ir * (ir - 1) + ip : 12 * nside * nside - 2 * ir * (ir + 1) + ip;
}

long long clr_hp_vec2pix_ring(long long nside, const double *v) {
  const double vlen = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  return clr_hp_ang2pix_ring(nside, v[2] / vlen, atan2(v[1], v[0]));
}

static long long isq(long long v) {
  long long r = (long long)sqrt((double)v + 0.5);
  while (r * r > v) r--;
  while ((r + 1) * (r + 1) <= v) r++;
  return r;
}

static void nest_xyf(long long nside, long long pix, long long *ix, long long *iy, int *f) {
  const long long np = nside * nside;
  long long ipf = pix % np, x = 0, y = 0;
  *f = (int)(pix / np);
  for (int b = 0; ipf; b++, ipf >>= 2) { x |= (ipf & 1) << b; y |= ((ipf >> 1) & 1) << b; }
  *ix = x; *iy = y;
}

static long long xyf_nest(long long nside, long long ix, long long iy, int f) {
  long long r = 0;
  for (int b = 0; (ix >> b) || (iy >> b); b++) r |= (((ix >> b) & 1) << (2 * b)) | (((iy >> b) & 1) << (2 * b + 1));
  return (long long)f * nside * nside + r;
}

static void ring_xyf(long long nside, long long pix, long long *ix, long long *iy, int *f) {
  const long long ncap = 2 * nside * (nside - 1), npix = 12 * nside * nside, nl2 = 2 * nside;
  long long iring, iphi, kshift, nr;
  if (pix < ncap) {
    iring = (1 + isq(1 + 2 * pix)) >> 1;
    iphi = (pix + 1) - 2 * iring * (iring - 1);
    kshift = 0; nr = iring; *f = (int)((iphi - 1) / nr);
  } else if (pix < npix - ncap) {
    const long long ip = pix - ncap, tmp = ip / (4 * nside);
    iring = tmp + nside; iphi = ip - tmp * 4 * nside + 1; kshift = (iring + nside) & 1; nr = nside;
    const long long ire = tmp + 1, irm = nl2 + 1 - tmp;
    const long long ifm = (iphi - ire / 2 + nside - 1) / nside, ifp = (iphi - irm / 2 + nside - 1) / nside;
    *f = (ifp == ifm) ? (int)(ifp | 4) : ((ifp < ifm) ? (int)ifp : (int)(ifm + 8));
  } else {
    const long long ip = npix - pix;
    iring = (1 + isq(2 * ip - 1)) >> 1;
    iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
    kshift = 0; nr = iring; iring = 2 * nl2 - iring; *f = (int)(8 + (iphi - 1) / nr);
  }
  const long long irt = iring - ((2 + (*f >> 2)) * nside) + 1;
  long long ipt = 2 * iphi - jpll_t[*f] * nr - kshift - 1;
  if (ipt >= nl2) ipt -= 8 * nside;
  *ix = (ipt - irt) >> 1;
  *iy = (-ipt - irt) >> 1;
}

static long long xyf_ring(long long nside, long long ix, long long iy, int f) {
  const long long nl4 = 4 * nside, npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
  const long long jr = jrll_t[f] * nside - ix - iy - 1;
  long long nr, kshift, nb;
  if (jr < nside) { nr = jr; nb = 2 * nr * (nr - 1); kshift = 0; }
  else if (jr > 3 * nside) { nr = nl4 - jr; nb = npix - 2 * (nr + 1) * nr; kshift = 0; }
  else { nr = nside; nb = ncap + (jr - nside) * nl4; kshift = (jr - nside) & 1; }
  long long jp = (jpll_t[f] * nr + ix - iy + 1 + kshift) / 2;
  if (jp > nl4) jp -= nl4; else if (jp < 1) jp += nl4;
  return nb + jp - 1;
}

long long clr_hp_ring2nest(long long nside, long long pix) {
  long long ix, iy; int f;
  ring_xyf(nside, pix, &ix, &iy, &f);
  return xyf_nest(nside, ix, iy, f);
}

long long clr_hp_nest2ring(long long nside, long long pix) {
  long long ix, iy; int f;
  nest_xyf(nside, pix, &ix, &iy, &f);
  return xyf_ring(nside, ix, iy, f);
}

void clr_hp_pix2vec_nest(long long nside, long long pix, double *v) {
  long long ix, iy; int f;
  nest_xyf(nside, pix, &ix, &iy, &f);
  const long long nl4 = 4 * nside;
  const double fact2 = 4.0 / (12.0 * (double)nside * (double)nside);
  const long long jr = jrll_t[f] * nside - ix - iy - 1;
  long long nr, kshift;
  double z;
  if (jr < nside) { nr = jr; z = 1 - (double)(nr * nr) * fact2; kshift = 0; }
  else if (jr > 3 * nside) { nr = nl4 - jr; z = (double)(nr * nr) * fact2 - 1; kshift = 0; }
  else { nr = nside; z = (double)(2 * nside - jr) * ((double)(nside << 1) * fact2); kshift = (jr - nside) & 1; }
  long long jp = (jpll_t[f] * nr + ix - iy + 1 + kshift) / 2;
  if (jp > nl4) jp -= nl4;
  if (jp < 1) jp += nl4;
  const double phi = ((double)jp - (double)(kshift + 1) * 0.5) * (1.5707963267948966 / (double)nr);
  const double st = sqrt((1.0 - z) * (1.0 + z));
  v[0] = st * cos(phi);
  v[1] = st * sin(phi);
  v[2] = z;
}
