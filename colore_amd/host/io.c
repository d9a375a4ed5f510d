/* io.c -- the reference's stdout conventions (print_info, timer:
 * common.c:114-168, 276-288) and its output files (io.c:565-1236,
 * healpix_extra.c:4-57).  FITS files are written directly (2880-byte
 * blocks, big-endian binary tables) so no FITS library is needed; the
 * headers carry the keys the reference's cfitsio calls produce. */
#include "clr_host.h"
#include <stdarg.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

static double now_ms(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec * 1000.0 + tv.tv_usec / 1000.0;
}
static double relbeg, absbeg;

void clr_timer(int i) {
  if (i == 0) relbeg = now_ms();
  else if (i == 1) printf(">    Relative time ellapsed %.1lf ms\n", now_ms() - relbeg);
  else if (i == 2) { printf(">    Relative time ellapsed %.1lf ms\n", now_ms() - relbeg); relbeg = now_ms(); }
  else if (i == 4) absbeg = now_ms();
  else if (i == 5) printf(">    Total time ellapsed %.1lf ms\n", now_ms() - absbeg);
}

void clr_print_info(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vprintf(fmt, ap);
  va_end(ap);
}

/* native density format: int NNodes, int sizeof(flouble), double L, int N,
 * int nz, int iz0, then nz*N rows of N values (padding stripped) */
int clr_write_density_grid(const char *fname, int N, double L, const float *g) {
  FILE *fo = fopen(fname, "wb");
  if (!fo) return 1;
  const int nn = 1, sz = (int)sizeof(float), nz = N, iz0 = 0;
  int ok = fwrite(&nn, sizeof(int), 1, fo) == 1 && fwrite(&sz, sizeof(int), 1, fo) == 1 &&
           fwrite(&L, sizeof(double), 1, fo) == 1 && fwrite(&N, sizeof(int), 1, fo) == 1 &&
           fwrite(&nz, sizeof(int), 1, fo) == 1 && fwrite(&iz0, sizeof(int), 1, fo) == 1;
  const size_t ngx = 2 * (size_t)(N / 2 + 1);
  for (size_t row = 0; ok && row < (size_t)N * N; row++) ok = fwrite(g + row * ngx, sizeof(float), N, fo) == (size_t)N;
  fclose(fo);
  return ok ? 0 : 1;
}

/* ---- FITS ------------------------------------------------------------- */
typedef struct { char *buf; size_t n, cap; } hdr;

static void card(hdr *h, const char *s) {
  char c[81];
  memset(c, ' ', 80);
  c[80] = 0;
  memcpy(c, s, strlen(s) > 80 ? 80 : strlen(s));
  if (h->n + 80 > h->cap) { h->cap = h->cap ? 2 * h->cap : 2880; h->buf = realloc(h->buf, h->cap); }
  memcpy(h->buf + h->n, c, 80);
  h->n += 80;
}
static void kv_str(hdr *h, const char *k, const char *v, const char *cm) {
  char s[200], q[100];
  snprintf(q, sizeof(q), "'%-8s'", v);
  snprintf(s, sizeof(s), "%-8.8s= %-20s%s%s", k, q, cm ? " / " : "", cm ? cm : "");
  card(h, s);
}
static void kv_int(hdr *h, const char *k, long long v, const char *cm) {
  char s[200];
  snprintf(s, sizeof(s), "%-8.8s= %20lld%s%s", k, v, cm ? " / " : "", cm ? cm : "");
  card(h, s);
}
static void kv_log(hdr *h, const char *k, int v, const char *cm) {
  char s[200];
  snprintf(s, sizeof(s), "%-8.8s= %20s%s%s", k, v ? "T" : "F", cm ? " / " : "", cm ? cm : "");
  card(h, s);
}
static int flush_hdr(FILE *f, hdr *h) {
  card(h, "END");
  const size_t pad = (2880 - h->n % 2880) % 2880;
  int ok = fwrite(h->buf, 1, h->n, f) == h->n;
  char sp[2880];
  memset(sp, ' ', sizeof(sp));
  ok = ok && fwrite(sp, 1, pad, f) == pad;
  free(h->buf);
  h->buf = NULL; h->n = h->cap = 0;
  return ok;
}
static void be32(void *dst, const void *src) {
  const uint8_t *s = src;
  uint8_t *d = dst;
  d[0] = s[3]; d[1] = s[2]; d[2] = s[1]; d[3] = s[0];
}
static int write_primary(FILE *f) {
  hdr h = {0};
  kv_log(&h, "SIMPLE", 1, "file does conform to FITS standard");
  kv_int(&h, "BITPIX", 8, "number of bits per data pixel");
  kv_int(&h, "NAXIS", 0, "number of data axes");
  kv_log(&h, "EXTEND", 1, "FITS dataset may contain extensions");
  return flush_hdr(f, &h);
}
/* binary table of ncol 4-byte columns (types: 'J' int32 or 'E' float32) */
static int write_table(FILE *f, int ncol, const char **ttype, const char *tform, const char **tunit,
                       const char *extname, long long nrow, const void *const *cols, hdr *extra) {
  hdr h = {0};
  char k[16], fm[8];
  kv_str(&h, "XTENSION", "BINTABLE", "binary table extension");
  kv_int(&h, "BITPIX", 8, "8-bit bytes");
  kv_int(&h, "NAXIS", 2, "2-dimensional binary table");
  kv_int(&h, "NAXIS1", 4LL * ncol, "width of table in bytes");
  kv_int(&h, "NAXIS2", nrow, "number of rows in table");
  kv_int(&h, "PCOUNT", 0, "size of special data area");
  kv_int(&h, "GCOUNT", 1, "one data group (required keyword)");
  kv_int(&h, "TFIELDS", ncol, "number of fields in each row");
  for (int c = 0; c < ncol; c++) {
    snprintf(k, sizeof(k), "TTYPE%d", c + 1);
    kv_str(&h, k, ttype[c], NULL);
    snprintf(k, sizeof(k), "TFORM%d", c + 1);
    snprintf(fm, sizeof(fm), "1%c", tform[c]);
    kv_str(&h, k, fm, NULL);
    snprintf(k, sizeof(k), "TUNIT%d", c + 1);
    kv_str(&h, k, tunit[c], NULL);
  }
  if (extname) kv_str(&h, "EXTNAME", extname, "name of this binary table extension");
  if (extra)
    for (size_t i = 0; i < extra->n; i += 80) {
      char c80[81];
      memcpy(c80, extra->buf + i, 80);
      c80[80] = 0;
      card(&h, c80);
    }
  if (!flush_hdr(f, &h)) return 0;
  const size_t rowb = 4 * (size_t)ncol;
  uint8_t *row = malloc(rowb * 4096);
  size_t written = 0;
  for (long long r0 = 0; r0 < nrow; r0 += 4096) {
    const long long nr = nrow - r0 < 4096 ? nrow - r0 : 4096;
    for (long long r = 0; r < nr; r++)
      for (int c = 0; c < ncol; c++) be32(row + r * rowb + 4 * c, (const uint8_t *)cols[c] + 4 * (r0 + r));
    if (fwrite(row, rowb, nr, f) != (size_t)nr) { free(row); return 0; }
    written += rowb * nr;
  }
  free(row);
  const size_t pad = (2880 - written % 2880) % 2880;
  char z[2880];
  memset(z, 0, sizeof(z));
  return fwrite(z, 1, pad, f) == pad;
}

int clr_write_healpix_map(const char *fname, long long nside, const float *map) {
  FILE *f = fopen(fname, "wb");
  if (!f) return 1;
  hdr ex = {0};
  kv_str(&ex, "PIXTYPE", "HEALPIX", "HEALPIX Pixelisation");
  kv_str(&ex, "ORDERING", "RING", "Pixel ordering scheme, either RING or NESTED");
  kv_int(&ex, "NSIDE", nside, "Resolution parameter for HEALPIX");
  kv_str(&ex, "COORDSYS", "G", "Pixelisation coordinate system");
  card(&ex, "COMMENT G = Galactic, E = ecliptic, C = celestial = equatorial");
  const char *tt[] = {"map 1"}, *tu[] = {"uK"};
  const void *cols[] = {map};
  int ok = write_primary(f) && write_table(f, 1, tt, "E", tu, "BINTABLE", 12 * nside * nside, cols, &ex);
  free(ex.buf);
  fclose(f);
  return ok ? 0 : 1;
}

int clr_write_catalog_fits(const char *fname, int ipop, long long n, const float *rdz, const float *lens) {
  FILE *f = fopen(fname, "wb");
  if (!f) return 1;
  const int nc = lens ? 9 : 4;
  int32_t *ty = malloc((n > 0 ? n : 1) * sizeof(int32_t));
  float *c[9];
  for (int k = 0; k < nc; k++) c[k] = malloc((n > 0 ? n : 1) * sizeof(float));
  for (long long i = 0; i < n; i++) {
    ty[i] = ipop;
    for (int k = 0; k < 4; k++) c[k][i] = rdz[4 * i + k];
    for (int k = 4; k < nc; k++) c[k][i] = lens[5 * i + k - 4];
  }
  hdr ex = {0};
  kv_str(&ex, "CONTENTS", "Source catalog", NULL);
  const char *tt[] = {"TYPE", "RA", "DEC", "Z_COSMO", "DZ_RSD", "E1", "E2", "KAPPA", "DRA", "DDEC"};
  const char *tu[] = {"NA", "DEG", "DEG", "NA", "NA", "NA", "NA", "NA", "DEG", "DEG"};
  const void *cols[] = {ty, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8]};
  int ok = write_primary(f) && write_table(f, nc + 1, tt, "JEEEEEEEEE", tu, NULL, n, cols, &ex);
  free(ex.buf);
  free(ty);
  for (int k = 0; k < nc; k++) free(c[k]);
  fclose(f);
  return ok ? 0 : 1;
}

int clr_write_catalog_ascii(const char *fname, int ipop, long long n, const float *rdz, const float *lens) {
  FILE *f = fopen(fname, "w");
  if (!f) return 1;
  fprintf(f, "#[1]type [2]RA, [3]dec, [4]z0, [5]dz_RSD ");
  fprintf(f, lens ? "#[6]e1, [7]e2 [8]kappa [9]d_ra [10]d_dec\n" : "\n");
  for (long long i = 0; i < n; i++) {
    fprintf(f, "%d %E %E %E %E ", ipop, rdz[4 * i], rdz[4 * i + 1], rdz[4 * i + 2], rdz[4 * i + 3]);
    if (lens) {
      const float *l = lens + 5 * i;
      fprintf(f, "%E %E %E %E %E \n", l[0], l[1], l[2], l[3], l[4]);
    } else {
      fprintf(f, "\n");
    }
  }
  fclose(f);
  return 0;
}

/* FITS image extension of ny rows of nx floats (fits_create_img FLOAT_IMG) */
static int write_image(FILE *f, long long nx, long long ny, const float *data, const char *contents) {
  hdr h = {0};
  kv_str(&h, "XTENSION", "IMAGE", "IMAGE extension");
  kv_int(&h, "BITPIX", -32, "number of bits per data pixel");
  kv_int(&h, "NAXIS", 2, "number of data axes");
  kv_int(&h, "NAXIS1", nx, "length of data axis 1");
  kv_int(&h, "NAXIS2", ny, "length of data axis 2");
  kv_int(&h, "PCOUNT", 0, "required keyword; must = 0");
  kv_int(&h, "GCOUNT", 1, "required keyword; must = 1");
  kv_str(&h, "CONTENTS", contents, NULL);
  if (!flush_hdr(f, &h)) return 0;
  const size_t tot = (size_t)nx * ny;
  uint8_t buf[4 * 4096];
  for (size_t i0 = 0; i0 < tot; i0 += 4096) {
    const size_t m = tot - i0 < 4096 ? tot - i0 : 4096;
    for (size_t i = 0; i < m; i++) be32(buf + 4 * i, data + i0 + i);
    if (fwrite(buf, 4, m, f) != m) return 0;
  }
  const size_t pad = (2880 - (4 * tot) % 2880) % 2880;
  char z[2880];
  memset(z, 0, sizeof(z));
  return fwrite(z, 1, pad, f) == pad;
}

int clr_append_skewers_fits(const char *fname, int gaussian, long long n, int nr, const float *dens_or_gauss,
                            const float *v, const float *bg4, float sigma_g) {
  FILE *f = fopen(fname, "ab");
  if (!f) return 1;
  int ok = write_image(f, nr, n, dens_or_gauss, gaussian ? "gaussian skewers" : "density skewers") &&
           write_image(f, nr, n, v, "velocity skewers");
  hdr ex = {0};
  kv_str(&ex, "CONTENTS", "Background cosmology", NULL);
  {
    char s[100];
    snprintf(s, sizeof(s), "%-8.8s= %20.7G", "SIGMA_G", (double)sigma_g);
    card(&ex, s);
  }
  const char *tt[] = {"R", "Z", "D", "V"}, *tu[] = {"MPC_H", "NA", "NA", "NA"};
  const void *cols[] = {bg4, bg4 + nr, bg4 + 2 * (size_t)nr, bg4 + 3 * (size_t)nr};
  ok = ok && write_table(f, 4, tt, "EEEE", tu, NULL, nr, cols, &ex);
  free(ex.buf);
  fclose(f);
  return ok ? 0 : 1;
}

/* write_lpt (io.c:597-695): a Gadget-2 snapshot of the LPT particles
 * (header with the reference's extra flag word, positions, zero velocities,
 * 64-bit IDs 0..n-1) */
int clr_write_lpt_gadget(const char *fname, long long n, double l_box, double omega_m, double omega_l,
                         double hhub, const float *xyz) {
  FILE *f = fopen(fname, "wb");
  if (!f) return 1;
  unsigned char h[256];
  memset(h, 0, sizeof(h));
  int np1 = (int)n;
  double mass1 = 27.7455 * omega_m * l_box * l_box * l_box / (double)n, one = 1.0;
  unsigned int lo = (unsigned int)n, hi = (unsigned int)((unsigned long long)n >> 32);
  int nfiles = 1, gflag = 1;
  memcpy(h + 4, &np1, 4);            /* np[1] */
  memcpy(h + 24 + 8, &mass1, 8);     /* mass[1] */
  memcpy(h + 72, &one, 8);           /* time (redshift 0 at 80) */
  memcpy(h + 96 + 4, &lo, 4);        /* np_total[1] */
  memcpy(h + 124, &nfiles, 4);       /* num_files */
  memcpy(h + 128, &l_box, 8);        /* boxsize */
  memcpy(h + 136, &omega_m, 8);
  memcpy(h + 144, &omega_l, 8);
  memcpy(h + 152, &hhub, 8);
  memcpy(h + 168 + 4, &hi, 4);       /* np_total_highword[1] */
  memcpy(h + 196, &gflag, 4);        /* flag_gadgetformat */
  int blk = 256, ok = 1;
  ok = ok && fwrite(&blk, 4, 1, f) == 1 && fwrite(h, 1, 256, f) == 256 && fwrite(&blk, 4, 1, f) == 1;
  blk = (int)(n * 12);
  ok = ok && fwrite(&blk, 4, 1, f) == 1 && fwrite(xyz, 12, (size_t)n, f) == (size_t)n && fwrite(&blk, 4, 1, f) == 1;
  const float zero3[3] = {0, 0, 0};
  ok = ok && fwrite(&blk, 4, 1, f) == 1;
  for (long long i = 0; ok && i < n; i++) ok = fwrite(zero3, 12, 1, f) == 1;
  ok = ok && fwrite(&blk, 4, 1, f) == 1;
  blk = (int)(n * 8);
  ok = ok && fwrite(&blk, 4, 1, f) == 1;
  for (long long i = 0; ok && i < n; i++) {
    const unsigned long long id = (unsigned long long)i;
    ok = fwrite(&id, 8, 1, f) == 1;
  }
  ok = ok && fwrite(&blk, 4, 1, f) == 1;
  fclose(f);
  return ok ? 0 : 1;
}
