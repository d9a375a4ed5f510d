/* pipeline.c -- the stage sequence of main.c:61-141 on the device, with the
 * reference's progress report and output files. */
#include "clr_host.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CHK(call)                                                                   \
  do {                                                                              \
    int rc_ = (call);                                                               \
    if (rc_) {                                                                      \
      fprintf(stderr, "MPI task 0, Fatal: %s\n", clr_last_error(ctx));              \
      clr_destroy(ctx);                                                             \
      return rc_;                                                                   \
    }                                                                               \
  } while (0)

static void stage_report(const clr_ctx *ctx, int id) {
  double ms[16];
  clr_stage_times(ctx, ms, 16);
  printf(">    Relative time ellapsed %.1lf ms\n", ms[id]);
}

static int dump_grid(clr_ctx *ctx, const clr_run *r, const char *tag) {
  const size_t n = (size_t)r->n_grid * r->n_grid * 2 * (r->n_grid / 2 + 1);
  float *g = malloc(n * sizeof(float));
  char fname[CLR_PATH + 64];
  if (!g) return 1;
  clr_timer(0);
  printf("*** Writing density field (native format)\n");
  int rc = clr_fetch_grid(ctx, CLR_GRID_DENS, g);
  snprintf(fname, sizeof(fname), "%s_dens_%s_0.dat", r->prefix_out, tag);
  if (!rc) rc = clr_write_density_grid(fname, r->n_grid, r->l_box, g);
  free(g);
  clr_timer(2);
  printf("\n");
  return rc;
}

static void write_maps(const char *pattern, const clr_run *r, int nside, int nplanes, const float *nest_maps) {
  const long long npix = 12LL * nside * nside;
  float *ring = malloc(npix * sizeof(float));
  char fname[CLR_PATH + 64];
  for (int ir = 0; ir < nplanes; ir++) {
    const float *m = nest_maps + (size_t)ir * npix;
    for (long long ip = 0; ip < npix; ip++) ring[clr_hp_nest2ring(nside, ip)] = m[ip];
    snprintf(fname, sizeof(fname), pattern, r->prefix_out, ir);
    if (clr_write_healpix_map(fname, nside, ring)) fprintf(stderr, "MPI task 0, Warning: could not write %s\n", fname);
  }
  free(ring);
}

int clr_run_execute(clr_run *r, int device) {
  clr_ctx *ctx = NULL;
  char err[1024];
  printf("Seed : %u\n", (unsigned)r->seed);
  if (r->do_pred && clr_write_predictions(r)) fprintf(stderr, "MPI task 0, Warning: could not write the predictions\n");
  if (r->just_do_pred) return 0;
  if (clr_run_device(r, device, &ctx, err, sizeof(err))) {
    fprintf(stderr, "MPI task 0, Fatal: %s\n", err);
    clr_destroy(ctx);
    return 1;
  }

  /* create_cartesian_fields */
  double sigma2 = 0;
  printf("*** Creating Gaussian density field \n");
  printf("Creating Fourier-space density and Newtonian potential \n");
  CHK(clr_gaussian_fields(ctx, (uint32_t)r->seed, r->r_smooth, r->smooth_potential, CLR_RNG_PHILOX, NULL, &sigma2));
  stage_report(ctx, 0);
  printf("Transforming density and Newtonian potential\n");
  stage_report(ctx, 1);
  printf("Normalizing density and Newtonian potential \n");
  stage_report(ctx, 2);
  printf(" <d^2>=%.3lE\n\n", sqrt(sigma2));
  if (r->output_density && dump_grid(ctx, r, "gaussian")) fprintf(stderr, "MPI task 0, Warning: grid dump failed\n");

  /* compute_physical_density_field */
  printf("*** Creating physical matter density\n");
  printf("Lognormalizin'\n");
  if (r->dens_type == 1 || r->dens_type == 2) {
    /* lpt_1 / lpt_2 (density.c:376-1031) */
    printf(" %dLPT\n", r->dens_type);
    const long long np = (long long)r->n_grid * r->n_grid * r->n_grid;
    float *xyz = r->output_lpt ? malloc((size_t)np * 3 * sizeof(float)) : NULL;
    if (r->output_lpt && !xyz) { fprintf(stderr, "MPI task 0, Fatal: Ran out of memory\n"); clr_destroy(ctx); return 1; }
    const int lrc = clr_lpt_density(ctx, r->dens_type, r->lpt_interp_type, xyz);
    if (lrc) free(xyz);
    CHK(lrc);
    if (xyz) {
      char lname[CLR_PATH + 32];
      printf(" - Writing LPT positions\n");
      snprintf(lname, sizeof(lname), "%s_lpt_out.0", r->prefix_out);
      if (clr_write_lpt_gadget(lname, np, r->l_box, r->OmegaM, r->OmegaL, r->hhub, xyz))
        fprintf(stderr, "MPI task 0, Warning: could not write %s\n", lname);
      free(xyz);
    }
  } else {
    CHK(clr_density(ctx, r->dens_type, sigma2));
  }
  stage_report(ctx, 3);
  printf("\n");
  if (r->output_density && dump_grid(ctx, r, "lightcone")) fprintf(stderr, "MPI task 0, Warning: grid dump failed\n");

  /* compute_density_normalization */
  printf("*** Computing normalization of density field\n");
  int nz = 0;
  const int ntr = r->n_tr[0] + r->n_tr[1] + r->n_tr[2];
  unsigned long long cnt[256];
  double zsum[256];
  double *sums = malloc((size_t)(ntr > 0 ? ntr : 1) * 256 * sizeof(double));
  CHK(clr_shell_norms(ctx, 0, sigma2, 256, &nz, cnt, zsum, sums));
  {
    double *tab = malloc((size_t)(ntr > 0 ? ntr : 1) * CLR_NA * sizeof(double));
    double e0[3 * CLR_NPOP_MAX], ef[3 * CLR_NPOP_MAX];
    int t = 0;
    for (int k = 0; k < 3; k++)
      for (int i = 0; i < r->n_tr[k]; i++, t++) CHK(clr_get_norm_table(ctx, k, i, tab + (size_t)t * CLR_NA, &e0[t], &ef[t]));
    for (int iz = 0; iz < nz; iz++) {
      double z = cnt[iz] > 0 ? zsum[iz] / cnt[iz] : 0;
      if (iz == 0) z = 0;
      if (iz == nz - 1) z = clr_run_get_bg(r, 0.5 * r->l_box, CLR_BG_Z, 0);
      const double rz = clr_run_r_of_z(r, z);
      printf("z=%.3lE, ", z);
      for (t = 0; t < ntr; t++) {
        const double *f = tab + (size_t)t * CLR_NA;
        double v;
        if (rz <= 0) v = e0[t];
        else if (rz >= r->r_arr_r2z[CLR_NA - 1]) v = ef[t];
        else {
          const int ir = (int)(rz * r->glob_idr);
          v = f[ir] + (f[ir + 1] - f[ir]) * (rz - r->r_arr_r2z[ir]) * r->glob_idr;
        }
        printf("<d^2_%d>=%.3lE, ", t, v);
      }
      printf("%011llu\n", cnt[iz]);
    }
    free(tab);
  }
  free(sums);
  stage_report(ctx, 4);
  printf("\n");

  /* srcs_set_cartesian + srcs_get_local_properties */
  long long nsrc[CLR_NPOP_MAX] = {0};
  float *props[CLR_NPOP_MAX] = {0}, *lens[CLR_NPOP_MAX] = {0};
  float *skw_d[CLR_NPOP_MAX] = {0}, *skw_v[CLR_NPOP_MAX] = {0};
  /* need_beaming (io.c:460-461; fast lensing is not part of this build) */
  int need_beaming = r->do_kappa || r->do_isw || r->n_tr[2] > 0;
  for (int ip = 0; ip < r->n_tr[0]; ip++) need_beaming |= r->lensing_srcs[ip] || r->skw_srcs[ip];
  if (r->n_tr[0] > 0) {
    printf("*** Getting point sources\n");
    for (int ip = 0; ip < r->n_tr[0]; ip++) {
      printf(" %d-th galaxy population\n", ip);
      printf("   Poisson-sampling\n");
      CHK(clr_sample_sources(ctx, ip, (uint32_t)r->seed, r->nside_base, &nsrc[ip]));
      printf("   There will be %ld objects in total \n", (long)nsrc[ip]);
      printf("   Assigning coordinates\n");
      stage_report(ctx, 5);
      props[ip] = malloc((nsrc[ip] > 0 ? nsrc[ip] : 1) * 4 * sizeof(float));
      CHK(clr_source_properties(ctx, ip, props[ip], nsrc[ip]));
      if (need_beaming) {
        /* srcs_beams_preproc/get_beam_properties/postproc (srcs.c:425-737):
         * the CIC velocity replaces dz_rsd; lensing when asked for */
        printf("   Line-of-sight properties\n");
        float *dz = malloc((nsrc[ip] > 0 ? nsrc[ip] : 1) * sizeof(float));
        if (r->lensing_srcs[ip]) lens[ip] = malloc((nsrc[ip] > 0 ? nsrc[ip] : 1) * 5 * sizeof(float));
        CHK(clr_source_beams(ctx, ip, clr_run_a2r(r), dz, lens[ip], nsrc[ip]));
        for (long long i = 0; i < nsrc[ip]; i++) props[ip][4 * i + 3] = dz[i];
        free(dz);
      }
      if (r->skw_srcs[ip]) {
        if (r->skw_gauss[ip] && r->dens_type != CLR_DENS_LOGNORMAL) {
          fprintf(stderr, "MPI task 0, Fatal: Cannot write Gaussian skewers with density type %d\n", r->dens_type);
          clr_destroy(ctx);
          return 1;
        }
        const size_t ns = (size_t)(nsrc[ip] > 0 ? nsrc[ip] : 1) * (r->n_grid / 2);
        skw_d[ip] = malloc(ns * sizeof(float));
        skw_v[ip] = malloc(ns * sizeof(float));
        CHK(clr_source_skewers(ctx, ip, r->skw_gauss[ip], sigma2, skw_d[ip], skw_v[ip], nsrc[ip]));
      }
    }
    printf("\n");
  }

  /* imap_set_cartesian */
  float *imd[CLR_NPOP_MAX] = {0};
  int *imn[CLR_NPOP_MAX] = {0};
  if (r->n_tr[1] > 0) {
    printf("*** Filling up intensity maps\n");
    for (int ip = 0; ip < r->n_tr[1]; ip++) {
      const size_t n = (size_t)12 * r->nside_imap[ip] * r->nside_imap[ip] * r->imap_nr[ip];
      printf(" %d-th intensity mapping species\n", ip);
      imd[ip] = malloc(n * sizeof(float));
      imn[ip] = malloc(n * sizeof(int));
      CHK(clr_imap(ctx, ip, r->nside_imap[ip], r->imap_nr[ip], r->imap_r0[ip], r->imap_rf[ip], imd[ip], imn[ip]));
      stage_report(ctx, 8);
    }
    printf("\n");
  }

  /* get_beam_properties: kappa, isw, custom maps */
  float *kmap = NULL, *imap_isw = NULL, *cmap[CLR_NPOP_MAX] = {0};
  if (r->do_kappa || r->do_isw || r->n_tr[2] > 0) {
    printf("*** Getting LOS information\n");
    clr_timer(0);
    for (int ip = 0; ip < r->n_tr[2]; ip++) {
      cmap[ip] = malloc((size_t)12 * r->nside_cstm[ip] * r->nside_cstm[ip] * sizeof(float));
      CHK(clr_custom_map(ctx, ip, r->nside_cstm[ip], cmap[ip]));
    }
    if (r->do_kappa) {
      kmap = malloc((size_t)12 * r->nside_kappa * r->nside_kappa * r->n_kappa * sizeof(float));
      CHK(clr_kappa(ctx, r->nside_kappa, r->n_kappa, r->r_kappa, kmap));
    }
    if (r->do_isw) {
      imap_isw = malloc((size_t)12 * r->nside_isw * r->nside_isw * r->n_isw * sizeof(float));
      CHK(clr_isw(ctx, r->nside_isw, r->n_isw, r->r_isw, imap_isw));
    }
    clr_timer(2);
    printf("\n");
  }

  /* writers */
  char fname[CLR_PATH + 64];
  if (r->do_kappa) {
    clr_timer(0);
    printf("*** Writing kappa source maps\n");
    char pat[CLR_PATH + 64];
    snprintf(pat, sizeof(pat), "%%s_kappa_z%%03d.fits");
    write_maps(pat, r, r->nside_kappa, r->n_kappa, kmap);
    clr_timer(2);
    printf("\n");
  }
  if (r->do_isw) {
    clr_timer(0);
    printf("*** Writing isw source maps\n");
    write_maps("%s_isw_z%03d.fits", r, r->nside_isw, r->n_isw, imap_isw);
    clr_timer(2);
    printf("\n");
  }
  if (r->n_tr[0] > 0) {
    printf("*** Writing source catalogs\n");
    for (int ip = 0; ip < r->n_tr[0]; ip++) {
      clr_timer(0);
      if (r->output_format == 1) {
        printf(" %d-th population (FITS)\n", ip);
        snprintf(fname, sizeof(fname), "%s_srcs_s%d_0.fits", r->prefix_out, ip + 1);
        if (clr_write_catalog_fits(fname, ip, nsrc[ip], props[ip], lens[ip])) fprintf(stderr, "MPI task 0, Warning: could not write %s\n", fname);
        if (skw_d[ip]) {
          /* slicing information, io.c:1164-1180 */
          const int nr = r->n_grid / 2;
          const double dr = r->r_max / nr;
          float *bg4 = malloc(4 * (size_t)nr * sizeof(float));
          for (int ir = 0; ir < nr; ir++) {
            const double rr = dr * (ir + 0.5);
            bg4[ir] = (float)rr;
            bg4[nr + ir] = (float)clr_run_get_bg(r, rr, CLR_BG_Z, 0);
            bg4[2 * nr + ir] = (float)clr_run_get_bg(r, rr, CLR_BG_D1, 0);
            bg4[3 * nr + ir] = (float)clr_run_get_bg(r, rr, CLR_BG_V1, 0);
          }
          if (clr_append_skewers_fits(fname, r->skw_gauss[ip], nsrc[ip], nr, skw_d[ip], skw_v[ip], bg4, (float)sqrt(sigma2)))
            fprintf(stderr, "MPI task 0, Warning: could not write skewers to %s\n", fname);
          free(bg4);
        }
      } else {
        printf(" %d-th population (ASCII)\n", ip);
        snprintf(fname, sizeof(fname), "%s_srcs_s%d_0.txt", r->prefix_out, ip + 1);
        if (clr_write_catalog_ascii(fname, ip, nsrc[ip], props[ip], lens[ip])) fprintf(stderr, "MPI task 0, Warning: could not write %s\n", fname);
      }
      clr_timer(2);
    }
    printf("\n");
  }
  if (r->n_tr[1] > 0) {
    clr_timer(0);
    printf("*** Writing intensity maps\n");
    for (int ip = 0; ip < r->n_tr[1]; ip++) {
      const long long npix = 12LL * r->nside_imap[ip] * r->nside_imap[ip];
      float *m = malloc(npix * sizeof(float));
      printf(" %d-th species\n", ip);
      for (int ir = 0; ir < r->imap_nr[ip]; ir++) {
        for (long long p = 0; p < npix; p++) {
          const size_t o = (size_t)ir * npix + p;
          m[p] = imn[ip][o] > 0 ? imd[ip][o] / imn[ip][o] : imd[ip][o];
        }
        snprintf(fname, sizeof(fname), "%s_imap_s%d_nu%03d.fits", r->prefix_out, ip + 1, ir);
        if (clr_write_healpix_map(fname, r->nside_imap[ip], m)) fprintf(stderr, "MPI task 0, Warning: could not write %s\n", fname);
      }
      free(m);
    }
    clr_timer(2);
    printf("\n");
  }
  if (r->n_tr[2] > 0) {
    /* write_cstm (io.c:754-806): nadd = 1, NEST converted to RING */
    clr_timer(0);
    printf("*** Writing custom maps\n");
    for (int ip = 0; ip < r->n_tr[2]; ip++) {
      char pat[CLR_PATH + 64];
      printf(" %d-th species\n", ip);
      snprintf(pat, sizeof(pat), "%%s_custom_s%d.fits", ip + 1);
      write_maps(pat, r, r->nside_cstm[ip], 1, cmap[ip]);
    }
    clr_timer(2);
    printf("\n");
  }
  for (int i = 0; i < CLR_NPOP_MAX; i++) {
    free(props[i]); free(lens[i]); free(skw_d[i]); free(skw_v[i]); free(imd[i]); free(imn[i]); free(cmap[i]);
  }
  free(kmap);
  free(imap_isw);
  clr_destroy(ctx);
  return 0;
}
