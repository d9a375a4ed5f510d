/* main.c -- `CoLoRe [--test-memory] <param_file>` (reference main.c:24-154):
 * the same invocation, param file and output files, on one GPU. */
#include "clr_host.h"
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv) {
  int test_memory = 0;
  const char *fname;
  char err[1024];
  if (argc < 2) {
    fprintf(stderr, "Usage: ./CoLoRe file_name\n");
    exit(0);
  }
  if (!strcmp(argv[1], "--test-memory")) {
    if (argc < 3) { fprintf(stderr, "Usage: ./CoLoRe --test-memory file_name\n"); exit(0); }
    test_memory = 1;
    fname = argv[2];
  } else
    fname = argv[1];
  setbuf(stdout, NULL);
  printf("\n");
  printf("|-------------------------------------------------|\n");
  printf("|                      CoLoRe                     |\n");
  printf("|-------------------------------------------------|\n\n");
  clr_timer(4);
  clr_run *r = clr_run_read(fname, err, sizeof(err));
  if (!r) {
    fprintf(stderr, "MPI task 0, Fatal: %s\n", err);
    exit(1);
  }
  const double dk = 2 * 3.14159265358979323846 / r->l_box;
  printf("Run parameters: \n");
  printf("  %.3lf < z < %.3lf\n", r->z_min, r->z_max);
  printf("  %.3lf < r/(Mpc/h) < %.3lf\n", r->r_min, r->r_max);
  printf("  L_box = %.3lf Mpc/h, N_grid = %d \n", r->l_box, r->n_grid);
  printf("  Scales resolved: %.3lE < k < %.3lE h/Mpc\n", dk, 0.5 * (r->n_grid - 1) * dk);
  printf("  Fourier-space resolution: dk = %.3lE h/Mpc\n", dk);
  printf("  Real-space resolution: dx = %.3lE Mpc/h\n", r->l_box / r->n_grid);
  if (r->r_smooth > 0) printf("  Density field pre-smoothed on scales: x_s = %.3lE Mpc/h\n", r->r_smooth);
  else printf("  No extra smoothing\n");
  if (r->n_tr[0] > 0) printf("  %d galaxy populations\n", r->n_tr[0]);
  if (r->n_tr[1] > 0) printf("  %d intensity mapping species\n", r->n_tr[1]);
  if (r->n_tr[2] > 0) printf("  %d custom maps\n", r->n_tr[2]);
  if (r->do_kappa) printf("  %d lensing source planes\n", r->n_kappa);
  if (r->do_isw) printf("  %d ISW source planes\n", r->n_isw);
  printf("\n");
  int rc = 0;
  if (test_memory) {
    const double grids = 2.0 * r->n_grid * (double)r->n_grid * 2 * (r->n_grid / 2 + 1) * 4;
    printf("Device will allocate %.3lf GB for the two grids\n", grids / (1024.0 * 1024.0 * 1024.0));
  } else {
    rc = clr_run_execute(r, 0);
  }
  printf("\n");
  printf("|-------------------------------------------------|\n\n");
  clr_run_free(r);
  clr_timer(5);
  return rc;
}
