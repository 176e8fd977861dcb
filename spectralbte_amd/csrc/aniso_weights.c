/* aniso_weights.c -- the reference's aniso_weights.h entry points; the Landau
 * table is generated on the GPU.
 *
 * Replaces src/aniso_weights.c for boltz_: initialize_weights_AnIso
 * (:37-146) with its file name, load-or-generate rule and messages, gHat3L
 * (:371-398) as a scalar host function, and generate_conv_weights_AnIso
 * (:400-428) through spbte_op_generate_weights_landau (k_weights.hip).
 *
 * With glance > 0 the reference's file-scope glance (:31), which selects
 * gHat3_AnIso in generate_conv_weights_AnIso (:417), is never assigned:
 * initialize_weights_AnIso's parameter of the same name only picks the file
 * name (:65-68).  A missing glance file is therefore filled with the Landau
 * table under the glance name, and so it is here.  The glancing-collision
 * table itself is what exec/MPIWeightGenerator.c writes
 * (spbte_weightgen_main, spbte_op_generate_weights_glance); present, it is
 * loaded.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "aniso_weights.h"
#include "spbte.h"
#include "spbte_internal.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static int N;
static double *eta; /* borrowed */
static double L_v, lambda_;
static double gl_x[32], gl_w[32]; /* 64-point Gauss-Legendre: positive nodes ascending, weights */
static int gl_ready;

/* Newton's iteration on P_64 in long double; nodes stored ascending. */
static void gl64(void) {
  if (gl_ready) return;
  const int n = 64;
  for (int i = 0; i < n / 2; i++) {
    long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L)), pp = 0.0L;
    for (int it = 0; it < 100; it++) {
      long double p1 = 1.0L, p2 = 0.0L;
      for (int j = 1; j <= n; j++) {
        const long double p3 = p2;
        p2 = p1;
        p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0L);
      const long double z1 = z;
      z = z1 - p1 / pp;
      if (fabsl(z - z1) < 1e-19L) break;
    }
    gl_x[n / 2 - 1 - i] = (double)z;
    gl_w[n / 2 - 1 - i] = (double)(2.0L / ((1.0L - z * z) * pp * pp));
  }
  gl_ready = 1;
}

/* ghatL2 integrand (:352-357) at theta, for a = {xiperp, |zeta|^2, xi.zeta/|zeta|, |zeta|} and radius r. */
static double landau_theta(double th, const double *a, double r) {
  return sin(th) * j0(r * a[0] * sin(th)) *
         (-r * r * a[1] * sin(th) * sin(th) * cos(r * a[2] * cos(th)) + 4 * r * a[3] * sin(r * a[2] * cos(th)) * cos(th));
}

/* r^(lambda+2) * int_0^pi ghatL2 (:359-369), the rule summed in symmetric pairs. */
static double landau_r(double r, const double *a) {
  const double A = 0.5 * (M_PI - 0.0), B = 0.5 * (M_PI + 0.0);
  double s = 0.0;
  for (int k = 0; k < 32; k++) {
    const double Ax = A * gl_x[k];
    s += gl_w[k] * (landau_theta(B + Ax, a, r) + landau_theta(B - Ax, a, r));
  }
  return pow(r, lambda_ + 2) * (A * s);
}

double gHat3L(double zeta1, double zeta2, double zeta3, double xi1, double xi2, double xi3) {
  gl64();
  const double zetalen2 = zeta3 * zeta3 + zeta2 * zeta2 + zeta1 * zeta1;
  const double xilen2 = xi1 * xi1 + xi2 * xi2 + xi3 * xi3;
  const double xizeta = xi1 * zeta1 + xi2 * zeta2 + xi3 * zeta3;
  const double zetalen = sqrt(zetalen2);
  double a[4];
  a[0] = (((xilen2 - xizeta * xizeta / zetalen2) < 0) || (zetalen2 == 0)) ? 0.0 : sqrt(xilen2 - xizeta * xizeta / zetalen2);
  a[1] = zetalen2;
  a[2] = (zetalen != 0) ? xizeta / zetalen : 0.0;
  a[3] = zetalen;
  const double A = 0.5 * (L_v - 0.0), B = 0.5 * (L_v + 0.0);
  double s = 0.0;
  for (int i = 0; i < 32; i++) {
    const double Ax = A * gl_x[i];
    s += gl_w[i] * (landau_r(B + Ax, a) + landau_r(B - Ax, a));
  }
  return 2.0 * M_PI * (A * s);
}

static void die(const char *what) {
  fprintf(stderr, "spectralbte_amd: %s: %s\n", what, spbte_last_error());
  exit(1);
}

void generate_conv_weights_AnIso(double **conv_weights) {
  spbte_abi_forget_weights(conv_weights);
  double *v = malloc(sizeof(double) * N);
  if (!v) die("generate_conv_weights_AnIso");
  for (int i = 0; i < N; i++) v[i] = -L_v + i * (2 * L_v / (N - 1));
  spbte_op *op = NULL;
  if (spbte_op_create(&op, N, L_v, v, eta) != SPBTE_OK || spbte_op_generate_weights_landau(op, lambda_) != SPBTE_OK ||
      spbte_op_get_weights_rows(op, conv_weights) != SPBTE_OK)
    die("generate_conv_weights_AnIso");
  spbte_op_destroy(op);
  free(v);
}

static void store(const char *name, double **conv_weights) {
  const size_t n3 = (size_t)N * N * N;
  FILE *fp = fopen(name, "w");
  if (!fp) {
    printf("Something is wrong with storing the weights");
    exit(0);
  }
  for (size_t i = 0; i < n3; i++) fwrite(conv_weights[i], sizeof(double), n3, fp);
  if (fflush(fp) != 0) {
    printf("Something is wrong with storing the weights");
    exit(0);
  }
  fclose(fp);
}

void initialize_weights_AnIso(int nodes, double *zeta, double Lv, double lam, int weightFlag, double **conv_weights,
                              double glance) {
  char name[256];
  N = nodes;
  eta = zeta;
  L_v = Lv;
  lambda_ = lam;
  gl64();
  const size_t n3 = (size_t)N * N * N;
  /* The rows are (re)filled below: drop any device copy made from them. */
  spbte_abi_forget_weights(conv_weights);
  for (size_t i = 0; i < n3; i++)
    if (!(conv_weights[i] = malloc(n3 * sizeof(double)))) {
      fprintf(stderr, "spectralbte_amd: out of host memory allocating weight rows\n");
      exit(1);
    }
  if (glance == 0)
    snprintf(name, sizeof name, "Weights/N%d_Aniso_L_v%g_lambda%g_Landau.wts", N, L_v, lambda_);
  else
    snprintf(name, sizeof name, "Weights/N%d_AnIso_L_v%g_lambda%g_glance%g.wts", N, L_v, lambda_, glance);
  if (weightFlag == 0) {
    FILE *fp = fopen(name, "r");
    if (fp) {
      printf("Loading weights from file %s\n", name);
      for (size_t i = 0; i < n3; i++)
        if (fread(conv_weights[i], sizeof(double), n3, fp) != n3) {
          printf("Error reading weight file\n");
          exit(1);
        }
      fclose(fp);
      return;
    }
    printf("Stored weights not found for this configuration, generating ...\n");
  } else {
    printf("Fresh version of weights being computed and stored for this configuration\n");
  }
  generate_conv_weights_AnIso(conv_weights);
  store(name, conv_weights);
}
