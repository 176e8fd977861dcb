/* weights.c -- the reference's weights.h interface; tables are generated on the GPU.
 *
 * Replaces src/weights.c.  The host layout (N^3 malloc'd rows of N^3 doubles),
 * the file names and the raw FP64 file format are the reference's
 * (src/weights.c:26-31, :61-75, :78-124).  Generation runs on the GPU
 * (k_weights.hip): closed forms of the radial integral for lambda = 0 and 1,
 * Gauss-Legendre quadrature otherwise, instead of one GSL QAG call per entry.
 * ghat and gHat3 remain scalar host functions, as in the reference API.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "weights.h"
#include "spbte.h"
#include "spbte_internal.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static int N;
static double L_v, lambda_, deta, prefactor, diam_i, diam_j, mass_i, mass_j;
static double *zeta; /* borrowed */

static void *xmalloc(size_t bytes) {
  void *p = malloc(bytes);
  if (!p) {
    fprintf(stderr, "spectralbte_amd: out of host memory allocating %zu bytes of weights\n", bytes);
    exit(1);
  }
  return p;
}

void alloc_weights(int n, double ****conv_weights, int total_species) {
  *conv_weights = xmalloc(total_species * sizeof(double **));
  for (int i = 0; i < total_species; i++) (*conv_weights)[i] = xmalloc((size_t)n * n * n * sizeof(double *));
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

void initialize_weights(int nodes, double *eta, double Lv, double lam, int weightFlag, int isoFlag, double **conv_weights,
                        species species_i, species species_j) {
  char name[256];
  N = nodes;
  zeta = eta;
  deta = eta[1] - eta[0];
  L_v = Lv;
  lambda_ = lam;
  prefactor = 16.0 * M_PI * M_PI * deta * deta * deta / pow(2.0 * M_PI, 1.5) / (4.0 * M_PI);
  diam_i = species_i.d_ref;
  diam_j = species_j.d_ref;
  mass_i = species_i.mass;
  mass_j = species_j.mass;
  printf("%g %g %g %g %s %s \n", diam_i, diam_j, mass_i, mass_j, species_i.name, species_j.name);

  const size_t n3 = (size_t)N * N * N;
  /* The rows are (re)filled below: a device copy made from this pointer
   * earlier by ComputeQ no longer matches them. */
  spbte_abi_forget_weights(conv_weights);
  for (size_t i = 0; i < n3; i++) conv_weights[i] = xmalloc(n3 * sizeof(double));

  if (!isoFlag) {
    if (strcmp(species_i.name, "default") == 0)
      snprintf(name, sizeof name, "Weights/N%d_isotropic_L_v%g_lambda%g.wts", N, L_v, lambda_);
    else
      snprintf(name, sizeof name, "Weights/N%d_isotropic_L_v%g_HS_%s_%zd_%s_%zd.wts", N, L_v, species_i.name, species_i.id,
               species_j.name, species_j.id);
  } else {
    snprintf(name, sizeof name, "Weights/N%d_AnIso_L_v%g_lambda%g_Landau.wts", N, L_v, lambda_);
  }

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
    if (isoFlag) {
      printf("Please use the MPI Weight generator to build the weights for this anisotropic function\n");
      exit(1);
    }
    generate_conv_weights_iso(conv_weights);
    store(name, conv_weights);
  } else {
    printf("Fresh version of weights being computed and stored for this configuration\n");
    generate_conv_weights_iso(conv_weights);
    store(name, conv_weights);
  }
}

void generate_conv_weights_iso(double **conv_weights) {
  spbte_abi_forget_weights(conv_weights);
  double *v = xmalloc(sizeof(double) * N);
  for (int i = 0; i < N; i++) v[i] = -L_v + i * (2 * L_v / (N - 1));
  spbte_op *op = NULL;
  if (spbte_op_create(&op, N, L_v, v, zeta) != SPBTE_OK ||
      spbte_op_generate_weights_iso(op, lambda_, diam_i, diam_j, mass_i, mass_j) != SPBTE_OK ||
      spbte_op_get_weights_rows(op, conv_weights) != SPBTE_OK) {
    fprintf(stderr, "spectralbte_amd: generate_conv_weights_iso: %s\n", spbte_last_error());
    exit(1);
  }
  spbte_op_destroy(op);
  free(v);
}

void dealloc_weights(int n, double **conv_weights) {
  spbte_abi_forget_weights(conv_weights);
  const size_t n3 = (size_t)n * n * n;
  for (size_t i = 0; i < n3; i++) free(conv_weights[i]);
  free(conv_weights);
}

static double sinc(double x) { return (x != 0.0) ? sin(x) / x : 1.0; }
static double fcos(double x) { return (fabs(x) > 1e-12) ? (cos(x) - 1.0) / (x * x) : -0.5 + x * x / 24.0; }

/* src/weights.c:156-160; args = {mu|zeta|, |xi|, |xi - mu zeta|}. */
double ghat(double r, void *args) {
  const double *a = (const double *)args;
  return pow(r, lambda_ + 2) * (sinc(r * a[0]) * sinc(r * a[2]) - sinc(r * a[1]));
}

/* int_0^L ghat dr: closed forms (gain - loss) for lambda 0 and 1, composite
 * 64-panel 20-point Gauss-Legendre otherwise. */
static double radial_integral(const double *a) {
  const double L = L_v, tiny = 1e-12, a0 = a[0], a1 = a[1], a2 = a[2];
  if (lambda_ == 0.0 || lambda_ == 1.0) {
    double gain, loss;
    if (lambda_ == 0.0) {
      if (a0 > tiny && a2 > tiny) gain = L / (2 * a0 * a2) * (sinc((a0 - a2) * L) - sinc((a0 + a2) * L));
      else if (a0 > tiny || a2 > tiny) {
        double c = a0 > tiny ? a0 : a2;
        gain = (sin(c * L) - c * L * cos(c * L)) / (c * c * c);
      } else gain = L * L * L / 3.0;
      loss = (a1 > tiny) ? (sin(a1 * L) - a1 * L * cos(a1 * L)) / (a1 * a1 * a1) : L * L * L / 3.0;
    } else {
      if (a0 > tiny && a2 > tiny) {
        double am = (a0 - a2) * L, ap = (a0 + a2) * L;
        if (fabs(a0 - a2) < 1e-10) am = 0.0;
        gain = (L * L * 0.5 / (a0 * a2)) * (sinc(am) - sinc(ap) + fcos(am) - fcos(ap));
      } else if (a0 > tiny || a2 > tiny) {
        double c = a0 > tiny ? a0 : a2, x = c * L;
        gain = (2 * x * sin(x) - x * x * cos(x) + 2 * cos(x) - 2) / (c * c * c * c);
      } else gain = 0.25 * L * L * L * L;
      if (a1 > tiny) {
        double x = a1 * L;
        loss = (cos(x) * (2.0 - x * x) + 2.0 * x * sin(x) - 2.0) / (a1 * a1 * a1 * a1);
      } else loss = 0.25 * L * L * L * L;
    }
    return gain - loss;
  }
  static double gx[20], gw[20];
  static int init = 0;
  if (!init) {
    for (int i = 0; i < 10; i++) {
      double z = cos(M_PI * (i + 0.75) / 20.5), pp = 0.0;
      for (int it = 0; it < 100; it++) {
        double p1 = 1.0, p2 = 0.0;
        for (int j = 1; j <= 20; j++) {
          double p3 = p2;
          p2 = p1;
          p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
        }
        pp = 20 * (z * p1 - p2) / (z * z - 1.0);
        double z1 = z;
        z = z1 - p1 / pp;
        if (fabs(z - z1) < 1e-16) break;
      }
      gx[i] = -z;
      gx[19 - i] = z;
      gw[i] = gw[19 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
    init = 1;
  }
  const int panels = 64;
  double h = L / panels, sum = 0.0;
  for (int p = 0; p < panels; p++) {
    double lo = p * h, ps = 0.0;
    for (int q = 0; q < 20; q++) ps += gw[q] * ghat(lo + 0.5 * h * (gx[q] + 1.0), (void *)a);
    sum += 0.5 * h * ps;
  }
  return sum;
}

/* src/weights.c:181-262 */
double gHat3(double ki1, double ki2, double ki3, double zeta1, double zeta2, double zeta3) {
  double mu = mass_j / (mass_i + mass_j), args[3];
  args[0] = mu * sqrt(zeta1 * zeta1 + zeta2 * zeta2 + zeta3 * zeta3);
  args[1] = sqrt(ki1 * ki1 + ki2 * ki2 + ki3 * ki3);
  args[2] = sqrt((ki1 - mu * zeta1) * (ki1 - mu * zeta1) + (ki2 - mu * zeta2) * (ki2 - mu * zeta2) +
                 (ki3 - mu * zeta3) * (ki3 - mu * zeta3));
  return prefactor * radial_integral(args);
}
