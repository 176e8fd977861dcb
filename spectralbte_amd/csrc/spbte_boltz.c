/* spbte_boltz.c -- the boltz_ front end over the batched runs (C99, host only).
 *
 * Reads the reference's input keywords (src/input.c:10-183), mesh files
 * (src/mesh_setup.c:9-200) and species files (src/species.c:28-160), builds
 * the velocity and Fourier grids and the initial state (src/initializer.c),
 * loads, or generates on the GPU and stores, the weight table
 * (src/weights.c:65-124; the Landau table, src/aniso_weights.c:37-146), and
 * runs the time loop of exec/boltz.c:181-398 for one or several species as
 * whole steps on the GPU, with the moment output of src/output.c and the
 * restart dumps of src/restart.c.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "spbte.h"
#include "spbte_boltz.h"
#include "spbte_internal.h"

#define KB_JOULES_PER_KELVIN 1.380658e-23 /* src/constants.h:6 */

#define CHK(expr)               \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != SPBTE_OK) return rc_; \
  } while (0)

/* ------------------------------------------------------------ input file */
static int token(FILE *fp, char *tok) { return fscanf(fp, "%79s", tok) == 1; }

static int read_count(FILE *fp, int *out) {
  unsigned long u;
  if (fscanf(fp, "%lu", &u) != 1 || u > 0x7fffffffUL) return 0;
  *out = (int)u;
  return 1;
}

static int read_real(FILE *fp, double *out) { return fscanf(fp, "%lf", out) == 1; }

int spbte_read_input(const char *path, spbte_input *in) {
  if (!path || !in) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  /* set_default_values, src/input.c:132-183 */
  memset(in, 0, sizeof *in);
  in->N = 16;
  in->L_v = 5.0;
  in->Kn = 1.0;
  in->lambda = 1.0;
  in->dt = -1.0;
  in->nT = 1000;
  in->order = 1;
  in->data_freq = 10;
  in->restart_time = 85500.0;
  in->init_flag = -1;
  in->num_species = 1;
  strcpy(in->mesh_file, "not set");
  strcpy(in->species_names[0], "default");

  FILE *fp = fopen(path, "r");
  if (!fp) return spbte_fail(SPBTE_ERR_IO, "Error - input file not found: %s", path);
  char line[SPBTE_NAME_LEN] = "dummy";
  int ok = 1;
  while (ok && strcmp(line, "Stop") != 0) {
    if (!token(fp, line)) { ok = 0; break; }
    if (!strcmp(line, "N")) ok = read_count(fp, &in->N);
    else if (!strcmp(line, "L_v")) ok = read_real(fp, &in->L_v);
    else if (!strcmp(line, "Knudsen")) ok = read_real(fp, &in->Kn);
    else if (!strcmp(line, "Lambda")) ok = read_real(fp, &in->lambda);
    else if (!strcmp(line, "Time_step")) ok = read_real(fp, &in->dt);
    else if (!strcmp(line, "Number_of_time_steps")) ok = read_count(fp, &in->nT);
    else if (!strcmp(line, "Space_order")) ok = read_count(fp, &in->order);
    else if (!strcmp(line, "Data_writing_frequency")) ok = read_count(fp, &in->data_freq);
    else if (!strcmp(line, "Restart")) ok = read_count(fp, &in->restart);
    else if (!strcmp(line, "Restart_time")) ok = read_real(fp, &in->restart_time);
    else if (!strcmp(line, "Init_field")) ok = read_count(fp, &in->init_flag);
    else if (!strcmp(line, "Bound_cond")) ok = read_count(fp, &in->bc_flag);
    else if (!strcmp(line, "SpaceInhom")) ok = read_count(fp, &in->inhom);
    else if (!strcmp(line, "Recompute_weights")) ok = read_count(fp, &in->weight_flag);
    else if (!strcmp(line, "Anisotropic")) ok = read_count(fp, &in->aniso);
    else if (!strcmp(line, "mesh_file")) ok = token(fp, in->mesh_file);
    else if (!strcmp(line, "num_species")) {
      ok = read_count(fp, &in->num_species);
      if (ok && (in->num_species < 1 || in->num_species > SPBTE_MAX_SPECIES)) {
        fclose(fp);
        return spbte_fail(SPBTE_ERR_UNSUPPORTED, "num_species %d in %s: 1 to %d species can be named", in->num_species,
                          path, SPBTE_MAX_SPECIES);
      }
      for (int i = 0; ok && i < in->num_species; i++) ok = token(fp, in->species_names[i]);
    }
  }
  fclose(fp);
  if (!ok) return spbte_fail(SPBTE_ERR_IO, "Error while reading the input file %s (a value is missing or there is no Stop)", path);
  if (!strcmp(in->mesh_file, "not set") && in->inhom == 1)
    return spbte_fail(SPBTE_ERR_ARG, "Error: please specify the mesh (mesh_file) in %s", path);
  return SPBTE_OK;
}

/* --------------------------------------------------------------- species */
void spbte_species_defaults(species *sp, int id) {
  memset(sp, 0, sizeof *sp);
  sp->id = (size_t)id;
  sp->Rgas = 1.0;
  sp->mass = 1.0;
  sp->mm = 1.0;
  sp->d_ref = 2.0;
  sp->T_ref = 1.0;
  sp->mu_ref = 1.0;
  sp->E0 = 0.0;
  strcpy(sp->name, "default");
}

int spbte_read_species(const char *path, species *sp) {
  if (!path || !sp) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  FILE *fp = fopen(path, "r");
  if (!fp) return spbte_fail(SPBTE_ERR_IO, "Error - species file not found: %s", path);
  char line[SPBTE_NAME_LEN] = "dummy";
  int ok = 1, levels = 0;
  while (ok && strcmp(line, "Stop") != 0) {
    if (!token(fp, line)) { ok = 0; break; }
    if (!strcmp(line, "name")) ok = token(fp, sp->name);
    else if (!strcmp(line, "num_en_levels")) { ok = read_count(fp, &levels); sp->num_levels = (size_t)levels; }
    else if (!strcmp(line, "mass")) ok = read_real(fp, &sp->mass);
    else if (!strcmp(line, "En_for")) ok = read_real(fp, &sp->E0);
    else if (!strcmp(line, "d_ref")) ok = read_real(fp, &sp->d_ref);
    else if (!strcmp(line, "T_ref")) ok = read_real(fp, &sp->T_ref);
    else if (!strcmp(line, "mu_ref")) ok = read_real(fp, &sp->mu_ref);
    else if (!strcmp(line, "omega")) ok = read_real(fp, &sp->omega);
  }
  fclose(fp);
  if (!ok) return spbte_fail(SPBTE_ERR_IO, "Error while reading the species file %s", path);
  if (!(sp->mass > 0)) return spbte_fail(SPBTE_ERR_ARG, "species file %s: mass must be positive", path);
  return SPBTE_OK;
}

/* ------------------------------------------------------------------ mesh */
int spbte_read_mesh(const char *path, int order, int *nX, double *x, double *dx, double *dx_min) {
  if (!path || !nX) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  if (order != 1 && order != 2) return spbte_fail(SPBTE_ERR_ARG, "order must be 1 or 2");
  FILE *fp = fopen(path, "r");
  if (!fp) return spbte_fail(SPBTE_ERR_IO, "cannot open mesh file %s", path);
  char expl[256];
  int n = 0, nz = 0, rc = SPBTE_OK;
  /* line 1: one explanation token; then points, zones, (points, length) per zone */
  if (fscanf(fp, "%255s", expl) != 1 || fscanf(fp, "%d", &n) != 1 || fscanf(fp, "%d", &nz) != 1)
    rc = spbte_fail(SPBTE_ERR_IO, "mesh file %s: cannot read the point and zone counts", path);
  else if (nz < 1 || n < 1)
    rc = spbte_fail(SPBTE_ERR_IO, "mesh file %s: bad number of zones (%d) or points (%d)", path, nz, n);
  int *zn = NULL;
  double *zdx = NULL;
  if (!rc) {
    zn = malloc(sizeof(int) * nz);
    zdx = malloc(sizeof(double) * nz);
    if (!zn || !zdx) rc = spbte_fail(SPBTE_ERR_ARG, "out of host memory");
  }
  long total = 0;
  double dmin = INFINITY;
  for (int i = 0; i < nz && !rc; i++) {
    double len;
    if (fscanf(fp, "%d", &zn[i]) != 1 || fscanf(fp, "%lf", &len) != 1 || zn[i] < 1 || !(len > 0))
      rc = spbte_fail(SPBTE_ERR_IO, "mesh file %s: bad zone %d", path, i);
    else {
      zdx[i] = len / (double)zn[i]; /* src/mesh_setup.c:69 */
      if (zdx[i] < dmin) dmin = zdx[i];
      total += zn[i];
    }
  }
  fclose(fp);
  if (!rc && total != n)
    rc = spbte_fail(SPBTE_ERR_IO, "mesh file %s: the zones hold %ld points, the header says %d", path, total, n);
  if (!rc) {
    *nX = n;
    if (dx_min) *dx_min = dmin;
  }
  if (!rc && x && dx) {
    /* cell centres from the accumulated zone edges (src/mesh_setup.c:121-133) */
    double edge = 0.0;
    int c = order;
    for (int i = 0; i < nz; i++)
      for (int j = 0; j < zn[i]; j++, c++) {
        dx[c] = zdx[i];
        x[c] = edge + 0.5 * zdx[i];
        edge += zdx[i];
      }
    /* ghosts extend the end cells (left :134-143, right as the last rank's :165-175) */
    for (int g = order - 1; g >= 0; g--) {
      dx[g] = dx[g + 1];
      x[g] = x[g + 1] - dx[g + 1];
    }
    for (int g = order + n; g < n + 2 * order; g++) {
      dx[g] = dx[g - 1];
      x[g] = x[g - 1] + dx[g - 1];
    }
  }
  free(zn);
  free(zdx);
  return rc;
}

/* ----------------------------------------------------- grids and states */
void spbte_velocity_grid(int N, double L_v, int inhom, double *v, double *eta) {
  const double dv = 2 * L_v / (N - 1);
  for (int i = 0; i < N; i++) v[i] = -L_v + i * dv;
  double deta, L_eta;
  if (!inhom) { /* src/initializer.c:73-80 */
    deta = (2 * M_PI / N) / dv;
    L_eta = (N % 2 == 0) ? 0.5 * N * deta : 0.5 * (N - 1) * deta;
  } else {      /* src/initializer.c:262-266 */
    L_eta = 0.5 * (N - 1) * M_PI / L_v;
    deta = M_PI * (N - 1) / (N * L_v);
  }
  for (int i = 0; i < N; i++) eta[i] = -L_eta + i * deta;
}

/* Boltzmann constant of the initial states: 1 for the "default" species, SI
 * otherwise (src/initializer.c:52-55, :249-252). */
static double state_kb(const species *sp) { return strcmp(sp->name, "default") ? KB_JOULES_PER_KELVIN : 1.0; }

/* rho exp(-|v - (ux,0,0)|^2 / T) / (pi T)^3/2, the 1D states' Maxwellian
 * (src/initializer.c:366, :388, :395, :402, :410, :417). */
static void max_1d(int N, const double *v, double rho, double ux, double T, double *f) {
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < N; k++)
        f[k + N * (j + N * i)] =
            rho * exp(-((v[i] - ux) * (v[i] - ux) + v[j] * v[j] + v[k] * v[k]) / T) / ((T * M_PI) * sqrt(T * M_PI));
}

/* (rho/m) (m / (2 pi KB T))^3/2 exp(-m |v|^2 / (2 KB T)) (src/initializer.c:166, :381) */
static void max_species(int N, const double *v, double rho, double T, double m, double kb, double *f) {
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < N; k++)
        f[k + N * (j + N * i)] = (rho / m) * pow(0.5 * m / (M_PI * kb * T), 1.5) *
                                 exp(-(0.5 * m / (kb * T)) * ((v[i]) * (v[i]) + v[j] * v[j] + v[k] * v[k]));
}

static int state_0d(int N, double L_v, int init, const species *sp, const double *v, double *f) {
  switch (init) {
    case 0: { /* shifted isotropic, :94-104 */
      const double sigma = 0.3 * L_v, S = 10.0;
      for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
          for (int k = 0; k < N; k++) {
            const double r = sqrt(v[i] * v[i] + v[j] * v[j] + v[k] * v[k]);
            f[k + N * (j + N * i)] = exp(-1 * S * (r - sigma) * (r - sigma) / (sigma * sigma)) / (S * S);
          }
      return SPBTE_OK;
    }
    case 1: /* discontinuous Maxwellian, :107-120 */
      for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
          for (int k = 0; k < N; k++) {
            const double m = exp(-(v[i] * v[i] + v[j] * v[j] + v[k] * v[k])) / (M_PI * sqrt(M_PI));
            f[k + N * (j + N * i)] = i < N / 2 ? m : m / 2;
          }
      return SPBTE_OK;
    case 2: { /* BKW at t = 5.5, :123-136 */
      const double K = 1 - exp(-5.5 / 6.0), T = 1.0;
      for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
          for (int k = 0; k < N; k++) {
            const double V2 = v[i] * v[i] + v[j] * v[j] + v[k] * v[k];
            f[k + N * (j + N * i)] = (exp(-V2 / (2 * K * T * T))) / (2.0 * pow(2 * M_PI * K * T * T, 1.5)) *
                                     ((5 * K - 3) / K + (1 - K) * V2 / (K * K * T * T));
          }
      return SPBTE_OK;
    }
    case 3: { /* two Maxwellians at +-2 sigma along v_x, :139-149 */
      const double sigma = M_PI * L_v / 10.0, pre = 0.5 / pow(2.0 * M_PI * sigma * sigma, 1.5);
      for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
          for (int k = 0; k < N; k++) {
            const double a = v[i] - 2.0 * sigma, b = v[i] + 2 * sigma, yz = v[j] * v[j] + v[k] * v[k];
            f[k + N * (j + N * i)] = pre * (exp(-(a * a + yz) / (2.0 * sigma * sigma)) +
                                            exp(-(b * b + yz) / (2.0 * sigma * sigma)));
          }
      return SPBTE_OK;
    }
    case 4: /* Maxwellian rho = 1, T = 1 of the species, :152-168 */
      max_species(N, v, 1.0, 1.0, sp->mass, state_kb(sp), f);
      return SPBTE_OK;
    case 5: /* perturbed Maxwellian, :172-180 */
      for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
          for (int k = 0; k < N; k++) {
            const double V2 = v[i] * v[i] + v[j] * v[j] + v[k] * v[k];
            f[k + N * (j + N * i)] = (1 + 0.1 * sin(V2)) * exp(-(V2)) / (M_PI * sqrt(M_PI));
          }
      return SPBTE_OK;
    default:
      return spbte_fail(SPBTE_ERR_UNSUPPORTED, "Init_field %d has no space-homogeneous initial state (0-5)", init);
  }
}

static int state_1d(int N, int init, int nX, int order, const species *sp, const double *v, double *f) {
  const size_t n3 = (size_t)N * N * N;
  double rho_l = 1.0, ux_l = 0.0, T_l = 1.0, rho_r = 1.0, ux_r = 0.0, T_r = 1.0;
  const double TWall = 1.0, Ma = 1.0; /* src/initializer.c:247, :253 */
  int split = 0;                      /* left state for cells l < nX/2, l counting the ghosts */
  switch (init) {
    case 0: /* :307-315 (the velocities are not used by the state) */
      rho_l = 4.0 * Ma * Ma / (Ma * Ma + 3.0);
      T_l = (5.0 * Ma * Ma - 1.0) * (Ma * Ma + 3.0) / (16.0 * Ma * Ma);
      split = 1;
      break;
    case 1: /* sudden heating, :317-320, :377-383 */
      for (int c = 0; c < nX; c++) max_species(N, v, rho_r, TWall, sp->mass, state_kb(sp), f + c * n3);
      return SPBTE_OK;
    case 2: /* every cell at the left state, :322-328, :384-390 */
      ux_l = -1.0;
      break;
    case 3: /* heat transfer, every cell at rho 1, T 1.5, :330-334, :391-397 */
      T_l = 1.5;
      break;
    case 5: /* Poiseuille, :336-339 */
      break;
    case 6: /* Mach 1.2 shock, :341-350 */
      ux_l = 1.2972;
      rho_r = 1.297;
      ux_r = 1.0;
      T_r = 1.195;
      split = 1;
      break;
    default:
      return spbte_fail(SPBTE_ERR_UNSUPPORTED, "Init_field %d has no space-inhomogeneous initial state (0-3, 5, 6)",
                        init);
  }
  for (int c = 0; c < nX; c++) {
    const int left = !split || (c + order) < nX / 2;
    max_1d(N, v, left ? rho_l : rho_r, left ? ux_l : ux_r, left ? T_l : T_r, f + c * n3);
  }
  return SPBTE_OK;
}

int spbte_initial_state(const spbte_input *in, const species *sp, const double *v, int nX, double *f) {
  if (!in || !sp || !v || !f) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  if (!in->inhom) return state_0d(in->N, in->L_v, in->init_flag, sp, v, f);
  if (nX < 1) return spbte_fail(SPBTE_ERR_ARG, "nX must be positive");
  return state_1d(in->N, in->init_flag, nX, in->order, sp, v, f);
}

/* ---------------------------------------------------------------- driver */
static double wall_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int form_from_env(void) {
  const char *s = getenv("SPBTE_FORM");
  if (!s || !strcmp(s, "fp64")) return SPBTE_FORM_FP64;
  if (!strcmp(s, "fp32")) return SPBTE_FORM_FP32;
  if (!strcmp(s, "fp16x2")) return SPBTE_FORM_FP16X2;
  if (!strcmp(s, "fp32a")) return SPBTE_FORM_FP32A;
  return -1;
}

/* Walls of a 1D run from Init_field, as upwindOne / upwindTwo choose them
 * (src/transportroutines.c:117-172, :359-404): T0 = 1, T1 = 2 (:46-47). */
static int walls_of(int init, int order, double *T0, double *T1) {
  *T0 = 1.0;
  *T1 = 2.0;
  if (init == 3 || init == 5) return SPBTE_BC_DIFFUSE;
  if (init == 1) {
    *T0 = 2.0 * 1.0; /* 2 TWall with TWall = 1 */
    return SPBTE_BC_DIFFUSE_LEFT;
  }
  if (init == 6 && order == 1) return SPBTE_BC_PERIODIC;
  return SPBTE_BC_COPY;
}

/* The weight table as exec/boltz.c:149-163 obtains it, from Weights/ or
 * generated on the GPU and stored there: isotropic (src/weights.c:65-124) or,
 * for Anisotropic 1, the glance = 0 Landau table (src/aniso_weights.c:37-146,
 * which boltz_ calls with glance 0). */
static int setup_weights(spbte_op *op, const spbte_input *in, const species *si, const species *sj) {
  char name[512];
  if (in->aniso)
    snprintf(name, sizeof name, "Weights/N%d_Aniso_L_v%g_lambda%g_Landau.wts", in->N, in->L_v, in->lambda);
  else if (!strcmp(si->name, "default"))
    snprintf(name, sizeof name, "Weights/N%d_isotropic_L_v%g_lambda%g.wts", in->N, in->L_v, in->lambda);
  else
    snprintf(name, sizeof name, "Weights/N%d_isotropic_L_v%g_HS_%s_%zu_%s_%zu.wts", in->N, in->L_v, si->name, si->id,
             sj->name, sj->id);
  if (in->weight_flag == 0) {
    FILE *fp = fopen(name, "rb");
    if (fp) {
      fclose(fp);
      printf("Loading weights from file %s\n", name);
      return spbte_op_load_weights_file(op, name);
    }
    printf("Stored weights not found for this configuration, generating ...\n");
  } else {
    printf("Fresh version of weights being computed and stored for this configuration\n");
  }
  if (in->aniso) CHK(spbte_op_generate_weights_landau(op, in->lambda));
  else CHK(spbte_op_generate_weights_iso(op, in->lambda, si->d_ref, sj->d_ref, si->mass, sj->mass));
  return spbte_op_store_weights_file(op, name);
}

/* exec/MPIWeightGenerator.c main on one GPU: N glance beta lambda L_v (beta
 * is read and unused, as there).  The grid is the generator's own (h_eta =
 * 2 pi / (N h_v), L_eta = N h_eta / 2, :63-74); an existing file is kept
 * (:87-90).  glance = 0 writes the Landau table to
 * Weights/N%d_Aniso_L_v%g_lambda%g_Landau.wts, glance > 0 the glancing-
 * collision table (exec/MPIcollisionRoutines.c:33-203) to
 * Weights/N%d_AnIso_L_v%g_lambda%g_glance%g.wts (:80-83). */
int spbte_weightgen_main(int argc, char **argv) {
  if (argc < 6 || !argv) {
    fprintf(stderr,
            "usage: %s N glance beta lambda L_v   (writes Weights/N<N>_Aniso_L_v<L_v>_lambda<lambda>_Landau.wts for "
            "glance 0, Weights/N<N>_AnIso_L_v<L_v>_lambda<lambda>_glance<glance>.wts otherwise)\n",
            (argc > 0 && argv && argv[0]) ? argv[0] : "weightgen_amd");
    return 1;
  }
  const int N = atoi(argv[1]);
  const double glance = atof(argv[2]), lambda = atof(argv[4]), L_v = atof(argv[5]);
  int rc = SPBTE_OK;
  if (N < 2 || !(L_v > 0) || !isfinite(lambda))
    rc = spbte_fail(SPBTE_ERR_ARG, "need N >= 2, L_v > 0 and a finite lambda");
  else if (glance != 0 && !(glance > 0 && glance < M_PI))
    rc = spbte_fail(SPBTE_ERR_ARG, "glance %g outside (0, pi), where log(sin(glance/2)) < 0", glance);
  char name[512];
  if (glance == 0) snprintf(name, sizeof name, "Weights/N%d_Aniso_L_v%g_lambda%g_Landau.wts", N, L_v, lambda);
  else snprintf(name, sizeof name, "Weights/N%d_AnIso_L_v%g_lambda%g_glance%g.wts", N, L_v, lambda, glance);
  FILE *fp = rc ? NULL : fopen(name, "rb");
  if (fp) {
    fclose(fp);
    printf("Stored weights already computed! \n");
    return 0;
  }
  double *v = rc ? NULL : malloc(sizeof(double) * N), *eta = rc ? NULL : malloc(sizeof(double) * N);
  spbte_op *op = NULL;
  if (!rc && (!v || !eta)) rc = spbte_fail(SPBTE_ERR_ARG, "out of host memory");
  if (!rc) {
    const double h_v = 2.0 * L_v / (double)(N - 1), h_eta = (2.0 * M_PI / N) / h_v, L_eta = 0.5 * N * h_eta;
    for (int i = 0; i < N; i++) {
      eta[i] = -L_eta + (double)i * h_eta;
      v[i] = -L_v + (double)i * h_v;
    }
    printf("Stored weights not found, generating...\n");
    rc = spbte_op_create(&op, N, L_v, v, eta);
  }
  if (!rc) rc = glance == 0 ? spbte_op_generate_weights_landau(op, lambda)
                            : spbte_op_generate_weights_glance(op, lambda, glance);
  if (!rc) rc = spbte_op_store_weights_file(op, name);
  spbte_op_destroy(op);
  free(v);
  free(eta);
  if (rc) {
    fprintf(stderr, "spectralbte_amd: %s\n", spbte_last_error());
    return 1;
  }
  printf("Wrote %s\n", name);
  return 0;
}

static int boltz_run(const char *input_name, const char *flags_name) {
  char path[1024], tpath[1024];
  spbte_input in;
  snprintf(path, sizeof path, "./input/%s", input_name);
  printf("Opening input file %s\n", path);
  CHK(spbte_read_input(path, &in));
  const int Ns = in.num_species;
  if (Ns > 1 && in.aniso)
    return spbte_fail(SPBTE_ERR_UNSUPPORTED,
                      "Anisotropic 1 with %d species: the reference loads one table for the first pair only "
                      "(exec/boltz.c:151-153)",
                      Ns);
  if (in.N < 2 || !(in.L_v > 0) || !(in.Kn > 0) || !(in.dt > 0) || in.nT < 0 || in.data_freq < 1 ||
      (in.order != 1 && in.order != 2) || in.inhom < 0 || in.inhom > 1)
    return spbte_fail(SPBTE_ERR_ARG,
                      "%s: need N >= 2, L_v, Knudsen, Time_step > 0, Data_writing_frequency >= 1, "
                      "Space_order 1 or 2, SpaceInhom 0 or 1",
                      path);
  if (in.init_flag < 0) return spbte_fail(SPBTE_ERR_ARG, "%s: Init_field is not set", path);
  const int form = form_from_env();
  if (form < 0) return spbte_fail(SPBTE_ERR_ARG, "SPBTE_FORM must be fp64, fp32, fp16x2 or fp32a");

  /* load_and_allocate_spec, src/species.c:45-160 */
  species sp[SPBTE_MAX_SPECIES];
  double masses[SPBTE_MAX_SPECIES];
  for (int l = 0; l < Ns; l++) {
    spbte_species_defaults(&sp[l], l);
    if (strcmp(in.species_names[l], "default") != 0) {
      snprintf(path, sizeof path, "./input/species/%s", in.species_names[l]);
      printf("Opening species input file %s\n", path);
      CHK(spbte_read_species(path, &sp[l]));
    }
    masses[l] = sp[l].mass;
  }

  const int N = in.N;
  const size_t n3 = (size_t)N * N * N;
  int nX = 1, rc = SPBTE_OK;
  double *x = NULL, *dx = NULL, *f = NULL, dx_min = 0.0;
  double *v = malloc(sizeof(double) * N), *eta = malloc(sizeof(double) * N);
  spbte_op *op = NULL;
  spbte_run *run = NULL;
  if (!v || !eta) rc = spbte_fail(SPBTE_ERR_ARG, "out of host memory");
  if (!rc && in.inhom) {
    snprintf(path, sizeof path, "./input/%s", in.mesh_file);
    printf("Loading mesh %s\n", path);
    rc = spbte_read_mesh(path, in.order, &nX, NULL, NULL, NULL);
    if (!rc) {
      x = malloc(sizeof(double) * (nX + 2 * in.order));
      dx = malloc(sizeof(double) * (nX + 2 * in.order));
      rc = (x && dx) ? spbte_read_mesh(path, in.order, &nX, x, dx, &dx_min)
                     : spbte_fail(SPBTE_ERR_ARG, "out of host memory");
    }
  }
  if (!rc) {
    spbte_velocity_grid(N, in.L_v, in.inhom, v, eta);
    f = malloc(sizeof(double) * n3 * nX * Ns);
    if (!f) rc = spbte_fail(SPBTE_ERR_ARG, "out of host memory");
    for (int l = 0; l < Ns && !rc; l++) {
      /* the states' Boltzmann constant follows species 0's name (src/initializer.c:52-55) */
      species sl = sp[l];
      strcpy(sl.name, sp[0].name);
      rc = spbte_initial_state(&in, &sl, v, nX, f + (size_t)l * nX * n3);
    }
  }
  if (!rc) rc = spbte_op_create(&op, N, in.L_v, v, eta);
  if (!rc) rc = spbte_op_set_species(op, Ns, masses);
  /* the table of pair (i, j) in slot j*Ns + i, exec/boltz.c:157-160 */
  for (int i = 0; i < Ns && !rc; i++)
    for (int j = 0; j < Ns && !rc; j++) {
      rc = spbte_op_select_table(op, j * Ns + i);
      if (!rc) rc = setup_weights(op, &in, &sp[i], &sp[j]);
    }
  if (!rc) rc = spbte_op_select_table(op, 0);

  double T0 = 1.0, T1 = 2.0;
  if (!rc && !in.inhom) {
    rc = spbte_run0d_create_species(&run, op, Ns, 1, in.order, in.dt, in.Kn, form, 1);
  } else if (!rc) {
    const int bc = walls_of(in.init_flag, in.order, &T0, &T1);
    if (in.init_flag == 5 && in.order == 1)
      rc = spbte_fail(SPBTE_ERR_UNSUPPORTED,
                      "Init_field 5 (Poiseuille) needs Space_order 2: the reference's order-1 body force is applied "
                      "outside the cell (src/transportroutines.c:218-226)");
    if (!rc) rc = spbte_run1d_create_species(&run, op, Ns, nX, in.order, x, dx, in.dt, in.Kn, bc, T0, T1, form);
    if (!rc && in.init_flag == 5) rc = spbte_run_set_forcing(run, 1.0); /* Ma = 1, :176-179 */
  }
  if (!rc) rc = spbte_run_set_state(run, f);

  /* restart (1D, src/initializer.c:287-296 and src/restart.c): one pdf per species */
  int t = 0;
  char pdf[SPBTE_MAX_SPECIES][1024];
  for (int l = 0; l < Ns; l++) snprintf(pdf[l], sizeof pdf[l], "Restart/%s_rank0_%s.plt", input_name, sp[l].name);
  snprintf(tpath, sizeof tpath, "Restart/%s_time.plt", input_name);
  if (!rc && in.inhom && in.restart == 1) {
    printf("Loading from previously generated data %s\n", pdf[0]);
    rc = spbte_run_load_restart(run, pdf[0], tpath, &t);
    for (int l = 1; l < Ns && !rc; l++) rc = spbte_run_load_restart_species(run, l, pdf[l]);
    if (!rc && (t < 0 || t > in.nT)) rc = spbte_fail(SPBTE_ERR_IO, "%s holds step %d of %d", tpath, t, in.nT);
  }
  if (!rc) rc = spbte_run_use_graph(run, 1);

  /* output (src/output.c:64-157, :215-293): one file per species, named
   * moments_<input> for one species and moments_<input>_<name> for several */
  int flags = 0, opened = 0;
  if (!rc) {
    snprintf(path, sizeof path, "./input/%s", flags_name);
    rc = spbte_read_output_flags(path, &flags);
  }
  for (int l = 0; l < Ns && !rc; l++) {
    if (Ns > 1) snprintf(path, sizeof path, "Data/moments_%s_%s", input_name, sp[l].name);
    else snprintf(path, sizeof path, "Data/moments_%s", input_name);
    printf("Opening output file %s\n", path);
    rc = spbte_run_open_output_species(run, l, path, flags, in.restart != 0);
    opened |= !rc;
  }
  if (!rc && (!in.inhom || !in.restart)) rc = spbte_run_write_output(run, 0.0);

  /* time loop (exec/boltz.c:181-398): output every data_freq steps at dt (t+1) */
  if (!rc) printf("Done with all setup, starting main loop\n");
  const double t_start = wall_seconds();
  double w_start = t_start;
  int count = 0;
  while (!rc && t < in.nT) {
    int n = in.data_freq - count;
    if (n > in.nT - t) n = in.nT - t;
    rc = spbte_run_step(run, n);
    t += n;
    count += n;
    if (rc || count < in.data_freq) continue;
    count = 0;
    if (in.inhom && in.restart_time > 0) { /* exec/boltz.c:360-388 */
      rc = spbte_op_synchronize(op);
      const double now = wall_seconds(), write_time = now - w_start, tot = now - t_start;
      w_start = now;
      if (!rc && tot + write_time > 0.95 * in.restart_time) {
        printf("RESTART TIME REACHED - STORING CURRENT DISTRIBUTION DATA\n");
        /* the reference stores the index of the step just taken, t - 1 here */
        rc = spbte_run_store_restart(run, pdf[0], tpath, t - 1);
        for (int l = 1; l < Ns && !rc; l++) rc = spbte_run_store_restart_species(run, l, pdf[l]);
        if (!rc) printf("Time stored: %d\n", t - 1);
        break;
      }
    }
    if (!rc) rc = spbte_run_write_output(run, in.dt * t);
    if (!rc) printf("In step %d of %d\n", t, in.nT);
  }
  if (!rc) rc = spbte_op_synchronize(op);
  if (opened) {
    const int rc2 = spbte_run_close_output(run);
    if (!rc) rc = rc2;
  }
  if (!rc) printf("Wrapping up\n");
  fflush(stdout);
  spbte_run_destroy(run);
  spbte_op_destroy(op);
  free(v);
  free(eta);
  free(x);
  free(dx);
  free(f);
  return rc;
}

int spbte_boltz_main(int argc, char **argv) {
  if (argc < 3 || !argv || !argv[1] || !argv[2]) {
    fprintf(stderr, "usage: %s <input file> <output flags file>   (both under ./input/)\n",
            (argc > 0 && argv && argv[0]) ? argv[0] : "boltz_amd");
    return 1;
  }
  const int rc = boltz_run(argv[1], argv[2]);
  if (rc) {
    fprintf(stderr, "spectralbte_amd: %s\n", spbte_last_error());
    fflush(stderr);
    return 1;
  }
  return 0;
}
