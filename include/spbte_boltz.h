/* spbte_boltz.h -- the boltz_ front end over the batched runs.
 *
 * The reference's driver (exec/boltz.c) reads an input file of keywords, an
 * optional mesh file and species files, builds the grids and the initial
 * state, and advances one or several species in time with the moment output
 * of src/output.c and the restart dumps of src/restart.c.  spbte_boltz_main does
 * the same with every time step on the GPU (spbte_run0d / spbte_run1d), and
 * the pieces are exported for hosts that build their own driver.
 *
 * Files are found where boltz_ finds them, relative to the working directory:
 * input/<input file>, input/<output flags file>, input/<mesh file>,
 * input/species/<name>, Weights/, Data/moments_<input file>, Restart/.
 */
#ifndef SPBTE_BOLTZ_H
#define SPBTE_BOLTZ_H

#include "species.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SPBTE_NAME_LEN 80
#define SPBTE_MAX_SPECIES 8

/* The keywords of an input file (src/input.c:10-183), with the defaults of
 * set_default_values (:132-183).  init_flag is -1 until Init_field is read. */
typedef struct {
  int N;               /* N */
  double L_v;          /* L_v */
  double Kn;           /* Knudsen */
  double lambda;       /* Lambda */
  double dt;           /* Time_step (-1: not set) */
  int nT;              /* Number_of_time_steps */
  int order;           /* Space_order (0D: order of the time integrator) */
  int data_freq;       /* Data_writing_frequency */
  int restart;         /* Restart */
  double restart_time; /* Restart_time [s] */
  int init_flag;       /* Init_field */
  int bc_flag;         /* Bound_cond (read; the walls follow Init_field) */
  int inhom;           /* SpaceInhom */
  int weight_flag;     /* Recompute_weights */
  int aniso;           /* Anisotropic */
  char mesh_file[SPBTE_NAME_LEN];                      /* mesh_file ("not set") */
  int num_species;                                     /* num_species */
  char species_names[SPBTE_MAX_SPECIES][SPBTE_NAME_LEN]; /* names after num_species */
} spbte_input;

/* Replaces read_input (src/input.c:10-127): path is the input file itself. */
int spbte_read_input(const char *path, spbte_input *in);

/* Replaces set_spec_default_values (src/species.c:28-43) for one species. */
void spbte_species_defaults(species *sp, int id);
/* Replaces the per-file part of load_and_allocate_spec (src/species.c:64-159):
 * reads name, num_en_levels, mass, En_for, d_ref, T_ref, mu_ref, omega. */
int spbte_read_species(const char *path, species *sp);

/* Replaces make_mesh (src/mesh_setup.c:9-200) for one rank: the zones of the
 * mesh file give nX cells; x and dx (nX + 2*order entries, ghost cells by
 * extension of the end cells on both sides) may be NULL to query nX first. */
int spbte_read_mesh(const char *path, int order, int *nX, double *x, double *dx, double *dx_min);

/* Velocity grid v[i] = -L_v + i dv, dv = 2 L_v / (N-1), and the Fourier grid
 * eta: 0D as initialize_hom (src/initializer.c:66-82), 1D as initialize_inhom
 * (:256-266); the two differ in rounding only. */
void spbte_velocity_grid(int N, double L_v, int inhom, double *v, double *eta);

/* Initial state of Init_field: 0D (inhom = 0) cases 0-5 of initialize_hom
 * (src/initializer.c:92-199) into N^3 doubles; 1D cases 0, 1, 2, 3, 5, 6 of
 * initialize_inhom (:300-423) into nX * N^3 doubles (interior cells). */
int spbte_initial_state(const spbte_input *in, const species *sp, const double *v, int nX, double *f);

/* exec/boltz.c main: argv[1] = input file, argv[2] = output flags file, both
 * under input/.  SPBTE_FORM = fp32 | fp16x2 | fp32a in the environment selects
 * a reduced form of the weight table (FP64 by default).  Returns 0 on success;
 * messages go to stdout, errors to stderr. */
int spbte_boltz_main(int argc, char **argv);

/* exec/MPIWeightGenerator.c main on one GPU: argv = N glance beta lambda L_v.
 * glance 0 writes Weights/N%d_Aniso_L_v%g_lambda%g_Landau.wts, 0 < glance < pi
 * the glancing-collision table Weights/N%d_AnIso_L_v%g_lambda%g_glance%g.wts
 * (exec/MPIWeightGenerator.c:80-83); an existing file is kept.  Returns 0 on success. */
int spbte_weightgen_main(int argc, char **argv);

#ifdef __cplusplus
}
#endif

#endif
