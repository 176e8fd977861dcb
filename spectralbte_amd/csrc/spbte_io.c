/* spbte_io.c -- moment output and restart dumps of a batched run (C99).
 *
 * Output follows src/output.c: a header line, then per output time one line
 * per cell of "%le " columns -- 0D (write_streams, :159-213): time, density,
 * velocity_x, temperature, pressure, " %le " positive/negative energy ratio,
 * " %le " slice f[N/2 + N (N/2 + N l)]; 1D (write_streams_inhom, :295-404):
 * time, position, density, velocity_x, temperature, pressure.  Which columns
 * appear is set by the flags of the reference's output-flags file
 * (read_flags, :417-443).  An ensemble of 0D problems writes one line per
 * problem at each time.  Restart dumps follow src/restart.c: the interior cells'
 * N^3 doubles back to back, and the step counter as one int.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spbte.h"
#include "spbte_internal.h"

#define MAX_FILES_OPEN 64

typedef struct {
  spbte_run *run;
  int species;
  FILE *fp;
  int flags;
} out_slot;

static out_slot slots[MAX_FILES_OPEN];

static out_slot *slot_of(spbte_run *run, int species) {
  for (int i = 0; i < MAX_FILES_OPEN; i++)
    if (slots[i].run == run && slots[i].fp && slots[i].species == species) return &slots[i];
  return NULL;
}

int spbte_read_output_flags(const char *path, int *flags) {
  if (!path || !flags) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  FILE *fp = fopen(path, "r");
  if (!fp) return spbte_fail(SPBTE_ERR_IO, "cannot open output flags file %s", path);
  /* defaults of src/output.c:32-38: everything on */
  int f = SPBTE_OUT_DENSITY | SPBTE_OUT_VELOCITY | SPBTE_OUT_TEMPERATURE | SPBTE_OUT_PRESSURE | SPBTE_OUT_ENTROPY |
          SPBTE_OUT_SLICE;
  static const struct {
    const char *key;
    int bit;
  } keys[] = {{"density", SPBTE_OUT_DENSITY},   {"velocity", SPBTE_OUT_VELOCITY}, {"temperature", SPBTE_OUT_TEMPERATURE},
              {"pressure", SPBTE_OUT_PRESSURE}, {"slice", SPBTE_OUT_SLICE},       {"entropy", SPBTE_OUT_ENTROPY}};
  char tok[80];
  int found_stop = 0;
  while (fscanf(fp, "%79s", tok) == 1) {
    if (!strcmp(tok, "Stop")) { found_stop = 1; break; }
    for (size_t k = 0; k < sizeof keys / sizeof keys[0]; k++)
      if (!strcmp(tok, keys[k].key)) {
        long val = 0;
        if (fscanf(fp, "%ld", &val) != 1) {
          fclose(fp);
          return spbte_fail(SPBTE_ERR_IO, "no value after %s in %s", tok, path);
        }
        f = val ? (f | keys[k].bit) : (f & ~keys[k].bit);
      }
  }
  fclose(fp);
  if (!found_stop) return spbte_fail(SPBTE_ERR_IO, "no Stop line in %s", path);
  *flags = f;
  return SPBTE_OK;
}

int spbte_run_open_output_species(spbte_run *run, int species, const char *path, int flags, int append) {
  if (!run || !path) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  if (species < 0 || species >= spbte_run_nspecies(run)) return spbte_fail(SPBTE_ERR_ARG, "no species %d", species);
  if (slot_of(run, species)) return spbte_fail(SPBTE_ERR_ARG, "species %d already has an open output file", species);
  out_slot *sl = NULL;
  for (int i = 0; i < MAX_FILES_OPEN && !sl; i++)
    if (!slots[i].fp) sl = &slots[i];
  if (!sl) return spbte_fail(SPBTE_ERR_ARG, "too many open output files");
  spbte_op *op;
  int dim, ncell;
  const double *d_state, *d_x;
  int rc = spbte_run_geometry(run, &op, &dim, &ncell, &d_state, &d_x);
  if (rc) return rc;
  FILE *fp = fopen(path, append ? "a" : "w");
  if (!fp) return spbte_fail(SPBTE_ERR_IO, "cannot open %s", path);
  {  /* the header is written on append too, as src/output.c:100-135 does */
    fprintf(fp, "#Time ");
    if (dim == 1) fprintf(fp, "Position ");
    if (flags & SPBTE_OUT_DENSITY) fprintf(fp, "Density ");
    if (flags & SPBTE_OUT_VELOCITY) fprintf(fp, "Velocity_x ");
    if (flags & SPBTE_OUT_TEMPERATURE) fprintf(fp, "Temperature ");
    if (flags & SPBTE_OUT_PRESSURE) fprintf(fp, "Pressure ");
    if (dim == 0) {
      const spbte_grid_dev *g = spbte_op_grid(op);
      if (flags & SPBTE_OUT_ENTROPY) fprintf(fp, "Pos/neg_energy_ratio ");
      if (flags & SPBTE_OUT_SLICE) {
        double *v = malloc(sizeof(double) * g->N);
        if (!v || hipMemcpy(v, g->d_v, sizeof(double) * g->N, hipMemcpyDeviceToHost) != hipSuccess) {
          free(v);
          fclose(fp);
          return spbte_fail(SPBTE_ERR_HIP, "velocity grid copy failed");
        }
        for (int l = 0; l < g->N; l++) fprintf(fp, "v:%le ", v[l]);
        free(v);
      }
    }
    fprintf(fp, "\n");
  }
  sl->run = run;
  sl->species = species;
  sl->fp = fp;
  sl->flags = flags;
  return SPBTE_OK;
}

int spbte_run_open_output(spbte_run *run, const char *path, int flags, int append) {
  return spbte_run_open_output_species(run, 0, path, flags, append);
}

/* Lines of one species at one time into its file. */
static int write_species(out_slot *sl, spbte_run *run, double time) {
  spbte_op *op;
  int dim, ncell;
  const double *d_state0, *d_x;
  int rc = spbte_run_geometry(run, &op, &dim, &ncell, &d_state0, &d_x);
  if (rc) return rc;
  double *d_state = NULL;
  rc = spbte_run_species_state(run, sl->species, &d_state);
  if (rc) return rc;
  const spbte_grid_dev *g = spbte_op_grid(op);
  const int N = g->N, per = 2 + N, ns = spbte_run_nspecies(run);
  double *mom = malloc(sizeof(double) * 5 * ncell * ns), *q = malloc(sizeof(double) * per * ncell);
  double *x = dim == 1 ? malloc(sizeof(double) * ncell) : NULL;
  double *d_q = NULL;
  if (!mom || !q || (dim == 1 && !x)) rc = spbte_fail(SPBTE_ERR_ARG, "out of host memory");
  if (!rc) rc = spbte_run_moments(run, mom); /* per species rho, u_x, u_y, u_z, T; synchronous */
  if (!rc && dim == 0) {
    if (hipMalloc((void **)&d_q, sizeof(double) * per * ncell) != hipSuccess)
      rc = spbte_fail(SPBTE_ERR_HIP, "output buffer allocation failed");
    if (!rc && spbte_launch_out_quantities(g, d_state, ncell, d_q, spbte_op_stream(op)) != 0)
      rc = spbte_fail(SPBTE_ERR_HIP, "output kernel launch failed");
    if (!rc && (hipStreamSynchronize(spbte_op_stream(op)) != hipSuccess ||
                hipMemcpy(q, d_q, sizeof(double) * per * ncell, hipMemcpyDeviceToHost) != hipSuccess))
      rc = spbte_fail(SPBTE_ERR_HIP, "output copy failed");
  }
  if (!rc && dim == 1 && hipMemcpy(x, d_x, sizeof(double) * ncell, hipMemcpyDeviceToHost) != hipSuccess)
    rc = spbte_fail(SPBTE_ERR_HIP, "mesh copy failed");
  /* species-0 mass m for every species, as src/output.c:171-177 and
   * :311-313 call the moment routines with species 0: rho = m int f, u
   * unchanged, T = (m / KB) times the mass-1 temperature
   * (src/momentRoutines.c:58-183); exact for m = KB = 1 */
  const double m = g->mass, tfac = g->mass / g->kb;
  const double *ms = mom + (size_t)5 * ncell * sl->species;
  for (int c = 0; c < ncell && !rc; c++) {
    const double rho = m * ms[5 * c], ux = ms[5 * c + 1], T = tfac * ms[5 * c + 4];
    FILE *fp = sl->fp;
    fprintf(fp, "%le ", time);
    if (dim == 1) fprintf(fp, "%le ", x[c]);
    if (sl->flags & SPBTE_OUT_DENSITY) fprintf(fp, "%le ", rho);
    if (sl->flags & SPBTE_OUT_VELOCITY) fprintf(fp, "%le ", ux);
    if (sl->flags & SPBTE_OUT_TEMPERATURE) fprintf(fp, "%le ", T);
    if (sl->flags & SPBTE_OUT_PRESSURE) fprintf(fp, "%le ", rho * T); /* getPressure, src/momentRoutines.c:187-190 */
    if (dim == 0) {
      const double *qc = q + (size_t)per * c;
      if (sl->flags & SPBTE_OUT_ENTROPY) fprintf(fp, " %le ", qc[1] / qc[0]);
      if (sl->flags & SPBTE_OUT_SLICE)
        for (int l = 0; l < N; l++) fprintf(fp, " %le ", qc[2 + l]);
    }
    fprintf(fp, "\n");
  }
  if (!rc && fflush(sl->fp) != 0) rc = spbte_fail(SPBTE_ERR_IO, "output write failed");
  if (d_q) hipFree(d_q);
  free(mom);
  free(q);
  free(x);
  return rc;
}

int spbte_run_write_output(spbte_run *run, double time) {
  int any = 0;
  for (int l = 0; l < spbte_run_nspecies(run); l++) {
    out_slot *sl = slot_of(run, l);
    if (!sl) continue;
    any = 1;
    int rc = write_species(sl, run, time);
    if (rc) return rc;
  }
  return any ? SPBTE_OK : spbte_fail(SPBTE_ERR_ARG, "the run has no open output file");
}

int spbte_run_close_output(spbte_run *run) {
  int any = 0, rc = SPBTE_OK;
  for (int i = 0; i < MAX_FILES_OPEN; i++)
    if (slots[i].fp && slots[i].run == run) {
      any = 1;
      if (fclose(slots[i].fp) != 0 && !rc) rc = spbte_fail(SPBTE_ERR_IO, "closing an output file failed");
      slots[i].fp = NULL;
      slots[i].run = NULL;
    }
  return any ? rc : spbte_fail(SPBTE_ERR_ARG, "the run has no open output file");
}

/* One species' interior cells to or from a raw dump (src/restart.c:41-54, :80-95). */
static int species_dump(spbte_run *run, int species, const char *pdf_path, int load) {
  if (!run || !pdf_path) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  spbte_op *op;
  int dim, ncell;
  const double *d_state0, *d_x;
  int rc = spbte_run_geometry(run, &op, &dim, &ncell, &d_state0, &d_x);
  if (rc) return rc;
  double *d_state = NULL;
  rc = spbte_run_species_state(run, species, &d_state);
  if (rc) return rc;
  const size_t n = (size_t)ncell * spbte_op_n(op) * spbte_op_n(op) * spbte_op_n(op);
  double *f = malloc(sizeof(double) * n);
  if (!f) return spbte_fail(SPBTE_ERR_ARG, "out of host memory");
  if (hipStreamSynchronize(spbte_op_stream(op)) != hipSuccess) rc = spbte_fail(SPBTE_ERR_HIP, "stream synchronisation failed");
  FILE *fp = NULL;
  if (!rc && !load) {
    if (hipMemcpy(f, d_state, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
      rc = spbte_fail(SPBTE_ERR_HIP, "state copy failed");
    fp = rc ? NULL : fopen(pdf_path, "wb");
    if (!rc && !fp) rc = spbte_fail(SPBTE_ERR_IO, "cannot open %s", pdf_path);
    if (!rc && (fwrite(f, sizeof(double), n, fp) != n || fflush(fp) != 0))
      rc = spbte_fail(SPBTE_ERR_IO, "writing %s failed", pdf_path);
  } else if (!rc) {
    fp = fopen(pdf_path, "rb");
    if (!fp) rc = spbte_fail(SPBTE_ERR_IO, "Error: unable to open restarted file %s", pdf_path);
    size_t got = rc ? 0 : fread(f, sizeof(double), n, fp);
    if (!rc && got != n)
      rc = spbte_fail(SPBTE_ERR_IO, "Error reloading pdf file %s (%zu of %zu values)", pdf_path, got, n);
    if (!rc && hipMemcpy(d_state, f, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
      rc = spbte_fail(SPBTE_ERR_HIP, "state copy failed");
  }
  if (fp) fclose(fp);
  free(f);
  return rc;
}

int spbte_run_store_restart_species(spbte_run *run, int species, const char *pdf_path) {
  return species_dump(run, species, pdf_path, 0);
}

int spbte_run_load_restart_species(spbte_run *run, int species, const char *pdf_path) {
  return species_dump(run, species, pdf_path, 1);
}

int spbte_run_store_restart(spbte_run *run, const char *pdf_path, const char *time_path, int step) {
  int rc = species_dump(run, 0, pdf_path, 0);
  if (!rc && time_path) {
    FILE *ft = fopen(time_path, "wb");
    if (!ft) return spbte_fail(SPBTE_ERR_IO, "cannot open %s", time_path);
    const int ok = fwrite(&step, sizeof(int), 1, ft) == 1 && fflush(ft) == 0;
    fclose(ft);
    if (!ok) return spbte_fail(SPBTE_ERR_IO, "writing %s failed", time_path);
  }
  return rc;
}

int spbte_run_load_restart(spbte_run *run, const char *pdf_path, const char *time_path, int *step) {
  int rc = species_dump(run, 0, pdf_path, 1);
  if (!rc && time_path && step) {
    FILE *ft = fopen(time_path, "rb");
    if (!ft) return spbte_fail(SPBTE_ERR_IO, "cannot open %s", time_path);
    const int ok = fread(step, sizeof(int), 1, ft) == 1;
    fclose(ft);
    if (!ok) return spbte_fail(SPBTE_ERR_IO, "reading %s failed", time_path);
  }
  return rc;
}
