/* collisions.c -- the reference's collisions.h interface over the MI355X operator.
 *
 * Replaces src/collisions.c.  Module state stays file-static and one grid per
 * process, as in the reference (src/collisions.c:11-30), but the operator runs
 * on the GPU: each call uploads f and g, evaluates the operator with the
 * batched path of spbte_core.c for one cell and downloads Q.  The copies go
 * through a pinned host buffer as asynchronous copies on the operator's
 * stream, so a call waits for the GPU once.  Weight tables
 * are cached in HBM keyed by the conv_weights pointer (the reference passes the
 * same pointer every step, exec/boltz.c:196,288), so an N=32 table (8.6 GB)
 * crosses PCIe once.
 *
 * The precision form defaults to FP64 (the reference method); the environment
 * variable SPBTE_FORM = fp32 | fp16x2 | fp32a selects another (other values
 * warn and keep FP64).  HIP failures print a message and exit(1), which is how
 * the reference handles fatal errors (src/weights.c:84-86).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "collisions.h"
#include "spbte.h"
#include "spbte_internal.h"

#define CACHE_SLOTS (SPBTE_MAX_NS * SPBTE_MAX_NS)

static spbte_op *ops[CACHE_SLOTS]; /* one operator (grid + table) per weight table */
static double **keys[CACHE_SLOTS];
static unsigned long ages[CACHE_SLOTS], clock_;
static spbte_op *op_plain;          /* grid only, for fft3D */
static int N = 0;
static double L_v;
static double *v, *eta;             /* borrowed */
static double *d_buf;               /* device: f, g, Q (3 N^3) or one complex cube (2 N^3) */
static double *h_buf;               /* pinned host staging of the same layout */
static double g_mass = 1.0;

static void die(const char *where) {
  fprintf(stderr, "spectralbte_amd: %s: %s\n", where, spbte_last_error());
  fflush(stderr);
  exit(1);
}

static void hip_ok(hipError_t e, const char *where) {
  if (e == hipSuccess) return;
  fprintf(stderr, "spectralbte_amd: %s: %s\n", where, hipGetErrorString(e));
  fflush(stderr);
  exit(1);
}

static int form_from_env(void) {
  const char *s = getenv("SPBTE_FORM");
  if (!s) return SPBTE_FORM_FP64;
  if (!strcmp(s, "fp32")) return SPBTE_FORM_FP32;
  if (!strcmp(s, "fp16x2")) return SPBTE_FORM_FP16X2;
  if (!strcmp(s, "fp32a")) return SPBTE_FORM_FP32A;
  if (strcmp(s, "fp64")) {
    static int warned;
    if (!warned) fprintf(stderr, "spectralbte_amd: SPBTE_FORM=%s is not fp64, fp32, fp16x2 or fp32a; using fp64\n", s);
    warned = 1;
  }
  return SPBTE_FORM_FP64;
}

/* Mass of the Maxwellian in ComputeQ_maxPreserve; the reference reads it from
 * the species table of src/momentRoutines.c (mixture[0].mass).  conserve.c
 * forwards initialize_conservation's mixture[0].mass here. */
void spbte_abi_set_mass(double mass) {
  g_mass = mass;
  for (int i = 0; i < CACHE_SLOTS; i++)
    if (ops[i]) spbte_op_set_species(ops[i], 1, &g_mass);
}

/* Drop the device copy of a table (called by dealloc_weights). */
void spbte_abi_forget_weights(double **conv_weights) {
  for (int i = 0; i < CACHE_SLOTS; i++)
    if (keys[i] == conv_weights && ops[i]) {
      spbte_op_destroy(ops[i]);
      ops[i] = NULL;
      keys[i] = NULL;
    }
}

void initialize_coll(int nodes, double length, double *vel, double *zeta) {
  if (N) dealloc_coll();
  N = nodes;
  L_v = length;
  v = vel;
  eta = zeta;
  if (spbte_op_create(&op_plain, N, L_v, v, eta) != SPBTE_OK) die("initialize_coll");
  if (spbte_malloc((void **)&d_buf, sizeof(double) * 3 * (size_t)N * N * N) != SPBTE_OK) die("initialize_coll");
  hip_ok(hipHostMalloc((void **)&h_buf, sizeof(double) * 3 * (size_t)N * N * N, hipHostMallocDefault), "initialize_coll");
}

void dealloc_coll(void) {
  for (int i = 0; i < CACHE_SLOTS; i++) {
    if (ops[i]) spbte_op_destroy(ops[i]);
    ops[i] = NULL;
    keys[i] = NULL;
  }
  if (op_plain) spbte_op_destroy(op_plain);
  op_plain = NULL;
  if (d_buf) spbte_free(d_buf);
  if (h_buf) hipHostFree(h_buf);
  d_buf = NULL;
  h_buf = NULL;
  N = 0;
}

static spbte_op *op_for(double **conv_weights) {
  if (!N) {
    fprintf(stderr, "spectralbte_amd: ComputeQ called before initialize_coll\n");
    exit(1);
  }
  int slot = -1, oldest = 0;
  for (int i = 0; i < CACHE_SLOTS; i++) {
    if (keys[i] == conv_weights && ops[i]) slot = i;
    if (ages[i] < ages[oldest]) oldest = i;
  }
  if (slot < 0) {
    /* a free slot, unless HBM cannot hold one more operator (its FP64 table,
     * the copy the streaming kernel reads and scratch): then evict the least
     * recently used tables until it can */
    const size_t n3 = (size_t)N * N * N, need = 2 * sizeof(double) * n3 * n3 + ((size_t)256 << 20);
    slot = -1;
    for (int i = 0; i < CACHE_SLOTS && slot < 0; i++)
      if (!ops[i]) slot = i;
    for (;;) {
      size_t free_b = 0, total_b = 0;
      int used = 0;
      for (int i = 0; i < CACHE_SLOTS; i++) used += ops[i] != NULL;
      if (slot >= 0 && (used == 0 || (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b >= need))) break;
      int lru = -1;
      for (int i = 0; i < CACHE_SLOTS; i++)
        if (ops[i] && (lru < 0 || ages[i] < ages[lru])) lru = i;
      if (lru < 0) break;
      static int warned;
      if (!warned)
        fprintf(stderr, "spectralbte_amd: device memory holds %d weight tables of N=%d; evicting the least recently "
                "used one (each eviction re-uploads a table on its next use)\n", used, N);
      warned = 1;
      spbte_op_destroy(ops[lru]);
      ops[lru] = NULL;
      keys[lru] = NULL;
      if (slot < 0) slot = lru;
    }
    if (slot < 0) slot = oldest;
    if (spbte_op_create(&ops[slot], N, L_v, v, eta) != SPBTE_OK) die("ComputeQ");
    if (spbte_op_set_species(ops[slot], 1, &g_mass) != SPBTE_OK) die("ComputeQ");
    if (spbte_op_set_weights_rows(ops[slot], (const double *const *)conv_weights) != SPBTE_OK) die("ComputeQ");
    keys[slot] = conv_weights;
  }
  ages[slot] = ++clock_;
  return ops[slot];
}

/* Host -> device through the pinned buffer, ordered on stream s. */
static void stage_in(double *d, double *h, const double *src, size_t bytes, hipStream_t s, const char *where) {
  memcpy(h, src, bytes);
  hip_ok(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s), where);
}

/* Device -> host through the pinned buffer after the work queued on s. */
static void finish_out(double *dst, double *h, const double *d, size_t bytes, hipStream_t s, const char *where) {
  hip_ok(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s), where);
  hip_ok(hipStreamSynchronize(s), where);
  memcpy(dst, h, bytes);
}

static void run_q(double *f, double *g, double *Q, double **conv_weights, int flags) {
  spbte_op *op = op_for(conv_weights);
  hipStream_t s = spbte_op_stream(op);
  const size_t n3 = (size_t)N * N * N, b = sizeof(double) * n3;
  double *df = d_buf, *dg = d_buf + n3, *dQ = d_buf + 2 * n3;
  stage_in(df, h_buf, f, b, s, "ComputeQ");
  if (g != f) stage_in(dg, h_buf + n3, g, b, s, "ComputeQ");
  if (spbte_compute_q(op, df, g == f ? df : dg, dQ, 1, form_from_env(), flags | (g == f ? SPBTE_SELF : 0)) != SPBTE_OK)
    die("ComputeQ");
  finish_out(Q, h_buf + 2 * n3, dQ, b, s, "ComputeQ");
}

void ComputeQ(double *f, double *g, double *Q, double **conv_weights) { run_q(f, g, Q, conv_weights, 0); }

void ComputeQ_maxPreserve(double *f, double *g, double *Q, double **conv_weights) {
  run_q(f, g, Q, conv_weights, SPBTE_MAXPRESERVE);
}

void fft3D(fftw_complex *in, fftw_complex *out, int invert) {
  if (!N) {
    fprintf(stderr, "spectralbte_amd: fft3D called before initialize_coll\n");
    exit(1);
  }
  const size_t b = 2 * sizeof(double) * (size_t)N * N * N;
  hipStream_t s = spbte_op_stream(op_plain);
  stage_in(d_buf, h_buf, (const double *)in, b, s, "fft3D");
  if (spbte_transform(op_plain, d_buf, d_buf, 1, invert ? 1 : 0) != SPBTE_OK) die("fft3D");
  finish_out((double *)out, h_buf, d_buf, b, s, "fft3D");
}

void ifft3D(fftw_complex *in, fftw_complex *out) { fft3D(in, out, 1); }
