/* spbte_core.c -- host side of the batched MI355X collision operator (C99).
 *
 * Owns the per-grid device state (transform matrices, weight tables in each
 * precision form, projection factors, scratch), sequences the kernels of one
 * operator evaluation on the operator's HIP stream, picks and tunes the tile
 * of the weighted-sum kernel, and implements the public interface of spbte.h.
 * The reference-compatible entry points (collisions.c, conserve.c, weights.c)
 * are thin wrappers over this file.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spbte.h"
#include "spbte_internal.h"

/* ------------------------------------------------------------------ errors */
static __thread char g_err[512];

const char *spbte_last_error(void) { return g_err; }

static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

/* The same, for the other host modules of the library. */
int spbte_fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(expr)                                                                                  \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) return fail(SPBTE_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                      __FILE__, __LINE__);                                            \
  } while (0)

#define LCHK(expr)                                                                                           \
  do {                                                                                                       \
    int e_ = (expr);                                                                                         \
    if (e_ != 0) return fail(SPBTE_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString((hipError_t)e_),      \
                             __FILE__, __LINE__);                                                            \
  } while (0)

/* --------------------------------------------------------------- op state */
#define NFORMS 4
#define MAXJOBS 4

struct spbte_op {
  spbte_grid_dev g;
  int N;
  size_t n3;
  double *v, *eta; /* host copies */
  hipStream_t s;
  /* weight tables: one slot per species pair of a multi-species run; t is the
   * selected slot, on which the table calls and evaluations act */
  struct wtable {
    double *dW64;      /* FP64 master copy */
    void *dWf[NFORMS]; /* [0] aliases dW64; [1] float (forms 1, 3); [2] (hi,lo) pairs + row scales */
    void *dWp[3][9];   /* tables permuted for the streaming kernel, per storage form and rows per lane */
    int have_w;
  } tab[SPBTE_MAX_NS * SPBTE_MAX_NS], *t;
  /* conservation */
  int Ns;
  double masses[SPBTE_MAX_NS];
  double LU[(SPBTE_MAX_NS + 4) * (SPBTE_MAX_NS + 4)];
  int pivot[SPBTE_MAX_NS + 4];
  /* scratch, sized for cap cells */
  int cap;
  double *d_mom;     /* 2 * cap * SPBTE_MOM_STRIDE */
  void *d_xs;        /* transform scratch: MAXJOBS * cap * n3 double2 */
  void *d_FG;        /* MAXJOBS * cap * n3 double2 (FP32 operands with their part sums use 12 bytes a value) */
  void *d_Qh;        /* cap * n3 double2 */
  double *d_part;    /* 10 * N * cap: per-plane partial sums of the moments (f's, g's) and the projection */
  /* tiles pinned per (form, nprod-1); tr == 0 -> automatic */
  spbte_tile pin[NFORMS][2];
  /* autotuned tiles per (form, nprod-1, ncell bucket) */
  spbte_tile tuned[NFORMS][2][16];
  /* partial sums of the split streaming kernel */
  void *d_Qp;
  size_t qp_bytes;
  /* the fixed workspace of spbte_compute_q_noalloc on this operator's GPU */
  void *nac;
  size_t nac_bytes;
  int nac_err; /* hipError_t of resolving it */
  /* bumped whenever a device buffer, the table, the species or a tile changes:
   * launch sequences captured in a graph are valid for one generation */
  unsigned gen;
};

int spbte_device_count(int *n) {
  HIPCHK(hipGetDeviceCount(n));
  return SPBTE_OK;
}
int spbte_malloc(void **p, size_t bytes) {
  HIPCHK(hipMalloc(p, bytes ? bytes : 1));
  return SPBTE_OK;
}
int spbte_free(void *p) {
  if (p) HIPCHK(hipFree(p));
  return SPBTE_OK;
}
int spbte_memcpy_htod(void *d, const void *s, size_t b) {
  HIPCHK(hipMemcpy(d, s, b, hipMemcpyHostToDevice));
  return SPBTE_OK;
}
int spbte_memcpy_dtoh(void *d, const void *s, size_t b) {
  HIPCHK(hipMemcpy(d, s, b, hipMemcpyDeviceToHost));
  return SPBTE_OK;
}
int spbte_memcpy_dtod(void *d, const void *s, size_t b) {
  HIPCHK(hipMemcpy(d, s, b, hipMemcpyDeviceToDevice));
  return SPBTE_OK;
}
int spbte_device_copy(void *d_dst, const void *d_src, size_t bytes) {
  LCHK(spbte_launch_copy(d_src, d_dst, bytes, 0));
  HIPCHK(hipStreamSynchronize(0));
  return SPBTE_OK;
}
int spbte_op_synchronize(spbte_op *op) {
  HIPCHK(hipStreamSynchronize(op->s));
  return SPBTE_OK;
}

hipStream_t spbte_op_stream(spbte_op *op) { return op->s; }
const spbte_grid_dev *spbte_op_grid(spbte_op *op) { return &op->g; }
unsigned spbte_op_generation(spbte_op *op) { return op->gen; }
/* Qhat of the last chunk evaluated by spbte_compute_q (ncell * N^3 double2). */
const void *spbte_op_qhat_buffer(spbte_op *op) { return op->d_Qh; }
int spbte_op_chunk_cells(spbte_op *op);
int spbte_op_n(spbte_op *op) { return op->N; }

/* ------------------------------------------------------ transform matrices */

/* Per-axis matrix of the transform of src/collisions.c:232-283 (see
 * k_transform.hip): M[k][j] = exp(i s L_end a_k) exp(-+2 pi i (jk mod N)/N)
 * exp(i s j L_start delta) w_j, accumulated phase in long double. */
static void build_matrix(int N, const double *a, double delta, double Lstart, double Lend, double sign, double dsgn,
                         double *M /* 2*N*N */) {
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++) {
      long double ph = (long double)sign * Lend * a[k] + (long double)dsgn * 2.0L * pi * (long double)((j * k) % N) / N +
                       (long double)sign * j * Lstart * delta;
      double w = (j == 0 || j == N - 1) ? 0.5 : 1.0;
      M[2 * (k * N + j)] = (double)(cosl(ph) * w);
      M[2 * (k * N + j) + 1] = (double)(sinl(ph) * w);
    }
}

/* FFT factors of build_matrix's matrix, M = diag(q) DFT diag(p):
 * p[j] = exp(i s j L_start delta) w_j, q[k] = exp(i s L_end a_k), and the DFT
 * twiddles tw[m] = exp(i dsgn 2 pi m / N), in long double. */
static void build_fft_factors(int N, const double *a, double delta, double Lstart, double Lend, double sign, double dsgn,
                              double *out /* 2 * 3N: p, q, tw */) {
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int j = 0; j < N; j++) {
    const double w = (j == 0 || j == N - 1) ? 0.5 : 1.0;
    const long double ph = (long double)sign * j * Lstart * delta;
    const long double qh = (long double)sign * Lend * a[j];
    const long double th = (long double)dsgn * 2.0L * pi * (long double)j / N;
    out[2 * j] = (double)(cosl(ph) * w);
    out[2 * j + 1] = (double)(sinl(ph) * w);
    out[2 * (N + j)] = (double)cosl(qh);
    out[2 * (N + j) + 1] = (double)sinl(qh);
    out[2 * (2 * N + j)] = (double)cosl(th);
    out[2 * (2 * N + j) + 1] = (double)sinl(th);
  }
}

/* ---------------------------------------------- LU (src/conserve.c:89-168) */

/* Scaled partial pivoting after Atkinson p.520, as the reference: the pivot
 * row is the first row below whose scaled entry exceeds the diagonal's. */
int spbte_host_lu_factor(int n, double *A, int *pivot) {
  double s[SPBTE_MAX_NS + 4];
  for (int i = 0; i < n; i++) {
    s[i] = fabs(A[i * n]);
    for (int j = 0; j < n; j++)
      if (s[i] < fabs(A[i * n + j])) s[i] = fabs(A[i * n + j]);
  }
  for (int k = 0; k < n - 1; k++) {
    double ck = fabs(A[k * n + k] / s[k]);
    int i0 = k, found = 0;
    for (int i = k; i < n; i++)
      if (ck < fabs(A[i * n + k] / s[i])) {
        ck = fabs(A[i * n + k] / s[i]);
        if (!found) { i0 = i; found = 1; }
      }
    pivot[k] = i0;
    if (ck == 0.0) return -1;
    if (i0 != k) {
      for (int j = k; j < n; j++) { double t = A[k * n + j]; A[k * n + j] = A[i0 * n + j]; A[i0 * n + j] = t; }
      double t = s[k]; s[k] = s[i0]; s[i0] = t;
    }
    for (int i = k + 1; i < n; i++) {
      double m = A[i * n + k] / A[k * n + k];
      A[i * n + k] = m;
      for (int j = k + 1; j < n; j++) A[i * n + j] -= m * A[k * n + j];
    }
  }
  pivot[n - 1] = n - 1;
  return 0;
}

/* src/conserve.c:173-202 */
void spbte_host_lu_solve(int n, const double *LU, const int *pivot, double *b) {
  for (int k = 0; k < n - 1; k++) {
    int i = pivot[k];
    if (i != k) { double t = b[i]; b[i] = b[k]; b[k] = t; }
    for (i = k + 1; i < n; i++) b[i] -= LU[i * n + k] * b[k];
  }
  b[n - 1] = b[n - 1] / LU[(n - 1) * n + n - 1];
  for (int i = n - 2; i >= 0; i--) {
    double sum = 0.0;
    for (int j = i + 1; j < n; j++) sum += LU[i * n + j] * b[j];
    b[i] = 1.0 / LU[i * n + i] * (b[i] - sum);
  }
}

/* C C^T of src/conserve.c:268-305 (unfactored), n = Ns + 4. */
void spbte_host_cct(int N, const double *v, int Ns, const double *masses, double *CCt) {
  int n = Ns + 4;
  double dv = v[1] - v[0];
  for (int a = 0; a < n * n; a++) CCt[a] = 0.0;
  for (int m = 0; m < Ns; m++)
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        for (int k = 0; k < N; k++) {
          double wi = (i == 0 || i == N - 1) ? 0.5 : 1.0, wj = (j == 0 || j == N - 1) ? 0.5 : 1.0,
                 wk = (k == 0 || k == N - 1) ? 0.5 : 1.0;
          double p = wi * wj * wk * dv * dv * dv * masses[m];
          double c[SPBTE_MAX_NS + 4];
          for (int q = 0; q < n; q++) c[q] = 0.0;
          c[m] = p;
          c[Ns] = p * v[i];
          c[Ns + 1] = p * v[j];
          c[Ns + 2] = p * v[k];
          c[Ns + 3] = p * 0.5 * (v[i] * v[i] + v[j] * v[j] + v[k] * v[k]);
          for (int a = 0; a < n; a++)
            for (int b = 0; b < n; b++) CCt[a * n + b] += c[a] * c[b];
        }
}

int spbte_op_set_species(spbte_op *op, int Ns, const double *masses) {
  if (Ns < 1 || Ns > SPBTE_MAX_NS) return fail(SPBTE_ERR_ARG, "Ns=%d out of range [1,%d]", Ns, SPBTE_MAX_NS);
  op->Ns = Ns;
  for (int m = 0; m < Ns; m++) op->masses[m] = masses ? masses[m] : 1.0;
  spbte_host_cct(op->N, op->v, Ns, op->masses, op->LU);
  if (spbte_host_lu_factor(Ns + 4, op->LU, op->pivot) != 0) return fail(SPBTE_ERR_SINGULAR, "C C^T is singular");
  op->g.mass = op->masses[0];
  op->g.kb = (op->masses[0] == 1.0) ? 1.0 : 1.380658e-23;
  op->gen++;
  return SPBTE_OK;
}

/* ----------------------------------------------------------- create/destroy */
int spbte_op_create(spbte_op **out, int N, double L_v, const double *v, const double *eta) {
  if (!out) return fail(SPBTE_ERR_ARG, "null output handle");
  *out = NULL;
  if (N < 2 || N > SPBTE_MAX_N || !v || !eta) return fail(SPBTE_ERR_ARG, "bad grid (N=%d)", N);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(SPBTE_ERR_NODEV, "no GPU visible to HIP");
  spbte_op *op = calloc(1, sizeof *op);
  if (!op) return fail(SPBTE_ERR_ARG, "out of host memory");
  op->t = &op->tab[0];
  op->N = N;
  op->n3 = (size_t)N * N * N;
  op->v = malloc(sizeof(double) * N);
  op->eta = malloc(sizeof(double) * N);
  if (!op->v || !op->eta) {
    spbte_op_destroy(op);
    return fail(SPBTE_ERR_ARG, "out of host memory");
  }
  memcpy(op->v, v, sizeof(double) * N);
  memcpy(op->eta, eta, sizeof(double) * N);
  spbte_grid_dev *g = &op->g;
  g->N = N;
  g->Lv = L_v;
  g->hv = v[1] - v[0];
  g->heta = eta[1] - eta[0];
  const double scale3 = pow(1.0 / sqrt(2.0 * M_PI), 3.0); /* src/collisions.c:46 */
  g->pre_fwd = scale3 * g->hv * g->hv * g->hv;
  g->pre_inv = scale3 * g->heta * g->heta * g->heta;
  double *Mf = malloc(sizeof(double) * 2 * N * N), *Mi = malloc(sizeof(double) * 2 * N * N), w[SPBTE_MAX_N];
  if (!Mf || !Mi) {
    free(Mf);
    free(Mi);
    spbte_op_destroy(op);
    return fail(SPBTE_ERR_ARG, "out of host memory");
  }
  const double Leta = -eta[0];
  build_matrix(N, eta, g->hv, Leta, L_v, 1.0, -1.0, Mf);
  build_matrix(N, v, g->heta, L_v, Leta, -1.0, 1.0, Mi);
  for (int i = 0; i < N; i++) w[i] = (i == 0 || i == N - 1) ? 0.5 : 1.0;
  int rc = SPBTE_OK;
  if (hipStreamCreate(&op->s) != hipSuccess || hipMalloc(&g->d_mat_fwd, sizeof(double) * 2 * N * N) != hipSuccess ||
      hipMalloc(&g->d_mat_inv, sizeof(double) * 2 * N * N) != hipSuccess ||
      hipMalloc(&g->d_v, sizeof(double) * N) != hipSuccess || hipMalloc(&g->d_w, sizeof(double) * N) != hipSuccess ||
      hipMemcpy(g->d_mat_fwd, Mf, sizeof(double) * 2 * N * N, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(g->d_mat_inv, Mi, sizeof(double) * 2 * N * N, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(g->d_v, v, sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(g->d_w, w, sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(SPBTE_ERR_HIP, "device setup failed: %s", hipGetErrorString(hipGetLastError()));
  free(Mf);
  free(Mi);
  if (rc == SPBTE_OK && spbte_fft_supported(N)) {
    double ff[2 * 6 * SPBTE_MAX_N];
    build_fft_factors(N, eta, g->hv, Leta, L_v, 1.0, -1.0, ff);
    build_fft_factors(N, v, g->heta, L_v, Leta, -1.0, 1.0, ff + 6 * N);
    if (hipMalloc(&g->d_fft, sizeof(double) * 12 * N) != hipSuccess ||
        hipMemcpy(g->d_fft, ff, sizeof(double) * 12 * N, hipMemcpyHostToDevice) != hipSuccess)
      rc = fail(SPBTE_ERR_HIP, "device setup failed: %s", hipGetErrorString(hipGetLastError()));
  }
  if (rc != SPBTE_OK) {
    spbte_op_destroy(op);
    return rc;
  }
  /* the allocation-free entry point's workspace on this GPU (resolved once per
   * device; a failure is reported by that entry point) */
  op->nac_err = spbte_noalloc_scratch(&op->nac, &op->nac_bytes);
  double one = 1.0;
  rc = spbte_op_set_species(op, 1, &one);
  if (rc != SPBTE_OK) {
    spbte_op_destroy(op);
    return rc;
  }
  *out = op;
  return SPBTE_OK;
}

static void free_weights(spbte_op *op) {
  for (int k = 0; k < SPBTE_MAX_NS * SPBTE_MAX_NS; k++) {
    struct wtable *w = &op->tab[k];
    if (w->dW64) hipFree(w->dW64);
    for (int f = 1; f < NFORMS; f++)
      if (w->dWf[f]) hipFree(w->dWf[f]);
    w->dW64 = NULL;
    for (int f = 0; f < NFORMS; f++) w->dWf[f] = NULL;
    for (int f = 0; f < 3; f++)
      for (int t = 0; t < 9; t++) {
        if (w->dWp[f][t]) hipFree(w->dWp[f][t]);
        w->dWp[f][t] = NULL;
      }
    w->have_w = 0;
  }
}

int spbte_op_select_table(spbte_op *op, int slot) {
  if (!op || slot < 0 || slot >= SPBTE_MAX_NS * SPBTE_MAX_NS)
    return fail(SPBTE_ERR_ARG, "table slot must be 0 to %d", SPBTE_MAX_NS * SPBTE_MAX_NS - 1);
  op->t = &op->tab[slot];
  return SPBTE_OK;
}

int spbte_op_table_slot(spbte_op *op) { return (int)(op->t - op->tab); }

static void free_scratch(spbte_op *op) {
  if (op->d_mom) hipFree(op->d_mom);
  if (op->d_xs) hipFree(op->d_xs);
  if (op->d_FG) hipFree(op->d_FG);
  if (op->d_Qh) hipFree(op->d_Qh);
  if (op->d_part) hipFree(op->d_part);
  op->d_mom = NULL;
  op->d_part = NULL;
  op->d_xs = op->d_FG = op->d_Qh = NULL;
  op->cap = 0;
}

void spbte_op_destroy(spbte_op *op) {
  if (!op) return;
  if (op->s) hipStreamSynchronize(op->s);
  if (op->d_Qp) hipFree(op->d_Qp);
  free_weights(op);
  free_scratch(op);
  if (op->g.d_mat_fwd) hipFree(op->g.d_mat_fwd);
  if (op->g.d_mat_inv) hipFree(op->g.d_mat_inv);
  if (op->g.d_v) hipFree(op->g.d_v);
  if (op->g.d_w) hipFree(op->g.d_w);
  if (op->g.d_fft) hipFree(op->g.d_fft);
  if (op->s) hipStreamDestroy(op->s);
  free(op->v);
  free(op->eta);
  free(op);
}

/* Largest batch evaluated in one pass; bigger batches are processed in chunks. */
static int chunk_cells(const spbte_op *op) {
  size_t per_cell = op->n3 * 16 * (2 * MAXJOBS + 1);
  size_t budget = (size_t)6 << 30; /* 6 GiB of scratch */
  size_t c = budget / per_cell;
  if (c > 4096) c = 4096;
  if (c < 1) c = 1;
  return (int)c;
}

int spbte_op_chunk_cells(spbte_op *op) { return chunk_cells(op); }

static int ensure_scratch(spbte_op *op, int ncell) {
  if (ncell <= op->cap) return SPBTE_OK;
  HIPCHK(hipStreamSynchronize(op->s));
  free_scratch(op);
  size_t n = (size_t)ncell * op->n3;
  HIPCHK(hipMalloc((void **)&op->d_mom, sizeof(double) * 2 * ncell * SPBTE_MOM_STRIDE));
  HIPCHK(hipMalloc(&op->d_xs, 16 * n * MAXJOBS));
  HIPCHK(hipMalloc(&op->d_FG, 16 * n * MAXJOBS));
  HIPCHK(hipMalloc(&op->d_Qh, 16 * n));
  HIPCHK(hipMalloc((void **)&op->d_part, sizeof(double) * 10 * (size_t)op->N * ncell));
  op->cap = ncell;
  op->gen++;
  return SPBTE_OK;
}

/* Buffers one evaluation works in.  The operator's own (grown on demand) for
 * spbte_compute_q; carved from the fixed workspace for the allocation-free
 * entry point, where nothing may be allocated: then the tables the tile streams
 * must be resident and Qp must already hold qp_bytes. */
typedef struct {
  double *mom; /* 2 * ncell * SPBTE_MOM_STRIDE */
  double *part; /* 10 * N * ncell, per-plane partial sums (two sets) */
  void *xs;    /* transform scratch */
  void *FG;    /* transformed operands */
  void *Qh;    /* Qhat, ncell * N^3 double2 */
  void *Qp;    /* partial sums of the split streaming kernel (fixed sets only) */
  size_t qp_bytes;
  int fixed;   /* 1: allocate nothing */
} scratch_set;

static scratch_set own_scratch(const spbte_op *op) {
  scratch_set ss = {op->d_mom, op->d_part, op->d_xs, op->d_FG, op->d_Qh, NULL, 0, 0};
  return ss;
}

/* ------------------------------------------------------------------ weights */
static int alloc_w64(spbte_op *op) {
  op->gen++;
  if (op->t->dW64) {
    HIPCHK(hipStreamSynchronize(op->s));
    for (int f = 1; f < NFORMS; f++)
      if (op->t->dWf[f]) { hipFree(op->t->dWf[f]); op->t->dWf[f] = NULL; }
    for (int f = 0; f < 3; f++)
      for (int t = 0; t < 9; t++)
        if (op->t->dWp[f][t]) { hipFree(op->t->dWp[f][t]); op->t->dWp[f][t] = NULL; }
    return SPBTE_OK;
  }
  HIPCHK(hipMalloc((void **)&op->t->dW64, sizeof(double) * op->n3 * op->n3));
  op->t->dWf[0] = op->t->dW64;
  return SPBTE_OK;
}

int spbte_op_set_weights(spbte_op *op, const double *W) {
  int rc = alloc_w64(op);
  if (rc) return rc;
  HIPCHK(hipMemcpy(op->t->dW64, W, sizeof(double) * op->n3 * op->n3, hipMemcpyHostToDevice));
  op->t->have_w = 1;
  return SPBTE_OK;
}

int spbte_op_set_weights_rows(spbte_op *op, const double *const *rows) {
  int rc = alloc_w64(op);
  if (rc) return rc;
  const size_t n3 = op->n3, stage_rows = (((size_t)64 << 20) / (n3 * 8)) ? ((size_t)64 << 20) / (n3 * 8) : 1;
  double *stage = NULL;
  HIPCHK(hipHostMalloc((void **)&stage, sizeof(double) * n3 * stage_rows, hipHostMallocDefault));
  for (size_t r0 = 0; r0 < n3; r0 += stage_rows) {
    size_t nr = (r0 + stage_rows <= n3) ? stage_rows : n3 - r0;
    for (size_t r = 0; r < nr; r++) memcpy(stage + r * n3, rows[r0 + r], sizeof(double) * n3);
    hipError_t e = hipMemcpy(op->t->dW64 + r0 * n3, stage, sizeof(double) * n3 * nr, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      hipHostFree(stage);
      return fail(SPBTE_ERR_HIP, "weight upload: %s", hipGetErrorString(e));
    }
  }
  hipHostFree(stage);
  op->t->have_w = 1;
  return SPBTE_OK;
}

int spbte_op_load_weights_file(spbte_op *op, const char *path) {
  FILE *fp = fopen(path, "rb");
  if (!fp) return fail(SPBTE_ERR_IO, "cannot open weight file %s", path);
  int rc = alloc_w64(op);
  if (rc) { fclose(fp); return rc; }
  const size_t n3 = op->n3, stage_rows = (((size_t)64 << 20) / (n3 * 8)) ? ((size_t)64 << 20) / (n3 * 8) : 1;
  double *stage = NULL;
  if (hipHostMalloc((void **)&stage, sizeof(double) * n3 * stage_rows, hipHostMallocDefault) != hipSuccess) {
    fclose(fp);
    return fail(SPBTE_ERR_HIP, "pinned staging allocation failed");
  }
  for (size_t r0 = 0; r0 < n3; r0 += stage_rows) {
    size_t nr = (r0 + stage_rows <= n3) ? stage_rows : n3 - r0;
    if (fread(stage, sizeof(double), n3 * nr, fp) != n3 * nr) {
      hipHostFree(stage);
      fclose(fp);
      return fail(SPBTE_ERR_IO, "weight file %s is short", path);
    }
    hipError_t e = hipMemcpy(op->t->dW64 + r0 * n3, stage, sizeof(double) * n3 * nr, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      hipHostFree(stage);
      fclose(fp);
      return fail(SPBTE_ERR_HIP, "weight upload: %s", hipGetErrorString(e));
    }
  }
  hipHostFree(stage);
  fclose(fp);
  op->t->have_w = 1;
  return SPBTE_OK;
}

int spbte_op_generate_weights_iso(spbte_op *op, double lambda, double d_i, double d_j, double m_i, double m_j) {
  int rc = alloc_w64(op);
  if (rc) return rc;
  double *d_eta = NULL;
  HIPCHK(hipMalloc((void **)&d_eta, sizeof(double) * op->N));
  HIPCHK(hipMemcpy(d_eta, op->eta, sizeof(double) * op->N, hipMemcpyHostToDevice));
  const double dfac = 0.25 * pow(0.5 * (d_i + d_j), 2); /* src/weights.c:277 */
  const double mu = m_j / (m_i + m_j);                  /* src/weights.c:190 */
  int e = spbte_launch_weights_iso(&op->g, d_eta, lambda, dfac, mu, op->t->dW64, op->s);
  hipError_t e2 = hipStreamSynchronize(op->s);
  hipFree(d_eta);
  if (e != 0 || e2 != hipSuccess)
    return fail(SPBTE_ERR_HIP, "weight generation: %s", hipGetErrorString(e ? (hipError_t)e : e2));
  op->t->have_w = 1;
  return SPBTE_OK;
}

int spbte_op_generate_weights_landau(spbte_op *op, double lambda) {
  if (!op || !isfinite(lambda)) return fail(SPBTE_ERR_ARG, "bad arguments");
  int rc = alloc_w64(op);
  if (rc) return rc;
  double *d_eta = NULL;
  HIPCHK(hipMalloc((void **)&d_eta, sizeof(double) * op->N));
  HIPCHK(hipMemcpy(d_eta, op->eta, sizeof(double) * op->N, hipMemcpyHostToDevice));
  int e = spbte_launch_weights_landau(&op->g, d_eta, lambda, op->t->dW64, op->s);
  hipError_t e2 = hipStreamSynchronize(op->s);
  hipFree(d_eta);
  if (e != 0 || e2 != hipSuccess)
    return fail(SPBTE_ERR_HIP, "Landau weight generation: %s", hipGetErrorString(e ? (hipError_t)e : e2));
  op->t->have_w = 1;
  return SPBTE_OK;
}

int spbte_op_generate_weights_glance(spbte_op *op, double lambda, double glance) {
  if (!op || !isfinite(lambda) || !isfinite(glance)) return fail(SPBTE_ERR_ARG, "bad arguments");
  if (!(glance > 0.0 && glance < M_PI))
    return fail(SPBTE_ERR_ARG, "glance %g outside (0, pi), where log(sin(glance/2)) < 0", glance);
  int rc = alloc_w64(op);
  if (rc) return rc;
  double *d_eta = NULL;
  HIPCHK(hipMalloc((void **)&d_eta, sizeof(double) * op->N));
  HIPCHK(hipMemcpy(d_eta, op->eta, sizeof(double) * op->N, hipMemcpyHostToDevice));
  int e = spbte_launch_weights_glance(&op->g, op->eta, d_eta, lambda, glance, op->t->dW64, op->s);
  hipError_t e2 = hipStreamSynchronize(op->s);
  hipFree(d_eta);
  if (e != 0 || e2 != hipSuccess)
    return fail(SPBTE_ERR_HIP, "glancing-collision weight generation: %s", hipGetErrorString(e ? (hipError_t)e : e2));
  op->t->have_w = 1;
  return SPBTE_OK;
}

int spbte_op_get_weights(spbte_op *op, double *W) {
  if (!op->t->have_w) return fail(SPBTE_ERR_NOWEIGHTS, "no weight table");
  HIPCHK(hipStreamSynchronize(op->s));
  HIPCHK(hipMemcpy(W, op->t->dW64, sizeof(double) * op->n3 * op->n3, hipMemcpyDeviceToHost));
  return SPBTE_OK;
}

int spbte_op_get_weights_rows(spbte_op *op, double **rows) {
  if (!op->t->have_w) return fail(SPBTE_ERR_NOWEIGHTS, "no weight table");
  HIPCHK(hipStreamSynchronize(op->s));
  for (size_t r = 0; r < op->n3; r++)
    HIPCHK(hipMemcpy(rows[r], op->t->dW64 + r * op->n3, sizeof(double) * op->n3, hipMemcpyDeviceToHost));
  return SPBTE_OK;
}

int spbte_op_store_weights_file(spbte_op *op, const char *path) {
  if (!op || !path) return fail(SPBTE_ERR_ARG, "null argument");
  if (!op->t->have_w) return fail(SPBTE_ERR_NOWEIGHTS, "no weight table");
  HIPCHK(hipStreamSynchronize(op->s));
  const size_t n3 = op->n3, stage_rows = (((size_t)64 << 20) / (n3 * 8)) ? ((size_t)64 << 20) / (n3 * 8) : 1;
  double *stage = NULL;
  if (hipHostMalloc((void **)&stage, sizeof(double) * n3 * stage_rows, hipHostMallocDefault) != hipSuccess)
    return fail(SPBTE_ERR_HIP, "pinned staging allocation failed");
  FILE *fp = fopen(path, "wb");
  int rc = fp ? SPBTE_OK : fail(SPBTE_ERR_IO, "cannot open %s for writing", path);
  for (size_t r0 = 0; r0 < n3 && !rc; r0 += stage_rows) {
    const size_t nr = (r0 + stage_rows <= n3) ? stage_rows : n3 - r0;
    hipError_t e = hipMemcpy(stage, op->t->dW64 + r0 * n3, sizeof(double) * n3 * nr, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(SPBTE_ERR_HIP, "weight download: %s", hipGetErrorString(e));
    else if (fwrite(stage, sizeof(double), n3 * nr, fp) != n3 * nr) rc = fail(SPBTE_ERR_IO, "writing %s failed", path);
  }
  if (fp && fclose(fp) != 0 && !rc) rc = fail(SPBTE_ERR_IO, "writing %s failed", path);
  hipHostFree(stage);
  return rc;
}

/* Storage of the table each form streams: FP64, FP32 (forms 1 and 3) or FP16x2. */
static int storage_of(int form) { return form == SPBTE_FORM_FP32A ? SPBTE_FORM_FP32 : form; }

/* F/G element size of a form: FP32 complex only when the arithmetic is FP32. */
static size_t fg_bytes(int form) { return form == SPBTE_FORM_FP32A ? 8 : 16; }

/* Bytes of one transformed operand of n values: FP64 complex, or FP32 complex
 * followed by the n sums of their parts (the FP32-product sum's a + b). */
static size_t fg_stride(int form, size_t n) {
  return form == SPBTE_FORM_FP32A ? ((12 * n + 255) & ~(size_t)255) : 16 * n;
}

static int weights_form(spbte_op *op, int form, int build, const void **W, const float **rowscale) {
  if (!op->t->have_w) return fail(SPBTE_ERR_NOWEIGHTS, "no weight table set");
  *rowscale = NULL;
  const int st = storage_of(form);
  if (st == SPBTE_FORM_FP64) {
    *W = op->t->dW64;
    return SPBTE_OK;
  }
  const size_t n = op->n3 * op->n3;
  if (!op->t->dWf[st] && !build)
    return fail(SPBTE_ERR_ARG, "the table in form %d is not resident (spbte_compute_q builds it on first use)", form);
  if (!op->t->dWf[st]) {
    size_t bytes = (st == SPBTE_FORM_FP32) ? 4 * n : 4 * n + 4 * op->n3;
    HIPCHK(hipMalloc(&op->t->dWf[st], bytes));
    op->gen++;
    LCHK(spbte_launch_convert_weights(op->t->dW64, op->t->dWf[st], st, op->n3, op->n3, op->s));
  }
  *W = op->t->dWf[st];
  if (st == SPBTE_FORM_FP16X2) *rowscale = (const float *)((const unsigned *)op->t->dWf[st] + n);
  return SPBTE_OK;
}

/* The table permuted for the streaming kernel with tr rows per lane. */
static int weights_permuted(spbte_op *op, int form, int tr, int build, const void **Wp, const float **rowscale) {
  const int st = storage_of(form);
  if (tr < 1 || tr > 8) return fail(SPBTE_ERR_ARG, "bad rows per lane %d", tr);
  const void *W;
  int rc = weights_form(op, form, build, &W, rowscale); /* FP16x2: the row scales come from here */
  if (rc) return rc;
  if (!op->t->dWp[st][tr] && !build)
    return fail(SPBTE_ERR_ARG,
                "the table in form %d permuted for %d rows per lane is not resident (spbte_compute_q builds it on "
                "first use)", form, tr);
  if (!op->t->dWp[st][tr]) {
    const size_t n = op->n3 * op->n3;
    HIPCHK(hipMalloc(&op->t->dWp[st][tr], n * (st == SPBTE_FORM_FP64 ? 8 : 4)));
    op->gen++;
    LCHK(spbte_launch_permute_weights(op->t->dW64, op->t->dWp[st][tr], *rowscale, st, tr, op->n3, op->s));
  }
  *Wp = op->t->dWp[st][tr];
  return SPBTE_OK;
}

/* ------------------------------------------------------------------ tiles */
static int bucket(int ncell) {
  int b = 0;
  while ((1 << b) < ncell && b < 15) b++;
  return b;
}

/* Default tile: enough workgroups to fill 256 CUs, cells per workgroup as the
 * batch allows (each streamed W value then feeds tb cells). */
static int default_tile(int N, int form, int nprod, int ncell, spbte_tile *t) {
  static const spbte_tile c32_1[] = {{4, 4, 4, 0}, {4, 2, 4, 0}, {4, 4, 2, 0}, {4, 2, 2, 0}, {4, 2, 1, 0}, {4, 1, 1, 0}};
  static const spbte_tile c32_1_64[] = {{4, 4, 2, 0}, {4, 2, 2, 0}, {4, 2, 1, 0}, {4, 1, 1, 0}};
  static const spbte_tile c32_2[] = {{4, 4, 2, 0}, {4, 2, 2, 0}, {4, 2, 1, 0}, {4, 1, 1, 0}};
  static const spbte_tile c24_1[] = {{6, 8, 4, 0}, {6, 8, 2, 0}, {6, 8, 1, 0}};
  static const spbte_tile c24_2[] = {{6, 8, 2, 0}, {6, 8, 1, 0}};
  static const spbte_tile c16_1[] = {{4, 8, 4, 0}, {4, 8, 2, 0}, {4, 4, 2, 0}, {4, 4, 1, 0}};
  static const spbte_tile c16_2[] = {{4, 8, 2, 0}, {4, 4, 2, 0}, {4, 4, 1, 0}};
  static const spbte_tile c8[] = {{2, 8, 2, 0}};
  const spbte_tile *c = NULL;
  int nc = 0;
  switch (N) {
    case 32:
      if (nprod == 2) { c = c32_2; nc = 4; }
      else if (form != SPBTE_FORM_FP32A) { c = c32_1_64; nc = 4; }
      else { c = c32_1; nc = 6; }
      break;
    case 24: if (nprod == 2) { c = c24_2; nc = 2; } else { c = c24_1; nc = 3; } break;
    case 16: if (nprod == 2) { c = c16_2; nc = 3; } else { c = c16_1; nc = 4; } break;
    case 8: c = c8; nc = 1; break;
    default: return 0;
  }
  for (int i = 0; i < nc; i++) {
    if (!spbte_wsum_tile_supported(N, form, nprod, &c[i])) continue;
    int ncg = (ncell + c[i].tb - 1) / c[i].tb;
    long wgs = (long)N * (N / c[i].lines) * ncg;
    if (c[i].tb > ncell && i + 1 < nc) continue;
    if (wgs >= 1024 || i + 1 == nc) { *t = c[i]; return 1; }
  }
  for (int i = nc - 1; i >= 0; i--)
    if (spbte_wsum_tile_supported(N, form, nprod, &c[i])) { *t = c[i]; return 1; }
  return 0;
}

/* Workgroups of a streaming tile without the x-plane split. */
static long stream_wgs(int N, int ncell, const spbte_tile *t) {
  return (long)N * N * N / (64L * t->lines * t->tr) * ((ncell + t->tb - 1) / t->tb);
}

/* Streaming kernel defaults (measured on 256-512 cells, MI355X): FP64 weights
 * favour 8 rows per lane for the two-product sum, FP32 weights 4 rows and 4
 * waves; the one-product (ComputeQ) sum at N=32 four cells per workgroup when
 * the batch has them; the FP32-product maxPreserve sum at N=24 takes 3 rows x 3 waves with
 * two problems per lane (31.7 ms for 512 problems against 44.2 ms on the tiled
 * kernel).  The 16-wave tile of the FP32-product maxPreserve sum at N=32 (four
 * cells per workgroup: half the W refetch) is left to the autotuner: on 512
 * problems it ran 5.8% faster than the 4-wave tile on one MI355X and 4.7%
 * slower on another.  Unsupported candidates are skipped.
 * Batches too small to give 2048 workgroups split the summed x-planes into up
 * to 8 slices; at N=16, where one to four cells take one row per lane, into
 * as many as give 2048 workgroups, cutting x-planes into y-line ranges if
 * need be (one cell: 32 slices of half an x-plane, its FP64 table streamed in
 * 41.6 us against 48.3 with 16 whole planes and 99 with 8 slices of 4 rows;
 * 64 slices took 56.8 us, the F staging outweighing the extra waves; with the
 * later kernel, two one-row waves a workgroup over the same 32 slices took
 * 26.8 us against 29.5 with one). */
static int default_stream_tile(int N, int form, int nprod, int ncell, spbte_tile *t) {
  static const spbte_tile s32[] = {{4, 4, 4, 1}, {4, 4, 2, 1}, {4, 4, 1, 1}}, s32m[] = {{4, 4, 1, 1}},
                         s32m64[] = {{8, 2, 1, 1}};
  static const spbte_tile s24[] = {{4, 2, 2, 1}, {4, 2, 1, 1}}, s24m[] = {{4, 2, 1, 1}};
  static const spbte_tile s24ma[] = {{3, 3, 2, 1}, {3, 3, 1, 1}};
  static const spbte_tile s16[] = {{4, 1, 2, 1}, {4, 1, 1, 1}}, s16m[] = {{4, 1, 2, 1}, {4, 1, 1, 1}};
  static const spbte_tile s16one[] = {{1, 2, 1, 1}};
  const spbte_tile *c = NULL;
  int nc = 0;
  if (N == 32) {
    if (nprod == 1) { c = s32; nc = 3; }
    else if (form == SPBTE_FORM_FP64) { c = s32m64; nc = 1; }
    else { c = s32m; nc = 1; }
  } else if (N == 24) {
    if (nprod == 2 && form == SPBTE_FORM_FP32A) { c = s24ma; nc = 2; }
    else { c = nprod == 2 ? s24m : s24; nc = nprod == 2 ? 1 : 2; }
  }
  else if (N == 16) {
    if (ncell <= 4 && spbte_stream_tile_supported(N, form, nprod, 1, 2, 1)) { c = s16one; nc = 1; }
    else { c = nprod == 2 ? s16m : s16; nc = 2; }
  }
  else return 0;
  for (int i = 0; i < nc; i++) {
    if (!spbte_stream_tile_supported(N, form, nprod, c[i].tr, c[i].lines, c[i].tb)) continue;
    if (c[i].tb > ncell && i + 1 < nc) continue;
    *t = c[i];
    const long wgs = stream_wgs(N, ncell, t);
    int k = 1;
    /* N <= 16 counts waves, not workgroups (its one-row tiles run 2 waves a workgroup) */
    while ((N <= 16 ? wgs * k * t->lines : wgs * k) < 2048 && spbte_csplit_valid(N, 2 * k) && (N <= 16 || 2 * k <= 8))
      k *= 2;
    t->kind = k;
    return 1;
  }
  return 0;
}

static int pick_tile(spbte_op *op, int ncell, int form, int nprod, spbte_tile *t) {
  if (op->pin[form][nprod - 1].tr) { *t = op->pin[form][nprod - 1]; return 1; }
  spbte_tile *tt = &op->tuned[form][nprod - 1][bucket(ncell)];
  if (tt->tr) { *t = *tt; return 1; }
  if (default_stream_tile(op->N, form, nprod, ncell, t)) return 1;
  return default_tile(op->N, form, nprod, ncell, t);
}

static int tile_supported(int N, int form, int nprod, const spbte_tile *t) {
  if (t->kind >= 1)
    return spbte_csplit_valid(N, t->kind) && spbte_stream_tile_supported(N, form, nprod, t->tr, t->lines, t->tb);
  return spbte_wsum_tile_supported(N, form, nprod, t);
}


/* Partial sums of the split streaming kernel: csplit * ncell * N^3 double2. */
static int ensure_partials(spbte_op *op, size_t bytes) {
  if (bytes <= op->qp_bytes) return SPBTE_OK;
  HIPCHK(hipStreamSynchronize(op->s));
  if (op->d_Qp) hipFree(op->d_Qp);
  op->d_Qp = NULL;
  op->qp_bytes = 0;
  HIPCHK(hipMalloc(&op->d_Qp, bytes));
  op->qp_bytes = bytes;
  op->gen++;
  return SPBTE_OK;
}

int spbte_set_tile(spbte_op *op, int form, int maxpreserve, int kind, int tr, int lines, int tb) {
  if (form < 0 || form >= NFORMS) return fail(SPBTE_ERR_ARG, "bad form %d", form);
  int np = maxpreserve ? 2 : 1;
  spbte_tile t = {tr, lines, tb, kind};
  if (tr && !tile_supported(op->N, form, np, &t))
    return fail(SPBTE_ERR_UNSUPPORTED, "tile kind %d (%d,%d,%d) not instantiated for N=%d form %d", kind, tr, lines, tb,
                op->N, form);
  op->pin[form][np - 1] = t;
  op->gen++;
  return SPBTE_OK;
}

int spbte_get_tile(spbte_op *op, int ncell, int form, int maxpreserve, int *kind, int *tr, int *lines, int *tb) {
  spbte_tile t = {0, 0, 0, 0};
  if (form < 0 || form >= NFORMS) return fail(SPBTE_ERR_ARG, "bad form %d", form);
  if (!pick_tile(op, ncell, form, maxpreserve ? 2 : 1, &t)) t.tr = t.lines = t.tb = t.kind = 0;
  *kind = t.kind; *tr = t.tr; *lines = t.lines; *tb = t.tb;
  return SPBTE_OK;
}

/* ---------------------------------------------------------- the operator */
/* The weighted sum's result is left at *qsrc as *nsl sets to add: Qh (one set),
 * or, with a->keep_slices, a split tile's unsummed slices in Qp. */
static int launch_tile(spbte_op *op, int form, const spbte_tile *t, const spbte_wsum_args *a, const scratch_set *ss,
                       const void **qsrc, int *nsl) {
  const float *rs;
  *qsrc = a->Qh;
  *nsl = 1;
  if (t && t->kind >= 1) {
    const void *Wp = NULL;
    int rc = weights_permuted(op, form, t->tr, !ss->fixed, &Wp, &rs);
    if (rc) return rc;
    void *Qp = NULL;
    if (t->kind > 1) {
      const size_t need = (size_t)t->kind * a->ncell * op->n3 * 16;
      if (ss->fixed) {
        if (need > ss->qp_bytes) return fail(SPBTE_ERR_ARG, "partial sums need %zu bytes of workspace", need);
        Qp = ss->Qp;
      } else {
        rc = ensure_partials(op, need);
        if (rc) return rc;
        Qp = op->d_Qp;
      }
    }
    LCHK(spbte_launch_wsum_stream(op->N, form, t->tr, t->lines, t->tb, a, Wp, rs, t->kind, Qp, op->s));
    if (a->keep_slices && t->kind > 1) {
      *qsrc = Qp;
      *nsl = t->kind;
    }
    return SPBTE_OK;
  }
  const void *W;
  int rc = weights_form(op, form, !ss->fixed, &W, &rs);
  if (rc) return rc;
  spbte_wsum_args b = *a;
  b.W = W;
  LCHK(spbte_launch_wsum_scaled(op->N, form, t, &b, rs, op->s));
  return SPBTE_OK;
}

/* keep_slices: a split tile may leave its slices unsummed (see launch_tile);
 * qsrc/nsl (NULL unless keep_slices) say where the result is. */
static int run_wsum(spbte_op *op, int form, int nprod, const void *F0, const void *G0, const void *F1, const void *G1,
                    int ncell, const scratch_set *ss, int keep_slices, const void **qsrc, int *nsl) {
  if (!op->t->have_w) return fail(SPBTE_ERR_NOWEIGHTS, "no weight table set");
  spbte_tile t;
  int have = pick_tile(op, ncell, form, nprod, &t);
  if (!have && form != SPBTE_FORM_FP64)
    return fail(SPBTE_ERR_UNSUPPORTED, "no reduced-precision weighted-sum kernel for N=%d", op->N);
  spbte_wsum_args a = {NULL, F0, G0, F1, G1, ss->Qh, ncell, nprod, keep_slices};
  const void *q;
  int n;
  return launch_tile(op, form, have ? &t : NULL, &a, ss, qsrc ? qsrc : &q, nsl ? nsl : &n);
}

/* One chunk of spbte_compute_q.  ev_a/ev_b (optional) bracket the weighted-sum
 * launch so its duration can be timed on the operator's stream. */
static int compute_chunk(spbte_op *op, const double *d_f, const double *d_g, double *d_Q, int ncell, int form, int flags,
                         hipEvent_t ev_a, hipEvent_t ev_b, const scratch_set *ss, const spbte_epilogue *epi) {
  const int self = (flags & SPBTE_SELF) != 0 || d_g == d_f;
  const int mp = (flags & SPBTE_MAXPRESERVE) != 0;
  const size_t n = (size_t)ncell * op->n3;
  const size_t stride = fg_stride(form, n);
  char *FG = (char *)ss->FG;
  void *B[MAXJOBS];
  for (int q = 0; q < MAXJOBS; q++) B[q] = FG + (size_t)q * stride;
  spbte_xform_job jobs[MAXJOBS];
  int nj = 0;
  const void *F0, *G0, *F1 = NULL, *G1 = NULL;
  double *mom_f = ss->mom, *mom_g = ss->mom + (size_t)ncell * SPBTE_MOM_STRIDE;
  /* FFT grids: the forward transform computes the Maxwellians' moments itself
   * from per-plane partial sums (f's, then g's, in ss->part), and the inverse
   * transform writes the projection's partial sums (over f's, dead by then) */
  const int fused = op->g.d_fft != NULL && spbte_fft_supported(op->N);
  /* small grids: for a few cells the forward transform's workgroups take the
   * Maxwellians' moments from the cells themselves (every Maxwellian
   * workgroup reads its whole cell, 1 MB per cell at N=16, against a launch
   * of the partial-sum pass), and the inverse transform's last pass runs over
   * whole cubes and projects in the same launch */
  const int mom_inl = fused && spbte_fft_moments_inline(op->N) && ncell <= 32;
  const int cube_i = fused && spbte_fft_cube(op->N, ncell);
  const double *part_f = fused && !mom_inl ? ss->part : NULL;
  const double *part_g = fused && !mom_inl ? ss->part + (size_t)5 * op->N * ncell : NULL;
  const double *msrc_f = mom_inl ? d_f : NULL, *msrc_g = mom_inl ? d_g : NULL;
  if (!mp) {
    jobs[nj++] = (spbte_xform_job){d_f, NULL, 0, B[0], NULL, NULL};
    F0 = B[0];
    if (self) G0 = B[0];
    else { jobs[nj++] = (spbte_xform_job){d_g, NULL, 0, B[1], NULL, NULL}; G0 = B[1]; }
  } else {
    /* Q = sum W [ T(g - M_f)[c] T(f)[s] + T(M_g)[c] T(f - M_f)[s] ] */
    if (fused) {
      if (!mom_inl) {
        LCHK(spbte_launch_moment_partials(&op->g, d_f, ncell, (double *)part_f, op->s));
        if (!self) LCHK(spbte_launch_moment_partials(&op->g, d_g, ncell, (double *)part_g, op->s));
      }
      mom_f = mom_g = NULL;
    } else {
      LCHK(spbte_launch_moments(&op->g, d_f, mom_f, ncell, ss->part, op->s));
      if (!self) LCHK(spbte_launch_moments(&op->g, d_g, mom_g, ncell, ss->part, op->s));
    }
    jobs[nj++] = (spbte_xform_job){d_f, NULL, 0, B[0], NULL, NULL};           /* F0 = T(f)        */
    jobs[nj++] = (spbte_xform_job){d_f, mom_f, 1, B[1], part_f, msrc_f};      /* F1 = T(f - M_f)  */
    F0 = B[0];
    F1 = B[1];
    if (self) {
      G0 = B[1];
      jobs[nj++] = (spbte_xform_job){d_f, mom_f, 2, B[2], part_f, msrc_f};    /* G1 = T(M_f)      */
      G1 = B[2];
    } else {
      jobs[nj++] = (spbte_xform_job){d_g, mom_f, 1, B[2], part_f, msrc_f};    /* G0 = T(g - M_f)  */
      jobs[nj++] = (spbte_xform_job){d_g, mom_g, 2, B[3], part_g, msrc_g};    /* G1 = T(M_g)      */
      G0 = B[2];
      G1 = B[3];
    }
  }
  LCHK(spbte_launch_xform_fwd(&op->g, jobs, nj, ncell, fg_bytes(form) == 8, ss->xs, op->s));
  if (ev_a) HIPCHK(hipEventRecord(ev_a, op->s));
  /* FFT grids: the inverse transform adds the split sum's slices itself */
  const void *qsrc;
  int nsl;
  int rc = run_wsum(op, form, mp ? 2 : 1, F0, G0, F1, G1, ncell, ss, fused, &qsrc, &nsl);
  if (ev_b) HIPCHK(hipEventRecord(ev_b, op->s));
  if (rc) return rc;
  const int cons = (flags & SPBTE_CONSERVE) != 0;
  if (cons && op->Ns != 1) return fail(SPBTE_ERR_UNSUPPORTED, "per-cell projection needs one species (Ns=%d)", op->Ns);
  if (cube_i) {
    LCHK(spbte_launch_xform_inv_real_cube(&op->g, qsrc, nsl, d_Q, ncell, ss->xs, cons ? op->masses : NULL, op->LU,
                                          op->pivot, epi, op->s));
    return SPBTE_OK;
  }
  if (epi) return fail(SPBTE_ERR_UNSUPPORTED, "no fused update on this grid and batch");
  LCHK(spbte_launch_xform_inv_real(&op->g, qsrc, nsl, d_Q, ncell, ss->xs, cons && fused ? ss->part : NULL, op->s));
  if (cons && fused) LCHK(spbte_launch_project_apply(&op->g, op->masses, op->LU, op->pivot, d_Q, ncell, ss->part, op->s));
  else if (cons) LCHK(spbte_launch_project(&op->g, 1, op->masses, op->LU, op->pivot, d_Q, ncell, ss->part, op->s));
  return SPBTE_OK;
}

int spbte_compute_q(spbte_op *op, const double *d_f, const double *d_g, double *d_Q, int ncell, int form, int flags) {
  if (!op || ncell < 0 || form < 0 || form >= NFORMS) return fail(SPBTE_ERR_ARG, "bad arguments");
  if (ncell == 0) return SPBTE_OK;
  if (!d_f || !d_Q) return fail(SPBTE_ERR_ARG, "null buffer");
  if (!op->t->have_w) return fail(SPBTE_ERR_NOWEIGHTS, "no weight table set");
  const int ch = chunk_cells(op);
  int rc = ensure_scratch(op, ncell < ch ? ncell : ch);
  if (rc) return rc;
  for (int c0 = 0; c0 < ncell; c0 += ch) {
    int nc = (ncell - c0 < ch) ? ncell - c0 : ch;
    const size_t off = (size_t)c0 * op->n3;
    const scratch_set ss = own_scratch(op);
    rc = compute_chunk(op, d_f + off, d_g ? d_g + off : NULL, d_Q + off, nc, form, flags | (d_g ? 0 : SPBTE_SELF), NULL,
                       NULL, &ss, NULL);
    if (rc) return rc;
  }
  return SPBTE_OK;
}

int spbte_compute_q_update(spbte_op *op, const double *d_f, double *d_Q, int ncell, int form, int flags,
                           const spbte_epilogue *epi, int *fused) {
  if (!op || !epi || !fused || ncell < 1 || form < 0 || form >= NFORMS || !d_f || !d_Q)
    return fail(SPBTE_ERR_ARG, "bad arguments");
  *fused = 0;
  if (!op->t->have_w) return fail(SPBTE_ERR_NOWEIGHTS, "no weight table set");
  if (ncell > chunk_cells(op) || !op->g.d_fft || !spbte_fft_supported(op->N) || !spbte_fft_cube(op->N, ncell) ||
      ((flags & SPBTE_CONSERVE) && op->Ns != 1))
    return spbte_compute_q(op, d_f, NULL, d_Q, ncell, form, flags);
  int rc = ensure_scratch(op, ncell);
  if (rc) return rc;
  const scratch_set ss = own_scratch(op);
  rc = compute_chunk(op, d_f, NULL, d_Q, ncell, form, flags | SPBTE_SELF, NULL, NULL, &ss, epi);
  if (!rc) *fused = 1;
  return rc;
}

/* Bytes rounded up to the workspace's 256-byte alignment. */
static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

int spbte_compute_q_noalloc(spbte_op *op, const double *d_f, const double *d_g, double *d_Q, int ncell, int form,
                            int flags) {
  if (!op || ncell < 0 || form < 0 || form >= NFORMS) return fail(SPBTE_ERR_ARG, "bad arguments");
  if (ncell == 0) return SPBTE_OK;
  if (!d_f || !d_Q) return fail(SPBTE_ERR_ARG, "null buffer");
  if (!op->t->have_w) return fail(SPBTE_ERR_NOWEIGHTS, "no weight table set");
  if (!op->nac)
    return fail(SPBTE_ERR_HIP, "no fixed workspace on this GPU: %s", hipGetErrorString((hipError_t)op->nac_err));
  if ((flags & SPBTE_CONSERVE) && op->Ns != 1)
    return fail(SPBTE_ERR_UNSUPPORTED, "per-cell projection needs one species (Ns=%d)", op->Ns);
  if (!d_g) flags |= SPBTE_SELF;
  const int self = (flags & SPBTE_SELF) != 0 || d_g == d_f;
  const int mp = (flags & SPBTE_MAXPRESERVE) != 0;
  /* the tile compute_chunk will pick, and its table, which must be resident */
  spbte_tile t;
  const int have = pick_tile(op, ncell, form, mp ? 2 : 1, &t);
  if (!have && form != SPBTE_FORM_FP64)
    return fail(SPBTE_ERR_UNSUPPORTED, "no reduced-precision weighted-sum kernel for N=%d", op->N);
  const void *W;
  const float *rs;
  int rc = (have && t.kind >= 1) ? weights_permuted(op, form, t.tr, 0, &W, &rs) : weights_form(op, form, 0, &W, &rs);
  if (rc) return rc;
  /* workspace: moments | partial sums | transformed operands | Qhat | transform scratch,
   * which also holds the split sum's partial sums (they are dead by the time
   * the inverse transform writes it) */
  const int nj = mp ? (self ? 3 : 4) : (self ? 1 : 2);
  const size_t n = (size_t)ncell * op->n3;
  const size_t b_mom = al256(sizeof(double) * 2 * (size_t)ncell * SPBTE_MOM_STRIDE);
  const size_t b_part = al256(sizeof(double) * 10 * (size_t)op->N * ncell);
  const size_t b_fg = al256((size_t)nj * fg_stride(form, n)), b_qh = al256(16 * n);
  const size_t b_xs = 16 * n * (size_t)nj, b_qp = (have && t.kind > 1) ? 16 * n * (size_t)t.kind : 0;
  const size_t b_x = al256(b_xs > b_qp ? b_xs : b_qp);
  const size_t need = b_mom + b_part + b_fg + b_qh + b_x;
  if (need > op->nac_bytes)
    return fail(SPBTE_ERR_ARG, "a batch of %d cells at N=%d needs %zu bytes of workspace, more than the %zu of the "
                "fixed workspace (SPBTE_NOALLOC_SCRATCH_BYTES)", ncell, op->N, need, op->nac_bytes);
  char *base = (char *)op->nac;
  scratch_set ss;
  ss.mom = (double *)base;
  ss.part = (double *)(base + b_mom);
  base += b_mom + b_part;
  ss.FG = base;
  ss.Qh = base + b_fg;
  ss.xs = base + b_fg + b_qh;
  ss.Qp = ss.xs;
  ss.qp_bytes = b_x;
  ss.fixed = 1;
  return compute_chunk(op, d_f, d_g, d_Q, ncell, form, flags, NULL, NULL, &ss, NULL);
}

/* The operator's partial-sum scratch when it holds the split projection's
 * (Ns+4) N doubles per group, else NULL (one workgroup per group). */
static double *proj_part(const spbte_op *op, int Ns, int ngroups) {
  return op->d_part && (size_t)(Ns + 4) * ngroups <= (size_t)10 * op->cap ? op->d_part : NULL;
}

int spbte_conserve(spbte_op *op, double *d_Q, int ncell) {
  if (!op || ncell < 0) return fail(SPBTE_ERR_ARG, "bad arguments");
  if (ncell > 0 && !d_Q) return fail(SPBTE_ERR_ARG, "null buffer");
  if (op->Ns != 1) return fail(SPBTE_ERR_UNSUPPORTED, "per-cell projection needs one species (Ns=%d)", op->Ns);
  if (ncell <= 0) return SPBTE_OK;
  LCHK(spbte_launch_project(&op->g, 1, op->masses, op->LU, op->pivot, d_Q, ncell, proj_part(op, 1, ncell), op->s));
  return SPBTE_OK;
}

int spbte_conserve_pairs(spbte_op *op, double *d_Q, int ncell) {
  if (!op || !d_Q || ncell < 0) return fail(SPBTE_ERR_ARG, "bad arguments");
  if (ncell == 0) return SPBTE_OK;
  LCHK(spbte_launch_project_strided(&op->g, op->Ns, op->masses, op->LU, op->pivot, d_Q, ncell, op->n3,
                                    (size_t)ncell * op->n3, proj_part(op, op->Ns, ncell), op->s));
  return SPBTE_OK;
}

int spbte_op_get_species(spbte_op *op, int *Ns, double *masses) {
  if (!op || !Ns) return fail(SPBTE_ERR_ARG, "null argument");
  *Ns = op->Ns;
  if (masses)
    for (int m = 0; m < op->Ns; m++) masses[m] = op->masses[m];
  return SPBTE_OK;
}

int spbte_conserve_species(spbte_op *op, double *d_Q, int ngroups) {
  if (!op || ngroups < 0) return fail(SPBTE_ERR_ARG, "bad arguments");
  if (ngroups == 0) return SPBTE_OK;
  if (!d_Q) return fail(SPBTE_ERR_ARG, "null buffer");
  LCHK(spbte_launch_project(&op->g, op->Ns, op->masses, op->LU, op->pivot, d_Q, ngroups, proj_part(op, op->Ns, ngroups),
                            op->s));
  return SPBTE_OK;
}

int spbte_transform(spbte_op *op, const double *d_in, double *d_out, int ncell, int invert) {
  if (!op || ncell < 0) return fail(SPBTE_ERR_ARG, "bad arguments");
  if (ncell == 0) return SPBTE_OK;
  if (!d_in || !d_out) return fail(SPBTE_ERR_ARG, "null buffer");
  const int ch = chunk_cells(op);
  int rc = ensure_scratch(op, ncell < ch ? ncell : ch);
  if (rc) return rc;
  for (int c0 = 0; c0 < ncell; c0 += ch) {
    int nc = (ncell - c0 < ch) ? ncell - c0 : ch;
    const size_t off = (size_t)c0 * op->n3 * 2;
    LCHK(spbte_launch_xform_fwd_complex(&op->g, d_in + off, d_out + off, nc, invert, op->d_xs, op->s));
  }
  return SPBTE_OK;
}

int spbte_qhat(spbte_op *op, const double *d_Fhat, const double *d_Ghat, double *d_Qhat, int ncell, int form) {
  if (!op || ncell < 0) return fail(SPBTE_ERR_ARG, "bad arguments");
  if (form < 0 || form >= NFORMS) return fail(SPBTE_ERR_ARG, "bad form %d", form);
  if (ncell == 0) return SPBTE_OK;
  if (!d_Fhat || !d_Ghat || !d_Qhat) return fail(SPBTE_ERR_ARG, "null buffer");
  if (!op->t->have_w) return fail(SPBTE_ERR_NOWEIGHTS, "no weight table set");
  int rc = ensure_scratch(op, ncell);
  if (rc) return rc;
  const void *F = d_Fhat, *G = d_Ghat;
  const size_t n = (size_t)ncell * op->n3;
  if (fg_bytes(form) == 8) {
    char *FG = (char *)op->d_FG;
    const size_t stride = fg_stride(form, n);
    LCHK(spbte_launch_c64_to_c32(d_Fhat, FG, n, op->s));
    LCHK(spbte_launch_c64_to_c32(d_Ghat, FG + stride, n, op->s));
    F = FG;
    G = FG + stride;
  }
  const scratch_set ss = own_scratch(op);
  rc = run_wsum(op, form, 1, F, G, NULL, NULL, ncell, &ss, 0, NULL, NULL);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(d_Qhat, op->d_Qh, 16 * n, hipMemcpyDeviceToDevice, op->s));
  return SPBTE_OK;
}

/* ------------------------------------------------------------ timing/tuning */
int spbte_time_compute_q(spbte_op *op, const double *d_f, const double *d_g, double *d_Q, int ncell, int form, int flags,
                         int iters, double *ms_per_call, double *ms_wsum) {
  if (!op || !d_f || !d_Q || !ms_per_call || !ms_wsum || ncell < 1 || form < 0 || form >= NFORMS)
    return fail(SPBTE_ERR_ARG, "bad arguments");
  if (iters < 1 || iters > 100000) return fail(SPBTE_ERR_ARG, "iters must be 1 to 100000");
  const int ch = chunk_cells(op);
  if (ncell > ch) return fail(SPBTE_ERR_ARG, "timed batch larger than one chunk (%d cells)", ch);
  int rc = ensure_scratch(op, ncell);
  if (rc) return rc;
  /* events: [0] start, [1] end, then a begin/end pair around each weighted sum */
  const int nev = 2 + 2 * iters;
  hipEvent_t *ev = calloc((size_t)nev, sizeof(hipEvent_t));
  if (!ev) return fail(SPBTE_ERR_ARG, "out of host memory");
  int made = 0;
  for (; made < nev; made++)
    if (hipEventCreate(&ev[made]) != hipSuccess) {
      rc = fail(SPBTE_ERR_HIP, "hipEventCreate failed");
      break;
    }
  float tot = 0.f, k = 0.f, ksum = 0.f;
  if (!rc && (hipStreamSynchronize(op->s) != hipSuccess || hipEventRecord(ev[0], op->s) != hipSuccess))
    rc = fail(SPBTE_ERR_HIP, "timing set-up failed");
  const scratch_set ss = own_scratch(op);
  for (int i = 0; i < iters && !rc; i++)
    rc = compute_chunk(op, d_f, d_g, d_Q, ncell, form, flags | (d_g ? 0 : SPBTE_SELF), ev[2 + 2 * i], ev[3 + 2 * i],
                       &ss, NULL);
  if (!rc && (hipEventRecord(ev[1], op->s) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
              hipEventElapsedTime(&tot, ev[0], ev[1]) != hipSuccess))
    rc = fail(SPBTE_ERR_HIP, "timing failed: %s", hipGetErrorString(hipGetLastError()));
  for (int i = 0; i < iters && !rc; i++) {
    if (hipEventElapsedTime(&k, ev[2 + 2 * i], ev[3 + 2 * i]) != hipSuccess)
      rc = fail(SPBTE_ERR_HIP, "timing failed: %s", hipGetErrorString(hipGetLastError()));
    ksum += k;
  }
  for (int i = 0; i < made; i++) hipEventDestroy(ev[i]);
  free(ev);
  if (rc) return rc;
  *ms_per_call = tot / iters;
  *ms_wsum = ksum / iters;
  return SPBTE_OK;
}

/* Candidate tiles for the autotuner: kind 0 from k_wsum.hip's table, kind 1
 * from k_wsum_stream.hip's (unsupported ones are skipped). */
static const spbte_tile k_cand[] = {
    {4, 1, 1, 0}, {4, 2, 1, 0}, {4, 2, 2, 0}, {4, 4, 2, 0}, {4, 2, 4, 0}, {4, 4, 4, 0}, {8, 4, 2, 0}, {8, 8, 2, 0},
    {8, 4, 4, 0}, {6, 8, 1, 0}, {6, 8, 2, 0}, {6, 8, 4, 0}, {4, 8, 2, 0}, {4, 8, 4, 0}, {4, 4, 1, 0}, {2, 8, 2, 0},
    {8, 2, 1, 1}, {8, 2, 2, 1}, {4, 4, 2, 1}, {4, 2, 2, 1}, {4, 4, 4, 1}, {4, 4, 1, 1}, {6, 3, 1, 1}, {3, 3, 1, 1},
    {3, 3, 2, 1}, {4, 2, 1, 1}, {8, 1, 1, 1}, {4, 1, 1, 1}, {4, 1, 2, 1}, {4, 9, 1, 1}, {4, 16, 1, 1},
    /* streaming with the summed x-planes split in 2-8 slices, for small batches */
    {4, 4, 1, 2}, {4, 4, 1, 4}, {4, 4, 1, 8}, {4, 4, 2, 2}, {4, 4, 2, 4}, {4, 4, 2, 8}, {4, 2, 1, 4}, {4, 2, 1, 8},
    {4, 2, 2, 4}, {8, 2, 1, 4}, {8, 2, 1, 8}, {4, 2, 1, 2}, {4, 9, 1, 2}, {4, 9, 1, 4},
    /* one or a few cells at N=16: narrower rows, up to one x-plane per slice */
    {4, 1, 1, 8}, {4, 1, 1, 16}, {2, 1, 1, 8}, {2, 1, 1, 16}, {1, 1, 1, 8}, {1, 1, 1, 16}, {8, 2, 1, 16},
    /* and x-planes cut into y-line ranges */
    {1, 1, 1, 32}, {1, 1, 1, 64}, {2, 1, 1, 32}, {2, 1, 1, 64}, {4, 1, 1, 64}, {4, 1, 2, 64},
    {1, 2, 1, 32}, {1, 2, 1, 64}, {1, 4, 1, 64}, {1, 4, 1, 128}};

int spbte_autotune(spbte_op *op, int ncell, int form, int maxpreserve, int *kind, int *tr, int *lines, int *tb) {
  if (form < 0 || form >= NFORMS || ncell < 1) return fail(SPBTE_ERR_ARG, "bad arguments");
  if (!op->t->have_w) return fail(SPBTE_ERR_NOWEIGHTS, "no weight table set");
  const int np = maxpreserve ? 2 : 1;
  int rc = ensure_scratch(op, ncell);
  if (rc) return rc;
  /* inputs: a fixed pseudo-random pattern, not zeros: under a power-limited
   * FP64 load the clock depends on the operands, and zeros ranked the tiles
   * differently from real transforms (12% apart on a 256-cell shock batch) */
  const size_t n = (size_t)ncell * op->n3;
  const int single = fg_bytes(form) == 8;
  const size_t stride = fg_stride(form, n);
  char *FG = (char *)op->d_FG;
  LCHK(spbte_launch_fill_pattern(FG, 4 * stride / (single ? 4 : 8), single, op->s));
  spbte_wsum_args args = {NULL, FG, FG + stride, np == 2 ? FG + 2 * stride : NULL, np == 2 ? FG + 3 * stride : NULL,
                          op->d_Qh, ncell, np, 0};
  const scratch_set ss = own_scratch(op);
  spbte_tile best = {0, 0, 0, 0};
  float best_ms = 1e30f;
  const void *qsrc;
  int nsl;
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  for (size_t i = 0; i < sizeof k_cand / sizeof k_cand[0]; i++) {
    const spbte_tile *t = &k_cand[i];
    if (!tile_supported(op->N, form, np, t)) continue;
    if (t->tb > 2 * ncell) continue;
    if (t->kind > 1 && stream_wgs(op->N, ncell, t) >= 2048) continue; /* the chip is full without a split */
    /* warm-up (builds the table layout it needs); a tile whose warm-up alone
     * takes twice the best time so far is not timed further */
    HIPCHK(hipEventRecord(a, op->s));
    rc = launch_tile(op, form, t, &args, &ss, &qsrc, &nsl);
    HIPCHK(hipEventRecord(b, op->s));
    HIPCHK(hipEventSynchronize(b));
    if (rc) break;
    float warm = 0.f;
    HIPCHK(hipEventElapsedTime(&warm, a, b));
    if (warm > 2.0f * best_ms) continue;
    HIPCHK(hipEventRecord(a, op->s));
    const int reps = 2;
    for (int r = 0; r < reps && rc == SPBTE_OK; r++) rc = launch_tile(op, form, t, &args, &ss, &qsrc, &nsl);
    HIPCHK(hipEventRecord(b, op->s));
    HIPCHK(hipEventSynchronize(b));
    if (rc) break;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    ms /= reps;
    if (ms < best_ms) { best_ms = ms; best = *t; }
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (rc) return rc;
  if (!best.tr) return fail(SPBTE_ERR_UNSUPPORTED, "no tiled kernel for N=%d", op->N);
  op->tuned[form][np - 1][bucket(ncell)] = best;
  op->gen++;
  *kind = best.kind; *tr = best.tr; *lines = best.lines; *tb = best.tb;
  return SPBTE_OK;
}
