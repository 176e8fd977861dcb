/* spbte_net.c -- the learned collision surrogate on one GPU (C99 host side).
 *
 * No counterpart in the reference.  A spectral mixing network maps the kept
 * block of Kb^3 lowest transform modes of a cell (Kb even, about eta = 0) to
 * the same block of Qhat:
 *
 *   features   x_m = (T(f - M_f), T(M_f))[m] / rho            (4 reals per mode)
 *   lift       H0[m][b][:] = x_m P + E[m] + pb                 (C channels)
 *   layer l    Hm = H + U_l (V_l^T H)                          (low-rank mode mixing)
 *              H' = Hm + GELU(Hm W1_l + b1_l) W2_l + b2_l      (channel MLP, 2C hidden)
 *   head       y_m = H_L[m][b][:] O + ob                        (Re, Im of Qhat / rho^2)
 *
 * with the collision operator's own scalings: Q(rho f) = rho^2 Q(f).  The
 * prediction is zero-padded to N^3 modes, inverse-transformed with the
 * operator's transform and projected with the FP64 conservation projection.
 * Activations are FP32 in HBM in mode-major layout [m][b][c], so every product
 * is one plain GEMM (rows m*B + b for the channel products, K = m for the mode
 * mixing); GEMM operands are rounded to BF16 and accumulated in FP32 on MFMA.
 * Targets are the exact Qhat of the operator (ComputeQ or the fused
 * maxPreserve sum), loss = mean over cells of |y - t|^2 / |t|^2.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spbte.h"
#include "spbte_internal.h"

#define HIPCHK(expr)                                                                                        \
  do {                                                                                                      \
    hipError_t e_ = (expr);                                                                                 \
    if (e_ != hipSuccess) return spbte_fail(SPBTE_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                            __FILE__, __LINE__);                                            \
  } while (0)
#define LCHK(expr)                                                                                             \
  do {                                                                                                         \
    int e_ = (expr);                                                                                           \
    if (e_ != 0) return spbte_fail(SPBTE_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString((hipError_t)e_), \
                                   __FILE__, __LINE__);                                                        \
  } while (0)
#define CHK(expr)           \
  do {                      \
    int r_ = (expr);        \
    if (r_ != 0) return r_; \
  } while (0)

#define MAXL 8

typedef struct {
  long U, V, W1, b1, W2, b2;
} layer_off;

struct spbte_net {
  spbte_op *op;
  hipStream_t s;
  int N, Kb, C, R, L, maxpreserve;
  long M;
  double fscale, tscale;
  float loss_eps; /* added to |t|^2 in each cell's loss denominator */
  /* parameters: one FP32 vector, with gradient and Adam moments */
  long npar, oP, opb, oE, oO, oob;
  layer_off lo[MAXL];
  float *par, *grad, *m1, *m2;
  int step;
  /* activations for up to cap cells */
  int cap;
  float *X, *T, *Y, *dY, *loss;
  float *H[MAXL + 1], *Hm[MAXL], *S[MAXL], *Z[MAXL], *G[MAXL];
  float *dH, *dHm, *dZ, *dG, *dS;
  float *work;
  size_t work_floats;
  double *mom;
  void *Fd, *Fm, *xs, *Qh;
  /* surrogate-assisted collisions: sampled cells */
  int scap, ccap;
  double *fs, *Qs, *Qn, *err;
  int *d_cells;
  /* dataset (sample-major) */
  long ds_cap, ds_n;
  float *dsX, *dsT;
  int *d_idx;
};

static uint64_t rng_next(uint64_t *s) {
  *s ^= *s << 13;
  *s ^= *s >> 7;
  *s ^= *s << 17;
  return *s;
}
static double rng_normal(uint64_t *s) {
  double u1 = ((rng_next(s) >> 11) + 1.0) / 9007199254740994.0, u2 = (rng_next(s) >> 11) / 9007199254740992.0;
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

static void free_acts(spbte_net *n) {
  float **bufs[] = {&n->X, &n->T, &n->Y, &n->dY, &n->loss, &n->dH, &n->dHm, &n->dZ, &n->dG, &n->dS};
  for (size_t i = 0; i < sizeof bufs / sizeof bufs[0]; i++)
    if (*bufs[i]) { hipFree(*bufs[i]); *bufs[i] = NULL; }
  for (int l = 0; l <= MAXL; l++)
    if (n->H[l]) { hipFree(n->H[l]); n->H[l] = NULL; }
  for (int l = 0; l < MAXL; l++) {
    float **b[] = {&n->Hm[l], &n->S[l], &n->Z[l], &n->G[l]};
    for (int i = 0; i < 4; i++)
      if (*b[i]) { hipFree(*b[i]); *b[i] = NULL; }
  }
  if (n->mom) hipFree(n->mom);
  if (n->Fd) hipFree(n->Fd);
  if (n->Fm) hipFree(n->Fm);
  if (n->xs) hipFree(n->xs);
  if (n->Qh) hipFree(n->Qh);
  if (n->d_idx) hipFree(n->d_idx);
  n->mom = NULL;
  n->Fd = n->Fm = n->xs = n->Qh = NULL;
  n->d_idx = NULL;
  n->cap = 0;
}

static int falloc(float **p, size_t n) {
  HIPCHK(hipMalloc((void **)p, sizeof(float) * (n ? n : 1)));
  return SPBTE_OK;
}

static int ensure_cap(spbte_net *n, int ncell) {
  if (ncell <= n->cap) return SPBTE_OK;
  HIPCHK(hipStreamSynchronize(n->s));
  free_acts(n);
  const size_t R = (size_t)n->M * ncell, C = n->C, n3 = (size_t)n->N * n->N * n->N;
  CHK(falloc(&n->X, R * 4));
  CHK(falloc(&n->T, R * 2));
  CHK(falloc(&n->Y, R * 2));
  CHK(falloc(&n->dY, R * 2));
  CHK(falloc(&n->loss, ncell));
  for (int l = 0; l <= n->L; l++) CHK(falloc(&n->H[l], R * C));
  for (int l = 0; l < n->L; l++) {
    CHK(falloc(&n->Hm[l], R * C));
    CHK(falloc(&n->S[l], (size_t)n->R * ncell * C));
    CHK(falloc(&n->Z[l], R * 2 * C));
    CHK(falloc(&n->G[l], R * 2 * C));
  }
  CHK(falloc(&n->dH, R * C));
  CHK(falloc(&n->dHm, R * C));
  CHK(falloc(&n->dZ, R * 2 * C));
  CHK(falloc(&n->dG, R * 2 * C));
  CHK(falloc(&n->dS, (size_t)n->R * ncell * C));
  HIPCHK(hipMalloc((void **)&n->mom, sizeof(double) * SPBTE_MOM_STRIDE * ncell));
  HIPCHK(hipMalloc(&n->Fd, 16 * n3 * ncell));
  HIPCHK(hipMalloc(&n->Fm, 16 * n3 * ncell));
  HIPCHK(hipMalloc(&n->xs, 16 * n3 * ncell * 2));
  HIPCHK(hipMalloc(&n->Qh, 16 * n3 * ncell));
  HIPCHK(hipMalloc((void **)&n->d_idx, sizeof(int) * ncell));
  n->cap = ncell;
  return SPBTE_OK;
}

void spbte_net_destroy(spbte_net *n) {
  if (!n) return;
  hipStreamSynchronize(n->s);
  free_acts(n);
  float *b[] = {n->par, n->grad, n->m1, n->m2, n->work, n->dsX, n->dsT};
  for (size_t i = 0; i < sizeof b / sizeof b[0]; i++)
    if (b[i]) hipFree(b[i]);
  double *d[] = {n->fs, n->Qs, n->Qn, n->err};
  for (size_t i = 0; i < sizeof d / sizeof d[0]; i++)
    if (d[i]) hipFree(d[i]);
  if (n->d_cells) hipFree(n->d_cells);
  free(n);
}

int spbte_net_create(spbte_net **out, spbte_op *op, int Kb, int width, int rank, int layers, int maxpreserve,
                     unsigned long long seed) {
  if (!out || !op) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  const int N = spbte_op_n(op);
  if (Kb < 2 || Kb > N || (Kb & 1)) return spbte_fail(SPBTE_ERR_ARG, "mode block %d must be even and <= N=%d", Kb, N);
  if (width < 8 || width > 256 || 256 % width) return spbte_fail(SPBTE_ERR_ARG, "width %d must divide 256", width);
  if (rank < 1 || layers < 1 || layers > MAXL) return spbte_fail(SPBTE_ERR_ARG, "bad rank or layer count");
  spbte_net *n = calloc(1, sizeof *n);
  if (!n) return spbte_fail(SPBTE_ERR_ARG, "out of host memory");
  n->op = op;
  n->s = spbte_op_stream(op);
  n->N = N;
  n->Kb = Kb;
  n->M = (long)Kb * Kb * Kb;
  n->C = width;
  n->R = rank;
  n->L = layers;
  n->maxpreserve = maxpreserve != 0;
  n->fscale = n->tscale = 1.0;
  n->loss_eps = (float)(0.02 * (double)n->M); /* per-value RMS floor 0.1 in target units */
  const long M = n->M, C = width, r = rank;
  long o = 0;
  n->oP = o; o += 4 * C;
  n->opb = o; o += C;
  n->oE = o; o += M * C;
  for (int l = 0; l < layers; l++) {
    n->lo[l].U = o; o += M * r;
    n->lo[l].V = o; o += M * r;
    n->lo[l].W1 = o; o += C * 2 * C;
    n->lo[l].b1 = o; o += 2 * C;
    n->lo[l].W2 = o; o += 2 * C * C;
    n->lo[l].b2 = o; o += C;
  }
  n->oO = o; o += 2 * C;
  n->oob = o; o += 2;
  n->npar = o;
  /* initialisation: fan-in scaled normals; the mode-mixing U starts small so
   * each layer begins near the identity plus its channel MLP */
  float *h = malloc(sizeof(float) * o);
  if (!h) { free(n); return spbte_fail(SPBTE_ERR_ARG, "out of host memory"); }
  uint64_t st = seed ? seed : 0x5EED5EEDull;
  for (long i = 0; i < o; i++) h[i] = 0.0f;
#define FILL(off, cnt, sd) \
  for (long i_ = 0; i_ < (cnt); i_++) h[(off) + i_] = (float)((sd) * rng_normal(&st))
  FILL(n->oP, 4 * C, 0.5);
  FILL(n->oE, M * C, 0.02);
  for (int l = 0; l < layers; l++) {
    FILL(n->lo[l].U, M * r, 0.01 / sqrt((double)r));
    FILL(n->lo[l].V, M * r, 1.0 / sqrt((double)M));
    FILL(n->lo[l].W1, C * 2 * C, 1.0 / sqrt((double)C));
    FILL(n->lo[l].W2, 2 * C * C, 0.5 / sqrt((double)(2 * C)));
  }
  FILL(n->oO, 2 * C, 1.0 / sqrt((double)C));
#undef FILL
  int rc = falloc(&n->par, o);
  if (!rc) rc = falloc(&n->grad, o);
  if (!rc) rc = falloc(&n->m1, o);
  if (!rc) rc = falloc(&n->m2, o);
  n->work_floats = (size_t)1 << 24;
  if ((size_t)M * 5 * C * 2 > n->work_floats) n->work_floats = (size_t)M * 5 * C * 2;
  if ((size_t)M * r * 4 > n->work_floats) n->work_floats = (size_t)M * r * 4;
  if (!rc) rc = falloc(&n->work, n->work_floats);
  if (!rc && hipMemcpy(n->par, h, sizeof(float) * o, hipMemcpyHostToDevice) != hipSuccess)
    rc = spbte_fail(SPBTE_ERR_HIP, "parameter upload failed");
  if (!rc && (hipMemset(n->m1, 0, sizeof(float) * o) != hipSuccess || hipMemset(n->m2, 0, sizeof(float) * o) != hipSuccess))
    rc = spbte_fail(SPBTE_ERR_HIP, "moment clear failed");
  free(h);
  if (rc) { spbte_net_destroy(n); return rc; }
  *out = n;
  return SPBTE_OK;
}

long spbte_net_num_params(spbte_net *n) { return n ? n->npar : 0; }

int spbte_net_get_params(spbte_net *n, float *p) {
  HIPCHK(hipStreamSynchronize(n->s));
  HIPCHK(hipMemcpy(p, n->par, sizeof(float) * n->npar, hipMemcpyDeviceToHost));
  return SPBTE_OK;
}

int spbte_net_set_params(spbte_net *n, const float *p) {
  HIPCHK(hipStreamSynchronize(n->s));
  HIPCHK(hipMemcpy(n->par, p, sizeof(float) * n->npar, hipMemcpyHostToDevice));
  return SPBTE_OK;
}

int spbte_net_set_scales(spbte_net *n, double fscale, double tscale) {
  if (!(fscale > 0) || !(tscale > 0)) return spbte_fail(SPBTE_ERR_ARG, "scales must be positive");
  n->fscale = fscale;
  n->tscale = tscale;
  return SPBTE_OK;
}

int spbte_net_set_loss_floor(spbte_net *n, double rms_floor) {
  if (!n || !(rms_floor >= 0)) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  n->loss_eps = (float)(rms_floor * rms_floor * 2.0 * (double)n->M);
  if (n->loss_eps <= 0.0f) n->loss_eps = 1e-30f;
  return SPBTE_OK;
}

int spbte_net_get_scales(spbte_net *n, double *fscale, double *tscale) {
  *fscale = n->fscale;
  *tscale = n->tscale;
  return SPBTE_OK;
}

/* ------------------------------------------------------------------ forward */
#define P_(off) (n->par + (off))
#define G_(off) (n->grad + (off))

/* C = op(A) op(B) (+ add) (+ bias), G = GELU(C) if G; add has C's layout. */
static int gemm(spbte_net *n, const float *A, long sa0, long sa1, const float *B, long sb0, long sb1, float *C, long sc0,
                long sc1, int M, int N, int K, const float *add, const float *bias, float *G) {
  LCHK(spbte_net_gemm(A, sa0, sa1, B, sb0, sb1, C, sc0, sc1, M, N, K, add, bias, G, n->work, n->work_floats, n->s));
  return SPBTE_OK;
}

/* Network on the features n->X of B cells -> n->Y. */
static int forward(spbte_net *n, int B) {
  const long M = n->M, R = M * B, C = n->C, BC = (long)B * C;
  LCHK(spbte_net_lift(M, B, (int)C, n->X, P_(n->oP), P_(n->oE), P_(n->opb), n->H[0], n->s));
  for (int l = 0; l < n->L; l++) {
    const layer_off *q = &n->lo[l];
    /* S = V^T H  (r x BC) */
    CHK(gemm(n, P_(q->V), 1, n->R, n->H[l], BC, 1, n->S[l], BC, 1, n->R, (int)BC, (int)M, NULL, NULL, NULL));
    /* Hm = H + U S */
    CHK(gemm(n, P_(q->U), n->R, 1, n->S[l], BC, 1, n->Hm[l], BC, 1, (int)M, (int)BC, n->R, n->H[l], NULL, NULL));
    /* Z = Hm W1 + b1, G = GELU(Z) */
    CHK(gemm(n, n->Hm[l], C, 1, P_(q->W1), 2 * C, 1, n->Z[l], 2 * C, 1, (int)R, (int)(2 * C), (int)C, NULL, P_(q->b1),
             n->G[l]));
    /* H' = Hm + G W2 + b2 */
    CHK(gemm(n, n->G[l], 2 * C, 1, P_(q->W2), C, 1, n->H[l + 1], C, 1, (int)R, (int)C, (int)(2 * C), n->Hm[l],
             P_(q->b2), NULL));
  }
  LCHK(spbte_net_head(R, (int)C, n->H[n->L], P_(n->oO), P_(n->oob), n->Y, n->s));
  return SPBTE_OK;
}

/* Gradient of the loss (given n->dY) with respect to every parameter -> n->grad. */
static int backward(spbte_net *n, int B) {
  const long M = n->M, R = M * B, C = n->C, BC = (long)B * C;
  LCHK(spbte_net_head_bwd(R, (int)C, n->H[n->L], P_(n->oO), n->dY, n->dH, G_(n->oO), G_(n->oob), n->work,
                          n->work_floats, n->s));
  for (int l = n->L - 1; l >= 0; l--) {
    const layer_off *q = &n->lo[l];
    /* channel MLP: dG = dH W2^T, dW2 = G^T dH, db2 = colsum dH */
    CHK(gemm(n, n->dH, C, 1, P_(q->W2), 1, C, n->dG, 2 * C, 1, (int)R, (int)(2 * C), (int)C, NULL, NULL, NULL));
    CHK(gemm(n, n->G[l], 1, 2 * C, n->dH, C, 1, G_(q->W2), C, 1, (int)(2 * C), (int)C, (int)R, NULL, NULL, NULL));
    LCHK(spbte_net_colsum(R, (int)C, n->dH, G_(q->b2), 0, n->work, n->work_floats, n->s));
    LCHK(spbte_net_gelu_bwd(R * 2 * C, n->Z[l], n->dG, n->dZ, n->s));
    CHK(gemm(n, n->Hm[l], 1, C, n->dZ, 2 * C, 1, G_(q->W1), 2 * C, 1, (int)C, (int)(2 * C), (int)R, NULL, NULL, NULL));
    LCHK(spbte_net_colsum(R, (int)(2 * C), n->dZ, G_(q->b1), 0, n->work, n->work_floats, n->s));
    /* dHm = dH + dZ W1^T */
    CHK(gemm(n, n->dZ, 2 * C, 1, P_(q->W1), 1, 2 * C, n->dHm, C, 1, (int)R, (int)C, (int)(2 * C), n->dH, NULL, NULL));
    /* mode mixing: dU = dHm S^T, dS = U^T dHm, dH = dHm + V dS, dV = H dS^T */
    CHK(gemm(n, n->dHm, BC, 1, n->S[l], 1, BC, G_(q->U), n->R, 1, (int)M, n->R, (int)BC, NULL, NULL, NULL));
    CHK(gemm(n, P_(q->U), 1, n->R, n->dHm, BC, 1, n->dS, BC, 1, n->R, (int)BC, (int)M, NULL, NULL, NULL));
    CHK(gemm(n, P_(q->V), n->R, 1, n->dS, BC, 1, n->dH, BC, 1, (int)M, (int)BC, n->R, n->dHm, NULL, NULL));
    CHK(gemm(n, n->H[l], BC, 1, n->dS, 1, BC, G_(q->V), n->R, 1, (int)M, n->R, (int)BC, NULL, NULL, NULL));
  }
  LCHK(spbte_net_lift_bwd(M, B, (int)C, n->X, n->dH, G_(n->oE), G_(n->oP), G_(n->opb), n->work, n->work_floats, n->s));
  return SPBTE_OK;
}

/* ------------------------------------------------------ features and targets */
static int moments_and_features(spbte_net *n, const double *d_f, int B) {
  const spbte_grid_dev *g = spbte_op_grid(n->op);
  LCHK(spbte_launch_moments(g, d_f, n->mom, B, NULL, n->s));
  spbte_xform_job jobs[2] = {{d_f, n->mom, 1, n->Fd, NULL, NULL}, {d_f, n->mom, 2, n->Fm, NULL, NULL}};
  LCHK(spbte_launch_xform_fwd(g, jobs, 2, B, 0, n->xs, n->s));
  LCHK(spbte_net_features(n->N, n->Kb, B, n->Fd, n->Fm, n->mom, n->fscale, n->X, n->s));
  return SPBTE_OK;
}

/* Exact Qhat of the operator (ComputeQ or fused maxPreserve, FP64 weights) -> n->T. */
static int exact_targets(spbte_net *n, const double *d_f, double *d_Qscratch, int B) {
  const int flags = SPBTE_SELF | (n->maxpreserve ? SPBTE_MAXPRESERVE : 0);
  CHK(spbte_compute_q(n->op, d_f, NULL, d_Qscratch, B, SPBTE_FORM_FP64, flags));
  LCHK(spbte_net_targets(n->N, n->Kb, B, spbte_op_qhat_buffer(n->op), n->mom, n->tscale, n->T, n->s));
  return SPBTE_OK;
}

int spbte_net_calibrate(spbte_net *n, const double *d_f, double *d_Qscratch, int ncell) {
  if (!n || !d_f || !d_Qscratch || ncell < 1) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  if (ncell > spbte_op_chunk_cells(n->op)) return spbte_fail(SPBTE_ERR_ARG, "batch exceeds one operator chunk");
  CHK(ensure_cap(n, ncell));
  n->fscale = n->tscale = 1.0;
  CHK(moments_and_features(n, d_f, ncell));
  CHK(exact_targets(n, d_f, d_Qscratch, ncell));
  HIPCHK(hipStreamSynchronize(n->s));
  const size_t nx = (size_t)n->M * ncell * 4, nt = (size_t)n->M * ncell * 2;
  float *hx = malloc(sizeof(float) * nx), *ht = malloc(sizeof(float) * nt);
  if (!hx || !ht) { free(hx); free(ht); return spbte_fail(SPBTE_ERR_ARG, "out of host memory"); }
  hipError_t e1 = hipMemcpy(hx, n->X, sizeof(float) * nx, hipMemcpyDeviceToHost);
  hipError_t e2 = hipMemcpy(ht, n->T, sizeof(float) * nt, hipMemcpyDeviceToHost);
  double sx = 0, stt = 0;
  for (size_t i = 0; i < nx; i++) sx += (double)hx[i] * hx[i];
  for (size_t i = 0; i < nt; i++) stt += (double)ht[i] * ht[i];
  free(hx);
  free(ht);
  if (e1 != hipSuccess || e2 != hipSuccess) return spbte_fail(SPBTE_ERR_HIP, "calibration copies failed");
  n->fscale = sx > 0 ? 1.0 / sqrt(sx / nx) : 1.0;
  n->tscale = stt > 0 ? 1.0 / sqrt(stt / nt) : 1.0;
  return SPBTE_OK;
}

int spbte_net_dataset_reserve(spbte_net *n, long nsamples) {
  if (!n || nsamples < 1) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  HIPCHK(hipStreamSynchronize(n->s));
  if (n->dsX) hipFree(n->dsX);
  if (n->dsT) hipFree(n->dsT);
  n->dsX = n->dsT = NULL;
  n->ds_cap = n->ds_n = 0;
  CHK(falloc(&n->dsX, (size_t)nsamples * n->M * 4));
  CHK(falloc(&n->dsT, (size_t)nsamples * n->M * 2));
  n->ds_cap = nsamples;
  return SPBTE_OK;
}

long spbte_net_dataset_size(spbte_net *n) { return n ? n->ds_n : 0; }

static int upload_idx(spbte_net *n, const int *idx, int B) {
  HIPCHK(hipMemcpyAsync(n->d_idx, idx, sizeof(int) * B, hipMemcpyHostToDevice, n->s));
  HIPCHK(hipStreamSynchronize(n->s)); /* idx is host memory the caller may reuse */
  return SPBTE_OK;
}

int spbte_net_dataset_add(spbte_net *n, const double *d_f, double *d_Qscratch, int ncell) {
  if (!n || !d_f || !d_Qscratch || ncell < 1) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  if (n->ds_n + ncell > n->ds_cap) return spbte_fail(SPBTE_ERR_ARG, "dataset full (%ld of %ld)", n->ds_n, n->ds_cap);
  if (ncell > spbte_op_chunk_cells(n->op)) return spbte_fail(SPBTE_ERR_ARG, "batch exceeds one operator chunk");
  CHK(ensure_cap(n, ncell));
  CHK(moments_and_features(n, d_f, ncell));
  CHK(exact_targets(n, d_f, d_Qscratch, ncell));
  int *idx = malloc(sizeof(int) * ncell);
  if (!idx) return spbte_fail(SPBTE_ERR_ARG, "out of host memory");
  for (int b = 0; b < ncell; b++) idx[b] = (int)(n->ds_n + b);
  int rc = upload_idx(n, idx, ncell);
  free(idx);
  CHK(rc);
  LCHK(spbte_net_relayout(n->M, ncell, 4, n->X, n->dsX, 1, n->d_idx, n->s));
  LCHK(spbte_net_relayout(n->M, ncell, 2, n->T, n->dsT, 1, n->d_idx, n->s));
  HIPCHK(hipStreamSynchronize(n->s));
  n->ds_n += ncell;
  return SPBTE_OK;
}

/* Loss and gradient of a batch of dataset samples (no update). */
static int batch_grad(spbte_net *n, const int *idx, int B, double *loss) {
  CHK(ensure_cap(n, B));
  CHK(upload_idx(n, idx, B));
  LCHK(spbte_net_relayout(n->M, B, 4, n->dsX, n->X, 0, n->d_idx, n->s));
  LCHK(spbte_net_relayout(n->M, B, 2, n->dsT, n->T, 0, n->d_idx, n->s));
  CHK(forward(n, B));
  LCHK(spbte_net_loss(n->M, B, n->Y, n->T, n->dY, n->loss, n->loss_eps, n->s));
  CHK(backward(n, B));
  if (loss) {
    float *lb = malloc(sizeof(float) * B);
    if (!lb) return spbte_fail(SPBTE_ERR_ARG, "out of host memory");
    hipError_t e = hipMemcpyAsync(lb, n->loss, sizeof(float) * B, hipMemcpyDeviceToHost, n->s);
    if (e == hipSuccess) e = hipStreamSynchronize(n->s);
    double s = 0;
    for (int b = 0; b < B; b++) s += lb[b];
    free(lb);
    if (e != hipSuccess) return spbte_fail(SPBTE_ERR_HIP, "loss copy failed");
    *loss = s / B;
  }
  return SPBTE_OK;
}

int spbte_net_grad(spbte_net *n, const int *idx, int B, float *grad, double *loss) {
  if (!n || !idx || B < 1) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  for (int b = 0; b < B; b++)
    if (idx[b] < 0 || idx[b] >= n->ds_n) return spbte_fail(SPBTE_ERR_ARG, "sample %d outside the dataset", idx[b]);
  CHK(batch_grad(n, idx, B, loss));
  if (grad) {
    HIPCHK(hipStreamSynchronize(n->s));
    HIPCHK(hipMemcpy(grad, n->grad, sizeof(float) * n->npar, hipMemcpyDeviceToHost));
  }
  return SPBTE_OK;
}

int spbte_net_train(spbte_net *n, int nsteps, int batch, float lr, unsigned long long seed, float *losses) {
  if (!n || nsteps < 0 || batch < 1) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  if (n->ds_n < batch) return spbte_fail(SPBTE_ERR_ARG, "dataset (%ld) smaller than a batch (%d)", n->ds_n, batch);
  int *idx = malloc(sizeof(int) * n->ds_n);
  if (!idx) return spbte_fail(SPBTE_ERR_ARG, "out of host memory");
  for (long i = 0; i < n->ds_n; i++) idx[i] = (int)i;
  uint64_t st = seed ? seed : 0xBA7C4ull;
  long pos = n->ds_n;
  int rc = SPBTE_OK;
  for (int t = 0; t < nsteps && !rc; t++) {
    if (pos + batch > n->ds_n) { /* new epoch: shuffle */
      for (long i = n->ds_n - 1; i > 0; i--) {
        long j = (long)(rng_next(&st) % (uint64_t)(i + 1));
        int tmp = idx[i];
        idx[i] = idx[j];
        idx[j] = tmp;
      }
      pos = 0;
    }
    double loss = 0;
    rc = batch_grad(n, idx + pos, batch, losses ? &loss : NULL);
    pos += batch;
    if (!rc) {
      n->step++;
      rc = spbte_net_adam(n->npar, n->par, n->grad, n->m1, n->m2, lr, 0.9f, 0.999f, 1e-8f, n->step, 0.0f, n->s);
      if (rc) rc = spbte_fail(SPBTE_ERR_HIP, "adam launch: %s", hipGetErrorString((hipError_t)rc));
    }
    if (!rc && losses) losses[t] = (float)loss;
  }
  free(idx);
  if (!rc && hipStreamSynchronize(n->s) != hipSuccess) rc = spbte_fail(SPBTE_ERR_HIP, "training failed");
  return rc;
}

/* Relative L2 error of the network on dataset samples (per sample), no gradient. */
int spbte_net_eval(spbte_net *n, const int *idx, int B, float *err) {
  if (!n || !idx || B < 1 || !err) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  for (int b = 0; b < B; b++)
    if (idx[b] < 0 || idx[b] >= n->ds_n) return spbte_fail(SPBTE_ERR_ARG, "sample %d outside the dataset", idx[b]);
  CHK(ensure_cap(n, B));
  CHK(upload_idx(n, idx, B));
  LCHK(spbte_net_relayout(n->M, B, 4, n->dsX, n->X, 0, n->d_idx, n->s));
  LCHK(spbte_net_relayout(n->M, B, 2, n->dsT, n->T, 0, n->d_idx, n->s));
  CHK(forward(n, B));
  LCHK(spbte_net_loss(n->M, B, n->Y, n->T, NULL, n->loss, n->loss_eps, n->s));
  HIPCHK(hipStreamSynchronize(n->s));
  HIPCHK(hipMemcpy(err, n->loss, sizeof(float) * B, hipMemcpyDeviceToHost));
  for (int b = 0; b < B; b++) err[b] = sqrtf(err[b]);
  return SPBTE_OK;
}

/* Surrogate Q of ncell cells: features -> network -> zero-padded Qhat ->
 * inverse transform (real part) -> optional FP64 projection. */
#define PREDICT_CHUNK 256

int spbte_net_predict(spbte_net *n, const double *d_f, double *d_Q, int ncell, int conserve) {
  if (!n || !d_f || !d_Q || ncell < 1) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  const spbte_grid_dev *g = spbte_op_grid(n->op);
  const size_t n3 = (size_t)n->N * n->N * n->N;
  const int ch = ncell < PREDICT_CHUNK ? ncell : PREDICT_CHUNK;
  CHK(ensure_cap(n, ch > n->cap ? ch : n->cap));
  for (int c0 = 0; c0 < ncell; c0 += ch) {
    const int B = ncell - c0 < ch ? ncell - c0 : ch;
    CHK(moments_and_features(n, d_f + c0 * n3, B));
    CHK(forward(n, B));
    LCHK(spbte_net_scatter(n->N, n->Kb, B, n->Y, n->mom, n->tscale, n->Qh, n->s));
    LCHK(spbte_launch_xform_inv_real(g, n->Qh, 1, d_Q + c0 * n3, B, n->xs, NULL, n->s));
  }
  if (conserve) CHK(spbte_conserve(n->op, d_Q, ncell));
  return SPBTE_OK;
}

/* Forward of the network alone on host features (mode-major [m][b][4]) -> host
 * outputs [m][b][2]; used to check the kernels against a CPU restatement. */
int spbte_net_forward_features(spbte_net *n, const float *X, int B, float *Y) {
  if (!n || !X || !Y || B < 1) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  CHK(ensure_cap(n, B));
  HIPCHK(hipMemcpy(n->X, X, sizeof(float) * n->M * B * 4, hipMemcpyHostToDevice));
  CHK(forward(n, B));
  HIPCHK(hipStreamSynchronize(n->s));
  HIPCHK(hipMemcpy(Y, n->Y, sizeof(float) * n->M * B * 2, hipMemcpyDeviceToHost));
  return SPBTE_OK;
}

/* Loss and gradient on host features and targets (mode-major), no update. */
int spbte_net_grad_features(spbte_net *n, const float *X, const float *T, int B, float *grad, double *loss) {
  if (!n || !X || !T || B < 1) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  CHK(ensure_cap(n, B));
  HIPCHK(hipMemcpy(n->X, X, sizeof(float) * n->M * B * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(n->T, T, sizeof(float) * n->M * B * 2, hipMemcpyHostToDevice));
  CHK(forward(n, B));
  LCHK(spbte_net_loss(n->M, B, n->Y, n->T, n->dY, n->loss, n->loss_eps, n->s));
  CHK(backward(n, B));
  HIPCHK(hipStreamSynchronize(n->s));
  if (grad) HIPCHK(hipMemcpy(grad, n->grad, sizeof(float) * n->npar, hipMemcpyDeviceToHost));
  if (loss) {
    float *lb = malloc(sizeof(float) * B);
    if (!lb) return spbte_fail(SPBTE_ERR_ARG, "out of host memory");
    hipError_t e = hipMemcpy(lb, n->loss, sizeof(float) * B, hipMemcpyDeviceToHost);
    double s = 0;
    for (int b = 0; b < B; b++) s += lb[b];
    free(lb);
    if (e != hipSuccess) return spbte_fail(SPBTE_ERR_HIP, "loss copy failed");
    *loss = s / B;
  }
  return SPBTE_OK;
}

int spbte_net_layout(spbte_net *n, long *offsets) {
  /* P, pb, E, then per layer U, V, W1, b1, W2, b2, then O, ob */
  int k = 0;
  offsets[k++] = n->oP;
  offsets[k++] = n->opb;
  offsets[k++] = n->oE;
  for (int l = 0; l < n->L; l++) {
    offsets[k++] = n->lo[l].U;
    offsets[k++] = n->lo[l].V;
    offsets[k++] = n->lo[l].W1;
    offsets[k++] = n->lo[l].b1;
    offsets[k++] = n->lo[l].W2;
    offsets[k++] = n->lo[l].b2;
  }
  offsets[k++] = n->oO;
  offsets[k++] = n->oob;
  return k;
}

/* Surrogate-assisted collision operator for ncell cells, checked per group.
 * The cells are taken in groups of check_every consecutive cells.  One cell of
 * every group -- the one at `phase` within the group, or the group's last cell
 * if the group is shorter -- is evaluated exactly (form `form`), keeps the
 * exact Q and judges its group: if the surrogate's error there exceeds tol,
 * every cell of the group is evaluated exactly; otherwise the group's other
 * cells keep the surrogate's Q.  Q is projected in every case.  n_exact: cells
 * evaluated exactly (checked cells included); n_failed: groups that fell
 * back; max_err: largest error over the checked cells. */
int spbte_net_collide_groups(spbte_net *n, const double *d_f, double *d_Q, int ncell, int form, int check_every,
                             int phase, double tol, double floor_rel, int *n_exact, int *n_failed, double *max_err) {
  if (!n || !d_f || !d_Q || ncell < 1 || check_every < 1) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  const size_t n3 = (size_t)n->N * n->N * n->N;
  const int flags = SPBTE_SELF | SPBTE_CONSERVE | (n->maxpreserve ? SPBTE_MAXPRESERVE : 0);
  const int ce = check_every;
  phase = ((phase % ce) + ce) % ce;
  const int k = (ncell + ce - 1) / ce; /* groups, one checked cell each */
  if (k > n->scap) {
    HIPCHK(hipStreamSynchronize(n->s));
    double **d[] = {&n->fs, &n->Qs, &n->Qn, &n->err};
    for (size_t i = 0; i < 4; i++)
      if (*d[i]) { hipFree(*d[i]); *d[i] = NULL; }
    if (n->d_cells) { hipFree(n->d_cells); n->d_cells = NULL; }
    HIPCHK(hipMalloc((void **)&n->fs, sizeof(double) * n3 * k));
    HIPCHK(hipMalloc((void **)&n->Qs, sizeof(double) * n3 * k));
    HIPCHK(hipMalloc((void **)&n->Qn, sizeof(double) * n3 * k));
    HIPCHK(hipMalloc((void **)&n->err, sizeof(double) * k));
    n->scap = k;
  }
  if (ncell > n->ccap) {
    HIPCHK(hipStreamSynchronize(n->s));
    if (n->d_cells) hipFree(n->d_cells);
    n->d_cells = NULL;
    HIPCHK(hipMalloc((void **)&n->d_cells, sizeof(int) * ncell));
    n->ccap = ncell;
  }
  int *idx = malloc(sizeof(int) * (size_t)ncell);
  double *he = malloc(sizeof(double) * 3 * (size_t)k);
  if (!idx || !he) { free(idx); free(he); return spbte_fail(SPBTE_ERR_ARG, "out of host memory"); }
  int rc = spbte_net_predict(n, d_f, d_Q, ncell, 1);
  /* the checked cells: gather f and the surrogate's Q, evaluate exactly */
  for (int j = 0; j < k; j++) {
    const int len = (j + 1) * ce <= ncell ? ce : ncell - j * ce;
    idx[j] = j * ce + (phase < len ? phase : len - 1);
  }
  hipError_t e = rc ? hipSuccess : hipMemcpyAsync(n->d_cells, idx, sizeof(int) * k, hipMemcpyHostToDevice, n->s);
  if (!rc && e == hipSuccess) e = (hipError_t)spbte_net_move_cells((long)n3, k, n->d_cells, d_f, n->fs, 0, n->s);
  if (!rc && e == hipSuccess) e = (hipError_t)spbte_net_move_cells((long)n3, k, n->d_cells, d_Q, n->Qn, 0, n->s);
  if (!rc && e == hipSuccess) e = hipStreamSynchronize(n->s); /* idx is reused below */
  if (!rc && e != hipSuccess) rc = spbte_fail(SPBTE_ERR_HIP, "gathering the checked cells: %s", hipGetErrorString(e));
  if (!rc) rc = spbte_compute_q(n->op, n->fs, NULL, n->Qs, k, form, flags);
  /* err_j = |Qn_j - Qs_j| / max(|Qs_j|, floor_rel * max_i |Qs_i|) */
  if (!rc) {
    e = (hipError_t)spbte_net_cell_relerr((long)n3, k, n->Qn, n->Qs, n->err, n->s);
    if (e == hipSuccess) e = hipMemcpyAsync(he, n->err, sizeof(double) * k, hipMemcpyDeviceToHost, n->s);
    if (e == hipSuccess) e = hipStreamSynchronize(n->s);
    if (e == hipSuccess && floor_rel > 0) {
      e = (hipError_t)spbte_net_cell_norm((long)n3, k, n->Qs, n->err, n->s);
      if (e == hipSuccess) e = hipMemcpyAsync(he + k, n->err, sizeof(double) * k, hipMemcpyDeviceToHost, n->s);
      if (e == hipSuccess) e = hipStreamSynchronize(n->s);
      if (e == hipSuccess) e = (hipError_t)spbte_net_cell_norm((long)n3, k, n->Qn, n->err, n->s);
      if (e == hipSuccess) e = hipMemcpyAsync(he + 2 * k, n->err, sizeof(double) * k, hipMemcpyDeviceToHost, n->s);
      if (e == hipSuccess) e = hipStreamSynchronize(n->s);
      double qmax = 0.0;
      for (int j = 0; j < k; j++) qmax = he[k + j] > qmax ? he[k + j] : qmax;
      const double fl = floor_rel * qmax;
      for (int j = 0; j < k && e == hipSuccess; j++) {
        const double q = he[k + j];
        if (q >= fl) continue;
        const double dq = q > 0 ? he[j] * q : he[2 * k + j]; /* |Qn - Qs|; Qs = 0 gives |Qn| */
        he[j] = fl > 0 ? dq / fl : he[j];
      }
    }
    if (e != hipSuccess) rc = spbte_fail(SPBTE_ERR_HIP, "error check failed: %s", hipGetErrorString(e));
  }
  double worst = 0.0;
  int nf = 0, nb = 0; /* failed groups, cells to re-evaluate */
  if (!rc) {
    for (int j = 0; j < k; j++) {
      if (!(he[j] <= worst)) worst = he[j]; /* NaN counts as a failure */
      if (!(he[j] <= tol)) {
        nf++;
        const int end = (j + 1) * ce <= ncell ? (j + 1) * ce : ncell;
        for (int c = j * ce; c < end; c++)
          if (c != idx[j]) idx[k + nb++] = c; /* after the checked cells' indices */
      }
    }
  }
  /* checked cells keep the exact Q */
  if (!rc) {
    e = (hipError_t)spbte_net_move_cells((long)n3, k, n->d_cells, n->Qs, d_Q, 1, n->s);
    if (e != hipSuccess) rc = spbte_fail(SPBTE_ERR_HIP, "scattering the checked cells: %s", hipGetErrorString(e));
  }
  /* the other cells of failed groups: exact, in batches of at most k cells
   * through the checked cells' buffers */
  for (int b0 = 0; b0 < nb && !rc; b0 += k) {
    const int m = nb - b0 < k ? nb - b0 : k;
    e = hipMemcpyAsync(n->d_cells, idx + k + b0, sizeof(int) * m, hipMemcpyHostToDevice, n->s);
    if (e == hipSuccess) e = (hipError_t)spbte_net_move_cells((long)n3, m, n->d_cells, d_f, n->fs, 0, n->s);
    if (e != hipSuccess) { rc = spbte_fail(SPBTE_ERR_HIP, "gathering fallback cells: %s", hipGetErrorString(e)); break; }
    rc = spbte_compute_q(n->op, n->fs, NULL, n->Qs, m, form, flags);
    if (!rc) e = (hipError_t)spbte_net_move_cells((long)n3, m, n->d_cells, n->Qs, d_Q, 1, n->s);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(n->s); /* before d_cells is rewritten */
    if (!rc && e != hipSuccess) rc = spbte_fail(SPBTE_ERR_HIP, "fallback cells: %s", hipGetErrorString(e));
  }
  if (!rc && hipStreamSynchronize(n->s) != hipSuccess) rc = spbte_fail(SPBTE_ERR_HIP, "surrogate step failed");
  free(idx);
  free(he);
  if (rc) return rc;
  if (n_exact) *n_exact = k + nb;
  if (n_failed) *n_failed = nf;
  if (max_err) *max_err = worst;
  return SPBTE_OK;
}

int spbte_net_collide(spbte_net *n, const double *d_f, double *d_Q, int ncell, int form, int check_every, int phase,
                      double tol, double floor_rel, int *n_exact, double *max_err) {
  return spbte_net_collide_groups(n, d_f, d_Q, ncell, form, check_every, phase, tol, floor_rel, n_exact, NULL, max_err);
}
