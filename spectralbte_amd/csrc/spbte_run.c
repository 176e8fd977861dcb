/* spbte_run.c -- batched time integration on one GPU (C99).
 *
 * Replaces the per-cell driver loops of exec/boltz.c on one rank, for one
 * species or several: the 0D loop (:193-241: ComputeQ_maxPreserve + conserveAllMoments, then
 * forward Euler or Heun) and the 1D loop (:285-354: transport, then per-cell
 * ComputeQ + conserveAllMoments with Euler or Heun, then for order 2 a second
 * transport stage).  All cells are one batch: a step is a fixed sequence of
 * launches on the operator's stream (transport kernels, one batched collision
 * evaluation per collision stage, elementwise updates), and it can be captured
 * once as a HIP graph and replayed.  The state never leaves HBM between steps.
 */
#include <math.h>
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

#define CHK(expr)               \
  do {                          \
    int r_ = (expr);            \
    if (r_ != 0) return r_;     \
  } while (0)

#define LCHK(expr)                                                                                             \
  do {                                                                                                         \
    int e_ = (expr);                                                                                           \
    if (e_ != 0) return spbte_fail(SPBTE_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString((hipError_t)e_), \
                                   __FILE__, __LINE__);                                                        \
  } while (0)

struct spbte_run {
  spbte_op *op;
  int ns;       /* species; species l's cells follow species l-1's in every state buffer */
  double wmass[SPBTE_MAX_NS]; /* species masses (diffuse walls of a multi-species run) */
  int dim;      /* 0: ensemble of 0D problems; 1: 1D mesh */
  int ncell;    /* problems (0D) or interior cells nX (1D) */
  int order;    /* 1: Euler (+ upwindOne); 2: Heun (+ advectTwo) */
  int form, flags;
  double dt, Kn;
  int bc;       /* 1D walls: SPBTE_BC_* */
  double T0, T1;
  double force_ma; /* Poiseuille body force along v_y (order 2), 0 = none */
  size_t n3;
  int nghost;   /* ghost cells per side (1D: order; 0D: 0) */
  double *d_x, *d_dx;            /* 1D: nX + 2 order */
  double *d_f, *d_fc, *d_tmp;    /* state (with ghosts), transported state, advectTwo stage (one species) */
  double *d_f1, *d_Q;            /* Heun stage; collision operator, pair (l, m) at (m*ns + l) * ncell * N^3 */
  double *d_fl, *d_fr;           /* order-2 wall states */
  double *d_mom;
  int use_graph;
  /* optional learned surrogate with its sampled check */
  spbte_net *net;
  int check_every, phase;
  double tol, floor_rel;
  long s_cells, s_exact, s_fallbacks;
  double s_max_err;
  hipGraph_t graph;
  hipGraphExec_t exec;
  unsigned gen;
};

static int dev_alloc(double **p, size_t n) {
  HIPCHK(hipMalloc((void **)p, sizeof(double) * (n ? n : 1)));
  HIPCHK(hipMemset(*p, 0, sizeof(double) * (n ? n : 1)));
  return SPBTE_OK;
}

static void drop_graph(spbte_run *r) {
  if (r->exec) hipGraphExecDestroy(r->exec);
  if (r->graph) hipGraphDestroy(r->graph);
  r->exec = NULL;
  r->graph = NULL;
}

void spbte_run_destroy(spbte_run *r) {
  if (!r) return;
  hipStreamSynchronize(spbte_op_stream(r->op));
  drop_graph(r);
  double *bufs[] = {r->d_x, r->d_dx, r->d_f, r->d_fc, r->d_tmp, r->d_f1, r->d_Q, r->d_fl, r->d_fr, r->d_mom};
  for (size_t i = 0; i < sizeof bufs / sizeof bufs[0]; i++)
    if (bufs[i]) hipFree(bufs[i]);
  free(r);
}

static int run_new(spbte_run **out, spbte_op *op, int dim, int ncell, int order, double dt, double Kn, int form) {
  if (!out || !op) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  if (ncell < 1) return spbte_fail(SPBTE_ERR_ARG, "ncell=%d must be positive", ncell);
  if (order != 1 && order != 2) return spbte_fail(SPBTE_ERR_ARG, "order=%d must be 1 or 2", order);
  if (form < 0 || form > SPBTE_FORM_FP32A) return spbte_fail(SPBTE_ERR_ARG, "bad form %d", form);
  if (!(Kn > 0.0)) return spbte_fail(SPBTE_ERR_ARG, "Kn must be positive");
  spbte_run *r = calloc(1, sizeof *r);
  if (!r) return spbte_fail(SPBTE_ERR_ARG, "out of host memory");
  const int N = spbte_op_n(op);
  r->op = op;
  r->ns = 1;
  r->dim = dim;
  r->ncell = ncell;
  r->order = order;
  r->dt = dt;
  r->Kn = Kn;
  r->form = form;
  r->n3 = (size_t)N * N * N;
  *out = r;
  return SPBTE_OK;
}

/* Species of a multi-species run from the operator (spbte_op_set_species). */
static int take_species(spbte_run *r, int Ns) {
  int ons = 0;
  CHK(spbte_op_get_species(r->op, &ons, r->wmass));
  if (Ns < 1 || Ns > SPBTE_MAX_NS) return spbte_fail(SPBTE_ERR_ARG, "Ns=%d out of range", Ns);
  if (Ns > 1 && ons != Ns)
    return spbte_fail(SPBTE_ERR_ARG, "the operator holds %d species (spbte_op_set_species), the run %d", ons, Ns);
  r->ns = Ns;
  return SPBTE_OK;
}

int spbte_run0d_create_species(spbte_run **out, spbte_op *op, int Ns, int ncell, int order, double dt, double Kn,
                               int form, int maxpreserve) {
  CHK(run_new(out, op, 0, ncell, order, dt, Kn, form));
  spbte_run *r = *out;
  r->flags = SPBTE_SELF | SPBTE_CONSERVE | (maxpreserve ? SPBTE_MAXPRESERVE : 0);
  int rc = take_species(r, Ns);
  const size_t n = (size_t)ncell * r->n3;
  if (!rc) rc = dev_alloc(&r->d_f, Ns * n);
  if (!rc) rc = dev_alloc(&r->d_Q, (size_t)Ns * Ns * n);
  if (!rc && order == 2) rc = dev_alloc(&r->d_f1, Ns * n);
  if (!rc) rc = dev_alloc(&r->d_mom, (size_t)Ns * ncell * SPBTE_MOM_STRIDE);
  if (rc) { spbte_run_destroy(r); *out = NULL; }
  return rc;
}

int spbte_run0d_create(spbte_run **out, spbte_op *op, int ncell, int order, double dt, double Kn, int form,
                       int maxpreserve) {
  return spbte_run0d_create_species(out, op, 1, ncell, order, dt, Kn, form, maxpreserve);
}

int spbte_run1d_create_species(spbte_run **out, spbte_op *op, int Ns, int nX, int order, const double *x,
                               const double *dx, double dt, double Kn, int bc, double T0, double T1, int form) {
  if (!x || !dx) return spbte_fail(SPBTE_ERR_ARG, "x and dx are required");
  if (bc < SPBTE_BC_COPY || bc > SPBTE_BC_DIFFUSE_LEFT) return spbte_fail(SPBTE_ERR_ARG, "bad boundary condition %d", bc);
  if (order == 2 && bc == SPBTE_BC_PERIODIC)
    return spbte_fail(SPBTE_ERR_UNSUPPORTED, "periodic ghosts exist for order 1 only (src/transportroutines.c:156-172)");
  if (order == 2 && nX < 2) return spbte_fail(SPBTE_ERR_ARG, "order 2 needs nX >= 2");
  CHK(run_new(out, op, 1, nX, order, dt, Kn, form));
  spbte_run *r = *out;
  r->flags = SPBTE_SELF | SPBTE_CONSERVE;
  r->bc = bc;
  r->T0 = T0;
  r->T1 = T1;
  r->nghost = order;
  int rc = take_species(r, Ns);
  const size_t nc = (size_t)nX + 2 * order;
  if (!rc) rc = dev_alloc(&r->d_x, nc);
  if (!rc) rc = dev_alloc(&r->d_dx, nc);
  if (!rc) rc = dev_alloc(&r->d_f, Ns * nc * r->n3);
  if (!rc) rc = dev_alloc(&r->d_fc, Ns * nc * r->n3);
  if (!rc) rc = dev_alloc(&r->d_Q, (size_t)Ns * Ns * nX * r->n3);
  if (!rc && order == 2) rc = dev_alloc(&r->d_tmp, nc * r->n3);
  if (!rc && order == 2) rc = dev_alloc(&r->d_f1, (size_t)Ns * nX * r->n3);
  if (!rc && order == 2) rc = dev_alloc(&r->d_fl, r->n3);
  if (!rc && order == 2) rc = dev_alloc(&r->d_fr, r->n3);
  if (!rc) rc = dev_alloc(&r->d_mom, (size_t)Ns * nX * SPBTE_MOM_STRIDE);
  if (!rc && hipMemcpy(r->d_x, x, sizeof(double) * nc, hipMemcpyHostToDevice) != hipSuccess)
    rc = spbte_fail(SPBTE_ERR_HIP, "copy of x failed");
  if (!rc && hipMemcpy(r->d_dx, dx, sizeof(double) * nc, hipMemcpyHostToDevice) != hipSuccess)
    rc = spbte_fail(SPBTE_ERR_HIP, "copy of dx failed");
  if (rc) { spbte_run_destroy(r); *out = NULL; }
  return rc;
}

int spbte_run1d_create(spbte_run **out, spbte_op *op, int nX, int order, const double *x, const double *dx, double dt,
                       double Kn, int bc, double T0, double T1, int form) {
  return spbte_run1d_create_species(out, op, 1, nX, order, x, dx, dt, Kn, bc, T0, T1, form);
}

static double *interior(const spbte_run *r, double *buf) { return buf + (size_t)r->nghost * r->n3; }
/* Cells of one species in a ghosted buffer (d_f, d_fc) and in an interior-only one (d_f1). */
static size_t gblock(const spbte_run *r) { return ((size_t)r->ncell + 2 * r->nghost) * r->n3; }
static size_t iblock(const spbte_run *r) { return (size_t)r->ncell * r->n3; }

int spbte_run_set_state(spbte_run *r, const double *f) {
  if (!r || !f) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(spbte_op_stream(r->op)));
  for (int l = 0; l < r->ns; l++)
    HIPCHK(hipMemcpy(interior(r, r->d_f + l * gblock(r)), f + l * iblock(r), sizeof(double) * iblock(r),
                     hipMemcpyHostToDevice));
  return SPBTE_OK;
}

int spbte_run_get_state(spbte_run *r, double *f) {
  if (!r || !f) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(spbte_op_stream(r->op)));
  for (int l = 0; l < r->ns; l++)
    HIPCHK(hipMemcpy(f + l * iblock(r), interior(r, r->d_f + l * gblock(r)), sizeof(double) * iblock(r),
                     hipMemcpyDeviceToHost));
  return SPBTE_OK;
}

int spbte_run_state(spbte_run *r, double **d_f) {
  if (!r || !d_f) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  *d_f = interior(r, r->d_f);
  return SPBTE_OK;
}

int spbte_run_species_state(spbte_run *r, int l, double **d_f) {
  if (!r || !d_f || l < 0 || l >= r->ns) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  *d_f = interior(r, r->d_f + l * gblock(r));
  return SPBTE_OK;
}

int spbte_run_moments(spbte_run *r, double *out) {
  if (!r || !out) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  hipStream_t s = spbte_op_stream(r->op);
  const size_t nm = (size_t)r->ns * r->ncell;
  for (int l = 0; l < r->ns; l++)
    LCHK(spbte_launch_moments(spbte_op_grid(r->op), interior(r, r->d_f + l * gblock(r)),
                              r->d_mom + (size_t)l * r->ncell * SPBTE_MOM_STRIDE, r->ncell, NULL, s));
  HIPCHK(hipStreamSynchronize(s));
  double *m = malloc(sizeof(double) * nm * SPBTE_MOM_STRIDE);
  if (!m) return spbte_fail(SPBTE_ERR_ARG, "out of host memory");
  hipError_t e = hipMemcpy(m, r->d_mom, sizeof(double) * nm * SPBTE_MOM_STRIDE, hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    for (size_t c = 0; c < nm; c++)
      for (int q = 0; q < 5; q++) out[5 * c + q] = m[SPBTE_MOM_STRIDE * c + q];
  free(m);
  if (e != hipSuccess) return spbte_fail(SPBTE_ERR_HIP, "moments copy: %s", hipGetErrorString(e));
  return SPBTE_OK;
}

/* ---------------------------------------------------------------- transport */
/* One upwindTwo stage src -> dst (src/transportroutines.c:241-470): ghost
 * extrapolation in src, wall states, diffuse walls, MUSCL half-step. */
static int upwind_two(spbte_run *r, const double *d_x, const double *d_dx, double *src, double *dst, int l) {
  const spbte_grid_dev *g = spbte_op_grid(r->op);
  spbte_grid_dev gw = *g; /* walls with species l's mass (setDiffuseReflectionBC's mixture[id]) */
  if (r->ns > 1) gw.mass = r->wmass[l];
  hipStream_t s = spbte_op_stream(r->op);
  const int N = g->N, nX = r->ncell;
  const int diffuse = r->bc == SPBTE_BC_DIFFUSE ? 3 : r->bc == SPBTE_BC_DIFFUSE_LEFT ? 1 : 0;
  LCHK(spbte_launch_ghosts(N, nX, 3, src, s));
  LCHK(spbte_launch_walls2(N, nX, diffuse, src, r->d_fl, r->d_fr, d_x, d_dx, s));
  if (diffuse) {
    const double *ws[2] = {r->d_fl, r->d_fr};
    double *wd[2] = {r->d_fl, r->d_fr};
    const double TW[2] = {r->T0, r->T1};
    const int right[2] = {0, 1};
    LCHK(spbte_launch_diffuse(&gw, diffuse == 3 ? 2 : 1, ws, wd, TW, right, s));
  }
  /* Ma*0.5*dt/(2*h_v) with h_v = 2 L_v / (N - 1), src/transportroutines.c:37,429 */
  const double fy = r->force_ma != 0.0 ? r->force_ma * 0.5 * r->dt / (2 * (2 * g->Lv / (N - 1))) : 0.0;
  LCHK(spbte_launch_upwind2(N, nX, src, dst, r->d_fl, r->d_fr, (const double *)g->d_v, d_x, d_dx, r->dt, fy, s));
  return SPBTE_OK;
}

/* Transport of species l, src -> dst over the interior cells: upwindOne
 * (order 1, src/transportroutines.c:94-238) or advectTwo (order 2, :477-492). */
static int advect(spbte_run *r, double *src, double *dst, int l) {
  const spbte_grid_dev *g = spbte_op_grid(r->op);
  spbte_grid_dev gw = *g;
  if (r->ns > 1) gw.mass = r->wmass[l];
  hipStream_t s = spbte_op_stream(r->op);
  const int N = g->N, nX = r->ncell;
  if (r->order == 1) {
    if (r->bc == SPBTE_BC_DIFFUSE) {
      const double *ws[2] = {src + r->n3, src + (size_t)nX * r->n3};
      double *wd[2] = {src, src + (size_t)(nX + 1) * r->n3};
      const double TW[2] = {r->T0, r->T1};
      const int right[2] = {0, 1};
      LCHK(spbte_launch_diffuse(&gw, 2, ws, wd, TW, right, s));
    } else if (r->bc == SPBTE_BC_DIFFUSE_LEFT) {
      /* right ghost copies cell nX; the left ghost's incoming half is the wall's */
      const double *ws[1] = {src + r->n3};
      double *wd[1] = {src};
      const double TW[1] = {r->T0};
      const int right[1] = {0};
      LCHK(spbte_launch_ghosts(N, nX, 0, src, s));
      LCHK(spbte_launch_diffuse(&gw, 1, ws, wd, TW, right, s));
    } else {
      LCHK(spbte_launch_ghosts(N, nX, r->bc == SPBTE_BC_PERIODIC ? 2 : 0, src, s));
    }
    LCHK(spbte_launch_upwind1(N, nX, src, dst, (const double *)g->d_v, r->d_dx, r->dt, s));
    return SPBTE_OK;
  }
  CHK(upwind_two(r, r->d_x, r->d_dx, src, r->d_tmp, l));
  CHK(upwind_two(r, r->d_x, r->d_dx, r->d_tmp, dst, l));
  const size_t n = (size_t)nX * r->n3;
  LCHK(spbte_launch_lincomb(n, 0.5, interior(r, src), 0.5, interior(r, dst), 0.0, NULL, interior(r, dst), s));
  return SPBTE_OK;
}

/* ------------------------------------------------------------------- a step */
/* Collision stage of all species: d_in is species 0's first cell, species l's
 * at d_in + l * stride.  One species: Q (+ projection) in one batched call.
 * Several (exec/boltz.c:193-198, :286-291): Q(f_l, f_m) with the table of pair
 * (l, m) in operator slot m*Ns + l, then the joint Ns + 4 projection
 * (src/conserve.c:207-264) over all pairs of each cell. */
static int collide(spbte_run *r, const double *d_in, size_t stride) {
  if (r->ns > 1) {
    spbte_op *op = r->op;
    const int ns = r->ns, slot0 = spbte_op_table_slot(op);
    const size_t n = iblock(r);
    int rc = SPBTE_OK;
    for (int l = 0; l < ns && !rc; l++)
      for (int m = 0; m < ns && !rc; m++) {
        rc = spbte_op_select_table(op, m * ns + l);
        if (!rc)
          rc = spbte_compute_q(op, d_in + l * stride, l == m ? NULL : d_in + m * stride, r->d_Q + (size_t)(m * ns + l) * n,
                               r->ncell, r->form, r->flags & SPBTE_MAXPRESERVE);
      }
    spbte_op_select_table(op, slot0);
    return rc ? rc : spbte_conserve_pairs(op, r->d_Q, r->ncell);
  }
  if (!r->net) return spbte_compute_q(r->op, d_in, NULL, r->d_Q, r->ncell, r->form, r->flags);
  int n_exact = 0, n_failed = 0;
  double err = 0.0;
  CHK(spbte_net_collide_groups(r->net, d_in, r->d_Q, r->ncell, r->form, r->check_every, r->phase, r->tol,
                               r->floor_rel, &n_exact, &n_failed, &err));
  r->phase = (r->phase + 1) % r->check_every;
  r->s_cells += r->ncell;
  r->s_exact += n_exact;
  r->s_fallbacks += n_failed;
  if (!(err <= r->s_max_err)) r->s_max_err = err;
  return SPBTE_OK;
}

int spbte_run_set_forcing(spbte_run *r, double Ma) {
  if (!r || !isfinite(Ma)) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  if (r->dim != 1 || r->order != 2)
    return spbte_fail(SPBTE_ERR_UNSUPPORTED,
                      "the v_y body force exists for 1D order-2 runs only: the reference's order-1 term is applied "
                      "after its k loop and writes outside the cell (src/transportroutines.c:218-226)");
  HIPCHK(hipStreamSynchronize(spbte_op_stream(r->op)));
  drop_graph(r);
  r->force_ma = Ma;
  return SPBTE_OK;
}

int spbte_run_set_surrogate(spbte_run *r, spbte_net *net, int check_every, double tol, double floor_rel) {
  if (!r || check_every < 1 || !(tol > 0) || !(floor_rel >= 0)) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  if (net && r->ns > 1) return spbte_fail(SPBTE_ERR_UNSUPPORTED, "the surrogate stands in for one species' Q");
  HIPCHK(hipStreamSynchronize(spbte_op_stream(r->op)));
  drop_graph(r);
  r->net = net;
  r->check_every = check_every;
  r->tol = tol;
  r->floor_rel = floor_rel;
  r->phase = 0;
  r->s_cells = r->s_exact = r->s_fallbacks = 0;
  r->s_max_err = 0.0;
  return SPBTE_OK;
}

int spbte_run_surrogate_stats(spbte_run *r, long *cells, long *exact, long *fallbacks, double *max_err) {
  if (!r) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  if (cells) *cells = r->s_cells;
  if (exact) *exact = r->s_exact;
  if (fallbacks) *fallbacks = r->s_fallbacks;
  if (max_err) *max_err = r->s_max_err;
  return SPBTE_OK;
}

/* out_l = a x_l + b y_l + c sum_m Q(l, m) for every species l, the sum taken
 * pair by pair in m order as exec/boltz.c:205-240, :308-341 add them
 * (y may be null with b = 0; xs, ys, os: species strides). */
static int update(spbte_run *r, double a, const double *x, size_t xs, double b, const double *y, size_t ys, double c,
                  double *out, size_t os) {
  hipStream_t s = spbte_op_stream(r->op);
  const size_t n = iblock(r);
  for (int l = 0; l < r->ns; l++) {
    const double *q0 = r->d_Q + (size_t)l * n; /* pair (l, 0) */
    double *o = out + l * os;
    if (y) LCHK(spbte_launch_lincomb(n, a, x + l * xs, b, y + l * ys, c, q0, o, s));
    else LCHK(spbte_launch_lincomb(n, a, x + l * xs, c, q0, 0.0, NULL, o, s));
    for (int m = 1; m < r->ns; m++)
      LCHK(spbte_launch_lincomb(n, 1.0, o, c, r->d_Q + (size_t)(m * r->ns + l) * n, 0.0, NULL, o, s));
  }
  return SPBTE_OK;
}

/* collide(d_in), then update(...); with one species and the exact operator the
 * update rides on the evaluation's last pass where the operator can fuse it. */
static int collide_update(spbte_run *r, const double *d_in, size_t stride, double a, const double *x, size_t xs,
                          double b, const double *y, size_t ys, double c, double *out, size_t os) {
  if (r->ns == 1 && !r->net) {
    const spbte_epilogue ep = {a, b, c, x, y, out};
    int fused = 0;
    CHK(spbte_compute_q_update(r->op, d_in, r->d_Q, r->ncell, r->form, r->flags, &ep, &fused));
    return fused ? SPBTE_OK : update(r, a, x, xs, b, y, ys, c, out, os);
  }
  CHK(collide(r, d_in, stride));
  return update(r, a, x, xs, b, y, ys, c, out, os);
}

static int advect_all(spbte_run *r, double *src, double *dst) {
  for (int l = 0; l < r->ns; l++) CHK(advect(r, src + l * gblock(r), dst + l * gblock(r), l));
  return SPBTE_OK;
}

static int enqueue_step(spbte_run *r) {
  const double h = r->dt / r->Kn;
  const size_t G = gblock(r), I = iblock(r);
  double *f = interior(r, r->d_f);
  if (r->dim == 0) {
    /* exec/boltz.c:193-241 */
    if (r->order == 1) return collide_update(r, f, G, 1.0, f, G, 0.0, NULL, 0, h, f, G);
    CHK(collide_update(r, f, G, 1.0, f, G, 0.0, NULL, 0, h, r->d_f1, I));
    return collide_update(r, r->d_f1, I, 0.5, f, G, 0.5, r->d_f1, I, 0.5 * h, f, G);
  }
  /* exec/boltz.c:285-354 */
  double *fc = interior(r, r->d_fc);
  CHK(advect_all(r, r->d_f, r->d_fc));
  if (r->order == 1) return collide_update(r, fc, G, 1.0, fc, G, 0.0, NULL, 0, h, f, G);
  CHK(collide_update(r, fc, G, 1.0, fc, G, 0.0, NULL, 0, h, r->d_f1, I));
  CHK(collide_update(r, r->d_f1, I, 0.5, fc, G, 0.5, r->d_f1, I, 0.5 * h, fc, G));
  return advect_all(r, r->d_fc, r->d_f);
}

/* Capture one step as a graph.  Everything a step allocates on first use
 * (operator scratch, derived weight tables) is allocated by an evaluation of
 * the operator before the capture starts, so the captured sequence contains
 * kernel launches only. */
static int capture(spbte_run *r) {
  hipStream_t s = spbte_op_stream(r->op);
  drop_graph(r);
  CHK(collide(r, interior(r, r->d_f), gblock(r)));
  HIPCHK(hipStreamSynchronize(s));
  const unsigned gen = spbte_op_generation(r->op);
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_step(r);
  hipGraph_t graph = NULL;
  hipError_t e = hipStreamEndCapture(s, &graph);
  if (rc) {
    if (graph) hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return spbte_fail(SPBTE_ERR_HIP, "stream capture: %s", hipGetErrorString(e));
  r->graph = graph;
  HIPCHK(hipGraphInstantiate(&r->exec, r->graph, NULL, NULL, 0));
  if (spbte_op_generation(r->op) != gen) {
    drop_graph(r);
    return spbte_fail(SPBTE_ERR_HIP, "operator state changed during capture");
  }
  r->gen = gen;
  return SPBTE_OK;
}

int spbte_run_use_graph(spbte_run *r, int on) {
  if (!r) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  r->use_graph = on != 0;
  if (!r->use_graph) {
    HIPCHK(hipStreamSynchronize(spbte_op_stream(r->op)));
    drop_graph(r);
  }
  return SPBTE_OK;
}

int spbte_run_step(spbte_run *r, int nsteps) {
  if (!r || nsteps < 0) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  hipStream_t s = spbte_op_stream(r->op);
  for (int t = 0; t < nsteps; t++) {
    if (r->use_graph && !r->net) { /* the surrogate's check synchronises with the host: no graph */
      if (!r->exec || r->gen != spbte_op_generation(r->op)) CHK(capture(r));
      HIPCHK(hipGraphLaunch(r->exec, s));
    } else {
      CHK(enqueue_step(r));
    }
  }
  return SPBTE_OK;
}

/* ------------------------------------------------------ standalone transport */
int spbte_advect(spbte_op *op, double *d_f, double *d_fc, int nX, int order, const double *x, const double *dx,
                 double dt, int bc, double T0, double T1) {
  if (!op || !d_f || !d_fc || !x || !dx || nX < 1) return spbte_fail(SPBTE_ERR_ARG, "bad arguments");
  spbte_run *r = NULL;
  CHK(spbte_run1d_create(&r, op, nX, order, x, dx, dt, 1.0, bc, T0, T1, SPBTE_FORM_FP64));
  int rc = advect(r, d_f, d_fc, 0);
  if (!rc && hipStreamSynchronize(spbte_op_stream(op)) != hipSuccess) rc = spbte_fail(SPBTE_ERR_HIP, "advect failed");
  spbte_run_destroy(r);
  return rc;
}

int spbte_run_nspecies(spbte_run *r) { return r ? r->ns : 0; }

int spbte_run_geometry(spbte_run *r, spbte_op **op, int *dim, int *ncell, const double **d_state, const double **d_x) {
  if (!r) return spbte_fail(SPBTE_ERR_ARG, "null argument");
  *op = r->op;
  *dim = r->dim;
  *ncell = r->ncell;
  *d_state = interior(r, r->d_f);
  *d_x = r->d_x ? r->d_x + r->nghost : NULL;
  return SPBTE_OK;
}
