/* spbte.h -- batched, device-resident interface of the MI355X collision operator.
 *
 * No counterpart in the reference, which evaluates one cell per call on the CPU
 * (src/collisions.c:178-221) and spreads cells over MPI ranks
 * (exec/boltz.c:285-345, src/mesh_setup.c:46-53).  This interface is the
 * single-GPU batched driver: one call evaluates the operator for a whole batch
 * of cells (a 1D-3V mesh or a 0D ensemble) that share a velocity grid and a
 * weight table, so every row of the table streamed from HBM serves many cells.
 *
 * Conventions: arrays named d_* are device pointers (from spbte_malloc or any
 * HIP allocation in this process); a batch is ncell contiguous blocks of N^3
 * doubles in the reference's (i,j,k) -> k + N(j + N i) order.  Functions return
 * SPBTE_OK (0) or a negative SPBTE_ERR_* code; spbte_last_error() describes the
 * last failure of the calling thread.  All work runs on the operator's stream;
 * batched calls return once the work is queued unless stated otherwise.
 */
#ifndef SPBTE_H
#define SPBTE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SPBTE_OK = 0,
  SPBTE_ERR_ARG = -1,       /* bad argument */
  SPBTE_ERR_HIP = -2,       /* HIP runtime or launch failure */
  SPBTE_ERR_NOWEIGHTS = -3, /* no weight table set */
  SPBTE_ERR_IO = -4,        /* weight file missing or short */
  SPBTE_ERR_SINGULAR = -5,  /* C C^T singular */
  SPBTE_ERR_NODEV = -6,     /* no GPU visible */
  SPBTE_ERR_UNSUPPORTED = -7
};

/* Precision forms of the weighted sum.  FP64 is the reference method.  FP32
 * streams W as float (half the HBM bytes) and widens it to FP64 in registers;
 * FP16X2 streams W as two FP16 halves (hi + lo) with a power-of-two scale per
 * row, widened to FP64; FP32A streams W as float and also computes the products
 * in FP32: the batched (streaming) kernel sums 8 to 32 steps in FP32 and adds
 * each partial sum into an FP64 accumulator (its maxPreserve tiles with one
 * problem per lane take each complex product in Gauss's three-multiplication
 * form), the tiled kernel uses compensated (TwoSum) FP32 sums.  Transforms,
 * moments and the projection are FP64 in every form. */
enum { SPBTE_FORM_FP64 = 0, SPBTE_FORM_FP32 = 1, SPBTE_FORM_FP16X2 = 2, SPBTE_FORM_FP32A = 3 };

/* flags of spbte_compute_q */
enum {
  SPBTE_MAXPRESERVE = 1, /* ComputeQ_maxPreserve semantics (fused single pass) */
  SPBTE_CONSERVE = 2,    /* apply the single-species projection to each cell */
  SPBTE_SELF = 4         /* g == f: d_g is ignored */
};

typedef struct spbte_op spbte_op;

const char *spbte_last_error(void);

/* Device memory and synchronisation helpers (callers need no HIP headers). */
int spbte_device_count(int *n);
int spbte_malloc(void **d_ptr, size_t bytes);
int spbte_free(void *d_ptr);
int spbte_memcpy_htod(void *d_dst, const void *src, size_t bytes);
int spbte_memcpy_dtoh(void *dst, const void *d_src, size_t bytes);
int spbte_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes);
/* Copy with a plain streaming kernel (16-byte aligned, size a multiple of 16);
 * synchronous.  Also the known-size reference for HBM byte counters. */
int spbte_device_copy(void *d_dst, const void *d_src, size_t bytes);
int spbte_op_synchronize(spbte_op *op);
/* Vector ALU issue ceiling, measured (no counterpart in the reference): a
 * kernel of independent FMA chains over every CU (fp64 != 0: FP64, else
 * unpacked FP32, the instructions this library issues; one operand
 * wave-uniform) returns its rate in TFLOP/s (2 flops per FMA) and the shader
 * clock during the run in GHz.  The roofline of the batched weighted sum, which
 * is bound by FMA issue.  Synchronous, about 10 ms. */
int spbte_valu_peak(int fp64, double *tflops, double *clock_ghz);

/* One velocity grid: N nodes per axis on [-L_v, L_v] with the caller's v and
 * eta arrays (either grid convention of src/initializer.c:66-82 or :256-266). */
int spbte_op_create(spbte_op **op, int N, double L_v, const double *v, const double *eta);
void spbte_op_destroy(spbte_op *op);

/* Species: Ns masses (Ns = 1 for the batched projection).  Builds C C^T and
 * its LU factors (src/conserve.c:268-317); masses[0] also sets the mass of the
 * maxPreserve Maxwellian.  Default: one species of mass 1. */
int spbte_op_set_species(spbte_op *op, int Ns, const double *masses);

/* Weight table.  set_weights takes N^3*N^3 contiguous host doubles (the file
 * layout); set_weights_rows takes the reference's N^3 row pointers;
 * load_weights_file reads a Weights/ file; generate_weights_iso builds the
 * isotropic table on the GPU (src/weights.c:265-281 closed forms for lambda 0
 * and 1, quadrature otherwise) for species diameters d_i, d_j and masses m_i, m_j.
 * get_weights copies the FP64 table back to host memory (contiguous). */
/* Table slots: an operator holds up to 64 tables on its grid (8 species),
 * one per species pair of a multi-species run (slot m*Ns + l for the pair
 * (l, m), the reference's conv_weights[m*Ns + l], exec/boltz.c:157-160).
 * The table calls below and spbte_compute_q act on the selected slot
 * (slot 0 until another is selected). */
int spbte_op_select_table(spbte_op *op, int slot);
int spbte_op_set_weights(spbte_op *op, const double *W);
int spbte_op_set_weights_rows(spbte_op *op, const double *const *rows);
int spbte_op_load_weights_file(spbte_op *op, const char *path);
int spbte_op_generate_weights_iso(spbte_op *op, double lambda, double d_i, double d_j, double m_i, double m_j);
/* The anisotropic generator's table with glance = 0, the Landau-Fokker-Planck
 * limit: gHat3L (exec/MPIcollisionRoutines.c:205-281, src/aniso_weights.c:352-428)
 * at every (row, column) pair of the operator's eta grid and L_v, by nested
 * 64-point Gauss-Legendre rules, on the GPU.  As the reference writes it, the
 * table carries no trapezoid or h_eta^3 factors. */
int spbte_op_generate_weights_landau(spbte_op *op, double lambda);
/* The anisotropic generator's glancing-collision table, glance > 0: gHat3
 * (exec/MPIcollisionRoutines.c:33-203, the table exec/MPIWeightGenerator.c
 * writes as Weights/N%d_AnIso_L_v%g_lambda%g_glance%g.wts) with the eps-linear
 * cross section (8/pi) glance / theta^4 on [sqrt(glance), pi] and the Taylor
 * term for [0, sqrt(glance)], at every (row, column) pair, on the GPU.  The
 * reference's nested adaptive rules (tolerance 1e-6) are replaced by a reduced
 * form of the same integral, within about 1e-12 of max|W| of it.  0 < glance < pi, else
 * SPBTE_ERR_ARG. */
int spbte_op_generate_weights_glance(spbte_op *op, double lambda, double glance);
int spbte_op_get_weights(spbte_op *op, double *W);
/* Writes the FP64 table as a Weights/ file (N^3 rows of N^3 doubles, the
 * layout src/weights.c:95-104 stores). */
int spbte_op_store_weights_file(spbte_op *op, const char *path);

/* Q for ncell cells: d_Q[b] = Q(d_f[b], d_g[b]) (ComputeQ, or maxPreserve with
 * SPBTE_MAXPRESERVE), optionally projected (SPBTE_CONSERVE). */
int spbte_compute_q(spbte_op *op, const double *d_f, const double *d_g, double *d_Q, int ncell, int form, int flags);

/* Allocation-free variant of spbte_compute_q.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as spbte_compute_q (bit for bit on the
 * same tile), but every device buffer of the evaluation beyond d_f, d_g and
 * d_Q lives in a fixed workspace of SPBTE_NOALLOC_SCRATCH_BYTES per GPU,
 * held by the library from the first spbte_op_create on that GPU until the
 * process ends, and the call allocates and frees nothing.  Limits: the batch
 * must fit the workspace (about 10 cells at N=32 and 80 at N=16 for
 * maxPreserve in FP64); larger batches, and tables not yet resident in the
 * layout the batch's tile streams (spbte_compute_q and spbte_autotune build
 * them on first use), return SPBTE_ERR_ARG without launching anything.  All
 * operators on one GPU share its workspace and each call overwrites it, so
 * calls that overlap on one GPU (other threads, other streams) race: order
 * them on one stream or synchronise between them. */
#define SPBTE_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int spbte_compute_q_noalloc(spbte_op *op, const double *d_f, const double *d_g, double *d_Q, int ncell, int form,
                            int flags);

/* Single-species projection of ncell cells in place (src/conserve.c:207-264). */
int spbte_conserve(spbte_op *op, double *d_Q, int ncell);

/* General projection of one group of Ns*Ns arrays (array n*Ns+m), in place. */
int spbte_conserve_species(spbte_op *op, double *d_Q, int ngroups);
/* The same projection for a pair-major batch: pair (l, m) of cell c at
 * d_Q + ((m*Ns + l) * ncell + c) * N^3, the layout a multi-species run keeps. */
int spbte_conserve_pairs(spbte_op *op, double *d_Q, int ncell);
/* Species count and masses set by spbte_op_set_species. */
int spbte_op_get_species(spbte_op *op, int *Ns, double *masses);

/* The transform of src/collisions.c:232-283 on ncell complex cubes
 * (interleaved re/im doubles); invert 0 forward, 1 inverse.  d_in may equal d_out. */
int spbte_transform(spbte_op *op, const double *d_in, double *d_out, int ncell, int invert);

/* The weighted sum alone, on transformed inputs (interleaved complex FP64):
 * d_Qhat[b][r] = sum_c W[r][c] d_Ghat[b][c] d_Fhat[b][s(r,c)]. */
int spbte_qhat(spbte_op *op, const double *d_Fhat, const double *d_Ghat, double *d_Qhat, int ncell, int form);

/* Tiles of the weighted-sum kernels.  kind 0 is the tiled kernel (tr output rows
 * per lane, `lines` z-lines of output modes per workgroup, tb cells per
 * workgroup).  kind k >= 1 is the streaming kernel (tr rows per lane, `lines` =
 * waves per workgroup, tb cells per lane; N=32 with FP32 products also has
 * lines = 16: four groups of 4 waves on the same rows, each group its own
 * cell, so that one workgroup's W loads serve four cells) with the summed
 * x-planes split into k slices whose partial sums are added afterwards (k > 1
 * gives small batches more workgroups; k divides N, or is N times a divisor of
 * N, cutting x-planes into y-line ranges).  set_tile pins a tile for (form,
 * maxpreserve); tr = 0 restores automatic choice.  autotune times every
 * instantiated tile on the given batch shape and pins the fastest for that
 * batch size; get_tile reports the tile a batch would use. */
int spbte_set_tile(spbte_op *op, int form, int maxpreserve, int kind, int tr, int lines, int tb);
int spbte_autotune(spbte_op *op, int ncell, int form, int maxpreserve, int *kind, int *tr, int *lines, int *tb);
int spbte_get_tile(spbte_op *op, int ncell, int form, int maxpreserve, int *kind, int *tr, int *lines, int *tb);

/* Timing on the operator's stream with HIP events: runs spbte_compute_q iters
 * times and returns the mean wall time per call and the mean duration of the
 * weighted-sum kernel alone (both in milliseconds).  The batch must fit one
 * pass (spbte_compute_q's chunk) and iters is 1 to 100000; SPBTE_ERR_ARG
 * otherwise. */
int spbte_time_compute_q(spbte_op *op, const double *d_f, const double *d_g, double *d_Q, int ncell, int form, int flags,
                         int iters, double *ms_per_call, double *ms_wsum);

/* ---------------------------------------------------------------------------
 * Batched time integration (replaces the per-cell loops of exec/boltz.c for one
 * species on one rank).  A run owns the state of all cells in HBM and advances
 * it by whole steps on the operator's stream.
 *
 * 0D (exec/boltz.c:193-241): ncell independent problems; each step evaluates
 * ComputeQ_maxPreserve (or ComputeQ) + projection for all of them, then
 * f += dt Q / Kn (order 1) or Heun (order 2).
 *
 * 1D (exec/boltz.c:285-354): nX interior cells with x[], dx[] of nX + 2*order
 * entries (ghost cells included, src/mesh_setup.c).  Order 1: upwindOne
 * (src/transportroutines.c:94-238), then per cell ComputeQ + projection and
 * forward Euler.  Order 2: advectTwo (:477-492, minmod MUSCL), Heun collision
 * step, advectTwo.  Walls: SPBTE_BC_DIFFUSE = diffuse reflection at T0 (left)
 * and T1 (right) (src/boundaryConditions.c:39-85, ICChoice 3); SPBTE_BC_COPY =
 * ghosts copy their neighbours at order 1 and the no-flux wall states at
 * order 2; SPBTE_BC_PERIODIC = periodic ghosts (order 1, ICChoice 6);
 * SPBTE_BC_DIFFUSE_LEFT = diffuse reflection at T0 on the left wall only, the
 * right wall as SPBTE_BC_COPY (ICChoice 1, src/transportroutines.c:120-121,
 * :363-364).
 *
 * spbte_run_use_graph(run, 1) captures one step as a HIP graph on first use
 * and replays it; the graph is captured again whenever the operator's buffers,
 * table, species or tiles change.  The run must not outlive its operator. */
enum { SPBTE_BC_COPY = 0, SPBTE_BC_DIFFUSE = 1, SPBTE_BC_PERIODIC = 2, SPBTE_BC_DIFFUSE_LEFT = 3 };

/* Several species (exec/boltz.c:193-241, :254-354 with Ns > 1): the operator
 * holds Ns species (spbte_op_set_species) and the table of pair (l, m) in slot
 * m*Ns + l; a step evaluates Q(f_l, f_m) for every pair, projects them jointly
 * (Ns + 4 moments, src/conserve.c:207-264) and updates f_l with the sum over m
 * of Q(l, m).  Walls reflect each species with its own mass.  The state of the
 * run (set_state, get_state, moments) is species-major: species l's cells
 * follow species l-1's. */

typedef struct spbte_run spbte_run;

int spbte_run0d_create(spbte_run **run, spbte_op *op, int ncell, int order, double dt, double Kn, int form,
                       int maxpreserve);
int spbte_run1d_create(spbte_run **run, spbte_op *op, int nX, int order, const double *x, const double *dx, double dt,
                       double Kn, int bc, double T0, double T1, int form);
int spbte_run0d_create_species(spbte_run **run, spbte_op *op, int Ns, int ncell, int order, double dt, double Kn,
                               int form, int maxpreserve);
int spbte_run1d_create_species(spbte_run **run, spbte_op *op, int Ns, int nX, int order, const double *x,
                               const double *dx, double dt, double Kn, int bc, double T0, double T1, int form);
void spbte_run_destroy(spbte_run *run);
/* Interior cells, Ns * ncell * N^3 host doubles (synchronous). */
int spbte_run_set_state(spbte_run *run, const double *f);
int spbte_run_get_state(spbte_run *run, double *f);
/* Device pointer to the interior cells of the state (species 0, or species l). */
int spbte_run_state(spbte_run *run, double **d_f);
int spbte_run_species_state(spbte_run *run, int l, double **d_f);
int spbte_run_use_graph(spbte_run *run, int on);
/* Queue nsteps steps (asynchronous; spbte_op_synchronize waits). */
int spbte_run_step(spbte_run *run, int nsteps);
/* Poiseuille body force of magnitude Ma along v_y in every transport stage of
 * a 1D order-2 run (ICChoice 5, src/transportroutines.c:429-437, :458-466);
 * Ma = 0 turns it off.  Order-1 runs are refused: the reference applies that
 * term after its k loop, outside the cell (:218-226). */
int spbte_run_set_forcing(spbte_run *run, double Ma);
/* Per species and cell rho, u_x, u_y, u_z, T with unit mass
 * (src/momentRoutines.c:58-183): 5 * Ns * ncell host doubles. */
int spbte_run_moments(spbte_run *run, double *out);

/* Moment output in the reference's format (src/output.c:64-213, :215-404):
 * a header line, then per spbte_run_write_output call one "%le "-column line
 * per cell (0D: time, density, velocity_x, temperature, pressure, positive /
 * negative energy ratio, slice f[N/2 + N(N/2 + N l)]; 1D: time, position,
 * density, velocity_x, temperature, pressure).  flags select the columns as
 * the reference's output-flags file does (read_flags, src/output.c:417-443);
 * spbte_read_output_flags parses such a file.  One open file per run; append
 * opens an existing file for a restarted run (the header line is written in
 * both cases, as the reference does). */
enum {
  SPBTE_OUT_DENSITY = 1,
  SPBTE_OUT_VELOCITY = 2,
  SPBTE_OUT_TEMPERATURE = 4,
  SPBTE_OUT_PRESSURE = 8,
  SPBTE_OUT_ENTROPY = 16, /* 0D: positive / negative energy ratio */
  SPBTE_OUT_SLICE = 32    /* 0D: the slice f[N/2 + N(N/2 + N l)], l = 0..N-1 */
};
int spbte_read_output_flags(const char *path, int *flags);
int spbte_run_open_output(spbte_run *run, const char *path, int flags, int append);
/* The file of species l of a multi-species run (src/output.c:138-157 names
 * them moments_<input>_<species name>); write_output writes every open file
 * of the run and close_output closes them all. */
int spbte_run_open_output_species(spbte_run *run, int l, const char *path, int flags, int append);
int spbte_run_write_output(spbte_run *run, double time);
int spbte_run_close_output(spbte_run *run);

/* Restart dumps (src/restart.c:24-109): species 0's interior cells' N^3
 * doubles back to back in pdf_path, and the step counter as one int in
 * time_path (time_path may be null). */
int spbte_run_store_restart(spbte_run *run, const char *pdf_path, const char *time_path, int step);
int spbte_run_load_restart(spbte_run *run, const char *pdf_path, const char *time_path, int *step);
/* The same for species l's cells only (one file per species, src/restart.c:41-54). */
int spbte_run_store_restart_species(spbte_run *run, int l, const char *pdf_path);
int spbte_run_load_restart_species(spbte_run *run, int l, const char *pdf_path);

/* One transport stage alone, synchronous: order 1 = upwindOne, order 2 =
 * advectTwo, from d_f to d_fc, both (nX + 2*order) * N^3 device doubles with
 * ghost cells; d_f's ghost cells are overwritten as in the reference. */
int spbte_advect(spbte_op *op, double *d_f, double *d_fc, int nX, int order, const double *x, const double *dx,
                 double dt, int bc, double T0, double T1);

/* ---------------------------------------------------------------------------
 * Learned collision surrogate (no counterpart in the reference).  A spectral
 * mixing network maps the block of Kb^3 lowest transform modes of a cell to
 * the same block of Qhat (features (T(f - M_f), T(M_f)) / rho; output
 * Qhat / rho^2); its prediction is zero-padded, inverse-transformed and passed
 * through the FP64 projection.  Layers: low-rank mode mixing H += U (V^T H)
 * (rank `rank`) and a GELU channel MLP of `width` -> 2 width -> `width`
 * channels.  GEMMs run on BF16 MFMA with FP32 accumulation; parameters, Adam
 * moments and activations are FP32.  Targets are the operator's exact Qhat
 * (ComputeQ, or the fused maxPreserve sum when `maxpreserve`), computed with
 * the FP64 table the operator holds.  Loss: mean over cells of
 * |y - t|^2 / |t|^2 on the block.  The network shares the operator's stream
 * and must not outlive it. */
typedef struct spbte_net spbte_net;

int spbte_net_create(spbte_net **net, spbte_op *op, int Kb, int width, int rank, int layers, int maxpreserve,
                     unsigned long long seed);
void spbte_net_destroy(spbte_net *net);
long spbte_net_num_params(spbte_net *net);
/* offsets of P, pb, E, {U, V, W1, b1, W2, b2} per layer, O, ob in the parameter vector; returns their count */
int spbte_net_layout(spbte_net *net, long *offsets);
int spbte_net_get_params(spbte_net *net, float *params);
int spbte_net_set_params(spbte_net *net, const float *params);
int spbte_net_set_scales(spbte_net *net, double fscale, double tscale);
int spbte_net_get_scales(spbte_net *net, double *fscale, double *tscale);
/* Feature and target scales from ncell cells (d_Q: ncell * N^3 device scratch). */
int spbte_net_calibrate(spbte_net *net, const double *d_f, double *d_Q, int ncell);
/* Training set in HBM: reserve, then append (features, exact targets) of cells. */
int spbte_net_dataset_reserve(spbte_net *net, long nsamples);
int spbte_net_dataset_add(spbte_net *net, const double *d_f, double *d_Q, int ncell);
long spbte_net_dataset_size(spbte_net *net);
/* nsteps Adam steps on shuffled batches of the dataset; losses[t] if non-null. */
int spbte_net_train(spbte_net *net, int nsteps, int batch, float lr, unsigned long long seed, float *losses);
/* Loss and gradient (host copy) of dataset samples idx[0..B), no update. */
int spbte_net_grad(spbte_net *net, const int *idx, int B, float *grad, double *loss);
/* Relative L2 error of the block prediction for dataset samples. */
int spbte_net_eval(spbte_net *net, const int *idx, int B, float *err);
/* Network alone on host features [m][b][4] -> host outputs [m][b][2]; and the
 * loss gradient for host features and targets [m][b][2]. */
int spbte_net_forward_features(spbte_net *net, const float *X, int B, float *Y);
int spbte_net_grad_features(spbte_net *net, const float *X, const float *T, int B, float *grad, double *loss);
/* Surrogate Q of ncell cells (device arrays), optionally projected. */
int spbte_net_predict(spbte_net *net, const double *d_f, double *d_Q, int ncell, int conserve);
/* Surrogate-assisted Q with a sampled full-spectral check, per group: the
 * cells are taken in groups of check_every consecutive cells; one cell of each
 * group (the one at `phase` within it, or its last cell if the group is
 * shorter) is also evaluated exactly (precision form `form`), keeps the exact
 * Q and judges its group: if the surrogate's error there exceeds tol, every
 * cell of the group falls back to the exact operator, otherwise the group's
 * other cells keep the surrogate's Q.  Q is projected in every case.  The
 * error of a checked cell is |Q_net - Q| / max(|Q|, floor_rel * max over
 * checked cells of |Q|): floor_rel = 0 is the plain relative L2 error, which
 * equilibrium cells (whose exact Q is only the discretisation residual) fail
 * by construction.  n_exact: cells evaluated exactly (checked cells
 * included); max_err: largest checked error.  Synchronous. */
int spbte_net_collide(spbte_net *net, const double *d_f, double *d_Q, int ncell, int form, int check_every, int phase,
                      double tol, double floor_rel, int *n_exact, double *max_err);
/* Loss denominator floor: each cell's loss is |y - t|^2 / (|t|^2 + 2 Kb^3 rms_floor^2)
 * in the calibrated target units (default rms_floor 0.1). */
int spbte_net_set_loss_floor(spbte_net *net, double rms_floor);
/* Runs whose collision stages use spbte_net_collide (the checked cell of each
 * group rotates by one each stage); the statistics count cells, exactly
 * evaluated cells, groups that fell back, and the largest sampled error.  Such
 * runs do not use graphs. */
int spbte_run_set_surrogate(spbte_run *run, spbte_net *net, int check_every, double tol, double floor_rel);
int spbte_run_surrogate_stats(spbte_run *run, long *cells, long *exact, long *fallbacks, double *max_err);

#ifdef __cplusplus
}
#endif

#endif
