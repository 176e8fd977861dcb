/* spbte_internal.h -- declarations shared by the host C layer and the HIP
 * kernel launchers of libspectralbte_amd.  Plain C types only. */
#ifndef SPBTE_INTERNAL_H
#define SPBTE_INTERNAL_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPBTE_MAX_N 64
#define SPBTE_MAX_NS 8

/* Per-grid device tables for the transform, moments and projection. */
typedef struct {
  int N;
  double Lv, hv, heta;
  double pre_fwd, pre_inv; /* (2 pi)^-3/2 h^3 of each direction */
  double mass, kb;         /* Maxwellian of the maxPreserve split */
  void *d_mat_fwd;         /* double2[N*N]: per-axis forward matrix */
  void *d_mat_inv;         /* double2[N*N]: per-axis inverse matrix */
  void *d_v;               /* double[N] */
  void *d_w;               /* double[N] trapezoid weights */
  void *d_fft;             /* double2[6N]: forward p, q, twiddles, then inverse p, q, twiddles
                              (FFT factors of the per-axis matrices), or NULL: matrix form */
} spbte_grid_dev;

/* Moments of one cell used by the Maxwellian split:
 * rho, u_x, u_y, u_z, T, prefactor rho (0.5/(pi T))^1.5. */
#define SPBTE_MOM_STRIDE 6

/* One input of the batched forward transform. mode: 0 = src as is,
 * 1 = src - M(mom), 2 = M(mom) only (src unused).  Real input, N^3 per cell. */
typedef struct {
  const double *src;
  const double *mom; /* SPBTE_MOM_STRIDE per cell, or NULL for mode 0 */
  int mode;
  void *dst;         /* complex output, double2 or float2 per element */
  const double *part; /* instead of mom (FFT grids): the per-plane moment partial
                         sums of spbte_launch_moment_partials, 5 N per cell */
  const double *msrc; /* instead of mom and part (grids where
                         spbte_fft_moments_inline holds): the cells whose
                         moments the transform takes itself */
} spbte_xform_job;

/* Weighted-sum kernel arguments.  F/G arrays hold ncell * N^3 complex values:
 * double2 for the FP64-arithmetic forms; for the FP32-product form float2,
 * each array followed by the ncell * N^3 sums of its values' parts (x + y),
 * which its two-product streaming tiles read for G.  Qh gets ncell * N^3
 * double2. */
typedef struct {
  const void *W;      /* N^3 x N^3: double, float, or (hi, lo) half pairs */
  const void *F0, *G0, *F1, *G1;
  void *Qh;
  int ncell;
  int nprod;          /* 1: ComputeQ, 2: fused maxPreserve */
  int keep_slices;    /* split streaming tiles: leave the slices' partial sums in
                         Qp for the consumer to add (Qh is then not written) */
} spbte_wsum_args;

/* Tile of a weighted-sum kernel.  kind 0: k_wsum.hip (lines = z-lines of output
 * modes per workgroup); kind 1: k_wsum_stream.hip (lines = waves per workgroup). */
typedef struct {
  int tr;    /* output rows (consecutive z modes) per lane */
  int lines; /* z-lines (kind 0) or waves (kind >= 1) per workgroup */
  int tb;    /* cells per workgroup */
  int kind;  /* 0: tiled kernel; k >= 1: streaming kernel with the summed x-planes in k slices */
} spbte_tile;

/* ----- launchers (k_*.hip); all return hipError_t of the launch ----- */
/* part: NULL (one workgroup per cell) or 5 * N * ncell doubles of scratch
 * (one workgroup per x-plane and cell, then a combining pass); the same for
 * the one-species projection. */
int spbte_launch_moments(const spbte_grid_dev *g, const double *f, double *mom, int ncell, double *part,
                         hipStream_t s);
int spbte_fft_supported(int N);
int spbte_launch_xform_fwd(const spbte_grid_dev *g, const spbte_xform_job *jobs, int njobs, int ncell, int out_fp32,
                           void *scratch, hipStream_t s);
int spbte_launch_xform_fwd_complex(const spbte_grid_dev *g, const void *in, void *out, int ncell, int invert,
                                   void *scratch, hipStream_t s);
/* Qh: nsl sets of ncell * N^3 double2, ncell * N^3 apart, whose sum is
 * transformed (nsl > 1 on FFT grids only: the split weighted sum's slices; Qh
 * may then alias scratch).  proj_part (FFT grids only, else NULL): also writes
 * the one-species projection's per-plane partial sums of q (5 N per cell) for
 * spbte_launch_project_apply. */
/* 1 if forward transform jobs on an N grid may give msrc for their moments. */
int spbte_fft_moments_inline(int N);
/* 1 if spbte_launch_xform_inv_real_cube serves ncubes cells of an N grid
 * (one workgroup per cube for the last pass). */
int spbte_fft_cube(int N, int ncubes);
/* An elementwise update fused into an evaluation's last pass: out = a x +
 * b y + c Q over the evaluated cells (y may be NULL; out may be x). */
typedef struct {
  double a, b, c;
  const double *x, *y;
  double *out;
} spbte_epilogue;
/* spbte_compute_q (f alone, as g) with the update epi fused into its last pass
 * where the grid and batch allow it (*fused = 1); otherwise the plain
 * evaluation, and the caller applies the update (*fused = 0). */
struct spbte_op;
int spbte_compute_q_update(struct spbte_op *op, const double *d_f, double *d_Q, int ncell, int form, int flags,
                           const spbte_epilogue *epi, int *fused);
/* Inverse transform of the sum of nsl sets to the real part on whole-cube
 * grids, projected per cell (one species) when masses is not NULL, then the
 * update epi when it is not NULL. */
int spbte_launch_xform_inv_real_cube(const spbte_grid_dev *g, const void *Qh, int nsl, double *q, int ncell,
                                     void *scratch, const double *masses, const double *LU, const int *pivot,
                                     const spbte_epilogue *epi, hipStream_t s);
int spbte_launch_xform_inv_real(const spbte_grid_dev *g, const void *Qh, int nsl, double *q, int ncell, void *scratch,
                                double *proj_part, hipStream_t s);
/* The first pass of spbte_launch_moments alone (part: 5 N per cell), whose
 * partials FFT transform jobs read in place of the moments. */
int spbte_launch_moment_partials(const spbte_grid_dev *g, const double *f, int ncell, double *part, hipStream_t s);
/* The second pass of the one-species projection, on partials written by the
 * inverse transform. */
int spbte_launch_project_apply(const spbte_grid_dev *g, const double *masses, const double *LU, const int *pivot,
                               double *Q, int ncell, const double *part, hipStream_t s);
int spbte_launch_wsum(int N, int form, const spbte_tile *tile, const spbte_wsum_args *a, hipStream_t s);
int spbte_launch_wsum_scaled(int N, int form, const spbte_tile *tile, const spbte_wsum_args *a, const float *rowscale,
                             hipStream_t s);
/* FP32 operands: n float2 values at out, then the n sums of their parts. */
int spbte_launch_c64_to_c32(const void *in, void *out, size_t n, hipStream_t s);
int spbte_launch_copy(const void *in, void *out, size_t bytes, hipStream_t s);
/* n values of a fixed pseudo-random pattern in [-1, 1): float if single, else double */
int spbte_launch_fill_pattern(void *p, size_t n, int single, hipStream_t s);
int spbte_wsum_tile_supported(int N, int form, int nprod, const spbte_tile *tile);
int spbte_stream_tile_supported(int N, int form, int nprod, int tr, int waves, int tb);
/* 1 if the streaming kernel can split the summed modes into csplit slices:
 * a divisor of N (x-planes) or N times a divisor of N (y-line ranges). */
int spbte_csplit_valid(int N, int csplit);
int spbte_launch_wsum_stream(int N, int form, int tr, int waves, int tb, const spbte_wsum_args *a, const void *Wp,
                             const float *rowscale, int csplit, void *Qp, hipStream_t s);
int spbte_launch_permute_weights(const double *W64, void *dst, const float *rowscale, int form, int TR, size_t n3,
                                 hipStream_t s);
/* Projection of ncell groups of Ns*Ns arrays (group-major).  part: (Ns+4) N
 * doubles per group of scratch for the split (x-plane, group) kernels, or
 * NULL for one workgroup per group. */
int spbte_launch_project(const spbte_grid_dev *g, int Ns, const double *masses, const double *LU, const int *pivot,
                         double *Q, int ncell, double *part, hipStream_t s);
int spbte_launch_project_strided(const spbte_grid_dev *g, int Ns, const double *masses, const double *LU, const int *pivot,
                                 double *Q, int ngroups, size_t gstride, size_t pstride, double *part, hipStream_t s);
int spbte_launch_convert_weights(const double *W64, void *dst, int form, size_t n_rows, size_t n_cols, hipStream_t s);
int spbte_launch_weights_iso(const spbte_grid_dev *g, const double *eta, double lam, double dfac, double mu,
                             double *W, hipStream_t s);
int spbte_launch_weights_landau(const spbte_grid_dev *g, const double *eta, double lam, double *W, hipStream_t s);
int spbte_launch_weights_glance(const spbte_grid_dev *g, const double *eta_host, const double *eta, double lam,
                                double glance, double *W, hipStream_t s);

/* transport and time-integration kernels (k_transport.hip) */
int spbte_launch_upwind1(int N, int nX, const double *f, double *fc, const double *v, const double *dx, double dt,
                         hipStream_t s);
int spbte_launch_ghosts(int N, int nX, int mode, double *f, hipStream_t s);
int spbte_launch_diffuse(const spbte_grid_dev *g, int nwall, const double *const *src, double *const *dst,
                         const double *TW, const int *right, hipStream_t s);
int spbte_launch_walls2(int N, int nX, int diffuse, const double *f, double *fl, double *fr, const double *x,
                        const double *dx, hipStream_t s);
int spbte_launch_upwind2(int N, int nX, const double *f, double *fc, const double *fl, const double *fr,
                         const double *v, const double *x, const double *dx, double dt, double fy, hipStream_t s);
int spbte_launch_lincomb(size_t n, double a, const double *x, double b, const double *y, double c, const double *z,
                         double *out, hipStream_t s);
int spbte_launch_out_quantities(const spbte_grid_dev *g, const double *f, int ncell, double *out, hipStream_t s);
/* run internals used by the output module (spbte_run.c) */
struct spbte_run;
int spbte_run_nspecies(struct spbte_run *run);
int spbte_run_geometry(struct spbte_run *run, struct spbte_op **op, int *dim, int *ncell, const double **d_state,
                       const double **d_x);

/* learned surrogate kernels (k_net.hip) */
int spbte_net_gemm(const float *A, long sa0, long sa1, const float *B, long sb0, long sb1, float *C, long sc0, long sc1,
                   int M, int N, int K, const float *add, const float *bias, float *G, float *work, size_t work_floats,
                   hipStream_t s);
int spbte_net_gelu_bwd(long n, const float *Z, const float *dG, float *dZ, hipStream_t s);
int spbte_net_colsum(long rows, int cols, const float *X, float *out, int accumulate, float *work, size_t work_floats,
                     hipStream_t s);
int spbte_net_features(int N, int Kb, int ncell, const void *Fd, const void *Fm, const double *mom, double fscale,
                       float *X, hipStream_t s);
int spbte_net_targets(int N, int Kb, int ncell, const void *Qh, const double *mom, double tscale, float *T,
                      hipStream_t s);
int spbte_net_scatter(int N, int Kb, int ncell, const float *Y, const double *mom, double tscale, void *Qh,
                      hipStream_t s);
int spbte_net_lift(long M, int ncell, int C, const float *X, const float *P, const float *E, const float *pb, float *H,
                   hipStream_t s);
int spbte_net_lift_bwd(long M, int ncell, int C, const float *X, const float *dH, float *dE, float *dP, float *dpb,
                       float *work, size_t work_floats, hipStream_t s);
int spbte_net_head(long R, int C, const float *H, const float *O, const float *ob, float *Y, hipStream_t s);
int spbte_net_head_bwd(long R, int C, const float *H, const float *O, const float *dY, float *dH, float *dO, float *dob,
                       float *work, size_t work_floats, hipStream_t s);
int spbte_net_loss(long M, int ncell, const float *Y, const float *T, float *dY, float *loss_b, float eps,
                   hipStream_t s);
int spbte_net_adam(long n, float *p, const float *g, float *m1, float *m2, float lr, float b1, float b2, float eps,
                   int step, float wd, hipStream_t s);
int spbte_net_axpy(long n, float a, const float *x, float *y, hipStream_t s);
int spbte_net_cell_relerr(long n, int ncell, const double *a, const double *b, double *out, hipStream_t s);
int spbte_net_cell_norm(long n, int ncell, const double *a, double *out, hipStream_t s);
/* spbte_net_collide with the count of groups that fell back (spbte_net.c). */
struct spbte_net;
int spbte_net_collide_groups(struct spbte_net *n, const double *d_f, double *d_Q, int ncell, int form, int check_every,
                             int phase, double tol, double floor_rel, int *n_exact, int *n_failed, double *max_err);
int spbte_net_move_cells(long n, int ncell, const int *idx, const double *src, double *dst, int scatter,
                         hipStream_t s);
int spbte_net_relayout(long M, int nsamp, int width, const float *src, float *dst, int to_samples, const int *idx,
                       hipStream_t s);

/* k_scratch.hip: device address and size of the current GPU's fixed workspace
 * of spbte_compute_q_noalloc (resolved once per device). */
int spbte_noalloc_scratch(void **ptr, size_t *bytes);

/* host helpers shared by the reference-compatible modules (spbte_core.c) */
int spbte_host_lu_factor(int n, double *A, int *pivot);
void spbte_host_lu_solve(int n, const double *LU, const int *pivot, double *b);
void spbte_host_cct(int N, const double *v, int Ns, const double *masses, double *CCt);
struct spbte_op;
hipStream_t spbte_op_stream(struct spbte_op *op);
const spbte_grid_dev *spbte_op_grid(struct spbte_op *op);
int spbte_op_get_weights_rows(struct spbte_op *op, double **rows);
unsigned spbte_op_generation(struct spbte_op *op);
const void *spbte_op_qhat_buffer(struct spbte_op *op);
int spbte_op_chunk_cells(struct spbte_op *op);
int spbte_op_n(struct spbte_op *op);
int spbte_op_table_slot(struct spbte_op *op);
int spbte_fail(int code, const char *fmt, ...);

/* collisions.c: the collisions.h module's device table cache.  Forget drops
 * the copy made from a conv_weights pointer (its rows were freed or
 * rewritten); set_mass passes the species mass that conserve.h receives. */
void spbte_abi_forget_weights(double **conv_weights);
void spbte_abi_set_mass(double mass);

#ifdef __cplusplus
}
#endif

#endif
