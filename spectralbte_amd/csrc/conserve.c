/* conserve.c -- the reference's conserve.h interface over the MI355X projection.
 *
 * Replaces src/conserve.c.  C C^T (Ns+4 square) is built and LU-factored on
 * the host in FP64 once (src/conserve.c:268-317, :89-168), as the reference
 * does; conserveAllMoments uploads the Ns*Ns operator arrays, runs the moment
 * reduction, the solve and the correction on the GPU in FP64 (k_proj_part and
 * k_proj_apply in k_transform.hip, one workgroup per x-plane) and downloads the
 * result.  The arrays pass through a pinned host buffer, so each direction is
 * one asynchronous copy and the call waits once.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "conserve.h"
#include "spbte.h"
#include "spbte_internal.h"

static int N = 0, Ns = 0;
static double dv;
static double *v;       /* borrowed */
static double *CCt;     /* (Ns+4)^2, LU factors after createCCtAndPivot */
static int *pivot;
static double masses[SPBTE_MAX_NS];
static spbte_grid_dev grid; /* device copies of v and the trapezoid weights */
static double *d_Q;
static double *d_part;  /* (Ns+4) N per-plane partial sums */
static double *h_Q;     /* pinned staging, Ns*Ns N^3 */
static hipStream_t stream;

static void die(const char *where, hipError_t e) {
  fprintf(stderr, "spectralbte_amd: %s: %s\n", where, hipGetErrorString(e));
  fflush(stderr);
  exit(1);
}

static void setup(int nodes, double h_v, double *vel, int numspec, const double *m) {
  if (N) dealloc_conservation();
  if (numspec < 1 || numspec > SPBTE_MAX_NS) {
    fprintf(stderr, "spectralbte_amd: initialize_conservation: %d species (at most %d)\n", numspec, SPBTE_MAX_NS);
    exit(1);
  }
  N = nodes;
  dv = h_v;
  v = vel;
  Ns = numspec;
  for (int i = 0; i < Ns; i++) masses[i] = m[i];
  CCt = calloc((size_t)(Ns + 4) * (Ns + 4), sizeof(double));
  pivot = calloc((size_t)(Ns + 4), sizeof(int));
  double w[SPBTE_MAX_N];
  for (int i = 0; i < N; i++) w[i] = (i == 0 || i == N - 1) ? 0.5 : 1.0;
  memset(&grid, 0, sizeof grid);
  grid.N = N;
  grid.hv = dv;
  hipError_t e;
  if ((e = hipStreamCreate(&stream)) != hipSuccess) die("initialize_conservation", e);
  if ((e = hipMalloc(&grid.d_v, sizeof(double) * N)) != hipSuccess) die("initialize_conservation", e);
  if ((e = hipMalloc(&grid.d_w, sizeof(double) * N)) != hipSuccess) die("initialize_conservation", e);
  if ((e = hipMemcpy(grid.d_v, v, sizeof(double) * N, hipMemcpyHostToDevice)) != hipSuccess) die("initialize_conservation", e);
  if ((e = hipMemcpy(grid.d_w, w, sizeof(double) * N, hipMemcpyHostToDevice)) != hipSuccess) die("initialize_conservation", e);
  const size_t qbytes = sizeof(double) * Ns * Ns * (size_t)N * N * N;
  if ((e = hipMalloc((void **)&d_Q, qbytes)) != hipSuccess) die("initialize_conservation", e);
  if ((e = hipMalloc((void **)&d_part, sizeof(double) * (Ns + 4) * (size_t)N)) != hipSuccess)
    die("initialize_conservation", e);
  if ((e = hipHostMalloc((void **)&h_Q, qbytes, hipHostMallocDefault)) != hipSuccess) die("initialize_conservation", e);
  spbte_abi_set_mass(masses[0]);
  createCCtAndPivot();
}

void initialize_conservation(int nodes, double h_v, double *vel, species *mix, int numspec) {
  double m[SPBTE_MAX_NS];
  for (int i = 0; i < numspec && i < SPBTE_MAX_NS; i++) m[i] = mix[i].mass;
  setup(nodes, h_v, vel, numspec, m);
}

void initialize_conservation_fast(int nodes, double h_v, double *vel) {
  double one = 1.0;
  setup(nodes, h_v, vel, 1, &one);
}

void dealloc_conservation(void) {
  if (!N) return;
  free(CCt);
  free(pivot);
  CCt = NULL;
  pivot = NULL;
  hipFree(grid.d_v);
  hipFree(grid.d_w);
  hipFree(d_Q);
  hipFree(d_part);
  hipHostFree(h_Q);
  hipStreamDestroy(stream);
  d_Q = NULL;
  d_part = NULL;
  h_Q = NULL;
  N = 0;
}

/* C with rows w h_v^3 m {delta, v, |v|^2/2}: CCt = sum_m C_m C_m^T
 * (src/conserve.c:268-305), then the LU of :89-168.  The reference builds C
 * with dv = h_v; v[1]-v[0] equals it for the callers. */
void createCCtAndPivot(void) {
  double *vv = malloc(sizeof(double) * N);
  memcpy(vv, v, sizeof(double) * N);
  /* spbte_host_cct takes its spacing from v; rebuild v with spacing dv only
   * if the caller's h_v differs from v[1]-v[0] (it never does in the reference). */
  if (vv[1] - vv[0] != dv)
    for (int i = 1; i < N; i++) vv[i] = vv[0] + i * dv;
  spbte_host_cct(N, vv, Ns, masses, CCt);
  free(vv);
  factorMatrixIntoLU(Ns + 4, 1.0);
}

void factorMatrixIntoLU(int nElem, double det) {
  (void)det; /* passed by value and unused, as in the reference */
  if (spbte_host_lu_factor(nElem, CCt, pivot) != 0) {
    puts("Exiting factorMatrixintoLU....");
    exit(0);
  }
}

void solveWithCCt(int nElem, double *b) { spbte_host_lu_solve(nElem, CCt, pivot, b); }

void conserveAllMoments(double **Q) {
  if (!N) {
    fprintf(stderr, "spectralbte_amd: conserveAllMoments called before initialize_conservation\n");
    exit(1);
  }
  const size_t n3 = (size_t)N * N * N, na = (size_t)Ns * Ns;
  hipError_t e;
  for (size_t a = 0; a < na; a++) memcpy(h_Q + a * n3, Q[a], sizeof(double) * n3);
  if ((e = hipMemcpyAsync(d_Q, h_Q, sizeof(double) * na * n3, hipMemcpyHostToDevice, stream)) != hipSuccess)
    die("conserveAllMoments", e);
  int le = spbte_launch_project(&grid, Ns, masses, CCt, pivot, d_Q, 1, d_part, stream);
  if (le != 0) die("conserveAllMoments", (hipError_t)le);
  if ((e = hipMemcpyAsync(h_Q, d_Q, sizeof(double) * na * n3, hipMemcpyDeviceToHost, stream)) != hipSuccess)
    die("conserveAllMoments", e);
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) die("conserveAllMoments", e);
  for (size_t a = 0; a < na; a++) memcpy(Q[a], h_Q + a * n3, sizeof(double) * n3);
}
