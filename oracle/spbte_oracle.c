/*
 * spbte_oracle.c -- CPU restatement of SpectralBTE's collision-operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
 * smoke() entry point and bench.py's CPU-baseline leg may load it.  The
 * product library (libspectralbte_amd.so) never links or calls it.
 *
 * Everything here is FP64 (long double where a transform accumulates), written
 * from the mathematics of the reference's C sources; every function cites the
 * file:line of the reference it restates (paths relative to the reference repo).
 *
 * Array conventions (same as the reference):
 *   - velocity / Fourier multi-index (i,j,k) -> linear k + N*(j + N*i), i slowest
 *     (src/collisions.c:100-102);
 *   - complex arrays are interleaved (re, im) doubles, like fftw_complex;
 *   - the weight table W is N^3 rows x N^3 columns, row = output mode
 *     ("zeta" in src/collisions.c:128), column = summed mode ("xi", :136),
 *     stored row after row (src/weights.c:81-82).
 *
 * Pinning (tests/test_oracle_*.py):
 *   - orc_weights_iso: both .wts fixtures of the reference tests;
 *   - the 0D pipeline (orc_compute_q_maxpreserve + orc_conserve + Heun):
 *     tests/BKW8/target/moments_BKW8.test.in;
 *   - the 1D order-1 pipeline (orc_upwind_one + diffuse walls + orc_compute_q
 *     + orc_conserve + Euler): tests/heat_transport/target/moments_heat_transport.test.in;
 *   - orc_fft3d: numpy.fft with the explicit twiddles, and Gaussian closed form;
 *   - orc_qhat: the convolution theorem (W == 1) and single-column W;
 *   - orc_conserve: C.Q == 0, idempotence, numpy least-squares projection;
 *   - orc_glance_entry: zero row mode, the xi = 0 and xi || zeta reductions
 *     (scipy quadrature of the plane-wave identity), and the psi-average form
 *     plus Lam/4 x orc_landau_entry for generic entries;
 *   - orc_upwind_two / orc_advect_two: no reference fixture exists; pinned by
 *     what the scheme must satisfy: exact translation of data linear in x on a
 *     non-uniform mesh, total variation and range of step data not increased,
 *     second-order convergence on smooth data, the body-force momentum
 *     identity, and the one-wall case composed of the two-wall and copy cases.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define IDX3(N, i, j, k) ((k) + (N) * ((j) + (N) * (i)))

/* ------------------------------------------------------------------ grids */

/* src/initializer.c:66-82 (0D grid): h_v = 2L_v/(N-1), v_i = -L_v + i h_v,
 * h_eta = (2 pi / N)/h_v, L_eta = N h_eta / 2 (even N) or (N-1) h_eta / 2. */
void orc_grid_0d(int N, double Lv, double *v, double *eta) {
  double dv = 2 * Lv / (N - 1);
  for (int i = 0; i < N; i++) v[i] = -Lv + i * dv;
  double deta = (2 * M_PI / N) / dv;
  double Leta = (N % 2 == 0) ? 0.5 * N * deta : 0.5 * (N - 1) * deta;
  for (int i = 0; i < N; i++) eta[i] = -Leta + i * deta;
}

/* src/initializer.c:256-266 (1D grid): L_eta = (N-1) pi / (2 L_v),
 * h_eta = pi (N-1) / (N L_v). */
void orc_grid_1d(int N, double Lv, double *v, double *eta) {
  double dv = 2 * Lv / (N - 1);
  for (int i = 0; i < N; i++) v[i] = -Lv + i * dv;
  double Leta = 0.5 * (N - 1) * M_PI / Lv;
  double deta = M_PI * (N - 1) / (N * Lv);
  for (int i = 0; i < N; i++) eta[i] = -Leta + i * deta;
}

/* src/collisions.c:48-54: trapezoid end weights 1/2. */
static void trap_weights(int N, double *w) {
  for (int i = 0; i < N; i++) w[i] = 1.0;
  w[0] = 0.5;
  w[N - 1] = 0.5;
}

/* ------------------------------------------------------------- transform */

/* src/collisions.c:232-283 (fft3D).  The reference multiplies by a pre-twiddle
 * exp(i s (j1+j2+j3) L_start delta) and the trapezoid weights, runs an
 * unnormalised FFTW transform (FORWARD exp(-2 pi i jk/N) when invert==0,
 * BACKWARD exp(+2 pi i jk/N) when invert==1), then a post-twiddle
 * exp(i s L_end (a_k1+a_k2+a_k3)).  All three factors are separable, so the
 * oracle applies, per axis, the N x N matrix
 *   M[k][j] = exp(i s L_end a_k) exp(-+2 pi i jk/N) exp(i s j L_start delta) w_j
 * as a direct DFT accumulated in long double, times (2 pi)^-3/2 delta^3.
 * Forward: delta = h_v, L_start = L_eta = -eta[0], L_end = L_v = -v[0], a = eta, s = +1.
 * Inverse: delta = h_eta, L_start = L_v, L_end = L_eta, a = v, s = -1.
 * `in` and `out` may alias (as in the reference). */
void orc_fft3d(int N, const double *v, const double *eta, const double *in, double *out, int invert) {
  double w[64];
  if (N > 64) return;
  trap_weights(N, w);
  double dv = v[1] - v[0], deta = eta[1] - eta[0];
  double Lv = -v[0], Leta = -eta[0];
  double delta, Lstart, Lend, sign;
  const double *a;
  if (!invert) { delta = dv; Lstart = Leta; Lend = Lv; a = eta; sign = 1.0; }
  else { delta = deta; Lstart = Lv; Lend = Leta; a = v; sign = -1.0; }
  double dsgn = invert ? 1.0 : -1.0; /* FFTW sign */
  long double *Mr = malloc(sizeof(long double) * N * N), *Mi = malloc(sizeof(long double) * N * N);
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++) {
      long double ph = (long double)sign * Lend * a[k] + (long double)dsgn * 2.0L * (long double)M_PI * (long double)((j * k) % N) / N +
                       (long double)sign * j * Lstart * delta;
      Mr[k * N + j] = cosl(ph) * w[j];
      Mi[k * N + j] = sinl(ph) * w[j];
    }
  size_t n3 = (size_t)N * N * N;
  long double *A = malloc(sizeof(long double) * 2 * n3), *B = malloc(sizeof(long double) * 2 * n3);
  for (size_t t = 0; t < n3; t++) { A[2 * t] = in[2 * t]; A[2 * t + 1] = in[2 * t + 1]; }
  /* axis k (fastest) */
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int kk = 0; kk < N; kk++) {
        long double sr = 0, si = 0;
        for (int k = 0; k < N; k++) {
          size_t s = IDX3(N, i, j, k);
          long double xr = A[2 * s], xi = A[2 * s + 1];
          sr += Mr[kk * N + k] * xr - Mi[kk * N + k] * xi;
          si += Mr[kk * N + k] * xi + Mi[kk * N + k] * xr;
        }
        size_t d = IDX3(N, i, j, kk);
        B[2 * d] = sr; B[2 * d + 1] = si;
      }
  /* axis j */
  for (int i = 0; i < N; i++)
    for (int jj = 0; jj < N; jj++)
      for (int k = 0; k < N; k++) {
        long double sr = 0, si = 0;
        for (int j = 0; j < N; j++) {
          size_t s = IDX3(N, i, j, k);
          long double xr = B[2 * s], xi = B[2 * s + 1];
          sr += Mr[jj * N + j] * xr - Mi[jj * N + j] * xi;
          si += Mr[jj * N + j] * xi + Mi[jj * N + j] * xr;
        }
        size_t d = IDX3(N, i, jj, k);
        A[2 * d] = sr; A[2 * d + 1] = si;
      }
  /* axis i (slowest) */
  long double pre = powl(1.0L / sqrtl(2.0L * (long double)M_PI), 3.0L) * (long double)delta * delta * delta;
  for (int ii = 0; ii < N; ii++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < N; k++) {
        long double sr = 0, si = 0;
        for (int i = 0; i < N; i++) {
          size_t s = IDX3(N, i, j, k);
          long double xr = A[2 * s], xi = A[2 * s + 1];
          sr += Mr[ii * N + i] * xr - Mi[ii * N + i] * xi;
          si += Mr[ii * N + i] * xi + Mi[ii * N + i] * xr;
        }
        size_t d = IDX3(N, ii, j, k);
        B[2 * d] = sr * pre; B[2 * d + 1] = si * pre;
      }
  for (size_t t = 0; t < 2 * n3; t++) out[t] = (double)B[t];
  free(A); free(B); free(Mr); free(Mi);
}

/* ---------------------------------------------------------- weighted sum */

/* src/collisions.c:125-165: Qhat[r] = sum_c W[r][c] * Ghat[c] * Fhat[s(r,c)],
 * s_d = (r_d + N/2 - c_d) wrapped once into [0, N).  Rows are independent;
 * each row is summed in the reference's column order.  nthreads > 1 splits
 * rows over OpenMP threads, as the reference does (:127). */
void orc_qhat(int N, const double *W, const double *Fhat, const double *Ghat, double *Qhat, int nthreads) {
  long n3 = (long)N * N * N;
  int n2 = N / 2;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
  for (long r = 0; r < n3; r++) {
    int rx = (int)(r / ((long)N * N)), ry = (int)((r / N) % N), rz = (int)(r % N);
    const double *Wr = W + r * n3;
    double qr = 0.0, qi = 0.0;
    for (long c = 0; c < n3; c++) {
      int cx = (int)(c / ((long)N * N)), cy = (int)((c / N) % N), cz = (int)(c % N);
      int x = rx + n2 - cx, y = ry + n2 - cy, z = rz + n2 - cz;
      if (x < 0) x += N; else if (x > N - 1) x -= N;
      if (y < 0) y += N; else if (y > N - 1) y -= N;
      if (z < 0) z += N; else if (z > N - 1) z -= N;
      long s = IDX3(N, x, y, z);
      double w = Wr[c];
      qr += w * (Ghat[2 * c] * Fhat[2 * s] - Ghat[2 * c + 1] * Fhat[2 * s + 1]);
      qi += w * (Ghat[2 * c] * Fhat[2 * s + 1] + Ghat[2 * c + 1] * Fhat[2 * s]);
    }
    Qhat[2 * r] = qr;
    Qhat[2 * r + 1] = qi;
  }
}

/* src/collisions.c:108-169 (compute_Qhat) + :212-221 (ComputeQ):
 * Q = Re( inverse( sum_c W * fft(g)[c] * fft(f)[s] ) ). */
void orc_compute_q(int N, const double *v, const double *eta, const double *W, const double *f, const double *g, double *Q,
                   int nthreads) {
  long n3 = (long)N * N * N;
  double *Fi = calloc(2 * n3, sizeof(double)), *Gi = calloc(2 * n3, sizeof(double));
  double *Qh = calloc(2 * n3, sizeof(double));
  for (long t = 0; t < n3; t++) { Fi[2 * t] = f[t]; Gi[2 * t] = g[t]; }
  orc_fft3d(N, v, eta, Fi, Fi, 0);
  orc_fft3d(N, v, eta, Gi, Gi, 0);
  orc_qhat(N, W, Fi, Gi, Qh, nthreads);
  orc_fft3d(N, v, eta, Qh, Qh, 1);
  for (long t = 0; t < n3; t++) Q[t] = Qh[2 * t];
  free(Fi); free(Gi); free(Qh);
}

/* ---------------------------------------------------------------- moments */

/* src/momentRoutines.c:26-29: KB = 1 iff mass == 1, else Boltzmann's constant. */
static double kb_for_mass(double mass) { return (mass == 1.0) ? 1.0 : 1.380658e-23; }

/* src/momentRoutines.c:58-72 */
double orc_density(int N, const double *v, double mass, const double *f) {
  double w[64], dv = v[1] - v[0], dv3 = dv * dv * dv, res = 0.0;
  trap_weights(N, w);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < N; k++) res += dv3 * w[i] * w[j] * w[k] * f[IDX3(N, i, j, k)];
  return mass * res;
}

/* src/momentRoutines.c:116-142 */
void orc_bulk_velocity(int N, const double *v, double mass, const double *f, double rho, double *u) {
  double w[64], dv = v[1] - v[0], dv3 = dv * dv * dv;
  trap_weights(N, w);
  u[0] = u[1] = u[2] = 0.0;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < N; k++) {
        double ww = dv3 * w[i] * w[j] * w[k] / rho, x = f[IDX3(N, i, j, k)];
        u[0] += v[i] * ww * x;
        u[1] += v[j] * ww * x;
        u[2] += v[k] * ww * x;
      }
  u[0] *= mass; u[1] *= mass; u[2] *= mass;
}

/* src/momentRoutines.c:168-183 */
double orc_temperature(int N, const double *v, double mass, const double *f, const double *u, double rho) {
  double w[64], dv = v[1] - v[0], dv3 = dv * dv * dv, res = 0.0;
  trap_weights(N, w);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < N; k++) {
        double t = (v[i] - u[0]) * (v[i] - u[0]) + (v[j] - u[1]) * (v[j] - u[1]) + (v[k] - u[2]) * (v[k] - u[2]);
        res += t * dv3 * w[i] * w[j] * w[k] * f[IDX3(N, i, j, k)] / (3.0 * rho);
      }
  return (mass * mass / kb_for_mass(mass)) * res;
}

/* src/momentRoutines.c:146-165: positive and negative parts of sum w |v|^2 f. */
void orc_energy_parts(int N, const double *v, const double *f, double *out) {
  double w[64], dv = v[1] - v[0], dv3 = dv * dv * dv, pos = 0.0, neg = 0.0;
  trap_weights(N, w);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < N; k++) {
        double e = dv3 * w[i] * w[j] * w[k] * f[IDX3(N, i, j, k)] * (v[i] * v[i] + v[j] * v[j] + v[k] * v[k]);
        if (e > 0) pos += e; else neg -= e;
      }
  out[0] = pos; out[1] = neg;
}

/* ------------------------------------------------------ maxPreserve form */

/* src/collisions.c:91-106 (find_maxwellians).  M = rho (2 pi T)^-3/2 exp(-|v-u|^2/2T)
 * from the discrete moments of `mat` (mass 1 in the reference call, spec_id 0).
 * The reference subtracts M_i -- the Maxwellian of the FIRST argument of
 * ComputeQ_maxPreserve -- in both calls (:104); `Msub` carries that array. */
static void find_maxwellian(int N, const double *v, double mass, const double *mat, double *M) {
  double rho = orc_density(N, v, mass, mat), u[3];
  orc_bulk_velocity(N, v, mass, mat, rho, u);
  double T = orc_temperature(N, v, mass, mat, u, rho);
  double pre = rho * pow(0.5 / (M_PI * T), 1.5);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < N; k++)
        M[IDX3(N, i, j, k)] = pre * exp(-(0.5 / T) * ((v[i] - u[0]) * (v[i] - u[0]) + (v[j] - u[1]) * (v[j] - u[1]) +
                                                      (v[k] - u[2]) * (v[k] - u[2])));
}

/* src/collisions.c:178-210 (ComputeQ_maxPreserve): with M_i, M_j the
 * Maxwellians of f and g, g_i = f - M_i, g_j = g - M_i (sic, :104),
 * Q = Q(M_i, g_j) + Q(g_i, M_j) + Q(g_i, g_j), three separate passes whose
 * real parts are summed in that order; Q(M,M) is omitted (:202-209). */
void orc_compute_q_maxpreserve(int N, const double *v, const double *eta, double mass, const double *W, const double *f,
                               const double *g, double *Q, int nthreads) {
  long n3 = (long)N * N * N;
  double *Mi = malloc(sizeof(double) * n3), *Mj = malloc(sizeof(double) * n3);
  double *gi = malloc(sizeof(double) * n3), *gj = malloc(sizeof(double) * n3), *T = malloc(sizeof(double) * n3);
  find_maxwellian(N, v, mass, f, Mi);
  for (long t = 0; t < n3; t++) gi[t] = f[t] - Mi[t];
  find_maxwellian(N, v, mass, g, Mj);
  for (long t = 0; t < n3; t++) gj[t] = g[t] - Mi[t];
  orc_compute_q(N, v, eta, W, Mi, gj, T, nthreads);
  for (long t = 0; t < n3; t++) Q[t] = T[t];
  orc_compute_q(N, v, eta, W, gi, Mj, T, nthreads);
  for (long t = 0; t < n3; t++) Q[t] += T[t];
  orc_compute_q(N, v, eta, W, gi, gj, T, nthreads);
  for (long t = 0; t < n3; t++) Q[t] += T[t];
  free(Mi); free(Mj); free(gi); free(gj); free(T);
}

/* ----------------------------------------------------------- conservation */

/* src/conserve.c:89-168 (factorMatrixIntoLU, after Atkinson p.520): in-place LU
 * with scaled partial pivoting.  The pivot row is the FIRST row i >= k whose
 * scaled entry |A[i][k]|/s_i exceeds the diagonal's (:115-126), not the
 * largest.  Returns -1 when a pivot column is entirely zero (:130-135). */
int orc_lu_factor(int n, double *A, int *pivot) {
  double s[16];
  for (int i = 0; i < n; i++) {
    s[i] = fabs(A[i * n]);
    for (int j = 0; j < n; j++)
      if (s[i] < fabs(A[i * n + j])) s[i] = fabs(A[i * n + j]);
  }
  for (int k = 0; k < n - 1; k++) {
    double ck = fabs(A[k * n + k] / s[k]);
    int i0 = k, found = 0;
    for (int i = k; i < n; i++) {
      if (ck < fabs(A[i * n + k] / s[i])) {
        ck = fabs(A[i * n + k] / s[i]);
        if (!found) { i0 = i; found = 1; }
      }
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
  return 0;
}

/* src/conserve.c:173-202 (solveWithCCt, Atkinson p.521). */
void orc_lu_solve(int n, const double *LU, const int *pivot, double *b) {
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

/* src/conserve.c:268-317 (createCCtAndPivot): C_m rows w_i w_j w_k h_v^3 m_m
 * {delta_mn, v_i, v_j, v_k, |v|^2/2}; CCt = sum_m C_m C_m^T, then LU.
 * LU has (Ns+4)^2 entries. Returns orc_lu_factor's status. */
int orc_conserve_setup(int N, const double *v, int Ns, const double *masses, double *LU, int *pivot) {
  int n = Ns + 4;
  long n3 = (long)N * N * N;
  double w[64], dv = v[1] - v[0];
  trap_weights(N, w);
  double *C = calloc((size_t)Ns * n * n3, sizeof(double));
  for (int m = 0; m < Ns; m++)
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        for (int k = 0; k < N; k++) {
          double p = w[i] * w[j] * w[k] * dv * dv * dv * masses[m];
          long t = IDX3(N, i, j, k);
          double *Cm = C + (size_t)m * n * n3;
          Cm[(size_t)m * n3 + t] = p;
          Cm[(size_t)Ns * n3 + t] = p * v[i];
          Cm[(size_t)(Ns + 1) * n3 + t] = p * v[j];
          Cm[(size_t)(Ns + 2) * n3 + t] = p * v[k];
          Cm[(size_t)(Ns + 3) * n3 + t] = p * 0.5 * (v[i] * v[i] + v[j] * v[j] + v[k] * v[k]);
        }
  for (int a = 0; a < n; a++)
    for (int b = 0; b < n; b++) {
      double s = 0.0;
      for (int m = 0; m < Ns; m++) {
        const double *Cm = C + (size_t)m * n * n3;
        for (long t = 0; t < n3; t++) s += Cm[(size_t)a * n3 + t] * Cm[(size_t)b * n3 + t];
      }
      LU[a * n + b] = s;
    }
  free(C);
  return orc_lu_factor(n, LU, pivot);
}

/* src/conserve.c:207-264 (conserveAllMoments): Q holds Ns*Ns arrays of N^3,
 * array n*Ns+m being the operator of species m against n.  lambda =
 * (CC^T)^-1 sum C_m Q, then Q -= C_m^T lambda / Ns. */
void orc_conserve(int N, const double *v, int Ns, const double *masses, const double *LU, const int *pivot, double *Q) {
  int n = Ns + 4;
  long n3 = (long)N * N * N;
  double w[64], dv = v[1] - v[0], mq[16], tc[16];
  trap_weights(N, w);
  for (int l = 0; l < n; l++) mq[l] = 0.0;
  for (int m = 0; m < Ns; m++)
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        for (int k = 0; k < N; k++) {
          double p = w[i] * w[j] * w[k] * dv * dv * dv * masses[m];
          tc[m] = p;
          tc[Ns] = p * v[i];
          tc[Ns + 1] = p * v[j];
          tc[Ns + 2] = p * v[k];
          tc[Ns + 3] = p * 0.5 * (v[i] * v[i] + v[j] * v[j] + v[k] * v[k]);
          for (int nn = 0; nn < Ns; nn++) {
            double q = Q[(size_t)(nn * Ns + m) * n3 + IDX3(N, i, j, k)];
            mq[m] += q * tc[m];
            mq[Ns] += q * tc[Ns];
            mq[Ns + 1] += q * tc[Ns + 1];
            mq[Ns + 2] += q * tc[Ns + 2];
            mq[Ns + 3] += q * tc[Ns + 3];
          }
        }
  orc_lu_solve(n, LU, pivot, mq);
  for (int m = 0; m < Ns; m++)
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        for (int k = 0; k < N; k++) {
          double p = w[i] * w[j] * w[k] * dv * dv * dv * masses[m] / Ns;
          tc[m] = p;
          tc[Ns] = p * v[i];
          tc[Ns + 1] = p * v[j];
          tc[Ns + 2] = p * v[k];
          tc[Ns + 3] = p * 0.5 * (v[i] * v[i] + v[j] * v[j] + v[k] * v[k]);
          for (int nn = 0; nn < Ns; nn++)
            Q[(size_t)(nn * Ns + m) * n3 + IDX3(N, i, j, k)] -=
                (tc[m] * mq[m] + tc[Ns] * mq[Ns] + tc[Ns + 1] * mq[Ns + 1] + tc[Ns + 2] * mq[Ns + 2] + tc[Ns + 3] * mq[Ns + 3]);
        }
}

/* ---------------------------------------------------------------- weights */

static double sinc(double x) { return (x != 0.0) ? sin(x) / x : 1.0; }
static double func_cos(double x) { return (fabs(x) > 1e-12) ? (cos(x) - 1.0) / (x * x) : -0.5 + x * x / 24.0; }

/* Integrand of src/weights.c:156-160 (ghat): r^(lam+2) [sinc(a0 r) sinc(a2 r) - sinc(a1 r)]. */
static double ghat_integrand(double r, double a0, double a1, double a2, double lam) {
  return pow(r, lam + 2) * (sinc(r * a0) * sinc(r * a2) - sinc(r * a1));
}

/* Gauss-Legendre nodes on [-1,1] by Newton iteration on P_n (Golub-Welsch is
 * not needed at these orders). */
static void gauss_legendre(int n, double *x, double *w) {
  for (int i = 0; i < (n + 1) / 2; i++) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
    for (int it = 0; it < 100; it++) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; j++) { double p3 = p2; p2 = p1; p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j; }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      double z1 = z;
      z = z1 - p1 / pp;
      if (fabs(z - z1) < 1e-16) break;
    }
    x[i] = -z; x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
}

/* I_lam(a0,a1,a2) = int_0^L r^(lam+2)[sinc(a0 r) sinc(a2 r) - sinc(a1 r)] dr,
 * the integral src/weights.c:203 evaluates with GSL QAG (GAUSS21, 1e-8/1e-8).
 * lam == 0 and lam == 1 use closed forms (gain - loss); lam == 1's is the one
 * the reference keeps commented out at src/weights.c:209-253, re-derived:
 *   int_0^L r^2 sin(ar) sin(br)/(ab) dr = L/(2ab) [sinc((a-b)L) - sinc((a+b)L)]
 *   int_0^L r^2 sinc(cr) dr            = (sin cL - cL cos cL)/c^3
 *   int_0^L r^3 sin(ar) sin(br)/(ab) dr = L^2/(2ab) [sinc(am)-sinc(ap)+fc(am)-fc(ap)],
 *        am=(a-b)L, ap=(a+b)L, fc(x)=(cos x - 1)/x^2
 *   int_0^L r^3 sinc(cr) dr            = (cos cL (2 - (cL)^2) + 2 cL sin cL - 2)/c^4.
 * Other lam: composite Gauss-Legendre (64 panels x 20 nodes). */
double orc_ghat_integral(double a0, double a1, double a2, double L, double lam) {
  const double tiny = 1e-12;
  if (lam == 0.0) {
    double gain, loss;
    if (a0 > tiny && a2 > tiny) gain = L / (2 * a0 * a2) * (sinc((a0 - a2) * L) - sinc((a0 + a2) * L));
    else if (a0 > tiny || a2 > tiny) { double c = (a0 > tiny) ? a0 : a2; gain = (sin(c * L) - c * L * cos(c * L)) / (c * c * c); }
    else gain = L * L * L / 3.0;
    if (a1 > tiny) loss = (sin(a1 * L) - a1 * L * cos(a1 * L)) / (a1 * a1 * a1);
    else loss = L * L * L / 3.0;
    return gain - loss;
  }
  if (lam == 1.0) {
    double gain, loss;
    if (a0 > tiny && a2 > tiny) {
      double am = (a0 - a2) * L, ap = (a0 + a2) * L;
      if (fabs(a0 - a2) < 1e-10) am = 0.0;
      gain = (L * L * 0.5 / (a0 * a2)) * (sinc(am) - sinc(ap) + func_cos(am) - func_cos(ap));
    } else if (a0 > tiny || a2 > tiny) {
      double c = (a0 > tiny) ? a0 : a2, x = c * L;
      gain = (2 * x * sin(x) - x * x * cos(x) + 2 * cos(x) - 2) / (c * c * c * c);
    } else gain = 0.25 * L * L * L * L;
    if (a1 > tiny) { double x = a1 * L; loss = (cos(x) * (2.0 - x * x) + 2.0 * x * sin(x) - 2.0) / (a1 * a1 * a1 * a1); }
    else loss = 0.25 * L * L * L * L;
    return gain - loss;
  }
  static double gx[20], gw[20];
  static int init = 0;
  if (!init) { gauss_legendre(20, gx, gw); init = 1; }
  const int panels = 64;
  double h = L / panels, sum = 0.0;
  for (int p = 0; p < panels; p++) {
    double a = p * h, ps = 0.0;
    for (int q = 0; q < 20; q++) ps += gw[q] * ghat_integrand(a + 0.5 * h * (gx[q] + 1.0), a0, a1, a2, lam);
    sum += 0.5 * h * ps;
  }
  return sum;
}

/* src/weights.c:265-281 (generate_conv_weights_iso) with gHat3 (:181-262):
 * W[r][c] = w_c1 w_c2 w_c3 * (d_i+d_j)^2/16 * P * I_lam(mu|eta_r|, |eta_c|, |eta_c - mu eta_r|),
 * P = 16 pi^2 h_eta^3 (2 pi)^-3/2 / (4 pi) (:44), mu = m_j/(m_i+m_j) (:190).
 * W is N^3 x N^3 row-major (row = r). */
void orc_weights_iso(int N, const double *eta, double Lv, double lam, double d_i, double d_j, double m_i, double m_j, double *W,
                     int nthreads) {
  double w[64];
  trap_weights(N, w);
  double deta = eta[1] - eta[0];
  double P = 16.0 * M_PI * M_PI * deta * deta * deta / pow(2.0 * M_PI, 1.5) / (4.0 * M_PI);
  double dfac = 0.25 * pow(0.5 * (d_i + d_j), 2);
  double mu = m_j / (m_i + m_j);
  long n3 = (long)N * N * N;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
#endif
  for (long r = 0; r < n3; r++) {
    int i = (int)(r / ((long)N * N)), j = (int)((r / N) % N), k = (int)(r % N);
    double z1 = eta[i], z2 = eta[j], z3 = eta[k];
    double a0 = mu * sqrt(z1 * z1 + z2 * z2 + z3 * z3);
    for (long c = 0; c < n3; c++) {
      int l = (int)(c / ((long)N * N)), m = (int)((c / N) % N), nn = (int)(c % N);
      double k1 = eta[l], k2 = eta[m], k3 = eta[nn];
      double a1 = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
      double a2 = sqrt((k1 - mu * z1) * (k1 - mu * z1) + (k2 - mu * z2) * (k2 - mu * z2) + (k3 - mu * z3) * (k3 - mu * z3));
      W[r * n3 + c] = w[l] * w[m] * w[nn] * dfac * (P * orc_ghat_integral(a0, a1, a2, Lv, lam));
    }
  }
}

/* 64-point Gauss-Legendre rule on [-1, 1]: the 32 positive nodes in
 * ascending order and their weights, from Newton's iteration on P_64 in long
 * double (the tables gsl_integration_glfixed and gauss_legendre use). */
void orc_gauss_legendre64(double *x, double *w) {
  const int n = 64;
  for (int i = 0; i < n / 2; i++) {
    long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L)), pp = 0.0L;
    for (int it = 0; it < 100; it++) {
      long double p1 = 1.0L, p2 = 0.0L;
      for (int j = 1; j <= n; j++) {
        long double p3 = p2;
        p2 = p1;
        p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0L);
      long double z1 = z;
      z = z1 - p1 / pp;
      if (fabsl(z - z1) < 1e-19L) break;
    }
    /* i = 0 is the largest root; store ascending */
    x[n / 2 - 1 - i] = (double)z;
    w[n / 2 - 1 - i] = (double)(2.0L / ((1.0L - z * z) * pp * pp));
  }
}

/* gsl_integration_glfixed / gauss_legendre for even n: A = (b-a)/2,
 * B = (b+a)/2, A * sum_i w_i (f(B + A x_i) + f(B - A x_i)). */
typedef double (*orc_fn)(double, const double *);
static double gl64_integrate(orc_fn f, const double *args, double a, double b, const double *x, const double *w) {
  double A = 0.5 * (b - a), B = 0.5 * (b + a), s = 0.0;
  for (int i = 0; i < 32; i++) {
    double Ax = A * x[i];
    s += w[i] * (f(B + Ax, args) + f(B - Ax, args));
  }
  return A * s;
}

/* ghatL2 (exec/MPIcollisionRoutines.c:205-210, src/aniso_weights.c:352-357);
 * args = {xiperp, |zeta|^2, xi.zeta/|zeta|, |zeta|, r}. */
static double landau_theta(double th, const double *a) {
  double r = a[4];
  return sin(th) * j0(r * a[0] * sin(th)) *
         (-r * r * a[1] * sin(th) * sin(th) * cos(r * a[2] * cos(th)) + 4 * r * a[3] * sin(r * a[2] * cos(th)) * cos(th));
}

static double orc_landau_lam;
static const double *orc_glx, *orc_glw;

/* ghatL (:212-217, :359-369): r^(lam+2) * int_0^pi ghatL2 dtheta */
static double landau_r(double r, const double *a) {
  double b[5] = {a[0], a[1], a[2], a[3], r};
  return pow(r, orc_landau_lam + 2) * gl64_integrate(landau_theta, b, 0, M_PI, orc_glx, orc_glw);
}

/* gHat3L (:219-248, :371-398) for row mode zeta and column mode xi. */
double orc_landau_entry(double z1, double z2, double z3, double x1, double x2, double x3, double Lv, double lam) {
  double gx[32], gw[32], a[4];
  orc_gauss_legendre64(gx, gw);
  double zetalen2 = z3 * z3 + z2 * z2 + z1 * z1;
  double xilen2 = x1 * x1 + x2 * x2 + x3 * x3;
  double xizeta = x1 * z1 + x2 * z2 + x3 * z3;
  double zetalen = sqrt(zetalen2), xiperp;
  if (((xilen2 - xizeta * xizeta / zetalen2) < 0) || (zetalen2 == 0)) xiperp = 0;
  else xiperp = sqrt(xilen2 - xizeta * xizeta / zetalen2);
  a[0] = xiperp;
  a[1] = zetalen2;
  a[2] = (zetalen != 0) ? xizeta / zetalen : 0.0;
  a[3] = zetalen;
  orc_landau_lam = lam;
  orc_glx = gx;
  orc_glw = gw;
  return 2.0 * M_PI * gl64_integrate(landau_r, a, 0, Lv, gx, gw);
}

/* generate_conv_weights with glance = 0 (exec/MPIcollisionRoutines.c:251-281,
 * src/aniso_weights.c:400-428): W[r][c] = gHat3L(eta_r, eta_c) for rows
 * r0 .. r0+nrows-1 (no trapezoid or h_eta^3 factors, as the reference writes
 * the table).  W holds nrows x N^3.  Single-threaded (the integrand keeps its
 * rule in statics). */
void orc_weights_landau_rows(int N, const double *eta, double Lv, double lam, long r0, long nrows, double *W) {
  long n3 = (long)N * N * N;
  for (long r = r0; r < r0 + nrows; r++) {
    int i = (int)(r / ((long)N * N)), j = (int)((r / N) % N), k = (int)(r % N);
    for (long c = 0; c < n3; c++) {
      int l = (int)(c / ((long)N * N)), m = (int)((c / N) % N), nn = (int)(c % N);
      W[(r - r0) * n3 + c] = orc_landau_entry(eta[i], eta[j], eta[k], eta[l], eta[m], eta[nn], Lv, lam);
    }
  }
}

/* ---- Glancing-collision table (glance > 0), gHat3 of
 * exec/MPIcollisionRoutines.c:149-203 with its integrands ghat_theta (:33-57,
 * the eps-linear cross section b = (8/pi) glance / theta^4), ghat_phi (:78-121,
 * the Taylor term C (2/pi) log(glance) / log(sin(glance/2)) for the excluded
 * [0, sqrt(glance)] part) and ghat_r (:124-140):
 *   W = 4 pi^2 int_0^L r^(lam+2) int_0^pi sin(phi) J0(r sin(phi) xi_perp) (R1 + R2) dphi dr,
 *   R1 = int_sqrt(g)^pi b(theta) [cos(a7 (1 - cos theta) - a6) J0(a8 sin theta) - cos a6] dtheta,
 *   a6 = r cos(phi) xi.zeta/|zeta|, a7 = r cos(phi) |zeta| / 2, a8 = r sin(phi) |zeta| / 2.
 * The reference integrates adaptively (gsl_integration_qag, 61-point
 * Gauss-Kronrod, for r and phi; gsl_integration_cquad for theta; absolute and
 * relative tolerance 1e-6 each).  Here every level is a fixed composite rule
 * of 16-point Gauss-Legendre panels: pr panels on [0, L], pphi on [0, pi];
 * theta in log(theta) on [sqrt(g), 1] (pt1 panels; the integrand grows like
 * 1/theta^2 towards the lower end) and in theta on [max(1, sqrt(g)), pi] (pt2
 * panels), so the sum converges to the integral the reference approximates. */
static void gl16_half(double *x, double *w) {
  const int n = 16;
  for (int i = 0; i < n / 2; i++) {
    long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L)), pp = 0.0L;
    for (int it = 0; it < 100; it++) {
      long double p1 = 1.0L, p2 = 0.0L;
      for (int j = 1; j <= n; j++) {
        long double p3 = p2;
        p2 = p1;
        p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0L);
      long double z1 = z;
      z = z1 - p1 / pp;
      if (fabsl(z - z1) < 1e-19L) break;
    }
    x[n / 2 - 1 - i] = (double)z;
    w[n / 2 - 1 - i] = (double)(2.0L / ((1.0L - z * z) * pp * pp));
  }
}

/* nodes and weights of P composite 16-point panels on [a, b] */
static int composite16(double a, double b, int P, const double *gx, const double *gw, double *x, double *w) {
  const double h = (b - a) / P;
  int n = 0;
  for (int p = 0; p < P; p++) {
    const double c = a + h * (p + 0.5);
    for (int i = 0; i < 8; i++) {
      x[n] = c - 0.5 * h * gx[i];
      w[n++] = 0.5 * h * gw[i];
      x[n] = c + 0.5 * h * gx[i];
      w[n++] = 0.5 * h * gw[i];
    }
  }
  return n;
}

double orc_glance_entry(double z1, double z2, double z3, double x1, double x2, double x3, double Lv, double lam,
                        double glance, int pr, int pphi, int pt1, int pt2) {
  double gx[8], gw[8];
  gl16_half(gx, gw);
  /* geometry as gHat3 (:165-183) */
  const double zetalen2 = z1 * z1 + z2 * z2 + z3 * z3;
  const double xilen2 = x1 * x1 + x2 * x2 + x3 * x3;
  const double xizeta = x1 * z1 + x2 * z2 + x3 * z3;
  const double zetalen = sqrt(zetalen2);
  double xiperp;
  if (((xilen2 - xizeta * xizeta / zetalen2) < 0) || (zetalen2 == 0)) xiperp = 0;
  else xiperp = sqrt(xilen2 - xizeta * xizeta / zetalen2);
  const double arg0 = zetalen, arg1 = (zetalen != 0) ? xizeta / zetalen : 0.0, arg2 = xiperp;

  /* theta rule, with b(theta) and the Jacobian folded into the weights */
  const double t0 = sqrt(glance);
  const int nt_max = 16 * (pt1 + pt2);
  double *tx = malloc(sizeof(double) * nt_max), *tw = malloc(sizeof(double) * nt_max);
  double *ct = malloc(sizeof(double) * nt_max), *st = malloc(sizeof(double) * nt_max);
  int nt = 0;
  if (t0 < 1.0) {
    nt = composite16(log(t0), 0.0, pt1, gx, gw, tx, tw);
    for (int k = 0; k < nt; k++) {
      tx[k] = exp(tx[k]);
      tw[k] *= tx[k];
    }
  }
  if (t0 < M_PI) nt += composite16(t0 > 1.0 ? t0 : 1.0, M_PI, pt2, gx, gw, tx + nt, tw + nt);
  for (int k = 0; k < nt; k++) {
    const double th = tx[k];
    tw[k] *= (8.0 / M_PI) * (glance / (th * th)) * pow(th, -2.0); /* ghat_theta :51 */
    ct[k] = cos(th);
    st[k] = sin(th);
  }
  const double Lfac = (2.0 / M_PI) * log(glance) / log(sin(0.5 * glance)); /* :119 */

  double *rx = malloc(sizeof(double) * 16 * pr), *rw = malloc(sizeof(double) * 16 * pr);
  double *px = malloc(sizeof(double) * 16 * pphi), *pw = malloc(sizeof(double) * 16 * pphi);
  const int nr = composite16(0.0, Lv, pr, gx, gw, rx, rw);
  const int np = composite16(0.0, M_PI, pphi, gx, gw, px, pw);
  double sr = 0.0;
  for (int i = 0; i < nr; i++) {
    const double r = rx[i];
    double sp = 0.0;
    for (int j = 0; j < np; j++) {
      const double arg4 = cos(px[j]), arg5 = sin(px[j]); /* ghat_phi :97-102 */
      const double arg6 = r * arg4 * arg1, arg7 = 0.5 * r * arg4 * arg0, arg8 = 0.5 * r * arg5 * arg0;
      const double arg9 = cos(r * arg4 * arg1);
      double R1 = 0.0;
      for (int k = 0; k < nt; k++) /* ghat_theta :55 */
        R1 += tw[k] * (cos(arg7 * (1 - ct[k]) - arg6) * j0(arg8 * st[k]) - arg9);
      const double c1 = 0.5 * r * arg0 * arg4, c2 = 0.5 * r * arg0 * arg5, c3 = r * arg1 * arg4;
      const double C = (-0.25 * c2 * c2 * cos(c3) + 0.5 * c1 * sin(c3));
      sp += pw[j] * (arg5 * j0(r * arg5 * arg2) * (R1 + C * Lfac)); /* :121 */
    }
    sr += rw[i] * (pow(r, lam + 2) * sp); /* ghat_r :139 */
  }
  free(tx); free(tw); free(ct); free(st); free(rx); free(rw); free(px); free(pw);
  return 4 * M_PI * M_PI * sr; /* :202 */
}

/* generate_conv_weights with glance > 0 (exec/MPIcollisionRoutines.c:251-281)
 * at n chosen entries: out[e] = gHat3(eta_row, eta_col) for row rows[e] and
 * column cols[e] (row = output mode zeta, column = summed mode xi).  OpenMP
 * over entries. */
void orc_weights_glance_entries(int N, const double *eta, double Lv, double lam, double glance, long n,
                                const long *rows, const long *cols, int pr, int pphi, int pt1, int pt2, double *out,
                                int nthreads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (long e = 0; e < n; e++) {
    const long r = rows[e], c = cols[e];
    const int i = (int)(r / ((long)N * N)), j = (int)((r / N) % N), k = (int)(r % N);
    const int l = (int)(c / ((long)N * N)), m = (int)((c / N) % N), nn = (int)(c % N);
    out[e] = orc_glance_entry(eta[i], eta[j], eta[k], eta[l], eta[m], eta[nn], Lv, lam, glance, pr, pphi, pt1, pt2);
  }
  (void)nthreads;
}

/* src/weights.c:68,78-88: raw native-endian FP64, N^3 rows of N^3, no header.
 * Returns 0 on success, -1 if the file is missing, -2 if it is short. */
int orc_weights_load(const char *path, int N, double *W) {
  FILE *fp = fopen(path, "rb");
  if (!fp) return -1;
  long n3 = (long)N * N * N;
  for (long r = 0; r < n3; r++)
    if ((long)fread(W + r * n3, sizeof(double), n3, fp) != n3) { fclose(fp); return -2; }
  fclose(fp);
  return 0;
}

/* -------------------------------------------------------------- transport */

/* src/transportroutines.c:80-90 */
static double minmod3(double a, double b, double c) {
  if (a > 0 && b > 0 && c > 0) return fmin(fmin(a, b), c);
  if (a < 0 && b < 0 && c < 0) return fmax(fmax(a, b), c);
  return 0.0;
}

/* src/boundaryConditions.c:39-85 (setDiffuseReflectionBC): the outgoing flux
 * of `in` through the wall sets sigma_W; `out` gets sigma_W times a wall
 * Maxwellian exp(-m|v|^2/(2 KB T_W)) on the incoming half (left wall: i >= N/2;
 * right wall: i < N/2).  The other half of `out` is left untouched. */
/* KB of the walls: src/boundaryConditions.c:31-34 takes it from species 0's
 * mass (1 for mass 1, SI otherwise) for every species; orc_set_wall_kb sets it
 * for the species-mass walls of a multi-species run (0: from `mass`). */
static double orc_wall_kb = 0.0;
void orc_set_wall_kb(double kb) { orc_wall_kb = kb; }

void orc_diffuse_bc(int N, const double *v, double mass, const double *in, double *out, double TW, int right) {
  double w[64], hv = v[1] - v[0], KB = orc_wall_kb > 0 ? orc_wall_kb : kb_for_mass(mass), sig = 0.0;
  trap_weights(N, w);
  int o0 = right ? N / 2 : 0, o1 = right ? N : N / 2;
  for (int i = o0; i < o1; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < N; k++) sig += v[i] * w[i] * w[j] * w[k] * hv * hv * hv * in[IDX3(N, i, j, k)];
  sig *= (right ? 1.0 : -1.0) * sqrt(2.0 * M_PI * mass / (KB * TW));
  int i0 = right ? 0 : N / 2, i1 = right ? N / 2 : N;
  for (int i = i0; i < i1; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < N; k++)
        out[IDX3(N, i, j, k)] = sig * pow(0.5 * mass / (M_PI * KB * TW), 1.5) *
                                exp(-0.5 * mass / (KB * TW) * (v[i] * v[i] + v[j] * v[j] + v[k] * v[k]));
}

/* src/transportroutines.c:94-238 (upwindOne) for a single rank.  f and fc
 * hold nX+2 cells of N^3 (ghosts 0 and nX+1).  bc: 0 = copy neighbour
 * (f[0] = f[1], f[nX+1] = f[nX]), 1 = diffuse walls at T0 (left) and T1
 * (right) (ICChoice 3/5), 2 = periodic (ICChoice 6), 3 = diffuse left wall at
 * T0 and copied right ghost (ICChoice 1, :120-121, :136-137). Writes
 * fc[1..nX].  f's ghost cells are overwritten. */
void orc_upwind_one(int N, int nX, const double *v, double mass, const double *dx, double dt, int bc, double T0, double T1,
                    double *f, double *fc) {
  long n3 = (long)N * N * N;
  if (bc == 1) {
    orc_diffuse_bc(N, v, mass, f + 1 * n3, f + 0, T0, 0);
    orc_diffuse_bc(N, v, mass, f + nX * n3, f + (nX + 1) * n3, T1, 1);
  } else if (bc == 2) {
    memcpy(f, f + nX * n3, sizeof(double) * n3);
    memcpy(f + (nX + 1) * n3, f + 1 * n3, sizeof(double) * n3);
  } else if (bc == 3) {
    orc_diffuse_bc(N, v, mass, f + 1 * n3, f + 0, T0, 0);
    memcpy(f + (nX + 1) * n3, f + nX * n3, sizeof(double) * n3);
  } else {
    memcpy(f, f + 1 * n3, sizeof(double) * n3);
    memcpy(f + (nX + 1) * n3, f + nX * n3, sizeof(double) * n3);
  }
  for (int l = 1; l < nX + 1; l++)
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        for (int k = 0; k < N; k++) {
          long t = IDX3(N, i, j, k);
          double cfl = dt * v[i] / dx[l];
          if (i < N / 2) fc[l * n3 + t] = (1.0 + cfl) * f[l * n3 + t] - cfl * f[(l + 1) * n3 + t];
          else fc[l * n3 + t] = (1.0 - cfl) * f[l * n3 + t] + cfl * f[(l - 1) * n3 + t];
        }
}

/* src/transportroutines.c:241-470 (upwindTwo) for a single rank: MUSCL with
 * minmod slopes and half dt.  f and fc hold nX+4 cells (ghosts 0,1 and
 * nX+2,nX+3); x and dx hold nX+4 entries (the right ghosts follow the
 * last-rank formula of src/mesh_setup.c:165-175; the reference leaves them
 * unset on a single rank).  Ghost 1 and nX+2 are linearly extrapolated
 * (:271-274, :293-296).  bc: 1 = diffuse walls (ICChoice 3/5), 3 = diffuse
 * left wall only (ICChoice 1, :363-364), otherwise the no-net-flux wall
 * states of :366-377 and :392-403.  Ma != 0 adds the Poiseuille body force
 * of :429-437 and :458-466 (coefficient Ma*0.5*dt/(2*h_v), h_v = 2 L_v/(N-1),
 * with its one-sided ends as written there).  Writes fc[2..nX+1]. */
static void force_vy(int N, double fy, const double *f0, long t, int j, double *out) {
  if (j == 0) *out = *out - fy * f0[t + N];
  else if (j == N - 1) *out = *out - fy * f0[t - N];
  else *out = *out - fy * (f0[t + N] - f0[t - N]);
}

void orc_upwind_two(int N, int nX, const double *v, double mass, const double *x, const double *dx, double dt, int bc,
                    double T0, double T1, double Ma, double *f, double *fc) {
  long n3 = (long)N * N * N;
  const double h_v = 2 * (-v[0]) / (N - 1), fy = Ma * 0.5 * dt / (2 * h_v);
  double *fl = malloc(sizeof(double) * n3), *fr = malloc(sizeof(double) * n3);
  for (long t = 0; t < n3; t++) {
    f[1 * n3 + t] = 2 * f[2 * n3 + t] - f[3 * n3 + t];
    f[(nX + 2) * n3 + t] = 2 * f[(nX + 1) * n3 + t] - f[nX * n3 + t];
  }
  for (int l = 2; l < nX + 2; l++) {
    if (l == 2) {
      for (int i = 0; i < N / 2; i++)
        for (int j = 0; j < N; j++)
          for (int k = 0; k < N; k++) {
            long t = IDX3(N, i, j, k);
            double s = minmod3((f[2 * n3 + t] - f[1 * n3 + t]) / (x[2] - x[1]), (f[3 * n3 + t] - f[2 * n3 + t]) / (x[3] - x[2]),
                               (f[3 * n3 + t] - f[1 * n3 + t]) / (x[3] - x[1]));
            fl[t] = f[2 * n3 + t] - 0.5 * dx[2] * s;
          }
      if (bc == 1 || bc == 3) orc_diffuse_bc(N, v, mass, fl, fl, T0, 0);
      else
        for (int j = 0; j < N; j++)
          for (int k = 0; k < N; k++)
            for (int i = N / 2; i < N; i++) {
              long t = IDX3(N, i, j, k);
              double s = minmod3((f[2 * n3 + t] - f[1 * n3 + t]) / (x[2] - x[1]), (f[3 * n3 + t] - f[2 * n3 + t]) / (x[3] - x[2]),
                                 (f[3 * n3 + t] - f[1 * n3 + t]) / (x[3] - x[1]));
              fl[t] = f[2 * n3 + t] + 0.5 * dx[2] * s;
            }
    }
    if (l == nX + 1) {
      for (int i = N / 2; i < N; i++)
        for (int j = 0; j < N; j++)
          for (int k = 0; k < N; k++) {
            long t = IDX3(N, i, j, k);
            double s = minmod3((f[(nX + 1) * n3 + t] - f[nX * n3 + t]) / (x[nX + 1] - x[nX]),
                               (f[(nX + 2) * n3 + t] - f[(nX + 1) * n3 + t]) / (x[nX + 2] - x[nX + 1]),
                               (f[(nX + 2) * n3 + t] - f[nX * n3 + t]) / (x[nX + 2] - x[nX]));
            fr[t] = f[(nX + 1) * n3 + t] + 0.5 * dx[nX + 1] * s;
          }
      if (bc == 1) orc_diffuse_bc(N, v, mass, fr, fr, T1, 1);
      else
        for (int i = 0; i < N / 2; i++)
          for (int j = 0; j < N; j++)
            for (int k = 0; k < N; k++) {
              long t = IDX3(N, i, j, k);
              double s = minmod3((f[(nX + 1) * n3 + t] - f[nX * n3 + t]) / (x[nX + 1] - x[nX]),
                                 (f[(nX + 2) * n3 + t] - f[(nX + 1) * n3 + t]) / (x[nX + 2] - x[nX + 1]),
                                 (f[(nX + 2) * n3 + t] - f[nX * n3 + t]) / (x[nX + 2] - x[nX]));
              fr[t] = f[(nX + 1) * n3 + t] - 0.5 * dx[nX + 1] * s;
            }
    }
    for (int j = 0; j < N; j++)
      for (int k = 0; k < N; k++) {
        for (int i = N / 2; i < N; i++) {
          long t = IDX3(N, i, j, k);
          const double *fm2 = f + (l - 2) * n3, *fm1 = f + (l - 1) * n3, *f0 = f + l * n3, *fp1 = f + (l + 1) * n3;
          double s1 = minmod3((f0[t] - fm1[t]) / (x[l] - x[l - 1]), (fp1[t] - f0[t]) / (x[l + 1] - x[l]),
                              (fp1[t] - fm1[t]) / (x[l + 1] - x[l - 1]));
          double cfl = 0.5 * dt * v[i] / dx[l];
          if (l == 2) fc[l * n3 + t] = f0[t] - cfl * (f0[t] + 0.5 * dx[l] * s1 - fl[t]);
          else {
            double s0 = minmod3((fm1[t] - fm2[t]) / (x[l - 1] - x[l - 2]), (f0[t] - fm1[t]) / (x[l] - x[l - 1]),
                                (f0[t] - fm2[t]) / (x[l] - x[l - 2]));
            fc[l * n3 + t] = f0[t] - cfl * (f0[t] + 0.5 * dx[l] * s1 - (fm1[t] + 0.5 * dx[l - 1] * s0));
          }
          if (Ma != 0.0) force_vy(N, fy, f0, t, j, &fc[l * n3 + t]);
        }
        for (int i = 0; i < N / 2; i++) {
          long t = IDX3(N, i, j, k);
          const double *fm1 = f + (l - 1) * n3, *f0 = f + l * n3, *fp1 = f + (l + 1) * n3, *fp2 = f + (l + 2) * n3;
          double s1 = minmod3((f0[t] - fm1[t]) / (x[l] - x[l - 1]), (fp1[t] - f0[t]) / (x[l + 1] - x[l]),
                              (fp1[t] - fm1[t]) / (x[l + 1] - x[l - 1]));
          double cfl = 0.5 * dt * v[i] / dx[l];
          if (l == nX + 1) fc[l * n3 + t] = f0[t] - cfl * (fr[t] - (f0[t] - 0.5 * dx[l] * s1));
          else {
            double s2 = minmod3((fp1[t] - f0[t]) / (x[l + 1] - x[l]), (fp2[t] - fp1[t]) / (x[l + 2] - x[l + 1]),
                                (fp2[t] - f0[t]) / (x[l + 2] - x[l]));
            fc[l * n3 + t] = f0[t] - cfl * (fp1[t] - 0.5 * dx[l + 1] * s2 - (f0[t] - 0.5 * dx[l] * s1));
          }
          if (Ma != 0.0) force_vy(N, fy, f0, t, j, &fc[l * n3 + t]);
        }
      }
  }
  free(fl); free(fr);
}

/* src/transportroutines.c:477-492 (advectTwo): two upwindTwo stages and the
 * average fc = (f + stage2)/2 over interior cells 2..nX+1.  tmp holds nX+4 cells. */
void orc_advect_two(int N, int nX, const double *v, double mass, const double *x, const double *dx, double dt, int bc,
                    double T0, double T1, double Ma, double *f, double *fc, double *tmp) {
  long n3 = (long)N * N * N;
  orc_upwind_two(N, nX, v, mass, x, dx, dt, bc, T0, T1, Ma, f, tmp);
  orc_upwind_two(N, nX, v, mass, x, dx, dt, bc, T0, T1, Ma, tmp, fc);
  for (int l = 2; l < nX + 2; l++)
    for (long t = 0; t < n3; t++) fc[l * n3 + t] = 0.5 * (f[l * n3 + t] + fc[l * n3 + t]);
}
