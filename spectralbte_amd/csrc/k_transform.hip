// k_transform.hip -- batched twiddled 3D transform, Maxwellian split and the
// FP64 mass/momentum/energy projection for the MI355X collision operator.
//
// Transform (reference: src/collisions.c:232-283, fft3D).  The reference's
// pre-twiddle, FFTW transform and post-twiddle factor per axis, so the whole
// transform is prefactor * (M (x) M (x) M) applied to the N^3 cube, with the
// per-axis N x N complex matrix
//   M[k][j] = exp(i s L_end a_k) exp(-+2 pi i jk/N) exp(i s j L_start delta) w_j
// built once on the host in long double (spbte_core.c).  On the GPU the cube of
// one cell never sits in LDS whole (512 KiB FP64 at N=32); it goes through two
// LDS-resident passes:
//   pass A, one workgroup per (x-plane i, cell, input): the plane X[i][:][:]
//     (N^2 complex) is staged in LDS, transformed along z then y in LDS, and
//     written to an FP64 scratch cube;
//   pass B, one workgroup per (y index b, cell, input): the N^2 values
//     T[:][b][:] are staged in LDS, transformed along x, scaled and written in
//     the consumer's format (FP64 complex, FP32 complex, or the real part).
// Work is O(3 N^4) complex MACs per cube: 0.2% of the O(N^6) weighted sum at
// N=32, so the matrix form (which also serves odd and non-power-of-two N, as
// FFTW does for the reference) costs nothing measurable.
#include <hip/hip_runtime.h>
#include <math.h>
#include "spbte_internal.h"

namespace {

constexpr int XT = 256;  // threads per transform workgroup

__device__ __forceinline__ double2 cmac(double2 acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
  return acc;
}

struct XformJobs {
  spbte_xform_job j[4];
};

// Maxwellian of src/collisions.c:98-103 at grid point (i, j, k).
__device__ __forceinline__ double maxwellian_at(const double *mom, const double *v, int i, int j, int k) {
  const double ux = mom[1], uy = mom[2], uz = mom[3], T = mom[4], pre = mom[5];
  const double dx = v[i] - ux, dy = v[j] - uy, dz = v[k] - uz;
  return pre * exp(-(0.5 / T) * (dx * dx + dy * dy + dz * dz));
}

// z then y transform of one x-plane; T2 goes to scratch (FP64 complex).
__device__ void plane_zy(int N, const double2 *__restrict__ M, double2 *X, double2 *T1, double2 *__restrict__ dst) {
  const int n2 = N * N;
  for (int e = threadIdx.x; e < n2; e += XT) {
    const int j = e / N, c = e - j * N;
    double2 acc = make_double2(0.0, 0.0);
    for (int k = 0; k < N; ++k) acc = cmac(acc, M[c * N + k], X[j * N + k]);
    T1[e] = acc;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n2; e += XT) {
    const int b = e / N, c = e - b * N;
    double2 acc = make_double2(0.0, 0.0);
    for (int j = 0; j < N; ++j) acc = cmac(acc, M[b * N + j], T1[j * N + c]);
    dst[e] = acc;
  }
}

// Pass A for real inputs (with the optional Maxwellian split).
__global__ void __launch_bounds__(XT) k_xform_a_real(int N, XformJobs jobs, int ncell, const double2 *__restrict__ M,
                                                     const double *__restrict__ v, double2 *__restrict__ scratch) {
  extern __shared__ double2 sm[];
  const int n2 = N * N;
  const size_t n3 = (size_t)n2 * N;
  const int i = blockIdx.x, cell = blockIdx.y, jb = blockIdx.z;
  const spbte_xform_job &job = jobs.j[jb];
  double2 *X = sm, *T1 = sm + n2;
  const double *mom = job.mom ? job.mom + (size_t)cell * SPBTE_MOM_STRIDE : nullptr;
  for (int e = threadIdx.x; e < n2; e += XT) {
    const int j = e / N, k = e - j * N;
    double val;
    if (job.mode == 0) val = job.src[cell * n3 + (size_t)i * n2 + e];
    else if (job.mode == 1) val = job.src[cell * n3 + (size_t)i * n2 + e] - maxwellian_at(mom, v, i, j, k);
    else val = maxwellian_at(mom, v, i, j, k);
    X[e] = make_double2(val, 0.0);
  }
  __syncthreads();
  plane_zy(N, M, X, T1, scratch + ((size_t)jb * ncell + cell) * n3 + (size_t)i * n2);
}

// Pass A for complex FP64 inputs.
__global__ void __launch_bounds__(XT) k_xform_a_cplx(int N, const double2 *__restrict__ in, int ncell,
                                                     const double2 *__restrict__ M, double2 *__restrict__ scratch) {
  extern __shared__ double2 sm[];
  const int n2 = N * N;
  const size_t n3 = (size_t)n2 * N;
  const int i = blockIdx.x, cell = blockIdx.y;
  double2 *X = sm, *T1 = sm + n2;
  for (int e = threadIdx.x; e < n2; e += XT) X[e] = in[cell * n3 + (size_t)i * n2 + e];
  __syncthreads();
  plane_zy(N, M, X, T1, scratch + (size_t)cell * n3 + (size_t)i * n2);
}

// Pass B: x transform of the y-plane b; OUT = 0 double2, 1 float2, 2 real double.
template <int OUT>
__global__ void __launch_bounds__(XT) k_xform_b(int N, int ncell, const double2 *__restrict__ M, double pre,
                                                const double2 *__restrict__ scratch, XformJobs jobs, void *out0) {
  extern __shared__ double2 sm[];
  const int n2 = N * N;
  const size_t n3 = (size_t)n2 * N;
  const int b = blockIdx.x, cell = blockIdx.y, jb = blockIdx.z;
  const double2 *src = scratch + ((size_t)jb * ncell + cell) * n3;
  double2 *P = sm;
  for (int e = threadIdx.x; e < n2; e += XT) {
    const int i = e / N, c = e - i * N;
    P[e] = src[(size_t)i * n2 + b * N + c];
  }
  __syncthreads();
  void *dst = jobs.j[jb].dst ? jobs.j[jb].dst : out0;
  for (int e = threadIdx.x; e < n2; e += XT) {
    const int a = e / N, c = e - a * N;
    const size_t o = cell * n3 + (size_t)a * n2 + b * N + c;
    if (OUT == 2) {
      double acc = 0.0;
      for (int i = 0; i < N; ++i) {
        const double2 m = M[a * N + i], p = P[i * N + c];
        acc = fma(m.x, p.x, acc);
        acc = fma(-m.y, p.y, acc);
      }
      ((double *)dst)[o] = pre * acc;
    } else {
      double2 acc = make_double2(0.0, 0.0);
      for (int i = 0; i < N; ++i) acc = cmac(acc, M[a * N + i], P[i * N + c]);
      acc.x *= pre;
      acc.y *= pre;
      if (OUT == 0) ((double2 *)dst)[o] = acc;
      else ((float2 *)dst)[o] = make_float2((float)acc.x, (float)acc.y);
    }
  }
}

// ---------------------------------------------------------------- reductions
template <int NV>
__device__ void block_sum(double (&x)[NV], double *red /* NV * (XT/64) */) {
  for (int q = 0; q < NV; ++q) {
    double s = x[q];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    x[q] = s;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
    for (int q = 0; q < NV; ++q) red[q * (XT / 64) + wave] = x[q];
  __syncthreads();
  for (int q = 0; q < NV; ++q) {
    double s = 0.0;
    for (int w = 0; w < XT / 64; ++w) s += red[q * (XT / 64) + w];
    x[q] = s;
  }
  __syncthreads();
}

// Moments for the Maxwellian split (src/collisions.c:95-98 with
// src/momentRoutines.c:58-72, :116-142, :168-183): rho, u, T, prefactor.
__global__ void __launch_bounds__(XT) k_moments(int N, const double *__restrict__ f, const double *__restrict__ v,
                                                const double *__restrict__ w, double hv, double mass, double kb,
                                                double *__restrict__ mom) {
  __shared__ double red[4 * (XT / 64)];
  const int cell = blockIdx.x;
  const int n2 = N * N;
  const size_t n3 = (size_t)n2 * N;
  const double *fc = f + cell * n3;
  const double dv3 = hv * hv * hv;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int e = threadIdx.x; e < (int)n3; e += XT) {
    const int i = e / n2, j = (e / N) % N, k = e % N;
    const double ww = w[i] * w[j] * w[k] * fc[e];
    s[0] += ww;
    s[1] += v[i] * ww;
    s[2] += v[j] * ww;
    s[3] += v[k] * ww;
  }
  block_sum<4>(s, red);
  const double rho = mass * dv3 * s[0];
  const double ux = mass * dv3 * s[1] / rho, uy = mass * dv3 * s[2] / rho, uz = mass * dv3 * s[3] / rho;
  double t[1] = {0.0};
  for (int e = threadIdx.x; e < (int)n3; e += XT) {
    const int i = e / n2, j = (e / N) % N, k = e % N;
    const double dx = v[i] - ux, dy = v[j] - uy, dz = v[k] - uz;
    t[0] += (dx * dx + dy * dy + dz * dz) * w[i] * w[j] * w[k] * fc[e];
  }
  block_sum<1>(t, red);
  if (threadIdx.x == 0) {
    const double T = (mass * mass / kb) * dv3 * t[0] / (3.0 * rho);
    double *m = mom + (size_t)cell * SPBTE_MOM_STRIDE;
    m[0] = rho;
    m[1] = ux;
    m[2] = uy;
    m[3] = uz;
    m[4] = T;
    m[5] = rho * pow(0.5 / (M_PI * T), 1.5);
  }
}

// Projection (src/conserve.c:207-264 with the solve of :173-202).  One
// workgroup per group of Ns*Ns arrays; array n*Ns+m is species m's operator
// against species n.  LU/pivots: (Ns+4)^2 scaled-pivot factors from the host.
struct ProjArgs {
  double LU[(SPBTE_MAX_NS + 4) * (SPBTE_MAX_NS + 4)];
  double masses[SPBTE_MAX_NS];
  int pivot[SPBTE_MAX_NS + 4];
};

__global__ void __launch_bounds__(XT) k_project(int N, int Ns, const double *__restrict__ v, const double *__restrict__ w,
                                                double hv, ProjArgs pa, double *__restrict__ Q, size_t gstride,
                                                size_t pstride) {
  __shared__ double red[(SPBTE_MAX_NS + 4) * (XT / 64)];
  __shared__ double lam[SPBTE_MAX_NS + 4];
  const int grp = blockIdx.x;
  const int n2 = N * N;
  const size_t n3 = (size_t)n2 * N;
  const int nel = Ns + 4;
  double *Qg = Q + (size_t)grp * gstride;  // pair p of this group at Qg + p * pstride
  const double dv3 = hv * hv * hv;
  double s[SPBTE_MAX_NS + 4];
  for (int q = 0; q < SPBTE_MAX_NS + 4; ++q) s[q] = 0.0;
  for (int e = threadIdx.x; e < (int)n3; e += XT) {
    const int i = e / n2, j = (e / N) % N, k = e % N;
    const double wv = w[i] * w[j] * w[k] * dv3;
    const double e2 = 0.5 * (v[i] * v[i] + v[j] * v[j] + v[k] * v[k]);
    for (int m = 0; m < Ns; ++m) {
      double qs = 0.0;
      for (int n = 0; n < Ns; ++n) qs += Qg[(size_t)(n * Ns + m) * pstride + e];
      const double p = wv * pa.masses[m] * qs;
      s[m] += p;
      s[Ns] += p * v[i];
      s[Ns + 1] += p * v[j];
      s[Ns + 2] += p * v[k];
      s[Ns + 3] += p * e2;
    }
  }
  block_sum<SPBTE_MAX_NS + 4>(s, red);
  if (threadIdx.x == 0) {
    double b[SPBTE_MAX_NS + 4];
    for (int q = 0; q < nel; ++q) b[q] = s[q];
    // forward elimination with the recorded row interchanges, then back substitution
    for (int k = 0; k < nel - 1; ++k) {
      const int p = pa.pivot[k];
      if (p != k) { const double t = b[p]; b[p] = b[k]; b[k] = t; }
      for (int i = k + 1; i < nel; ++i) b[i] -= pa.LU[i * nel + k] * b[k];
    }
    b[nel - 1] = b[nel - 1] / pa.LU[(nel - 1) * nel + nel - 1];
    for (int i = nel - 2; i >= 0; --i) {
      double sum = 0.0;
      for (int j = i + 1; j < nel; ++j) sum += pa.LU[i * nel + j] * b[j];
      b[i] = 1.0 / pa.LU[i * nel + i] * (b[i] - sum);
    }
    for (int q = 0; q < nel; ++q) lam[q] = b[q];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < (int)n3; e += XT) {
    const int i = e / n2, j = (e / N) % N, k = e % N;
    const double wv = w[i] * w[j] * w[k] * dv3;
    const double e2 = 0.5 * (v[i] * v[i] + v[j] * v[j] + v[k] * v[k]);
    for (int m = 0; m < Ns; ++m) {
      const double p = wv * pa.masses[m] / Ns;
      const double corr = p * lam[m] + p * v[i] * lam[Ns] + p * v[j] * lam[Ns + 1] + p * v[k] * lam[Ns + 2] + p * e2 * lam[Ns + 3];
      for (int n = 0; n < Ns; ++n) Qg[(size_t)(n * Ns + m) * pstride + e] -= corr;
    }
  }
}

}  // namespace

extern "C" int spbte_launch_moments(const spbte_grid_dev *g, const double *f, double *mom, int ncell, hipStream_t s) {
  hipLaunchKernelGGL(k_moments, dim3(ncell), dim3(XT), 0, s, g->N, f, (const double *)g->d_v, (const double *)g->d_w, g->hv,
                     g->mass, g->kb, mom);
  return (int)hipGetLastError();
}

extern "C" int spbte_launch_xform_fwd(const spbte_grid_dev *g, const spbte_xform_job *jobs, int njobs, int ncell,
                                      int out_fp32, void *scratch, hipStream_t s) {
  if (njobs < 1 || njobs > 4) return (int)hipErrorInvalidValue;
  XformJobs J = {};
  for (int q = 0; q < njobs; ++q) J.j[q] = jobs[q];
  const int N = g->N;
  const size_t lds_a = 2 * (size_t)N * N * sizeof(double2), lds_b = (size_t)N * N * sizeof(double2);
  hipLaunchKernelGGL(k_xform_a_real, dim3(N, ncell, njobs), dim3(XT), lds_a, s, N, J, ncell,
                     (const double2 *)g->d_mat_fwd, (const double *)g->d_v, (double2 *)scratch);
  if (out_fp32)
    hipLaunchKernelGGL(k_xform_b<1>, dim3(N, ncell, njobs), dim3(XT), lds_b, s, N, ncell, (const double2 *)g->d_mat_fwd,
                       g->pre_fwd, (const double2 *)scratch, J, nullptr);
  else
    hipLaunchKernelGGL(k_xform_b<0>, dim3(N, ncell, njobs), dim3(XT), lds_b, s, N, ncell, (const double2 *)g->d_mat_fwd,
                       g->pre_fwd, (const double2 *)scratch, J, nullptr);
  return (int)hipGetLastError();
}

extern "C" int spbte_launch_xform_fwd_complex(const spbte_grid_dev *g, const void *in, void *out, int ncell, int invert,
                                              void *scratch, hipStream_t s) {
  const int N = g->N;
  const size_t lds_a = 2 * (size_t)N * N * sizeof(double2), lds_b = (size_t)N * N * sizeof(double2);
  const double2 *M = (const double2 *)(invert ? g->d_mat_inv : g->d_mat_fwd);
  const double pre = invert ? g->pre_inv : g->pre_fwd;
  XformJobs J = {};
  hipLaunchKernelGGL(k_xform_a_cplx, dim3(N, ncell, 1), dim3(XT), lds_a, s, N, (const double2 *)in, ncell, M,
                     (double2 *)scratch);
  hipLaunchKernelGGL(k_xform_b<0>, dim3(N, ncell, 1), dim3(XT), lds_b, s, N, ncell, M, pre, (const double2 *)scratch, J, out);
  return (int)hipGetLastError();
}

extern "C" int spbte_launch_xform_inv_real(const spbte_grid_dev *g, const void *Qh, double *q, int ncell, void *scratch,
                                           hipStream_t s) {
  const int N = g->N;
  const size_t lds_a = 2 * (size_t)N * N * sizeof(double2), lds_b = (size_t)N * N * sizeof(double2);
  XformJobs J = {};
  hipLaunchKernelGGL(k_xform_a_cplx, dim3(N, ncell, 1), dim3(XT), lds_a, s, N, (const double2 *)Qh, ncell,
                     (const double2 *)g->d_mat_inv, (double2 *)scratch);
  hipLaunchKernelGGL(k_xform_b<2>, dim3(N, ncell, 1), dim3(XT), lds_b, s, N, ncell, (const double2 *)g->d_mat_inv,
                     g->pre_inv, (const double2 *)scratch, J, (void *)q);
  return (int)hipGetLastError();
}

extern "C" int spbte_launch_project(const spbte_grid_dev *g, int Ns, const double *masses, const double *LU, const int *pivot,
                                    double *Q, int ncell, hipStream_t s) {
  const size_t n3 = (size_t)g->N * g->N * g->N;
  return spbte_launch_project_strided(g, Ns, masses, LU, pivot, Q, ncell, (size_t)Ns * Ns * n3, n3, s);
}

// Groups at Q + grp * gstride, the Ns^2 pairs of a group pstride apart
// (group-major: gstride = Ns^2 N^3, pstride = N^3; pair-major over ncell
// cells: gstride = N^3, pstride = ncell N^3).
extern "C" int spbte_launch_project_strided(const spbte_grid_dev *g, int Ns, const double *masses, const double *LU,
                                            const int *pivot, double *Q, int ngroups, size_t gstride, size_t pstride,
                                            hipStream_t s) {
  if (Ns < 1 || Ns > SPBTE_MAX_NS) return (int)hipErrorInvalidValue;
  const int ncell = ngroups;
  ProjArgs pa = {};
  const int nel = Ns + 4;
  for (int q = 0; q < nel * nel; ++q) pa.LU[q] = LU[q];
  for (int q = 0; q < nel; ++q) pa.pivot[q] = pivot[q];
  for (int q = 0; q < Ns; ++q) pa.masses[q] = masses[q];
  hipLaunchKernelGGL(k_project, dim3(ncell), dim3(XT), 0, s, g->N, Ns, (const double *)g->d_v, (const double *)g->d_w, g->hv,
                     pa, Q, gstride, pstride);
  return (int)hipGetLastError();
}
