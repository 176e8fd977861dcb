// k_transport.hip -- on-device transport stage of the 1D-3V solver and the
// elementwise updates of the time integrators.
//
// Reference: src/transportroutines.c (upwindOne :94-238, upwindTwo :241-470,
// advectTwo :477-492, minmod :80-90), src/boundaryConditions.c:39-85
// (setDiffuseReflectionBC), exec/boltz.c:193-241 and :285-354 (updates).
// Every cell of the mesh, ghosts included, is one contiguous N^3 block of FP64;
// the kernels are elementwise over (cell, velocity) and bound by HBM, apart from
// the wall kernel, which reduces the outgoing half of one N^3 block per wall.
#include <hip/hip_runtime.h>
#include <math.h>
#include "spbte_internal.h"

namespace {

constexpr int ET = 256;   // threads of the elementwise kernels
constexpr int WT = 1024;  // threads of the wall kernel

__device__ __forceinline__ double minmod3(double a, double b, double c) {
  if (a > 0 && b > 0 && c > 0) return fmin(fmin(a, b), c);
  if (a < 0 && b < 0 && c < 0) return fmax(fmax(a, b), c);
  return 0.0;
}

// Slope of cell l (minmod of the left, right and centred differences).
__device__ __forceinline__ double slope(const double *f, size_t n3, size_t t, int l, const double *x) {
  const double fm = f[(size_t)(l - 1) * n3 + t], f0 = f[(size_t)l * n3 + t], fp = f[(size_t)(l + 1) * n3 + t];
  return minmod3((f0 - fm) / (x[l] - x[l - 1]), (fp - f0) / (x[l + 1] - x[l]), (fp - fm) / (x[l + 1] - x[l - 1]));
}

// ---- order 1 (upwindOne): fc[l] for l = 1..nX from cells l-1, l, l+1.
__global__ void __launch_bounds__(ET) k_upwind1(int N, int nX, const double *__restrict__ f, double *__restrict__ fc,
                                                const double *__restrict__ v, const double *__restrict__ dx, double dt) {
  const size_t n3 = (size_t)N * N * N, total = (size_t)nX * n3;
  const size_t e = (size_t)blockIdx.x * ET + threadIdx.x;
  if (e >= total) return;
  const int l = (int)(e / n3) + 1;
  const size_t t = e - (size_t)(l - 1) * n3;
  const int i = (int)(t / ((size_t)N * N));
  const double cfl = dt * v[i] / dx[l];
  const double *fl = f + (size_t)l * n3 + t;
  fc[(size_t)l * n3 + t] = (i < N / 2) ? (1.0 + cfl) * fl[0] - cfl * fl[n3] : (1.0 - cfl) * fl[0] + cfl * fl[-(long)n3];
}

// ---- ghost cells.  mode 0: copy neighbours (f[0] = f[1], f[nX+1] = f[nX]);
// mode 2: periodic (f[0] = f[nX], f[nX+1] = f[1]); mode 3: order-2 linear
// extrapolation f[1] = 2 f[2] - f[3], f[nX+2] = 2 f[nX+1] - f[nX].
__global__ void __launch_bounds__(ET) k_ghosts(int N, int nX, int mode, double *__restrict__ f) {
  const size_t n3 = (size_t)N * N * N;
  const size_t e = (size_t)blockIdx.x * ET + threadIdx.x;
  if (e >= 2 * n3) return;
  const int side = (int)(e / n3);
  const size_t t = e - side * n3;
  double *c = f + t;
  if (mode == 3) {
    if (side == 0) c[n3] = 2.0 * c[2 * n3] - c[3 * n3];
    else c[(size_t)(nX + 2) * n3] = 2.0 * c[(size_t)(nX + 1) * n3] - c[(size_t)nX * n3];
  } else if (mode == 2) {
    if (side == 0) c[0] = c[(size_t)nX * n3];
    else c[(size_t)(nX + 1) * n3] = c[n3];
  } else {
    if (side == 0) c[0] = c[n3];
    else c[(size_t)(nX + 1) * n3] = c[(size_t)nX * n3];
  }
}

// ---- diffuse reflection (setDiffuseReflectionBC): one workgroup per wall.
// The outgoing half of `src` (left wall: v_x < 0, i < N/2; right wall: i >= N/2)
// sets sigma_W; the incoming half of `dst` becomes sigma_W times the wall
// Maxwellian.  src may equal dst: the halves are disjoint.
struct WallJobs {
  const double *src[2];
  double *dst[2];
  double TW[2];
  int right[2];
};

__global__ void __launch_bounds__(WT) k_diffuse(int N, WallJobs J, const double *__restrict__ v,
                                                const double *__restrict__ w, double hv, double mass, double kb) {
  __shared__ double red[WT / 64];
  const int wi = blockIdx.x;
  const double *in = J.src[wi];
  double *out = J.dst[wi];
  const int right = J.right[wi];
  const double TW = J.TW[wi];
  const int n2 = N * N, half = (N / 2) * n2;
  const int o0 = right ? half : 0, i0 = right ? 0 : half;
  const double h3 = hv * hv * hv;
  double s = 0.0;
  for (int e = threadIdx.x; e < half; e += WT) {
    const int t = o0 + e, i = t / n2, j = (t / N) % N, k = t % N;
    s += v[i] * w[i] * w[j] * w[k] * h3 * in[t];
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  double sig = 0.0;
  for (int q = 0; q < WT / 64; ++q) sig += red[q];
  sig *= (right ? 1.0 : -1.0) * sqrt(2.0 * M_PI * mass / (kb * TW));
  const double pre = sig * pow(0.5 * mass / (M_PI * kb * TW), 1.5);
  const double a = -0.5 * mass / (kb * TW);
  for (int e = threadIdx.x; e < half; e += WT) {
    const int t = i0 + e, i = t / n2, j = (t / N) % N, k = t % N;
    out[t] = pre * exp(a * (v[i] * v[i] + v[j] * v[j] + v[k] * v[k]));
  }
}

// ---- order 2, wall states (upwindTwo :349-404).  fl: left-wall state from
// cell 2 (outgoing half always; incoming half too unless diffuse walls, where
// k_diffuse sets it), fr: right-wall state from cell nX+1, symmetrically.
// diffuse: bit 0 = the left wall reflects diffusely, bit 1 = the right wall;
// a diffuse wall's incoming half is left for k_diffuse.
__global__ void __launch_bounds__(ET) k_walls2(int N, int nX, int diffuse, const double *__restrict__ f,
                                               double *__restrict__ fl, double *__restrict__ fr,
                                               const double *__restrict__ x, const double *__restrict__ dx) {
  const size_t n3 = (size_t)N * N * N;
  const size_t e = (size_t)blockIdx.x * ET + threadIdx.x;
  if (e >= 2 * n3) return;
  const int side = (int)(e / n3);
  const size_t t = e - side * n3;
  const int i = (int)(t / ((size_t)N * N));
  if (side == 0) {
    const bool out_half = i < N / 2;
    if (!out_half && (diffuse & 1)) return;
    const double s = slope(f, n3, t, 2, x);
    fl[t] = out_half ? f[2 * n3 + t] - 0.5 * dx[2] * s : f[2 * n3 + t] + 0.5 * dx[2] * s;
  } else {
    const bool out_half = i >= N / 2;
    if (!out_half && (diffuse & 2)) return;
    const double s = slope(f, n3, t, nX + 1, x);
    const double c = f[(size_t)(nX + 1) * n3 + t];
    fr[t] = out_half ? c + 0.5 * dx[nX + 1] * s : c - 0.5 * dx[nX + 1] * s;
  }
}

// ---- order 2, MUSCL half-step (upwindTwo :406-468) for l = 2..nX+1.
// fy != 0 adds the Poiseuille body force along v_y (:429-437, :458-466):
// fc -= fy (f[j+1] - f[j-1]), with -fy f[1] at j = 0 and -fy f[N-2] at
// j = N-1 as the reference writes them; fy = Ma dt / (4 h_v).
__global__ void __launch_bounds__(ET) k_upwind2(int N, int nX, const double *__restrict__ f, double *__restrict__ fc,
                                                const double *__restrict__ fl, const double *__restrict__ fr,
                                                const double *__restrict__ v, const double *__restrict__ x,
                                                const double *__restrict__ dx, double dt, double fy) {
  const size_t n3 = (size_t)N * N * N, total = (size_t)nX * n3;
  const size_t e = (size_t)blockIdx.x * ET + threadIdx.x;
  if (e >= total) return;
  const int l = (int)(e / n3) + 2;
  const size_t t = e - (size_t)(l - 2) * n3;
  const int i = (int)(t / ((size_t)N * N));
  const double f0 = f[(size_t)l * n3 + t];
  const double s1 = slope(f, n3, t, l, x);
  const double cfl = 0.5 * dt * v[i] / dx[l];
  double r;
  if (i >= N / 2) {  // from the left
    if (l == 2) {
      r = f0 - cfl * (f0 + 0.5 * dx[l] * s1 - fl[t]);
    } else {
      const double s0 = slope(f, n3, t, l - 1, x);
      r = f0 - cfl * (f0 + 0.5 * dx[l] * s1 - (f[(size_t)(l - 1) * n3 + t] + 0.5 * dx[l - 1] * s0));
    }
  } else {           // from the right
    if (l == nX + 1) {
      r = f0 - cfl * (fr[t] - (f0 - 0.5 * dx[l] * s1));
    } else {
      const double s2 = slope(f, n3, t, l + 1, x);
      r = f0 - cfl * (f[(size_t)(l + 1) * n3 + t] - 0.5 * dx[l + 1] * s2 - (f0 - 0.5 * dx[l] * s1));
    }
  }
  if (fy != 0.0) {
    const int j = (int)((t / N) % N);
    const double *fj = f + (size_t)l * n3 + t;
    if (j == 0) r = r - fy * fj[N];
    else if (j == N - 1) r = r - fy * fj[-N];
    else r = r - fy * (fj[N] - fj[-N]);
  }
  fc[(size_t)l * n3 + t] = r;
}

// ---- out = a x + b y + c z over n elements (y, z may be null when b, c are 0).
__global__ void __launch_bounds__(ET) k_lincomb(size_t n, double a, const double *__restrict__ x, double b,
                                                const double *__restrict__ y, double c, const double *__restrict__ z,
                                                double *out) {
  const size_t e = (size_t)blockIdx.x * ET + threadIdx.x;
  if (e >= n) return;
  double r = a * x[e];
  if (y) r += b * y[e];
  if (z) r += c * z[e];
  out[e] = r;
}

inline unsigned blocks(size_t n) { return (unsigned)((n + ET - 1) / ET); }

}  // namespace

extern "C" int spbte_launch_upwind1(int N, int nX, const double *f, double *fc, const double *v, const double *dx,
                                    double dt, hipStream_t s) {
  const size_t n = (size_t)nX * N * N * N;
  hipLaunchKernelGGL(k_upwind1, dim3(blocks(n)), dim3(ET), 0, s, N, nX, f, fc, v, dx, dt);
  return (int)hipGetLastError();
}

extern "C" int spbte_launch_ghosts(int N, int nX, int mode, double *f, hipStream_t s) {
  const size_t n = 2 * (size_t)N * N * N;
  hipLaunchKernelGGL(k_ghosts, dim3(blocks(n)), dim3(ET), 0, s, N, nX, mode, f);
  return (int)hipGetLastError();
}

extern "C" int spbte_launch_diffuse(const spbte_grid_dev *g, int nwall, const double *const *src, double *const *dst,
                                    const double *TW, const int *right, hipStream_t s) {
  if (nwall < 1 || nwall > 2) return (int)hipErrorInvalidValue;
  WallJobs J = {};
  for (int q = 0; q < nwall; ++q) {
    J.src[q] = src[q];
    J.dst[q] = dst[q];
    J.TW[q] = TW[q];
    J.right[q] = right[q];
  }
  hipLaunchKernelGGL(k_diffuse, dim3(nwall), dim3(WT), 0, s, g->N, J, (const double *)g->d_v, (const double *)g->d_w,
                     g->hv, g->mass, g->kb);
  return (int)hipGetLastError();
}

extern "C" int spbte_launch_walls2(int N, int nX, int diffuse, const double *f, double *fl, double *fr, const double *x,
                                   const double *dx, hipStream_t s) {
  const size_t n = 2 * (size_t)N * N * N;
  hipLaunchKernelGGL(k_walls2, dim3(blocks(n)), dim3(ET), 0, s, N, nX, diffuse, f, fl, fr, x, dx);
  return (int)hipGetLastError();
}

extern "C" int spbte_launch_upwind2(int N, int nX, const double *f, double *fc, const double *fl, const double *fr,
                                    const double *v, const double *x, const double *dx, double dt, double fy,
                                    hipStream_t s) {
  const size_t n = (size_t)nX * N * N * N;
  hipLaunchKernelGGL(k_upwind2, dim3(blocks(n)), dim3(ET), 0, s, N, nX, f, fc, fl, fr, v, x, dx, dt, fy);
  return (int)hipGetLastError();
}

extern "C" int spbte_launch_lincomb(size_t n, double a, const double *x, double b, const double *y, double c,
                                    const double *z, double *out, hipStream_t s) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_lincomb, dim3(blocks(n)), dim3(ET), 0, s, n, a, x, b, y, c, z, out);
  return (int)hipGetLastError();
}

// ---- output quantities of src/output.c per cell: out[c] = {E+, E-, slice[0..N)}
// with E+- the positive and negative parts of sum dv^3 w w w f |v|^2
// (src/momentRoutines.c:146-165) and slice[l] = f[N/2 + N (N/2 + N l)]
// (src/output.c:203-206).  One workgroup per cell.
__global__ void __launch_bounds__(ET) k_out_quantities(int N, const double *__restrict__ f, const double *__restrict__ v,
                                                       const double *__restrict__ w, double hv,
                                                       double *__restrict__ out) {
  __shared__ double red[2][ET / 64];
  const int c = blockIdx.x;
  const size_t n3 = (size_t)N * N * N;
  const double *fc = f + c * n3;
  const double h3 = hv * hv * hv;
  double pos = 0.0, neg = 0.0;
  for (int e = threadIdx.x; e < (int)n3; e += ET) {
    const int i = e / (N * N), j = (e / N) % N, k = e % N;
    const double en = h3 * w[i] * w[j] * w[k] * fc[e] * (v[i] * v[i] + v[j] * v[j] + v[k] * v[k]);
    if (en > 0) pos += en;
    else neg -= en;
  }
  for (int off = 32; off > 0; off >>= 1) {
    pos += __shfl_down(pos, off, 64);
    neg += __shfl_down(neg, off, 64);
  }
  if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = pos; red[1][threadIdx.x >> 6] = neg; }
  __syncthreads();
  double *o = out + (size_t)c * (2 + N);
  if (threadIdx.x == 0) {
    double p = 0.0, q = 0.0;
    for (int s = 0; s < ET / 64; ++s) { p += red[0][s]; q += red[1][s]; }
    o[0] = p;
    o[1] = q;
  }
  for (int l = threadIdx.x; l < N; l += ET) o[2 + l] = fc[N / 2 + (size_t)N * (N / 2 + (size_t)N * l)];
}

extern "C" int spbte_launch_out_quantities(const spbte_grid_dev *g, const double *f, int ncell, double *out,
                                           hipStream_t s) {
  hipLaunchKernelGGL(k_out_quantities, dim3(ncell), dim3(ET), 0, s, g->N, f, (const double *)g->d_v,
                     (const double *)g->d_w, g->hv, out);
  return (int)hipGetLastError();
}
