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
// For N = 2^a 3^b up to 64 (every grid of the benchmark configurations) each
// axis is a Stockham FFT in LDS instead of the matrix: M factors as
// diag(q) DFT diag(p) with p_j = exp(i s j L_start delta) w_j (the reference's
// pre-twiddle and trapezoid weight) and q_k = exp(i s L_end a_k) (its
// post-twiddle), so pass A multiplies the plane by p(i) p(j) p(k), runs the z
// and y FFTs, and pass B runs the x FFT and multiplies by the prefactor and
// q(a) q(b) q(c): O(N^3 log N) per cube, as FFTW's.  Other N keep the matrix
// form (O(3 N^4) complex MACs per cube), which serves any N as FFTW does.
#include <hip/hip_runtime.h>
#include <math.h>
#include <type_traits>
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

// rho, u, T and the Maxwellian prefactor of one cell from the sums over its
// grid of w f, w v_x f, w v_y f, w v_z f, w |v|^2 f
// (src/momentRoutines.c:58-72, :116-142, :168-183).  The temperature's
// sum of w |v - u|^2 f is taken as sum w |v|^2 f - |sum w v f|^2 / sum w f,
// which equals the reference's two-pass sum up to rounding.
__device__ void moments_finish(const double (&s)[5], double hv, double mass, double kb, double *m) {
  const double dv3 = hv * hv * hv;
  const double rho = mass * dv3 * s[0];
  const double ux = s[1] / s[0], uy = s[2] / s[0], uz = s[3] / s[0];
  const double c2 = s[4] - (s[1] * ux + s[2] * uy + s[3] * uz);
  const double T = (mass * mass / kb) * dv3 * c2 / (3.0 * rho);
  m[0] = rho;
  m[1] = ux;
  m[2] = uy;
  m[3] = uz;
  m[4] = T;
  m[5] = rho * pow(0.5 / (M_PI * T), 1.5);
}

// The same from the cell's N (<= 64) per-plane partials, summed across the
// lanes of the calling wave; m is written by lane 0.
__device__ void moments_wave(int N, const double *__restrict__ part, double hv, double mass, double kb, double *m) {
  const int l = threadIdx.x & 63;
  double s[5];
  for (int q = 0; q < 5; ++q) s[q] = l < N ? part[(size_t)l * 5 + q] : 0.0;
  for (int q = 0; q < 5; ++q)
    for (int off = 32; off > 0; off >>= 1) s[q] += __shfl_xor(s[q], off, 64);
  if (l == 0) moments_finish(s, hv, mass, kb, m);
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
      else {
        const float2 h = make_float2((float)acc.x, (float)acc.y);
        ((float2 *)dst)[o] = h;
        ((float *)((float2 *)dst + ncell * n3))[o] = h.x + h.y;  // the parts' sums after the values
      }
    }
  }
}

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

// ------------------------------------------------------------------- FFTs
// Stage plan of a length-n FFT: radix-4 stages, then radix 2, then radix 3.
constexpr int fft_nstages(int n) {
  int s = 0;
  while (n % 4 == 0) { n /= 4; ++s; }
  while (n % 2 == 0) { n /= 2; ++s; }
  while (n % 3 == 0) { n /= 3; ++s; }
  return n == 1 ? s : -1;
}
constexpr int fft_radix(int n, int st) {
  int s = 0;
  while (n % 4 == 0) { if (s == st) return 4; n /= 4; ++s; }
  while (n % 2 == 0) { if (s == st) return 2; n /= 2; ++s; }
  while (n % 3 == 0) { if (s == st) return 3; n /= 3; ++s; }
  return 0;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}

// R-point DFT y_t = sum_r v_r exp(dsgn 2 pi i r t / R) in place.
template <int R>
__device__ __forceinline__ void butterfly(double2 (&v)[R], double dsgn) {
  if constexpr (R == 2) {
    const double2 a = v[0], b = v[1];
    v[0] = make_double2(a.x + b.x, a.y + b.y);
    v[1] = make_double2(a.x - b.x, a.y - b.y);
  } else if constexpr (R == 4) {
    // J = exp(dsgn i pi / 2) = dsgn i
    const double2 s02 = make_double2(v[0].x + v[2].x, v[0].y + v[2].y);
    const double2 d02 = make_double2(v[0].x - v[2].x, v[0].y - v[2].y);
    const double2 s13 = make_double2(v[1].x + v[3].x, v[1].y + v[3].y);
    const double2 d13 = make_double2(v[1].x - v[3].x, v[1].y - v[3].y);
    const double2 jd13 = make_double2(-dsgn * d13.y, dsgn * d13.x);  // J (v1 - v3)
    v[0] = make_double2(s02.x + s13.x, s02.y + s13.y);
    v[1] = make_double2(d02.x + jd13.x, d02.y + jd13.y);
    v[2] = make_double2(s02.x - s13.x, s02.y - s13.y);
    v[3] = make_double2(d02.x - jd13.x, d02.y - jd13.y);
  } else {
    // omega = exp(dsgn 2 pi i / 3) = -1/2 + dsgn i sqrt(3)/2
    const double h = 0.86602540378443864676 * dsgn;
    const double2 a = v[0], b = v[1], c = v[2];
    const double2 sbc = make_double2(b.x + c.x, b.y + c.y), dbc = make_double2(b.x - c.x, b.y - c.y);
    const double2 m = make_double2(a.x - 0.5 * sbc.x, a.y - 0.5 * sbc.y);
    v[0] = make_double2(a.x + sbc.x, a.y + sbc.y);
    v[1] = make_double2(m.x - h * dbc.y, m.y + h * dbc.x);
    v[2] = make_double2(m.x + h * dbc.y, m.y - h * dbc.x);
  }
}

// One Stockham stage of radix R (span Ns = product of the earlier radices) on
// nseq length-N sequences in LDS: element m of sequence q at q * qs + m * ms
// of src, results at the same places of dst.  tw[m] = exp(dsgn 2 pi i m / N).
template <int N, int R>
__device__ __forceinline__ void fft_stage(const double2 *src, double2 *dst, int qs, int ms, int nseq, int Ns,
                                          const double2 *tw, double dsgn) {
  constexpr int NB = N / R;
  const int step = N / (Ns * R);
  for (int it = threadIdx.x; it < nseq * NB; it += XT) {
    const int q = it / NB, j = it - q * NB;
    const int k = j % Ns;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[q * qs + (j + r * NB) * ms];
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[k * r * step]);
    butterfly<R>(v, dsgn);
    const int o = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[q * qs + (o + r * Ns) * ms] = v[r];
  }
}

// All stages of a length-N FFT along one axis of the plane in a[]; b[] is the
// other buffer.  Returns the buffer holding the result.
template <int N>
__device__ double2 *fft_axis(double2 *a, double2 *b, int qs, int ms, const double2 *tw, double dsgn) {
  constexpr int NS = fft_nstages(N);
  static_assert(NS > 0, "N must be 2^a 3^b");
  int Ns = 1;
#pragma unroll
  for (int st = 0; st < NS; ++st) {
    const int R = fft_radix(N, st);
    if (R == 4) fft_stage<N, 4>(a, b, qs, ms, N, Ns, tw, dsgn);
    else if (R == 2) fft_stage<N, 2>(a, b, qs, ms, N, Ns, tw, dsgn);
    else fft_stage<N, 3>(a, b, qs, ms, N, Ns, tw, dsgn);
    __syncthreads();
    Ns *= R;
    double2 *t = a;
    a = b;
    b = t;
  }
  return a;
}

// Per-axis factors of one direction: p[N], q[N], tw[N] (double2 each), in the
// grid's d_fft block as forward p, q, tw, then inverse p, q, tw.
struct FftTabs {
  double2 p[64], q[64], tw[64];
};

template <int N>
__device__ void load_tabs(FftTabs &t, const double2 *__restrict__ g) {
  for (int e = threadIdx.x; e < 3 * N; e += XT) {
    const int which = e / N, m = e - which * N;
    const double2 x = g[e];
    if (which == 0) t.p[m] = x;
    else if (which == 1) t.q[m] = x;
    else t.tw[m] = x;
  }
}

// Pass A with FFTs, real inputs (with the optional Maxwellian split): x-plane i
// of (cell, job) times p(i) p(j) p(k), then z and y FFTs, to scratch.
template <int N>
__global__ void __launch_bounds__(XT) k_fft_a_real(XformJobs jobs, int ncell, const double2 *__restrict__ tabs,
                                                   const double *__restrict__ v, const double *__restrict__ w,
                                                   double2 *__restrict__ scratch, double dsgn, double hv, double mass,
                                                   double kb) {
  __shared__ double2 A[N * N], B[N * N];
  __shared__ FftTabs t;
  __shared__ double smom[SPBTE_MOM_STRIDE];
  __shared__ double red[5 * (XT / 64)];
  __shared__ double sv[N], sw[N];
  constexpr int n2 = N * N;
  constexpr size_t n3 = (size_t)n2 * N;
  const int i = blockIdx.x, cell = blockIdx.y, jb = blockIdx.z;
  const spbte_xform_job &job = jobs.j[jb];
  load_tabs<N>(t, tabs);
  // the Maxwellian's moments: from the cell's per-plane partial sums, from
  // the whole cell of msrc (small grids: every workgroup of the job sums it),
  // or given
  const bool inl = N <= 16 && job.mode != 0 && !job.part && !job.mom;  // spbte_fft_moments_inline
  if (job.part && job.mode != 0 && threadIdx.x < 64)
    moments_wave(N, job.part + (size_t)cell * N * 5, hv, mass, kb, smom);
  if constexpr (N <= 16) if (inl) {  // all of the cell's loads issued at once, v and w from LDS
    constexpr int R = (int)((n3 + XT - 1) / XT);
    const double *ms = job.msrc + cell * n3;
    double fv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int e = threadIdx.x + r * XT;
      fv[r] = e < (int)n3 ? ms[e] : 0.0;
    }
    if (threadIdx.x < N) {
      sv[threadIdx.x] = v[threadIdx.x];
      sw[threadIdx.x] = w[threadIdx.x];
    }
    __syncthreads();
    double s5[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int e = threadIdx.x + r * XT;
      if (e < (int)n3) {
        const int a = e / n2, b = (e / N) % N, c = e % N;
        const double ww = sw[a] * sw[b] * sw[c] * fv[r];
        s5[0] += ww;
        s5[1] += sv[a] * ww;
        s5[2] += sv[b] * ww;
        s5[3] += sv[c] * ww;
        s5[4] += (sv[a] * sv[a] + sv[b] * sv[b] + sv[c] * sv[c]) * ww;
      }
    }
    block_sum<5>(s5, red);
    if (threadIdx.x == 0) moments_finish(s5, hv, mass, kb, smom);
  }
  __syncthreads();
  const double *mom = job.part || inl ? smom : (job.mom ? job.mom + (size_t)cell * SPBTE_MOM_STRIDE : nullptr);
  const double2 pi_ = t.p[i];
  for (int e = threadIdx.x; e < n2; e += XT) {
    const int j = e / N, k = e - j * N;
    double val;
    if (job.mode == 0) val = job.src[cell * n3 + (size_t)i * n2 + e];
    else if (job.mode == 1) val = job.src[cell * n3 + (size_t)i * n2 + e] - maxwellian_at(mom, v, i, j, k);
    else val = maxwellian_at(mom, v, i, j, k);
    const double2 pjk = cmul(cmul(pi_, t.p[j]), t.p[k]);
    A[e] = make_double2(val * pjk.x, val * pjk.y);
  }
  __syncthreads();
  double2 *r = fft_axis<N>(A, B, N, 1, t.tw, dsgn);        // z: sequences j (stride N), elements k
  double2 *o = (r == A) ? B : A;
  r = fft_axis<N>(r, o, 1, N, t.tw, dsgn);                 // y: sequences k, elements j (stride N)
  double2 *dst = scratch + ((size_t)jb * ncell + cell) * n3 + (size_t)i * n2;
  for (int e = threadIdx.x; e < n2; e += XT) dst[e] = r[e];
}

// Pass A with FFTs, complex FP64 inputs: the sum of nsl sets ncell N^3 apart
// (added in set order).  in may alias scratch: a workgroup's plane of set 0 is
// the only part of in that overlaps what it writes, and it reads all of its
// input before it writes.
template <int N>
__global__ void __launch_bounds__(XT) k_fft_a_cplx(const double2 *in, int nsl, int ncell,
                                                   const double2 *__restrict__ tabs, double2 *scratch, double dsgn) {
  __shared__ double2 A[N * N], B[N * N];
  __shared__ FftTabs t;
  constexpr int n2 = N * N;
  constexpr size_t n3 = (size_t)n2 * N;
  const int i = blockIdx.x, cell = blockIdx.y;
  load_tabs<N>(t, tabs);
  __syncthreads();
  const double2 pi_ = t.p[i];
  const size_t sstride = (size_t)ncell * n3;
  for (int e = threadIdx.x; e < n2; e += XT) {
    const int j = e / N, k = e - j * N;
    const size_t o = cell * n3 + (size_t)i * n2 + e;
    double2 x = in[o];
    for (int sl = 1; sl < nsl; sl += 8) {  // up to eight loads in flight, added in set order
      double2 y[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) y[u] = sl + u < nsl ? in[(size_t)(sl + u) * sstride + o] : make_double2(0.0, 0.0);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (sl + u < nsl) {
          x.x += y[u].x;
          x.y += y[u].y;
        }
    }
    A[e] = cmul(x, cmul(cmul(pi_, t.p[j]), t.p[k]));
  }
  __syncthreads();
  double2 *r = fft_axis<N>(A, B, N, 1, t.tw, dsgn);
  double2 *o = (r == A) ? B : A;
  r = fft_axis<N>(r, o, 1, N, t.tw, dsgn);
  double2 *dst = scratch + (size_t)cell * n3 + (size_t)i * n2;
  for (int e = threadIdx.x; e < n2; e += XT) dst[e] = r[e];
}

// Pass B with FFTs: the y index b plane (x, z) of scratch, x FFT, times
// pre q(a) q(b) q(c); OUT = 0 double2, 1 float2, 2 real part.
template <int N, int OUT>
__global__ void __launch_bounds__(XT) k_fft_b(int ncell, const double2 *__restrict__ tabs, double pre,
                                              const double2 *__restrict__ scratch, XformJobs jobs, void *out0,
                                              double dsgn, const double *__restrict__ v, const double *__restrict__ w,
                                              double pm, double *__restrict__ ppart) {
  __shared__ double2 A[N * N], B[N * N];
  __shared__ FftTabs t;
  constexpr int n2 = N * N;
  constexpr size_t n3 = (size_t)n2 * N;
  const int b = blockIdx.x, cell = blockIdx.y, jb = blockIdx.z;
  load_tabs<N>(t, tabs);
  const double2 *src = scratch + ((size_t)jb * ncell + cell) * n3;
  for (int e = threadIdx.x; e < n2; e += XT) {
    const int i = e / N, c = e - i * N;
    A[e] = src[(size_t)i * n2 + b * N + c];
  }
  __syncthreads();
  double2 *r = fft_axis<N>(A, B, 1, N, t.tw, dsgn);        // x: sequences c, elements i (stride N)
  void *dst = jobs.j[jb].dst ? jobs.j[jb].dst : out0;
  const double2 qb = t.q[b];
  // OUT == 2 with ppart: the one-species projection's partial sums of this
  // y-plane, p {1, v_x, v_y, v_z, |v|^2 / 2} q with p = w h_v^3 m
  double ps[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int e = threadIdx.x; e < n2; e += XT) {
    const int a = e / N, c = e - a * N;
    const size_t o = cell * n3 + (size_t)a * n2 + b * N + c;
    double2 f = cmul(cmul(cmul(t.q[a], qb), t.q[c]), r[e]);
    f.x *= pre;
    f.y *= pre;
    if (OUT == 0) ((double2 *)dst)[o] = f;
    else if (OUT == 1) {
      // FP32 operands: the complex values, then (after ncell N^3 of them) the
      // sums of their parts, which the FP32-product sum takes as G's a + b
      const float2 h = make_float2((float)f.x, (float)f.y);
      ((float2 *)dst)[o] = h;
      ((float *)((float2 *)dst + ncell * n3))[o] = h.x + h.y;
    } else {
      ((double *)dst)[o] = f.x;
      if (ppart) {
        const double p = w[a] * w[b] * w[c] * pm * f.x;
        ps[0] += p;
        ps[1] += p * v[a];
        ps[2] += p * v[b];
        ps[3] += p * v[c];
        ps[4] += p * 0.5 * (v[a] * v[a] + v[b] * v[b] + v[c] * v[c]);
      }
    }
  }
  if (OUT == 2 && ppart) {
    __shared__ double red[5 * (XT / 64)];
    block_sum<5>(ps, red);
    if (threadIdx.x == 0)
      for (int q = 0; q < 5; ++q) ppart[((size_t)cell * N + b) * 5 + q] = ps[q];
  }
}

// ---------------------------------------------------------------- reductions

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
  double pm[SPBTE_MAX_NS];  // h_v^3 m per species
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

// ---- the same moments and projection for small batches: one workgroup per
// (x-plane, cell) sums its plane into a partial (PS values), and a second pass
// combines the N partials of a cell, so one cell keeps N workgroups busy
// instead of one.
constexpr int PS = 5;

// Partial sums of w f, w v_x f, w v_y f, w v_z f, w |v|^2 f over x-plane i.
__global__ void __launch_bounds__(XT) k_moments_part(int N, const double *__restrict__ f, const double *__restrict__ v,
                                                     const double *__restrict__ w, double *__restrict__ part) {
  __shared__ double red[PS * (XT / 64)];
  const int i = blockIdx.x, cell = blockIdx.y;
  const int n2 = N * N;
  const double *fp = f + (size_t)cell * n2 * N + (size_t)i * n2;
  const double vi = v[i], wi = w[i];
  double s[PS] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int e = threadIdx.x; e < n2; e += XT) {
    const int j = e / N, k = e - j * N;
    const double ww = wi * w[j] * w[k] * fp[e];
    s[0] += ww;
    s[1] += vi * ww;
    s[2] += v[j] * ww;
    s[3] += v[k] * ww;
    s[4] += (vi * vi + v[j] * v[j] + v[k] * v[k]) * ww;
  }
  block_sum<PS>(s, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < PS; ++q) part[((size_t)cell * N + i) * PS + q] = s[q];
}

// rho, u, T and the Maxwellian prefactor from a cell's N partials.
__global__ void __launch_bounds__(64) k_moments_fin(int N, int ncell, const double *__restrict__ part, double hv,
                                                    double mass, double kb, double *__restrict__ mom) {
  const int cell = blockIdx.x * 64 + threadIdx.x;
  if (cell >= ncell) return;
  double s[PS] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = 0; i < N; ++i)
    for (int q = 0; q < PS; ++q) s[q] += part[((size_t)cell * N + i) * PS + q];
  moments_finish(s, hv, mass, kb, mom + (size_t)cell * SPBTE_MOM_STRIDE);
}

// b <- (C C^T)^-1 b from the host's LU factors with their row interchanges
// (forward elimination, then back substitution; src/conserve.c:173-202).
template <int NE>
__device__ __forceinline__ void proj_solve(double (&b)[NE], const ProjArgs &pa) {
  for (int k = 0; k < NE - 1; ++k) {
    const int p = pa.pivot[k];
    if (p != k) { const double t = b[p]; b[p] = b[k]; b[k] = t; }
    for (int r = k + 1; r < NE; ++r) b[r] -= pa.LU[r * NE + k] * b[k];
  }
  b[NE - 1] = b[NE - 1] / pa.LU[(NE - 1) * NE + NE - 1];
  for (int r = NE - 2; r >= 0; --r) {
    double sum = 0.0;
    for (int c = r + 1; c < NE; ++c) sum += pa.LU[r * NE + c] * b[c];
    b[r] = 1.0 / pa.LU[r * NE + r] * (b[r] - sum);
  }
}

// Ns species, groups of Ns*Ns arrays at Q + grp * gstride, array n*Ns+m
// (species m's operator against species n) pstride after the group's first.
// Partial sums over x-plane i of p_m {1 (species m), v_x, v_y, v_z, |v|^2/2}
// sum_n Q_{n*Ns+m}, p_m = w h_v^3 m_m: the rows of C (src/conserve.c:282-296),
// NS + 4 values per (plane, group).
template <int NS>
__global__ void __launch_bounds__(XT) k_proj_part(int N, const double *__restrict__ v, const double *__restrict__ w,
                                                  ProjArgs pa, const double *__restrict__ Q, size_t gstride,
                                                  size_t pstride, double *__restrict__ part) {
  constexpr int NE = NS + 4;
  __shared__ double red[NE * (XT / 64)];
  const int i = blockIdx.x, grp = blockIdx.y;
  const int n2 = N * N;
  const double *qp = Q + (size_t)grp * gstride + (size_t)i * n2;
  const double vi = v[i];
  double wi[NS];
  for (int m = 0; m < NS; ++m) wi[m] = w[i] * pa.pm[m];
  double s[NE];
  for (int q = 0; q < NE; ++q) s[q] = 0.0;
  for (int e = threadIdx.x; e < n2; e += XT) {
    const int j = e / N, k = e - j * N;
    for (int m = 0; m < NS; ++m) {
      double qs = qp[(size_t)m * pstride + e];
      for (int n = 1; n < NS; ++n) qs += qp[(size_t)(n * NS + m) * pstride + e];
      const double p = wi[m] * w[j] * w[k] * qs;
      s[m] += p;
      s[NS] += p * vi;
      s[NS + 1] += p * v[j];
      s[NS + 2] += p * v[k];
      s[NS + 3] += p * 0.5 * (vi * vi + v[j] * v[j] + v[k] * v[k]);
    }
  }
  block_sum<NE>(s, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < NE; ++q) part[((size_t)grp * N + i) * NE + q] = s[q];
}

// Q <- Q - C^T (C C^T)^-1 C Q on x-plane i (src/conserve.c:207-264, the solve
// of :173-202): every workgroup of a group combines the group's N partials and
// solves the (NS+4)-square system itself; species m's correction is shared
// equally by its NS operators.
template <int NS>
__global__ void __launch_bounds__(XT) k_proj_apply(int N, const double *__restrict__ v, const double *__restrict__ w,
                                                   ProjArgs pa, double *__restrict__ Q, size_t gstride,
                                                   size_t pstride, const double *__restrict__ part) {
  constexpr int NE = NS + 4;
  __shared__ double lam[NE];
  const int i = blockIdx.x, grp = blockIdx.y;
  const int n2 = N * N;
  if (threadIdx.x < 64) {  // wave 0 adds the partials (N <= 64), lane 0 solves
    const int l = threadIdx.x;
    double b[NE];
    for (int q = 0; q < NE; ++q) b[q] = l < N ? part[((size_t)grp * N + l) * NE + q] : 0.0;
    for (int q = 0; q < NE; ++q)
      for (int off = 32; off > 0; off >>= 1) b[q] += __shfl_xor(b[q], off, 64);
    if (l == 0) {
      proj_solve<NE>(b, pa);
      for (int q = 0; q < NE; ++q) lam[q] = b[q];
    }
  }
  __syncthreads();
  double *qp = Q + (size_t)grp * gstride + (size_t)i * n2;
  const double vi = v[i];
  double wi[NS];
  for (int m = 0; m < NS; ++m) wi[m] = w[i] * pa.pm[m];
  for (int e = threadIdx.x; e < n2; e += XT) {
    const int j = e / N, k = e - j * N;
    const double e2 = 0.5 * (vi * vi + v[j] * v[j] + v[k] * v[k]);
    for (int m = 0; m < NS; ++m) {
      const double p = wi[m] * w[j] * w[k] / NS;
      const double corr = p * lam[m] + p * vi * lam[NS] + p * v[j] * lam[NS + 1] + p * v[k] * lam[NS + 2] +
                          p * e2 * lam[NS + 3];
      for (int n = 0; n < NS; ++n) qp[(size_t)(n * NS + m) * pstride + e] -= corr;
    }
  }
}

// ---- the inverse transform's last pass over whole cubes for small grids
// (N^2 <= XT): one workgroup per cube, thread t owning the x-line t in
// registers, so the real part can be projected in the same launch (the
// projection needs the whole cube's sums).

// In-register length-N FFT (the Stockham stages of fft_axis, unrolled):
// x_t <- sum_r x_r exp(dsgn 2 pi i r t / N), tw[m] = exp(dsgn 2 pi i m / N).
template <int N, int ST, int SPAN>
__device__ __forceinline__ void dft_reg_stage(double2 (&a)[N], double2 (&b)[N], const double2 *tw, double dsgn) {
  if constexpr (ST < fft_nstages(N)) {
    constexpr int R = fft_radix(N, ST);
    constexpr int NB = N / R;
    constexpr int step = N / (SPAN * R);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int k = j % SPAN;
      double2 v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = a[j + r * NB];
#pragma unroll
      for (int r = 1; r < R; ++r)
        if (k * r * step != 0) v[r] = cmul(v[r], tw[k * r * step]);
      butterfly<R>(v, dsgn);
      const int o = (j / SPAN) * SPAN * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) b[o + r * SPAN] = v[r];
    }
    dft_reg_stage<N, ST + 1, SPAN * R>(b, a, tw, dsgn);
  }
}

template <int N>
__device__ __forceinline__ void dft_reg(double2 (&x)[N], const double2 *tw, double dsgn) {
  double2 y[N];
  dft_reg_stage<N, 0, 1>(x, y, tw, dsgn);
  if constexpr (fft_nstages(N) % 2 == 1) {
#pragma unroll
    for (int m = 0; m < N; ++m) x[m] = y[m];
  }
}

// Last pass of the inverse transform of whole cubes: pass A (k_fft_a_cplx)
// has done z and y; workgroup `cell` does x, keeps the real part and, with
// PROJ, projects it (src/conserve.c:207-264 with one species: the moments'
// sums, the 5 x 5 solve and the correction, all in this workgroup) before the
// one store of Q.  With ep.out, also out = a x + b y + c Q over the same
// cells (a time step's update; x and out may be the same array).
template <int N, bool PROJ>
__global__ void __launch_bounds__(XT) k_fft_cube_inv_x(int ncell, const double2 *__restrict__ tabs, double pre,
                                                       const double2 *__restrict__ scratch, double *__restrict__ q,
                                                       const double *__restrict__ v, const double *__restrict__ w,
                                                       double pm, ProjArgs pa, spbte_epilogue ep) {
  static_assert(N * N <= XT, "one line per thread");
  constexpr int n2 = N * N;
  constexpr size_t n3 = (size_t)n2 * N;
  constexpr double dsgn = 1.0;
  __shared__ FftTabs t;
  __shared__ double red[5 * (XT / 64)];
  __shared__ double lam[5];
  const int cell = blockIdx.x, tid = threadIdx.x;
  const bool own = tid < n2;
  const int b = own ? tid / N : 0, c = own ? tid - b * N : 0;
  load_tabs<N>(t, tabs);
  double2 x[N];
#pragma unroll
  for (int a = 0; a < N; ++a) x[a] = own ? scratch[cell * n3 + (size_t)a * n2 + tid] : make_double2(0.0, 0.0);
  // the update's a x + b y, loaded with the transform's input (every load
  // before the first store: out may alias x)
  double r[N];
  if (ep.out && own) {
#pragma unroll
    for (int a = 0; a < N; ++a) r[a] = ep.a * ep.x[cell * n3 + (size_t)a * n2 + tid];
    if (ep.y) {
#pragma unroll
      for (int a = 0; a < N; ++a) r[a] += ep.b * ep.y[cell * n3 + (size_t)a * n2 + tid];
    }
  }
  __syncthreads();
  double qv[N];
  dft_reg<N>(x, t.tw, dsgn);
#pragma unroll
  for (int a = 0; a < N; ++a) qv[a] = pre * cmul(cmul(cmul(t.q[a], t.q[b]), t.q[c]), x[a]).x;
  if constexpr (PROJ) {
    double ps[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (own) {
#pragma unroll
      for (int a = 0; a < N; ++a) {
        const double p = w[a] * w[b] * w[c] * pm * qv[a];
        ps[0] += p;
        ps[1] += p * v[a];
        ps[2] += p * v[b];
        ps[3] += p * v[c];
        ps[4] += p * 0.5 * (v[a] * v[a] + v[b] * v[b] + v[c] * v[c]);
      }
    }
    block_sum<5>(ps, red);
    if (tid == 0) {
      proj_solve<5>(ps, pa);
      for (int m = 0; m < 5; ++m) lam[m] = ps[m];
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < N; ++a) {
      const double p = w[a] * pm * w[b] * w[c];
      const double e2 = 0.5 * (v[a] * v[a] + v[b] * v[b] + v[c] * v[c]);
      qv[a] -= p * lam[0] + p * v[a] * lam[1] + p * v[b] * lam[2] + p * v[c] * lam[3] + p * e2 * lam[4];
    }
  }
  if (own) {
#pragma unroll
    for (int a = 0; a < N; ++a) q[cell * n3 + (size_t)a * n2 + tid] = qv[a];
    if (ep.out) {
#pragma unroll
      for (int a = 0; a < N; ++a) ep.out[cell * n3 + (size_t)a * n2 + tid] = r[a] + ep.c * qv[a];
    }
  }
}

// FFT-capable grids (N = 2^a 3^b with both LDS buffers of a plane in 64 KiB).
constexpr int kFftN[] = {6, 8, 9, 12, 16, 18, 24, 32, 36, 48};

template <int N>
void launch_fft_fwd(const spbte_grid_dev *g, const XformJobs &J, int njobs, int ncell, int out_fp32, void *scratch,
                    hipStream_t s) {
  const double2 *tf = (const double2 *)g->d_fft;  // forward p, q, tw
  hipLaunchKernelGGL(k_fft_a_real<N>, dim3(N, ncell, njobs), dim3(XT), 0, s, J, ncell, tf, (const double *)g->d_v,
                     (const double *)g->d_w, (double2 *)scratch, -1.0, g->hv, g->mass, g->kb);
  if (out_fp32)
    hipLaunchKernelGGL((k_fft_b<N, 1>), dim3(N, ncell, njobs), dim3(XT), 0, s, ncell, tf, g->pre_fwd,
                       (const double2 *)scratch, J, nullptr, -1.0, nullptr, nullptr, 0.0, nullptr);
  else
    hipLaunchKernelGGL((k_fft_b<N, 0>), dim3(N, ncell, njobs), dim3(XT), 0, s, ncell, tf, g->pre_fwd,
                       (const double2 *)scratch, J, nullptr, -1.0, nullptr, nullptr, 0.0, nullptr);
}

// Complex cube transform (invert 0/1, OUT 0 = complex, 2 = real part, with
// the projection's partials into ppart when it is not null).
template <int N>
void launch_fft_cplx(const spbte_grid_dev *g, const void *in, int nsl, void *out, int ncell, int invert, int real_out,
                     void *scratch, double *ppart, hipStream_t s) {
  const double2 *t = (const double2 *)g->d_fft + (invert ? 3 * N : 0);
  const double dsgn = invert ? 1.0 : -1.0, pre = invert ? g->pre_inv : g->pre_fwd;
  XformJobs J = {};
  hipLaunchKernelGGL(k_fft_a_cplx<N>, dim3(N, ncell, 1), dim3(XT), 0, s, (const double2 *)in, nsl, ncell, t,
                     (double2 *)scratch, dsgn);
  const double pm = g->hv * g->hv * g->hv * g->mass;
  if (real_out)
    hipLaunchKernelGGL((k_fft_b<N, 2>), dim3(N, ncell, 1), dim3(XT), 0, s, ncell, t, pre, (const double2 *)scratch, J,
                       out, dsgn, (const double *)g->d_v, (const double *)g->d_w, pm, ppart);
  else
    hipLaunchKernelGGL((k_fft_b<N, 0>), dim3(N, ncell, 1), dim3(XT), 0, s, ncell, t, pre, (const double2 *)scratch, J,
                       out, dsgn, nullptr, nullptr, 0.0, nullptr);
}

// Grids with the whole-cube last pass (N^2 <= XT), used when a launch has at
// most kCubeMaxWg cubes: more cubes fill the GPU with the plane passes as well.
constexpr int kCubeN[] = {6, 8, 9, 12, 16};
constexpr int kCubeMaxWg = 512;

template <typename F, int I = 0>
bool with_cube_n(int N, F &&fn) {
  if constexpr (I < (int)(sizeof kCubeN / sizeof kCubeN[0])) {
    if (N == kCubeN[I]) {
      fn(std::integral_constant<int, kCubeN[I]>{});
      return true;
    }
    return with_cube_n<F, I + 1>(N, static_cast<F &&>(fn));
  } else {
    return false;
  }
}

template <int N>
void launch_cube_inv(const spbte_grid_dev *g, const void *Qh, int nsl, double *q, int ncell, void *scratch,
                     const ProjArgs *pa, const spbte_epilogue &ep, hipStream_t s) {
  const double2 *t = (const double2 *)g->d_fft + 3 * N;  // inverse p, q, tw
  hipLaunchKernelGGL(k_fft_a_cplx<N>, dim3(N, ncell, 1), dim3(XT), 0, s, (const double2 *)Qh, nsl, ncell, t,
                     (double2 *)scratch, 1.0);
  const double *v = (const double *)g->d_v, *w = (const double *)g->d_w;
  if (pa)
    hipLaunchKernelGGL((k_fft_cube_inv_x<N, true>), dim3(ncell), dim3(XT), 0, s, ncell, t, g->pre_inv,
                       (const double2 *)scratch, q, v, w, pa->pm[0], *pa, ep);
  else
    hipLaunchKernelGGL((k_fft_cube_inv_x<N, false>), dim3(ncell), dim3(XT), 0, s, ncell, t, g->pre_inv,
                       (const double2 *)scratch, q, v, w, 0.0, ProjArgs{}, ep);
}

// Runs fn<N>() for the grid's N if it has FFT kernels; false otherwise.
template <typename F, int I = 0>
bool with_fft_n(int N, F &&fn) {
  if constexpr (I < (int)(sizeof kFftN / sizeof kFftN[0])) {
    if (N == kFftN[I]) {
      fn(std::integral_constant<int, kFftN[I]>{});
      return true;
    }
    return with_fft_n<F, I + 1>(N, static_cast<F &&>(fn));
  } else {
    return false;
  }
}

}  // namespace

extern "C" int spbte_launch_moments(const spbte_grid_dev *g, const double *f, double *mom, int ncell, double *part,
                                    hipStream_t s) {
  if (!part) {
    hipLaunchKernelGGL(k_moments, dim3(ncell), dim3(XT), 0, s, g->N, f, (const double *)g->d_v, (const double *)g->d_w,
                       g->hv, g->mass, g->kb, mom);
    return (int)hipGetLastError();
  }
  hipLaunchKernelGGL(k_moments_part, dim3(g->N, ncell), dim3(XT), 0, s, g->N, f, (const double *)g->d_v,
                     (const double *)g->d_w, part);
  hipLaunchKernelGGL(k_moments_fin, dim3((ncell + 63) / 64), dim3(64), 0, s, g->N, ncell, (const double *)part, g->hv,
                     g->mass, g->kb, mom);
  return (int)hipGetLastError();
}

extern "C" int spbte_launch_moment_partials(const spbte_grid_dev *g, const double *f, int ncell, double *part,
                                            hipStream_t s) {
  hipLaunchKernelGGL(k_moments_part, dim3(g->N, ncell), dim3(XT), 0, s, g->N, f, (const double *)g->d_v,
                     (const double *)g->d_w, part);
  return (int)hipGetLastError();
}

extern "C" int spbte_launch_xform_fwd(const spbte_grid_dev *g, const spbte_xform_job *jobs, int njobs, int ncell,
                                      int out_fp32, void *scratch, hipStream_t s) {
  if (njobs < 1 || njobs > 4) return (int)hipErrorInvalidValue;
  XformJobs J = {};
  for (int q = 0; q < njobs; ++q) J.j[q] = jobs[q];
  for (int q = 0; q < njobs; ++q)
    if (jobs[q].mode != 0 && !jobs[q].mom && !jobs[q].part &&
        (!jobs[q].msrc || !g->d_fft || !spbte_fft_moments_inline(g->N)))
      return (int)hipErrorInvalidValue;  // msrc: FFT grids with N <= 16 only
  if (g->d_fft && with_fft_n(g->N, [&](auto n) {
        launch_fft_fwd<decltype(n)::value>(g, J, njobs, ncell, out_fp32, scratch, s);
      }))
    return (int)hipGetLastError();
  for (int q = 0; q < njobs; ++q)
    if (jobs[q].part) return (int)hipErrorInvalidValue;  // the matrix form takes moments, not partials
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
  if (g->d_fft && with_fft_n(g->N, [&](auto n) {
        launch_fft_cplx<decltype(n)::value>(g, in, 1, out, ncell, invert, 0, scratch, nullptr, s);
      }))
    return (int)hipGetLastError();
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

extern "C" int spbte_launch_xform_inv_real(const spbte_grid_dev *g, const void *Qh, int nsl, double *q, int ncell,
                                           void *scratch, double *proj_part, hipStream_t s) {
  if (nsl < 1) return (int)hipErrorInvalidValue;
  if (g->d_fft && with_fft_n(g->N, [&](auto n) {
        launch_fft_cplx<decltype(n)::value>(g, Qh, nsl, q, ncell, 1, 1, scratch, proj_part, s);
      }))
    return (int)hipGetLastError();
  if (proj_part || nsl != 1) return (int)hipErrorInvalidValue;  // the matrix form takes one set, writes no partials
  const int N = g->N;
  const size_t lds_a = 2 * (size_t)N * N * sizeof(double2), lds_b = (size_t)N * N * sizeof(double2);
  XformJobs J = {};
  hipLaunchKernelGGL(k_xform_a_cplx, dim3(N, ncell, 1), dim3(XT), lds_a, s, N, (const double2 *)Qh, ncell,
                     (const double2 *)g->d_mat_inv, (double2 *)scratch);
  hipLaunchKernelGGL(k_xform_b<2>, dim3(N, ncell, 1), dim3(XT), lds_b, s, N, ncell, (const double2 *)g->d_mat_inv,
                     g->pre_inv, (const double2 *)scratch, J, (void *)q);
  return (int)hipGetLastError();
}

static ProjArgs proj_args(const spbte_grid_dev *g, int Ns, const double *masses, const double *LU, const int *pivot) {
  ProjArgs pa = {};
  const int nel = Ns + 4;
  for (int q = 0; q < nel * nel; ++q) pa.LU[q] = LU[q];
  for (int q = 0; q < nel; ++q) pa.pivot[q] = pivot[q];
  for (int q = 0; q < Ns; ++q) {
    pa.masses[q] = masses[q];
    pa.pm[q] = g->hv * g->hv * g->hv * masses[q];
  }
  return pa;
}

template <int NS>
void launch_proj_split(const spbte_grid_dev *g, const ProjArgs &pa, double *Q, int ngroups, size_t gstride,
                       size_t pstride, double *part, hipStream_t s) {
  hipLaunchKernelGGL(k_proj_part<NS>, dim3(g->N, ngroups), dim3(XT), 0, s, g->N, (const double *)g->d_v,
                     (const double *)g->d_w, pa, (const double *)Q, gstride, pstride, part);
  hipLaunchKernelGGL(k_proj_apply<NS>, dim3(g->N, ngroups), dim3(XT), 0, s, g->N, (const double *)g->d_v,
                     (const double *)g->d_w, pa, Q, gstride, pstride, (const double *)part);
}

extern "C" int spbte_launch_project(const spbte_grid_dev *g, int Ns, const double *masses, const double *LU, const int *pivot,
                                    double *Q, int ncell, double *part, hipStream_t s) {
  const size_t n3 = (size_t)g->N * g->N * g->N;
  return spbte_launch_project_strided(g, Ns, masses, LU, pivot, Q, ncell, (size_t)Ns * Ns * n3, n3, part, s);
}

extern "C" int spbte_launch_project_apply(const spbte_grid_dev *g, const double *masses, const double *LU,
                                          const int *pivot, double *Q, int ncell, const double *part, hipStream_t s) {
  const size_t n3 = (size_t)g->N * g->N * g->N;
  const ProjArgs pa = proj_args(g, 1, masses, LU, pivot);
  hipLaunchKernelGGL(k_proj_apply<1>, dim3(g->N, ncell), dim3(XT), 0, s, g->N, (const double *)g->d_v,
                     (const double *)g->d_w, pa, Q, n3, n3, part);
  return (int)hipGetLastError();
}

// Groups at Q + grp * gstride, the Ns^2 pairs of a group pstride apart
// (group-major: gstride = Ns^2 N^3, pstride = N^3; pair-major over ncell
// cells: gstride = N^3, pstride = ncell N^3).  With a partial-sum buffer
// ((Ns+4) N doubles per group) every (x-plane, group) is a workgroup of its
// own; without one, each group is one workgroup.
extern "C" int spbte_launch_project_strided(const spbte_grid_dev *g, int Ns, const double *masses, const double *LU,
                                            const int *pivot, double *Q, int ngroups, size_t gstride, size_t pstride,
                                            double *part, hipStream_t s) {
  if (Ns < 1 || Ns > SPBTE_MAX_NS) return (int)hipErrorInvalidValue;
  const ProjArgs pa = proj_args(g, Ns, masses, LU, pivot);
  if (part && ngroups <= 65535) {
    switch (Ns) {
      case 1: launch_proj_split<1>(g, pa, Q, ngroups, gstride, pstride, part, s); break;
      case 2: launch_proj_split<2>(g, pa, Q, ngroups, gstride, pstride, part, s); break;
      case 3: launch_proj_split<3>(g, pa, Q, ngroups, gstride, pstride, part, s); break;
      case 4: launch_proj_split<4>(g, pa, Q, ngroups, gstride, pstride, part, s); break;
      case 5: launch_proj_split<5>(g, pa, Q, ngroups, gstride, pstride, part, s); break;
      case 6: launch_proj_split<6>(g, pa, Q, ngroups, gstride, pstride, part, s); break;
      case 7: launch_proj_split<7>(g, pa, Q, ngroups, gstride, pstride, part, s); break;
      default: launch_proj_split<8>(g, pa, Q, ngroups, gstride, pstride, part, s); break;
    }
    return (int)hipGetLastError();
  }
  hipLaunchKernelGGL(k_project, dim3(ngroups), dim3(XT), 0, s, g->N, Ns, (const double *)g->d_v, (const double *)g->d_w, g->hv,
                     pa, Q, gstride, pstride);
  return (int)hipGetLastError();
}

// Grids small enough for every forward-transform workgroup of a Maxwellian job
// to sum its cell's moments itself (msrc jobs), which saves the partial-sum pass.
extern "C" int spbte_fft_moments_inline(int N) { return N <= 16 && spbte_fft_supported(N); }

extern "C" int spbte_fft_cube(int N, int ncubes) {
  return ncubes >= 1 && ncubes <= kCubeMaxWg && with_cube_n(N, [](auto) {}) ? 1 : 0;
}

extern "C" int spbte_launch_xform_inv_real_cube(const spbte_grid_dev *g, const void *Qh, int nsl, double *q, int ncell,
                                                void *scratch, const double *masses, const double *LU, const int *pivot,
                                                const spbte_epilogue *epi, hipStream_t s) {
  if (nsl < 1 || !g->d_fft || !spbte_fft_cube(g->N, ncell)) return (int)hipErrorInvalidValue;
  const ProjArgs pa = masses ? proj_args(g, 1, masses, LU, pivot) : ProjArgs{};
  const spbte_epilogue ep = epi ? *epi : spbte_epilogue{};
  with_cube_n(g->N, [&](auto n) {
    launch_cube_inv<decltype(n)::value>(g, Qh, nsl, q, ncell, scratch, masses ? &pa : nullptr, ep, s);
  });
  return (int)hipGetLastError();
}

// 1 if the transform of an N grid runs as FFTs (the grid then needs the factor
// tables of spbte_grid_dev.d_fft).
extern "C" int spbte_fft_supported(int N) {
  return with_fft_n(N, [](auto) {}) ? 1 : 0;
}
