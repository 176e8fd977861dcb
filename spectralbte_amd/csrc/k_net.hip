// k_net.hip -- kernels of the learned collision surrogate: a BF16 MFMA GEMM with
// FP32 accumulation (forward and backward products of the network), and the
// elementwise kernels around it (spectral features, lift, GELU, head, loss,
// column sums, Adam).  No counterpart in the reference.
//
// GEMM: C[i][j] (+)= sum_k A(i,k) B(k,j), A(i,k) = A[i*sa0 + k*sa1],
// B(k,j) = B[k*sb0 + j*sb1], FP32 in global memory.  Operands are rounded to
// BF16 (round to nearest even) on their way into LDS and multiplied with
// v_mfma_f32_16x16x32_bf16; sums are FP32.  A workgroup computes a 64x64 tile
// with 4 waves (32x32 each, 2x2 MFMA tiles); K advances 32 per LDS stage.
// Split-K: grid.z slices of K write FP32 partial tiles that k_splitk_reduce
// sums with the epilogue.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include "spbte_internal.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 32;
constexpr int KP = TK + 8;  // padded k extent of an LDS row (bf16): 80-byte rows

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

struct GemmArgs {
  const float *A, *B;
  float *C;           // output (or the split-K partials)
  const float *add;   // addend with C's layout (may equal C), or null
  const float *bias;  // per column j, or null
  float *G;           // GELU(C) output (epilogue 1), or null
  long sa0, sa1, sb0, sb1, sc0, sc1;
  int M, N, K;
  int kslice;         // K per split (multiple of TK)
  int vecA, vecB;     // 16-byte aligned rows along the contiguous dimension
};

// Register image of one thread's share of a K-tile: 8 values of A and 8 of B.
struct Stage {
  float a[8], b[8];
};

// Thread t covers, for A: row-major (sa1 == 1) row t/4, k (t%4)*8 .. +8;
// column-major row (t%8)*8 .. +8, k t/8.  For B: k-contiguous (sb0 == 1)
// column t/4, k (t%4)*8 .. +8; row-major column (t%8)*8 .. +8, k t/8.
__device__ __forceinline__ void load_stage(const GemmArgs &g, Stage &r, int i0, int j0, int k0, int kend) {
  const int t = threadIdx.x;
  if (g.sa1 == 1) {
    const int gi = i0 + (t >> 2), gk = k0 + (t & 3) * 8;
    const float *p = g.A + gi * g.sa0 + gk;
    if (g.vecA && gi < g.M && gk + 8 <= kend) {
      const float4 u = *(const float4 *)p, v = *(const float4 *)(p + 4);
      r.a[0] = u.x; r.a[1] = u.y; r.a[2] = u.z; r.a[3] = u.w; r.a[4] = v.x; r.a[5] = v.y; r.a[6] = v.z; r.a[7] = v.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) r.a[e] = (gi < g.M && gk + e < kend) ? p[e] : 0.0f;
    }
  } else {
    const int gi = i0 + (t & 7) * 8, gk = k0 + (t >> 3);
    const float *p = g.A + gi + gk * g.sa1;
    if (g.vecA && gi + 8 <= g.M && gk < kend) {
      const float4 u = *(const float4 *)p, v = *(const float4 *)(p + 4);
      r.a[0] = u.x; r.a[1] = u.y; r.a[2] = u.z; r.a[3] = u.w; r.a[4] = v.x; r.a[5] = v.y; r.a[6] = v.z; r.a[7] = v.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) r.a[e] = (gi + e < g.M && gk < kend) ? p[e] : 0.0f;
    }
  }
  if (g.sb0 == 1) {
    const int gj = j0 + (t >> 2), gk = k0 + (t & 3) * 8;
    const float *p = g.B + gk + gj * g.sb1;
    if (g.vecB && gj < g.N && gk + 8 <= kend) {
      const float4 u = *(const float4 *)p, v = *(const float4 *)(p + 4);
      r.b[0] = u.x; r.b[1] = u.y; r.b[2] = u.z; r.b[3] = u.w; r.b[4] = v.x; r.b[5] = v.y; r.b[6] = v.z; r.b[7] = v.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) r.b[e] = (gj < g.N && gk + e < kend) ? p[e] : 0.0f;
    }
  } else {
    const int gj = j0 + (t & 7) * 8, gk = k0 + (t >> 3);
    const float *p = g.B + gk * g.sb0 + gj;
    if (g.vecB && gj + 8 <= g.N && gk < kend) {
      const float4 u = *(const float4 *)p, v = *(const float4 *)(p + 4);
      r.b[0] = u.x; r.b[1] = u.y; r.b[2] = u.z; r.b[3] = u.w; r.b[4] = v.x; r.b[5] = v.y; r.b[6] = v.z; r.b[7] = v.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) r.b[e] = (gj + e < g.N && gk < kend) ? p[e] : 0.0f;
    }
  }
}

// Round to BF16 and write into the k-contiguous LDS tiles sA[i][k], sB[j][k].
__device__ __forceinline__ void store_stage(const GemmArgs &g, const Stage &r, __bf16 (*sA)[KP], __bf16 (*sB)[KP]) {
  const int t = threadIdx.x;
  if (g.sa1 == 1) {
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (__bf16)r.a[e];
    *(bf16x8 *)&sA[t >> 2][(t & 3) * 8] = v;
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) sA[(t & 7) * 8 + e][t >> 3] = (__bf16)r.a[e];
  }
  if (g.sb0 == 1) {
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (__bf16)r.b[e];
    *(bf16x8 *)&sB[t >> 2][(t & 3) * 8] = v;
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) sB[(t & 7) * 8 + e][t >> 3] = (__bf16)r.b[e];
  }
}

__device__ __forceinline__ void epilogue(const GemmArgs &g, int i, int j, float v) {
  const long o = i * g.sc0 + j * g.sc1;
  float out = g.add ? g.add[o] + v : v;
  if (g.bias) out += g.bias[j];
  g.C[o] = out;
  if (g.G) g.G[o] = gelu(out);
}

// The K-loop keeps the next tile's global loads in registers while the MFMAs
// consume the current LDS buffer; two LDS buffers, one barrier per K-tile.
// split: write FP32 partials [z][i][j] instead of the epilogue.
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g, int split) {
  __shared__ __attribute__((aligned(16))) __bf16 sA[2][TM][KP];
  __shared__ __attribute__((aligned(16))) __bf16 sB[2][TN][KP];
  const int i0 = blockIdx.y * TM, j0 = blockIdx.x * TN;
  const int kbeg = blockIdx.z * g.kslice;
  const int kend = min(g.K, kbeg + g.kslice);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
  Stage st;
  load_stage(g, st, i0, j0, kbeg, kend);
  store_stage(g, st, sA[0], sB[0]);
  __syncthreads();
  int cur = 0;
  for (int k0 = kbeg; k0 < kend; k0 += TK) {
    const bool more = k0 + TK < kend;
    if (more) load_stage(g, st, i0, j0, k0 + TK, kend);
    // lane l: row/col (l % 16), k group 8 (l / 16) .. + 8
    const int r = lane & 15, kg = (lane >> 4) * 8;
    bf16x8 av[2], bv[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) av[a] = *(const bf16x8 *)&sA[cur][wi + 16 * a + r][kg];
#pragma unroll
    for (int b = 0; b < 2; ++b) bv[b] = *(const bf16x8 *)&sB[cur][wj + 16 * b + r][kg];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
    if (more) store_stage(g, st, sA[cur ^ 1], sB[cur ^ 1]);
    __syncthreads();
    cur ^= 1;
  }
  // D layout: lane l holds rows 4 (l / 16) + q, column l % 16 of each 16x16 tile
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = i0 + wi + 16 * a + 4 * (lane >> 4) + q, j = j0 + wj + 16 * b + (lane & 15);
        if (i >= g.M || j >= g.N) continue;
        const float v = acc[a][b][q];
        if (split) g.C[((long)blockIdx.z * g.M + i) * g.N + j] = v;
        else epilogue(g, i, j, v);
      }
}

__global__ void __launch_bounds__(256) k_splitk_reduce(GemmArgs g, const float *__restrict__ part, int nsplit) {
  const long n = (long)g.M * g.N;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const int i = (int)(e / g.N), j = (int)(e - (long)i * g.N);
    float v = 0.f;
    for (int z = 0; z < nsplit; ++z) v += part[(long)z * n + e];
    epilogue(g, i, j, v);
  }
}

// ---- elementwise kernels -------------------------------------------------
// dZ = dG * GELU'(Z)
__global__ void __launch_bounds__(256) k_gelu_bwd(long n, const float *__restrict__ Z, const float *__restrict__ dG,
                                                  float *__restrict__ dZ) {
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) dZ[e] = dG[e] * gelu_grad(Z[e]);
}

// Column sums of X (rows x cols, row-major) into partial[blk][col]; second stage sums blocks.
// A workgroup covers 64 columns with 4 row groups (256 threads) and reduces
// the groups in LDS; grid (column chunks, row blocks).
__global__ void __launch_bounds__(256) k_colsum_part(long rows, int cols, const float *__restrict__ X, float *__restrict__ part,
                                                     long rows_per_blk) {
  __shared__ float red[4][64];
  const int lc = threadIdx.x & 63, rg = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + lc;
  const long r0 = (long)blockIdx.y * rows_per_blk, r1 = min(rows, r0 + rows_per_blk);
  float s = 0.f;
  if (c < cols)
    for (long r = r0 + rg; r < r1; r += 4) s += X[r * cols + c];
  red[rg][lc] = s;
  __syncthreads();
  if (rg == 0 && c < cols) part[(long)blockIdx.y * cols + c] = red[0][lc] + red[1][lc] + red[2][lc] + red[3][lc];
}
__global__ void __launch_bounds__(256) k_colsum_final(int nblk, int cols, const float *__restrict__ part, float *out,
                                                      int accumulate) {
  __shared__ float red[4][64];
  const int lc = threadIdx.x & 63, rg = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + lc;
  float s = 0.f;
  if (c < cols)
    for (int b = rg; b < nblk; b += 4) s += part[(long)b * cols + c];
  red[rg][lc] = s;
  __syncthreads();
  if (rg == 0 && c < cols) {
    const float t = red[0][lc] + red[1][lc] + red[2][lc] + red[3][lc];
    out[c] = accumulate ? out[c] + t : t;
  }
}

// Features of the kept mode block (K^3 modes about eta = 0) of each cell, mode-major:
// X[m][b][0..3] = (T(f - M)/rho, T(M)/rho)[mode m] * fscale, from the complex
// transforms Fd (f - M) and Fm (M) of the cells (N^3 double2 each).
__global__ void __launch_bounds__(256) k_net_features(int N, int Kb, int ncell, const double2 *__restrict__ Fd,
                                                      const double2 *__restrict__ Fm, const double *__restrict__ mom,
                                                      double fscale, float *__restrict__ X) {
  const long n3 = (long)N * N * N, M = (long)Kb * Kb * Kb;
  const long total = M * ncell;
  const int lo = N / 2 - Kb / 2;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const long m = e / ncell;
    const int b = (int)(e - m * ncell);
    const int mz = (int)(m % Kb), my = (int)((m / Kb) % Kb), mx = (int)(m / ((long)Kb * Kb));
    const long idx = ((long)(lo + mx) * N + (lo + my)) * N + (lo + mz);
    const double s = fscale / mom[(long)b * SPBTE_MOM_STRIDE];
    const double2 d = Fd[(long)b * n3 + idx], q = Fm[(long)b * n3 + idx];
    float *x = X + e * 4;
    x[0] = (float)(d.x * s);
    x[1] = (float)(d.y * s);
    x[2] = (float)(q.x * s);
    x[3] = (float)(q.y * s);
  }
}

// Targets: T[m][b][0..1] = Qhat[mode m] / rho^2 * tscale.
__global__ void __launch_bounds__(256) k_net_targets(int N, int Kb, int ncell, const double2 *__restrict__ Qh,
                                                     const double *__restrict__ mom, double tscale, float *__restrict__ T) {
  const long n3 = (long)N * N * N, M = (long)Kb * Kb * Kb;
  const int lo = N / 2 - Kb / 2;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < M * ncell; e += (long)gridDim.x * 256) {
    const long m = e / ncell;
    const int b = (int)(e - m * ncell);
    const int mz = (int)(m % Kb), my = (int)((m / Kb) % Kb), mx = (int)(m / ((long)Kb * Kb));
    const long idx = ((long)(lo + mx) * N + (lo + my)) * N + (lo + mz);
    const double rho = mom[(long)b * SPBTE_MOM_STRIDE];
    const double2 q = Qh[(long)b * n3 + idx];
    T[e * 2] = (float)(q.x * tscale / (rho * rho));
    T[e * 2 + 1] = (float)(q.y * tscale / (rho * rho));
  }
}

// Network output -> zero-padded N^3 spectrum of each cell: Qh = Y * rho^2 / tscale on the block.
__global__ void __launch_bounds__(256) k_net_scatter(int N, int Kb, int ncell, const float *__restrict__ Y,
                                                     const double *__restrict__ mom, double tscale, double2 *__restrict__ Qh) {
  const long n3 = (long)N * N * N, M = (long)Kb * Kb * Kb;
  const int lo = N / 2 - Kb / 2;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n3 * ncell; e += (long)gridDim.x * 256) {
    const int b = (int)(e / n3);
    const long idx = e - (long)b * n3;
    const int iz = (int)(idx % N) - lo, iy = (int)((idx / N) % N) - lo, ix = (int)(idx / ((long)N * N)) - lo;
    double2 out = make_double2(0.0, 0.0);
    if (ix >= 0 && ix < Kb && iy >= 0 && iy < Kb && iz >= 0 && iz < Kb) {
      const long m = ((long)ix * Kb + iy) * Kb + iz;
      const double rho = mom[(long)b * SPBTE_MOM_STRIDE];
      const double s = rho * rho / tscale;
      out = make_double2((double)Y[(m * ncell + b) * 2] * s, (double)Y[(m * ncell + b) * 2 + 1] * s);
    }
    Qh[e] = out;
  }
}

// Lift: H[m][b][c] = sum_q X[m][b][q] P[q][c] + E[m][c] + pb[c]   (C channels, 4 features)
__global__ void __launch_bounds__(256) k_net_lift(long M, int ncell, int C, const float *__restrict__ X,
                                                  const float *__restrict__ P, const float *__restrict__ E,
                                                  const float *__restrict__ pb, float *__restrict__ H) {
  const long total = M * ncell * C;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int c = (int)(e % C);
    const long mb = e / C, m = mb / ncell;
    const float *x = X + mb * 4;
    H[e] = x[0] * P[c] + x[1] * P[C + c] + x[2] * P[2 * C + c] + x[3] * P[3 * C + c] + E[m * C + c] + pb[c];
  }
}

// Lift backward: dE[m][c] = sum_b dH[m][b][c]; dP[q][c] and dpb[c] need sums
// over (m, b): one workgroup per LIFT_MB consecutive modes writes dE for them
// and one row of partial sums [5C] (4 rows of dP, then dpb).
constexpr int LIFT_MB = 64;
__global__ void __launch_bounds__(256) k_net_lift_bwd(long M, int ncell, int C, const float *__restrict__ X,
                                                      const float *__restrict__ dH, float *__restrict__ dE,
                                                      float *__restrict__ part /* [ceil(M / LIFT_MB)][5C] */) {
  const long m0 = (long)blockIdx.x * LIFT_MB, m1 = min(M, m0 + LIFT_MB);
  for (int c = threadIdx.x; c < C; c += 256) {
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, sb = 0.f;
    for (long m = m0; m < m1; ++m) {
      float se = 0.f;
      for (int b = 0; b < ncell; ++b) {
        const float d = dH[(m * ncell + b) * C + c];
        const float *x = X + (m * ncell + b) * 4;
        se += d;
        s0 += d * x[0];
        s1 += d * x[1];
        s2 += d * x[2];
        s3 += d * x[3];
      }
      dE[m * C + c] = se;
      sb += se;
    }
    float *p = part + (long)blockIdx.x * 5 * C;
    p[c] = s0;
    p[C + c] = s1;
    p[2 * C + c] = s2;
    p[3 * C + c] = s3;
    p[4 * C + c] = sb;
  }
}

// Head: Y[r][0..1] = sum_c H[r][c] O[c][0..1] + ob   (r = (m, b))
__global__ void __launch_bounds__(256) k_net_head(long R, int C, const float *__restrict__ H, const float *__restrict__ O,
                                                  const float *__restrict__ ob, float *__restrict__ Y) {
  for (long r = (long)blockIdx.x * 256 + threadIdx.x; r < R; r += (long)gridDim.x * 256) {
    float y0 = ob[0], y1 = ob[1];
    const float *h = H + r * C;
    for (int c = 0; c < C; ++c) {
      y0 = fmaf(h[c], O[2 * c], y0);
      y1 = fmaf(h[c], O[2 * c + 1], y1);
    }
    Y[2 * r] = y0;
    Y[2 * r + 1] = y1;
  }
}

// Head backward: dH[r][c] = dY[r] . O[c]; per-block partials of dO[c][0..1] and dob.
__global__ void __launch_bounds__(256) k_net_head_bwd(long R, int C, const float *__restrict__ H,
                                                      const float *__restrict__ O, const float *__restrict__ dY,
                                                      float *__restrict__ dH, float *__restrict__ part /* [blk][2C+2] */,
                                                      long rows_per_blk) {
  const long r0 = (long)blockIdx.x * rows_per_blk, r1 = min(R, r0 + rows_per_blk);
  for (long r = r0 + threadIdx.x / C; r < r1; r += 256 / C) {
    const int c = threadIdx.x % C;
    dH[r * C + c] = dY[2 * r] * O[2 * c] + dY[2 * r + 1] * O[2 * c + 1];
  }
  for (int c = threadIdx.x; c < C + 1; c += 256) {
    float s0 = 0.f, s1 = 0.f;
    for (long r = r0; r < r1; ++r) {
      const float h = c < C ? H[r * C + c] : 1.0f;
      s0 += h * dY[2 * r];
      s1 += h * dY[2 * r + 1];
    }
    part[(long)blockIdx.x * (2 * C + 2) + 2 * c] = s0;
    part[(long)blockIdx.x * (2 * C + 2) + 2 * c + 1] = s1;
  }
}

// Per-cell relative squared error and the gradient of the mean over cells:
// L_b = sum_m |Y - T|^2 / (sum_m |T|^2 + eps),  dY = 2 (Y - T) / (den_b * ncell).
__global__ void __launch_bounds__(256) k_net_loss(long M, int ncell, const float *__restrict__ Y, const float *__restrict__ T,
                                                  float *__restrict__ dY, float *__restrict__ loss_b, float eps) {
  __shared__ float red[2][4];
  const int b = blockIdx.x;
  float num = 0.f, den = 0.f;
  for (long m = threadIdx.x; m < M; m += 256) {
    const long r = m * ncell + b;
    const float d0 = Y[2 * r] - T[2 * r], d1 = Y[2 * r + 1] - T[2 * r + 1];
    num += d0 * d0 + d1 * d1;
    den += T[2 * r] * T[2 * r] + T[2 * r + 1] * T[2 * r + 1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    num += __shfl_down(num, off, 64);
    den += __shfl_down(den, off, 64);
  }
  if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = num; red[1][threadIdx.x >> 6] = den; }
  __syncthreads();
  num = red[0][0] + red[0][1] + red[0][2] + red[0][3];
  den = red[1][0] + red[1][1] + red[1][2] + red[1][3] + eps;
  if (threadIdx.x == 0) loss_b[b] = num / den;
  if (dY) {
    const float s = 2.0f / (den * (float)ncell);
    for (long m = threadIdx.x; m < M; m += 256) {
      const long r = m * ncell + b;
      dY[2 * r] = s * (Y[2 * r] - T[2 * r]);
      dY[2 * r + 1] = s * (Y[2 * r + 1] - T[2 * r + 1]);
    }
  }
}

__global__ void __launch_bounds__(256) k_adam(long n, float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m1, float *__restrict__ m2, float lr, float b1,
                                              float b2, float eps, float c1, float c2, float wd) {
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const float gr = g[e];
    const float a = b1 * m1[e] + (1.f - b1) * gr, v = b2 * m2[e] + (1.f - b2) * gr * gr;
    m1[e] = a;
    m2[e] = v;
    p[e] -= lr * ((a / c1) / (sqrtf(v / c2) + eps) + wd * p[e]);
  }
}

__global__ void __launch_bounds__(256) k_axpy_f(long n, float a, const float *__restrict__ x, float *__restrict__ y) {
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) y[e] += a * x[e];
}

inline unsigned grid_for(long n) {
  long b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

}  // namespace

// ---- launchers ---------------------------------------------------------------
extern "C" int spbte_net_gemm(const float *A, long sa0, long sa1, const float *B, long sb0, long sb1, float *C, long sc0,
                              long sc1, int M, int N, int K, const float *add, const float *bias, float *G, float *work,
                              size_t work_floats, hipStream_t s) {
  if (M <= 0 || N <= 0) return 0;
  if (!((sa1 == 1) || (sa0 == 1)) || !((sb0 == 1) || (sb1 == 1))) return (int)hipErrorInvalidValue;
  GemmArgs g = {A, B, C, add, bias, G, sa0, sa1, sb0, sb1, sc0, sc1, M, N, K, K, 0, 0};
  const long lda = sa1 == 1 ? sa0 : sa1, ldb = sb0 == 1 ? sb1 : sb0;
  g.vecA = ((uintptr_t)A % 16 == 0) && (lda % 4 == 0);
  g.vecB = ((uintptr_t)B % 16 == 0) && (ldb % 4 == 0);
  const unsigned gx = (N + TN - 1) / TN, gy = (M + TM - 1) / TM;
  // split K when the output tile grid alone cannot fill the chip
  int nsplit = 1;
  const long tiles = (long)gx * gy;
  if (tiles < 512 && K > 4 * TK && work) {
    nsplit = (int)((1024 + tiles - 1) / tiles);
    const int maxs = K / (2 * TK);
    if (nsplit > maxs) nsplit = maxs;
    while (nsplit > 1 && (size_t)nsplit * M * N > work_floats) nsplit--;
  }
  if (nsplit <= 1) {
    g.kslice = ((K + TK - 1) / TK) * TK;
    hipLaunchKernelGGL(k_gemm, dim3(gx, gy, 1), dim3(256), 0, s, g, 0);
    return (int)hipGetLastError();
  }
  const int kt = (K + TK - 1) / TK;
  g.kslice = ((kt + nsplit - 1) / nsplit) * TK;
  nsplit = (K + g.kslice - 1) / g.kslice;
  GemmArgs gp = g;
  gp.C = work;
  hipLaunchKernelGGL(k_gemm, dim3(gx, gy, nsplit), dim3(256), 0, s, gp, 1);
  hipLaunchKernelGGL(k_splitk_reduce, dim3(grid_for((long)M * N)), dim3(256), 0, s, g, (const float *)work, nsplit);
  return (int)hipGetLastError();
}

extern "C" int spbte_net_gelu_bwd(long n, const float *Z, const float *dG, float *dZ, hipStream_t s) {
  hipLaunchKernelGGL(k_gelu_bwd, dim3(grid_for(n)), dim3(256), 0, s, n, Z, dG, dZ);
  return (int)hipGetLastError();
}

extern "C" int spbte_net_colsum(long rows, int cols, const float *X, float *out, int accumulate, float *work,
                                size_t work_floats, hipStream_t s) {
  long nblk = 1024;
  if ((size_t)nblk * cols > work_floats) nblk = (long)(work_floats / cols);
  if (nblk < 1) return (int)hipErrorInvalidValue;
  const long rpb = (rows + nblk - 1) / nblk;
  nblk = (rows + rpb - 1) / rpb;
  hipLaunchKernelGGL(k_colsum_part, dim3((cols + 63) / 64, (unsigned)nblk), dim3(256), 0, s, rows, cols, X, work, rpb);
  hipLaunchKernelGGL(k_colsum_final, dim3((cols + 63) / 64), dim3(256), 0, s, (int)nblk, cols, (const float *)work, out,
                     accumulate);
  return (int)hipGetLastError();
}

extern "C" int spbte_net_features(int N, int Kb, int ncell, const void *Fd, const void *Fm, const double *mom,
                                  double fscale, float *X, hipStream_t s) {
  hipLaunchKernelGGL(k_net_features, dim3(grid_for((long)Kb * Kb * Kb * ncell)), dim3(256), 0, s, N, Kb, ncell,
                     (const double2 *)Fd, (const double2 *)Fm, mom, fscale, X);
  return (int)hipGetLastError();
}

extern "C" int spbte_net_targets(int N, int Kb, int ncell, const void *Qh, const double *mom, double tscale, float *T,
                                 hipStream_t s) {
  hipLaunchKernelGGL(k_net_targets, dim3(grid_for((long)Kb * Kb * Kb * ncell)), dim3(256), 0, s, N, Kb, ncell,
                     (const double2 *)Qh, mom, tscale, T);
  return (int)hipGetLastError();
}

extern "C" int spbte_net_scatter(int N, int Kb, int ncell, const float *Y, const double *mom, double tscale, void *Qh,
                                 hipStream_t s) {
  hipLaunchKernelGGL(k_net_scatter, dim3(grid_for((long)N * N * N * ncell)), dim3(256), 0, s, N, Kb, ncell, Y, mom,
                     tscale, (double2 *)Qh);
  return (int)hipGetLastError();
}

extern "C" int spbte_net_lift(long M, int ncell, int C, const float *X, const float *P, const float *E, const float *pb,
                              float *H, hipStream_t s) {
  hipLaunchKernelGGL(k_net_lift, dim3(grid_for(M * ncell * C)), dim3(256), 0, s, M, ncell, C, X, P, E, pb, H);
  return (int)hipGetLastError();
}

/* dE written; dP (4 x C) and dpb (C) from per-block partials (work >= (nblk + 1) * 5C floats). */
extern "C" int spbte_net_lift_bwd(long M, int ncell, int C, const float *X, const float *dH, float *dE, float *dP,
                                  float *dpb, float *work, size_t work_floats, hipStream_t s) {
  const long nblk = (M + LIFT_MB - 1) / LIFT_MB;
  if ((size_t)(nblk + 1) * 5 * C > work_floats) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_net_lift_bwd, dim3((unsigned)nblk), dim3(256), 0, s, M, ncell, C, X, dH, dE, work);
  // column sums of the [nblk][5C] partials: first 4C are dP, last C dpb
  float *fin = work + (size_t)nblk * 5 * C;
  hipLaunchKernelGGL(k_colsum_final, dim3((5 * C + 63) / 64), dim3(256), 0, s, (int)nblk, 5 * C, (const float *)work,
                     fin, 0);
  hipError_t e = hipMemcpyAsync(dP, fin, sizeof(float) * 4 * C, hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return (int)e;
  e = hipMemcpyAsync(dpb, fin + 4 * C, sizeof(float) * C, hipMemcpyDeviceToDevice, s);
  return (int)e;
}

extern "C" int spbte_net_head(long R, int C, const float *H, const float *O, const float *ob, float *Y, hipStream_t s) {
  hipLaunchKernelGGL(k_net_head, dim3(grid_for(R)), dim3(256), 0, s, R, C, H, O, ob, Y);
  return (int)hipGetLastError();
}

/* dH written; dO (C x 2) and dob (2) from block partials (work >= 1024 * (2C + 2) + 2C + 2 floats). */
extern "C" int spbte_net_head_bwd(long R, int C, const float *H, const float *O, const float *dY, float *dH, float *dO,
                                  float *dob, float *work, size_t work_floats, hipStream_t s) {
  if (C > 256 || 256 % C) return (int)hipErrorInvalidValue;
  long nblk = 1024;
  const size_t need = (size_t)nblk * (2 * C + 2) + 2 * C + 2;
  if (need > work_floats) return (int)hipErrorInvalidValue;
  const long rpb = (R + nblk - 1) / nblk;
  nblk = (R + rpb - 1) / rpb;
  hipLaunchKernelGGL(k_net_head_bwd, dim3((unsigned)nblk), dim3(256), 0, s, R, C, H, O, dY, dH, work, rpb);
  float *fin = work + (size_t)1024 * (2 * C + 2);
  hipLaunchKernelGGL(k_colsum_final, dim3((2 * C + 2 + 63) / 64), dim3(256), 0, s, (int)nblk, 2 * C + 2,
                     (const float *)work, fin, 0);
  hipError_t e = hipMemcpyAsync(dO, fin, sizeof(float) * 2 * C, hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return (int)e;
  e = hipMemcpyAsync(dob, fin + 2 * C, sizeof(float) * 2, hipMemcpyDeviceToDevice, s);
  return (int)e;
}

extern "C" int spbte_net_loss(long M, int ncell, const float *Y, const float *T, float *dY, float *loss_b, float eps,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_net_loss, dim3(ncell), dim3(256), 0, s, M, ncell, Y, T, dY, loss_b, eps);
  return (int)hipGetLastError();
}

extern "C" int spbte_net_adam(long n, float *p, const float *g, float *m1, float *m2, float lr, float b1, float b2,
                              float eps, int step, float wd, hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(256), 0, s, n, p, g, m1, m2, lr, b1, b2, eps, c1, c2, wd);
  return (int)hipGetLastError();
}

extern "C" int spbte_net_axpy(long n, float a, const float *x, float *y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy_f, dim3(grid_for(n)), dim3(256), 0, s, n, a, x, y);
  return (int)hipGetLastError();
}

// Layout changes between the mode-major batch arrays ([m][b][w]) and the
// sample-major dataset ([s][m][w]).  to_samples: dst[(idx[b] M + m) w + q] =
// src[(m B + b) w + q]; otherwise the gather dst[(m B + b) w + q] = src[(idx[b] M + m) w + q].
__global__ void __launch_bounds__(256) k_relayout(long M, int B, int w, const float *__restrict__ src,
                                                  float *__restrict__ dst, int to_samples, const int *__restrict__ idx) {
  const long total = M * B * w;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int q = (int)(e % w);
    const long mb = e / w, m = mb / B;
    const int b = (int)(mb - m * B);
    const long se = ((long)idx[b] * M + m) * w + q;
    if (to_samples) dst[se] = src[e];
    else dst[e] = src[se];
  }
}

extern "C" int spbte_net_relayout(long M, int nsamp, int width, const float *src, float *dst, int to_samples,
                                  const int *idx, hipStream_t s) {
  const long n = M * nsamp * width;
  long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  hipLaunchKernelGGL(k_relayout, dim3((unsigned)b), dim3(256), 0, s, M, nsamp, width, src, dst, to_samples, idx);
  return (int)hipGetLastError();
}

// Per-cell relative L2 difference |a_b - b_b| / |b_b| of ncell blocks of n doubles.
__global__ void __launch_bounds__(256) k_cell_relerr(long n, const double *__restrict__ a, const double *__restrict__ b,
                                                     double *__restrict__ out) {
  __shared__ double red[2][4];
  const long c = blockIdx.x;
  double num = 0.0, den = 0.0;
  for (long e = threadIdx.x; e < n; e += 256) {
    const double x = a[c * n + e], y = b[c * n + e];
    num += (x - y) * (x - y);
    den += y * y;
  }
  for (int off = 32; off > 0; off >>= 1) {
    num += __shfl_down(num, off, 64);
    den += __shfl_down(den, off, 64);
  }
  if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = num; red[1][threadIdx.x >> 6] = den; }
  __syncthreads();
  if (threadIdx.x == 0) {
    num = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    den = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    out[c] = den > 0.0 ? sqrt(num / den) : (num > 0.0 ? INFINITY : 0.0);
  }
}

extern "C" int spbte_net_cell_relerr(long n, int ncell, const double *a, const double *b, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_cell_relerr, dim3(ncell), dim3(256), 0, s, n, a, b, out);
  return (int)hipGetLastError();
}

// Per-cell L2 norm of ncell blocks of n doubles.
__global__ void __launch_bounds__(256) k_cell_norm(long n, const double *__restrict__ a, double *__restrict__ out) {
  __shared__ double red[4];
  const long c = blockIdx.x;
  double s = 0.0;
  for (long e = threadIdx.x; e < n; e += 256) s += a[c * n + e] * a[c * n + e];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[c] = sqrt(red[0] + red[1] + red[2] + red[3]);
}

extern "C" int spbte_net_cell_norm(long n, int ncell, const double *a, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_cell_norm, dim3(ncell), dim3(256), 0, s, n, a, out);
  return (int)hipGetLastError();
}

// Cells by index: scatter == 0 copies src cell idx[b] to dst cell b (gather),
// scatter != 0 copies src cell b to dst cell idx[b]; cells of n doubles.
__global__ void __launch_bounds__(256) k_move_cells(long n, const int *__restrict__ idx, const double *__restrict__ src,
                                                    double *__restrict__ dst, int scatter) {
  const long b = blockIdx.y;
  const long from = scatter ? b : idx[b], to = scatter ? idx[b] : b;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) dst[to * n + e] = src[from * n + e];
}

extern "C" int spbte_net_move_cells(long n, int ncell, const int *idx, const double *src, double *dst, int scatter,
                                    hipStream_t s) {
  if (ncell < 1) return (int)hipSuccess;
  const long blocks = (n + 255) / 256 < 64 ? (n + 255) / 256 : 64;
  hipLaunchKernelGGL(k_move_cells, dim3((unsigned)blocks, (unsigned)ncell), dim3(256), 0, s, n, idx, src, dst, scatter);
  return (int)hipGetLastError();
}
