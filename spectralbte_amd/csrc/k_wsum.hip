// k_wsum.hip -- the O(N^6) weighted convolution sum of the spectral collision
// operator (reference: src/collisions.c:125-165):
//
//   Qhat_b[r] = sum_c W[r][c] * ( G0_b[c] F0_b[s(r,c)]  (+ G1_b[c] F1_b[s(r,c)]) )
//   s_d(r,c)  = (r_d + N/2 - c_d) mod N   for d = x, y, z
//
// for a batch of cells b that share one table W (N^3 x N^3, row = output mode r).
// The second product is the fused maxPreserve form (three reference passes in one).
//
// Design (MI355X, wave64, 160 KiB LDS/CU, HBM ~8 TB/s):
//  * W is the only large operand (8.6 GB FP64 / 4.3 GB FP32 at N=32).  It is
//    streamed from HBM straight into registers, each element exactly once per
//    workgroup, with 16-byte loads: the 8 (N=32) lanes of a z-row chunk read one
//    full 128-byte line W[r][cx,cy,0:N].  Each loaded W value feeds TB cells.
//  * Work split: a workgroup owns LINES z-lines of output modes (rx, ry..ry+LINES-1,
//    all rz) for TB cells and loops over every summed mode c.  A lane owns TR
//    consecutive rz of one line and a chunk of 4 consecutive cz; its partial sums
//    over its cz chunk are combined with the other chunks once, at the end.
//  * For a fixed (rx, cx) every output line needs the same shifted x-plane
//    F[sx][:][:] and the same G[cx][:][:]: both planes (per cell, per product)
//    are staged in LDS once per cx.  The F lines are stored extended by TR+2 so
//    a lane's shifted window F[sy][base .. base+TR+2] never wraps; base is fixed
//    per lane for the whole kernel, so the (r,c) -> s index map costs nothing in
//    the inner loop.  Each window element is reused by up to 4 (r,c) pairs.
//  * blockIdx -> (r-tile, cell group) is XCD-aware: consecutive workgroups that
//    land on one XCD (bid mod 8) share the r-tile, so the W lines they stream
//    are shared through that XCD's L2 when several cell groups run together.
//  * Precision forms (FORM): 0 = FP64 W, FP64 arithmetic (the reference method);
//    1 = FP32 W streamed (half the HBM bytes), widened to FP64 in registers, FP64
//    arithmetic; 2 = FP16x2 W: W = s_r (hi + lo), two FP16 halves per entry with a
//    power-of-two row scale s_r, widened to FP64; 3 = FP32 W and FP32 arithmetic
//    (F, G, products in FP32), partial sums over 4 cy folded into a compensated
//    (TwoSum) FP32 pair.  On CDNA4 an FP64 FMA issues at the same vector rate as an
//    unpacked FP32 FMA, so forms 1-2 cost the same VALU as form 3 and are accurate
//    to the rounding of W alone; form 3 halves LDS traffic and registers instead.
//    The final cross-lane combination and the output are FP64 in every form.
#include <hip/hip_runtime.h>
#include <type_traits>
#include "spbte_internal.h"

namespace {

typedef _Float16 h2v __attribute__((ext_vector_type(2)));

struct KArgs {
  const void *W;
  const float *rowscale;  // FP16x2 only
  const void *F0, *G0, *F1, *G1;
  double2 *Qh;
  int ncell;
};

template <typename T> struct cplx;
template <> struct cplx<double> { using t = double2; };
template <> struct cplx<float> { using t = float2; };

__device__ __forceinline__ void two_sum(float &hi, float &lo, float x) {
  const float s = hi + x;
  const float bb = s - hi;
  const float e = (hi - (s - bb)) + (x - bb);
  hi = s;
  lo += e;
}

// Load 4 consecutive weights of one row (storage WF) as CT.
template <int WF, typename CT>
__device__ __forceinline__ void load_w4(const void *p, CT (&w)[4]) {
  if constexpr (WF == 0) {
    const double2 *q = (const double2 *)p;
    const double2 a = q[0], b = q[1];
    w[0] = (CT)a.x; w[1] = (CT)a.y; w[2] = (CT)b.x; w[3] = (CT)b.y;
  } else if constexpr (WF == 1) {
    const float4 a = *(const float4 *)p;
    w[0] = (CT)a.x; w[1] = (CT)a.y; w[2] = (CT)a.z; w[3] = (CT)a.w;
  } else {
    const uint4 a = *(const uint4 *)p;
    const unsigned u[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const h2v hl = __builtin_bit_cast(h2v, u[j]);  // .x = hi, .y = lo
      w[j] = (CT)hl.x + (CT)hl.y;
    }
  }
}

// FORM -> (W storage, arithmetic): see the header comment.
template <int FORM> struct FormT;
template <> struct FormT<0> { static constexpr int WF = 0; using CT = double; };
template <> struct FormT<1> { static constexpr int WF = 1; using CT = double; };
template <> struct FormT<2> { static constexpr int WF = 2; using CT = double; };
template <> struct FormT<3> { static constexpr int WF = 1; using CT = float; };

template <int N, int TR, int LINES, int TB, int NP, int FORM>
struct Cfg {
  static constexpr int WF = FormT<FORM>::WF;
  using CT = typename FormT<FORM>::CT;
  using C2 = typename cplx<CT>::t;
  static constexpr bool A32 = std::is_same<CT, float>::value;
  static constexpr int FOLD = 4;              // cy per compensated fold (FP32 arithmetic)
  static constexpr int NCC = N / 4;           // cz chunks of 4
  static constexpr int NRG = N / TR;          // row groups per z-line
  static constexpr int PER_LINE = NCC * NRG;  // lanes per z-line
  static constexpr int NT = LINES * PER_LINE; // threads per workgroup
  static constexpr int LEXT = N + TR + 2;     // extended z-line length in LDS
  static constexpr int FPL = N * LEXT;        // one extended F x-plane
  static constexpr int GPL = N * N;           // one G x-plane
  static constexpr size_t WBYTES = (WF == 0) ? 8 : 4;
  static constexpr size_t LDS_FG = (size_t)NP * TB * (FPL + GPL) * sizeof(C2);
  static constexpr size_t LDS_RED = (size_t)NT * TR * sizeof(double2);
  static constexpr size_t LDS = LDS_FG > LDS_RED ? LDS_FG : LDS_RED;
  static_assert(N % 4 == 0 && N % TR == 0 && N % LINES == 0, "tile must divide N");
};

template <int N, int TR, int LINES, int TB, int NP, int FORM>
__global__ void __launch_bounds__((Cfg<N, TR, LINES, TB, NP, FORM>::NT))
k_wsum(KArgs a) {
  using C = Cfg<N, TR, LINES, TB, NP, FORM>;
  constexpr int WF = C::WF;
  using CT = typename C::CT;
  using C2 = typename C::C2;
  constexpr int NCC = C::NCC, NRG = C::NRG, PER_LINE = C::PER_LINE, NT = C::NT;
  constexpr int LEXT = C::LEXT, FPL = C::FPL, GPL = C::GPL;
  constexpr size_t n3 = (size_t)N * N * N;
  constexpr int NRT = N * (N / LINES);  // r-tiles

  __shared__ __attribute__((aligned(16))) char smem[C::LDS];
  C2 *sF = (C2 *)smem;                   // [NP][TB][N][LEXT]
  C2 *sG = sF + (size_t)NP * TB * FPL;    // [NP][TB][N][N]

  // ---- workgroup -> (r-tile, cell group), XCD-aware when the r-tiles split evenly
  const int ncg = (a.ncell + TB - 1) / TB;
  const int bid = blockIdx.x;
  int rt, cg;
  if ((NRT & 7) == 0) {
    const int xcd = bid & 7, q = bid >> 3;
    cg = q % ncg;
    rt = (q / ncg) * 8 + xcd;
  } else {
    cg = bid % ncg;
    rt = bid / ncg;
  }
  const int rx = rt / (N / LINES), ryb = rt - rx * (N / LINES);

  // ---- lane -> (line, row group, cz chunk)
  const int tid = threadIdx.x;
  const int cc = tid % NCC, rg = (tid / NCC) % NRG, ln = tid / PER_LINE;
  const int ry = ryb * LINES + ln, rz0 = rg * TR;
  const int base = (rz0 - 4 * cc - 3 + N / 2 + 2 * N) % N;  // window start in the shifted z-line
  const size_t row0 = ((size_t)(rx * N + ry)) * N + rz0;

  const char *Wb = (const char *)a.W + (row0 * n3 + 4 * cc) * C::WBYTES;
  const C2 *Fp[2] = {(const C2 *)a.F0, (const C2 *)a.F1};
  const C2 *Gp[2] = {(const C2 *)a.G0, (const C2 *)a.G1};

  // accumulators
  using A2 = C2;
  A2 acc[TB][TR];
  float2 chi[TB][TR], clo[TB][TR];  // compensated pairs (FP32 arithmetic only)
#pragma unroll
  for (int b = 0; b < TB; ++b)
#pragma unroll
    for (int t = 0; t < TR; ++t) {
      acc[b][t].x = 0; acc[b][t].y = 0;
      chi[b][t] = make_float2(0.f, 0.f);
      clo[b][t] = make_float2(0.f, 0.f);
    }

  CT wcur[TR][4], wnxt[TR][4];
#pragma unroll
  for (int t = 0; t < TR; ++t) load_w4<WF>(Wb + (size_t)t * n3 * C::WBYTES, wcur[t]);

  for (int cx = 0; cx < N; ++cx) {
    int sx = rx + N / 2 - cx;
    sx = sx < 0 ? sx + N : (sx > N - 1 ? sx - N : sx);
    __syncthreads();
    // ---- stage the shifted F x-plane (extended lines) and the G x-plane
    for (int e = tid; e < NP * TB * FPL; e += NT) {
      const int p = e / (TB * FPL), rem = e - p * (TB * FPL), b = rem / FPL, q = rem - b * FPL;
      const int sy = q / LEXT;
      int z = q - sy * LEXT;
      z = z >= N ? z - N : z;
      const int cell = cg * TB + b;
      C2 val;
      val.x = 0; val.y = 0;
      if (cell < a.ncell) val = Fp[p][(size_t)cell * n3 + ((size_t)sx * N + sy) * N + z];
      sF[e] = val;
    }
    for (int e = tid; e < NP * TB * GPL; e += NT) {
      const int p = e / (TB * GPL), rem = e - p * (TB * GPL), b = rem / GPL, q = rem - b * GPL;
      const int cell = cg * TB + b;
      C2 val;
      val.x = 0; val.y = 0;
      if (cell < a.ncell) val = Gp[p][(size_t)cell * n3 + (size_t)cx * N * N + q];
      sG[e] = val;
    }
    __syncthreads();

    for (int cy = 0; cy < N; ++cy) {
      // prefetch the next (cx, cy) weights
      const bool more = (cy + 1 < N) || (cx + 1 < N);
      const size_t coff_n = ((size_t)cx * N + cy + 1) * N;  // (cx, cy+1) == (cx+1, 0) when cy+1 == N
      if (more) {
#pragma unroll
        for (int t = 0; t < TR; ++t) load_w4<WF>(Wb + ((size_t)t * n3 + coff_n) * C::WBYTES, wnxt[t]);
      }
      int sy = ry + N / 2 - cy;
      sy = sy < 0 ? sy + N : (sy > N - 1 ? sy - N : sy);

#pragma unroll
      for (int b = 0; b < TB; ++b) {
        C2 g[NP][4], f[NP][TR + 3];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          const C2 *gl = sG + ((size_t)(p * TB + b) * N + cy) * N + 4 * cc;
          const C2 *fl = sF + ((size_t)(p * TB + b) * N + sy) * LEXT + base;
#pragma unroll
          for (int j = 0; j < 4; ++j) g[p][j] = gl[j];
#pragma unroll
          for (int u = 0; u < TR + 3; ++u) f[p][u] = fl[u];
        }
#pragma unroll
        for (int t = 0; t < TR; ++t)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int u = t - j + 3;
            CT pr = g[0][j].x * f[0][u].x, pi = g[0][j].x * f[0][u].y;
            pr = fma(-g[0][j].y, f[0][u].y, pr);
            pi = fma(g[0][j].y, f[0][u].x, pi);
            if constexpr (NP == 2) {
              pr = fma(g[1][j].x, f[1][u].x, pr);
              pi = fma(g[1][j].x, f[1][u].y, pi);
              pr = fma(-g[1][j].y, f[1][u].y, pr);
              pi = fma(g[1][j].y, f[1][u].x, pi);
            }
            acc[b][t].x = fma(wcur[t][j], pr, acc[b][t].x);
            acc[b][t].y = fma(wcur[t][j], pi, acc[b][t].y);
          }
      }
      if (more) {
#pragma unroll
        for (int t = 0; t < TR; ++t)
#pragma unroll
          for (int j = 0; j < 4; ++j) wcur[t][j] = wnxt[t][j];
      }
      if constexpr (C::A32) {
        if ((cy % C::FOLD) == C::FOLD - 1) {
#pragma unroll
          for (int b = 0; b < TB; ++b)
#pragma unroll
            for (int t = 0; t < TR; ++t) {
              two_sum(chi[b][t].x, clo[b][t].x, acc[b][t].x);
              two_sum(chi[b][t].y, clo[b][t].y, acc[b][t].y);
              acc[b][t].x = 0.f;
              acc[b][t].y = 0.f;
            }
        }
      }
    }
  }

  // ---- combine the NCC cz-chunk partials of each output row in FP64
  __syncthreads();
  double2 *red = (double2 *)smem;  // [NT][TR]
#pragma unroll
  for (int b = 0; b < TB; ++b) {
#pragma unroll
    for (int t = 0; t < TR; ++t) {
      double2 v;
      if constexpr (!C::A32) v = acc[b][t];
      else {
        v.x = (double)chi[b][t].x + (double)clo[b][t].x;
        v.y = (double)chi[b][t].y + (double)clo[b][t].y;
      }
      red[tid * TR + t] = v;
    }
    __syncthreads();
    const int cell = cg * TB + b;
    if (cc == 0 && cell < a.ncell) {
#pragma unroll
      for (int t = 0; t < TR; ++t) {
        double2 s = red[tid * TR + t];
        for (int q = 1; q < NCC; ++q) {
          const double2 o = red[(tid + q) * TR + t];
          s.x += o.x;
          s.y += o.y;
        }
        if constexpr (C::WF == 2) {
          const double sc = (double)a.rowscale[row0 + t];
          s.x *= sc;
          s.y *= sc;
        }
        a.Qh[(size_t)cell * n3 + row0 + t] = s;
      }
    }
    __syncthreads();
  }
}

// Generic form for any N (and the reference's one-row-at-a-time order): one
// thread per (output mode, cell), FP64 throughout.  Used for grids the tiled
// kernel is not instantiated for.
__global__ void __launch_bounds__(256)
k_wsum_generic(int N, const double *__restrict__ W, const double2 *__restrict__ F0, const double2 *__restrict__ G0,
               const double2 *__restrict__ F1, const double2 *__restrict__ G1, int nprod, double2 *__restrict__ Qh) {
  const size_t n3 = (size_t)N * N * N;
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t cell = blockIdx.y;
  if (r >= n3) return;
  const int rx = (int)(r / ((size_t)N * N)), ry = (int)((r / N) % N), rz = (int)(r % N);
  const int n2 = N / 2;
  const double *Wr = W + r * n3;
  const double2 *f0 = F0 + cell * n3, *g0 = G0 + cell * n3;
  const double2 *f1 = nprod == 2 ? F1 + cell * n3 : nullptr, *g1 = nprod == 2 ? G1 + cell * n3 : nullptr;
  double qr = 0.0, qi = 0.0;
  size_t c = 0;
  for (int cx = 0; cx < N; ++cx) {
    int x = rx + n2 - cx;
    x = x < 0 ? x + N : (x > N - 1 ? x - N : x);
    for (int cy = 0; cy < N; ++cy) {
      int y = ry + n2 - cy;
      y = y < 0 ? y + N : (y > N - 1 ? y - N : y);
      for (int cz = 0; cz < N; ++cz, ++c) {
        int z = rz + n2 - cz;
        z = z < 0 ? z + N : (z > N - 1 ? z - N : z);
        const size_t s = ((size_t)x * N + y) * N + z;
        const double w = Wr[c];
        double pr = g0[c].x * f0[s].x - g0[c].y * f0[s].y;
        double pi = g0[c].x * f0[s].y + g0[c].y * f0[s].x;
        if (nprod == 2) {
          pr += g1[c].x * f1[s].x - g1[c].y * f1[s].y;
          pi += g1[c].x * f1[s].y + g1[c].y * f1[s].x;
        }
        qr += w * pr;
        qi += w * pi;
      }
    }
  }
  Qh[cell * n3 + r] = make_double2(qr, qi);
}

template <int N, int TR, int LINES, int TB, int NP, int FORM>
int launch_t(const spbte_wsum_args *a, const float *rowscale, hipStream_t s) {
  using C = Cfg<N, TR, LINES, TB, NP, FORM>;
  KArgs k;
  k.W = a->W;
  k.rowscale = rowscale;
  k.F0 = a->F0; k.G0 = a->G0; k.F1 = a->F1; k.G1 = a->G1;
  k.Qh = (double2 *)a->Qh;
  k.ncell = a->ncell;
  const int ncg = (a->ncell + TB - 1) / TB;
  const int nrt = N * (N / LINES);
  hipLaunchKernelGGL((k_wsum<N, TR, LINES, TB, NP, FORM>), dim3(nrt * ncg), dim3(C::NT), 0, s, k);
  return (int)hipGetLastError();
}

// ---- instantiation table: (N, TR, LINES, TB) per form and product count
typedef int (*launch_fn)(const spbte_wsum_args *, const float *, hipStream_t);
struct Entry {
  int N, tr, lines, tb, np, form;
  launch_fn fn;
};

#define E(N, TR, L, TB, NP, F) {N, TR, L, TB, NP, F, &launch_t<N, TR, L, TB, NP, F>}
#define E_ALLFORMS(N, TR, L, TB, NP) E(N, TR, L, TB, NP, 0), E(N, TR, L, TB, NP, 1), E(N, TR, L, TB, NP, 2), \
                                     E(N, TR, L, TB, NP, 3)

const Entry kTable[] = {
    // N = 32: 64 lanes per z-line at TR=4, 32 at TR=8
    E_ALLFORMS(32, 4, 1, 1, 1), E_ALLFORMS(32, 4, 2, 1, 1), E_ALLFORMS(32, 4, 2, 2, 1), E_ALLFORMS(32, 4, 4, 2, 1),
    E_ALLFORMS(32, 4, 2, 4, 1), E_ALLFORMS(32, 4, 4, 4, 1), E_ALLFORMS(32, 8, 4, 2, 1), E_ALLFORMS(32, 8, 8, 2, 1),
    E_ALLFORMS(32, 8, 4, 4, 1),
    E_ALLFORMS(32, 4, 1, 1, 2), E_ALLFORMS(32, 4, 2, 1, 2), E_ALLFORMS(32, 4, 2, 2, 2), E_ALLFORMS(32, 4, 4, 2, 2),
    // N = 24: TR=6 -> 24 lanes per line
    E_ALLFORMS(24, 6, 8, 1, 1), E_ALLFORMS(24, 6, 8, 2, 1), E_ALLFORMS(24, 6, 8, 4, 1),
    E_ALLFORMS(24, 6, 8, 1, 2), E_ALLFORMS(24, 6, 8, 2, 2),
    // N = 16: TR=4 -> 16 lanes per line
    E_ALLFORMS(16, 4, 4, 1, 1), E_ALLFORMS(16, 4, 4, 2, 1), E_ALLFORMS(16, 4, 8, 2, 1), E_ALLFORMS(16, 4, 8, 4, 1),
    E_ALLFORMS(16, 4, 4, 1, 2), E_ALLFORMS(16, 4, 4, 2, 2), E_ALLFORMS(16, 4, 8, 2, 2),
    // N = 8: TR=2 -> 8 lanes per line (reduced forms; FP64 N=8 runs the generic kernel)
    E(8, 2, 8, 2, 1, 1), E(8, 2, 8, 2, 1, 2), E(8, 2, 8, 2, 1, 3), E(8, 2, 8, 2, 2, 1), E(8, 2, 8, 2, 2, 2),
    E(8, 2, 8, 2, 2, 3),
};
#undef E_ALLFORMS
#undef E

const Entry *find(int N, int form, int nprod, const spbte_tile *t) {
  for (const Entry &e : kTable)
    if (e.N == N && e.form == form && e.np == nprod && e.tr == t->tr && e.lines == t->lines && e.tb == t->tb) return &e;
  return nullptr;
}

}  // namespace

extern "C" int spbte_wsum_tile_supported(int N, int form, int nprod, const spbte_tile *tile) {
  return find(N, form, nprod, tile) != nullptr;
}

// Returns hipError_t; tile == NULL (or an uninstantiated tile) selects the
// generic FP64 kernel, which requires form 0.
extern "C" int spbte_launch_wsum_scaled(int N, int form, const spbte_tile *tile, const spbte_wsum_args *a,
                                        const float *rowscale, hipStream_t s) {
  const Entry *e = tile ? find(N, form, a->nprod, tile) : nullptr;
  if (e) return e->fn(a, rowscale, s);
  if (form != 0) return (int)hipErrorInvalidValue;
  const size_t n3 = (size_t)N * N * N;
  hipLaunchKernelGGL(k_wsum_generic, dim3((unsigned)((n3 + 255) / 256), a->ncell), dim3(256), 0, s, N, (const double *)a->W,
                     (const double2 *)a->F0, (const double2 *)a->G0, (const double2 *)a->F1, (const double2 *)a->G1, a->nprod,
                     (double2 *)a->Qh);
  return (int)hipGetLastError();
}

extern "C" int spbte_launch_wsum(int N, int form, const spbte_tile *tile, const spbte_wsum_args *a, hipStream_t s) {
  return spbte_launch_wsum_scaled(N, form, tile, a, nullptr, s);
}

// ---- weight-table conversions --------------------------------------------
namespace {
__global__ void k_w_to_f32(const double *__restrict__ W, float *__restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = (float)W[i];
}

// One workgroup per row: s_r = 2^(e - 14) where 2^e > max|W_r|, so |W/s_r| < 2^14
// sits well inside FP16 range; hi = fp16(W/s_r), lo = fp16(W/s_r - hi).
__global__ void __launch_bounds__(256) k_w_to_f16x2(const double *__restrict__ W, unsigned *__restrict__ out,
                                                    float *__restrict__ rowscale, size_t ncol) {
  __shared__ double red[4];
  const size_t r = blockIdx.x;
  const double *Wr = W + r * ncol;
  double m = 0.0;
  for (size_t c = threadIdx.x; c < ncol; c += 256) m = fmax(m, fabs(Wr[c]));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  int e = 0;
  if (m > 0.0) frexp(m, &e);  // m = f 2^e, f in [0.5, 1)
  const double sc = ldexp(1.0, e - 14);
  if (threadIdx.x == 0) rowscale[r] = (float)sc;
  const double inv = 1.0 / sc;
  for (size_t c = threadIdx.x; c < ncol; c += 256) {
    const double x = Wr[c] * inv;
    const _Float16 hi = (_Float16)x;
    const _Float16 lo = (_Float16)(x - (double)hi);
    const h2v pr = {hi, lo};
    out[r * ncol + c] = __builtin_bit_cast(unsigned, pr);
  }
}
}  // namespace

// form 1: dst = float[n_rows*n_cols]; form 2: dst = unsigned[n_rows*n_cols]
// (hi, lo) halves followed by float rowscale[n_rows].
extern "C" int spbte_launch_convert_weights(const double *W64, void *dst, int form, size_t n_rows, size_t n_cols,
                                            hipStream_t s) {
  const size_t n = n_rows * n_cols;
  if (form == 1) {
    hipLaunchKernelGGL(k_w_to_f32, dim3(4096), dim3(256), 0, s, W64, (float *)dst, n);
  } else if (form == 2) {
    unsigned *o = (unsigned *)dst;
    float *rs = (float *)(o + n);
    hipLaunchKernelGGL(k_w_to_f16x2, dim3((unsigned)n_rows), dim3(256), 0, s, W64, o, rs, n_cols);
  } else {
    return (int)hipErrorInvalidValue;
  }
  return (int)hipGetLastError();
}

namespace {
// FP32 operands: n complex values, then the n sums of their parts.
__global__ void k_c64_to_c32(const double2 *__restrict__ in, float2 *__restrict__ out, size_t n) {
  float *sums = (float *)(out + n);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float2 h = make_float2((float)in[i].x, (float)in[i].y);
    out[i] = h;
    sums[i] = h.x + h.y;
  }
}
}  // namespace

extern "C" int spbte_launch_c64_to_c32(const void *in, void *out, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_c64_to_c32, dim3(1024), dim3(256), 0, s, (const double2 *)in, (float2 *)out, n);
  return (int)hipGetLastError();
}

// ---- plain device copy (also the reference kernel for calibrating HBM byte counters)
namespace {
__global__ void __launch_bounds__(256) k_copy16(const uint4 *__restrict__ in, uint4 *__restrict__ out, size_t n16) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}
}  // namespace

extern "C" int spbte_launch_copy(const void *in, void *out, size_t bytes, hipStream_t s) {
  if (bytes % 16 || ((uintptr_t)in | (uintptr_t)out) % 16) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_copy16, dim3(8192), dim3(256), 0, s, (const uint4 *)in, (uint4 *)out, bytes / 16);
  return (int)hipGetLastError();
}

// ---- deterministic fill of the autotuner's inputs: values spread over
// [-1, 1) from a hash of the index.  The clock under a power-limited FP64 load
// depends on the operands (zeros switch less), so tiles are timed on data
// that loads the chip the way real transforms do.
namespace {
__device__ __forceinline__ unsigned fill_hash(size_t i) {
  unsigned h = (unsigned)i * 0x9E3779B1u ^ (unsigned)(i >> 32) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  return h ^ (h >> 15);
}
template <class T>
__global__ void __launch_bounds__(256) k_fill_pattern(T *__restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = (T)((double)(fill_hash(i) >> 8) * (2.0 / 16777216.0) - 1.0);
}
}  // namespace

extern "C" int spbte_launch_fill_pattern(void *p, size_t n, int single, hipStream_t s) {
  if (single) hipLaunchKernelGGL(k_fill_pattern<float>, dim3(4096), dim3(256), 0, s, (float *)p, n);
  else hipLaunchKernelGGL(k_fill_pattern<double>, dim3(4096), dim3(256), 0, s, (double *)p, n);
  return (int)hipGetLastError();
}
