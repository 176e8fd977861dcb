// k_wsum_stream.hip -- batched weighted convolution sum, streaming form.
//
// Same sum as k_wsum.hip (reference: src/collisions.c:125-165),
//   Qhat_b[r] = sum_c W[r][c] ( G0_b[c] F0_b[s(r,c)] (+ G1_b[c] F1_b[s(r,c)]) ),
//   s_d = (r_d + N/2 - c_d) mod N,
// organised for large batches, where the kernel is bound by VALU issue rather
// than HBM (each W element feeds TB cells, and cell groups that share an r-tile
// share W through L2):
//  * a lane owns TR consecutive output rows of one z-line and runs the whole
//    sum over c itself, so there is no cross-lane reduction;
//  * every lane of the workgroup handles the same summed mode c at the same
//    time, so G_b[c] is wave-uniform: it is read straight from global memory
//    with scalar loads (no LDS traffic, no vector registers; staging the G
//    x-planes in LDS held 32 KiB per workgroup and twice the LDS reads), except
//    with FP64 weights and one product, where the G lines passing through L2
//    crowd out the 8-byte W stream and the x-plane is staged in LDS instead;
//  * along cz the shifted operand F_b[sy][rz + N/2 - cz] of a lane's TR rows is a
//    window sliding by one element per cz step: it lives in registers and takes
//    one LDS load per step, shared by TR rows (the F x-planes of the cells sit in
//    LDS, one reload per cx);
//  * W is read from a copy permuted for this kernel: unit (wt, c4, t, lane) holds
//    W[row][4 c4 .. 4 c4 + 3] of row = 64 TR wt + TR lane + t, so every 16-byte
//    wave-wide load reads 1 KiB of contiguous HBM, in consumption order.
// Precision forms: 0 FP64 W, 1 FP32 W, 2 FP16x2 W with FP64 arithmetic; 3 FP32 W
// with FP32 F, G (8-byte LDS entries, half the LDS), FP32 products and FP32
// partial sums added into FP64 accumulators every 32 steps in 8-step loop
// bodies (N=32, N=16) and every 8 steps in longer ones (N=24).  Folding every
// 4 steps cost 6% more time on the 512-problem N=32 batch for a largest
// error against FP64 of 2.5e-7 instead of 2.9e-7.
#include <hip/hip_runtime.h>
#include <type_traits>
#include "spbte_internal.h"

namespace {

typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));  // one 16-byte unit of the permuted table

template <int N, int TR, int WAVES>
struct SGeom {
  static constexpr int RW = 64 * TR;                 // rows per wave tile
  static constexpr int RWG = WAVES * RW;             // rows per workgroup
  static constexpr int N2 = N * N;
  static constexpr int NC4 = N * N * N / 4;          // 4-column groups per row
  // x-planes of output rows a workgroup can touch: RWG / N2 when workgroups
  // tile whole planes, 1 when they tile within a plane, else the most planes a
  // run of RWG consecutive rows can straddle
  static constexpr int P = (RWG % N2 == 0) ? RWG / N2 : ((N2 % RWG == 0) ? 1 : (RWG - 1) / N2 + 2);
  // cz steps per loop body: a multiple of TR (the circular window returns to its
  // start) and of 8 (an even number of 4-column W chunks)
  static constexpr int BS = (TR % 8 == 0) ? TR : ((TR % 4 == 0) ? 2 * TR : (TR % 2 == 0 ? 4 * TR : 8 * TR));
  // a staged F line holds LW = N + BS values: entry m is F[(m - BS) mod N], so the
  // window refill of every step of a body reads at (base - j + BS) with j the
  // step's constant offset in the body, and never wraps
  static constexpr int LW = N + BS;
  // Line storage.  When every lane's window start is a multiple of G = 4 (TR
  // and N/2 multiples of 4) or of G = 3 (TR, N/2 and LW multiples of 3),
  // entry m is stored at (m mod G) * LW/G + m / G: the refills of the lanes of
  // one z-line (entries TR apart, a multiple of G) then fall on consecutive
  // 16-byte slots instead of one slot in G of the banks, and a refill offset
  // stays a compile-time constant per body step.  Otherwise entries are
  // stored in order (G = 1).
  static constexpr int G = ((TR % 4 == 0) && ((N / 2) % 4 == 0) && (LW % 4 == 0))
                               ? 4
                               : (((TR % 3 == 0) && ((N / 2) % 3 == 0) && (LW % 3 == 0)) ? 3 : 1);
  static constexpr bool TL = G > 1;
  static constexpr int LS = LW / G;
  __host__ __device__ static constexpr int pos(int m) { return TL ? (m % G) * LS + m / G : m; }
  static_assert(N % BS == 0, "cz body must divide N");
  static_assert(N % TR == 0, "TR must divide N");
  static_assert((N * N * N) % RWG == 0, "workgroups must tile the output rows");
};

// Weight j (0..3) of a 16-byte unit pair of the permuted table: a double, or
// the FP32 value itself in the FP32-product form (FORM 3).
template <int FORM>
__device__ __forceinline__ auto wget(const u32x4 &u, const u32x4 &u2, int j) {
  if constexpr (FORM == 0) {
    const u32x4 &q = j < 2 ? u : u2;
    return (j & 1) ? __builtin_bit_cast(double, make_uint2(q.z, q.w)) : __builtin_bit_cast(double, make_uint2(q.x, q.y));
  } else {
    const unsigned d = j == 0 ? u.x : (j == 1 ? u.y : (j == 2 ? u.z : u.w));
    if constexpr (FORM == 1) {
      return (double)__builtin_bit_cast(float, d);
    } else if constexpr (FORM == 3) {
      return __builtin_bit_cast(float, d);
    } else {
      const h2v hl = __builtin_bit_cast(h2v, d);
      return (double)hl.x + (double)hl.y;
    }
  }
}

// NTW: W read with non-temporal loads.  The launcher sets it when the batch is
// one cell group, so that every W element is read by one workgroup only: a
// streaming read then runs about 11% faster (an 8 GiB read measured 6.88 TB/s
// non-temporal against 6.18 TB/s), while batches keep W in L2 for the cell
// groups that share an r-tile.  A table that fits the 256 MB Infinity Cache
// with room to spare (N=16: 134 MB in FP64) keeps plain loads, so successive
// evaluations find it there (the N=16 single-cell sum measured 29.0 us against
// 30.8 us non-temporal).
// Four problems per lane with FP64 arithmetic: G by scalar loads and at most
// 256 VGPRs, so that two workgroups (80 KiB of LDS each) share a CU and every
// SIMD runs two waves (ComputeQ on 256 shock cells: 65.5 ms against 73.0 ms
// with G staged in LDS at one wave per SIMD).  (Capping the FP32-product tile
// of 4 rows x 4 waves at 128 VGPRs for a fourth wave per SIMD made it 1.8x
// slower: it runs at 141 VGPRs, three waves per SIMD.)
template <int TR, int WAVES, int TB, int FORM>
constexpr int kWavesPerEU = (TB >= 4 && FORM != 3) ? 2 : 1;

template <int N, int TR, int WAVES, int TB, int NP, int FORM, bool NTW, int CPW>
__global__ void __launch_bounds__(64 * WAVES * CPW) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<TR, WAVES, TB, FORM>)))
k_wsum_stream(const uint4 *__restrict__ Wp, const float *__restrict__ rowscale, const double2 *__restrict__ F0,
              const double2 *__restrict__ G0, const double2 *__restrict__ F1, const double2 *__restrict__ G1,
              double2 *__restrict__ Qh, int ncell, int csplit, double2 *__restrict__ Qp) {
  using Gm = SGeom<N, TR, WAVES>;
  constexpr int N2 = Gm::N2, P = Gm::P, RWG = Gm::RWG, NC4 = Gm::NC4;
  constexpr size_t n3 = (size_t)N * N * N;
  constexpr int NRT = (int)(n3 / RWG);                  // r-tiles
  constexpr int UPW = (FORM == 0) ? 2 : 1;              // 16-byte units per 4 weights
  // F, G elements: FP32 complex in the FP32-product form (products and
  // partial sums in FP32, folded into the FP64 accumulators), else FP64
  using C = std::conditional_t<FORM == 3, float2, double2>;
  constexpr int BS = Gm::BS;                            // cz steps per loop body

  constexpr int LW = Gm::LW;
  // CPW cell groups per workgroup, WAVES waves each, on the same rows: their
  // W loads hit the same lines at about the same time
  constexpr int NT = 64 * WAVES * CPW;                  // threads per workgroup
  constexpr int KB = CPW * TB;                          // cells per workgroup
  __shared__ __attribute__((aligned(16))) C sF[NP * KB * P * N * LW];  // [p][k][plane][sy][m]
  // FP32 products, two per pair (maxPreserve): each complex product
  // (a + ib)(c + id) with G = a + ib wave-uniform takes Gauss's three
  // multiplications, re = c (a + b) - b (c + d) and im = c (a + b) + a (d - c),
  // and the two products share one MUL + FMA for the c (a + b) terms: 8
  // instructions per pair with w instead of 10.  The staged F entry holds
  // (c, d - c) in sF and c + d in sP; a + b comes from the sums stored after
  // G's values.  One problem per lane only: with two, and with one product
  // (a saving of 1 in 6), the extra window registers cost more than the
  // instructions save (N=24, 512 problems, 3 rows x 3 waves x 2 problems:
  // 39.4 ms against 30.1 ms in the plain form).
  constexpr bool GS = FORM == 3 && NP == 2 && TB == 1;
  __shared__ float sP[GS ? NP * KB * P * N * LW : 1];
  constexpr bool SG = FORM != 0 || NP == 2 || TB >= 4;  // G by scalar loads (else staged in LDS)
  __shared__ __attribute__((aligned(16))) C sG[SG ? 1 : NP * KB * N2];  // [p][k][cy][cz]

  // ---- workgroup -> (r-tile, cx slice, cell group), XCD-aware: workgroups are
  // dispatched round-robin over the 8 XCDs, so XCD x (bid mod 8) takes the x-th
  // contiguous slice of the r-tile-major list of (r-tile, cx slice, cell group)
  // triples, and the workgroups it runs back to back share a piece of W in its
  // L2.  With csplit > 1 the summed x-planes cx are split into csplit slices
  // (small batches: more workgroups), each writing partial sums to Qp.
  const int ncg = (ncell + KB - 1) / KB;                // groups of KB cells
  const int total = NRT * csplit * ncg, per = (total + 7) >> 3;
  const int vid = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if ((blockIdx.x >> 3) >= per || vid >= total) return;  // the few padding workgroups of the last slice
  const int rt = vid / (csplit * ncg), rem = vid - rt * (csplit * ncg);
  const int cs = rem / ncg, cg = rem - cs * ncg;
  // slices: up to N take whole x-planes (N / csplit each); beyond N every
  // x-plane's summed y-lines are split as well (csplit / N pieces of it)
  const int ny = csplit > N ? csplit / N : 1;
  const int cx0 = csplit > N ? cs / ny : cs * (N / csplit);
  const int cx1 = csplit > N ? cx0 + 1 : cx0 + N / csplit;
  const int cy0 = csplit > N ? (cs - cx0 * ny) * (N / ny) : 0, cy1 = cy0 + N / ny;
  // wave and cell group are wave-uniform: say so, so that the cell's G lines
  // stay scalar loads
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) % WAVES);
  const int grp = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) / WAVES);
  const size_t wg_row0 = (size_t)rt * RWG;
  const int rx0 = (int)(wg_row0 / N2);                  // first x-plane of the workgroup
  const size_t row0 = wg_row0 + (size_t)TR * (threadIdx.x - grp * 64 * WAVES);
  const int rx = (int)(row0 / N2), ry = (int)((row0 / N) % N), rz0 = (int)(row0 % N);
  const int pl = rx - rx0;                               // this lane's plane in LDS
  const int k0 = (rz0 + N / 2) % N;                      // window start at cz = 0 (non-negative)
  const size_t wt = (size_t)rt * WAVES + wave;           // wave tile of the permuted table

  const C *Fp[2] = {(const C *)F0, (const C *)F1};
  const C *Gp[2] = {(const C *)G0, (const C *)G1};
  const float *Gsum[2] = {nullptr, nullptr};
  if constexpr (GS) {
    Gsum[0] = (const float *)(Gp[0] + (size_t)ncell * n3);
    if constexpr (NP == 2) Gsum[1] = (const float *)(Gp[1] + (size_t)ncell * n3);
  }
  int cells[TB];
#pragma unroll
  for (int b = 0; b < TB; ++b) {
    const int c = cg * KB + grp * TB + b;
    cells[b] = c < ncell ? c : ncell - 1;  // pad a ragged last group with a valid cell (not written)
  }

  double2 acc[TB][TR];
  float2 a32[FORM == 3 ? TB : 1][TR];  // FORM 3: FP32 partial sums between folds
#pragma unroll
  for (int b = 0; b < TB; ++b)
#pragma unroll
    for (int t = 0; t < TR; ++t) acc[b][t] = make_double2(0.0, 0.0);
#pragma unroll
  for (int b = 0; b < (FORM == 3 ? TB : 1); ++b)
#pragma unroll
    for (int t = 0; t < TR; ++t) a32[b][t] = make_float2(0.f, 0.f);

  // W chunk pointer: unit index ((wt*NC4 + c4)*TR + t)*UPW*64 + h*64 + lane
  const u32x4 *Wl = (const u32x4 *)Wp + (wt * NC4 * TR * UPW) * 64 + lane;
  auto wload = [&](int c4, u32x4 (&u)[TR][UPW]) {
#pragma unroll
    for (int t = 0; t < TR; ++t)
#pragma unroll
      for (int h = 0; h < UPW; ++h) {
        const u32x4 *q = Wl + (((size_t)c4 * TR + t) * UPW + h) * 64;
        if constexpr (NTW) u[t][h] = __builtin_nontemporal_load(q);
        else u[t][h] = *q;
      }
  };
  // Two static register sets alternate over the 4-column chunks of W (a body
  // holds an even number of chunks), each loaded four cz steps before use.
  u32x4 wA[TR][UPW], wB[TR][UPW];
  wload((cx0 * N + cy0) * (N / 4), wA);

  for (int cx = cx0; cx < cx1; ++cx) {
    __syncthreads();
    // ---- stage the shifted F x-planes of every cell, product and lane plane,
    // each z-line extended to LW values (entry m = F[(m - BS) mod N])
    for (int e = threadIdx.x; e < NP * KB * P * N * LW; e += NT) {
      const int ln = e / LW, m = e - ln * LW;                 // line (p, k, q, sy), entry
      const int p = ln / (KB * P * N), r1 = ln - p * (KB * P * N);
      const int k = r1 / (P * N), r2 = r1 - k * (P * N);
      const int q = r2 / N, y = r2 - q * N;
      int sx = (rx0 + q) % N + N / 2 - cx;
      sx = sx < 0 ? sx + N : (sx > N - 1 ? sx - N : sx);
      const int z = m < BS ? m - BS + N : m - BS;
      const int c = cg * KB + k < ncell ? cg * KB + k : ncell - 1;
      const C val = Fp[p][(size_t)c * n3 + (size_t)sx * N2 + (size_t)y * N + z];
      if constexpr (GS) {
        sF[ln * LW + Gm::pos(m)] = make_float2(val.x, val.y - val.x);
        sP[ln * LW + Gm::pos(m)] = val.x + val.y;
      } else {
        sF[ln * LW + Gm::pos(m)] = val;
      }
    }
    if constexpr (!SG)
      for (int e = threadIdx.x; e < NP * KB * N2; e += NT) {
        const int p = e / (KB * N2), r1 = e - p * (KB * N2), k = r1 / N2, yz = r1 - k * N2;
        const int c = cg * KB + k < ncell ? cg * KB + k : ncell - 1;
        sG[e] = Gp[p][(size_t)c * n3 + (size_t)cx * N2 + yz];
      }
    __syncthreads();

    for (int cy = cy0; cy < cy1; ++cy) {
      int sy = ry + N / 2 - cy;
      sy = sy < 0 ? sy + N : (sy > N - 1 ? sy - N : sy);
      const C *line[NP][TB];
      const float *lineP[NP][TB];
#pragma unroll
      for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int b = 0; b < TB; ++b) {
          const size_t off = ((size_t)((p * KB + grp * TB + b) * P + pl) * N + sy) * LW;
          line[p][b] = sF + off;
          lineP[p][b] = sP + (GS ? off : 0);
        }
      // circular window: at step cz, element t (row rz0 + t) is buf[(t - cz) mod TR]
      // and holds F_line[(k0 + t - cz) mod N] = line[BS + (k0 + t - cz) mod N]
      // (bufP: its c + d with FP32 products)
      C buf[NP][TB][TR];
      float bufP[NP][GS ? TB : 1][GS ? TR : 1];
#pragma unroll
      for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int b = 0; b < TB; ++b)
#pragma unroll
          for (int t = 0; t < TR; ++t) {
            buf[p][b][t] = line[p][b][Gm::pos(BS + ((k0 + t) % N))];
            if constexpr (GS) bufP[p][b][t] = lineP[p][b][Gm::pos(BS + ((k0 + t) % N))];
          }
      const int c4base = (cx * N + cy) * (N / 4);
      bool fold_body = true;
      // One 4-step chunk at body offset jb (a compile-time constant after
      // unrolling): wc holds its weights, wn receives the next chunk's.  `fl`
      // points at the refill entry of body step 0: line + BS + (k0 - 1 - cz_body) mod N.
      auto chunk = [&](u32x4 (&wc)[TR][UPW], u32x4 (&wn)[TR][UPW], const C *const (&fl)[NP][TB],
                       const float *const (&flP)[NP][TB], const C *const (&gz)[NP][TB],
                       const float *const (&gzs)[NP][TB], int jb, int c4n) __attribute__((always_inline)) {
        // The prefetch of the step after the last is clamped to the last chunk
        // (loaded and not used) rather than skipped: a conditional load would
        // make the compiler copy both register sets at every chunk.  The
        // empty asm marks this chunk's weights as defined here, so their
        // conversions (and the wait for their loads, issued four steps
        // earlier) are not moved up next to the loads; the scheduling barrier
        // keeps the next chunk's loads ahead of this chunk's arithmetic.
#pragma unroll
        for (int t = 0; t < TR; ++t)
#pragma unroll
          for (int h = 0; h < UPW; ++h) asm volatile("" : "+v"(wc[t][h]));
        wload(c4n < NC4 ? c4n : NC4 - 1, wn);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const int j = jb + jj;  // position in the body and in the window period
          // rows in descending order: row 0 reads the element loaded one step
          // earlier, last; the slot of row TR-1 is refilled right after its use
#pragma unroll
          for (int t = TR - 1; t >= 0; --t) {
            const int sl = ((t - j) % TR + TR) % TR;
            const auto w = wget<FORM>(wc[t][0], wc[t][UPW - 1], jj);
#pragma unroll
            for (int b = 0; b < TB; ++b) {
              if constexpr (GS) {
                // Gauss's form: 1 + 2 ops per product (the shared c (a + b)
                // terms of both products in one MUL + FMA), 2 for w
                const C g0 = gz[0][b][j], f0 = buf[0][b][sl];
                float k1 = f0.x * gzs[0][b][j];
                if constexpr (NP == 2) k1 = fmaf(buf[NP - 1][b][sl].x, gzs[NP - 1][b][j], k1);
                float qr = fmaf(-g0.y, bufP[0][b][sl], k1), qi = fmaf(g0.x, f0.y, k1);
                if constexpr (NP == 2) {
                  const C g1 = gz[NP - 1][b][j], f1 = buf[NP - 1][b][sl];
                  qr = fmaf(-g1.y, bufP[NP - 1][b][sl], qr);
                  qi = fmaf(g1.x, f1.y, qi);
                }
                a32[b][t].x = fmaf(w, qr, a32[b][t].x);
                a32[b][t].y = fmaf(w, qi, a32[b][t].y);
              } else {
                // complex products as explicit FMA chains: 4 ops per product, 2 for w
                const C g0 = gz[0][b][j], f0 = buf[0][b][sl];
                auto pr = g0.x * f0.x, pi = g0.x * f0.y;
                pr = fma(-g0.y, f0.y, pr);
                pi = fma(g0.y, f0.x, pi);
                if constexpr (NP == 2) {
                  const C g1 = gz[NP - 1][b][j], f1 = buf[NP - 1][b][sl];
                  pr = fma(g1.x, f1.x, pr);
                  pi = fma(g1.x, f1.y, pi);
                  pr = fma(-g1.y, f1.y, pr);
                  pi = fma(g1.y, f1.x, pi);
                }
                if constexpr (FORM == 3) {
                  a32[b][t].x = fmaf(w, pr, a32[b][t].x);
                  a32[b][t].y = fmaf(w, pi, a32[b][t].y);
                } else {
                  acc[b][t].x = fma(w, pr, acc[b][t].x);
                  acc[b][t].y = fma(w, pi, acc[b][t].y);
                }
              }
            }
            if (t == TR - 1) {
              const int slnew = ((-1 - j) % TR + TR) % TR;
#pragma unroll
              for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int b = 0; b < TB; ++b) {
                  // entry BS + base - j: with transposed lines BS + base = G BG + G - 1
                  // and fl points at slot BG of sub-row 0
                  const int ro = Gm::TL ? (Gm::G - 1 - j % Gm::G) * Gm::LS - j / Gm::G : -j;
                  buf[p][b][slnew] = fl[p][b][ro];
                  if constexpr (GS) bufP[p][b][slnew] = flP[p][b][ro];
                }
            }
          }
        }
        if constexpr (FORM == 3) {
          if ((jb + 4) % 8 != 0 || !fold_body) return;  // 8 (32 in 8-step bodies) steps per FP32 partial sum
          // fold the FP32 partial sums into FP64
#pragma unroll
          for (int b = 0; b < TB; ++b)
#pragma unroll
            for (int t = 0; t < TR; ++t) {
              acc[b][t].x += (double)a32[b][t].x;
              acc[b][t].y += (double)a32[b][t].y;
              a32[b][t] = make_float2(0.f, 0.f);
            }
        }
      };
      // The body is a whole number of window periods (TR) and of chunk pairs,
      // so window slots and the wA/wB roles are the same at every iteration.
      // (k0 - 1 - cz_body) mod N, cz_body = it * BS; never negative, so the
      // division by the line-storage group below is a plain shift
      unsigned base = (unsigned)(k0 - 1 + N);
#pragma unroll 1
      for (int it = 0; it < N / BS; ++it) {
        base = base >= (unsigned)N ? base - N : base;
        const C *fl[NP][TB];
        const float *flP[NP][TB];
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
          for (int b = 0; b < TB; ++b) {
            const unsigned o = Gm::TL ? ((unsigned)BS + base) / (unsigned)Gm::G : BS + base;
            fl[p][b] = line[p][b] + o;
            flP[p][b] = lineP[p][b] + (GS ? o : 0);
          }
        // G_b[cx][cy][it BS ...] of product p, cell b: uniform addresses (scalar
        // loads), or the staged x-plane in LDS
        const C *gz[NP][TB];
        const float *gzs[NP][TB];
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
          for (int b = 0; b < TB; ++b) {
            const size_t go = (size_t)cells[b] * n3 + (size_t)cx * N2 + (size_t)cy * N + it * BS;
            gz[p][b] = SG ? Gp[p] + go : sG + (size_t)(p * KB + grp * TB + b) * N2 + (size_t)cy * N + it * BS;
            gzs[p][b] = GS ? Gsum[p] + go : nullptr;
          }
        // 8-step bodies fold every fourth body (32-term partial sums), longer
        // bodies every 8 steps; the last body of a line always folds
        fold_body = (BS >= 16) || ((it & 3) == 3) || (it == N / BS - 1);
#pragma unroll
        for (int k = 0; k < BS / 4; k += 2) {
          const int c4 = c4base + it * (BS / 4) + k;
          chunk(wA, wB, fl, flP, gz, gzs, 4 * k, c4 + 1);
          chunk(wB, wA, fl, flP, gz, gzs, 4 * k + 4, c4 + 2);
        }
        base = base - BS + N;
      }
    }
  }

#pragma unroll
  for (int b = 0; b < TB; ++b) {
    const int cell = cg * KB + grp * TB + b;
    if (cell >= ncell) continue;
#pragma unroll
    for (int t = 0; t < TR; ++t) {
      double2 s = acc[b][t];
      if constexpr (FORM == 2) {
        const double sc = (double)rowscale[row0 + t];
        s.x *= sc;
        s.y *= sc;
      }
      if (csplit > 1) Qp[((size_t)cs * ncell + cell) * n3 + row0 + t] = s;
      else Qh[(size_t)cell * n3 + row0 + t] = s;
    }
  }
}

// Qh[b][r] = sum over the csplit slices of Qp[s][b][r].
__global__ void __launch_bounds__(256) k_slice_sum(size_t n, int csplit, const double2 *__restrict__ Qp,
                                                   double2 *__restrict__ Qh) {
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) {
    double2 v = Qp[e];
    for (int s = 1; s < csplit; ++s) {
      const double2 w = Qp[(size_t)s * n + e];
      v.x += w.x;
      v.y += w.y;
    }
    Qh[e] = v;
  }
}

template <int N, int TR, int WAVES, int TB, int NP, int FORM, int CPW>
int launch_s(const spbte_wsum_args *a, const void *Wp, const float *rowscale, int csplit, void *Qp, hipStream_t s) {
  using Gm = SGeom<N, TR, WAVES>;
  const int ncg = (a->ncell + TB * CPW - 1) / (TB * CPW);
  const int nrt = (int)((size_t)N * N * N / Gm::RWG);
  const int per = (nrt * csplit * ncg + 7) / 8;  // workgroups per XCD slice
  constexpr size_t n3 = (size_t)N * N * N, kCacheResident = (size_t)160 << 20;
  constexpr bool resident = n3 * n3 * (FORM == 0 ? 8 : 4) <= kCacheResident;
  bool nt = false;
  if constexpr (CPW == 1) nt = ncg == 1 && !resident;  // every W element read once: non-temporal loads
  if constexpr (CPW == 1)
    if (nt)
      hipLaunchKernelGGL((k_wsum_stream<N, TR, WAVES, TB, NP, FORM, true, 1>), dim3(8 * per), dim3(64 * WAVES), 0, s,
                         (const uint4 *)Wp, rowscale, (const double2 *)a->F0, (const double2 *)a->G0,
                         (const double2 *)a->F1, (const double2 *)a->G1, (double2 *)a->Qh, a->ncell, csplit,
                         (double2 *)Qp);
  if (!nt)
    hipLaunchKernelGGL((k_wsum_stream<N, TR, WAVES, TB, NP, FORM, false, CPW>), dim3(8 * per), dim3(64 * WAVES * CPW), 0,
                       s, (const uint4 *)Wp, rowscale, (const double2 *)a->F0, (const double2 *)a->G0,
                       (const double2 *)a->F1, (const double2 *)a->G1, (double2 *)a->Qh, a->ncell, csplit,
                       (double2 *)Qp);
  if (csplit > 1 && !a->keep_slices) {
    const size_t n = (size_t)a->ncell * N * N * N;
    hipLaunchKernelGGL(k_slice_sum, dim3((unsigned)((n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384)), dim3(256), 0,
                       s, n, csplit, (const double2 *)Qp, (double2 *)a->Qh);
  }
  return (int)hipGetLastError();
}

typedef int (*launch_fn)(const spbte_wsum_args *, const void *, const float *, int, void *, hipStream_t);
struct SEntry {
  int N, tr, waves, tb, np, form;
  launch_fn fn;
};

#define S(N, TR, WV, TB, NP, F) {N, TR, WV, TB, NP, F, &launch_s<N, TR, WV, TB, NP, F, 1>}
#define S3(N, TR, WV, TB, NP) S(N, TR, WV, TB, NP, 0), S(N, TR, WV, TB, NP, 1), S(N, TR, WV, TB, NP, 2)
#define S4(N, TR, WV, TB, NP) S3(N, TR, WV, TB, NP), S(N, TR, WV, TB, NP, 3)

const SEntry kSTable[] = {
    // N = 32: TR=8 (4 lanes per z-line, 2 waves per x-plane) and TR=4 (8 lanes per z-line)
    S3(32, 8, 2, 1, 1), S3(32, 8, 2, 2, 1), S4(32, 4, 4, 2, 1), S3(32, 4, 2, 2, 1), S4(32, 4, 4, 4, 1),
    S3(32, 4, 4, 1, 1), S3(32, 8, 2, 1, 2), S4(32, 4, 4, 2, 2), S4(32, 4, 4, 1, 2),
    // small batches: more, smaller r-tiles (half and quarter x-planes)
    S3(32, 4, 2, 1, 1), S3(32, 4, 2, 1, 2), S3(32, 4, 1, 1, 1), S3(32, 4, 1, 2, 1),
    // FP32 products, maxPreserve: four 4-wave groups of the same rows, one cell
    // each, per workgroup ("16 waves")
    {32, 4, 16, 1, 2, 3, &launch_s<32, 4, 4, 1, 2, 3, 4>},
    // N = 24 (rows per workgroup must divide 24^3): TR=3 x 3 waves = one x-plane,
    // TR=6 x 3 waves = 2 planes, TR=4 x 2 waves and TR=8 x 1 wave straddle planes
    S4(24, 3, 3, 1, 1), S4(24, 3, 3, 2, 1), S4(24, 3, 3, 1, 2), S3(24, 6, 3, 1, 1), S4(24, 6, 3, 1, 2),
    S4(24, 4, 2, 1, 1), S4(24, 4, 2, 2, 1), S4(24, 4, 2, 1, 2), S3(24, 8, 1, 1, 1), S4(24, 8, 1, 1, 2),
    S4(24, 4, 9, 1, 2), S(24, 3, 3, 2, 2, 3),
    // N = 16: TR=4 (1 wave = 1 x-plane); TR=2 and TR=1 (half and quarter planes)
    // give one cell more waves
    S3(16, 4, 1, 1, 1), S4(16, 4, 1, 2, 1), S3(16, 4, 1, 1, 2), S4(16, 4, 1, 2, 2),
    S3(16, 2, 1, 1, 1), S3(16, 2, 1, 1, 2), S3(16, 1, 1, 1, 1), S3(16, 1, 1, 1, 2),
    // one-row waves, two or four a workgroup (one cell: more waves, staging shared)
    S3(16, 1, 2, 1, 1), S3(16, 1, 2, 1, 2), S3(16, 1, 4, 1, 1), S3(16, 1, 4, 1, 2),
};
#undef S4
#undef S3
#undef S

const SEntry *sfind(int N, int form, int nprod, int tr, int waves, int tb) {
  for (const SEntry &e : kSTable)
    if (e.N == N && e.form == form && e.np == nprod && e.tr == tr && e.waves == waves && e.tb == tb) return &e;
  return nullptr;
}

// Natural FP64 table -> permuted copy for tile row count TR: unit (wt, c4, t, h, lane)
// holds W[64 TR wt + TR lane + t][4 c4 .. 4 c4 + 3] (h selects the first or second
// pair of doubles in the FP64 form).  One thread per 4 weights.
__global__ void __launch_bounds__(256) k_permute_w(const double *__restrict__ W, void *__restrict__ out,
                                                   const float *__restrict__ rowscale, int form, int TR, size_t n3) {
  const size_t nc4 = n3 / 4;
  const size_t total = n3 * nc4;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    // i enumerates the permuted units in storage order: ((wt*nc4 + c4)*TR + t)*64 + lane
    const size_t lane = i % 64, q1 = i / 64;
    const size_t t = q1 % TR, q2 = q1 / TR;
    const size_t c4 = q2 % nc4, wtile = q2 / nc4;
    const size_t row = (wtile * 64 + lane) * TR + t;
    const double *src = W + row * n3 + 4 * c4;
    if (form == 0) {
      double2 *o = (double2 *)out;
      o[((wtile * nc4 + c4) * TR + t) * 128 + lane] = make_double2(src[0], src[1]);
      o[((wtile * nc4 + c4) * TR + t) * 128 + 64 + lane] = make_double2(src[2], src[3]);
    } else if (form == 1) {
      ((float4 *)out)[i] = make_float4((float)src[0], (float)src[1], (float)src[2], (float)src[3]);
    } else {
      const double inv = 1.0 / (double)rowscale[row];
      unsigned u[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double x = src[j] * inv;
        const _Float16 hi = (_Float16)x;
        const _Float16 lo = (_Float16)(x - (double)hi);
        const h2v pr = {hi, lo};
        u[j] = __builtin_bit_cast(unsigned, pr);
      }
      ((uint4 *)out)[i] = make_uint4(u[0], u[1], u[2], u[3]);
    }
  }
}

}  // namespace

extern "C" int spbte_csplit_valid(int N, int csplit) {
  if (csplit < 1) return 0;
  if (csplit <= N) return N % csplit == 0;
  return csplit % N == 0 && N % (csplit / N) == 0;
}

extern "C" int spbte_stream_tile_supported(int N, int form, int nprod, int tr, int waves, int tb) {
  return sfind(N, form, nprod, tr, waves, tb) != nullptr;
}

// Launch with the permuted table Wp built for the same tr (spbte_launch_permute_weights).
// csplit: slices of the summed modes -- a divisor of N (whole x-planes) or N
// times a divisor of N (x-planes cut into y-line ranges); csplit > 1 needs Qp
// with room for csplit * ncell * N^3 double2.
extern "C" int spbte_launch_wsum_stream(int N, int form, int tr, int waves, int tb, const spbte_wsum_args *a, const void *Wp,
                                        const float *rowscale, int csplit, void *Qp, hipStream_t s) {
  const SEntry *e = sfind(N, form, a->nprod, tr, waves, tb);
  if (!e || !spbte_csplit_valid(N, csplit) || (csplit > 1 && !Qp)) return (int)hipErrorInvalidValue;
  return e->fn(a, Wp, rowscale, csplit, Qp, s);
}

// rowscale (FP16x2 only) must already hold the per-row scales.
extern "C" int spbte_launch_permute_weights(const double *W64, void *dst, const float *rowscale, int form, int TR, size_t n3,
                                            hipStream_t s) {
  if (TR < 1 || n3 % (64 * (size_t)TR) || n3 % 4) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_permute_w, dim3(8192), dim3(256), 0, s, W64, dst, rowscale, form, TR, n3);
  return (int)hipGetLastError();
}
