// k_issue.hip -- vector ALU issue ceiling of the device, measured.
//
// The weighted sum in its batched forms is bound by vector FMA issue, not by
// HBM.  The published FP32 vector peak counts packed FP32 instructions, which
// this library does not emit; this kernel measures what unpacked FMAs
// actually reach: every lane runs K independent FMA chains (one operand
// wave-uniform, as the transformed G values of the weighted sum are), enough
// waves per SIMD to hide the FMA latency, and the shader cycle counter is
// read against the constant 100 MHz clock to give the clock during the run.
#include <hip/hip_runtime.h>
#include "spbte.h"
#include "spbte_internal.h"

namespace {

constexpr int kChains = 16;
constexpr int kBlocks = 8192;
constexpr int kThreads = 256;

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_fma_chains(T *__restrict__ out, T a, const T *__restrict__ bsrc, int iters, unsigned long long *__restrict__ clk) {
  T x[kChains];
#pragma unroll
  for (int k = 0; k < kChains; ++k) x[k] = (T)(threadIdx.x * 1e-3) + (T)k;
  const T b = bsrc[blockIdx.x & 7];  // wave-uniform operand
  const unsigned long long c0 = clock64(), w0 = wall_clock64();
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < kChains; ++k) x[k] = fma(x[k], a, b);
  }
  const unsigned long long c1 = clock64(), w1 = wall_clock64();
  T s = 0;
#pragma unroll
  for (int k = 0; k < kChains; ++k) s += x[k];
  out[(size_t)blockIdx.x * kThreads + threadIdx.x] = s;  // keeps the chains live
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    clk[0] = c1 - c0;
    clk[1] = w1 - w0;
  }
}

template <typename T>
int measure(double *tflops, double *ghz) {
  T *out = nullptr, *b = nullptr;
  unsigned long long *clk = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&out, sizeof(T) * kBlocks * kThreads);
  if (e == hipSuccess) e = hipMalloc(&b, sizeof(T) * 8);
  if (e == hipSuccess) e = hipMalloc(&clk, 2 * sizeof(unsigned long long));
  T hb[8];
  for (int i = 0; i < 8; ++i) hb[i] = (T)(1e-3 * (i + 1));
  if (e == hipSuccess) e = hipMemcpy(b, hb, sizeof hb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  const int iters = 10000;
  float ms = 0.f;
  unsigned long long hc[2] = {0, 0};
  if (e == hipSuccess) {
    // warm-up (clock ramp, code load), then the timed launch
    hipLaunchKernelGGL(k_fma_chains<T>, dim3(kBlocks), dim3(kThreads), 0, 0, out, (T)0.999, b, iters / 10, clk);
    e = hipDeviceSynchronize();
  }
  if (e == hipSuccess) e = hipEventRecord(e0, 0);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_fma_chains<T>, dim3(kBlocks), dim3(kThreads), 0, 0, out, (T)0.999, b, iters, clk);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess) e = hipMemcpy(hc, clk, sizeof hc, hipMemcpyDeviceToHost);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (out) (void)hipFree(out);
  if (b) (void)hipFree(b);
  if (clk) (void)hipFree(clk);
  if (e != hipSuccess) return (int)e;
  const double flops = 2.0 * kChains * (double)iters * kBlocks * kThreads;
  *tflops = ms > 0.f ? flops / (ms * 1e-3) / 1e12 : 0.0;
  *ghz = hc[1] ? (double)hc[0] / (double)hc[1] * 0.1 : 0.0;  // wall clock counter at 100 MHz
  return (int)hipSuccess;
}

}  // namespace

extern "C" int spbte_valu_peak(int fp64, double *tflops, double *clock_ghz) {
  if (!tflops || !clock_ghz) return spbte_fail(SPBTE_ERR_ARG, "null output");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return spbte_fail(SPBTE_ERR_NODEV, "no GPU visible to HIP");
  const int e = fp64 ? measure<double>(tflops, clock_ghz) : measure<float>(tflops, clock_ghz);
  if (e != (int)hipSuccess) return spbte_fail(SPBTE_ERR_HIP, "issue-rate kernel: %s", hipGetErrorString((hipError_t)e));
  return SPBTE_OK;
}
