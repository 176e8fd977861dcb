// k_scratch.hip -- the fixed device workspace of spbte_compute_q_noalloc.
//
// One evaluation of the operator works in device memory beyond its inputs and
// outputs: the moments of the Maxwellian split, the transform scratch, the
// transformed operands, Qhat and the partial sums of the split weighted sum.
// spbte_compute_q allocates these per operator; the allocation-free entry point
// takes them from this file-scope array instead, so that it calls no
// allocator.  Every GPU the library runs on holds its own copy of the array
// (SPBTE_NOALLOC_SCRATCH_BYTES) until the process ends.
//
// In host code `scratch` names a host-side stand-in, not a device address: its
// device address is resolved once per device with hipGetSymbolAddress (with
// that device current) and only that address is handed to kernels.
#include <hip/hip_runtime.h>
#include <atomic>
#include "spbte.h"
#include "spbte_internal.h"

namespace {

static __device__ __align__(256) unsigned char scratch[SPBTE_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::atomic<void *> g_scratch_dev[kMaxDevices];

}  // namespace

// Device address of the current device's workspace (resolved on the first call
// per device; later calls return the stored address).  Returns hipError_t.
extern "C" int spbte_noalloc_scratch(void **ptr, size_t *bytes) {
  *ptr = nullptr;
  *bytes = 0;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
  void *p = g_scratch_dev[dev].load(std::memory_order_acquire);
  if (!p) {
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (!p) return (int)hipErrorInvalidSymbol;
    g_scratch_dev[dev].store(p, std::memory_order_release);
  }
  *ptr = p;
  *bytes = SPBTE_NOALLOC_SCRATCH_BYTES;
  return (int)hipSuccess;
}
