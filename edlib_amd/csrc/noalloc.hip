// Allocation-free inner product for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops): the reduction
// partials live in a file-scope device array instead of context buffers.
//
// Every context creation resolves the array's device address once per GPU
// (hipGetSymbolAddress with that GPU current); the entry point then only
// launches the two reduction kernels of edh_vv on the context's stream.
// All calls on one GPU share the array: calls that overlap on one GPU
// (different streams or threads) race and must be serialised by the caller.
#include <mutex>

#include "edh_ctx.h"
#include "edh_kernels.h"

namespace edh {

// Reduction partials of one inner product: one per block of the grid-stride
// partial kernel, at most kMaxPartials (16 KiB).
constexpr int kNoAllocPartials = kMaxPartials;
static __device__ __align__(256) double scratch[kNoAllocPartials];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Resolves this file's scratch array on `device` (the current device);
// called once per device from edh_create.
void noalloc_register_device(int device) {
  EDH_CHECK(device >= 0 && device < kMaxDevices, EDH_ERR_LIMIT, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* ptr = nullptr;
  EDH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  EDH_CHECK(ptr != nullptr, EDH_ERR_HIP, "scratch array has no device address");
  g_scratch_ptr[device] = static_cast<double*>(ptr);
}

// Inner product into the device scalar *out with the partials in this
// GPU's scratch array; stream-ordered on s, nothing allocated or cleared.
void noalloc_dot(int device, const double* a, const double* b, int64_t n, double* out, hipStream_t s) {
  EDH_CHECK(n >= 0, EDH_ERR_ARG, "negative vector length");
  const int g = reduce_grid(n);
  EDH_CHECK(g <= kNoAllocPartials, EDH_ERR_LIMIT, "reduction grid exceeds the scratch array");
  double* part = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device >= 0 && device < kMaxDevices) part = g_scratch_ptr[device];
  }
  EDH_CHECK(part, EDH_ERR_STATE, "scratch array not resolved for this device");
  // every partial the reduction reads is written by the partial kernel
  // first, so nothing needs clearing between calls
  dev_dot(a, b, n, part, out, s);
}

}  // namespace edh
