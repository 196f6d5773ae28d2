// tfce_batch_max without any allocation, for callers that cannot allocate
// device memory (code under graph capture, serving loops, harnesses that
// forbid hipMalloc).  The sweep's scratch (work counter, per-item records,
// per-workgroup union-find slots) lives in one file-scope device array of
// TFCE_NOALLOC_SCRATCH_BYTES that every GPU holds from module load until the
// process ends.  Its address on a device comes from hipGetSymbolAddress, once
// per device with that device current; host code uses only that address.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one GPU
// (two streams, two threads) race.  The work counter, the only part read
// before it is written, is cleared by a kernel at the start of each launch.
#include <hip/hip_runtime.h>

#include <mutex>

#include "tfcehip.h"
#include "capi_common.h"

static_assert(TFCE_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size must be a multiple of 256 bytes");

static __device__ __align__(256) int32_t scratch[TFCE_NOALLOC_SCRATCH_BYTES / sizeof(int32_t)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices];  // device address of `scratch` per device, null until looked up

// The array's address on the current device; the first call on a device looks
// it up and fills the per-device caches the launch path reads, so that later
// calls allocate nothing on the host either.
int scratch_address(void** out) {
  const int dev = tfcehip::current_device();
  if (dev < 0 || dev >= kMaxDevices)
    return tfcehip::set_error(TFCE_EINVAL, "device index beyond the no-allocation table");
  std::lock_guard<std::mutex> g(g_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return tfcehip::set_error(e != hipSuccess ? -(int)e : TFCE_EINVAL,
                                std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    (void)tfcehip::cu_count(dev);
    (void)tfcehip::device_memory(dev);
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return 0;
}

}  // namespace

extern "C" int tfce_batch_max_noalloc(tfce_adj_t* adj, const float* d_maps, int P, int sign_mask,
                                      const float* d_vdensity, float* d_max, hipStream_t stream) {
  void* p = nullptr;
  const int rc = scratch_address(&p);
  if (rc) return rc;
  const tfcehip::ExtScratch ext{p, (size_t)TFCE_NOALLOC_SCRATCH_BYTES};
  return tfcehip::batch_max_in(adj, d_maps, P, sign_mask, d_vdensity, d_max, &ext, stream);
}
