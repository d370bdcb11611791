// negf_rgf_noalloc: the RGF sweep of negf_rgf with its workspace in a
// file-scope device array instead of memory the library allocates, for
// callers that cannot allocate device memory (code under graph capture,
// serving loops, harnesses that forbid hipMalloc).
//
// The array `scratch` (NEGF_RGF_NOALLOC_SCRATCH_BYTES) exists once per GPU
// the library runs on and stays allocated until the process ends.  Host code
// never passes the array's name to a kernel (on the host that name is a
// stand-in, not a device address): on the first call on a device, with that
// device current, hipGetSymbolAddress gives the device address, which is kept
// per device and is what the sweep's kernels receive.
//
// Every call on a GPU uses the same array and leaves its contents behind.
// Nothing in it needs to start at zero: the sweep writes every workspace
// block before reading it, and clears its own flags and counters (the
// per-instance info accumulator and the inverse's pivot-used flags) with
// kernels at the start of each call.  Calls that overlap in time on one GPU
// (different streams, or host threads) race on the array and must be ordered
// by the caller.
#include "../../include/negf_hip.h"
#include "negf_internal.hpp"
#include "sweeps.hpp"

#include <mutex>

using namespace negf;

namespace {

constexpr size_t kScratchBytes = (size_t)NEGF_RGF_NOALLOC_SCRATCH_BYTES;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];   // device address of `scratch`, per device

// Device address of this device's copy of `scratch` (looked up once per
// device, with that device current).
int scratch_address(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return NEGF_HIP_ERROR;
  if (dev < 0 || dev >= kMaxDevices) return NEGF_HIP_ERROR;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return NEGF_HIP_ERROR;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return NEGF_OK;
}

}  // namespace

extern "C" {

int negf_rgf_noalloc(int32_t nblocks, const int32_t* block_sizes, const void* diag,
                     int64_t diag_bstride, const void* upper, int64_t upper_bstride,
                     const void* lower, int64_t lower_bstride, void* g_diag, void* g_upper,
                     void* g_lower, int32_t batch, int32_t* info, void* stream) {
  if (!block_sizes || nblocks <= 0 || batch < 0) return NEGF_BAD_ARG;
  for (int i = 0; i < nblocks; ++i)
    if (block_sizes[i] <= 0 || block_sizes[i] > 2048) return NEGF_BAD_ARG;
  if (batch == 0) return NEGF_OK;
  if (!diag || !g_diag || (nblocks > 1 && !upper)) return NEGF_BAD_ARG;
  Layout L(nblocks, block_sizes);
  // the sweep's workspace, then the info accumulator when the caller has none
  size_t sw = (sweep_work_bytes(L, batch, false) + 255) / 256 * 256;
  const size_t sw2 = (sweep_rgf_two_sided_bytes(L, batch) + 255) / 256 * 256;
  if (sw2 > sw) sw = sw2;
  const size_t need = sw + sizeof(int32_t) * (size_t)batch;
  if (need > kScratchBytes) return NEGF_TOO_LARGE;
  void* w = nullptr;
  const int rc = scratch_address(&w);
  if (rc) return rc;
  int32_t* inf = info ? info : (int32_t*)((char*)w + sw);
  BtdIn A{(const zc*)diag, diag_bstride, (const zc*)upper, upper_bstride, (const zc*)lower,
          lower_bstride};
  BtdOut G{(zc*)g_diag, (zc*)g_upper, (zc*)g_lower};
  return sweep_rgf(L, A, G, nullptr, nullptr, nullptr, nullptr, batch, inf, w,
                   (hipStream_t)stream, /*inplace_inverse=*/true, /*two_sided=*/true);
}

}  // extern "C"
