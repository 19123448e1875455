// pobtaf without any device allocation: rocbta_dpobtaf_noalloc.
//
// It runs exactly the kernels of rocbta_dpobtaf, with the workspace taken
// from the file-scope device array `scratch` (ROCBTA_NOALLOC_SCRATCH_BYTES
// per GPU, held until the process ends) instead of a caller buffer, for
// callers that cannot allocate device memory: code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc.
//
// In host code the name `scratch` is only a host-side handle.  Its device
// address is fetched once per device, with that device current, by
// hipGetSymbolAddress, kept per device, and only that address is handed to
// the kernels.  Every call on one GPU uses the same array and the kernels
// clear the flags and counters they need at the start of each call; calls
// that overlap on one GPU (two streams, two host threads) race.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/rocbta.h"

namespace {

static __device__ __align__(256) unsigned char scratch[ROCBTA_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;

// Device address of `scratch` on the current device (nullptr on error).
void* scratch_ptr() {
  static std::mutex mu;
  static void* ptrs[NOALLOC_MAX_DEVICES] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "rocbta: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
              hipGetErrorString(e));
      return nullptr;
    }
    ptrs[dev] = p;
  }
  return ptrs[dev];
}

}  // namespace

extern "C" {

size_t rocbta_noalloc_scratch_bytes(void) { return ROCBTA_NOALLOC_SCRATCH_BYTES; }

int rocbta_dpobtaf_noalloc(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R,
                           double* T, int factorize_last_block, int* info, void* stream) {
  const size_t need = rocbta_workspace_bytes(ROCBTA_OP_POBTAF, n, b, a, 0);
  if (need == 0) return ROCBTA_ERR_ARG;  // invalid shape
  if (need > ROCBTA_NOALLOC_SCRATCH_BYTES) {
    fprintf(stderr,
            "rocbta_dpobtaf_noalloc: n=%lld b=%lld a=%lld needs %zu bytes of workspace, more "
            "than the %zu-byte static scratch; use rocbta_dpobtaf with a workspace\n",
            (long long)n, (long long)b, (long long)a, need, (size_t)ROCBTA_NOALLOC_SCRATCH_BYTES);
    return ROCBTA_ERR_WORKSPACE;
  }
  void* ws = scratch_ptr();
  if (!ws) return ROCBTA_ERR_HIP;
  return rocbta_dpobtaf(n, b, a, D, L, R, T, factorize_last_block, info, ws,
                        ROCBTA_NOALLOC_SCRATCH_BYTES, stream);
}

}  // extern "C"
