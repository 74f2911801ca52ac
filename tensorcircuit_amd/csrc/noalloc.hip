// Pauli-sum expectation that allocates nothing: the fp64 partial sums live
// in a static device array instead of a caller-provided workspace, for
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).
//
// The array is per GPU (each device holds its own copy at its own address)
// and shared by every call on that GPU: calls that overlap on one GPU (two
// host threads, or two streams without ordering between them) race on it.
// Every partial-sum slot a call reads is written by that call's kernels
// first, so nothing has to be cleared between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/tcm.h"

namespace tcm {
int set_error(int code, const char* fmt, ...);
int pauli_expect_check(const tcm_pauli* ham, int32_t dtype, const void* state, int64_t batch, double* out,
                       int64_t scratch_bytes);
int pauli_expect_prepared(const tcm_pauli* ham, int32_t dtype, const void* state, int64_t batch, double* out,
                          void* scratch, int64_t scratch_bytes, void* stream);
}  // namespace tcm

namespace {

// 8 MiB: fp64 (re, im) partials for batch x slots <= 524288 (32 complex64
// states of the 24-qubit TFIM sum, 12288 slots each, use 6 MiB)
constexpr int64_t kScratchBytes = 8ll << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// device address of `scratch` on the current device, looked up once per
// device (the host-side name is not a device address)
int scratch_of_current_device(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return tcm::set_error((int)e, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return tcm::set_error(TCM_E_ARG, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return tcm::set_error(e != hipSuccess ? (int)e : TCM_E_ARG, "hipGetSymbolAddress(scratch): %s",
                            hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return 0;
}

}  // namespace

extern "C" int tcm_expect_pauli_sum_noalloc(const tcm_pauli* ham, int32_t dtype, const void* state, int64_t batch,
                                            double* out, void* stream) {
  if (int rc = tcm::pauli_expect_check(ham, dtype, state, batch, out, kScratchBytes)) return rc;
  if (batch == 0) return 0;
  void* p = nullptr;
  if (int rc = scratch_of_current_device(&p)) return rc;
  return tcm::pauli_expect_prepared(ham, dtype, state, batch, out, p, kScratchBytes, stream);
}
