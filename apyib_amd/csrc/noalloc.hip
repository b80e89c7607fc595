// apy_cisd_so_step for callers that cannot allocate device memory (code
// under HIP graph capture, serving loops, harnesses that forbid hipMalloc):
// the step's workspace is a file-scope array of this library, one copy per
// device, held until the process ends.  Host code never hands the array's
// name to a kernel: it uses the device address hipGetSymbolAddress returns
// for the current device, looked up once per device and kept.
//
// Every call on a device shares that device's array: calls that overlap on
// one GPU (different streams or host threads) race.  The step writes every
// workspace region it reads, so nothing has to start at zero.
#include "common.h"

#include <mutex>

namespace {

static __device__ __align__(256) unsigned char scratch[APY_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES];

int scratch_address(void** out) {
  int dev = -1;
  APY_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return APY_E_ARG;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    APY_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return APY_E_HIP;
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return APY_OK;
}

}  // namespace

extern "C" {

size_t apy_cisd_so_noalloc_bytes(const apy_cisd_so_args* a) { return apy_cisd_so_ws_exact(a); }

int apy_cisd_so_step_noalloc(const apy_cisd_so_args* a, void* stream) {
  if (!a || a->O < 2 || a->V < 2) return APY_E_ARG;
  if (apy_cisd_so_ws_exact(a) > (size_t)APY_NOALLOC_SCRATCH_BYTES) return APY_E_TOO_LARGE;
  void* p = nullptr;
  const int rc = scratch_address(&p);
  if (rc != APY_OK) return rc;
  return apy_cisd_so_step_ws(a, p, (size_t)APY_NOALLOC_SCRATCH_BYTES, stream, 1);
}

}  // extern "C"
