// Device scratch for the allocation-free batch entry point
// (pofd1d_loglike_batch_device_noalloc): a file-scope device array that the
// code object reserves on every GPU the library runs on, so a call needs
// no hipMalloc.  Its address differs per device and is resolved once per
// device by pofd_create (with that device current) through
// hipGetSymbolAddress; host code uses only that resolved address, never the
// array's name, which on the host is a stand-in and not a device address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"

static __device__ __align__(256) unsigned char scratch[pofd::kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

struct Slot {
  void* ptr = nullptr;
  bool resolved = false;
  hipError_t err = hipSuccess;
};

std::mutex g_mu;
Slot g_slot[kMaxDevices];

}  // namespace

namespace pofd {

int noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return POFD_E_ARG;
  std::lock_guard<std::mutex> lk(g_mu);
  Slot& s = g_slot[device];
  if (s.resolved) return s.err == hipSuccess ? POFD_OK : POFD_E_INTERNAL;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  s.resolved = true;
  s.err = e;
  s.ptr = (e == hipSuccess) ? p : nullptr;
  return e == hipSuccess ? POFD_OK : POFD_E_INTERNAL;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  const Slot& s = g_slot[device];
  return (s.resolved && s.err == hipSuccess) ? s.ptr : nullptr;
}

}  // namespace pofd
