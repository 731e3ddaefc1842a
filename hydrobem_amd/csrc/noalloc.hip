// The device memory of the allocation-free entry point (hbg_lu_solve_noalloc): one file-scope
// array per GPU, `scratch`, that holds the staged system [A | B] and the LU's whole workspace (the
// per-panel inverses, pivot records, flags, the tournament's candidates, the substitution copy).
//
// In host code `scratch` names a host-side stand-in, not the device array: its device address is
// taken once per device by hipGetSymbolAddress (noalloc_scratch_setup, called at session creation
// with the device current) and only that address is ever handed to kernels or copies. Every call
// on a GPU shares that GPU's array and it keeps the last call's contents (the LU fills every flag
// and counter it reads before reading it); calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "hbg_internal.h"

namespace hbg {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;  // 32 MiB per GPU

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_addr[MAX_DEVICES] = {};
}  // namespace

size_t noalloc_scratch_bytes() { return NOALLOC_SCRATCH_BYTES; }

// The device address of `scratch` on `device` (which must be current), looked up on the first call
// for that device and kept.
void* noalloc_scratch_setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) throw Error{6, "device index beyond the scratch table"};
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error{6, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
    g_addr[device] = p;
  }
  return g_addr[device];
}

}  // namespace hbg
