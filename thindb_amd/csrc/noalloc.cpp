// noalloc.cpp — tdbg_execute_noalloc: one grouped query with no device or
// pinned allocation (no counterpart in the reference; for callers that
// cannot allocate: code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// Every device buffer the execution's kernels use - group records, tiles,
// flags, compaction and sort scratch, radix partitions, distinct sets, the
// finalized result - is carved from `scratch`, a file-scope device array of
// TDBG_NOALLOC_SCRATCH_BYTES that each GPU holds for the life of the
// process.  In host code the array's name is only a stand-in for its device
// address, so the address is resolved once per device with
// hipGetSymbolAddress (at context creation, with the device current) and
// only that pointer is ever handed to kernels or copies.
#include <hip/hip_runtime.h>

#include <mutex>

#include "runtime.h"

namespace {

constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[TDBG_NOALLOC_SCRATCH_BYTES];

struct DeviceScratch {
  std::once_flag once;
  hipError_t status = hipErrorNotInitialized;
  uint8_t* ptr = nullptr;  // this device's copy of `scratch`
  std::mutex run_mu;       // one no-alloc execution at a time per device
};
DeviceScratch g_dev[kMaxDevices];

}  // namespace

namespace tdbg {
namespace rt {

// Resolves the calling thread's current device's copy of the array
// (idempotent); called by tdbg_context_create with `device` current.
int32_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(TDBG_ERR_INVALID, "device %d", device);
  DeviceScratch& d = g_dev[device];
  std::call_once(d.once, [&] {
    void* p = nullptr;
    d.status = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (d.status == hipSuccess) d.ptr = static_cast<uint8_t*>(p);
    else (void)hipGetLastError();
  });
  if (d.status != hipSuccess) return hip_fail(d.status, "hipGetSymbolAddress(scratch)");
  return TDBG_OK;
}

}  // namespace rt
}  // namespace tdbg

extern "C" int32_t tdbg_execute_noalloc(tdbg_plan* plan, const tdbg_rg_ref* rgs, uint32_t n_rgs, void* out,
                                        uint64_t out_capacity, tdbg_column_view* cols, uint32_t n_cols,
                                        uint64_t* n_rows) {
  using namespace tdbg::rt;
  if (!plan) return fail(TDBG_ERR_INVALID, "null plan");
  const int dev = plan->ctx->device;
  HIP_TRY(hipSetDevice(dev), "hipSetDevice");
  int32_t rc = noalloc_setup(dev);  // resolved at context creation; this only reads it
  if (rc) return rc;
  DeviceScratch& d = g_dev[dev];
  std::lock_guard<std::mutex> lk(d.run_mu);
  NoAllocOut na{d.ptr, TDBG_NOALLOC_SCRATCH_BYTES, out, out_capacity, cols, n_cols, n_rows};
  return execute_plan_noalloc(plan, rgs, n_rgs, na);
}
