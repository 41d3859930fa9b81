// scan_agg.hip — the ahead-of-time generic scan/filter/aggregate kernel and
// the record-initialisation kernels.
//
// The generic kernel reads the plan (column kinds, predicate terms, key
// parts, aggregate kinds) from its arguments at run time, so it covers every
// plan the runtime accepts.  Plans run faster through the kernels jit.cpp
// specialises per plan signature; this kernel is the path when run-time
// compilation is unavailable or disabled (THINDB_GPU_JIT=0).
#include <algorithm>

#include "strip_generic.h"

namespace tdbg {

template <uint32_t LMODE, bool FAST>
__global__ void __launch_bounds__(1024) scan_agg_kernel(const ScanParams P) {
  extern __shared__ __align__(16) uint8_t lds[];
  scan_body<LMODE, GenericStrip<FAST>>(P, lds);
}

template <uint32_t LMODE>
static hipError_t allow_large_lds_for() {
  const int bytes = 160 * 1024;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(scan_agg_kernel<LMODE, true>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(scan_agg_kernel<LMODE, false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  return e;
}

static hipError_t allow_large_lds() {
  hipError_t e = allow_large_lds_for<LDS_DIRECT>();
  return e == hipSuccess ? allow_large_lds_for<LDS_HASH>() : e;
}

template <uint32_t LMODE>
static void launch_mode(const ScanParams& p, uint32_t grid, uint32_t block, size_t lds, hipStream_t s) {
  if (p.n_cols <= (uint32_t)kFastCols)
    hipLaunchKernelGGL((scan_agg_kernel<LMODE, true>), dim3(grid), dim3(block), lds, s, p);
  else
    hipLaunchKernelGGL((scan_agg_kernel<LMODE, false>), dim3(grid), dim3(block), lds, s, p);
}

hipError_t launch_scan_agg(const ScanParams& p, uint32_t grid, uint32_t block, hipStream_t s) {
  static const hipError_t attr = allow_large_lds();
  (void)attr;  // best effort: a launch over the limit reports its own error
  switch (p.lds_mode) {
    case LDS_DIRECT: launch_mode<LDS_DIRECT>(p, grid, block, p.lds_bytes, s); break;
    case LDS_HASH: launch_mode<LDS_HASH>(p, grid, block, p.lds_bytes, s); break;
    default: launch_mode<LDS_NONE>(p, grid, block, 0, s); break;
  }
  return hipGetLastError();
}

// Initialise n_rec records from one template record (a fill kernel rather
// than hipMemsetAsync so the sequence stays graph-capturable).
__global__ void fill_records_kernel(uint64_t* state, uint64_t n_words, uint32_t rec_words,
                                    const uint64_t* __restrict__ tmpl, uint32_t* flags, uint32_t n_flags) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_flags) flags[t] = 0;
  for (uint64_t i = t; i < n_words; i += (uint64_t)gridDim.x * blockDim.x) state[i] = tmpl[i % rec_words];
}

hipError_t launch_fill_records(uint8_t* state, uint64_t n_rec, uint32_t rec_bytes, const uint64_t* tmpl,
                               hipStream_t s, uint32_t* flags, uint32_t n_flags) {
  const uint64_t words = n_rec * (rec_bytes / 8);
  uint64_t blocks = (std::max<uint64_t>(words, n_flags) + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_records_kernel, dim3((uint32_t)blocks), dim3(256), 0, s,
                     reinterpret_cast<uint64_t*>(state), words, rec_bytes / 8, tmpl, flags, n_flags);
  return hipGetLastError();
}

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace tdbg
