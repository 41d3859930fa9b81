// batch.hip — many GROUP BY plans in one persistent launch, ahead of time:
// the loop of batch_device.h with the plan-agnostic strip, the record-init
// kernel, and their launchers.  jit.cpp compiles the same loop with strips
// generated for a batch's plans.
#include "batch_device.h"
#include "strip_generic.h"

namespace tdbg {

// A scan item through the plan-agnostic strip (columns held in registers
// when the plan reads at most kFastCols).
struct GenericScan {
  static __device__ __forceinline__ void run(const BatchQuery& B, const ScanParams& P, uint8_t* lds,
                                             const BatchItem& it) {
    const uint64_t part_lo = (uint64_t)it.part * P.lds_cap;
    if (P.n_cols <= (uint32_t)kFastCols) {
      if (P.lds_mode == LDS_DIRECT) scan_range<LDS_DIRECT, GenericStrip<true>>(P, lds, it.a0, it.a1, part_lo);
      else if (P.lds_mode == LDS_HASH) scan_range<LDS_HASH, GenericStrip<true>>(P, lds, it.a0, it.a1, 0);
      else scan_range<LDS_NONE, GenericStrip<true>>(P, lds, it.a0, it.a1, 0);
    } else {
      if (P.lds_mode == LDS_DIRECT) scan_range<LDS_DIRECT, GenericStrip<false>>(P, lds, it.a0, it.a1, part_lo);
      else if (P.lds_mode == LDS_HASH) scan_range<LDS_HASH, GenericStrip<false>>(P, lds, it.a0, it.a1, 0);
      else scan_range<LDS_NONE, GenericStrip<false>>(P, lds, it.a0, it.a1, 0);
    }
  }
};

// Two workgroups per CU (4 wavefronts per SIMD): at most 128 VGPRs.
__global__ void __launch_bounds__(kBatchThreads, 4) batch_kernel(const BatchQuery* __restrict__ Q,
                                                                 const BatchItem* __restrict__ items,
                                                                 const BatchArgs* __restrict__ Ap) {
  batch_loop<GenericScan>(Q, items, Ap);
}

// Record template per block of kBatchChunk records (plans whose leaf pass
// writes every record are skipped).
__global__ void __launch_bounds__(256) batch_init_kernel(const BatchQuery* __restrict__ Q,
                                                         const BatchBlock* __restrict__ blocks) {
  const BatchBlock B = blocks[blockIdx.x];
  const BatchQuery& q = Q[B.q];
  if (!q.init_records) return;
  const uint32_t words = q.sp.g_rec / 8;
  const uint64_t n_rec = q.sp.g_cap - B.r0 < kBatchChunk ? q.sp.g_cap - B.r0 : kBatchChunk;
  uint64_t* dst = reinterpret_cast<uint64_t*>(q.sp.g_state + B.r0 * q.sp.g_rec);
  for (uint64_t w = threadIdx.x; w < n_rec * words; w += blockDim.x) dst[w] = q.tmpl[w % words];
}

hipError_t launch_batch_init(const BatchQuery* q, const BatchBlock* blocks, uint32_t n_blocks, hipStream_t s) {
  if (n_blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_init_kernel, dim3(n_blocks), dim3(256), 0, s, q, blocks);
  return hipGetLastError();
}

hipError_t launch_batch(const BatchQuery* Q, const BatchItem* items, const BatchArgs* args, uint32_t n_items,
                        uint32_t grid, uint32_t lds_bytes, hipStream_t s) {
  static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(batch_kernel),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
  if (attr != hipSuccess) return attr;
  if (grid == 0 || lds_bytes > 160 * 1024 - 1024) return hipErrorInvalidConfiguration;
  if (n_items == 0 || !args) return hipSuccess;
  hipLaunchKernelGGL(batch_kernel, dim3(grid), dim3(kBatchThreads), lds_bytes, s, Q, items, args);
  return hipGetLastError();
}

}  // namespace tdbg
