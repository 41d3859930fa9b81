// batch.hip — many GROUP BY plans in one persistent launch.
//
// A batch is a queue of work items (BatchItem) over the plans' tiles and
// radix leaves; every workgroup of the grid pulls the next item with one
// device atomic until the queue is empty, so tables of very different sizes
// share the GPU without a launch (or a drain) per plan.  Per item the
// workgroup runs the same routines as the single-plan kernels:
//   BI_SCAN        scan/filter/aggregate tiles [a0, a1) into an LDS table
//                  (direct: one key range `part`; hash) or HBM atomics
//   BI_RX_HIST     radix level-1 digit counts of the filtered rows of tiles
//                  [a0, a1); the plan's last histogram item turns every
//                  item's counts into its write offsets and the bucket ranges
//   BI_RX_SCATTER  radix scatter of tiles [a0, a1) at its item's offsets (no
//                  device atomics), after the offsets are published
//   BI_RX_AGG      radix leaves [a0, a1), after all of the plan's scatters,
//                  into the plan's dense record array
// The queue lists every histogram item first, then scans, then scatters, then
// leaves: an item only waits for items dequeued before it, by workgroups that
// are running and never wait on it, so every wait ends.  A wait that exceeds
// spin_limit polls raises ERR_BATCH_WAIT and abandons the item (the host then
// fails the batch) rather than hang.
#include "radix_device.h"
#include "strip_generic.h"

namespace tdbg {

// Polls *flag (device scope) until it reaches `target`; returns false after
// spin_limit polls.  Acquire: the caller then sees the producers' writes.
__device__ __forceinline__ bool batch_wait(const uint32_t* flag, uint32_t target, uint32_t spin_limit, uint32_t* err,
                                           uint32_t* s_ok) {
  if (threadIdx.x == 0) {
    uint32_t spins = 0;
    bool ok = true;
    while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(8);
      if (++spins >= spin_limit) {
        atomicOr(err, ERR_BATCH_WAIT);
        ok = false;
        break;
      }
    }
    *s_ok = ok ? 1u : 0u;
  }
  __syncthreads();
  __threadfence();
  return *s_ok != 0;
}

// Makes this workgroup's writes visible device-wide, then adds 1 to *counter;
// returns (to thread 0) the counter's previous value.
__device__ __forceinline__ uint32_t batch_signal(uint32_t* counter) {
  __threadfence();
  __syncthreads();
  uint32_t prev = 0;
  if (threadIdx.x == 0) prev = atomicAdd(counter, 1u);
  return prev;
}

__global__ void __launch_bounds__(kBatchThreads) batch_kernel(const BatchQuery* __restrict__ Q,
                                                              const BatchItem* __restrict__ items, uint32_t n_items,
                                                              uint32_t* head, uint32_t* err, uint32_t spin_limit) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ uint32_t s_item;
  __shared__ uint32_t s_ok;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(head, 1u);
    __syncthreads();
    const uint32_t idx = __builtin_amdgcn_readfirstlane(s_item);
    if (idx >= n_items) break;  // uniform: every workgroup leaves once the queue is drained
    const BatchItem it = items[idx];
    const BatchQuery& B = Q[it.q];
    const ScanParams& P = B.sp;
    switch (it.kind) {
      case BI_SCAN: {
        // columns held in registers when the plan reads at most kFastCols
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
        break;
      }
      case BI_RX_HIST: {
        const uint32_t nb = 1u << B.rx.bits1;
        rx_item_hist(P, B.rx, it.a0, it.a1, reinterpret_cast<uint32_t*>(lds), B.item_hist + (uint64_t)it.part * nb);
        const uint32_t prev = batch_signal(&B.sync[0]);
        if (threadIdx.x == 0) s_ok = prev + 1 == B.n_hist ? 1u : 0u;
        __syncthreads();
        if (s_ok) {  // uniform: the plan's last histogram item turns counts into offsets
          __threadfence();
          rx_item_offsets(B.rx, B.item_hist, B.n_hist, reinterpret_cast<uint64_t*>(lds));
          __threadfence();
          __syncthreads();
          if (threadIdx.x == 0) __hip_atomic_store(&B.sync[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        break;
      }
      case BI_RX_SCATTER: {
        if (batch_wait(&B.sync[1], 1u, spin_limit, err, &s_ok)) {
          const uint32_t nb = 1u << B.rx.bits1;
          rx_scatter_range(P, B.rx, it.a0, it.a1, lds, B.item_hist + (uint64_t)it.part * nb);
          batch_signal(&B.sync[2]);
        }
        break;
      }
      default: {  // BI_RX_AGG
        if (batch_wait(&B.sync[2], B.n_scatter, spin_limit, err, &s_ok))
          for (uint32_t leaf = it.a0; leaf < it.a1; leaf++) rx_agg_leaf<true>(P, B.rx, leaf, lds, nullptr);
        break;
      }
    }
    __syncthreads();  // s_item and the LDS are reused by the next item
  }
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

hipError_t launch_batch_scan(const BatchQuery* q, const BatchItem* items, uint32_t n_items, uint32_t* head,
                             uint32_t* err, uint32_t spin_limit, uint32_t grid, uint32_t lds_bytes, hipStream_t s) {
  static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(batch_kernel),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
  if (attr != hipSuccess) return attr;
  if (grid == 0 || lds_bytes > 160 * 1024 - 1024 || spin_limit == 0) return hipErrorInvalidConfiguration;
  if (n_items == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_kernel, dim3(grid), dim3(kBatchThreads), lds_bytes, s, q, items, n_items, head, err,
                     spin_limit);
  return hipGetLastError();
}

}  // namespace tdbg
