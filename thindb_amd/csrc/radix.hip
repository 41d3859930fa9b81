// radix.hip — high-cardinality GROUP BY by radix partitioning.
//
// Replaces the silo grid's route/stage and deferred radix parts
// (src/exec/silo_group_core.zig:5226-5235 routeHashRowBits, :9228-9400
// stageBucketIds/scatter, :10799-10830 deferPartIds, :10868-10960
// foldDeferredBucket): rows are routed by the high bits of their dense
// relative key (so a partition is a contiguous key range) in at most two
// scatter levels, then each leaf's keys fit one workgroup's LDS as a direct
// table, and each group is written to HBM exactly once.
//
//   rx_hist      level-1 digit histogram over all rows (bucket capacities)
//   rx_offsets   bucket starts from the histogram
//   rx_scatter1  filter + key + native-width payload; a tile of up to 4096
//                rows is staged in LDS grouped by digit and each digit's run is
//                written contiguously (one device atomic per (tile, digit))
//   rx_scatter2  one workgroup per level-1 bucket: digit histogram, leaf
//                ranges, and the same staged scatter with LDS cursors
//   rx_agg       persistent workgroups over leaves: LDS direct aggregation,
//                then occupied slots -> compacted group records
// Each pass reads and writes rows once, coalesced; no per-row HBM atomics.
#include "radix_device.h"

namespace tdbg {

__global__ void __launch_bounds__(kRxThreads) rx_hist_kernel(const ScanParams P, const RadixParams R,
                                                             unsigned long long* hist) {
  __shared__ uint32_t h[1u << kRxMaxBits1];
  uint32_t t0, t1;
  rx_tile_range(P, t0, t1);
  rx_hist_range(P, R, t0, t1, h, hist);
}

__global__ void __launch_bounds__(256) rx_offsets_kernel(const RadixParams R, const unsigned long long* hist) {
  __shared__ uint64_t tmp[256];
  rx_offsets(R, hist, tmp);
}

__global__ void __launch_bounds__(kRxThreads) rx_scatter1_kernel(const ScanParams P, const RadixParams R) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint32_t t0, t1;
  rx_tile_range(P, t0, t1);
  rx_scatter_range(P, R, t0, t1, lds);
}

__global__ void __launch_bounds__(kRxThreads) rx_scatter2_kernel(const RadixParams R) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint32_t* cnt = reinterpret_cast<uint32_t*>(lds + kRxCntOff);
  uint32_t* off = reinterpret_cast<uint32_t*>(lds + kRxOffOff);
  uint32_t* cur = reinterpret_cast<uint32_t*>(lds + kRxCurOff);
  uint32_t* tmp = reinterpret_cast<uint32_t*>(lds + kRxTmpOff);
  uint8_t* stage = lds + kRxStageOff;
  const uint32_t b = blockIdx.x;
  const uint32_t nb2 = 1u << R.bits2, m2 = nb2 - 1u, stride = R.row.stride;
  const uint64_t lo = R.bucket_start[b];
  const uint32_t n = (uint32_t)(R.bucket_fill[b] - lo);
  // leaf ranges of this bucket
  for (uint32_t d = threadIdx.x; d < nb2; d += blockDim.x) cnt[d] = 0;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t key = *reinterpret_cast<const uint32_t*>(R.rows1 + (lo + i) * stride);
    atomicAdd(&cnt[(key >> R.shift2) & m2], 1u);
  }
  __syncthreads();
  rx_prefix(cnt, off, nb2, tmp);
  for (uint32_t d = threadIdx.x; d < nb2; d += blockDim.x) {
    R.leaf_start[2 * ((uint64_t)b * nb2 + d)] = lo + off[d];
    R.leaf_start[2 * ((uint64_t)b * nb2 + d) + 1] = lo + off[d] + cnt[d];
    cur[d] = off[d];
  }
  __syncthreads();
  // staged scatter of the bucket's rows by level-2 digit
  for (uint32_t s0 = 0; s0 < n; s0 += R.tile_rows) {
    const uint32_t tn = n - s0 < R.tile_rows ? n - s0 : R.tile_rows;
    for (uint32_t d = threadIdx.x; d < nb2; d += blockDim.x) cnt[d] = 0;
    __syncthreads();
    uint32_t digit[kVec], rank[kVec];
#pragma unroll
    for (int j = 0; j < kVec; j++) {
      const uint32_t i = threadIdx.x + (uint32_t)j * blockDim.x;
      if (i < tn) {
        const uint32_t key = *reinterpret_cast<const uint32_t*>(R.rows1 + (lo + s0 + i) * stride);
        digit[j] = (key >> R.shift2) & m2;
        rank[j] = atomicAdd(&cnt[digit[j]], 1u);
      }
    }
    __syncthreads();
    rx_prefix(cnt, off, nb2, tmp);
#pragma unroll
    for (int j = 0; j < kVec; j++) {
      const uint32_t i = threadIdx.x + (uint32_t)j * blockDim.x;
      if (i < tn)
        rx_copy_row(stage + (size_t)(off[digit[j]] + rank[j]) * stride, R.rows1 + (lo + s0 + i) * stride, stride);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tn; i += blockDim.x) {
      const uint8_t* row = stage + (size_t)i * stride;
      const uint32_t d = (*reinterpret_cast<const uint32_t*>(row) >> R.shift2) & m2;
      rx_copy_row(R.rows2 + (lo + cur[d] + (i - off[d])) * stride, row, stride);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < nb2; d += blockDim.x) cur[d] += cnt[d];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(512) rx_agg_kernel(const ScanParams P, const RadixParams R) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ uint32_t tmp[512];
  __shared__ uint64_t out_base;
  for (uint32_t leaf = blockIdx.x; leaf < R.n_leaves; leaf += gridDim.x) {
    if (R.dense_out) rx_agg_leaf<true>(P, R, leaf, lds, tmp, &out_base);
    else rx_agg_leaf<false>(P, R, leaf, lds, tmp, &out_base);
  }
}

hipError_t launch_radix(const ScanParams& p, const RadixParams& r, uint32_t grid, unsigned long long* hist,
                        hipStream_t s, int phase) {
  static const hipError_t attr = [] {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(rx_scatter1_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(rx_scatter2_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e == hipSuccess)  // rx_agg also holds ~2 KiB of static LDS
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(rx_agg_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 4096);
    return e;
  }();
  if (attr != hipSuccess) return attr;
  if (r.bits1 > kRxMaxBits1 || r.bits2 > 10 || (r.row.stride & 7u) || r.tile_rows == 0 || r.tile_rows > kRxTile ||
      (r.tile_rows & (kVec - 1)) || rx_scatter_lds_bytes(r) > 160 * 1024 || grid == 0 ||
      p.lds_cap != (1u << r.leaf_bits) || r.n_leaves != (1u << (r.bits1 + r.bits2)) ||
      p.lds_bytes > 160 * 1024 - 4096)
    return hipErrorInvalidConfiguration;
  switch (phase) {
    case RX_HIST: hipLaunchKernelGGL(rx_hist_kernel, dim3(grid), dim3(kRxThreads), 0, s, p, r, hist); break;
    case RX_OFFSETS: hipLaunchKernelGGL(rx_offsets_kernel, dim3(1), dim3(256), 0, s, r, hist); break;
    case RX_SCATTER1:
      hipLaunchKernelGGL(rx_scatter1_kernel, dim3(grid), dim3(kRxThreads), rx_scatter_lds_bytes(r), s, p,
                         r);
      break;
    case RX_SCATTER2:
      hipLaunchKernelGGL(rx_scatter2_kernel, dim3(1u << r.bits1), dim3(kRxThreads),
                         rx_scatter_lds_bytes(r), s, r);
      break;
    case RX_AGG: hipLaunchKernelGGL(rx_agg_kernel, dim3(grid), dim3(512), p.lds_bytes, s, p, r); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace tdbg
