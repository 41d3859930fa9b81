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
//   rx_scatter2  level 2 over chunks of each level-1 bucket: per-chunk digit
//                histograms (rx_hist2), leaf ranges and chunk offsets per
//                bucket (rx_offsets2), then the same staged scatter with LDS
//                cursors from the chunk's offsets
//   rx_agg       persistent workgroups over leaves: LDS direct aggregation,
//                then occupied slots -> group records from the leaf's first
//                row index (or every key's record, dense output)
//   rx_gids      compacted leaves: positions of the groups, leaf by leaf
// Each pass reads and writes rows once, coalesced; no per-row HBM atomics.
#include "radix_device.h"

namespace tdbg {

template <bool HASHED>
__global__ void __launch_bounds__(kRxThreads) rx_hist_kernel(const ScanParams P, const RadixParams R,
                                                             unsigned long long* hist) {
  __shared__ uint32_t h[1u << kRxMaxBits1];
  uint32_t t0, t1;
  rx_tile_range(P, t0, t1);
  rx_hist_range<HASHED>(P, R, t0, t1, h, hist);
}

__global__ void __launch_bounds__(256) rx_offsets_kernel(const RadixParams R, const unsigned long long* hist) {
  __shared__ uint64_t tmp[256];
  rx_offsets(R, hist, tmp);
}

template <bool HASHED>
__global__ void __launch_bounds__(kRxThreads) rx_scatter1_kernel(const ScanParams P, const RadixParams R) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint32_t t0, t1;
  rx_tile_range(P, t0, t1);
  rx_scatter_range<HASHED>(P, R, t0, t1, lds);
}

// Level 2 runs on chunks2 workgroups per level-1 bucket; chunk c of bucket b
// covers the bucket's rows [n * c / chunks2, n * (c + 1) / chunks2).
__device__ __forceinline__ void rx_chunk2(const RadixParams& R, uint32_t& b, uint32_t& c, uint64_t& lo, uint32_t& c0,
                                          uint32_t& c1) {
  b = blockIdx.x / R.chunks2;
  c = blockIdx.x % R.chunks2;
  lo = R.bucket_start[b];
  const uint64_t n = R.bucket_fill[b] - lo;
  c0 = (uint32_t)(n * c / R.chunks2);
  c1 = (uint32_t)(n * (c + 1) / R.chunks2);
}

// Level-2 digit counts of one chunk into hist2[(b * chunks2 + c) * 2^bits2 + d].
__global__ void __launch_bounds__(kRxThreads) rx_hist2_kernel(const RadixParams R) {
  __shared__ uint32_t cnt[kRxMaxDigits];
  uint32_t b, c, c0, c1;
  uint64_t lo;
  rx_chunk2(R, b, c, lo, c0, c1);
  const uint32_t nb2 = 1u << R.bits2, m2 = nb2 - 1u, stride = R.row.stride;
  for (uint32_t d = threadIdx.x; d < nb2; d += blockDim.x) cnt[d] = 0;
  __syncthreads();
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
    const uint32_t key = R.digits1 ? R.digits1[lo + i] : *reinterpret_cast<const uint32_t*>(R.rows1 + (lo + i) * stride);
    atomicAdd(&cnt[(key >> R.shift2) & m2], 1u);
  }
  __syncthreads();
  uint32_t* out = R.hist2 + (uint64_t)blockIdx.x * nb2;
  for (uint32_t d = threadIdx.x; d < nb2; d += blockDim.x) out[d] = cnt[d];
}

// One workgroup per level-1 bucket: the leaf ranges (leaf_start) and, in
// place, each chunk's first row per digit relative to the bucket start.
__global__ void __launch_bounds__(kRxThreads) rx_offsets2_kernel(const RadixParams R) {
  __shared__ uint32_t tmp[kRxThreads / 64 + 1];
  const uint32_t b = blockIdx.x, nb2 = 1u << R.bits2, C = R.chunks2;
  const uint64_t lo = R.bucket_start[b];
  uint32_t* h = R.hist2 + (uint64_t)b * C * nb2;
  const uint32_t d = threadIdx.x;  // nb2 <= blockDim.x
  uint32_t tot = 0;
  if (d < nb2)
    for (uint32_t c = 0; c < C; c++) tot += h[(uint64_t)c * nb2 + d];
  uint32_t all;
  const uint32_t off = block_excl_scan(tot, tmp, &all);
  if (d < nb2) {
    const uint64_t leaf = (uint64_t)b * nb2 + d;
    R.leaf_start[2 * leaf] = lo + off;
    R.leaf_start[2 * leaf + 1] = lo + off + tot;
    uint32_t run = off;
    for (uint32_t c = 0; c < C; c++) {
      const uint32_t v = h[(uint64_t)c * nb2 + d];
      h[(uint64_t)c * nb2 + d] = run;
      run += v;
    }
  }
}

// Staged scatter of one chunk by level-2 digit, from its offsets (LDS
// cursors, no device atomics).  Tiles of up to 8192 rows (kRx2Per per
// thread) keep each digit's run per tile near a full cache line.  Each row
// is read from HBM once, into registers (W 8-byte words), then staged in
// LDS by digit and written out in runs.
constexpr uint32_t kRx2Per = 8;
inline size_t rx_scatter2_lds_bytes(const RadixParams& r) { return kRxStageOff + (size_t)r.tile_rows2 * r.row.stride; }

template <uint32_t W>
__device__ __forceinline__ void rx_scatter2_chunk(const RadixParams& R, uint8_t* lds, uint64_t lo, uint32_t c0,
                                                  uint32_t c1) {
  uint32_t* cnt = reinterpret_cast<uint32_t*>(lds + kRxCntOff);
  uint32_t* off = reinterpret_cast<uint32_t*>(lds + kRxOffOff);
  uint32_t* cur = reinterpret_cast<uint32_t*>(lds + kRxCurOff);
  uint32_t* tmp = reinterpret_cast<uint32_t*>(lds + kRxTmpOff);
  uint64_t* stage = reinterpret_cast<uint64_t*>(lds + kRxStageOff);
  const uint32_t nb2 = 1u << R.bits2, m2 = nb2 - 1u;
  const uint64_t* rows1 = reinterpret_cast<const uint64_t*>(R.rows1);
  uint64_t* rows2 = reinterpret_cast<uint64_t*>(R.rows2);
  for (uint32_t s0 = c0; s0 < c1; s0 += R.tile_rows2) {  // uniform
    const uint32_t tn = c1 - s0 < R.tile_rows2 ? c1 - s0 : R.tile_rows2;
    for (uint32_t d = threadIdx.x; d < nb2; d += blockDim.x) cnt[d] = 0;
    __syncthreads();
    uint64_t v[kRx2Per][W];
    uint32_t digit[kRx2Per], rank[kRx2Per];
#pragma unroll
    for (uint32_t j = 0; j < kRx2Per; j++) {
      const uint32_t i = threadIdx.x + j * blockDim.x;
      if (i < tn) {
        const uint64_t* src = rows1 + (lo + s0 + i) * W;
#pragma unroll
        for (uint32_t w = 0; w < W; w++) v[j][w] = src[w];
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < kRx2Per; j++) {
      const uint32_t i = threadIdx.x + j * blockDim.x;
      if (i < tn) {
        digit[j] = ((uint32_t)v[j][0] >> R.shift2) & m2;  // the u32 key leads the row
        rank[j] = atomicAdd(&cnt[digit[j]], 1u);
      }
    }
    __syncthreads();
    rx_prefix(cnt, off, nb2, tmp);
#pragma unroll
    for (uint32_t j = 0; j < kRx2Per; j++) {
      const uint32_t i = threadIdx.x + j * blockDim.x;
      if (i < tn) {
        uint64_t* dst = stage + (size_t)(off[digit[j]] + rank[j]) * W;
#pragma unroll
        for (uint32_t w = 0; w < W; w++) dst[w] = v[j][w];
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tn; i += blockDim.x) {
      const uint64_t* row = stage + (size_t)i * W;
      const uint32_t d = ((uint32_t)row[0] >> R.shift2) & m2;
      uint64_t* dst = rows2 + (lo + cur[d] + (i - off[d])) * W;
#pragma unroll
      for (uint32_t w = 0; w < W; w++) dst[w] = row[w];
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < nb2; d += blockDim.x) cur[d] += cnt[d];
    __syncthreads();
  }
}

// Rows wider than 4 words: the key is read first, the row again when staged.
__device__ __forceinline__ void rx_scatter2_chunk_wide(const RadixParams& R, uint8_t* lds, uint64_t lo, uint32_t c0,
                                                       uint32_t c1) {
  uint32_t* cnt = reinterpret_cast<uint32_t*>(lds + kRxCntOff);
  uint32_t* off = reinterpret_cast<uint32_t*>(lds + kRxOffOff);
  uint32_t* cur = reinterpret_cast<uint32_t*>(lds + kRxCurOff);
  uint32_t* tmp = reinterpret_cast<uint32_t*>(lds + kRxTmpOff);
  uint8_t* stage = lds + kRxStageOff;
  const uint32_t nb2 = 1u << R.bits2, m2 = nb2 - 1u, stride = R.row.stride;
  for (uint32_t s0 = c0; s0 < c1; s0 += R.tile_rows2) {  // uniform
    const uint32_t tn = c1 - s0 < R.tile_rows2 ? c1 - s0 : R.tile_rows2;
    for (uint32_t d = threadIdx.x; d < nb2; d += blockDim.x) cnt[d] = 0;
    __syncthreads();
    uint32_t digit[kRx2Per], rank[kRx2Per];
#pragma unroll
    for (uint32_t j = 0; j < kRx2Per; j++) {
      const uint32_t i = threadIdx.x + j * blockDim.x;
      if (i < tn) {
        const uint32_t key = *reinterpret_cast<const uint32_t*>(R.rows1 + (lo + s0 + i) * stride);
        digit[j] = (key >> R.shift2) & m2;
        rank[j] = atomicAdd(&cnt[digit[j]], 1u);
      }
    }
    __syncthreads();
    rx_prefix(cnt, off, nb2, tmp);
#pragma unroll
    for (uint32_t j = 0; j < kRx2Per; j++) {
      const uint32_t i = threadIdx.x + j * blockDim.x;
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

__global__ void __launch_bounds__(kRxThreads) rx_scatter2_kernel(const RadixParams R) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint32_t* cur = reinterpret_cast<uint32_t*>(lds + kRxCurOff);
  uint32_t b, c, c0, c1;
  uint64_t lo;
  rx_chunk2(R, b, c, lo, c0, c1);
  const uint32_t nb2 = 1u << R.bits2;
  const uint32_t* h = R.hist2 + (uint64_t)blockIdx.x * nb2;
  for (uint32_t d = threadIdx.x; d < nb2; d += blockDim.x) cur[d] = h[d];
  switch (R.row.stride / 8) {  // uniform
    case 1: rx_scatter2_chunk<1>(R, lds, lo, c0, c1); break;
    case 2: rx_scatter2_chunk<2>(R, lds, lo, c0, c1); break;
    case 3: rx_scatter2_chunk<3>(R, lds, lo, c0, c1); break;
    case 4: rx_scatter2_chunk<4>(R, lds, lo, c0, c1); break;
    default: rx_scatter2_chunk_wide(R, lds, lo, c0, c1); break;
  }
}

constexpr uint32_t kRxAggTmp = 512;  // LDS words for compaction counts: 9 + table slots / 64

// HAVING: compacted leaves write only the groups passing R.having (a
// separate instantiation, so the plain leaf pass keeps its registers).
template <bool HAVING>
__global__ void __launch_bounds__(512) rx_agg_kernel(const ScanParams P, const RadixParams R) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ uint32_t tmp[kRxAggTmp];
  __shared__ uint32_t chist[kRxCountBins];
  __shared__ uint32_t lhist[HAVING ? 1 : kRxCountBins];  // per-leaf top-k: the leaf's groups by count bin
  if (R.counts_out || R.leaf_topk) {  // uniform
    for (uint32_t b = threadIdx.x; b < kRxCountBins; b += blockDim.x) {
      chist[b] = 0;
      if (!HAVING) lhist[b] = 0;
    }
    __syncthreads();
  }
  for (uint32_t leaf = blockIdx.x; leaf < R.n_leaves; leaf += gridDim.x) {
    if (HAVING) rx_agg_leaf<false, true>(P, R, leaf, lds, tmp, chist);
    else if (R.dense_out) rx_agg_leaf<true>(P, R, leaf, lds, tmp);
    else rx_agg_leaf<false>(P, R, leaf, lds, tmp, chist, lhist);
  }
  if (R.counts_out) {
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < kRxCountBins; b += blockDim.x)
      if (chist[b]) atomicAdd(&R.count_hist[b], chist[b]);
  }
}

// Hash partitions: the same persistent leaf loop over LDS hash tables.
template <bool HAVING>
__global__ void __launch_bounds__(512) rx_agg_hashed_kernel(const ScanParams P, const RadixParams R) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ uint32_t tmp[kRxAggTmp];
  __shared__ uint32_t chist[kRxCountBins];
  __shared__ uint32_t n_claimed;
  if (R.counts_out) {  // uniform
    for (uint32_t b = threadIdx.x; b < kRxCountBins; b += blockDim.x) chist[b] = 0;
    __syncthreads();
  }
  rx_hashed_clear(P, R, lds, &n_claimed);  // each leaf leaves the table as clean as it found it
  for (uint32_t leaf = blockIdx.x; leaf < R.n_leaves; leaf += gridDim.x)
    rx_agg_leaf_hashed<HAVING>(P, R, leaf, lds, tmp, chist, &n_claimed);
  if (R.counts_out) {
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < kRxCountBins; b += blockDim.x)
      if (chist[b]) atomicAdd(&R.count_hist[b], chist[b]);
  }
}

// Compacted leaves: group positions, leaf by leaf.  The per-leaf counts
// are summed per block of 256 leaves (rx_gsum), the block sums scanned
// (rx_gscan: the total to *n_out and leaf_groups[n_leaves]), and each block
// then scans its own counts, turns them into the leaves' first positions in
// place and writes its leaves' positions, one wavefront per leaf (rx_gids).
// The block sums sit after leaf_groups[n_leaves].
constexpr uint32_t kRxGidsBlock = 256;

__global__ void __launch_bounds__(kRxGidsBlock) rx_gsum_kernel(const RadixParams R) {
  __shared__ uint32_t red[kRxGidsBlock / 64];
  const uint32_t leaf = blockIdx.x * kRxGidsBlock + threadIdx.x;
  uint32_t v = leaf < R.n_leaves ? R.leaf_groups[leaf] : 0u;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t k = 0; k < kRxGidsBlock / 64; k++) t += red[k];
    R.leaf_groups[R.n_leaves + 1 + blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(1024) rx_gscan_kernel(const RadixParams R, uint32_t n_blocks, uint64_t* n_out) {
  __shared__ uint32_t tmp[1024 / 64 + 1];
  uint32_t* sums = R.leaf_groups + R.n_leaves + 1;
  const uint32_t per = (n_blocks + blockDim.x - 1) / blockDim.x;
  const uint32_t lo = threadIdx.x * per;
  uint32_t s = 0;
  for (uint32_t i = 0; i < per; i++)
    if (lo + i < n_blocks) s += sums[lo + i];
  uint32_t total;
  uint32_t run = block_excl_scan(s, tmp, &total);
  for (uint32_t i = 0; i < per; i++)
    if (lo + i < n_blocks) {
      const uint32_t c = sums[lo + i];
      sums[lo + i] = run;
      run += c;
    }
  if (threadIdx.x == 0) {
    *n_out = total;
    R.leaf_groups[R.n_leaves] = total;
  }
}

__global__ void __launch_bounds__(kRxGidsBlock) rx_gids_kernel(const RadixParams R, uint32_t* gids) {
  __shared__ uint32_t tmp[kRxGidsBlock / 64 + 1];
  __shared__ uint32_t s_off[kRxGidsBlock], s_cnt[kRxGidsBlock];
  __shared__ uint64_t s_row[kRxGidsBlock];
  const uint32_t leaf = blockIdx.x * kRxGidsBlock + threadIdx.x;
  const uint32_t c = leaf < R.n_leaves ? R.leaf_groups[leaf] : 0u;
  uint32_t total;
  const uint32_t ex = block_excl_scan(c, tmp, &total);
  s_off[threadIdx.x] = R.leaf_groups[R.n_leaves + 1 + blockIdx.x] + ex;
  s_cnt[threadIdx.x] = c;
  if (leaf < R.n_leaves) R.leaf_groups[leaf] = s_off[threadIdx.x];  // counts become offsets
  s_row[threadIdx.x] = leaf < R.n_leaves ? (R.bits2 ? R.leaf_start[2 * (uint64_t)leaf] : R.bucket_start[leaf]) : 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, w0 = (threadIdx.x >> 6) * 64;
  for (uint32_t k = w0; k < w0 + 64; k++) {
    const uint32_t o = s_off[k], n = s_cnt[k];
    const uint64_t r = s_row[k];
    for (uint32_t j = lane; j < n; j += 64) gids[(uint64_t)o + j] = (uint32_t)(r + j);
  }
}

// ORDER BY COUNT(*): one wavefront per leaf (after rx_gids) appends the
// position index of each group whose count passes the cut.
__global__ void __launch_bounds__(256) rx_count_take_kernel(const RadixParams R, uint32_t cut, uint32_t desc,
                                                            uint32_t* idx, uint32_t cap, uint32_t* n_idx) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t leaf = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); leaf < R.n_leaves; leaf += waves) {
    const uint32_t o = R.leaf_groups[leaf];
    const uint32_t n = R.leaf_groups[leaf + 1] - o;  // offsets, and the total at n_leaves (rx_gids)
    const uint64_t r = R.bits2 ? R.leaf_start[2 * (uint64_t)leaf] : R.bucket_start[leaf];
    for (uint32_t j0 = 0; j0 < n; j0 += 64) {  // uniform per wavefront
      const uint32_t j = j0 + lane;
      const uint32_t c = j < n ? R.counts_out[r + j] : 0u;
      const bool take = j < n && (desc ? c >= cut : c <= cut);
      const uint64_t m = __ballot(take);
      if (m == 0) continue;
      const int first = __ffsll((long long)m) - 1;
      uint32_t base = 0;
      if ((int)lane == first) base = atomicAdd(n_idx, (uint32_t)__popcll(m));
      base = __shfl(base, first);
      const uint32_t at = base + (uint32_t)__popcll(m & ((1ull << lane) - 1));
      if (take && at < cap) idx[at] = o + j;
    }
  }
}

__global__ void rx_zero_kernel(uint32_t* p) { *p = 0; }

hipError_t launch_radix_count_take(const RadixParams& r, uint32_t cut, bool desc, uint32_t* idx, uint32_t cap,
                                   uint32_t* n_idx, hipStream_t s) {
  if (r.dense_out || !r.counts_out || !r.leaf_groups || r.n_leaves == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(rx_zero_kernel, dim3(1), dim3(1), 0, s, n_idx);
  const uint32_t grid = (r.n_leaves + 3) / 4 < 16384 ? (r.n_leaves + 3) / 4 : 16384;
  hipLaunchKernelGGL(rx_count_take_kernel, dim3(grid), dim3(256), 0, s, r, cut, desc ? 1u : 0u, idx, cap, n_idx);
  return hipGetLastError();
}

hipError_t launch_radix_gids(const RadixParams& r, uint32_t* gids, uint64_t* n_out, hipStream_t s) {
  if (r.dense_out || !r.leaf_groups || r.n_leaves == 0) return hipErrorInvalidValue;
  const uint32_t blocks = (r.n_leaves + kRxGidsBlock - 1) / kRxGidsBlock;
  hipLaunchKernelGGL(rx_gsum_kernel, dim3(blocks), dim3(kRxGidsBlock), 0, s, r);
  hipLaunchKernelGGL(rx_gscan_kernel, dim3(1), dim3(1024), 0, s, r, blocks, n_out);
  hipLaunchKernelGGL(rx_gids_kernel, dim3(blocks), dim3(kRxGidsBlock), 0, s, r, gids);
  return hipGetLastError();
}

hipError_t launch_radix(const ScanParams& p, const RadixParams& r, uint32_t grid, unsigned long long* hist,
                        hipStream_t s, int phase) {
  static const hipError_t attr = [] {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(rx_scatter1_kernel<false>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(rx_scatter1_kernel<true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(rx_agg_hashed_kernel<false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 4096);
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(rx_agg_hashed_kernel<true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 4096);
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(rx_scatter2_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e == hipSuccess)  // rx_agg also holds ~2 KiB of static LDS
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(rx_agg_kernel<false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 4096);
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(rx_agg_kernel<true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 4096);
    return e;
  }();
  if (attr != hipSuccess) return attr;
  // hash partitions: digits from a 32-bit hash, any table size (compacted,
  // keyed records); key-range partitions: a leaf's 2^leaf_bits keys fill the table
  const bool table_ok = r.hashed ? (r.bits1 + r.bits2 + r.leaf_bits == 32 && !r.dense_out && p.lds_cap >= 64 &&
                                    p.g_mode != G_DENSE && (r.row.key_words == 1 || r.row.key_words == 2) &&
                                    (r.row.key_off & 7u) == 0 && r.row.key_off + 8 * r.row.key_words <= r.row.stride &&
                                    (p.g_mode == G_WIDE) == (r.row.key_words == 2))
                                 : p.lds_cap == (1u << r.leaf_bits);
  if (r.bits1 > kRxMaxBits1 || r.bits2 > 10 || (r.row.stride & 7u) || r.tile_rows == 0 || r.tile_rows > kRxTile ||
      (r.tile_rows & (kVec - 1)) || rx_scatter_lds_bytes(r) > 160 * 1024 || grid == 0 || !table_ok ||
      r.n_leaves != (1u << (r.bits1 + r.bits2)) || 512 / 64 + 1 + (p.lds_cap + 511) / 512 * (512 / 64) > kRxAggTmp ||
      (!r.dense_out && !r.leaf_groups) ||
      (r.bits2 && (r.chunks2 == 0 || r.chunks2 > 64 || !r.hist2 || !r.leaf_start || r.tile_rows2 == 0 ||
                   r.tile_rows2 > kRxThreads * kRx2Per || rx_scatter2_lds_bytes(r) > 160 * 1024)) ||
      p.lds_bytes > 160 * 1024 - 4096)
    return hipErrorInvalidConfiguration;
  switch (phase) {
    case RX_HIST:
      if (r.hashed) hipLaunchKernelGGL(rx_hist_kernel<true>, dim3(grid), dim3(kRxThreads), 0, s, p, r, hist);
      else hipLaunchKernelGGL(rx_hist_kernel<false>, dim3(grid), dim3(kRxThreads), 0, s, p, r, hist);
      break;
    case RX_OFFSETS: hipLaunchKernelGGL(rx_offsets_kernel, dim3(1), dim3(256), 0, s, r, hist); break;
    case RX_SCATTER1:
      if (r.hashed)
        hipLaunchKernelGGL(rx_scatter1_kernel<true>, dim3(grid), dim3(kRxThreads), rx_scatter_lds_bytes(r), s, p, r);
      else
        hipLaunchKernelGGL(rx_scatter1_kernel<false>, dim3(grid), dim3(kRxThreads), rx_scatter_lds_bytes(r), s, p, r);
      break;
    case RX_SCATTER2: {
      const dim3 g((1u << r.bits1) * r.chunks2);
      hipLaunchKernelGGL(rx_hist2_kernel, g, dim3(kRxThreads), 0, s, r);
      hipLaunchKernelGGL(rx_offsets2_kernel, dim3(1u << r.bits1), dim3(kRxThreads), 0, s, r);
      hipLaunchKernelGGL(rx_scatter2_kernel, g, dim3(kRxThreads), rx_scatter2_lds_bytes(r), s, r);
      break;
    }
    case RX_AGG:
      if (r.counts_out && (r.dense_out || !r.count_hist)) return hipErrorInvalidValue;
      if (r.hashed && r.having)
        hipLaunchKernelGGL(rx_agg_hashed_kernel<true>, dim3(grid), dim3(512), p.lds_bytes, s, p, r);
      else if (r.hashed)
        hipLaunchKernelGGL(rx_agg_hashed_kernel<false>, dim3(grid), dim3(512), p.lds_bytes, s, p, r);
      else if (r.having && !r.dense_out)
        hipLaunchKernelGGL(rx_agg_kernel<true>, dim3(grid), dim3(512), p.lds_bytes, s, p, r);
      else
        hipLaunchKernelGGL(rx_agg_kernel<false>, dim3(grid), dim3(512), p.lds_bytes, s, p, r);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace tdbg
