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
#include "scan_common.h"

namespace tdbg {

constexpr uint32_t kRxThreads = 1024;
constexpr uint32_t kRxTile = kRxThreads * kVec;  // most rows staged per tile (R.tile_rows <= this)

// Filter and key of one strip (kinds read at run time).
__device__ __forceinline__ uint32_t rx_strip_keys(const ScanParams& P, uint64_t r, uint32_t cnt, uint64_t key[kVec],
                                                  bool with_preds) {
  uint32_t mask = (1u << cnt) - 1u;
  if (with_preds) {
    for (uint32_t i = 0; i < P.n_preds; i++) {
      const PredDesc& pd = P.preds[i];
      const ColDesc& c = P.cols[pd.col];
      mask &= pred_rt(pd, load_raw(c, r), c.kind) & valid4(c, r);
    }
  }
  for (uint32_t i = 0; i < P.n_parts; i++) {
    const PartDesc& pt = P.parts[i];
    int64_t v[kVec];
    decode_i(load_raw(P.cols[pt.col], r), P.cols[pt.col].kind, v);
    if (i == 0) key_part<true>(P, pt, v, key, mask);
    else key_part<false>(P, pt, v, key, mask);
  }
  return mask;
}

// The tiles this workgroup scans (contiguous share of the tile list).
__device__ __forceinline__ void rx_tile_range(const ScanParams& P, uint32_t& t0, uint32_t& t1) {
  t0 = (uint32_t)(((uint64_t)blockIdx.x * P.n_tiles) / gridDim.x);
  t1 = (uint32_t)(((uint64_t)(blockIdx.x + 1) * P.n_tiles) / gridDim.x);
}

constexpr uint32_t kRxMaxBits1 = 8;  // level-1 fanout <= 256

__global__ void __launch_bounds__(kRxThreads) rx_hist_kernel(const ScanParams P, const RadixParams R,
                                                             unsigned long long* hist) {
  __shared__ uint32_t h[1u << kRxMaxBits1];
  const uint32_t nb = 1u << R.bits1;
  for (uint32_t d = threadIdx.x; d < nb; d += blockDim.x) h[d] = 0;
  __syncthreads();
  uint32_t t0, t1;
  rx_tile_range(P, t0, t1);
  for (uint32_t t = t0; t < t1; t++) {
    const Tile T = P.tiles[t];
    for (uint32_t s = threadIdx.x * kVec; s < T.n; s += blockDim.x * kVec) {
      const uint32_t cnt = T.n - s < (uint32_t)kVec ? T.n - s : (uint32_t)kVec;
      uint64_t key[kVec];
      const uint32_t mask = rx_strip_keys(P, T.row0 + s, cnt, key, false);
#pragma unroll
      for (int j = 0; j < kVec; j++)
        if ((mask >> j) & 1u) atomicAdd(&h[(uint32_t)(key[j] >> R.shift1)], 1u);
    }
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < nb; d += blockDim.x)
    if (h[d]) atomicAdd(&hist[d], (unsigned long long)h[d]);
}

// Level-1 bucket starts from the histogram (one thread; at most 256 buckets):
// bucket_start[d] = rows in buckets < d, bucket_fill[d] = bucket_start[d].
__global__ void rx_offsets_kernel(const RadixParams R, const unsigned long long* hist) {
  if (threadIdx.x != 0) return;
  const uint32_t nb = 1u << R.bits1;
  uint64_t run = 0;
  for (uint32_t d = 0; d < nb; d++) {
    R.bucket_start[d] = run;
    R.bucket_fill[d] = run;
    run += hist[d];
  }
  R.bucket_start[nb] = run;
}

// Exclusive prefix over nb (<= blockDim.x * per) LDS counters into `off`;
// returns the total.
__device__ __forceinline__ uint32_t rx_prefix(const uint32_t* cnt, uint32_t* off, uint32_t nb, uint32_t* tmp) {
  const uint32_t per = (nb + blockDim.x - 1) / blockDim.x;
  const uint32_t lo = threadIdx.x * per;
  uint32_t s = 0;
  for (uint32_t i = 0; i < per; i++)
    if (lo + i < nb) s += cnt[lo + i];
  tmp[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t o = 1; o < blockDim.x; o <<= 1) {
    const uint32_t x = threadIdx.x >= o ? tmp[threadIdx.x - o] : 0;
    __syncthreads();
    tmp[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t run = tmp[threadIdx.x] - s;
  for (uint32_t i = 0; i < per; i++)
    if (lo + i < nb) {
      off[lo + i] = run;
      run += cnt[lo + i];
    }
  const uint32_t total = tmp[blockDim.x - 1];
  __syncthreads();
  return total;
}

// Stores row j's value (b bytes wide) out of a strip's raw words.
__device__ __forceinline__ void rx_put_value(uint8_t* dst, const Raw& x, uint32_t b, int j) {
  uint64_t w;
  if (b == 8) w = j == 0 ? x.w0 : (j == 1 ? x.w1 : (j == 2 ? x.w2 : x.w3));
  else if (b == 4) w = (j < 2 ? x.w0 : x.w1) >> (32 * (j & 1));
  else if (b == 2) w = x.w0 >> (16 * j);
  else w = x.w0 >> (8 * j);
  if (b == 8) *reinterpret_cast<uint64_t*>(dst) = w;
  else if (b == 4) *reinterpret_cast<uint32_t*>(dst) = (uint32_t)w;
  else if (b == 2) *reinterpret_cast<uint16_t*>(dst) = (uint16_t)w;
  else *dst = (uint8_t)w;
}

// Copies `n` staged rows to their destinations (stride is a multiple of 8).
__device__ __forceinline__ void rx_copy_row(uint8_t* dst, const uint8_t* src, uint32_t stride) {
  for (uint32_t w = 0; w < stride; w += 8)
    *reinterpret_cast<uint64_t*>(dst + w) = *reinterpret_cast<const uint64_t*>(src + w);
}

// LDS layout shared by the two scatter kernels (digits of at most 10 bits):
//   cnt[1024] off[1024] cur[1024] tmp[1024] (u32), base[1024] (u64), then
//   R.tile_rows staged rows (tile_rows * stride <= 136 KiB).
constexpr uint32_t kRxMaxDigits = 1024;
constexpr uint32_t kRxCntOff = 0;
constexpr uint32_t kRxOffOff = 4 * kRxMaxDigits;
constexpr uint32_t kRxCurOff = 8 * kRxMaxDigits;
constexpr uint32_t kRxTmpOff = 12 * kRxMaxDigits;
constexpr uint32_t kRxBaseOff = 16 * kRxMaxDigits;
constexpr uint32_t kRxStageOff = 24 * kRxMaxDigits;

size_t rx_scatter_lds_bytes(const RadixParams& r) { return kRxStageOff + (size_t)r.tile_rows * r.row.stride; }

__global__ void __launch_bounds__(kRxThreads) rx_scatter1_kernel(const ScanParams P, const RadixParams R) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint32_t* cnt = reinterpret_cast<uint32_t*>(lds + kRxCntOff);
  uint32_t* off = reinterpret_cast<uint32_t*>(lds + kRxOffOff);
  uint32_t* tmp = reinterpret_cast<uint32_t*>(lds + kRxTmpOff);
  uint64_t* base = reinterpret_cast<uint64_t*>(lds + kRxBaseOff);
  uint8_t* stage = lds + kRxStageOff;
  const uint32_t nb = 1u << R.bits1, stride = R.row.stride;
  uint32_t t0, t1;
  rx_tile_range(P, t0, t1);
  for (uint32_t t = t0; t < t1; t++) {
    const Tile T = P.tiles[t];
    for (uint32_t s0 = 0; s0 < T.n; s0 += R.tile_rows) {  // uniform sub-tile loop
      for (uint32_t d = threadIdx.x; d < nb; d += blockDim.x) cnt[d] = 0;
      __syncthreads();
      const uint32_t s = s0 + threadIdx.x * kVec;
      const uint64_t r = T.row0 + s;
      uint32_t mask = 0, digit[kVec], rank[kVec];
      uint64_t key[kVec];
      if (threadIdx.x * kVec < R.tile_rows && s < T.n) {
        const uint32_t c = T.n - s < (uint32_t)kVec ? T.n - s : (uint32_t)kVec;
        mask = rx_strip_keys(P, r, c, key, true);
#pragma unroll
        for (int j = 0; j < kVec; j++)
          if ((mask >> j) & 1u) {
            digit[j] = (uint32_t)(key[j] >> R.shift1);
            rank[j] = atomicAdd(&cnt[digit[j]], 1u);
          }
      }
      __syncthreads();
      const uint32_t total = rx_prefix(cnt, off, nb, tmp);
      for (uint32_t d = threadIdx.x; d < nb; d += blockDim.x)
        base[d] = cnt[d] ? atomicAdd(reinterpret_cast<unsigned long long*>(&R.bucket_fill[d]),
                                     (unsigned long long)cnt[d])
                         : 0;
      if (mask) {
        // stage the rows: key, then each aggregate input (the loads hit the
        // cache lines the key pass just fetched), then the validity word
        uint8_t* row[kVec];
        uint32_t flags[kVec] = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < kVec; j++) {
          row[j] = stage + (size_t)(((mask >> j) & 1u) ? off[digit[j]] + rank[j] : 0) * stride;
          if ((mask >> j) & 1u) *reinterpret_cast<uint32_t*>(row[j]) = (uint32_t)key[j];
        }
        for (uint32_t a = 0; a < P.n_aggs; a++) {
          const ColDesc& col = P.cols[P.aggs[a].col];
          const uint32_t vm = valid4(col, r);
          const uint32_t b = R.row.val_bytes[a];
          const Raw x = b ? load_raw(col, r) : Raw{0, 0, 0, 0};
#pragma unroll
          for (int j = 0; j < kVec; j++) {
            if (!((mask >> j) & 1u)) continue;
            flags[j] |= ((vm >> j) & 1u) << a;
            if (b) rx_put_value(row[j] + R.row.val_off[a], x, b, j);
          }
        }
        if (R.row.flags_off != ~0u) {
#pragma unroll
          for (int j = 0; j < kVec; j++)
            if ((mask >> j) & 1u) *reinterpret_cast<uint32_t*>(row[j] + R.row.flags_off) = flags[j];
        }
      }
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
        const uint8_t* row = stage + (size_t)i * stride;
        const uint32_t d = *reinterpret_cast<const uint32_t*>(row) >> R.shift1;
        rx_copy_row(R.rows1 + (base[d] + (i - off[d])) * stride, row, stride);
      }
      __syncthreads();
    }
  }
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

// One row of a leaf folded into the LDS table.
__device__ __forceinline__ void rx_fold_row(const ScanParams& P, const RadixParams& R, uint8_t* lds,
                                            const uint8_t* row, uint32_t slot) {
  l_add_u32(lds_at<uint32_t>(lds, P.lds_count_off) + slot, 1);
  const uint32_t flags = R.row.flags_off != ~0u ? *reinterpret_cast<const uint32_t*>(row + R.row.flags_off) : ~0u;
  for (uint32_t a = 0; a < P.n_aggs; a++) {
    const AggDesc& d = P.aggs[a];
    if (!((flags >> a) & 1u)) continue;
    const uint32_t b = R.row.val_bytes[a];
    const uint8_t* p = row + R.row.val_off[a];
    int64_t iv = 0;
    double fv = 0;
    if (d.kind == A_SUM_F || d.kind == A_MIN_F || d.kind == A_MAX_F) {
      fv = b == 8 ? *reinterpret_cast<const double*>(p) : (double)*reinterpret_cast<const float*>(p);
    } else if (b == 8) {
      iv = *reinterpret_cast<const int64_t*>(p);
    } else if (b == 4) {
      iv = *reinterpret_cast<const int32_t*>(p);
    } else if (b == 2) {
      iv = *reinterpret_cast<const int16_t*>(p);
    } else if (b == 1) {
      iv = P.cols[d.col].kind == K_U8 ? (int64_t)*p : (int64_t)*reinterpret_cast<const int8_t*>(p);
    }
    switch (d.kind) {
      case A_COUNT_COL: l_apply<A_COUNT_COL, false, false>(d, lds, P.lds_cap, slot, iv, fv); break;
      case A_SUM_I:
        if (d.split) {
          if (d.nullable) l_apply<A_SUM_I, true, true>(d, lds, P.lds_cap, slot, iv, fv);
          else l_apply<A_SUM_I, false, true>(d, lds, P.lds_cap, slot, iv, fv);
        } else {
          if (d.nullable) l_apply<A_SUM_I, true, false>(d, lds, P.lds_cap, slot, iv, fv);
          else l_apply<A_SUM_I, false, false>(d, lds, P.lds_cap, slot, iv, fv);
        }
        break;
#define TDBG_RX_CASE(AK)                                                       \
  case AK:                                                                     \
    if (d.nullable) l_apply<AK, true, false>(d, lds, P.lds_cap, slot, iv, fv); \
    else l_apply<AK, false, false>(d, lds, P.lds_cap, slot, iv, fv);           \
    break;
      TDBG_RX_CASE(A_SUM_F)
      TDBG_RX_CASE(A_MIN_I)
      TDBG_RX_CASE(A_MAX_I)
      TDBG_RX_CASE(A_MIN_F)
      TDBG_RX_CASE(A_MAX_F)
#undef TDBG_RX_CASE
      default: break;
    }
  }
}

// Writes slot s of the LDS table as a group record (key word, count, each
// aggregate's HBM state) at `rec`.
__device__ __forceinline__ void rx_write_record(const ScanParams& P, uint8_t* lds, uint32_t s, uint64_t key,
                                                uint8_t* rec) {
  const uint32_t cap = P.lds_cap;
  *reinterpret_cast<uint64_t*>(rec + P.g_key_off) = key;
  *reinterpret_cast<uint64_t*>(rec + P.g_count_off) = lds_at<uint32_t>(lds, P.lds_count_off)[s];
  for (uint32_t a = 0; a < P.n_aggs; a++) {
    const AggDesc& d = P.aggs[a];
    uint8_t* g = rec + d.g_off;
    if (d.kind == A_COUNT_COL) {
      *reinterpret_cast<uint64_t*>(g) = lds_at<uint32_t>(lds, d.lds_off)[s];
      continue;
    }
    const uint64_t w = lds_at<uint64_t>(lds, d.lds_off)[s];
    if (d.kind == A_SUM_I) {
      __int128 v;
      if (d.split) {
        const int64_t hs = (int64_t)lds_at<uint64_t>(lds, d.lds_off)[cap + s];
        v = (__int128)hs * ((__int128)1 << 32) + (__int128)w;
      } else {
        v = (__int128)(int64_t)w;
      }
      *reinterpret_cast<uint64_t*>(g) = (uint64_t)v;
      *reinterpret_cast<uint64_t*>(g + 8) = (uint64_t)(v >> 64);
    } else {
      *reinterpret_cast<uint64_t*>(g) = w;  // f64 bits, or (ordered) i64
    }
    if (d.nullable) *reinterpret_cast<uint64_t*>(rec + d.g_ns_off) = lds_at<uint32_t>(lds, d.lds_ns_off)[s];
  }
}

__global__ void __launch_bounds__(512) rx_agg_kernel(const ScanParams P, const RadixParams R) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ uint32_t tmp[512];
  __shared__ uint64_t out_base;
  const uint32_t cap = P.lds_cap, stride = R.row.stride;
  const uint64_t mask = (uint64_t)cap - 1u;
  for (uint32_t leaf = blockIdx.x; leaf < R.n_leaves; leaf += gridDim.x) {
    uint64_t s, e;
    if (R.bits2) {
      s = R.leaf_start[2 * (uint64_t)leaf];
      e = R.leaf_start[2 * (uint64_t)leaf + 1];
    } else {
      s = R.bucket_start[leaf];
      e = R.bucket_fill[leaf];
    }
    if (s == e) continue;  // uniform
    lds_init(P, lds);
    __syncthreads();
    for (uint64_t i = s + threadIdx.x; i < e; i += blockDim.x) {
      const uint8_t* row = R.rows2 + i * stride;
      const uint32_t key = *reinterpret_cast<const uint32_t*>(row);
      rx_fold_row(P, R, lds, row, (uint32_t)(key & mask));
    }
    __syncthreads();
    // occupied slots, in slot order, to consecutive records
    const uint32_t per = (cap + blockDim.x - 1) / blockDim.x;
    const uint32_t lo = threadIdx.x * per;
    uint32_t mine = 0;
    for (uint32_t k = 0; k < per; k++)
      if (lo + k < cap && lds_at<uint32_t>(lds, P.lds_count_off)[lo + k]) mine++;
    tmp[threadIdx.x] = mine;
    __syncthreads();
    for (uint32_t o = 1; o < blockDim.x; o <<= 1) {
      const uint32_t x = threadIdx.x >= o ? tmp[threadIdx.x - o] : 0;
      __syncthreads();
      tmp[threadIdx.x] += x;
      __syncthreads();
    }
    if (threadIdx.x == blockDim.x - 1)
      out_base = atomicAdd(reinterpret_cast<unsigned long long*>(R.out_cursor), (unsigned long long)tmp[blockDim.x - 1]);
    __syncthreads();
    uint64_t at = out_base + tmp[threadIdx.x] - mine;
    const uint64_t leaf_key0 = (uint64_t)leaf << R.leaf_bits;
    for (uint32_t k = 0; k < per; k++) {
      const uint32_t slot = lo + k;
      if (slot < cap && lds_at<uint32_t>(lds, P.lds_count_off)[slot])
        rx_write_record(P, lds, slot, leaf_key0 + slot, P.g_state + (at++) * P.g_rec);
    }
    __syncthreads();
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
    case RX_OFFSETS: hipLaunchKernelGGL(rx_offsets_kernel, dim3(1), dim3(64), 0, s, r, hist); break;
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
