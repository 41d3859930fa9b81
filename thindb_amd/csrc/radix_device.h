// radix_device.h — device routines of the radix-partitioned GROUP BY
// (radix.hip), shared by its own kernels and by the batched persistent
// kernel (batch.hip), which runs them on work items of many plans.
#pragma once

#include "order_device.h"
#include "scan_common.h"

namespace tdbg {


constexpr uint32_t kRxThreads = 1024;
constexpr uint32_t kRxTile = kRxThreads * kVec;  // most rows staged per tile (R.tile_rows <= this)

// Filter and key of one strip (kinds read at run time).
__device__ __forceinline__ uint32_t rx_strip_keys(const ScanParams& P, uint64_t r, uint32_t cnt, uint64_t key[kVec],
                                                  bool with_preds, bool with_parts = true) {
  uint32_t mask = (1u << cnt) - 1u;
  if (with_preds) {
    for (uint32_t i = 0; i < P.n_preds; i++) {
      const PredDesc& pd = P.preds[i];
      const ColDesc& c = P.cols[pd.col];
      mask &= pred_rt(pd, load_raw(c, r), c.kind) & valid4(c, r);
    }
    if (P.n_filt) mask &= filter_mask(P, r, [&](uint32_t col) { return load_raw(P.cols[col], r); });
  }
  if (!with_parts) return mask;
  for (uint32_t i = 0; i < P.n_parts; i++) {
    const PartDesc& pt = P.parts[i];
    int64_t v[kVec];
    decode_i(load_raw(P.cols[pt.col], r), P.cols[pt.col].kind, v);
    const uint32_t pv = pt.nullable ? valid4(P.cols[pt.col], r) : 0xfu;
    if (i == 0) key_part<true>(P, pt, v, key, mask, pv);
    else key_part<false>(P, pt, v, key, mask, pv);
  }
  return mask;
}

// The packed 128-bit key (G_WIDE) of a strip's live rows: each part's
// relative value (NULL = 0) as its bit field, the first part lowest.  A
// value outside the part's zone-map range raises ERR_KEY_RANGE and drops the
// row, as key_part does.
__device__ __forceinline__ uint32_t rx_strip_wide(const ScanParams& P, uint64_t r, uint32_t mask, uint64_t lo[kVec],
                                                  uint64_t hi[kVec]) {
#pragma unroll
  for (int j = 0; j < kVec; j++) lo[j] = hi[j] = 0;
  for (uint32_t i = 0; i < P.n_parts; i++) {
    const PartDesc& pt = P.parts[i];
    int64_t v[kVec];
    decode_i(load_raw(P.cols[pt.col], r), P.cols[pt.col].kind, v);
    const uint32_t pv = pt.nullable ? valid4(P.cols[pt.col], r) : 0xfu;
    const uint32_t o = pt.bit_off, w = pt.bit_width;
#pragma unroll
    for (int j = 0; j < kVec; j++) {
      const uint64_t rel = ((pv >> j) & 1u) ? (uint64_t)v[j] - (uint64_t)pt.base : 0ull;
      if (pt.span && rel >= pt.span && ((mask >> j) & 1u)) {  // span 0: all 2^64 values
        atomicOr(P.err, ERR_KEY_RANGE);
        mask &= ~(1u << j);
      }
      if (w == 0) continue;
      if (o >= 64) {
        hi[j] |= rel << (o - 64);
      } else {
        lo[j] |= rel << o;
        if (o && o + w > 64) hi[j] |= rel >> (64 - o);
      }
    }
  }
  return mask;
}

// 64-bit hash of a group key (lo, hi): hash partitions take their digits
// from its top 32 bits, leaf tables their first slot and tag from the rest.
__device__ __forceinline__ uint64_t rx_key_hash(uint64_t lo, uint64_t hi) {
  return mix64(lo ^ mix64(hi + 0x9e3779b97f4a7c15ull));
}

// Filter, group key (lo, and hi for a packed key) and the value the digits
// are taken from: the relative key itself, or with HASHED the top 32 bits
// of the key's hash.
template <bool HASHED>
__device__ __forceinline__ uint32_t rx_strip_route(const ScanParams& P, uint64_t r, uint32_t cnt, bool with_preds,
                                                   uint64_t key[kVec], uint64_t khi[kVec], uint64_t route[kVec]) {
  uint32_t mask;
  if (HASHED && P.g_mode == G_WIDE) {
    mask = rx_strip_wide(P, r, rx_strip_keys(P, r, cnt, key, with_preds, false), key, khi);
  } else {
    mask = rx_strip_keys(P, r, cnt, key, with_preds);
#pragma unroll
    for (int j = 0; j < kVec; j++) khi[j] = 0;
  }
#pragma unroll
  for (int j = 0; j < kVec; j++) route[j] = HASHED ? rx_key_hash(key[j], khi[j]) >> 32 : key[j];
  return mask;
}

// The tiles this workgroup scans (contiguous share of the tile list).
__device__ __forceinline__ void rx_tile_range(const ScanParams& P, uint32_t& t0, uint32_t& t1) {
  t0 = (uint32_t)(((uint64_t)blockIdx.x * P.n_tiles) / gridDim.x);
  t1 = (uint32_t)(((uint64_t)(blockIdx.x + 1) * P.n_tiles) / gridDim.x);
}

constexpr uint32_t kRxMaxBits1 = 10;  // level-1 fanout <= 1024

// Level-1 bucket starts from the histogram, by the whole workgroup (at most
// 1024 buckets, tmp: blockDim.x LDS words): bucket_start[d] = rows in buckets
// < d, bucket_fill[d] = bucket_start[d], bucket_start[2^bits1] = all rows.
__device__ __forceinline__ void rx_offsets(const RadixParams& R, const unsigned long long* hist, uint64_t* tmp) {
  const uint32_t nb = 1u << R.bits1;
  const uint32_t per = (nb + blockDim.x - 1) / blockDim.x;
  const uint32_t lo = threadIdx.x * per;
  uint64_t v[4] = {0, 0, 0, 0};  // per <= 4: nb <= 1024 over >= 256 threads
  uint64_t s = 0;
  for (uint32_t i = 0; i < per && i < 4; i++)
    if (lo + i < nb) {
      v[i] = hist[lo + i];
      s += v[i];
    }
  tmp[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t o = 1; o < blockDim.x; o <<= 1) {
    const uint64_t x = threadIdx.x >= o ? tmp[threadIdx.x - o] : 0;
    __syncthreads();
    tmp[threadIdx.x] += x;
    __syncthreads();
  }
  uint64_t run = tmp[threadIdx.x] - s;
  for (uint32_t i = 0; i < per && i < 4; i++)
    if (lo + i < nb) {
      R.bucket_start[lo + i] = run;
      R.bucket_fill[lo + i] = run;
      run += v[i];
    }
  if (threadIdx.x == blockDim.x - 1) R.bucket_start[nb] = tmp[blockDim.x - 1];
  __syncthreads();
}

// Per-item digit counts (batched radix): the rows of tiles [t0, t1) that
// pass the predicates, by level-1 digit, through `h` (2^bits1 LDS words) to
// `out`.  Counting after the filter makes the item's runs exact, so the
// scatter can write at precomputed offsets without gaps.
__device__ __forceinline__ void rx_item_hist(const ScanParams& P, const RadixParams& R, uint32_t t0, uint32_t t1,
                                             uint32_t* h, uint32_t* out) {
  const uint32_t nb = 1u << R.bits1;
  for (uint32_t d = threadIdx.x; d < nb; d += blockDim.x) h[d] = 0;
  __syncthreads();
  for (uint32_t t = t0; t < t1; t++) {
    const Tile T = P.tiles[t];
    for (uint32_t s = threadIdx.x * kVec; s < T.n; s += blockDim.x * kVec) {
      const uint32_t cnt = T.n - s < (uint32_t)kVec ? T.n - s : (uint32_t)kVec;
      uint64_t key[kVec];
      const uint32_t mask = rx_strip_keys(P, T.row0 + s, cnt, key, true);
#pragma unroll
      for (int j = 0; j < kVec; j++)
        if ((mask >> j) & 1u) atomicAdd(&h[(uint32_t)(key[j] >> R.shift1)], 1u);
    }
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < nb; d += blockDim.x) out[d] = h[d];
}

// From the per-item counts item_hist[i * 2^bits1 + d] of n_items items (one
// workgroup): the level-1 bucket ranges (bucket_start[d], bucket_fill[d] =
// its end) and, in place, each item's first row per digit: bucket start plus
// the rows of that digit in earlier items.  tmp: blockDim.x LDS words.
__device__ __forceinline__ void rx_item_offsets(const RadixParams& R, uint32_t* item_hist, uint32_t n_items,
                                                uint64_t* tmp) {
  const uint32_t nb = 1u << R.bits1;
  const uint32_t per = (nb + blockDim.x - 1) / blockDim.x;  // <= 4 for nb <= 1024 over >= 256 threads
  const uint32_t lo = threadIdx.x * per;
  uint64_t tot[4] = {0, 0, 0, 0};
  uint64_t s = 0;
  for (uint32_t k = 0; k < per && k < 4; k++) {
    if (lo + k >= nb) break;
    for (uint32_t i = 0; i < n_items; i++) tot[k] += item_hist[(uint64_t)i * nb + lo + k];
    s += tot[k];
  }
  tmp[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t o = 1; o < blockDim.x; o <<= 1) {
    const uint64_t x = threadIdx.x >= o ? tmp[threadIdx.x - o] : 0;
    __syncthreads();
    tmp[threadIdx.x] += x;
    __syncthreads();
  }
  uint64_t run = tmp[threadIdx.x] - s;
  for (uint32_t k = 0; k < per && k < 4; k++) {
    const uint32_t d = lo + k;
    if (d >= nb) break;
    R.bucket_start[d] = run;
    R.bucket_fill[d] = run + tot[k];
    uint64_t c = run;
    for (uint32_t i = 0; i < n_items; i++) {
      uint32_t* p = item_hist + (uint64_t)i * nb + d;
      const uint32_t v = *p;
      *p = (uint32_t)c;
      c += v;
    }
    run += tot[k];
  }
  if (threadIdx.x == blockDim.x - 1) R.bucket_start[nb] = tmp[blockDim.x - 1];
  __syncthreads();
}

// Exclusive prefix of one value per thread over the workgroup (blockDim.x a
// multiple of 64): wavefront scans by shuffles, one pass over the wavefront
// totals.  *total gets the sum; tmp: blockDim.x / 64 + 1 LDS words.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* tmp, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) tmp[wave] = x;
  __syncthreads();
  if (wave == 0) {
    const uint32_t t = lane < nw ? tmp[lane] : 0u;
    uint32_t u = t;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(u, o);
      if (lane >= (uint32_t)o) u += y;
    }
    if (lane < nw) tmp[lane] = u - t;
    if (lane == nw - 1) tmp[nw] = u;
  }
  __syncthreads();
  const uint32_t r = tmp[wave] + x - v;
  *total = tmp[nw];
  __syncthreads();  // tmp is free again
  return r;
}

// Exclusive prefix over nb (<= blockDim.x * per) LDS counters into `off`;
// returns the total.  tmp: blockDim.x / 64 + 1 LDS words.
__device__ __forceinline__ uint32_t rx_prefix(const uint32_t* cnt, uint32_t* off, uint32_t nb, uint32_t* tmp) {
  const uint32_t per = (nb + blockDim.x - 1) / blockDim.x;
  const uint32_t lo = threadIdx.x * per;
  uint32_t s = 0;
  for (uint32_t i = 0; i < per; i++)
    if (lo + i < nb) s += cnt[lo + i];
  uint32_t total;
  uint32_t run = block_excl_scan(s, tmp, &total);
  for (uint32_t i = 0; i < per; i++)
    if (lo + i < nb) {
      off[lo + i] = run;
      run += cnt[lo + i];
    }
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
static_assert(kRxStageOff == 24 * kRxMaxDigits, "digit tables end where the staged rows start");

inline size_t rx_scatter_lds_bytes(const RadixParams& r) { return kRxStageOff + (size_t)r.tile_rows * r.row.stride; }

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
__device__ __forceinline__ void rx_write_record(const ScanParams& P, uint8_t* lds, uint32_t s, GKey key,
                                                uint8_t* rec) {
  const uint32_t cap = P.lds_cap;
  if (P.g_mode != G_DENSE) *reinterpret_cast<uint64_t*>(rec + P.g_key_off) = key.lo;  // dense: key = index
  if (P.g_mode == G_WIDE) *reinterpret_cast<uint64_t*>(rec + P.g_key_off + 8) = key.hi;
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
      if (!d.lo_only) *reinterpret_cast<uint64_t*>(g + 8) = (uint64_t)(v >> 64);
    } else {
      *reinterpret_cast<uint64_t*>(g) = w;  // f64 bits, or (ordered) i64
    }
    if (d.nullable) *reinterpret_cast<uint64_t*>(rec + d.g_ns_off) = lds_at<uint32_t>(lds, d.lds_ns_off)[s];
  }
}

// Level-1 digit histogram of tiles [t0, t1) (every row, before filtering:
// the counts bound each bucket's size) through `h` (2^bits1 LDS counters)
// into `hist`.
template <bool HASHED = false>
__device__ __forceinline__ void rx_hist_range(const ScanParams& P, const RadixParams& R, uint32_t t0, uint32_t t1,
                                              uint32_t* h, unsigned long long* hist) {
  const uint32_t nb = 1u << R.bits1;
  for (uint32_t d = threadIdx.x; d < nb; d += blockDim.x) h[d] = 0;
  __syncthreads();
  for (uint32_t t = t0; t < t1; t++) {
    const Tile T = P.tiles[t];
    for (uint32_t s = threadIdx.x * kVec; s < T.n; s += blockDim.x * kVec) {
      const uint32_t cnt = T.n - s < (uint32_t)kVec ? T.n - s : (uint32_t)kVec;
      uint64_t key[kVec], khi[kVec], route[kVec];
      const uint32_t mask = rx_strip_route<HASHED>(P, T.row0 + s, cnt, false, key, khi, route);
#pragma unroll
      for (int j = 0; j < kVec; j++)
        if ((mask >> j) & 1u) atomicAdd(&h[(uint32_t)(route[j] >> R.shift1)], 1u);
    }
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < nb; d += blockDim.x)
    if (h[d]) atomicAdd(&hist[d], (unsigned long long)h[d]);
}

// Filters tiles [t0, t1) and moves each surviving row (key, validity word,
// native-width aggregate inputs) to its level-1 bucket: a tile of
// R.tile_rows rows (<= blockDim.x * kVec) is staged in LDS grouped by digit,
// and each digit's run is written contiguously after one device atomic on
// the bucket cursor.
//
// With `item_off` (the item's row offset per digit, from rx_item_offsets)
// the runs go to consecutive places from there, tracked in LDS, and no
// device atomic is needed.
template <bool HASHED = false>
__device__ __forceinline__ void rx_scatter_range(const ScanParams& P, const RadixParams& R, uint32_t t0, uint32_t t1,
                                                 uint8_t* lds, const uint32_t* item_off = nullptr) {
  uint32_t* cnt = reinterpret_cast<uint32_t*>(lds + kRxCntOff);
  uint32_t* off = reinterpret_cast<uint32_t*>(lds + kRxOffOff);
  uint32_t* cur = reinterpret_cast<uint32_t*>(lds + kRxCurOff);
  uint32_t* tmp = reinterpret_cast<uint32_t*>(lds + kRxTmpOff);
  uint64_t* base = reinterpret_cast<uint64_t*>(lds + kRxBaseOff);
  uint8_t* stage = lds + kRxStageOff;
  const uint32_t nb = 1u << R.bits1, stride = R.row.stride;
  if (item_off)  // written by another workgroup in the same launch: a device-coherent load
    for (uint32_t d = threadIdx.x; d < nb; d += blockDim.x)
      cur[d] = __hip_atomic_load(item_off + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (uint32_t t = t0; t < t1; t++) {
    const Tile T = P.tiles[t];
    for (uint32_t s0 = 0; s0 < T.n; s0 += R.tile_rows) {  // uniform sub-tile loop
      for (uint32_t d = threadIdx.x; d < nb; d += blockDim.x) cnt[d] = 0;
      __syncthreads();
      const uint32_t s = s0 + threadIdx.x * kVec;
      const uint64_t r = T.row0 + s;
      uint32_t mask = 0, digit[kVec], rank[kVec];
      uint64_t key[kVec], khi[kVec], route[kVec];
      if (threadIdx.x * kVec < R.tile_rows && s < T.n) {
        const uint32_t c = T.n - s < (uint32_t)kVec ? T.n - s : (uint32_t)kVec;
        mask = rx_strip_route<HASHED>(P, r, c, true, key, khi, route);
#pragma unroll
        for (int j = 0; j < kVec; j++)
          if ((mask >> j) & 1u) {
            digit[j] = (uint32_t)(route[j] >> R.shift1);
            rank[j] = atomicAdd(&cnt[digit[j]], 1u);
          }
      }
      __syncthreads();
      const uint32_t total = rx_prefix(cnt, off, nb, tmp);
      for (uint32_t d = threadIdx.x; d < nb; d += blockDim.x) {
        if (item_off) {  // this thread alone reads and advances cur[d]
          base[d] = cur[d];
          cur[d] += cnt[d];
        } else {
          base[d] = cnt[d] ? atomicAdd(reinterpret_cast<unsigned long long*>(&R.bucket_fill[d]),
                                       (unsigned long long)cnt[d])
                           : 0;
        }
      }
      if (mask) {
        // stage the rows: key, then each aggregate input (the loads hit the
        // cache lines the key pass just fetched), then the validity word
        uint8_t* row[kVec];
        uint32_t flags[kVec] = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < kVec; j++) {
          row[j] = stage + (size_t)(((mask >> j) & 1u) ? off[digit[j]] + rank[j] : 0) * stride;
          if ((mask >> j) & 1u) {
            *reinterpret_cast<uint32_t*>(row[j]) = (uint32_t)route[j];
            if (HASHED) {  // the key itself after the hash
              uint64_t* kp = reinterpret_cast<uint64_t*>(row[j] + R.row.key_off);
              kp[0] = key[j];
              if (R.row.key_words == 2) kp[1] = khi[j];
            }
          }
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
        const uint32_t route = *reinterpret_cast<const uint32_t*>(row);
        const uint32_t d = route >> R.shift1;
        const uint64_t at = base[d] + (i - off[d]);
        rx_copy_row(R.rows1 + at * stride, row, stride);
        if (R.digits1) R.digits1[at] = route;
      }
      __syncthreads();
    }
  }
}

// Output column c of LDS slot s (relative key `key`) as havingOutputNum
// reads it (src/exec/v2_shape_group_topn.zig:2461-2500): the same values
// rx_write_record would store and having_value would read back.
__device__ __forceinline__ HNum rx_having_value(const ScanParams& P, const HavingLeaf& L, uint32_t c, uint8_t* lds,
                                                uint32_t s, GKey key) {
  const HavingSrc& h = L.src[c];
  HNum v{0, 0.0, false, false};
  const uint32_t count = lds_at<uint32_t>(lds, P.lds_count_off)[s];
  switch (h.kind) {
    case O_KEY: {
      const PartDesc& pt = P.parts[h.part];
      if (part_null(pt, key)) v.null = true;
      else v.i = part_value(pt, key);
      return v;
    }
    case O_COUNT: v.i = count; return v;
    case O_NS: v.i = lds_at<uint32_t>(lds, P.aggs[h.agg].lds_off)[s]; return v;
    default: break;
  }
  const AggDesc& d = P.aggs[h.agg];
  const uint64_t ns = d.nullable ? lds_at<uint32_t>(lds, d.lds_ns_off)[s] : count;
  if (ns == 0) {
    v.null = true;
    return v;
  }
  const uint64_t w = lds_at<uint64_t>(lds, d.lds_off)[s];
  switch (h.kind) {
    case O_SUM_I64:
    case O_SUM_I128:
    case O_AVG_I: {
      __int128 x;
      if (d.split) {
        const int64_t hs = (int64_t)lds_at<uint64_t>(lds, d.lds_off)[P.lds_cap + s];
        x = (__int128)hs * ((__int128)1 << 32) + (__int128)w;
      } else {
        x = (__int128)(int64_t)w;
      }
      if (h.kind == O_SUM_I64) v.i = (int64_t)(uint64_t)x;  // the wrapped BIGINT
      else if (h.kind == O_SUM_I128) v.i = x;
      else { v.is_f = true; v.f = i128_to_f64(x) / (double)ns / h.scale_div; }
      break;
    }
    case O_SUM_F64: v.is_f = true; v.f = __longlong_as_double((long long)w); break;
    case O_AVG_F: v.is_f = true; v.f = __longlong_as_double((long long)w) / (double)ns; break;
    case O_MINMAX_I: v.i = (int64_t)w; break;
    default: v.is_f = true; v.f = unord_f64((int64_t)w); break;  // O_MINMAX_F
  }
  return v;
}

// havingPasses (:2502-2535) over LDS slot s.
__device__ __forceinline__ bool rx_having_pass(const ScanParams& P, const HavingLeaf& L, uint8_t* lds, uint32_t s,
                                               GKey key) {
  uint64_t st = 0;
  for (uint32_t k = 0; k < L.n; k++) {
    const HavingNode& h = L.nodes[k];
    bool r;
    switch (h.kind) {
      case H_CMP: {
        const HNum a = rx_having_value(P, L, h.target, lds, s, key);
        if (a.null) {
          r = false;
          break;
        }
        const int o = having_order(a, h);
        switch (h.op) {
          case 0: r = o == 0; break;
          case 1: r = o != 0; break;
          case 2: r = o < 0; break;
          case 3: r = o <= 0; break;
          case 4: r = o > 0; break;
          default: r = o >= 0; break;
        }
        break;
      }
      case H_IS_NULL:
      case H_IS_NOT_NULL: {
        const bool nul = rx_having_value(P, L, h.target, lds, s, key).null;
        r = h.kind == H_IS_NULL ? nul : !nul;
        break;
      }
      case H_NOT:
        st ^= 1ull;
        continue;
      default: {  // H_AND / H_OR
        const uint64_t m = h.arity >= 64 ? ~0ull : ((1ull << h.arity) - 1);
        const uint64_t top = st & m;
        st = h.arity >= 64 ? 0 : (st >> h.arity);
        r = h.kind == H_AND ? top == m : top != 0;
        break;
      }
    }
    st = (st << 1) | (r ? 1ull : 0ull);
  }
  return (st & 1ull) != 0;
}

// A written group's row count to R.counts_out[at] and its bin of chist
// (kRxCountBins LDS counters of min(count, kRxCountBins - 1)); every lane of
// the wavefront calls it (occ: this lane wrote a group).
__device__ __forceinline__ void rx_count_out(const RadixParams& R, uint64_t at, uint32_t c, bool occ,
                                             uint32_t* chist) {
  const uint32_t lane = threadIdx.x & 63;
  if (occ) R.counts_out[at] = c;
  const uint32_t b = c < kRxCountBins ? c : kRxCountBins - 1;
  // the common small counts: one LDS atomic per wavefront and value
  for (uint32_t v = 1; v <= 3; v++) {
    const uint64_t mv = __ballot(occ && b == v);
    if (lane == 0 && mv) atomicAdd(&chist[v], (uint32_t)__popcll(mv));
  }
  if (occ && b > 3) atomicAdd(&chist[b], 1u);
}

// A leaf's cut for ORDER BY COUNT(*) LIMIT K, by one wavefront (every lane
// gets it; hist: the leaf's groups by count bin): the bin at which the
// groups counted from the best end reach K (0: the leaf has at most K
// groups); all: the leaf's groups.
__device__ __forceinline__ void rx_leaf_cut(const RadixParams& R, const uint32_t* hist, uint32_t& cut,
                                            uint32_t& all) {
  static_assert(kRxCountBins == 256, "four bins per lane");
  const uint32_t lane = threadIdx.x & 63, K = R.leaf_topk;
  const bool desc = R.topk_desc != 0;
  uint32_t h[4];
#pragma unroll
  for (uint32_t j = 0; j < 4; j++) h[j] = hist[lane * 4 + j];
  const uint32_t s4 = h[0] + h[1] + h[2] + h[3];
  // inclusive sum over the lanes at least as good as this one (desc: the
  // lanes above, whose bins hold larger counts)
  uint32_t incl = s4;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = desc ? __shfl_down(incl, o) : __shfl_up(incl, o);
    if (desc ? lane + (uint32_t)o < 64 : lane >= (uint32_t)o) incl += y;
  }
  all = __shfl(incl, desc ? 0 : 63);
  uint32_t c = 0;
  if (all > K && incl - s4 < K && incl >= K) {  // the one lane where the count reaches K
    uint32_t cum = incl - s4;
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
      const uint32_t j = desc ? 3 - i : i;
      if (c == 0 && cum + h[j] >= K) c = lane * 4 + j;  // bin 0 holds no group: c != 0
      cum += h[j];
    }
  }
  const uint64_t m = __ballot(c != 0);
  cut = __shfl(c, m ? __ffsll((long long)m) - 1 : 0);
}

// The occupied slots of a leaf's LDS table (count != 0; with HAVING those
// passing it) as group records from row index s on, and their number to
// R.leaf_groups[leaf].  Positions p < n_pos name slot slot_of(p) (the table
// in slot order, or a list of slots); position k * blockDim.x + t is handled
// by thread t in round k, so each record store is coalesced; per (round,
// wavefront) counts are scanned in that order.  key_of(slot) gives a slot's
// group key.  tmp: blockDim.x / 64 + 1 + n_pos / 64 LDS words.  With
// R.counts_out each written group's row count goes out too (rx_count_out).
//
// With R.leaf_topk (lhist: kRxCountBins LDS words, zero on entry and on
// return) the leaf writes only the groups in count bins at least as good as
// its cut (the bin where its groups, counted from the best end, reach
// leaf_topk): every group of the query's top leaf_topk is among them, since
// fewer than leaf_topk of its own leaf's groups precede it.  The ORDER BY
// that follows sorts them by (count, relative key), so they go out in no
// particular order: one LDS atomic per wavefront places them.
template <bool HAVING, typename SlotOf, typename KeyOf>
__device__ __forceinline__ void rx_emit_leaf(const ScanParams& P, const RadixParams& R, uint32_t leaf, uint64_t s,
                                             uint32_t n_pos, uint8_t* lds, uint32_t* tmp, uint32_t* chist,
                                             SlotOf slot_of, KeyOf key_of, uint32_t* lhist = nullptr) {
  const uint32_t* count = lds_at<uint32_t>(lds, P.lds_count_off);
  const uint32_t nw = blockDim.x >> 6, wave = threadIdx.x >> 6;
  const uint32_t rounds = (n_pos + blockDim.x - 1) / blockDim.x;
  uint32_t* scan_tmp = tmp;
  uint32_t* wcnt = tmp + nw + 1;  // rounds * nw entries
  const HavingLeaf* H = HAVING ? R.having : nullptr;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t below = (1ull << lane) - 1ull;
  const bool topk = R.leaf_topk != 0 && lhist != nullptr && rounds <= 64;  // uniform
  auto bin_of = [&](uint32_t slot) { return count[slot] < kRxCountBins ? count[slot] : kRxCountBins - 1; };
  if (topk && threadIdx.x == 0) tmp[0] = 0;  // the leaf's written groups (ordered by the barrier below)
  uint64_t keep = 0;  // each round's outcome (rounds beyond 64 evaluate again below)
  for (uint32_t k = 0; k < rounds; k++) {
    const uint32_t p = k * blockDim.x + threadIdx.x;
    const uint32_t slot = p < n_pos ? slot_of(p) : 0u;
    bool occ = p < n_pos && count[slot] != 0;
    if (H && occ) occ = rx_having_pass(P, *H, lds, slot, key_of(slot));
    if (k < 64 && occ) keep |= 1ull << k;
    if (topk) {  // the leaf's groups by count bin (the common small counts: one LDS atomic per wavefront)
      const uint32_t b = occ ? bin_of(slot) : 0u;
      for (uint32_t v = 1; v <= 3; v++) {
        const uint64_t mv = __ballot(occ && b == v);
        if (lane == 0 && mv) atomicAdd(&lhist[v], (uint32_t)__popcll(mv));
      }
      if (occ && b > 3) atomicAdd(&lhist[b], 1u);
    } else {
      const uint64_t m = __ballot(occ);
      if (lane == 0) wcnt[k * nw + wave] = (uint32_t)__popcll(m);
    }
  }
  __syncthreads();
  if (topk) {
    uint32_t cut, all;
    rx_leaf_cut(R, lhist, cut, all);  // every wavefront, from the same histogram
    if (threadIdx.x == 0 && R.group_total)
      atomicAdd(&R.group_total[(leaf % kRxTotalSlots) * 8], (unsigned long long)all);
    for (uint32_t k = 0; k < rounds; k++) {
      const uint32_t slot = ((keep >> k) & 1ull) ? slot_of(k * blockDim.x + threadIdx.x) : 0u;
      const uint32_t b = ((keep >> k) & 1ull) ? bin_of(slot) : 0u;
      const bool take = ((keep >> k) & 1ull) && (cut == 0 || b == cut || (R.topk_desc ? b > cut : b < cut));
      const uint64_t m = __ballot(take);
      uint64_t at = 0;
      if (m) {  // uniform per wavefront
        const int first = __ffsll((long long)m) - 1;
        uint32_t base = 0;
        if ((int)lane == first) base = atomicAdd(&tmp[0], (uint32_t)__popcll(m));
        at = s + __shfl(base, first) + (uint32_t)__popcll(m & below);
        if (take) rx_write_record(P, lds, slot, key_of(slot), P.g_state + at * P.g_rec);
      }
      if (R.counts_out) rx_count_out(R, at, take ? count[slot] : 0u, take, chist);  // uniform
    }
    __syncthreads();
    if (threadIdx.x == 0) R.leaf_groups[leaf] = tmp[0];
    // every wavefront has its cut: clear the histogram for the next leaf
    // (whose first use is past that leaf's own barriers)
    for (uint32_t b = threadIdx.x; b < kRxCountBins; b += blockDim.x) lhist[b] = 0;
    __syncthreads();
    return;
  }
  uint32_t total = 0;
  for (uint32_t b = 0; b < rounds * nw; b += blockDim.x) {  // uniform; one pass while rounds * nw <= blockDim.x
    const uint32_t t = b + threadIdx.x;
    const uint32_t v = t < rounds * nw ? wcnt[t] : 0u;
    uint32_t part;
    const uint32_t ex = block_excl_scan(v, scan_tmp, &part);
    if (t < rounds * nw) wcnt[t] = total + ex;
    total += part;
  }
  __syncthreads();
  for (uint32_t k = 0; k < rounds; k++) {
    const uint32_t p = k * blockDim.x + threadIdx.x;
    const uint32_t slot = p < n_pos ? slot_of(p) : 0u;
    bool occ;
    if (k < 64) {
      occ = (keep >> k) & 1ull;
    } else {
      occ = p < n_pos && count[slot] != 0;
      if (H && occ) occ = rx_having_pass(P, *H, lds, slot, key_of(slot));
    }
    const uint64_t m = __ballot(occ);
    const uint64_t at = s + wcnt[k * nw + wave] + (uint32_t)__popcll(m & below);
    if (occ) rx_write_record(P, lds, slot, key_of(slot), P.g_state + at * P.g_rec);
    if (R.counts_out) rx_count_out(R, at, occ ? count[slot] : 0u, occ, chist);  // uniform
  }
  if (threadIdx.x == 0) R.leaf_groups[leaf] = total;
  __syncthreads();
}

// A leaf's row range in the final partitioned array.
__device__ __forceinline__ void rx_leaf_range(const RadixParams& R, uint32_t leaf, uint64_t& s, uint64_t& e) {
  if (R.bits2) {
    s = R.leaf_start[2 * (uint64_t)leaf];
    e = R.leaf_start[2 * (uint64_t)leaf + 1];
  } else {
    s = R.bucket_start[leaf];
    e = R.bucket_fill[leaf];
  }
}

constexpr uint32_t kRxLeafPer = 4;  // rows per thread whose keys a leaf pass loads at once

// Aggregates leaf `leaf` in an LDS direct table and writes its groups: with
// DENSE_OUT every key of the leaf's range below g_cap at record index key
// (empty keys get a zero count), else compacted (rx_emit_leaf; a leaf holds
// at least as many rows as groups, so leaves never overlap).  With HAVING
// (compacted only) a group is written, counted and histogrammed only when
// it passes R.having.  tmp: blockDim.x / 64 + 1 + cap / 64 LDS words
// (unused with DENSE_OUT).
// PER: rows per thread whose keys one memory round trip brings (a leaf of
// many rows wants more; more registers cost the leaf kernel occupancy).
template <bool DENSE_OUT, bool HAVING = false, uint32_t PER = kRxLeafPer>
__device__ __forceinline__ void rx_agg_leaf(const ScanParams& P, const RadixParams& R, uint32_t leaf, uint8_t* lds,
                                            uint32_t* tmp, uint32_t* chist = nullptr, uint32_t* lhist = nullptr) {
  const uint32_t cap = P.lds_cap, stride = R.row.stride;
  const uint64_t mask = (uint64_t)cap - 1u;
  uint64_t s, e;
  rx_leaf_range(R, leaf, s, e);
  const uint64_t leaf_key0 = (uint64_t)leaf << R.leaf_bits;
  if (s == e) {  // uniform
    if (DENSE_OUT) {
      for (uint32_t slot = threadIdx.x; slot < cap && leaf_key0 + slot < P.g_cap; slot += blockDim.x)
        *reinterpret_cast<uint64_t*>(P.g_state + (leaf_key0 + slot) * P.g_rec + P.g_count_off) = 0;
    } else if (threadIdx.x == 0) {
      R.leaf_groups[leaf] = 0;
    }
    return;
  }
  lds_init(P, lds);
  __syncthreads();
  // the keys of up to PER rows per thread are loaded first, so the
  // leaf's rows arrive in one memory round trip (their values, in the same
  // cache lines, are then cache hits)
  for (uint64_t b = s; b < e; b += (uint64_t)blockDim.x * PER) {  // uniform
    uint32_t key[PER];
#pragma unroll
    for (uint32_t k = 0; k < PER; k++) {
      const uint64_t i = b + threadIdx.x + (uint64_t)k * blockDim.x;
      key[k] = i < e ? *reinterpret_cast<const uint32_t*>(R.rows2 + i * stride) : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < PER; k++) {
      const uint64_t i = b + threadIdx.x + (uint64_t)k * blockDim.x;
      if (i < e) rx_fold_row(P, R, lds, R.rows2 + i * stride, (uint32_t)(key[k] & mask));
    }
  }
  __syncthreads();
  if (DENSE_OUT) {
    const uint32_t* count = lds_at<uint32_t>(lds, P.lds_count_off);
    for (uint32_t slot = threadIdx.x; slot < cap && leaf_key0 + slot < P.g_cap; slot += blockDim.x) {
      uint8_t* rec = P.g_state + (leaf_key0 + slot) * P.g_rec;
      if (count[slot]) rx_write_record(P, lds, slot, GKey{leaf_key0 + slot, 0}, rec);
      else *reinterpret_cast<uint64_t*>(rec + P.g_count_off) = 0;
    }
    __syncthreads();
    return;
  }
  rx_emit_leaf<HAVING>(P, R, leaf, s, cap, lds, tmp, chist, [](uint32_t p) { return p; },
                       [&](uint32_t slot) { return GKey{leaf_key0 + slot, 0}; }, lhist);
}

// Records a claimed slot in the leaf's list of claimed slots (list, *n
// entries): one LDS atomic per wavefront.  Every lane of the wavefront calls
// it (claimed: this lane claimed `slot`).
__device__ __forceinline__ void rx_list_claim(uint32_t* list, uint32_t* n, bool claimed, uint32_t slot) {
  const uint64_t m = __ballot(claimed);
  if (m == 0) return;
  const uint32_t lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)m) - 1;
  uint32_t base = 0;
  if ((int)lane == leader) base = atomicAdd(n, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  if (claimed) list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1))] = slot;
}

// Hash partitions, exact form: rows [s, e) of a leaf fold into the LDS hash
// table (u32 tags, 0 = free; each claimed slot's key words, and the list of
// claimed slots).  Rows are taken blockDim.x at a time, and each batch
// settles in rounds between barriers, so no thread ever waits on another: a
// row either claims a free slot by CAS on its tag (and writes the key), or
// stops at a slot holding its tag, whose key is compared after the barrier
// (the claimer's key is then in place); a mismatch probes on from the next
// slot in the next round.  A table with no free slot left for a row raises
// ERR_HASH_FULL (the host then partitions into more leaves).
__device__ __forceinline__ void rx_hashed_fold_exact(const ScanParams& P, const RadixParams& R, uint64_t s,
                                                     uint64_t e, uint8_t* lds, uint32_t* n_claimed) {
  const uint32_t cap = P.lds_cap, stride = R.row.stride, kw = R.row.key_words;
  uint32_t* tag = lds_at<uint32_t>(lds, R.tag_off);
  uint32_t* list = lds_at<uint32_t>(lds, R.list_off);
  uint64_t* keys = lds_at<uint64_t>(lds, P.lds_keys_off);
  for (uint64_t b = s; b < e; b += blockDim.x) {  // uniform
    const uint64_t i = b + threadIdx.x;
    bool pending = i < e;
    const uint8_t* row = R.rows2 + (pending ? i : s) * stride;
    uint64_t klo = 0, khi = 0;
    uint32_t slot = 0, t = 0, probes = 0;
    if (pending) {
      const uint64_t* kp = reinterpret_cast<const uint64_t*>(row + R.row.key_off);
      klo = kp[0];
      khi = kw == 2 ? kp[1] : 0ull;
      const uint64_t h = rx_key_hash(klo, khi);
      slot = __umulhi((uint32_t)h, cap);
      t = ((uint32_t)(h >> 16) & R.tag_mask) | 1u;
    }
    while (__syncthreads_or(pending)) {
      int cand = -1;
      bool claimed = false;
      if (pending) {
        for (;;) {
          uint32_t cur = __hip_atomic_load(tag + slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (cur == 0) {
            uint32_t expect = 0;
            if (__hip_atomic_compare_exchange_strong(tag + slot, &expect, t, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP)) {
              keys[(uint64_t)slot * kw] = klo;
              if (kw == 2) keys[(uint64_t)slot * kw + 1] = khi;
              rx_fold_row(P, R, lds, row, slot);
              pending = false;
              claimed = true;
              break;
            }
            cur = expect;
          }
          if (cur == t) {  // same tag: compare keys after the barrier
            cand = (int)slot;
            break;
          }
          slot = slot + 1 == cap ? 0 : slot + 1;
          if (++probes >= cap) {
            atomicOr(P.err, ERR_HASH_FULL);
            pending = false;
            break;
          }
        }
      }
      rx_list_claim(list, n_claimed, claimed, slot);
      __syncthreads();
      if (cand >= 0) {
        const uint64_t* kk = keys + (uint64_t)cand * kw;
        if (kk[0] == klo && (kw == 1 || kk[1] == khi)) {
          rx_fold_row(P, R, lds, row, (uint32_t)cand);
          pending = false;
        } else {
          slot = (uint32_t)cand + 1 == cap ? 0 : (uint32_t)cand + 1;
          if (++probes >= cap) {
            atomicOr(P.err, ERR_HASH_FULL);
            pending = false;
          }
        }
      }
    }
  }
}

// Clears the whole hash table: tags free, aggregate states initial, no
// claimed slots.
__device__ __forceinline__ void rx_hashed_clear(const ScanParams& P, const RadixParams& R, uint8_t* lds,
                                                uint32_t* n_claimed) {
  uint32_t* tag = lds_at<uint32_t>(lds, R.tag_off);
  for (uint32_t i = threadIdx.x; i < P.lds_cap; i += blockDim.x) tag[i] = 0;
  lds_init(P, lds);
  if (threadIdx.x == 0) *n_claimed = 0;
  __syncthreads();
}

// Hash partitions: leaf `leaf`'s rows fold into an LDS hash table, its groups
// go out as records from the leaf's first row on, and the slots it claimed
// are cleared again (the table starts clean: rx_hashed_clear once per
// workgroup), so a leaf costs in proportion to its rows and groups, not to
// the table's size.
//
// The fold is optimistic: each thread loads the keys of up to kRxHashPer
// rows at once (one memory round trip for most leaves), and each row folds
// straight into the slot it claims or the first slot holding its tag.
// After a barrier every row that followed a tag checks the slot's key; a
// mismatch (two keys of one 31-bit tag on one probe path: rare) voids the
// whole leaf, which is cleared and folded again in the exact form.  With
// R.hash_exact only the exact form runs.
//
// Without HAVING each claimed slot is a group: list position p goes to
// record s + p.  With HAVING, rx_emit_leaf compacts the passing ones.
constexpr uint32_t kRxHashPer = 4;

template <bool HAVING>
__device__ __forceinline__ void rx_agg_leaf_hashed(const ScanParams& P, const RadixParams& R, uint32_t leaf,
                                                   uint8_t* lds, uint32_t* tmp, uint32_t* chist,
                                                   uint32_t* n_claimed) {
  const uint32_t cap = P.lds_cap, stride = R.row.stride, kw = R.row.key_words;
  uint64_t s, e;
  rx_leaf_range(R, leaf, s, e);
  if (s == e) {  // uniform
    if (threadIdx.x == 0) R.leaf_groups[leaf] = 0;
    return;
  }
  uint32_t* tag = lds_at<uint32_t>(lds, R.tag_off);
  uint32_t* list = lds_at<uint32_t>(lds, R.list_off);
  uint64_t* keys = lds_at<uint64_t>(lds, P.lds_keys_off);
  bool exact = R.hash_exact != 0;
  for (uint64_t b = s; b < e && !exact; b += (uint64_t)blockDim.x * kRxHashPer) {  // uniform
    uint64_t klo[kRxHashPer], khi[kRxHashPer];
    uint32_t slot[kRxHashPer];
    uint32_t live = 0, follow = 0;
#pragma unroll
    for (uint32_t k = 0; k < kRxHashPer; k++) {
      const uint64_t i = b + threadIdx.x + (uint64_t)k * blockDim.x;
      klo[k] = khi[k] = 0;
      if (i < e) {
        const uint64_t* kp = reinterpret_cast<const uint64_t*>(R.rows2 + i * stride + R.row.key_off);
        live |= 1u << k;
        klo[k] = kp[0];
        if (kw == 2) khi[k] = kp[1];
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < kRxHashPer; k++) {
      bool claimed = false, placed = false;
      uint32_t sl = 0;
      if ((live >> k) & 1u) {
        const uint64_t h = rx_key_hash(klo[k], khi[k]);
        const uint32_t t = ((uint32_t)(h >> 16) & R.tag_mask) | 1u;
        uint32_t probes = 0;
        sl = __umulhi((uint32_t)h, cap);
        for (;;) {
          uint32_t cur = __hip_atomic_load(tag + sl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (cur == 0) {
            uint32_t expect = 0;
            if (__hip_atomic_compare_exchange_strong(tag + sl, &expect, t, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP)) {
              keys[(uint64_t)sl * kw] = klo[k];
              if (kw == 2) keys[(uint64_t)sl * kw + 1] = khi[k];
              claimed = placed = true;
              break;
            }
            cur = expect;
          }
          if (cur == t) {  // its key is compared after the barrier
            follow |= 1u << k;
            placed = true;
            break;
          }
          sl = sl + 1 == cap ? 0 : sl + 1;
          if (++probes >= cap) break;
        }
        if (placed) {
          const uint64_t i = b + threadIdx.x + (uint64_t)k * blockDim.x;
          rx_fold_row(P, R, lds, R.rows2 + i * stride, sl);
        } else {
          atomicOr(P.err, ERR_HASH_FULL);
        }
      }
      slot[k] = sl;
      rx_list_claim(list, n_claimed, claimed, sl);
    }
    __syncthreads();
    bool bad = false;
#pragma unroll
    for (uint32_t k = 0; k < kRxHashPer; k++) {
      if (!((follow >> k) & 1u)) continue;
      const uint64_t* kk = keys + (uint64_t)slot[k] * kw;
      if (kk[0] != klo[k] || (kw == 2 && kk[1] != khi[k])) bad = true;
    }
    if (__syncthreads_or(bad)) {  // a row went to another key's group: fold the leaf again, exactly
      rx_hashed_clear(P, R, lds, n_claimed);
      exact = true;
    }
  }
  if (exact) {
    rx_hashed_fold_exact(P, R, s, e, lds, n_claimed);
    __syncthreads();
  }
  const uint32_t n = *n_claimed;
  auto key_of = [&](uint32_t sl) {
    const uint64_t* kk = keys + (uint64_t)sl * kw;
    return GKey{kk[0], kw == 2 ? kk[1] : 0ull};
  };
  if (HAVING) {
    rx_emit_leaf<true>(P, R, leaf, s, n, lds, tmp, chist, [&](uint32_t p) { return list[p]; }, key_of);
  } else {
    const uint32_t* count = lds_at<uint32_t>(lds, P.lds_count_off);
    for (uint32_t p0 = 0; p0 < n; p0 += blockDim.x) {  // uniform
      const uint32_t p = p0 + threadIdx.x;
      const bool occ = p < n;
      const uint32_t sl = occ ? list[p] : 0u;
      if (occ) rx_write_record(P, lds, sl, key_of(sl), P.g_state + (s + p) * P.g_rec);
      if (R.counts_out) rx_count_out(R, s + p, occ ? count[sl] : 0u, occ, chist);
    }
    if (threadIdx.x == 0) R.leaf_groups[leaf] = n;
    __syncthreads();
  }
  // the claimed slots back to their initial state
  for (uint32_t p = threadIdx.x; p < n; p += blockDim.x) {
    const uint32_t sl = list[p];
    tag[sl] = 0;
    lds_init_slot(P, lds, sl);
  }
  if (threadIdx.x == 0) *n_claimed = 0;
  __syncthreads();
}

}  // namespace tdbg
