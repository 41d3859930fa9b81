// distinct.hip — COUNT(DISTINCT x) per group.
//
// The reference folds each row's (packed group key, value bits) composite
// into per-worker distinct sets and counts each group's members after the
// merge (foldDistinctKernel/foldDistinctValues/compositeOf,
// src/exec/v2_lowcard_group.zig:1348-1418, :1497-1499; partition merge
// :1538-1552).  Here one pass over the plan's tiles (the same strips,
// predicates and relative keys as the scan) inserts the composite
// key * span + (x - base) into an open-addressing set in HBM with one
// compare-and-swap; the insertion that finds the slot empty adds one to the
// group's counter, so every (group, value) pair is counted exactly once.
// NULL values are never inserted (a NULL is not a counted value).  A dense
// composite range (at most 2^33 pairs) sets one bit per pair in a bit array
// instead, counting the pair whose bit was clear.  Float
// values, and integer spans whose composite would not fit 63 bits, take the
// wide form: the set holds row indices of staged (group, value bits) entries.
#include "order_device.h"
#include "radix_device.h"

namespace tdbg {

// The one representative of a float's equality class, as f64 bits: -0 and
// +0 are one value and so is every NaN (canonicalFloatBits of the value
// widened to f64, src/types.zig:593-600, src/exec/silo_group_core.zig:8052).
__device__ __forceinline__ uint64_t canonical_bits(double v) {
  if (v != v) return 0x7ff8000000000000ull;
  if (v == 0.0) return 0ull;
  return (uint64_t)__double_as_longlong(v);
}

// Adds one to group `key`'s distinct count (record `rec` when TO_RECORDS).
template <bool LDS_COUNTS, bool TO_RECORDS>
__device__ __forceinline__ void distinct_count(const DistinctParams& D, uint32_t* lcnt, uint64_t key, uint8_t* rec) {
  if (LDS_COUNTS) atomicAdd(&lcnt[key], 1u);
  else if (TO_RECORDS) g_add_u64(rec + D.dslot_off, 1);
  else atomicAdd(&D.counts[key], 1u);
}

template <bool LDS_COUNTS, bool TO_RECORDS>
__device__ __forceinline__ void distinct_flush(const ScanParams& P, const DistinctParams& D, const uint32_t* lcnt,
                                               uint32_t n_keys) {
  if (!LDS_COUNTS) return;
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < n_keys; k += blockDim.x) {
    if (!lcnt[k]) continue;
    if (TO_RECORDS) g_add_u64(P.g_state + (uint64_t)k * P.g_rec + D.dslot_off, lcnt[k]);  // dense: index = key
    else atomicAdd(&D.counts[k], lcnt[k]);
  }
}

// LDS_COUNTS: a domain of at most kDistinctLdsKeys groups counts in LDS per
// workgroup and adds its totals once at the end (few groups with many
// distinct values would otherwise serialise on a handful of counters).
// TO_RECORDS: the plan's records are addressable by key (dense index or HBM
// hash table): the composite takes the record index instead of the relative
// key (so sparse key domains fit 63 bits) and counts go to the record.
template <bool LDS_COUNTS, bool TO_RECORDS>
__global__ void __launch_bounds__(256) distinct_insert_kernel(const ScanParams P, const DistinctParams D,
                                                              uint32_t n_keys) {
  __shared__ uint32_t lcnt[LDS_COUNTS ? kDistinctLdsKeys : 1];
  if (LDS_COUNTS) {
    for (uint32_t k = threadIdx.x; k < n_keys; k += blockDim.x) lcnt[k] = 0;
    __syncthreads();
  }
  for (uint32_t t = blockIdx.x; t < P.n_tiles; t += gridDim.x) {
    const Tile T = P.tiles[t];
    for (uint32_t s = threadIdx.x * kVec; s < T.n; s += blockDim.x * kVec) {
      const uint32_t cnt = T.n - s < (uint32_t)kVec ? T.n - s : (uint32_t)kVec;
      const uint64_t r = T.row0 + s;
      uint64_t key[kVec];
      uint32_t mask = rx_strip_keys(P, r, cnt, key, true) & valid4(D.x, r);
      if (!mask) continue;
      int64_t v[kVec];
      decode_i(load_raw(D.x, r), D.x.kind, v);
#pragma unroll
      for (int j = 0; j < kVec; j++) {
        if (!((mask >> j) & 1u)) continue;
        const uint64_t rel = (uint64_t)v[j] - (uint64_t)D.base;
        if (rel >= D.span) {  // outside the zone-map statistics
          atomicOr(P.err, ERR_KEY_RANGE);
          continue;
        }
        uint8_t* rec = nullptr;
        uint64_t group = key[j];
        if (TO_RECORDS) {  // the scan made the group's record: this finds it
          rec = g_record(P, key[j]);
          if (!rec) continue;
          group = (uint64_t)(rec - P.g_state) / P.g_rec;
        }
        const unsigned long long c = group * D.span + rel;
        if (D.bitmap) {  // a dense composite range: one bit per pair, set once
          const uint32_t bit = 1u << (c & 31u);
          if (!(atomicOr(reinterpret_cast<uint32_t*>(D.set) + (c >> 5), bit) & bit))
            distinct_count<LDS_COUNTS, TO_RECORDS>(D, lcnt, key[j], rec);
          continue;
        }
        // at least twice as many slots as rows scanned: an empty slot ends
        // every probe
        uint64_t h = mix64(c) & D.set_mask;
        for (;;) {
          const unsigned long long prev = atomicCAS(&D.set[h], ~0ull, c);
          if (prev == c) break;
          if (prev == ~0ull) {
            distinct_count<LDS_COUNTS, TO_RECORDS>(D, lcnt, key[j], rec);
            break;
          }
          h = (h + 1) & D.set_mask;
        }
      }
    }
  }
  distinct_flush<LDS_COUNTS, TO_RECORDS>(P, D, lcnt, n_keys);
}

// Wide form (float values, or composites that do not fit 63 bits), first
// pass: every scanned row writes its entry (group, value bits) at its row
// index, or group ~0 when the row does not count (filtered out, NULL value,
// no record).
template <bool TO_RECORDS>
__global__ void __launch_bounds__(256) distinct_stage_kernel(const ScanParams P, const DistinctParams D) {
  for (uint32_t t = blockIdx.x; t < P.n_tiles; t += gridDim.x) {
    const Tile T = P.tiles[t];
    for (uint32_t s = threadIdx.x * kVec; s < T.n; s += blockDim.x * kVec) {
      const uint32_t cnt = T.n - s < (uint32_t)kVec ? T.n - s : (uint32_t)kVec;
      const uint64_t r = T.row0 + s;
      uint64_t key[kVec];
      const uint32_t mask = rx_strip_keys(P, r, cnt, key, true) & valid4(D.x, r);
      uint64_t bits[kVec];
      if (D.is_float) {
        double f[kVec];
        decode_f(load_raw(D.x, r), D.x.kind, f);
#pragma unroll
        for (int j = 0; j < kVec; j++) bits[j] = canonical_bits(f[j]);
      } else {
        int64_t v[kVec];
        decode_i(load_raw(D.x, r), D.x.kind, v);
#pragma unroll
        for (int j = 0; j < kVec; j++) bits[j] = (uint64_t)v[j];
      }
#pragma unroll
      for (int j = 0; j < kVec; j++) {
        if ((uint32_t)j >= cnt) break;
        uint64_t group = ~0ull;
        if ((mask >> j) & 1u) {
          group = key[j];
          if (TO_RECORDS) {
            const uint8_t* rec = g_record(P, key[j]);
            group = rec ? (uint64_t)(rec - P.g_state) / P.g_rec : ~0ull;
          }
        }
        D.entries[2 * (r + j)] = group;
        D.entries[2 * (r + j) + 1] = bits[j];
      }
    }
  }
}

// Wide form, second pass: a slot holds the row index of the entry that
// claimed it (every entry was written by the first pass, so a claimed slot's
// entry is readable); a row whose (group, bits) pair no earlier claim holds
// counts once for its group.
template <bool LDS_COUNTS, bool TO_RECORDS>
__global__ void __launch_bounds__(256) distinct_wide_insert_kernel(const ScanParams P, const DistinctParams D,
                                                                   uint32_t n_keys) {
  __shared__ uint32_t lcnt[LDS_COUNTS ? kDistinctLdsKeys : 1];
  if (LDS_COUNTS) {
    for (uint32_t k = threadIdx.x; k < n_keys; k += blockDim.x) lcnt[k] = 0;
    __syncthreads();
  }
  const ulonglong2* __restrict__ ent = reinterpret_cast<const ulonglong2*>(D.entries);
  for (uint32_t t = blockIdx.x; t < P.n_tiles; t += gridDim.x) {
    const Tile T = P.tiles[t];
    for (uint32_t i = threadIdx.x; i < T.n; i += blockDim.x) {
      const uint64_t r = T.row0 + i;
      const ulonglong2 e = ent[r];
      if (e.x == ~0ull) continue;
      // at least twice as many slots as rows scanned: an empty slot ends
      // every probe
      uint64_t h = mix64(mix64(e.x) ^ e.y) & D.set_mask;
      for (;;) {
        const unsigned long long prev = atomicCAS(&D.set[h], ~0ull, (unsigned long long)r);
        if (prev == ~0ull) {  // group = record index (TO_RECORDS) or the relative key
          uint8_t* rec = TO_RECORDS ? P.g_state + e.x * P.g_rec : nullptr;
          distinct_count<LDS_COUNTS, TO_RECORDS>(D, lcnt, e.x, rec);
          break;
        }
        const ulonglong2 o = ent[prev];
        if (o.x == e.x && o.y == e.y) break;
        h = (h + 1) & D.set_mask;
      }
    }
  }
  distinct_flush<LDS_COUNTS, TO_RECORDS>(P, D, lcnt, n_keys);
}

// Each compacted group's count from its relative key (the dense index, or
// the key word of a keyed record); radix plans only (their records are not
// addressable by key while the set is built).
__global__ void __launch_bounds__(256) distinct_gather_kernel(const FinalParams F, const DistinctParams D,
                                                              const uint32_t* __restrict__ gids,
                                                              const uint64_t* __restrict__ n_groups, uint64_t max_n) {
  const uint64_t n = *n_groups < max_n ? *n_groups : max_n;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t gid = gids[i];
    uint8_t* rec = const_cast<uint8_t*>(F.g_state) + (uint64_t)gid * F.g_rec;
    const uint64_t key = group_key(rec, gid, F.g_mode, F.g_key_off).lo;  // never G_WIDE
    *reinterpret_cast<uint64_t*>(rec + D.dslot_off) = D.counts[key];
  }
}

hipError_t launch_distinct_insert(const ScanParams& sp, const DistinctParams& D, uint64_t n_keys, uint32_t grid,
                                  hipStream_t s) {
  if (sp.n_tiles == 0) return hipSuccess;
  if (grid == 0 || !D.set || (!D.to_records && !D.counts) || (!D.wide && D.span == 0) || (D.wide && !D.entries) ||
      n_keys == 0)
    return hipErrorInvalidValue;
  // LDS counters flushed to records need the dense layout (record index = key)
  const bool lds = n_keys <= kDistinctLdsKeys && (!D.to_records || sp.g_mode == G_DENSE);
  const uint32_t nk = lds ? (uint32_t)n_keys : 0u;
  if (D.wide) {
    if (D.to_records) hipLaunchKernelGGL((distinct_stage_kernel<true>), dim3(grid), dim3(256), 0, s, sp, D);
    else hipLaunchKernelGGL((distinct_stage_kernel<false>), dim3(grid), dim3(256), 0, s, sp, D);
  }
#define TDBG_DISTINCT_LAUNCH(K, L, T) hipLaunchKernelGGL((K<L, T>), dim3(grid), dim3(256), 0, s, sp, D, nk)
  const int sel = (D.wide ? 4 : 0) | (lds ? 2 : 0) | (D.to_records ? 1 : 0);
  switch (sel) {
    case 0: TDBG_DISTINCT_LAUNCH(distinct_insert_kernel, false, false); break;
    case 1: TDBG_DISTINCT_LAUNCH(distinct_insert_kernel, false, true); break;
    case 2: TDBG_DISTINCT_LAUNCH(distinct_insert_kernel, true, false); break;
    case 3: TDBG_DISTINCT_LAUNCH(distinct_insert_kernel, true, true); break;
    case 4: TDBG_DISTINCT_LAUNCH(distinct_wide_insert_kernel, false, false); break;
    case 5: TDBG_DISTINCT_LAUNCH(distinct_wide_insert_kernel, false, true); break;
    case 6: TDBG_DISTINCT_LAUNCH(distinct_wide_insert_kernel, true, false); break;
    default: TDBG_DISTINCT_LAUNCH(distinct_wide_insert_kernel, true, true); break;
  }
#undef TDBG_DISTINCT_LAUNCH
  return hipGetLastError();
}

hipError_t launch_distinct_gather(const FinalParams& fp, const DistinctParams& D, const uint32_t* gids,
                                  const uint64_t* n_groups, uint64_t max_n, hipStream_t s) {
  if (max_n == 0) return hipSuccess;
  const uint64_t g = (max_n + 255) / 256;
  hipLaunchKernelGGL(distinct_gather_kernel, dim3((uint32_t)(g < 8192 ? g : 8192)), dim3(256), 0, s, fp, D, gids,
                     n_groups, max_n);
  return hipGetLastError();
}

}  // namespace tdbg
