// scan_common.h — device code of the fused scan -> filter -> GROUP BY
// aggregate, shared by the ahead-of-time generic kernel (scan_agg.hip) and
// the plan-specialised kernels compiled at run time (jit.cpp).
//
// Replaces the CPU fold of thinDB's grouped operators: the scan/filter of
// src/exec/scan.zig:4663-4800 (filterRowGroup), the key packing of
// src/exec/v2_lowcard_group.zig:886-948, the gid resolution of :1054-1118 and
// the fold kernels of :1206-1347 (silo equivalents:
// src/exec/silo_group_core.zig:7290-8003).
//
// Structure of one launch:
//  * Each thread owns kVec=4 consecutive rows per strip and requests every
//    column of the strip at its start (one or two 16-byte loads per column),
//    so a wavefront reads 1-2 KiB contiguous per column and keeps all of a
//    strip's loads in flight.
//  * Rows are filtered, their keys folded into a relative mixed-radix index,
//    and aggregated into an LDS table with LDS atomics: direct-indexed (the
//    key span split over `lds_parts` workgroups that scan the same tiles), or
//    open addressing (a row that finds no slot spills to HBM).  At the end a
//    workgroup flushes each occupied slot to its HBM record once, with
//    64-bit atomics (i128 sums with carry).
//  * Without an LDS table (very high cardinality) rows update HBM records
//    directly.
// Integer sums are exact: LDS partials are i64 when rows-per-workgroup times
// max|v| stays below 2^62, else split into 32-bit halves; HBM totals are
// i128.  Float sums use f64 atomics (order-dependent rounding, like the
// reference's parallel merge).
#pragma once

#include "engine.h"

namespace tdbg {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint32_t hash32(uint64_t k) { return (uint32_t)(mix64(k) >> 32); }

// Order-preserving map of an f64 onto i64 (signed compare).
__device__ __forceinline__ int64_t f64_ord(double x) {
  const int64_t s = __double_as_longlong(x);
  return s >= 0 ? s : (s ^ 0x7fffffffffffffffll);
}

// ---- column access ---------------------------------------------------------------

// The raw bytes of one column for a thread's kVec rows (at most 32), as four
// 64-bit words: 64-bit kinds one row per word, narrower kinds packed
// little-endian from w0.
struct Raw {
  uint64_t w0, w1, w2, w3;
};

template <uint32_t K>
__device__ __forceinline__ Raw ld(const ColDesc& c, uint64_t r) {
  Raw x = {0, 0, 0, 0};
  if (K == K_I64 || K == K_F64) {
    const ulonglong2* p = reinterpret_cast<const ulonglong2*>(static_cast<const uint64_t*>(c.data) + r);
    const ulonglong2 a = p[0], b = p[1];
    x.w0 = a.x; x.w1 = a.y; x.w2 = b.x; x.w3 = b.y;
  } else if (K == K_I32 || K == K_F32) {
    const ulonglong2 a = *reinterpret_cast<const ulonglong2*>(static_cast<const uint32_t*>(c.data) + r);
    x.w0 = a.x; x.w1 = a.y;
  } else if (K == K_I16) {
    x.w0 = *reinterpret_cast<const uint64_t*>(static_cast<const uint16_t*>(c.data) + r);
  } else {
    x.w0 = *reinterpret_cast<const uint32_t*>(static_cast<const uint8_t*>(c.data) + r);
  }
  return x;
}

__device__ __forceinline__ Raw load_raw(const ColDesc& c, uint64_t r) {
  switch (c.kind) {
    case K_I64: case K_F64: return ld<K_I64>(c, r);
    case K_I32: case K_F32: return ld<K_I32>(c, r);
    case K_I16: return ld<K_I16>(c, r);
    default: return ld<K_U8>(c, r);
  }
}

template <uint32_t K>
__device__ __forceinline__ void dec_i(const Raw& x, int64_t v[kVec]) {
  if (K == K_I64) {
    v[0] = (int64_t)x.w0; v[1] = (int64_t)x.w1; v[2] = (int64_t)x.w2; v[3] = (int64_t)x.w3;
  } else if (K == K_I32) {
    v[0] = (int32_t)(uint32_t)x.w0; v[1] = (int32_t)(uint32_t)(x.w0 >> 32);
    v[2] = (int32_t)(uint32_t)x.w1; v[3] = (int32_t)(uint32_t)(x.w1 >> 32);
  } else if (K == K_I16) {
    v[0] = (int16_t)(uint16_t)x.w0; v[1] = (int16_t)(uint16_t)(x.w0 >> 16);
    v[2] = (int16_t)(uint16_t)(x.w0 >> 32); v[3] = (int16_t)(uint16_t)(x.w0 >> 48);
  } else if (K == K_I8) {
    v[0] = (int8_t)(uint8_t)x.w0; v[1] = (int8_t)(uint8_t)(x.w0 >> 8);
    v[2] = (int8_t)(uint8_t)(x.w0 >> 16); v[3] = (int8_t)(uint8_t)(x.w0 >> 24);
  } else {  // K_U8
    v[0] = (uint8_t)x.w0; v[1] = (uint8_t)(x.w0 >> 8); v[2] = (uint8_t)(x.w0 >> 16); v[3] = (uint8_t)(x.w0 >> 24);
  }
}

template <uint32_t K>
__device__ __forceinline__ void dec_f(const Raw& x, double v[kVec]) {
  if (K == K_F64) {
    v[0] = __longlong_as_double((int64_t)x.w0); v[1] = __longlong_as_double((int64_t)x.w1);
    v[2] = __longlong_as_double((int64_t)x.w2); v[3] = __longlong_as_double((int64_t)x.w3);
  } else {
    v[0] = __uint_as_float((uint32_t)x.w0); v[1] = __uint_as_float((uint32_t)(x.w0 >> 32));
    v[2] = __uint_as_float((uint32_t)x.w1); v[3] = __uint_as_float((uint32_t)(x.w1 >> 32));
  }
}

__device__ __forceinline__ void decode_i(const Raw& x, uint32_t kind, int64_t v[kVec]) {
  switch (kind) {
    case K_I64: dec_i<K_I64>(x, v); break;
    case K_I32: dec_i<K_I32>(x, v); break;
    case K_I16: dec_i<K_I16>(x, v); break;
    case K_I8: dec_i<K_I8>(x, v); break;
    default: dec_i<K_U8>(x, v); break;
  }
}

__device__ __forceinline__ void decode_f(const Raw& x, uint32_t kind, double v[kVec]) {
  if (kind == K_F64) dec_f<K_F64>(x, v);
  else dec_f<K_F32>(x, v);
}

// Validity bits of the strip's rows (bit j = row r+j is non-NULL).
__device__ __forceinline__ uint32_t valid4(const ColDesc& c, uint64_t r) {
  if (!c.valid) return 0xfu;
  return (c.valid[r >> 5] >> (r & 31)) & 0xfu;
}

// ---- predicates --------------------------------------------------------------------

template <uint32_t PK, uint32_t K>
__device__ __forceinline__ uint32_t pred_k(const PredDesc& p, const Raw& x) {
  uint32_t m = 0;
  if (PK == P_RANGE_I || PK == P_NEQ_I) {
    int64_t v[kVec];
    dec_i<K>(x, v);
#pragma unroll
    for (int j = 0; j < kVec; j++) {
      const bool ok = PK == P_RANGE_I ? (v[j] >= p.lo_i && v[j] <= p.hi_i) : (v[j] != p.lo_i);
      m |= (uint32_t)ok << j;
    }
  } else {
    double v[kVec];
    dec_f<K>(x, v);
#pragma unroll
    for (int j = 0; j < kVec; j++) {
      bool ok;
      if (PK == P_RANGE_F) {
        const bool lo = p.lo_incl ? v[j] >= p.lo_f : v[j] > p.lo_f;
        const bool hi = p.hi_incl ? v[j] <= p.hi_f : v[j] < p.hi_f;
        ok = lo && hi;
      } else {
        ok = v[j] != p.lo_f;
      }
      m |= (uint32_t)ok << j;
    }
  }
  return m;
}

__device__ __forceinline__ uint32_t pred_rt(const PredDesc& p, const Raw& x, uint32_t kind) {
  switch (p.kind) {
    case P_RANGE_I:
      switch (kind) {
        case K_I64: return pred_k<P_RANGE_I, K_I64>(p, x);
        case K_I32: return pred_k<P_RANGE_I, K_I32>(p, x);
        case K_I16: return pred_k<P_RANGE_I, K_I16>(p, x);
        case K_I8: return pred_k<P_RANGE_I, K_I8>(p, x);
        default: return pred_k<P_RANGE_I, K_U8>(p, x);
      }
    case P_NEQ_I: {
      int64_t v[kVec];
      decode_i(x, kind, v);
      uint32_t m = 0;
#pragma unroll
      for (int j = 0; j < kVec; j++) m |= (uint32_t)(v[j] != p.lo_i) << j;
      return m;
    }
    case P_RANGE_F: return kind == K_F64 ? pred_k<P_RANGE_F, K_F64>(p, x) : pred_k<P_RANGE_F, K_F32>(p, x);
    default: return kind == K_F64 ? pred_k<P_NEQ_F, K_F64>(p, x) : pred_k<P_NEQ_F, K_F32>(p, x);
  }
}

// The WHERE tree over one strip (evaluatePredicate,
// src/exec/predicate.zig:1727-1790): a comparison with a NULL value does not
// match, IS [NOT] NULL read the validity, NOT inverts its child, AND / OR
// combine (arity 0: all / none).  `fetch(col)` gives column col's strip.
// The stack holds one 4-bit strip mask per pending node (at most kMaxFilter).
template <typename Fetch>
__device__ __forceinline__ uint32_t filter_mask(const ScanParams& P, uint64_t r, Fetch fetch) {
  uint64_t st = 0;
  for (uint32_t k = 0; k < P.n_filt; k++) {
    const FilterNode& f = P.filt[k];
    uint32_t m;
    switch (f.kind) {
      case H_CMP: {
        const ColDesc& c = P.cols[f.pd.col];
        m = pred_rt(f.pd, fetch(f.pd.col), c.kind) & valid4(c, r);
        break;
      }
      case H_IS_NULL: m = ~valid4(P.cols[f.pd.col], r) & 0xfu; break;
      case H_IS_NOT_NULL: m = valid4(P.cols[f.pd.col], r); break;
      case H_NOT: st ^= 0xfull; continue;
      default: {  // H_AND / H_OR
        const bool all = f.kind == H_AND;
        m = all ? 0xfu : 0u;
        for (uint32_t a = 0; a < f.arity; a++) {
          const uint32_t x = (uint32_t)(st & 0xfull);
          st >>= 4;
          m = all ? (m & x) : (m | x);
        }
        break;
      }
    }
    st = (st << 4) | m;
  }
  return (uint32_t)(st & 0xfull);
}

// ---- HBM records --------------------------------------------------------------------

__device__ __forceinline__ void g_add_u64(uint8_t* p, uint64_t v) {
  __hip_atomic_fetch_add(reinterpret_cast<uint64_t*>(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exact i128 += (hi:lo) with two 64-bit atomics: the low word's carry goes
// into the high word.  Commutative, so concurrent adds compose exactly.
__device__ __forceinline__ void g_add_i128(uint8_t* p, uint64_t lo, uint64_t hi) {
  if (lo) {
    const uint64_t old = __hip_atomic_fetch_add(reinterpret_cast<uint64_t*>(p), lo, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
    hi += (old + lo < old) ? 1u : 0u;
  }
  if (hi) g_add_u64(p + 8, hi);
}

__device__ __forceinline__ void g_add_f64(uint8_t* p, double v) {
  __hip_atomic_fetch_add(reinterpret_cast<double*>(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void g_min_i64(uint8_t* p, int64_t v) {
  __hip_atomic_fetch_min(reinterpret_cast<int64_t*>(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void g_max_i64(uint8_t* p, int64_t v) {
  __hip_atomic_fetch_max(reinterpret_cast<int64_t*>(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The HBM record of a relative key: dense index, or open addressing with a
// CAS-claimed u64 key word (kEmptyKey = free).  Null when the table is full.
__device__ __forceinline__ uint8_t* g_record(const ScanParams& P, uint64_t key) {
  if (P.g_mode == G_DENSE) return P.g_state + key * P.g_rec;
  const uint64_t mask = P.g_cap - 1;
  uint64_t s = mix64(key) & mask;
#pragma nounroll
  for (uint32_t probe = 0; probe < 65536u; probe++) {
    uint8_t* rec = P.g_state + s * P.g_rec;
    uint64_t* kp = reinterpret_cast<uint64_t*>(rec + P.g_key_off);
    const uint64_t cur = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == key) return rec;
    if (cur == kEmptyKey) {
      uint64_t expect = kEmptyKey;
      if (__hip_atomic_compare_exchange_strong(kp, &expect, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT))
        return rec;
      if (expect == key) return rec;
    }
    s = (s + 1) & mask;
  }
  atomicOr(P.err, ERR_HASH_FULL);
  return nullptr;
}

// One row's contribution straight into its HBM record.
template <uint32_t AK, bool NULLABLE>
__device__ __forceinline__ void g_apply(const AggDesc& a, uint8_t* rec, int64_t iv, double fv) {
  if (AK == A_COUNT_COL) {
    g_add_u64(rec + a.g_off, 1);
    return;
  }
  if (AK == A_SUM_I) {
    if (a.lo_only) g_add_u64(rec + a.g_off, (uint64_t)iv);
    else g_add_i128(rec + a.g_off, (uint64_t)iv, iv < 0 ? ~0ull : 0ull);
  }
  if (AK == A_SUM_F) g_add_f64(rec + a.g_off, fv);
  if (AK == A_MIN_I) g_min_i64(rec + a.g_off, iv);
  if (AK == A_MAX_I) g_max_i64(rec + a.g_off, iv);
  if (AK == A_MIN_F) g_min_i64(rec + a.g_off, f64_ord(fv));
  if (AK == A_MAX_F) g_max_i64(rec + a.g_off, f64_ord(fv));
  if (NULLABLE) g_add_u64(rec + a.g_ns_off, 1);
}

// ---- LDS table -----------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ T* lds_at(uint8_t* lds, uint32_t off) {
  return reinterpret_cast<T*>(lds + off);
}

__device__ __forceinline__ void l_add_u32(uint32_t* p, uint32_t v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void l_add_u64(uint64_t* p, uint64_t v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <uint32_t AK, bool NULLABLE, bool SPLIT>
__device__ __forceinline__ void l_apply(const AggDesc& a, uint8_t* lds, uint32_t cap, uint32_t slot, int64_t iv,
                                        double fv) {
  if (AK == A_COUNT_COL) {
    l_add_u32(lds_at<uint32_t>(lds, a.lds_off) + slot, 1);
    return;
  }
  if (AK == A_SUM_I) {
    if (SPLIT) {
      l_add_u64(lds_at<uint64_t>(lds, a.lds_off) + slot, (uint64_t)(uint32_t)iv);
      l_add_u64(lds_at<uint64_t>(lds, a.lds_off) + cap + slot, (uint64_t)(int64_t)(int32_t)(iv >> 32));
    } else {
      l_add_u64(lds_at<uint64_t>(lds, a.lds_off) + slot, (uint64_t)iv);
    }
  }
  if (AK == A_SUM_F)
    __hip_atomic_fetch_add(lds_at<double>(lds, a.lds_off) + slot, fv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  if (AK == A_MIN_I)
    __hip_atomic_fetch_min(lds_at<int64_t>(lds, a.lds_off) + slot, iv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  if (AK == A_MAX_I)
    __hip_atomic_fetch_max(lds_at<int64_t>(lds, a.lds_off) + slot, iv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  if (AK == A_MIN_F)
    __hip_atomic_fetch_min(lds_at<int64_t>(lds, a.lds_off) + slot, f64_ord(fv), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
  if (AK == A_MAX_F)
    __hip_atomic_fetch_max(lds_at<int64_t>(lds, a.lds_off) + slot, f64_ord(fv), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
  if (NULLABLE) l_add_u32(lds_at<uint32_t>(lds, a.lds_ns_off) + slot, 1);
}

// LDS slots of the strip's rows, probing all rows together so their LDS
// round trips overlap; -1 = not found within kLdsProbe probes.
template <typename K>
__device__ __forceinline__ void lds_find4(K* keys, uint32_t cap, const uint64_t key[kVec], uint32_t mask,
                                          int slot[kVec]) {
  const K empty = (K)kEmptyKey;
  uint32_t s[kVec];
  uint32_t todo = mask;
#pragma unroll
  for (int j = 0; j < kVec; j++) {
    slot[j] = -1;
    s[j] = __umulhi(hash32(key[j]), cap);
  }
#pragma nounroll
  for (int probe = 0; probe < kLdsProbe && todo; probe++) {
    K cur[kVec];
#pragma unroll
    for (int j = 0; j < kVec; j++)
      cur[j] = ((todo >> j) & 1u) ? __hip_atomic_load(keys + s[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                                  : (K)0;
#pragma unroll
    for (int j = 0; j < kVec; j++) {
      if (!((todo >> j) & 1u)) continue;
      const K k = (K)key[j];
      bool hit = cur[j] == k;
      if (!hit && cur[j] == empty) {
        K expect = empty;
        hit = __hip_atomic_compare_exchange_strong(keys + s[j], &expect, k, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP) ||
              expect == k;
      }
      if (hit) {
        slot[j] = (int)s[j];
        todo &= ~(1u << j);
      } else {
        s[j] = s[j] + 1 == cap ? 0 : s[j] + 1;
      }
    }
  }
}

__device__ __forceinline__ void lds_init(const ScanParams& P, uint8_t* lds) {
  const uint32_t cap = P.lds_cap, tid = threadIdx.x, nt = blockDim.x;
  if (P.lds_mode == LDS_HASH) {
    if (P.lds_key64) {
      for (uint32_t i = tid; i < cap; i += nt) lds_at<uint64_t>(lds, P.lds_keys_off)[i] = kEmptyKey;
    } else {
      for (uint32_t i = tid; i < cap; i += nt) lds_at<uint32_t>(lds, P.lds_keys_off)[i] = 0xffffffffu;
    }
  }
  for (uint32_t i = tid; i < cap; i += nt) lds_at<uint32_t>(lds, P.lds_count_off)[i] = 0;
  for (uint32_t a = 0; a < P.n_aggs; a++) {
    const AggDesc& d = P.aggs[a];
    if (d.kind == A_COUNT_COL) {
      for (uint32_t i = tid; i < cap; i += nt) lds_at<uint32_t>(lds, d.lds_off)[i] = 0;
      continue;
    }
    uint64_t init = 0;
    if (d.kind == A_MIN_I || d.kind == A_MIN_F) init = 0x7fffffffffffffffull;
    if (d.kind == A_MAX_I || d.kind == A_MAX_F) init = 0x8000000000000000ull;
    const uint32_t words = (d.kind == A_SUM_I && d.split) ? 2 * cap : cap;
    for (uint32_t i = tid; i < words; i += nt) lds_at<uint64_t>(lds, d.lds_off)[i] = init;
    if (d.nullable)
      for (uint32_t i = tid; i < cap; i += nt) lds_at<uint32_t>(lds, d.lds_ns_off)[i] = 0;
  }
}

// One slot of the LDS table back to its initial state (as lds_init leaves it).
__device__ __forceinline__ void lds_init_slot(const ScanParams& P, uint8_t* lds, uint32_t slot) {
  const uint32_t cap = P.lds_cap;
  lds_at<uint32_t>(lds, P.lds_count_off)[slot] = 0;
  for (uint32_t a = 0; a < P.n_aggs; a++) {
    const AggDesc& d = P.aggs[a];
    if (d.kind == A_COUNT_COL) {
      lds_at<uint32_t>(lds, d.lds_off)[slot] = 0;
      continue;
    }
    uint64_t init = 0;
    if (d.kind == A_MIN_I || d.kind == A_MIN_F) init = 0x7fffffffffffffffull;
    if (d.kind == A_MAX_I || d.kind == A_MAX_F) init = 0x8000000000000000ull;
    lds_at<uint64_t>(lds, d.lds_off)[slot] = init;
    if (d.kind == A_SUM_I && d.split) lds_at<uint64_t>(lds, d.lds_off)[cap + slot] = init;
    if (d.nullable) lds_at<uint32_t>(lds, d.lds_ns_off)[slot] = 0;
  }
}

// Every occupied LDS slot goes to its HBM record once.
__device__ __forceinline__ void lds_flush(const ScanParams& P, uint8_t* lds, uint64_t part_lo) {
  const uint32_t cap = P.lds_cap;
  for (uint32_t s = threadIdx.x; s < cap; s += blockDim.x) {
    const uint32_t cnt = lds_at<uint32_t>(lds, P.lds_count_off)[s];
    if (cnt == 0) continue;
    uint64_t key = part_lo + s;
    if (P.lds_mode == LDS_HASH)
      key = P.lds_key64 ? lds_at<uint64_t>(lds, P.lds_keys_off)[s] : lds_at<uint32_t>(lds, P.lds_keys_off)[s];
    uint8_t* rec = g_record(P, key);
    if (!rec) continue;
    g_add_u64(rec + P.g_count_off, cnt);
    for (uint32_t a = 0; a < P.n_aggs; a++) {
      const AggDesc& d = P.aggs[a];
      if (d.kind == A_COUNT_COL) {
        const uint32_t ns = lds_at<uint32_t>(lds, d.lds_off)[s];
        if (ns) g_add_u64(rec + d.g_off, ns);
        continue;
      }
      const uint32_t ns = d.nullable ? lds_at<uint32_t>(lds, d.lds_ns_off)[s] : cnt;
      if (ns == 0) continue;
      if (d.nullable) g_add_u64(rec + d.g_ns_off, ns);
      const uint64_t w = lds_at<uint64_t>(lds, d.lds_off)[s];
      switch (d.kind) {
        case A_SUM_I:
          if (d.split) {
            // sum = hi_sum * 2^32 + lo_sum, as i128
            const int64_t hs = (int64_t)lds_at<uint64_t>(lds, d.lds_off)[cap + s];
            const __int128 v = (__int128)hs * ((__int128)1 << 32) + (__int128)w;
            if (d.lo_only) g_add_u64(rec + d.g_off, (uint64_t)v);
            else g_add_i128(rec + d.g_off, (uint64_t)v, (uint64_t)(v >> 64));
          } else if (d.lo_only) {
            g_add_u64(rec + d.g_off, w);
          } else {
            g_add_i128(rec + d.g_off, w, (int64_t)w < 0 ? ~0ull : 0ull);
          }
          break;
        case A_SUM_F: g_add_f64(rec + d.g_off, __longlong_as_double((int64_t)w)); break;
        case A_MIN_I: case A_MIN_F: g_min_i64(rec + d.g_off, (int64_t)w); break;
        default: g_max_i64(rec + d.g_off, (int64_t)w); break;
      }
    }
  }
}

// ---- strip building blocks -------------------------------------------------------------

// Folds key part `pt` (values `v`) into the strip's relative keys; a value
// outside the staged statistics raises ERR_KEY_RANGE and drops the row.
// `valid`: the part column's validity bits of the strip (0xf for a
// non-nullable part); a NULL value takes relative value 0.
template <bool FIRST>
__device__ __forceinline__ void key_part(const ScanParams& P, const PartDesc& pt, const int64_t v[kVec],
                                         uint64_t key[kVec], uint32_t& mask, uint32_t valid) {
#pragma unroll
  for (int j = 0; j < kVec; j++) {
    const uint64_t rel = ((valid >> j) & 1u) ? (uint64_t)v[j] - (uint64_t)pt.base : 0ull;
    if (rel >= pt.span && ((mask >> j) & 1u)) {
      atomicOr(P.err, ERR_KEY_RANGE);
      mask &= ~(1u << j);
    }
    key[j] = FIRST ? rel : key[j] + rel * pt.stride;
  }
}

// Resolves each live row to an LDS slot (slot >= 0) or an HBM record and
// counts the row; returns the rows still live.
template <uint32_t LMODE>
__device__ __forceinline__ uint32_t resolve(const ScanParams& P, uint8_t* lds, const uint64_t key[kVec],
                                            uint32_t mask, uint64_t part_lo, int slot[kVec],
                                            uint8_t* rec[kVec]) {
  if (LMODE == LDS_DIRECT) {
    // this workgroup owns keys [part_lo, part_lo + lds_cap)
#pragma unroll
    for (int j = 0; j < kVec; j++)
      if (key[j] - part_lo >= P.lds_cap) mask &= ~(1u << j);
  }
  if (LMODE == LDS_HASH) {
    if (P.lds_key64) lds_find4(lds_at<uint64_t>(lds, P.lds_keys_off), P.lds_cap, key, mask, slot);
    else lds_find4(lds_at<uint32_t>(lds, P.lds_keys_off), P.lds_cap, key, mask, slot);
  }
#pragma unroll
  for (int j = 0; j < kVec; j++) {
    if (LMODE == LDS_DIRECT) slot[j] = (int)(key[j] - part_lo);
    if (LMODE == LDS_NONE) slot[j] = -1;
    rec[j] = nullptr;
    if (!((mask >> j) & 1u)) continue;
    if (slot[j] < 0) {
      rec[j] = g_record(P, key[j]);
      if (!rec[j]) {
        mask &= ~(1u << j);
        continue;
      }
      g_add_u64(rec[j] + P.g_count_off, 1);
    } else {
      l_add_u32(lds_at<uint32_t>(lds, P.lds_count_off) + slot[j], 1);
    }
  }
  return mask;
}

// Aggregate `a` (kind AK over a column of kind K) for the strip's live rows.
template <uint32_t AK, uint32_t K, bool NULLABLE, bool SPLIT>
__device__ __forceinline__ void agg_k(const ScanParams& P, const AggDesc& a, const Raw& x, uint32_t m,
                                      const int slot[kVec], uint8_t* const rec[kVec], uint8_t* lds) {
  int64_t iv[kVec] = {0, 0, 0, 0};
  double fv[kVec] = {0, 0, 0, 0};
  if (AK == A_SUM_F || AK == A_MIN_F || AK == A_MAX_F) dec_f<K>(x, fv);
  else if (AK != A_COUNT_COL) dec_i<K>(x, iv);
#pragma unroll
  for (int j = 0; j < kVec; j++) {
    if (!((m >> j) & 1u)) continue;
    if (slot[j] >= 0) l_apply<AK, NULLABLE, SPLIT>(a, lds, P.lds_cap, (uint32_t)slot[j], iv[j], fv[j]);
    else g_apply<AK, NULLABLE>(a, rec[j], iv[j], fv[j]);
  }
}

// The same with kinds read at run time (the generic kernel).
template <uint32_t AK, bool NULLABLE, bool SPLIT>
__device__ __forceinline__ void agg_col(const ScanParams& P, const AggDesc& a, uint32_t kind, const Raw& x,
                                        uint32_t m, const int slot[kVec], uint8_t* const rec[kVec],
                                        uint8_t* lds) {
  if (AK == A_SUM_F || AK == A_MIN_F || AK == A_MAX_F) {
    if (kind == K_F64) agg_k<AK, K_F64, NULLABLE, SPLIT>(P, a, x, m, slot, rec, lds);
    else agg_k<AK, K_F32, NULLABLE, SPLIT>(P, a, x, m, slot, rec, lds);
    return;
  }
  switch (kind) {
    case K_I64: agg_k<AK, K_I64, NULLABLE, SPLIT>(P, a, x, m, slot, rec, lds); break;
    case K_I32: agg_k<AK, K_I32, NULLABLE, SPLIT>(P, a, x, m, slot, rec, lds); break;
    case K_I16: agg_k<AK, K_I16, NULLABLE, SPLIT>(P, a, x, m, slot, rec, lds); break;
    case K_I8: agg_k<AK, K_I8, NULLABLE, SPLIT>(P, a, x, m, slot, rec, lds); break;
    default: agg_k<AK, K_U8, NULLABLE, SPLIT>(P, a, x, m, slot, rec, lds); break;
  }
}

__device__ __forceinline__ void agg_rt(const ScanParams& P, const AggDesc& a, uint32_t kind, const Raw& x,
                                       uint32_t m, const int slot[kVec], uint8_t* const rec[kVec], uint8_t* lds) {
  switch (a.kind) {
    case A_COUNT_COL: agg_col<A_COUNT_COL, false, false>(P, a, kind, x, m, slot, rec, lds); break;
    case A_SUM_I:
      if (a.split) {
        if (a.nullable) agg_col<A_SUM_I, true, true>(P, a, kind, x, m, slot, rec, lds);
        else agg_col<A_SUM_I, false, true>(P, a, kind, x, m, slot, rec, lds);
      } else {
        if (a.nullable) agg_col<A_SUM_I, true, false>(P, a, kind, x, m, slot, rec, lds);
        else agg_col<A_SUM_I, false, false>(P, a, kind, x, m, slot, rec, lds);
      }
      break;
#define TDBG_AGG_CASE(AK)                                                               \
  case AK:                                                                              \
    if (a.nullable) agg_col<AK, true, false>(P, a, kind, x, m, slot, rec, lds);         \
    else agg_col<AK, false, false>(P, a, kind, x, m, slot, rec, lds);                   \
    break;
    TDBG_AGG_CASE(A_SUM_F)
    TDBG_AGG_CASE(A_MIN_I)
    TDBG_AGG_CASE(A_MAX_I)
    TDBG_AGG_CASE(A_MIN_F)
    TDBG_AGG_CASE(A_MAX_F)
#undef TDBG_AGG_CASE
    default: break;
  }
}

// ---- the kernel body ----------------------------------------------------------------------

// The LDS table's life cycle around a strip routine
// STRIP::run<LMODE>(P, lds, row, count, part_lo) over tiles [t0, t1); with a
// direct table the workgroup owns keys [part_lo, part_lo + lds_cap).
template <uint32_t LMODE, class STRIP>
__device__ __forceinline__ void scan_range(const ScanParams& P, uint8_t* lds, uint32_t t0, uint32_t t1,
                                           uint64_t part_lo) {
  if (LMODE != LDS_NONE) {
    lds_init(P, lds);
    __syncthreads();
  }
  const uint32_t step = blockDim.x * kVec;
  for (uint32_t t = t0; t < t1; t++) {
    const Tile T = P.tiles[t];
    for (uint32_t s = threadIdx.x * kVec; s < T.n; s += step) {
      const uint32_t cnt = T.n - s < (uint32_t)kVec ? T.n - s : (uint32_t)kVec;
      STRIP::template run<LMODE>(P, lds, T.row0 + s, cnt, part_lo);
    }
  }
  if (LMODE != LDS_NONE) {
    __syncthreads();
    lds_flush(P, lds, part_lo);
  }
}

// scan_range with the next strip's column loads issued before the current
// strip is filtered and aggregated, so each thread keeps a strip's loads in
// flight while it works (STRIP provides Data, load() and process<LMODE>()).
// A thread walks its strips of tiles [t0, t1) in the same order as
// scan_range and loads only strips it then processes.
template <uint32_t LMODE, class STRIP>
__device__ __forceinline__ void scan_range_pf(const ScanParams& P, uint8_t* lds, uint32_t t0, uint32_t t1,
                                              uint64_t part_lo) {
  if (LMODE != LDS_NONE) {
    lds_init(P, lds);
    __syncthreads();
  }
  const uint32_t step = blockDim.x * kVec, s0 = threadIdx.x * kVec;
  uint32_t t = t0, s = s0;
  Tile T{};
  while (t < t1) {  // this thread's first strip
    T = P.tiles[t];
    if (s < T.n) break;
    t++;
  }
  typename STRIP::Data cur, nxt;
  if (t < t1) STRIP::load(P, T.row0 + s, cur);
  while (t < t1) {
    uint32_t nt = t, ns = s + step;
    Tile NT = T;
    while (nt < t1 && ns >= NT.n) {
      nt++;
      ns = s0;
      if (nt < t1) NT = P.tiles[nt];
    }
    if (nt < t1) STRIP::load(P, NT.row0 + ns, nxt);
    const uint32_t cnt = T.n - s < (uint32_t)kVec ? T.n - s : (uint32_t)kVec;
    STRIP::template process<LMODE>(P, lds, T.row0 + s, cnt, part_lo, cur);
    t = nt;
    s = ns;
    T = NT;
    cur = nxt;
  }
  if (LMODE != LDS_NONE) {
    __syncthreads();
    lds_flush(P, lds, part_lo);
  }
}

template <uint32_t LMODE, class STRIP>
__device__ __forceinline__ void scan_body_pf(const ScanParams& P, uint8_t* lds) {
  const uint32_t parts = LMODE == LDS_DIRECT ? P.lds_parts : 1u;
  const uint32_t groups = gridDim.x / parts;
  const uint32_t g = blockIdx.x % groups;
  const uint64_t part_lo = (uint64_t)(blockIdx.x / groups) * P.lds_cap;
  const uint32_t t0 = (uint32_t)(((uint64_t)g * P.n_tiles) / groups);
  const uint32_t t1 = (uint32_t)(((uint64_t)(g + 1) * P.n_tiles) / groups);
  scan_range_pf<LMODE, STRIP>(P, lds, t0, t1, part_lo);
}

// One launch over all of P's tiles.  With a partitioned direct table, the
// `parts` workgroups of one group (blockIdx.x = part * groups + g; groups a
// multiple of 8 keeps them on one XCD, sharing its L2) scan the same tiles,
// each aggregating only the keys it owns, so the tiles stream from HBM once.
template <uint32_t LMODE, class STRIP>
__device__ __forceinline__ void scan_body(const ScanParams& P, uint8_t* lds) {
  const uint32_t parts = LMODE == LDS_DIRECT ? P.lds_parts : 1u;
  const uint32_t groups = gridDim.x / parts;
  const uint32_t g = blockIdx.x % groups;
  const uint64_t part_lo = (uint64_t)(blockIdx.x / groups) * P.lds_cap;
  const uint32_t t0 = (uint32_t)(((uint64_t)g * P.n_tiles) / groups);
  const uint32_t t1 = (uint32_t)(((uint64_t)(g + 1) * P.n_tiles) / groups);
  scan_range<LMODE, STRIP>(P, lds, t0, t1, part_lo);
}

}  // namespace tdbg
