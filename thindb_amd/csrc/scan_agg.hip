// scan_agg.hip — fused scan -> filter -> GROUP BY aggregate for gfx950.
//
// Replaces the CPU fold of thinDB's grouped operators: the scan/filter of
// src/exec/scan.zig:4663-4800 (filterRowGroup), the key packing of
// src/exec/v2_lowcard_group.zig:886-948, the gid resolution of :1054-1118 and
// the fold kernels of :1206-1347 (silo equivalents:
// src/exec/silo_group_core.zig:7290-8003).
//
// Design (one launch per query):
//  * Each thread owns kVec=4 consecutive rows per strip, so every column is
//    read with one 16-byte (i32) or two 16-byte (i64/f64) loads per thread
//    and a wavefront reads 1-2 KiB contiguous per column: fully coalesced.
//  * Predicates run first; a lane whose rows all fail skips the key and
//    aggregate loads, so their cache lines are fetched only when some row of
//    the line survives.
//  * Each workgroup aggregates into an LDS table (direct-indexed when the
//    key span fits, else open addressing on the relative key), with LDS
//    atomics; a row whose key finds no LDS slot within kLdsProbe probes
//    spills straight into the HBM table.  At the end the workgroup flushes
//    its occupied slots into the HBM table with 64-bit atomics (i128 sums
//    with carry), so HBM sees one update per (workgroup, group).
//  * Without an LDS table (very high cardinality) every row updates its HBM
//    record directly.
// Integer sums are exact: LDS partials are i64 when rows-per-workgroup times
// max|v| stays below 2^63, else split into 32-bit halves; HBM totals are
// i128.  Float sums use f64 atomics (order-dependent rounding, like the
// reference's parallel merge).
#include "engine.h"

namespace tdbg {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Order-preserving map of an f64 onto i64 (signed compare).
__device__ __forceinline__ int64_t f64_ord(double x) {
  int64_t s = __double_as_longlong(x);
  return s >= 0 ? s : (s ^ 0x7fffffffffffffffll);
}

__device__ __forceinline__ void load4_i(const ColDesc& c, uint64_t r, int64_t v[kVec]) {
  switch (c.kind) {
    case K_I64: {
      const longlong2* p = reinterpret_cast<const longlong2*>(static_cast<const int64_t*>(c.data) + r);
      longlong2 a = p[0], b = p[1];
      v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
      break;
    }
    case K_I32: {
      int4 a = *reinterpret_cast<const int4*>(static_cast<const int32_t*>(c.data) + r);
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
      break;
    }
    case K_I16: {
      short4 a = *reinterpret_cast<const short4*>(static_cast<const int16_t*>(c.data) + r);
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
      break;
    }
    case K_I8: {
      char4 a = *reinterpret_cast<const char4*>(static_cast<const int8_t*>(c.data) + r);
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
      break;
    }
    default: {  // K_U8 (BOOLEAN)
      uchar4 a = *reinterpret_cast<const uchar4*>(static_cast<const uint8_t*>(c.data) + r);
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
      break;
    }
  }
}

__device__ __forceinline__ void load4_f(const ColDesc& c, uint64_t r, double v[kVec]) {
  if (c.kind == K_F64) {
    const double2* p = reinterpret_cast<const double2*>(static_cast<const double*>(c.data) + r);
    double2 a = p[0], b = p[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
  } else {
    float4 a = *reinterpret_cast<const float4*>(static_cast<const float*>(c.data) + r);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  }
}

__device__ __forceinline__ uint32_t valid4(const ColDesc& c, uint64_t r) {
  if (!c.valid) return 0xfu;
  return (c.valid[r >> 5] >> (r & 31)) & 0xfu;
}

__device__ __forceinline__ uint32_t eval_pred(const PredDesc& p, const ColDesc& c, uint64_t r) {
  uint32_t m = 0;
  if (p.kind == P_RANGE_I || p.kind == P_NEQ_I) {
    int64_t v[kVec];
    load4_i(c, r, v);
#pragma unroll
    for (int j = 0; j < kVec; j++) {
      bool ok = p.kind == P_RANGE_I ? (v[j] >= p.lo_i && v[j] <= p.hi_i) : (v[j] != p.lo_i);
      m |= (uint32_t)ok << j;
    }
  } else {
    double v[kVec];
    load4_f(c, r, v);
#pragma unroll
    for (int j = 0; j < kVec; j++) {
      bool ok;
      if (p.kind == P_RANGE_F) {
        bool lo = p.lo_incl ? v[j] >= p.lo_f : v[j] > p.lo_f;
        bool hi = p.hi_incl ? v[j] <= p.hi_f : v[j] < p.hi_f;
        ok = lo && hi;
      } else {
        ok = v[j] != p.lo_f;
      }
      m |= (uint32_t)ok << j;
    }
  }
  return m & valid4(c, r);
}

// ---- HBM records -----------------------------------------------------------

__device__ __forceinline__ void g_add_u64(uint8_t* p, uint64_t v) {
  __hip_atomic_fetch_add(reinterpret_cast<uint64_t*>(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exact i128 += v with two 64-bit atomics: the low word's carry goes into
// the high word.  Commutative, so concurrent adds compose exactly.
__device__ __forceinline__ void g_add_i128(uint8_t* p, uint64_t lo, uint64_t hi) {
  if (lo) {
    uint64_t old = __hip_atomic_fetch_add(reinterpret_cast<uint64_t*>(p), lo, __ATOMIC_RELAXED,
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
// CAS-claimed u64 key word (kEmptyKey = free).
__device__ uint8_t* g_record(const ScanParams& P, uint64_t key) {
  if (P.g_mode == G_DENSE) return P.g_state + key * P.g_rec;
  const uint64_t mask = P.g_cap - 1;
  uint64_t s = mix64(key) & mask;
#pragma nounroll
  for (uint32_t probe = 0; probe < 65536u; probe++) {
    uint8_t* rec = P.g_state + s * P.g_rec;
    uint64_t* kp = reinterpret_cast<uint64_t*>(rec + P.g_key_off);
    uint64_t cur = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
__device__ __forceinline__ void g_apply_row(const ScanParams& P, const AggDesc& a, uint8_t* rec, int64_t iv,
                                            double fv) {
  switch (a.kind) {
    case A_COUNT_COL: g_add_u64(rec + a.g_off, 1); return;
    case A_SUM_I: g_add_i128(rec + a.g_off, (uint64_t)iv, iv < 0 ? ~0ull : 0ull); break;
    case A_SUM_F: g_add_f64(rec + a.g_off, fv); break;
    case A_MIN_I: g_min_i64(rec + a.g_off, iv); break;
    case A_MAX_I: g_max_i64(rec + a.g_off, iv); break;
    case A_MIN_F: g_min_i64(rec + a.g_off, f64_ord(fv)); break;
    default: g_max_i64(rec + a.g_off, f64_ord(fv)); break;
  }
  if (a.nullable) g_add_u64(rec + a.g_ns_off, 1);
}

// ---- LDS table -------------------------------------------------------------

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

__device__ __forceinline__ void l_apply_row(const ScanParams& P, const AggDesc& a, uint8_t* lds, uint32_t slot,
                                            int64_t iv, double fv) {
  const uint32_t cap = P.lds_cap;
  switch (a.kind) {
    case A_COUNT_COL: l_add_u32(lds_at<uint32_t>(lds, a.lds_off) + slot, 1); return;
    case A_SUM_I:
      if (a.split) {
        l_add_u64(lds_at<uint64_t>(lds, a.lds_off) + slot, (uint64_t)(uint32_t)iv);
        l_add_u64(lds_at<uint64_t>(lds, a.lds_off) + cap + slot, (uint64_t)(int64_t)(int32_t)(iv >> 32));
      } else {
        l_add_u64(lds_at<uint64_t>(lds, a.lds_off) + slot, (uint64_t)iv);
      }
      break;
    case A_SUM_F:
      __hip_atomic_fetch_add(lds_at<double>(lds, a.lds_off) + slot, fv, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
      break;
    case A_MIN_I:
      __hip_atomic_fetch_min(lds_at<int64_t>(lds, a.lds_off) + slot, iv, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
      break;
    case A_MAX_I:
      __hip_atomic_fetch_max(lds_at<int64_t>(lds, a.lds_off) + slot, iv, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
      break;
    case A_MIN_F:
      __hip_atomic_fetch_min(lds_at<int64_t>(lds, a.lds_off) + slot, f64_ord(fv), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
      break;
    default:
      __hip_atomic_fetch_max(lds_at<int64_t>(lds, a.lds_off) + slot, f64_ord(fv), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
      break;
  }
  if (a.nullable) l_add_u32(lds_at<uint32_t>(lds, a.lds_ns_off) + slot, 1);
}

__device__ __forceinline__ uint32_t hash32(uint64_t k) { return (uint32_t)(mix64(k) >> 32); }

// Open addressing in LDS on the relative key; -1 when no slot is found
// within kLdsProbe probes (the row then spills to HBM).
template <typename K>
__device__ __forceinline__ int lds_find(K* keys, uint32_t cap, uint64_t key64) {
  const K key = (K)key64;
  const K empty = (K)kEmptyKey;
  uint32_t s = __umulhi(hash32(key64), cap);
#pragma nounroll
  for (int probe = 0; probe < kLdsProbe; probe++) {
    K cur = __hip_atomic_load(keys + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == key) return (int)s;
    if (cur == empty) {
      K expect = empty;
      if (__hip_atomic_compare_exchange_strong(keys + s, &expect, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP))
        return (int)s;
      if (expect == key) return (int)s;
    }
    s = s + 1 == cap ? 0 : s + 1;
  }
  return -1;
}

__device__ void lds_init(const ScanParams& P, uint8_t* lds) {
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
    uint32_t words = (d.kind == A_SUM_I && d.split) ? 2 * cap : cap;
    for (uint32_t i = tid; i < words; i += nt) lds_at<uint64_t>(lds, d.lds_off)[i] = init;
    if (d.nullable)
      for (uint32_t i = tid; i < cap; i += nt) lds_at<uint32_t>(lds, d.lds_ns_off)[i] = 0;
  }
}

// Every occupied LDS slot goes to its HBM record once.
__device__ void lds_flush(const ScanParams& P, uint8_t* lds, uint64_t part_lo) {
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
        uint32_t ns = lds_at<uint32_t>(lds, d.lds_off)[s];
        if (ns) g_add_u64(rec + d.g_off, ns);
        continue;
      }
      uint32_t ns = d.nullable ? lds_at<uint32_t>(lds, d.lds_ns_off)[s] : cnt;
      if (ns == 0) continue;
      if (d.nullable) g_add_u64(rec + d.g_ns_off, ns);
      const uint64_t w = lds_at<uint64_t>(lds, d.lds_off)[s];
      switch (d.kind) {
        case A_SUM_I:
          if (d.split) {
            // sum = hi_sum * 2^32 + lo_sum, as i128
            const int64_t hs = (int64_t)lds_at<uint64_t>(lds, d.lds_off)[cap + s];
            const __int128 v = ((__int128)hs << 32) + (__int128)w;
            g_add_i128(rec + d.g_off, (uint64_t)v, (uint64_t)(v >> 64));
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

// ---- one strip: kVec rows of this thread -----------------------------------

// The raw bytes of one column for this thread's kVec rows (at most 32), as
// four 64-bit words: 64-bit kinds one row per word, narrower kinds packed
// little-endian from w0.
struct Raw {
  uint64_t w0, w1, w2, w3;
};

__device__ __forceinline__ Raw load_raw(const ColDesc& c, uint64_t r) {
  Raw x = {0, 0, 0, 0};
  switch (c.kind) {
    case K_I64:
    case K_F64: {
      const ulonglong2* p = reinterpret_cast<const ulonglong2*>(static_cast<const uint64_t*>(c.data) + r);
      const ulonglong2 a = p[0], b = p[1];
      x.w0 = a.x; x.w1 = a.y; x.w2 = b.x; x.w3 = b.y;
      break;
    }
    case K_I32:
    case K_F32: {
      const ulonglong2 a = *reinterpret_cast<const ulonglong2*>(static_cast<const uint32_t*>(c.data) + r);
      x.w0 = a.x; x.w1 = a.y;
      break;
    }
    case K_I16: x.w0 = *reinterpret_cast<const uint64_t*>(static_cast<const uint16_t*>(c.data) + r); break;
    default: x.w0 = *reinterpret_cast<const uint32_t*>(static_cast<const uint8_t*>(c.data) + r); break;
  }
  return x;
}

__device__ __forceinline__ void decode_i(const Raw& x, uint32_t kind, int64_t v[kVec]) {
  switch (kind) {
    case K_I64: v[0] = (int64_t)x.w0; v[1] = (int64_t)x.w1; v[2] = (int64_t)x.w2; v[3] = (int64_t)x.w3; break;
    case K_I32:
      v[0] = (int32_t)(uint32_t)x.w0; v[1] = (int32_t)(uint32_t)(x.w0 >> 32);
      v[2] = (int32_t)(uint32_t)x.w1; v[3] = (int32_t)(uint32_t)(x.w1 >> 32);
      break;
    case K_I16:
      v[0] = (int16_t)(uint16_t)x.w0; v[1] = (int16_t)(uint16_t)(x.w0 >> 16);
      v[2] = (int16_t)(uint16_t)(x.w0 >> 32); v[3] = (int16_t)(uint16_t)(x.w0 >> 48);
      break;
    case K_I8:
      v[0] = (int8_t)(uint8_t)x.w0; v[1] = (int8_t)(uint8_t)(x.w0 >> 8);
      v[2] = (int8_t)(uint8_t)(x.w0 >> 16); v[3] = (int8_t)(uint8_t)(x.w0 >> 24);
      break;
    default:  // K_U8
      v[0] = (uint8_t)x.w0; v[1] = (uint8_t)(x.w0 >> 8); v[2] = (uint8_t)(x.w0 >> 16); v[3] = (uint8_t)(x.w0 >> 24);
      break;
  }
}

__device__ __forceinline__ void decode_f(const Raw& x, uint32_t kind, double v[kVec]) {
  if (kind == K_F64) {
    v[0] = __longlong_as_double((int64_t)x.w0); v[1] = __longlong_as_double((int64_t)x.w1);
    v[2] = __longlong_as_double((int64_t)x.w2); v[3] = __longlong_as_double((int64_t)x.w3);
  } else {
    v[0] = __uint_as_float((uint32_t)x.w0); v[1] = __uint_as_float((uint32_t)(x.w0 >> 32));
    v[2] = __uint_as_float((uint32_t)x.w1); v[3] = __uint_as_float((uint32_t)(x.w1 >> 32));
  }
}

constexpr int kFastCols = 6;  // columns a strip holds in registers

// The strip's columns held in registers (named fields, never an indexed
// array, so they stay in VGPRs).
struct RawSet {
  Raw c0, c1, c2, c3, c4, c5;
};

// Column `c` of the strip: from the registers loaded at the strip's start
// (FAST) or straight from memory.  `c` is uniform across the launch, so the
// switch is a scalar branch.
template <bool FAST>
__device__ __forceinline__ Raw fetch(const ScanParams& P, const RawSet& raw, uint32_t c, uint64_t r) {
  if (FAST) {
    switch (__builtin_amdgcn_readfirstlane(c)) {
      case 0: return raw.c0;
      case 1: return raw.c1;
      case 2: return raw.c2;
      case 3: return raw.c3;
      case 4: return raw.c4;
      default: return raw.c5;
    }
  }
  return load_raw(P.cols[c], r);
}

__device__ __forceinline__ uint32_t eval_pred_raw(const PredDesc& p, const Raw& x, uint32_t kind) {
  uint32_t m = 0;
  if (p.kind == P_RANGE_I || p.kind == P_NEQ_I) {
    int64_t v[kVec];
    decode_i(x, kind, v);
#pragma unroll
    for (int j = 0; j < kVec; j++) {
      const bool ok = p.kind == P_RANGE_I ? (v[j] >= p.lo_i && v[j] <= p.hi_i) : (v[j] != p.lo_i);
      m |= (uint32_t)ok << j;
    }
  } else {
    double v[kVec];
    decode_f(x, kind, v);
#pragma unroll
    for (int j = 0; j < kVec; j++) {
      bool ok;
      if (p.kind == P_RANGE_F) {
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

template <uint32_t LMODE, bool FAST>
__device__ __forceinline__ void strip(const ScanParams& P, uint8_t* lds, uint64_t r, uint32_t cnt,
                                      uint64_t part_lo) {
  // Every column of the strip is requested before any is used, so a
  // thread keeps all its loads in flight at once.
  RawSet raw = {};
  if (FAST) {
    const uint32_t nc = P.n_cols;
    if (nc > 0) raw.c0 = load_raw(P.cols[0], r);
    if (nc > 1) raw.c1 = load_raw(P.cols[1], r);
    if (nc > 2) raw.c2 = load_raw(P.cols[2], r);
    if (nc > 3) raw.c3 = load_raw(P.cols[3], r);
    if (nc > 4) raw.c4 = load_raw(P.cols[4], r);
    if (nc > 5) raw.c5 = load_raw(P.cols[5], r);
  }
  uint32_t mask = (1u << cnt) - 1u;
  for (uint32_t i = 0; i < P.n_preds; i++) {
    const PredDesc& pd = P.preds[i];
    const ColDesc& c = P.cols[pd.col];
    mask &= eval_pred_raw(pd, fetch<FAST>(P, raw, pd.col, r), c.kind) & valid4(c, r);
  }
  if (!mask) return;
  uint64_t key[kVec] = {0, 0, 0, 0};
  for (uint32_t i = 0; i < P.n_parts; i++) {
    const PartDesc& pt = P.parts[i];
    int64_t v[kVec];
    decode_i(fetch<FAST>(P, raw, pt.col, r), P.cols[pt.col].kind, v);
#pragma unroll
    for (int j = 0; j < kVec; j++) {
      const uint64_t rel = (uint64_t)v[j] - (uint64_t)pt.base;
      if (rel >= pt.span && ((mask >> j) & 1u)) {
        atomicOr(P.err, ERR_KEY_RANGE);
        mask &= ~(1u << j);
      }
      key[j] += rel * pt.stride;
    }
  }
  if (LMODE == LDS_DIRECT) {
    // this workgroup owns keys [part_lo, part_lo + lds_cap)
#pragma unroll
    for (int j = 0; j < kVec; j++)
      if (key[j] - part_lo >= P.lds_cap) mask &= ~(1u << j);
  }
  if (!mask) return;
  int slot[kVec];
  uint8_t* rec[kVec];
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
  for (uint32_t a = 0; a < P.n_aggs; a++) {
    const AggDesc& d = P.aggs[a];
    const ColDesc& c = P.cols[d.col];
    const uint32_t m = mask & valid4(c, r);
    if (!m) continue;
    int64_t iv[kVec] = {0, 0, 0, 0};
    double fv[kVec] = {0, 0, 0, 0};
    if (d.kind == A_SUM_F || d.kind == A_MIN_F || d.kind == A_MAX_F) {
      decode_f(fetch<FAST>(P, raw, d.col, r), c.kind, fv);
    } else if (d.kind != A_COUNT_COL) {
      decode_i(fetch<FAST>(P, raw, d.col, r), c.kind, iv);
    }
#pragma unroll
    for (int j = 0; j < kVec; j++) {
      if (!((m >> j) & 1u)) continue;
      if (slot[j] >= 0) l_apply_row(P, d, lds, (uint32_t)slot[j], iv[j], fv[j]);
      else g_apply_row(P, d, rec[j], iv[j], fv[j]);
    }
  }
}

template <uint32_t LMODE, bool FAST>
__global__ void __launch_bounds__(1024) scan_agg_kernel(const ScanParams P) {
  extern __shared__ __align__(16) uint8_t lds[];
  if (LMODE != LDS_NONE) {
    lds_init(P, lds);
    __syncthreads();
  }
  // With a partitioned direct table, the `parts` workgroups of one group
  // (blockIdx.x = part * groups + g; groups a multiple of 8 keeps them on
  // one XCD, sharing its L2) scan the same tiles, each aggregating only the
  // keys it owns, so the tiles stream from HBM once.
  const uint32_t parts = LMODE == LDS_DIRECT ? P.lds_parts : 1u;
  const uint32_t groups = gridDim.x / parts;
  const uint32_t g = blockIdx.x % groups;
  const uint64_t part_lo = (uint64_t)(blockIdx.x / groups) * P.lds_cap;
  const uint32_t t0 = (uint32_t)(((uint64_t)g * P.n_tiles) / groups);
  const uint32_t t1 = (uint32_t)(((uint64_t)(g + 1) * P.n_tiles) / groups);
  const uint32_t step = blockDim.x * kVec;
  for (uint32_t t = t0; t < t1; t++) {
    const Tile T = P.tiles[t];
    for (uint32_t s = threadIdx.x * kVec; s < T.n; s += step) {
      const uint32_t cnt = T.n - s < (uint32_t)kVec ? T.n - s : (uint32_t)kVec;
      strip<LMODE, FAST>(P, lds, T.row0 + s, cnt, part_lo);
    }
  }
  if (LMODE != LDS_NONE) {
    __syncthreads();
    lds_flush(P, lds, part_lo);
  }
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
                                    const uint64_t* __restrict__ tmpl) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_words;
       i += (uint64_t)gridDim.x * blockDim.x)
    state[i] = tmpl[i % rec_words];
}

hipError_t launch_fill_records(uint8_t* state, uint64_t n_rec, uint32_t rec_bytes, const uint64_t* tmpl,
                               hipStream_t s) {
  const uint64_t words = n_rec * (rec_bytes / 8);
  uint64_t blocks = (words + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_records_kernel, dim3((uint32_t)blocks), dim3(256), 0, s,
                     reinterpret_cast<uint64_t*>(state), words, rec_bytes / 8, tmpl);
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
