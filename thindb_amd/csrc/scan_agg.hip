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
__device__ void lds_flush(const ScanParams& P, uint8_t* lds) {
  const uint32_t cap = P.lds_cap;
  for (uint32_t s = threadIdx.x; s < cap; s += blockDim.x) {
    const uint32_t cnt = lds_at<uint32_t>(lds, P.lds_count_off)[s];
    if (cnt == 0) continue;
    uint64_t key = s;
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

template <uint32_t LMODE>
__device__ __forceinline__ void strip(const ScanParams& P, uint8_t* lds, uint64_t r, uint32_t cnt) {
  uint32_t mask = (1u << cnt) - 1u;
  for (uint32_t i = 0; i < P.n_preds; i++) {
    const PredDesc& pd = P.preds[i];
    mask &= eval_pred(pd, P.cols[pd.col], r);
    if (!mask) return;
  }
  uint64_t key[kVec] = {0, 0, 0, 0};
  for (uint32_t i = 0; i < P.n_parts; i++) {
    const PartDesc& pt = P.parts[i];
    int64_t v[kVec];
    load4_i(P.cols[pt.col], r, v);
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
  if (!mask) return;
  int slot[kVec];
  uint8_t* rec[kVec];
#pragma unroll
  for (int j = 0; j < kVec; j++) {
    slot[j] = -1;
    rec[j] = nullptr;
    if (!((mask >> j) & 1u)) continue;
    if (LMODE == LDS_DIRECT) {
      slot[j] = (int)key[j];
    } else if (LMODE == LDS_HASH) {
      slot[j] = P.lds_key64 ? lds_find(lds_at<uint64_t>(lds, P.lds_keys_off), P.lds_cap, key[j])
                            : lds_find(lds_at<uint32_t>(lds, P.lds_keys_off), P.lds_cap, key[j]);
    }
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
      load4_f(c, r, fv);
    } else if (d.kind != A_COUNT_COL) {
      load4_i(c, r, iv);
    }
#pragma unroll
    for (int j = 0; j < kVec; j++) {
      if (!((m >> j) & 1u)) continue;
      if (slot[j] >= 0) l_apply_row(P, d, lds, (uint32_t)slot[j], iv[j], fv[j]);
      else g_apply_row(P, d, rec[j], iv[j], fv[j]);
    }
  }
}

template <uint32_t LMODE>
__global__ void __launch_bounds__(1024) scan_agg_kernel(const ScanParams P) {
  extern __shared__ __align__(16) uint8_t lds[];
  if (LMODE != LDS_NONE) {
    lds_init(P, lds);
    __syncthreads();
  }
  const uint32_t t0 = (uint32_t)(((uint64_t)blockIdx.x * P.n_tiles) / gridDim.x);
  const uint32_t t1 = (uint32_t)(((uint64_t)(blockIdx.x + 1) * P.n_tiles) / gridDim.x);
  const uint32_t step = blockDim.x * kVec;
  for (uint32_t t = t0; t < t1; t++) {
    const Tile T = P.tiles[t];
    for (uint32_t s = threadIdx.x * kVec; s < T.n; s += step) {
      const uint32_t cnt = T.n - s < (uint32_t)kVec ? T.n - s : (uint32_t)kVec;
      strip<LMODE>(P, lds, T.row0 + s, cnt);
    }
  }
  if (LMODE != LDS_NONE) {
    __syncthreads();
    lds_flush(P, lds);
  }
}

static hipError_t allow_large_lds() {
  const int bytes = 160 * 1024;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(scan_agg_kernel<LDS_DIRECT>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(scan_agg_kernel<LDS_HASH>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  return e;
}

hipError_t launch_scan_agg(const ScanParams& p, uint32_t grid, uint32_t block, hipStream_t s) {
  static const hipError_t attr = allow_large_lds();
  (void)attr;  // best effort: a launch over the limit reports its own error
  const size_t lds = p.lds_mode == LDS_NONE ? 0 : p.lds_bytes;
  switch (p.lds_mode) {
    case LDS_DIRECT:
      hipLaunchKernelGGL(scan_agg_kernel<LDS_DIRECT>, dim3(grid), dim3(block), lds, s, p);
      break;
    case LDS_HASH:
      hipLaunchKernelGGL(scan_agg_kernel<LDS_HASH>, dim3(grid), dim3(block), lds, s, p);
      break;
    default:
      hipLaunchKernelGGL(scan_agg_kernel<LDS_NONE>, dim3(grid), dim3(block), 0, s, p);
      break;
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
