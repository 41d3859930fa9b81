// order_device.h — order-preserving 64-bit words of ORDER BY items, shared
// by the sorter and result writer (order.hip) and the radix leaf pass
// (radix.hip), which emits the first word of each group it writes.
//
// Follows the reference's SortCtx.lessThan/compareBySpec
// (src/exec/v2_lowcard_group.zig:2069-2311) and compareAggregateValue
// (src/exec/v2_shape_group_topn.zig:1361): every ORDER BY item becomes one or
// two words whose unsigned order is the item's order (desc = bitwise NOT).
#pragma once

#include "engine.h"

namespace tdbg {

// ---- shared helpers ----------------------------------------------------------

__device__ __forceinline__ uint64_t ord_i64(int64_t v) { return (uint64_t)v ^ 0x8000000000000000ull; }

// types.floatOrder (src/types.zig:579-587): -0 == +0, NaN largest.
__device__ __forceinline__ uint64_t ord_f64(double x) {
  if (x != x) return ~0ull;
  if (x == 0.0) x = 0.0;
  uint64_t b = (uint64_t)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Inverse of the order-preserving i64 map used for float MIN/MAX states.
__device__ __forceinline__ double unord_f64(int64_t k) {
  int64_t s = k >= 0 ? k : (k ^ 0x7fffffffffffffffll);
  return __longlong_as_double(s);
}

// i128 -> f64 with round-to-nearest-even (Zig's @floatFromInt).
__device__ __forceinline__ double i128_to_f64(__int128 v) {
  const bool neg = v < 0;
  unsigned __int128 a = neg ? (unsigned __int128)(-(v + 1)) + 1u : (unsigned __int128)v;
  const uint64_t hi = (uint64_t)(a >> 64);
  double r;
  if (hi == 0) {
    r = (double)(uint64_t)a;  // u64 -> f64 is correctly rounded
  } else {
    const int lz = __clzll(hi);
    const int drop = 64 - lz;  // bits beyond a 64-bit window
    uint64_t m = (uint64_t)(a >> drop);
    const unsigned __int128 rest = a & ((((unsigned __int128)1) << drop) - 1u);
    if (rest != 0) m |= 1u;  // sticky bit below the rounding position
    r = ldexp((double)m, drop);
  }
  return neg ? -r : r;
}

__device__ __forceinline__ __int128 load_i128(const uint8_t* p) {
  const uint64_t lo = *reinterpret_cast<const uint64_t*>(p);
  const int64_t hi = *reinterpret_cast<const int64_t*>(p + 8);
  return (__int128)(((unsigned __int128)(uint64_t)hi << 64) | lo);
}

__device__ __forceinline__ uint64_t group_key(const uint8_t* rec, uint32_t gid, uint32_t g_mode, uint32_t key_off) {
  return g_mode == G_DENSE ? (uint64_t)gid : *reinterpret_cast<const uint64_t*>(rec + key_off);
}

__device__ __forceinline__ int64_t part_value(const PartDesc& pt, uint64_t key) {
  const uint64_t rel = (key / pt.stride) % pt.span;
  return (int64_t)((uint64_t)pt.base + rel);
}

__device__ __forceinline__ uint64_t rec_ns(const uint8_t* rec, uint32_t ns_off, uint32_t count_off) {
  return *reinterpret_cast<const uint64_t*>(rec + (ns_off != ~0u ? ns_off : count_off));
}

// One ORDER BY word of one group.
__device__ __forceinline__ uint64_t comp_word(const OrderParams& P, const CompDesc& c, uint32_t gid) {
  const uint8_t* rec = P.g_state + (uint64_t)gid * P.g_rec;
  uint64_t w = 0;
  switch (c.src) {
    case C_KEY: w = ord_i64(part_value(P.parts[c.idx], group_key(rec, gid, P.g_mode, P.g_key_off))); break;
    case C_COUNT: w = *reinterpret_cast<const uint64_t*>(rec + P.g_count_off); break;
    case C_U64_AT: w = *reinterpret_cast<const uint64_t*>(rec + c.off); break;
    case C_I64_AT:
    case C_I128_HI:
    case C_I128_LO:
    case C_F64_AT:
    case C_FORD_AT: {
      // A group with no non-null input compares as a zero slot
      // (compareBySpec reads the untouched slot).
      const bool empty = rec_ns(rec, c.ns_off, P.g_count_off) == 0;
      if (c.src == C_I64_AT) w = ord_i64(empty ? 0 : *reinterpret_cast<const int64_t*>(rec + c.off));
      else if (c.src == C_I128_HI) w = ord_i64(empty ? 0 : *reinterpret_cast<const int64_t*>(rec + c.off + 8));
      else if (c.src == C_I128_LO) w = empty ? 0 : *reinterpret_cast<const uint64_t*>(rec + c.off);
      else if (c.src == C_F64_AT) w = ord_f64(empty ? 0.0 : *reinterpret_cast<const double*>(rec + c.off));
      else w = ord_f64(empty ? 0.0 : unord_f64(*reinterpret_cast<const int64_t*>(rec + c.off)));
      break;
    }
    case C_AVG_I: {
      const uint64_t n = rec_ns(rec, c.ns_off, P.g_count_off);
      w = ord_f64(n == 0 ? 0.0 : i128_to_f64(load_i128(rec + c.off)) / (double)n);
      break;
    }
    default: {  // C_AVG_F
      const uint64_t n = rec_ns(rec, c.ns_off, P.g_count_off);
      w = ord_f64(n == 0 ? 0.0 : *reinterpret_cast<const double*>(rec + c.off) / (double)n);
      break;
    }
  }
  return c.desc ? ~w : w;
}

}  // namespace tdbg
