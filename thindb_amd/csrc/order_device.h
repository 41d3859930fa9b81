// order_device.h — order-preserving 64-bit words of ORDER BY items, shared
// by the sorter and result writer (order.hip) and the radix leaf pass
// (radix.hip), which emits the first word of each group it writes.
//
// Follows the silo's result order, rowOrderBefore/compareOrderKey/
// compareAggregateValue (src/exec/v2_shape_group_topn.zig:1327-1395); the
// low-cardinality core's SortCtx.lessThan/compareBySpec
// (src/exec/v2_lowcard_group.zig:2198-2271) orders by the same items (a
// DECIMAL AVG by its unscaled mean, the same order up to rounding ties) and
// breaks ties by gid, which its parallel merge leaves unspecified.  Each
// ORDER BY item becomes one to three words whose unsigned order is the item's
// order (desc = bitwise NOT), and the packed group key's fields (C_PKEY)
// break ties.
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

__device__ __forceinline__ GKey group_key(const uint8_t* rec, uint32_t gid, uint32_t g_mode, uint32_t key_off) {
  if (g_mode == G_DENSE) return GKey{(uint64_t)gid, 0};
  const uint64_t* k = reinterpret_cast<const uint64_t*>(rec + key_off);
  return GKey{k[0], g_mode == G_WIDE ? k[1] : 0ull};
}

// The part's relative value: its bit field of a packed key, else its digit
// of the mixed-radix key.
__device__ __forceinline__ uint64_t part_rel(const PartDesc& pt, GKey key) {
  if (pt.bit_width) {
    const unsigned __int128 k = ((unsigned __int128)key.hi << 64) | key.lo;
    const uint64_t f = (uint64_t)(k >> pt.bit_off);
    return pt.bit_width >= 64 ? f : f & ((1ull << pt.bit_width) - 1);
  }
  return (key.lo / pt.stride) % pt.span;
}

__device__ __forceinline__ int64_t part_value(const PartDesc& pt, GKey key) {
  return (int64_t)((uint64_t)pt.base + part_rel(pt, key));
}

// A nullable part's NULL is relative value 0.
__device__ __forceinline__ bool part_null(const PartDesc& pt, GKey key) {
  return pt.nullable && part_rel(pt, key) == 0;
}

__device__ __forceinline__ uint64_t rec_ns(const uint8_t* rec, uint32_t ns_off, uint32_t count_off) {
  return *reinterpret_cast<const uint64_t*>(rec + (ns_off != ~0u ? ns_off : count_off));
}

// One ORDER BY word of one group.
__device__ __forceinline__ uint64_t comp_word(const OrderParams& P, const CompDesc& c, uint32_t gid) {
  const uint8_t* rec = P.g_state + (uint64_t)gid * P.g_rec;
  uint64_t w = 0;
  switch (c.src) {
    case C_KEY: {
      // a nullable part orders by its relative value: NULL (0) first
      // ascending, last descending (compareBySpec, v2_lowcard_group.zig:2219-2224)
      const PartDesc& pt = P.parts[c.idx];
      const GKey k = group_key(rec, gid, P.g_mode, P.g_key_off);
      w = pt.nullable ? part_rel(pt, k) : ord_i64(part_value(pt, k));
      break;
    }
    case C_COUNT: w = *reinterpret_cast<const uint64_t*>(rec + P.g_count_off); break;
    case C_U64_AT: w = *reinterpret_cast<const uint64_t*>(rec + c.off); break;
    case C_I64_AT:
    case C_I128_HI:
    case C_I128_LO:
    case C_F64_AT:
    case C_FORD_AT: {
      // A group with no non-null input is NULL: a C_AGG_NULL word before
      // this one orders it, so its value word is a placeholder 0.
      const bool empty = rec_ns(rec, c.ns_off, P.g_count_off) == 0;
      if (c.src == C_I64_AT) w = ord_i64(empty ? 0 : *reinterpret_cast<const int64_t*>(rec + c.off));
      else if (c.src == C_I128_HI) w = ord_i64(empty ? 0 : *reinterpret_cast<const int64_t*>(rec + c.off + 8));
      else if (c.src == C_I128_LO) w = empty ? 0 : *reinterpret_cast<const uint64_t*>(rec + c.off);
      else if (c.src == C_F64_AT) w = ord_f64(empty ? 0.0 : *reinterpret_cast<const double*>(rec + c.off));
      else w = ord_f64(empty ? 0.0 : unord_f64(*reinterpret_cast<const int64_t*>(rec + c.off)));
      break;
    }
    case C_AVG_I: {  // the emitted value (avgValue, v2_shape_group_topn.zig:1436-1444)
      const uint64_t n = rec_ns(rec, c.ns_off, P.g_count_off);
      double p10 = 1.0;
      for (uint32_t s = 0; s < c.aux; s++) p10 *= 10.0;
      w = ord_f64(n == 0 ? 0.0 : i128_to_f64(load_i128(rec + c.off)) / (double)n / p10);
      break;
    }
    case C_AVG_F: {
      const uint64_t n = rec_ns(rec, c.ns_off, P.g_count_off);
      w = ord_f64(n == 0 ? 0.0 : *reinterpret_cast<const double*>(rec + c.off) / (double)n);
      break;
    }
    case C_AGG_NULL:  // an all-NULL aggregate orders first (compareAggregateValue :1361-1370)
      w = rec_ns(rec, c.ns_off, P.g_count_off) == 0 ? 0ull : 1ull;
      break;
    case C_PKEY_VALID: {
      const PartDesc& pt = P.parts[c.idx];
      w = part_null(pt, group_key(rec, gid, P.g_mode, P.g_key_off)) ? 0ull : 1ull;
      break;
    }
    default: {  // C_PKEY: the part's field of the silo's packed key, below its validity bit
      const PartDesc& pt = P.parts[c.idx];
      const GKey k = group_key(rec, gid, P.g_mode, P.g_key_off);
      const uint64_t mask = c.aux >= 64 ? ~0ull : ((1ull << c.aux) - 1);
      w = part_null(pt, k) ? 0ull : ((uint64_t)part_value(pt, k) & mask);
      break;
    }
  }
  return c.desc ? ~w : w;
}

// ---- HAVING ----------------------------------------------------------------------

// A HAVING operand: an integer (i128) or a double, or NULL (valueAsNum /
// havingOutputNum, src/exec/v2_shape_group_topn.zig:2391-2500).
struct HNum {
  __int128 i;
  double f;
  bool is_f;
  bool null;
};

// Output column o of a group as the emitted value: a BIGINT SUM is its
// wrapped 64-bit value (integerAggValue :1448-1452), AVG = f64(sum)/f64(n)
// /10^scale (avgValue :1438-1445), a float MIN/MAX the f64 state before any
// narrowing to FLOAT (rowStateFloat :1465-1468); SUM/AVG/MIN/MAX over no
// non-null input are NULL (aggIsNull :1411-1423).
__device__ __forceinline__ HNum having_value(const FinalParams& P, const OutDesc& o, const uint8_t* rec,
                                             uint32_t gid) {
  HNum v{0, 0.0, false, false};
  const uint64_t count = *reinterpret_cast<const uint64_t*>(rec + P.g_count_off);
  switch (o.kind) {
    case O_KEY: {
      const GKey k = group_key(rec, gid, P.g_mode, P.g_key_off);
      if (part_null(P.parts[o.idx], k)) v.null = true;  // keyPartIsNull (:1475-1479)
      else v.i = part_value(P.parts[o.idx], k);
      return v;
    }
    case O_COUNT: v.i = (__int128)count; return v;
    case O_NS: v.i = (__int128)*reinterpret_cast<const uint64_t*>(rec + o.off); return v;
    default: break;
  }
  const uint64_t ns = o.ns_off != ~0u ? *reinterpret_cast<const uint64_t*>(rec + o.ns_off) : count;
  if (ns == 0) {
    v.null = true;
    return v;
  }
  switch (o.kind) {
    case O_SUM_I64:
    case O_MINMAX_I: v.i = *reinterpret_cast<const int64_t*>(rec + o.off); break;
    case O_SUM_I128: v.i = load_i128(rec + o.off); break;
    case O_SUM_F64: v.is_f = true; v.f = *reinterpret_cast<const double*>(rec + o.off); break;
    case O_AVG_I: v.is_f = true; v.f = i128_to_f64(load_i128(rec + o.off)) / (double)ns / o.scale_div; break;
    case O_AVG_F: v.is_f = true; v.f = *reinterpret_cast<const double*>(rec + o.off) / (double)ns; break;
    default: v.is_f = true; v.f = unord_f64(*reinterpret_cast<const int64_t*>(rec + o.off)); break;  // O_MINMAX_F
  }
  return v;
}

// numOrder (:2409-2422): -1 / 0 / 1; in f64 under types.floatOrder when
// either side is a double.
__device__ __forceinline__ int having_order(const HNum& a, const HavingNode& h) {
  const __int128 b = (__int128)(((unsigned __int128)(uint64_t)h.lit_hi << 64) | (uint64_t)h.lit_lo);
  if (a.is_f || h.lit_float) {
    const uint64_t x = ord_f64(a.is_f ? a.f : i128_to_f64(a.i));
    const uint64_t y = ord_f64(h.lit_float ? h.lit_f64 : i128_to_f64(b));
    return x < y ? -1 : (x > y ? 1 : 0);
  }
  return a.i < b ? -1 : (a.i > b ? 1 : 0);
}

// havingPasses (:2502-2535) over the postfix nodes; the stack is one bit per
// pending entry (at most kMaxHaving).
__device__ __forceinline__ bool having_pass_rec(const FinalParams& P, const HavingNode* __restrict__ H, uint32_t n,
                                               const uint8_t* rec, uint32_t gid) {
  uint64_t st = 0;
  for (uint32_t k = 0; k < n; k++) {
    const HavingNode h = H[k];
    bool r;
    switch (h.kind) {
      case H_CMP: {
        const HNum a = having_value(P, P.outs[h.target], rec, gid);
        if (a.null) {
          r = false;
          break;
        }
        const int o = having_order(a, h);
        switch (h.op) {
          case 0: r = o == 0; break;   // eq
          case 1: r = o != 0; break;   // neq
          case 2: r = o < 0; break;    // lt
          case 3: r = o <= 0; break;   // lte
          case 4: r = o > 0; break;    // gt
          default: r = o >= 0; break;  // gte
        }
        break;
      }
      case H_IS_NULL:
      case H_IS_NOT_NULL: {
        const bool is_null = having_value(P, P.outs[h.target], rec, gid).null;
        r = h.kind == H_IS_NULL ? is_null : !is_null;
        break;
      }
      case H_NOT:
        st ^= 1ull;
        continue;
      default: {  // H_AND / H_OR over the top `arity` entries
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

__device__ __forceinline__ bool having_pass(const FinalParams& P, const HavingNode* __restrict__ H, uint32_t n,
                                           uint32_t gid) {
  return having_pass_rec(P, H, n, P.g_state + (uint64_t)gid * P.g_rec, gid);
}

// ---- result rows -------------------------------------------------------------------

// A record holds a group (its row count is not zero).
__device__ __forceinline__ bool occupied(const OrderParams& P, uint64_t i) {
  return *reinterpret_cast<const uint64_t*>(P.g_state + i * P.g_rec + P.g_count_off) != 0;
}

__device__ __forceinline__ void put(uint8_t* out, uint64_t off, uint32_t bytes, uint64_t row, uint64_t v) {
  uint8_t* p = out + off + row * bytes;
  switch (bytes) {
    case 1: *p = (uint8_t)v; break;
    case 2: *reinterpret_cast<uint16_t*>(p) = (uint16_t)v; break;
    case 4: *reinterpret_cast<uint32_t*>(p) = (uint32_t)v; break;
    default: *reinterpret_cast<uint64_t*>(p) = v; break;
  }
}

// Output row i from record gid, read at `rec` (the group table, or a copy of
// the record): every output column at its offset in `out`.
__device__ __forceinline__ void finalize_rec(const FinalParams& P, const uint8_t* rec, uint32_t gid, uint64_t i,
                                             uint8_t* out) {
  {
    const uint64_t count = *reinterpret_cast<const uint64_t*>(rec + P.g_count_off);
    for (uint32_t c = 0; c < P.n_outs; c++) {
      const OutDesc& o = P.outs[c];
      uint8_t valid = 1;
      switch (o.kind) {
        case O_KEY: {
          const GKey k = group_key(rec, gid, P.g_mode, P.g_key_off);
          if (part_null(P.parts[o.idx], k)) {
            valid = 0;
            put(out, o.data_off, o.bytes, i, 0);
          } else {
            put(out, o.data_off, o.bytes, i, (uint64_t)part_value(P.parts[o.idx], k));
          }
          break;
        }
        case O_COUNT: put(out, o.data_off, 8, i, count); break;
        case O_NS: put(out, o.data_off, 8, i, *reinterpret_cast<const uint64_t*>(rec + o.off)); break;
        default: {
          const uint64_t ns = o.ns_off != ~0u ? *reinterpret_cast<const uint64_t*>(rec + o.ns_off) : count;
          if (ns == 0) {
            valid = 0;
            if (o.bytes == 16) {
              put(out, o.data_off, 8, 2 * i, 0);
              put(out, o.data_off, 8, 2 * i + 1, 0);
            } else {
              put(out, o.data_off, o.bytes, i, 0);
            }
            break;
          }
          switch (o.kind) {
            case O_SUM_I64: put(out, o.data_off, 8, i, *reinterpret_cast<const uint64_t*>(rec + o.off)); break;
            case O_SUM_I128:
              put(out, o.data_off, 8, 2 * i, *reinterpret_cast<const uint64_t*>(rec + o.off));
              put(out, o.data_off, 8, 2 * i + 1, *reinterpret_cast<const uint64_t*>(rec + o.off + 8));
              break;
            case O_SUM_F64:
              put(out, o.data_off, 8, i, *reinterpret_cast<const uint64_t*>(rec + o.off));
              break;
            case O_AVG_I: {
              // AVG = f64(sum) / f64(n) / 10^scale (v2_lowcard_group.zig:2128-2139)
              double a = i128_to_f64(load_i128(rec + o.off)) / (double)ns;
              if (o.scale_div != 1.0) a = a / o.scale_div;
              put(out, o.data_off, 8, i, (uint64_t)__double_as_longlong(a));
              break;
            }
            case O_AVG_F: {
              const double a = *reinterpret_cast<const double*>(rec + o.off) / (double)ns;
              put(out, o.data_off, 8, i, (uint64_t)__double_as_longlong(a));
              break;
            }
            case O_MINMAX_I: put(out, o.data_off, o.bytes, i, *reinterpret_cast<const uint64_t*>(rec + o.off)); break;
            default: {  // O_MINMAX_F
              const double v = unord_f64(*reinterpret_cast<const int64_t*>(rec + o.off));
              if (o.is_f32) {
                const float f = (float)v;
                put(out, o.data_off, 4, i, (uint64_t)(uint32_t)__float_as_uint(f));
              } else {
                put(out, o.data_off, 8, i, (uint64_t)__double_as_longlong(v));
              }
              break;
            }
          }
        }
      }
      if (o.valid_off != ~0ull) out[o.valid_off + i] = valid;
    }
  }
}

// Output row i from record gid of the group table.
__device__ __forceinline__ void finalize_row(const FinalParams& P, uint32_t gid, uint64_t i, uint8_t* out) {
  finalize_rec(P, P.g_state + (uint64_t)gid * P.g_rec, gid, i, out);
}

// A batched plan's record i is a result row: occupied and passing its HAVING.
__device__ __forceinline__ bool batch_keep(const BatchQuery& q, uint64_t i) {
  return occupied(q.op, i) && (q.n_having == 0 || having_pass(q.fp, q.having, q.n_having, (uint32_t)i));
}

}  // namespace tdbg
