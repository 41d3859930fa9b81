// strip_generic.h — the plan-agnostic strip routine: column kinds,
// predicate terms, key parts and aggregate kinds are read from ScanParams at
// run time.  Used by the ahead-of-time scan kernel (scan_agg.hip) and by the
// batched persistent kernel (batch.hip), whose plans differ item by item.
#pragma once

#include "scan_common.h"

namespace tdbg {

constexpr int kFastCols = 6;  // columns a strip holds in registers

// The strip's columns held in registers (named fields, never an indexed
// array, so they stay in VGPRs).
struct RawSet {
  Raw c0, c1, c2, c3, c4, c5;
};

// Column `c` of the strip: from the registers loaded at the strip's start
// (FAST) or straight from memory.  `c` is uniform across the workgroup, so
// the switch is a scalar branch.
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

template <bool FAST>
struct GenericStrip {
  template <uint32_t LMODE>
  static __device__ __forceinline__ void run(const ScanParams& P, uint8_t* lds, uint64_t r, uint32_t cnt,
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
      mask &= pred_rt(pd, fetch<FAST>(P, raw, pd.col, r), c.kind) & valid4(c, r);
    }
    if (P.n_filt) mask &= filter_mask(P, r, [&](uint32_t col) { return fetch<FAST>(P, raw, col, r); });
    if (!mask) return;
    uint64_t key[kVec] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < P.n_parts; i++) {
      const PartDesc& pt = P.parts[i];
      int64_t v[kVec];
      decode_i(fetch<FAST>(P, raw, pt.col, r), P.cols[pt.col].kind, v);
      const uint32_t pv = pt.nullable ? valid4(P.cols[pt.col], r) : 0xfu;
      if (i == 0) key_part<true>(P, pt, v, key, mask, pv);
      else key_part<false>(P, pt, v, key, mask, pv);
    }
    if (!mask) return;
    int slot[kVec];
    uint8_t* rec[kVec];
    mask = resolve<LMODE>(P, lds, key, mask, part_lo, slot, rec);
    for (uint32_t a = 0; a < P.n_aggs; a++) {
      const AggDesc& d = P.aggs[a];
      const ColDesc& c = P.cols[d.col];
      const uint32_t m = mask & valid4(c, r);
      if (!m) continue;
      agg_rt(P, d, c.kind, fetch<FAST>(P, raw, d.col, r), m, slot, rec, lds);
    }
  }
};

}  // namespace tdbg
