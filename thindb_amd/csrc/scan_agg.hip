// scan_agg.hip — the ahead-of-time generic scan/filter/aggregate kernel and
// the record-initialisation kernels.
//
// The generic kernel reads the plan (column kinds, predicate terms, key
// parts, aggregate kinds) from its arguments at run time, so it covers every
// plan the runtime accepts.  Plans run faster through the kernels jit.cpp
// specialises per plan signature; this kernel is the path when run-time
// compilation is unavailable or disabled (THINDB_GPU_JIT=0).
#include "scan_common.h"

namespace tdbg {

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
    if (!mask) return;
    uint64_t key[kVec] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < P.n_parts; i++) {
      const PartDesc& pt = P.parts[i];
      int64_t v[kVec];
      decode_i(fetch<FAST>(P, raw, pt.col, r), P.cols[pt.col].kind, v);
      if (i == 0) key_part<true>(P, pt, v, key, mask);
      else key_part<false>(P, pt, v, key, mask);
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

template <uint32_t LMODE, bool FAST>
__global__ void __launch_bounds__(1024) scan_agg_kernel(const ScanParams P) {
  extern __shared__ __align__(16) uint8_t lds[];
  scan_body<LMODE, GenericStrip<FAST>>(P, lds);
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
