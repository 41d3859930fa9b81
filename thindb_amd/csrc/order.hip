// order.hip — group compaction, ORDER BY / LIMIT selection and result
// emission on gfx950.
//
// Replaces the reference's emit half: emitSorted/SortCtx.lessThan/
// compareBySpec and emitGroup (src/exec/v2_lowcard_group.zig:2069-2311), the
// silo's top-N (src/exec/silo_group_core.zig:4856-5210, adapter
// src/exec/v2_shape_group_topn.zig:1095-1180, :1399-1470) and the bounded
// exec.TopN (src/exec/topn.zig:1-60).
//
// Order semantics: every ORDER BY item becomes one to three 64-bit words
// whose unsigned order is the reference's order for that item (desc =
// bitwise NOT), followed by the silo's tie-break, the packed group key
// ascending (rowOrderBefore, src/exec/v2_shape_group_topn.zig:1327-1335): as
// words of their own, or - where the slot order already is that order (a
// dense table over key parts whose values share one sign) - the group's
// position in slot order.  The order is total, so the page at a LIMIT cut is
// the reference silo's page.
#include <cstdlib>
#include <vector>

#include "order_device.h"

namespace tdbg {

// ---- compaction: occupied records in slot order --------------------------------

constexpr uint32_t kCompactThreads = 256;
constexpr uint32_t kCompactPer = 16;
constexpr uint32_t kCompactChunk = kCompactThreads * kCompactPer;

__global__ void __launch_bounds__(kCompactThreads) compact_count_kernel(const OrderParams P, uint64_t n_rec,
                                                                        uint32_t* block_counts) {
  __shared__ uint32_t red[kCompactThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kCompactChunk + (uint64_t)threadIdx.x * kCompactPer;
  uint32_t c = 0;
  for (uint32_t j = 0; j < kCompactPer; j++)
    if (base + j < n_rec && occupied(P, base + j)) c++;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t w = 0; w < kCompactThreads / 64; w++) t += red[w];
    block_counts[blockIdx.x] = t;
  }
}

// Exclusive scan of `n` u32 counts by one workgroup; total -> *total.
__global__ void __launch_bounds__(1024) scan_counts_kernel(uint32_t* counts, uint32_t n, uint64_t* total) {
  __shared__ uint64_t part[1024];
  const uint32_t per = (n + 1023) / 1024;
  const uint32_t lo = threadIdx.x * per, hi = lo + per < n ? lo + per : n;
  uint64_t s = 0;
  for (uint32_t i = lo; i < hi; i++) s += counts[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t o = 1; o < 1024; o <<= 1) {
    uint64_t v = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint64_t run = part[threadIdx.x] - s;
  for (uint32_t i = lo; i < hi; i++) {
    uint32_t c = counts[i];
    counts[i] = (uint32_t)run;
    run += c;
  }
  if (threadIdx.x == 1023) *total = part[1023];
}

__global__ void __launch_bounds__(kCompactThreads) compact_write_kernel(const OrderParams P, uint64_t n_rec,
                                                                        const uint32_t* block_offs,
                                                                        uint32_t* gids) {
  __shared__ uint32_t pre[kCompactThreads];
  const uint64_t base = (uint64_t)blockIdx.x * kCompactChunk + (uint64_t)threadIdx.x * kCompactPer;
  uint32_t bits = 0, c = 0;
  for (uint32_t j = 0; j < kCompactPer; j++)
    if (base + j < n_rec && occupied(P, base + j)) { bits |= 1u << j; c++; }
  pre[threadIdx.x] = c;
  __syncthreads();
  for (uint32_t o = 1; o < kCompactThreads; o <<= 1) {
    uint32_t v = threadIdx.x >= o ? pre[threadIdx.x - o] : 0;
    __syncthreads();
    pre[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t out = block_offs[blockIdx.x] + pre[threadIdx.x] - c;
  for (uint32_t j = 0; j < kCompactPer; j++)
    if ((bits >> j) & 1u) gids[out++] = (uint32_t)(base + j);
}

hipError_t launch_compact(const OrderParams& p, uint64_t n_rec, uint32_t* block_counts, uint32_t* gids,
                          uint64_t* n_out, hipStream_t s) {
  const uint32_t blocks = (uint32_t)((n_rec + kCompactChunk - 1) / kCompactChunk);
  if (blocks == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(compact_count_kernel, dim3(blocks), dim3(kCompactThreads), 0, s, p, n_rec, block_counts);
  hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, s, block_counts, blocks, n_out);
  hipLaunchKernelGGL(compact_write_kernel, dim3(blocks), dim3(kCompactThreads), 0, s, p, n_rec,
                     (const uint32_t*)block_counts, gids);
  return hipGetLastError();
}

// ---- HAVING: survivors of the compacted positions, in order ----------------------
// (the reference filters the grouped rows before its final sort,
// src/exec/v2_shape_group_topn.zig:672 applyHaving, :2537 havingEmitPass)
//
// A workgroup owns kCompactChunk consecutive positions, visited in
// kCompactPer rounds of 256 lanes so that position, gid and record loads are
// coalesced.  Pass 1 evaluates the predicate once per group and keeps one
// 64-bit pass mask per (round, wavefront); pass 2 writes the survivors' gids
// at their ordered positions from the masks alone.

__global__ void __launch_bounds__(kCompactThreads) having_mask_kernel(const FinalParams P,
                                                                      const HavingNode* __restrict__ H,
                                                                      uint32_t n_nodes,
                                                                      const uint32_t* __restrict__ gids,
                                                                      const uint64_t* __restrict__ n_in,
                                                                      uint64_t max_n, uint64_t* __restrict__ masks,
                                                                      uint32_t* __restrict__ block_counts) {
  __shared__ uint32_t red[kCompactThreads / 64];
  const uint64_t n = *n_in < max_n ? *n_in : max_n;
  const uint64_t base = (uint64_t)blockIdx.x * kCompactChunk;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t c = 0;
  if (base < n) {  // uniform per workgroup
    for (uint32_t r = 0; r < kCompactPer; r++) {
      const uint64_t i = base + r * kCompactThreads + threadIdx.x;
      const bool ok = i < n && having_pass(P, H, n_nodes, gids[i]);
      const uint64_t m = __ballot(ok);
      if (lane == 0) masks[(base >> 6) + r * (kCompactThreads / 64) + wave] = m;
      c += (uint32_t)__popcll(m);
    }
  }
  if (lane == 0) red[wave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t w = 0; w < kCompactThreads / 64; w++) t += red[w];
    block_counts[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kCompactThreads) having_write_kernel(const uint32_t* __restrict__ gids,
                                                                       const uint64_t* __restrict__ n_in,
                                                                       uint64_t max_n,
                                                                       const uint64_t* __restrict__ masks,
                                                                       const uint32_t* __restrict__ block_offs,
                                                                       uint32_t* __restrict__ out) {
  constexpr uint32_t kMasks = kCompactChunk / 64;  // 64 masks per workgroup
  __shared__ uint64_t mk[kMasks];
  __shared__ uint32_t pre[kMasks];
  const uint64_t n = *n_in < max_n ? *n_in : max_n;
  const uint64_t base = (uint64_t)blockIdx.x * kCompactChunk;
  if (base >= n) return;  // uniform per workgroup
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x < kMasks) {  // wavefront 0: exclusive scan of the masks' counts
    const uint64_t m = masks[(base >> 6) + threadIdx.x];
    const uint32_t c = (uint32_t)__popcll(m);
    uint32_t x = c;
    for (uint32_t o = 1; o < 64; o <<= 1) {
      const uint32_t v = __shfl_up(x, o);
      if (lane >= o) x += v;
    }
    mk[threadIdx.x] = m;
    pre[threadIdx.x] = x - c;
  }
  __syncthreads();
  const uint32_t off = block_offs[blockIdx.x];
  const uint64_t below = (1ull << lane) - 1;
  for (uint32_t r = 0; r < kCompactPer; r++) {
    const uint32_t k = r * (kCompactThreads / 64) + wave;
    const uint64_t m = mk[k];
    if ((m >> lane) & 1ull)
      out[off + pre[k] + (uint32_t)__popcll(m & below)] = gids[base + r * kCompactThreads + threadIdx.x];
  }
}

size_t having_scratch_bytes(uint64_t max_n) {
  const uint64_t blocks = (max_n + kCompactChunk - 1) / kCompactChunk + 1;
  return blocks * (kCompactChunk / 64) * 8 + blocks * 4 + 256;
}

hipError_t launch_having(const FinalParams& p, const HavingNode* nodes, uint32_t n_nodes, const uint32_t* gids_in,
                         const uint64_t* n_in, uint64_t max_n, void* scratch, uint32_t* gids_out, uint64_t* n_kept,
                         hipStream_t s) {
  static_assert(kCompactChunk / 64 == 64, "one wavefront scans a workgroup's masks");
  const uint64_t nb = (max_n + kCompactChunk - 1) / kCompactChunk;
  const uint32_t blocks = (uint32_t)(nb ? nb : 1);
  if (n_nodes == 0 || n_nodes > kMaxHaving || !nodes) return hipErrorInvalidValue;
  uint64_t* masks = static_cast<uint64_t*>(scratch);
  uint32_t* counts = reinterpret_cast<uint32_t*>(masks + (uint64_t)blocks * (kCompactChunk / 64));
  hipLaunchKernelGGL(having_mask_kernel, dim3(blocks), dim3(kCompactThreads), 0, s, p, nodes, n_nodes, gids_in, n_in,
                     max_n, masks, counts);
  hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, s, counts, blocks, n_kept);
  hipLaunchKernelGGL(having_write_kernel, dim3(blocks), dim3(kCompactThreads), 0, s, gids_in, n_in, max_n,
                     (const uint64_t*)masks, (const uint32_t*)counts, gids_out);
  return hipGetLastError();
}

// ---- order words ---------------------------------------------------------------

__global__ void make_comps_kernel(const OrderParams P, const uint32_t* __restrict__ gids, uint64_t n,
                                  uint64_t* comps) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t gid = gids[i];
    for (uint32_t c = 0; c < P.n_comps; c++) comps[c * n + i] = comp_word(P, P.comps[c], gid);
  }
}

// Composite order of positions a, b: words 0..n_comps-1, then position.
__device__ __forceinline__ bool pos_less(const uint64_t* comps, uint32_t n_comps, uint64_t n, uint32_t a,
                                         uint64_t wa, uint32_t b, uint64_t wb) {
  if (wa != wb) return wa < wb;
  for (uint32_t c = 1; c < n_comps; c++) {
    const uint64_t xa = comps[c * n + a], xb = comps[c * n + b];
    if (xa != xb) return xa < xb;
  }
  return a < b;
}

// ---- small sort: one workgroup, bitonic over (word0, position) in LDS -----------

constexpr uint32_t kSmallSort = 4096;

// Composite order from word `first` on (the words before it are equal for
// every group): wa/wb are the two positions' word `first`.
__device__ __forceinline__ bool pos_less_from(const uint64_t* comps, uint32_t first, uint32_t n_comps, uint64_t n,
                                              uint32_t a, uint64_t wa, uint32_t b, uint64_t wb) {
  if (wa != wb) return wa < wb;
  for (uint32_t c = first + 1; c < n_comps; c++) {
    const uint64_t xa = comps[c * n + a], xb = comps[c * n + b];
    if (xa != xb) return xa < xb;
  }
  return a < b;
}

// Sorts n <= kSmallSort positions (pos_in[i], or base + i) by the composite
// order in LDS, from word `first` on; writes the first n_out to pos_out.
__device__ __forceinline__ void sort_block(const uint64_t* comps, uint32_t first, uint32_t n_comps, uint64_t n_all,
                                          const uint32_t* pos_in, uint32_t base, uint32_t n, uint32_t* pos_out,
                                          uint32_t n_out, uint32_t* sp, uint64_t* sw) {
  uint32_t m = 1;
  while (m < n) m <<= 1;
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
    if (i < n) {
      const uint32_t p = pos_in ? pos_in[i] : base + i;
      sp[i] = p;
      sw[i] = first < n_comps ? comps[(uint64_t)first * n_all + p] : 0ull;
    } else {
      sp[i] = 0xffffffffu;
      sw[i] = ~0ull;
    }
  }
  __syncthreads();
  for (uint32_t k = 2; k <= m; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        const uint32_t l = i ^ j;
        if (l > i) {
          const bool up = (i & k) == 0;
          const uint32_t pa = sp[i], pb = sp[l];
          bool less;  // element l < element i
          if (pb == 0xffffffffu) less = false;
          else if (pa == 0xffffffffu) less = true;
          else less = pos_less_from(comps, first, n_comps, n_all, pb, sw[l], pa, sw[i]);
          if (less == up) {
            sp[i] = pb; sp[l] = pa;
            const uint64_t t = sw[i]; sw[i] = sw[l]; sw[l] = t;
          }
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t i = threadIdx.x; i < n_out; i += blockDim.x) pos_out[i] = sp[i];
}

// The same for n <= blockDim.x positions, one per thread in registers: the
// compare-exchange steps within a wavefront (partner distance < 64) go
// through shuffles, only the wider ones through LDS, so a 1024-position sort
// waits at 20 barriers instead of 55.
__device__ __forceinline__ void sort_block_reg(const uint64_t* comps, uint32_t first, uint32_t n_comps, uint64_t n_all,
                                              const uint32_t* pos_in, uint32_t base, uint32_t n, uint32_t* pos_out,
                                              uint32_t n_out, uint32_t* sp, uint64_t* sw) {
  uint32_t m = 1;
  while (m < n) m <<= 1;
  const uint32_t i = threadIdx.x;
  uint32_t p = 0xffffffffu;  // none: after every position
  uint64_t w = ~0ull;
  if (i < n) {
    p = pos_in ? pos_in[i] : base + i;
    w = first < n_comps ? comps[(uint64_t)first * n_all + p] : 0ull;
  }
  for (uint32_t k = 2; k <= m; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      uint32_t q;
      uint64_t v;
      if (j >= 64) {  // uniform
        __syncthreads();
        sp[i] = p;
        sw[i] = w;
        __syncthreads();
        q = sp[i ^ j];
        v = sw[i ^ j];
      } else {
        q = __shfl_xor(p, (int)j);
        const uint32_t vh = __shfl_xor((uint32_t)(w >> 32), (int)j), vl = __shfl_xor((uint32_t)w, (int)j);
        v = ((uint64_t)vh << 32) | vl;
      }
      // the lower index of a pair keeps the smaller element in an ascending
      // run (i & k == 0), the larger in a descending one
      const bool keep_min = ((i & j) == 0) == ((i & k) == 0);
      bool take;
      if (q == p) take = false;  // both none
      else if (q == 0xffffffffu) take = !keep_min;
      else if (p == 0xffffffffu) take = keep_min;
      else take = keep_min ? pos_less_from(comps, first, n_comps, n_all, q, v, p, w)
                           : pos_less_from(comps, first, n_comps, n_all, p, w, q, v);
      if (take) {
        p = q;
        w = v;
      }
    }
  }
  __syncthreads();
  sp[i] = p;
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < n_out; t += blockDim.x) pos_out[t] = sp[t];
}

__global__ void __launch_bounds__(1024) small_sort_kernel(const uint64_t* comps, uint32_t n_comps, uint64_t n_all,
                                                          const uint32_t* pos_in, uint32_t n, uint32_t* pos_out) {
  __shared__ uint32_t sp[kSmallSort];
  __shared__ uint64_t sw[kSmallSort];
  if (n <= blockDim.x) sort_block_reg(comps, 0, n_comps, n_all, pos_in, 0, n, pos_out, n, sp, sw);
  else sort_block(comps, 0, n_comps, n_all, pos_in, 0, n, pos_out, n, sp, sw);
}

// The first word whose min and max (mm[2c], mm[2c + 1]) differ, or n_comps.
__device__ __forceinline__ uint32_t first_varying(const uint64_t* mm, uint32_t n_comps) {
  uint32_t c = 0;
  while (c < n_comps && mm[2 * c] == mm[2 * c + 1]) c++;
  return c;
}

// Workgroup b sorts positions [b * chunk, min(n, (b + 1) * chunk)) (chunk
// <= kSmallSort) from the first varying word on (mm: every word's min and
// max) and keeps its first `keep` at out[b * keep ...].
__global__ void __launch_bounds__(1024) chunk_sort_kernel(const uint64_t* comps, uint32_t n_comps, uint64_t n_all,
                                                          const uint64_t* mm, uint32_t n, uint32_t chunk,
                                                          uint32_t keep, uint32_t* out) {
  __shared__ uint32_t sp[kSmallSort];
  __shared__ uint64_t sw[kSmallSort];
  const uint32_t base = blockIdx.x * chunk;
  const uint32_t len = n - base < chunk ? n - base : chunk;
  if (len <= blockDim.x)
    sort_block_reg(comps, first_varying(mm, n_comps), n_comps, n_all, nullptr, base, len,
                   out + (uint64_t)blockIdx.x * keep, keep < len ? keep : len, sp, sw);
  else
    sort_block(comps, first_varying(mm, n_comps), n_comps, n_all, nullptr, base, len,
               out + (uint64_t)blockIdx.x * keep, keep < len ? keep : len, sp, sw);
}

// The candidates pos_in[0..n) sorted the same way.
__global__ void __launch_bounds__(1024) cand_sort_kernel(const uint64_t* comps, uint32_t n_comps, uint64_t n_all,
                                                         const uint64_t* mm, const uint32_t* pos_in, uint32_t n,
                                                         uint32_t* pos_out) {
  __shared__ uint32_t sp[kSmallSort];
  __shared__ uint64_t sw[kSmallSort];
  if (n <= blockDim.x) sort_block_reg(comps, first_varying(mm, n_comps), n_comps, n_all, pos_in, 0, n, pos_out, n, sp, sw);
  else sort_block(comps, first_varying(mm, n_comps), n_comps, n_all, pos_in, 0, n, pos_out, n, sp, sw);
}

// ---- per-workgroup top-k (k <= kTopMaxK) -----------------------------------------

constexpr uint32_t kTopThreads = 256;
constexpr uint32_t kTopPer = 16;
constexpr uint32_t kTopChunk = kTopThreads * kTopPer;
constexpr uint32_t kTopMaxK = 256;
constexpr uint32_t kChunkMax = 16;  // chunk path: up to 16 chunks of kSmallSort groups (16 x kTopMaxK candidates)
constexpr uint32_t kNone = 0xffffffffu;

// Each workgroup takes kTopChunk items of the position list `pos_in`
// (the identity when null, m items) and writes its k best, in order, to
// cand[b*k ...].  Items stay in registers; each round is one block-wide
// argmin under the composite order.  Every workgroup but the last holds at
// least k items, so the candidates land contiguously.
__global__ void __launch_bounds__(kTopThreads) block_topk_kernel(const uint64_t* comps, uint32_t n_comps,
                                                                 uint64_t n_all, const uint32_t* pos_in,
                                                                 uint64_t m, uint32_t k, uint32_t* cand) {
  __shared__ uint64_t rw[kTopThreads / 64];
  __shared__ uint32_t rp[kTopThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kTopChunk + threadIdx.x;
  uint64_t w[kTopPer];
  uint32_t p[kTopPer];
#pragma unroll
  for (uint32_t j = 0; j < kTopPer; j++) {
    const uint64_t i = base + (uint64_t)j * kTopThreads;
    p[j] = kNone;
    w[j] = ~0ull;
    if (i < m) {
      p[j] = pos_in ? pos_in[i] : (uint32_t)i;
      w[j] = comps[p[j]];
    }
  }
  for (uint32_t r = 0; r < k; r++) {
    uint64_t bw = ~0ull;
    uint32_t bp = kNone;
#pragma unroll
    for (uint32_t j = 0; j < kTopPer; j++)
      if (p[j] != kNone && (bp == kNone || pos_less(comps, n_comps, n_all, p[j], w[j], bp, bw))) {
        bw = w[j];
        bp = p[j];
      }
    for (int o = 32; o > 0; o >>= 1) {
      const uint64_t ow = __shfl_xor(bw, o);
      const uint32_t op = __shfl_xor(bp, o);
      if (op != kNone && (bp == kNone || pos_less(comps, n_comps, n_all, op, ow, bp, bw))) {
        bw = ow;
        bp = op;
      }
    }
    if ((threadIdx.x & 63) == 0) {
      rw[threadIdx.x >> 6] = bw;
      rp[threadIdx.x >> 6] = bp;
    }
    __syncthreads();
    uint64_t fw = rw[0];
    uint32_t fp = rp[0];
    for (uint32_t x = 1; x < kTopThreads / 64; x++)
      if (rp[x] != kNone && (fp == kNone || pos_less(comps, n_comps, n_all, rp[x], rw[x], fp, fw))) {
        fw = rw[x];
        fp = rp[x];
      }
    __syncthreads();
    if (fp == kNone) break;  // uniform: the workgroup ran out of items
    if (threadIdx.x == 0) cand[(uint64_t)blockIdx.x * k + r] = fp;
#pragma unroll
    for (uint32_t j = 0; j < kTopPer; j++)
      if (p[j] == fp) p[j] = kNone;  // positions are unique: only the owner matches
  }
}

// ---- one-workgroup radix select (n <= kSelect1) ---------------------------------------

constexpr uint64_t kSelect1 = 1u << 16;

// Exclusive prefix sum over the 1024 threads of a workgroup; *total gets the sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* tmp, uint32_t* total) {
  const uint32_t t = threadIdx.x;
  tmp[t] = v;
  __syncthreads();
  for (uint32_t o = 1; o < 1024; o <<= 1) {
    const uint32_t x = t >= o ? tmp[t - o] : 0;
    __syncthreads();
    tmp[t] += x;
    __syncthreads();
  }
  const uint32_t incl = tmp[t];
  *total = tmp[1023];
  __syncthreads();
  return incl - v;
}

// The `need` best positions by the first order word: radix-selects the
// need-th smallest word T (12-bit digits over the word's [min, max] range),
// then gathers, in position order, every position whose word is <= T —
// the best `need` plus any ties at T — into out[0 .. *m).
__global__ void __launch_bounds__(1024) select1_kernel(const uint64_t* __restrict__ w0, uint32_t n, const uint64_t* mm,
                                                       uint32_t need, uint32_t* out, uint32_t* m_out) {
  __shared__ uint32_t hist[4096];
  __shared__ uint32_t tmp[1024];
  __shared__ uint32_t sh_digit, sh_before;
  const uint32_t t = threadIdx.x;
  const uint64_t lo = mm[0], range = mm[1] - mm[0];
  const int bits = range == 0 ? 0 : 64 - __clzll(range);
  uint64_t prefix = 0;
  uint32_t remaining = need;  // 1-based rank inside the current prefix class
  for (int shift = (bits + 11) / 12 * 12 - 12; shift >= 0; shift -= 12) {
    for (uint32_t i = t; i < 4096; i += 1024) hist[i] = 0;
    __syncthreads();
    const int hs = shift + 12;
    for (uint32_t i = t; i < n; i += 1024) {
      const uint64_t v = w0[i] - lo;
      if (hs >= 64 || (v >> hs) == (prefix >> hs)) atomicAdd(&hist[(v >> shift) & 4095u], 1u);
    }
    __syncthreads();
    const uint32_t h0 = hist[4 * t], h1 = hist[4 * t + 1], h2 = hist[4 * t + 2], h3 = hist[4 * t + 3];
    uint32_t total;
    uint32_t before = block_excl_scan(h0 + h1 + h2 + h3, tmp, &total);
    if (before < remaining && before + h0 + h1 + h2 + h3 >= remaining) {
      uint32_t d = 4 * t;
      if (before + h0 < remaining) {
        before += h0;
        d++;
        if (before + h1 < remaining) {
          before += h1;
          d++;
          if (before + h2 < remaining) {
            before += h2;
            d++;
          }
        }
      }
      sh_digit = d;
      sh_before = before;
    }
    __syncthreads();
    prefix |= (uint64_t)sh_digit << shift;
    remaining -= sh_before;
    __syncthreads();
  }
  const uint64_t T = prefix + lo;
  uint32_t base = 0;
  for (uint32_t c = 0; c < n; c += 1024) {
    const uint32_t i = c + t;
    const uint32_t keep = (i < n && w0[i] <= T) ? 1u : 0u;
    uint32_t total;
    const uint32_t at = block_excl_scan(keep, tmp, &total);
    if (keep) out[base + at] = i;
    base += total;
  }
  if (t == 0) *m_out = base;
}

// ---- stable LSD radix sort of positions by the composite order --------------------

constexpr uint32_t kRsThreads = 256;
constexpr uint32_t kRsPer = 8;
constexpr uint32_t kRsTile = kRsThreads * kRsPer;

__global__ void gather_keys_kernel(const uint64_t* comp, const uint32_t* pos, uint64_t n, uint64_t* keys) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    keys[i] = comp[pos[i]];
}

__global__ void minmax_kernel(const uint64_t* keys, uint64_t n, uint64_t* mm) {
  uint64_t lo = ~0ull, hi = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = keys[i];
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t a = __shfl_xor(lo, o), b = __shfl_xor(hi, o);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  // one device atomic pair per workgroup (the two words are contended)
  __shared__ uint64_t wl[16], wh[16];
  if ((threadIdx.x & 63) == 0) {
    wl[threadIdx.x >> 6] = lo;
    wh[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < (blockDim.x + 63) / 64; w++) {
      lo = wl[w] < lo ? wl[w] : lo;
      hi = wh[w] > hi ? wh[w] : hi;
    }
    atomicMin((unsigned long long*)&mm[0], (unsigned long long)lo);
    atomicMax((unsigned long long*)&mm[1], (unsigned long long)hi);
  }
}

__global__ void init_mm_kernel(uint64_t* mm) {
  mm[0] = ~0ull;
  mm[1] = 0;
}

__device__ __forceinline__ uint32_t digit_of(uint64_t k, const uint64_t* mm, uint32_t shift) {
  return (uint32_t)(((k - mm[0]) >> shift) & 0xffu);
}

// hist[d * n_blocks + b] = number of digit d in tile b.  Skipped (all zero)
// when the key range needs no bits at this shift.
__global__ void __launch_bounds__(kRsThreads) rs_hist_kernel(const uint64_t* keys, uint64_t n, const uint64_t* mm,
                                                             uint32_t shift, uint32_t* hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kRsTile;
  for (uint32_t j = 0; j < kRsPer; j++) {
    const uint64_t i = base + (uint64_t)j * kRsThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[digit_of(keys[i], mm, shift)], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// Generic exclusive scan of u32 in three phases (chunk sums, scan, apply).
constexpr uint32_t kScanChunk = 2048;
__global__ void scan_chunk_sums_kernel(const uint32_t* a, uint64_t n, uint32_t* sums) {
  __shared__ uint32_t red[4];
  const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
  uint32_t s = 0;
  for (uint32_t j = threadIdx.x; j < kScanChunk; j += 256)
    if (base + j < n) s += a[base + j];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) sums[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
__global__ void __launch_bounds__(256) scan_apply_kernel(uint32_t* a, uint64_t n, const uint32_t* offs) {
  constexpr uint32_t per = kScanChunk / 256;
  __shared__ uint32_t pre[256];
  const uint64_t base = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * per;
  uint32_t v[per];
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t j = 0; j < per; j++) {
    v[j] = base + j < n ? a[base + j] : 0;
    sum += v[j];
  }
  pre[threadIdx.x] = sum;
  __syncthreads();
  for (uint32_t o = 1; o < 256; o <<= 1) {
    const uint32_t x = threadIdx.x >= o ? pre[threadIdx.x - o] : 0;
    __syncthreads();
    pre[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t run = offs[blockIdx.x] + pre[threadIdx.x] - sum;
#pragma unroll
  for (uint32_t j = 0; j < per; j++) {
    if (base + j < n) a[base + j] = run;
    run += v[j];
  }
}

// Stable scatter: round j handles items j*256 + t, so (round, wave, lane)
// order is item order; equal digits keep their relative order.
__global__ void __launch_bounds__(kRsThreads) rs_scatter_kernel(const uint64_t* keys_in, const uint32_t* pos_in,
                                                                uint64_t n, const uint64_t* mm, uint32_t shift,
                                                                const uint32_t* offs, uint64_t* keys_out,
                                                                uint32_t* pos_out) {
  __shared__ uint32_t wave_cnt[kRsThreads / 64][256];
  __shared__ uint32_t wave_off[kRsThreads / 64][256];
  __shared__ uint32_t run[256];
  const uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63;
  for (uint32_t x = 0; x < kRsThreads / 64; x++) wave_cnt[x][t] = 0;
  run[t] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kRsTile;
  const uint64_t lt = (1ull << lane) - 1ull;
  for (uint32_t j = 0; j < kRsPer; j++) {
    const uint64_t i = base + (uint64_t)j * kRsThreads + t;
    const bool live = i < n;
    uint64_t k = 0;
    uint32_t p = 0, d = 0;
    if (live) {
      k = keys_in[i];
      p = pos_in[i];
      d = digit_of(k, mm, shift);
    }
    uint64_t m = __ballot(live);
    for (uint32_t b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      m &= bit ? bal : ~bal;
    }
    const uint32_t rank = __popcll(m & lt);
    if (live && rank == 0) wave_cnt[w][d] = __popcll(m);
    __syncthreads();
    {
      uint32_t o = run[t];
      for (uint32_t x = 0; x < kRsThreads / 64; x++) {
        const uint32_t c = wave_cnt[x][t];
        wave_off[x][t] = o;
        o += c;
        wave_cnt[x][t] = 0;
      }
      run[t] = o;
    }
    __syncthreads();
    if (live) {
      const uint64_t dst = (uint64_t)offs[(uint64_t)d * gridDim.x + blockIdx.x] + wave_off[w][d] + rank;
      keys_out[dst] = k;
      pos_out[dst] = p;
    }
    __syncthreads();
  }
}

static inline uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap) {
  uint64_t g = (n + block - 1) / block;
  if (g > cap) g = cap;
  return g == 0 ? 1 : (uint32_t)g;
}

// Scratch layout for the sorter.
struct SortScratch {
  uint64_t* comps;     // n_comps * n
  uint64_t* keys_a;    // n
  uint64_t* keys_b;    // n
  uint32_t* pos_a;     // n
  uint32_t* pos_b;     // n
  uint32_t* hist;      // 256 * tiles
  uint32_t* sums;      // hist chunks
  uint32_t* sums2;     // chunk sums of sums (one scan level more)
  uint64_t* mm;        // min and max per order word, then a u32 count at word 2 * kMaxComps
  uint64_t* total;     // 1 word
};

static inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t sort_scratch_bytes(uint64_t n, uint32_t n_comps) {
  const uint64_t tiles = (n + kRsTile - 1) / kRsTile + 1;
  const uint64_t hist = 256 * tiles;
  const uint64_t chunks = (hist + kScanChunk - 1) / kScanChunk + 1;
  return al256(n_comps * n * 8) + 2 * al256(n * 8) + 2 * al256(n * 4) + al256(hist * 4) + al256(chunks * 4) +
         al256(((chunks + kScanChunk - 1) / kScanChunk + 1) * 4) + 4 * 256;
}

static SortScratch carve(void* base, uint64_t n, uint32_t n_comps) {
  const uint64_t tiles = (n + kRsTile - 1) / kRsTile + 1;
  const uint64_t hist = 256 * tiles;
  const uint64_t chunks = (hist + kScanChunk - 1) / kScanChunk + 1;
  uint8_t* p = static_cast<uint8_t*>(base);
  SortScratch s;
  s.comps = reinterpret_cast<uint64_t*>(p); p += al256(n_comps * n * 8);
  s.keys_a = reinterpret_cast<uint64_t*>(p); p += al256(n * 8);
  s.keys_b = reinterpret_cast<uint64_t*>(p); p += al256(n * 8);
  s.pos_a = reinterpret_cast<uint32_t*>(p); p += al256(n * 4);
  s.pos_b = reinterpret_cast<uint32_t*>(p); p += al256(n * 4);
  s.hist = reinterpret_cast<uint32_t*>(p); p += al256(hist * 4);
  s.sums = reinterpret_cast<uint32_t*>(p); p += al256(chunks * 4);
  s.sums2 = reinterpret_cast<uint32_t*>(p); p += al256(((chunks + kScanChunk - 1) / kScanChunk + 1) * 4);
  s.mm = reinterpret_cast<uint64_t*>(p); p += 512;  // 2 words per order word, then a count
  s.total = reinterpret_cast<uint64_t*>(p);
  return s;
}

// Exclusive scan of a u32 array of length m (m <= 2048^3).
static void scan_u32(uint32_t* a, uint64_t m, uint32_t* sums, uint32_t* sums2, uint64_t* total, hipStream_t s) {
  const uint32_t chunks = (uint32_t)((m + kScanChunk - 1) / kScanChunk);
  hipLaunchKernelGGL(scan_chunk_sums_kernel, dim3(chunks), dim3(256), 0, s, a, m, sums);
  if (chunks <= 1024 * 64) {
    hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, s, sums, chunks, total);
  } else {
    const uint32_t c2 = (chunks + kScanChunk - 1) / kScanChunk;
    hipLaunchKernelGGL(scan_chunk_sums_kernel, dim3(c2), dim3(256), 0, s, sums, (uint64_t)chunks, sums2);
    hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, s, sums2, c2, total);
    hipLaunchKernelGGL(scan_apply_kernel, dim3(c2), dim3(256), 0, s, sums, (uint64_t)chunks, (const uint32_t*)sums2);
  }
  hipLaunchKernelGGL(scan_apply_kernel, dim3(chunks), dim3(256), 0, s, a, m, (const uint32_t*)sums);
}

// Sorts m positions (pos_a) by the composite order; result in pos_a.
static hipError_t radix_sort_positions(SortScratch& S, uint64_t n_all, uint32_t n_comps, uint64_t m,
                                       hipStream_t s) {
  const uint32_t tiles = (uint32_t)((m + kRsTile - 1) / kRsTile);
  for (int c = (int)n_comps - 1; c >= 0; c--) {
    const uint64_t* comp = S.comps + (uint64_t)c * n_all;
    hipLaunchKernelGGL(gather_keys_kernel, dim3(grid_for(m, 256, 8192)), dim3(256), 0, s, comp,
                       (const uint32_t*)S.pos_a, m, S.keys_a);
    hipLaunchKernelGGL(init_mm_kernel, dim3(1), dim3(1), 0, s, S.mm);
    hipLaunchKernelGGL(minmax_kernel, dim3(grid_for(m, 256, 1024)), dim3(256), 0, s, (const uint64_t*)S.keys_a, m,
                       S.mm);
    uint64_t mm[2];
    hipError_t e = hipMemcpyAsync(mm, S.mm, sizeof mm, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    const uint64_t range = mm[1] - mm[0];
    int bits = range == 0 ? 0 : 64 - __builtin_clzll(range);
    for (int shift = 0; shift < bits; shift += 8) {
      hipLaunchKernelGGL(rs_hist_kernel, dim3(tiles), dim3(kRsThreads), 0, s, (const uint64_t*)S.keys_a, m,
                         (const uint64_t*)S.mm, (uint32_t)shift, S.hist);
      scan_u32(S.hist, (uint64_t)tiles * 256, S.sums, S.sums2, S.total, s);
      hipLaunchKernelGGL(rs_scatter_kernel, dim3(tiles), dim3(kRsThreads), 0, s, (const uint64_t*)S.keys_a,
                         (const uint32_t*)S.pos_a, m, (const uint64_t*)S.mm, (uint32_t)shift,
                         (const uint32_t*)S.hist, S.keys_b, S.pos_b);
      uint64_t* tk = S.keys_a; S.keys_a = S.keys_b; S.keys_b = tk;
      uint32_t* tp = S.pos_a; S.pos_a = S.pos_b; S.pos_b = tp;
    }
  }
  return hipGetLastError();
}

__global__ void iota_kernel(uint32_t* p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

hipError_t launch_iota(uint32_t* p, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

__global__ void select_gids_kernel(const uint32_t* gids, const uint32_t* pos, uint64_t k, uint32_t* out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = gids[pos[i]];
}

// ---- selection over many groups: the cut of the first word -------------------------
// For ORDER BY ... LIMIT over more groups than one workgroup handles: only
// the first order word is built for every group; the value of the
// need-th smallest word is found 12 bits at a time (multi-workgroup
// histograms), every group at or below it is gathered in slot order, and
// only those candidates get their full order words and a sort.

constexpr uint32_t kSelBins = 4096;
constexpr uint32_t kSelBlocks = 512;

// w[i] = order word c of group gids[i]; its min and max into mm (which
// init_mm_kernel set).
__global__ void __launch_bounds__(256) sel_word_kernel(const OrderParams P, uint32_t c,
                                                       const uint32_t* __restrict__ gids, uint64_t n, uint64_t* w,
                                                       uint64_t* mm) {
  uint64_t lo = ~0ull, hi = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = comp_word(P, P.comps[c], gids[i]);
    w[i] = v;
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t a = __shfl_xor(lo, o), b = __shfl_xor(hi, o);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  __shared__ uint64_t wl[4], wh[4];
  if ((threadIdx.x & 63) == 0) {
    wl[threadIdx.x >> 6] = lo;
    wh[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t k = 1; k < 4; k++) {
      lo = wl[k] < lo ? wl[k] : lo;
      hi = wh[k] > hi ? wh[k] : hi;
    }
    atomicMin((unsigned long long*)&mm[0], (unsigned long long)lo);
    atomicMax((unsigned long long*)&mm[1], (unsigned long long)hi);
  }
}

// Per-workgroup histogram of digit (rel >> shift) & mask over the words
// whose bits at and above `above` equal `prefix` (rel = w - min).  A thread
// adds a run of equal digits with one LDS atomic, so a skewed digit (most
// groups share one value) does not serialise on one counter.
__global__ void __launch_bounds__(256) sel_hist_kernel(const uint64_t* __restrict__ w, uint64_t n,
                                                       const uint64_t* __restrict__ mm, uint64_t prefix,
                                                       uint32_t above, uint32_t shift, uint32_t mask,
                                                       uint32_t* hist) {
  __shared__ uint32_t h[kSelBins];
  for (uint32_t d = threadIdx.x; d < kSelBins; d += blockDim.x) h[d] = 0;
  __syncthreads();
  const uint64_t lo = mm[0];
  uint32_t run_d = 0, run_n = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t rel = w[i] - lo;
    if (above >= 64 || (rel >> above) == prefix) {
      const uint32_t d = (uint32_t)(rel >> shift) & mask;
      if (d != run_d && run_n) {
        atomicAdd(&h[run_d], run_n);
        run_n = 0;
      }
      run_d = d;
      run_n++;
    }
  }
  if (run_n) atomicAdd(&h[run_d], run_n);
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < kSelBins; d += blockDim.x) hist[(uint64_t)blockIdx.x * kSelBins + d] = h[d];
}

__global__ void sel_reduce_kernel(const uint32_t* __restrict__ hist, uint32_t blocks, uint32_t* out) {
  const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= kSelBins) return;
  uint32_t s = 0;
  for (uint32_t b = 0; b < blocks; b++) s += hist[(uint64_t)b * kSelBins + d];
  out[d] = s;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  const uint32_t lane = threadIdx.x & 63;
  return (uint32_t)__popcll(mask & ((1ull << lane) - 1));
}

// Few candidates: the index i of every word with w[i] - min <= cut, in any
// order (one device atomic per wavefront).  *count starts at zero.
__global__ void __launch_bounds__(256) sel_take_kernel(const uint64_t* __restrict__ w, uint64_t n,
                                                       const uint64_t* __restrict__ mm, uint64_t cut, uint32_t cap,
                                                       uint32_t* idx, uint32_t* count) {
  const uint64_t lo = mm[0];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x; b < n; b += stride) {  // uniform per wavefront
    const uint64_t i = b + threadIdx.x;
    const bool take = i < n && w[i] - lo <= cut;
    const uint64_t m = __ballot(take);
    if (m == 0) continue;
    uint32_t base = 0;
    if ((threadIdx.x & 63) == (uint32_t)(__ffsll((long long)m) - 1)) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, __ffsll((long long)m) - 1);
    const uint32_t at = base + lanes_below(m);
    if (take && at < cap) idx[at] = (uint32_t)i;
  }
}

// Order words of the candidates plus a last word, the candidate's index in
// slot order, so that ties break by slot whatever order they were taken in.
__global__ void sel_comps_kernel(const OrderParams P, const uint32_t* __restrict__ gids, uint64_t n,
                                 const uint32_t* __restrict__ idx, uint32_t m, uint64_t* comps) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x) {
    const uint32_t gid = gids[idx[j] < n ? idx[j] : 0];  // idx[j] < n: taken from [0, n)
    for (uint32_t c = 0; c < P.n_comps; c++) comps[(uint64_t)c * m + j] = comp_word(P, P.comps[c], gid);
    comps[(uint64_t)P.n_comps * m + j] = idx[j];
  }
}

__global__ void sel_out_kernel(const uint32_t* __restrict__ gids, uint64_t n, const uint32_t* __restrict__ idx,
                               const uint32_t* __restrict__ pos, uint64_t k, uint32_t* out) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < k; r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t ix = idx[pos[r]];
    out[r] = gids[ix < n ? ix : 0];
  }
}

// Many candidates: ordered gather of gids[i] for every w[i] - min <= cut.
// Workgroup b covers items [b * kCompactChunk, (b + 1) * kCompactChunk), read
// as kCompactPer coalesced rows of kCompactThreads items.
__global__ void __launch_bounds__(kCompactThreads) sel_count_kernel(const uint64_t* __restrict__ w, uint64_t n,
                                                                    const uint64_t* __restrict__ mm, uint64_t cut,
                                                                    uint32_t* block_counts) {
  __shared__ uint32_t red[kCompactThreads / 64];
  const uint64_t lo = mm[0];
  const uint64_t base = (uint64_t)blockIdx.x * kCompactChunk + threadIdx.x;
  uint32_t c = 0;
  for (uint32_t k = 0; k < kCompactPer; k++) {
    const uint64_t i = base + (uint64_t)k * kCompactThreads;
    if (i < n && w[i] - lo <= cut) c++;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t k = 0; k < kCompactThreads / 64; k++) t += red[k];
    block_counts[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kCompactThreads) sel_write_kernel(const uint64_t* __restrict__ w, uint64_t n,
                                                                    const uint64_t* __restrict__ mm, uint64_t cut,
                                                                    const uint32_t* block_offs,
                                                                    const uint32_t* __restrict__ gids, uint32_t* out) {
  constexpr uint32_t kWaves = kCompactThreads / 64;
  __shared__ uint32_t cnt[kCompactPer * kWaves];
  const uint64_t lo = mm[0];
  const uint64_t base = (uint64_t)blockIdx.x * kCompactChunk + threadIdx.x;
  const uint32_t wave = threadIdx.x >> 6;
  uint64_t masks[kCompactPer];
  for (uint32_t k = 0; k < kCompactPer; k++) {
    const uint64_t i = base + (uint64_t)k * kCompactThreads;
    masks[k] = __ballot(i < n && w[i] - lo <= cut);
    if ((threadIdx.x & 63) == 0) cnt[k * kWaves + wave] = (uint32_t)__popcll(masks[k]);
  }
  __syncthreads();
  if (threadIdx.x == 0) {  // exclusive scan in item order: row k, then wavefront
    uint32_t run = block_offs[blockIdx.x];
    for (uint32_t e = 0; e < kCompactPer * kWaves; e++) {
      const uint32_t v = cnt[e];
      cnt[e] = run;
      run += v;
    }
  }
  __syncthreads();
  for (uint32_t k = 0; k < kCompactPer; k++) {
    if ((masks[k] >> (threadIdx.x & 63)) & 1ull)
      out[cnt[k * kWaves + wave] + lanes_below(masks[k])] = gids[base + (uint64_t)k * kCompactThreads];
  }
}

static_assert(kSortCandidatesMax == kSmallSort, "candidates are ordered by one workgroup");

size_t sort_candidates_scratch_bytes(uint32_t m, uint32_t n_comps) {
  return ((size_t)(n_comps + 1) * m * 8 + 255) / 256 * 256 + (size_t)kSmallSort * 4;
}

hipError_t launch_sort_candidates(const OrderParams& p, const uint32_t* gids, uint64_t n, const uint32_t* idx,
                                  uint32_t m, uint64_t need, void* scratch, uint32_t* out_sel, hipStream_t s) {
  if (m == 0 || need == 0) return hipSuccess;
  if (m > kSmallSort || need > m) return hipErrorInvalidValue;
  uint64_t* comps = static_cast<uint64_t*>(scratch);
  uint32_t* pos = reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(scratch) +
                                              ((size_t)(p.n_comps + 1) * m * 8 + 255) / 256 * 256);
  hipLaunchKernelGGL(sel_comps_kernel, dim3(grid_for(m, 256, 16)), dim3(256), 0, s, p, gids, n, idx, m, comps);
  hipLaunchKernelGGL(small_sort_kernel, dim3(1), dim3(1024), 0, s, (const uint64_t*)comps, p.n_comps + 1, (uint64_t)m,
                     (const uint32_t*)nullptr, m, pos);
  hipLaunchKernelGGL(sel_out_kernel, dim3(grid_for(need, 256, 16)), dim3(256), 0, s, gids, n, idx,
                     (const uint32_t*)pos, need, out_sel);
  return hipGetLastError();
}

static hipError_t sort_general(const OrderParams& p, const uint32_t* gids, uint64_t n, uint64_t need, SortScratch& S,
                               uint32_t* out_sel, hipStream_t s);

// The selection path; leaves *done false (nothing written) when it does not
// apply: every word constant, or so many candidates that the scratch below
// them cannot order them.
static hipError_t select_large(const OrderParams& p, const uint32_t* gids, uint64_t n, uint64_t need,
                               SortScratch& S, uint32_t* out_sel, hipStream_t s, bool* done) {
  *done = false;
  uint64_t* w = S.keys_a;
  // the first word that varies selects (constant words decide nothing)
  uint64_t mm[2] = {0, 0};
  for (uint32_t c = 0; c < p.n_comps && mm[0] == mm[1]; c++) {
    hipLaunchKernelGGL(init_mm_kernel, dim3(1), dim3(1), 0, s, S.mm);
    hipLaunchKernelGGL(sel_word_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, p, c, gids, n, w, S.mm);
    hipError_t e = hipMemcpyAsync(mm, S.mm, sizeof mm, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
  }
  if (mm[0] == mm[1]) return hipSuccess;
  const uint32_t bits = 64 - __builtin_clzll(mm[1] - mm[0]);
  // the need-th smallest relative word, 12 bits at a time from the top;
  // per-workgroup histograms in keys_b (n u64 words hold 2n counts)
  const uint64_t hb = 2 * n / kSelBins;
  const uint32_t hblocks = (uint32_t)(hb < kSelBlocks ? hb : kSelBlocks);
  uint32_t* hist = reinterpret_cast<uint32_t*>(S.keys_b);
  uint32_t* sums = S.pos_a;  // kSelBins counts (n > kSelBins)
  uint64_t prefix = 0, remaining = need, m = 0;
  uint32_t done_bits = 0;
  std::vector<uint32_t> h(kSelBins);
  while (done_bits < bits) {
    const uint32_t width = bits - done_bits < 12 ? bits - done_bits : 12;
    const uint32_t shift = bits - done_bits - width;
    hipLaunchKernelGGL(sel_hist_kernel, dim3(hblocks), dim3(256), 0, s, (const uint64_t*)w, n,
                       (const uint64_t*)S.mm, prefix, shift + width, shift, (1u << width) - 1, hist);
    hipLaunchKernelGGL(sel_reduce_kernel, dim3(kSelBins / 256), dim3(256), 0, s, (const uint32_t*)hist, hblocks,
                       sums);
    hipError_t e = hipMemcpyAsync(h.data(), sums, (4u << width), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    uint64_t below = 0;
    uint32_t v = 0;
    for (; v < (1u << width); v++) {
      if (below + h[v] >= remaining) break;
      below += h[v];
    }
    if (v == (1u << width)) return hipErrorUnknown;  // fewer than `need` words: need <= n rules it out
    remaining -= below;
    m += below;  // every word counted below the digit lies below the cut
    prefix = (prefix << width) | v;
    done_bits += width;
    if (done_bits == bits) m += h[v];  // and the words equal to it
  }
  const uint64_t cut = prefix;  // m words lie at or below it
  if (m < need) return hipErrorUnknown;
  if (m <= kSmallSort) {
    // take the candidates in any order, then sort them with their slot
    // index as the last word
    uint32_t* idx = S.pos_b;
    uint32_t* d_count = reinterpret_cast<uint32_t*>(S.total);
    hipError_t e = launch_fill_u32(d_count, 1, 0, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sel_take_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, (const uint64_t*)w, n,
                       (const uint64_t*)S.mm, cut, (uint32_t)m, idx, d_count);
    e = launch_sort_candidates(p, gids, n, idx, (uint32_t)m, need, S.comps, out_sel, s);
    if (e == hipSuccess) *done = true;
    return e;
  }
  // many ties at the cut: gather the candidates in slot order (ties still
  // break by slot) and order them with the general path, in scratch below them
  const uint32_t blocks = (uint32_t)((n + kCompactChunk - 1) / kCompactChunk);
  uint32_t* counts = S.pos_a;
  uint32_t* cg = S.pos_b;
  const size_t below = (size_t)(reinterpret_cast<uint8_t*>(cg) - reinterpret_cast<uint8_t*>(S.comps));
  if (sort_scratch_bytes(m, p.n_comps) > below) return hipSuccess;
  hipLaunchKernelGGL(sel_count_kernel, dim3(blocks), dim3(kCompactThreads), 0, s, (const uint64_t*)w, n,
                     (const uint64_t*)S.mm, cut, counts);
  hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, s, counts, blocks, S.total);
  hipLaunchKernelGGL(sel_write_kernel, dim3(blocks), dim3(kCompactThreads), 0, s, (const uint64_t*)w, n,
                     (const uint64_t*)S.mm, cut, (const uint32_t*)counts, gids, cg);
  SortScratch C = carve(S.comps, m, p.n_comps);
  hipError_t e = sort_general(p, cg, m, need, C, out_sel, s);
  if (e == hipSuccess) *done = true;
  return e;
}

// Writes the first `need` gids in ORDER BY order into out_sel.
hipError_t launch_sort(const OrderParams& p, uint32_t* gids, uint64_t n, uint64_t need, void* scratch,
                       size_t scratch_bytes, uint32_t* out_sel, hipStream_t s) {
  if (n == 0 || need == 0) return hipSuccess;
  if (scratch_bytes < sort_scratch_bytes(n, p.n_comps)) return hipErrorInvalidValue;
  SortScratch S = carve(scratch, n, p.n_comps);
  const char* sel_env = getenv("THINDB_GPU_TOPN_SELECT");  // "0": always order every group
  const bool select_on = !(sel_env && sel_env[0] == '0');
  if (select_on && n > kSelect1 && need <= kSmallSort && p.n_comps > 0) {
    bool done = false;
    hipError_t e = select_large(p, gids, n, need, S, out_sel, s, &done);
    if (e != hipSuccess || done) return e;
  }
  const char* ch_env = getenv("THINDB_GPU_TOPN_CHUNKS");  // "0": not by chunks
  if (!(ch_env && ch_env[0] == '0') && n > kSmallSort && n <= (uint64_t)kSmallSort * kChunkMax && need <= kTopMaxK &&
      p.n_comps > 0) {
    // every chunk of 4096 groups sorted by its own workgroup, its best
    // `need` kept, those (<= 4096) sorted again: no host round trip
    hipLaunchKernelGGL(make_comps_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, p, (const uint32_t*)gids,
                       n, S.comps);
    // words equal for every group decide nothing: the sorts start at the
    // first that varies (found on the device, no host round trip)
    for (uint32_t c = 0; c < p.n_comps; c++) {
      hipLaunchKernelGGL(init_mm_kernel, dim3(1), dim3(1), 0, s, S.mm + 2 * c);
      hipLaunchKernelGGL(minmax_kernel, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s,
                         (const uint64_t*)(S.comps + (uint64_t)c * n), n, S.mm + 2 * c);
    }
    // the smallest chunk (>= 1024: one element per thread) whose kept
    // candidates still fit one workgroup's sort
    uint32_t chunk = 1024;
    while ((n + chunk - 1) / chunk * need > kSmallSort) chunk <<= 1;  // ends at kSmallSort at the latest
    const uint32_t chunks = (uint32_t)((n + chunk - 1) / chunk);
    const uint64_t last = n - (uint64_t)(chunks - 1) * chunk;
    const uint32_t m = (uint32_t)((chunks - 1) * need + (last < need ? last : need));
    hipLaunchKernelGGL(chunk_sort_kernel, dim3(chunks), dim3(1024), 0, s, (const uint64_t*)S.comps, p.n_comps, n,
                       (const uint64_t*)S.mm, (uint32_t)n, chunk, (uint32_t)need, S.pos_b);
    hipLaunchKernelGGL(cand_sort_kernel, dim3(1), dim3(1024), 0, s, (const uint64_t*)S.comps, p.n_comps, n,
                       (const uint64_t*)S.mm, (const uint32_t*)S.pos_b, m, S.pos_a);
    hipLaunchKernelGGL(select_gids_kernel, dim3(grid_for(need, 256, 4096)), dim3(256), 0, s, (const uint32_t*)gids,
                       (const uint32_t*)S.pos_a, need, out_sel);
    return hipGetLastError();
  }
  return sort_general(p, gids, n, need, S, out_sel, s);
}

// Every order word of all n groups, then one of: slot order (every word
// constant), a one-workgroup sort, a one-workgroup select of the first word,
// repeated per-workgroup top-k, or an LSD radix sort of all positions.
static hipError_t sort_general(const OrderParams& p, const uint32_t* gids, uint64_t n, uint64_t need, SortScratch& S,
                               uint32_t* out_sel, hipStream_t s) {
  hipLaunchKernelGGL(make_comps_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, p, (const uint32_t*)gids,
                     n, S.comps);
  // Words equal across every group decide nothing (a DECIMAL SUM's high
  // half is usually constant): drop them so comparisons stay on the first
  // word that varies.
  uint32_t n_comps = 0;
  uint32_t first_active = 0;
  {
    uint64_t mm[2 * kMaxComps];
    for (uint32_t c = 0; c < p.n_comps; c++) {
      hipLaunchKernelGGL(init_mm_kernel, dim3(1), dim3(1), 0, s, S.mm + 2 * c);
      hipLaunchKernelGGL(minmax_kernel, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s,
                         (const uint64_t*)(S.comps + (uint64_t)c * n), n, S.mm + 2 * c);
    }
    hipError_t e = hipMemcpyAsync(mm, S.mm, 16 * p.n_comps, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    for (uint32_t c = 0; c < p.n_comps; c++) {
      if (mm[2 * c] == mm[2 * c + 1]) continue;
      if (n_comps == 0) first_active = c;
      if (c != n_comps) {
        e = hipMemcpyAsync(S.comps + (uint64_t)n_comps * n, S.comps + (uint64_t)c * n, n * 8,
                           hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
      }
      n_comps++;
    }
  }
  if (n_comps == 0) {
    // every group ties on every word: slot order is the order
    hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, S.pos_a, n);
  } else if (n <= kSmallSort) {
    hipLaunchKernelGGL(small_sort_kernel, dim3(1), dim3(1024), 0, s, (const uint64_t*)S.comps, n_comps, n,
                       (const uint32_t*)nullptr, (uint32_t)n, S.pos_a);
  } else if (n <= kSelect1 && need <= kSmallSort) {
    // radix-select the first word in one workgroup, then sort the few
    // candidates (the best `need` and the ties at the cut)
    uint32_t* d_m = reinterpret_cast<uint32_t*>(S.mm + 2 * kMaxComps);
    hipLaunchKernelGGL(select1_kernel, dim3(1), dim3(1024), 0, s, (const uint64_t*)S.comps, (uint32_t)n,
                       (const uint64_t*)(S.mm + 2 * first_active), (uint32_t)need, S.pos_a, d_m);
    uint32_t m = 0;
    hipError_t e = hipMemcpyAsync(&m, d_m, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    if (m <= kSmallSort) {
      hipLaunchKernelGGL(small_sort_kernel, dim3(1), dim3(1024), 0, s, (const uint64_t*)S.comps, n_comps, n,
                         (const uint32_t*)S.pos_a, m, S.pos_b);
      uint32_t* t = S.pos_a;
      S.pos_a = S.pos_b;
      S.pos_b = t;
    } else {
      e = radix_sort_positions(S, n, n_comps, m, s);
      if (e != hipSuccess) return e;
    }
  } else if (need <= kTopMaxK) {
    // per-workgroup top-k, repeated over the candidates until one
    // workgroup can sort them
    const uint32_t* cur = nullptr;
    uint32_t* dst = S.pos_a;
    uint32_t* other = S.pos_b;
    uint64_t m = n;
    while (m > kSmallSort) {
      const uint32_t blocks = (uint32_t)((m + kTopChunk - 1) / kTopChunk);
      hipLaunchKernelGGL(block_topk_kernel, dim3(blocks), dim3(kTopThreads), 0, s, (const uint64_t*)S.comps,
                         n_comps, n, cur, m, (uint32_t)need, dst);
      const uint64_t last = m - (uint64_t)(blocks - 1) * kTopChunk;
      m = (uint64_t)(blocks - 1) * need + (last < need ? last : need);
      cur = dst;
      uint32_t* t = dst; dst = other; other = t;
    }
    hipLaunchKernelGGL(small_sort_kernel, dim3(1), dim3(1024), 0, s, (const uint64_t*)S.comps, n_comps, n, cur,
                       (uint32_t)m, dst);
    S.pos_a = dst;
  } else {
    hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, S.pos_a, n);
    hipError_t e = radix_sort_positions(S, n, n_comps, n, s);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(select_gids_kernel, dim3(grid_for(need, 256, 4096)), dim3(256), 0, s, (const uint32_t*)gids,
                     (const uint32_t*)S.pos_a, need, out_sel);
  return hipGetLastError();
}

// ---- finalize: typed result columns (finalize_row: order_device.h) ------------------

__global__ void finalize_kernel(const FinalParams P, const uint32_t* __restrict__ sel, uint64_t n_out,
                                uint8_t* out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_out; i += (uint64_t)gridDim.x * blockDim.x)
    finalize_row(P, sel[i], i, out);
}

hipError_t launch_finalize(const FinalParams& p, const uint32_t* sel, uint64_t n_out, uint8_t* out, hipStream_t s) {
  if (n_out == 0) return hipSuccess;
  hipLaunchKernelGGL(finalize_kernel, dim3(grid_for(n_out, 256, 8192)), dim3(256), 0, s, p, sel, n_out, out);
  return hipGetLastError();
}

// ---- one workgroup: dense table -> ordered, sliced, typed result --------------------
//
// Thread t owns slots t, t + kFastThreads, ... (at most kFastJ of them).  The
// records are read twice at most per order word (occupancy, then each word
// until one varies); the first varying word of every slot then stays in LDS,
// where the radix select and the candidate gather read it.  Up to kFastRank
// candidates are ranked (each counts those ordering before it, every thread
// at once, no barrier rounds) and write their rows at their ranks; more take
// the bitonic sort.

constexpr uint32_t kFastThreads = 1024;
constexpr uint32_t kFastJ = kFastRec / kFastThreads;
constexpr uint32_t kFastBins = 4096;  // 12-bit digits of the radix select
constexpr uint32_t kFastRank = 256;   // candidates ranked without a sort

// Block-wide sum of one u32 per thread (result to every thread).
__device__ __forceinline__ uint32_t fast_block_sum(uint32_t v, uint32_t* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t t = 0;
  for (uint32_t w = 0; w < kFastThreads / 64; w++) t += red[w];
  return t;
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int o) {
  return ((uint64_t)__shfl_xor((uint32_t)(v >> 32), o) << 32) | __shfl_xor((uint32_t)v, o);
}

__global__ void __launch_bounds__(kFastThreads) dense_topn_kernel(const OrderParams P, const FinalParams F,
                                                                  uint32_t n_rec, uint32_t offset, uint32_t need,
                                                                  const uint32_t* err, uint8_t* out,
                                                                  FastHeader* hdr) {
  __shared__ uint64_t lw[kFastRec];  // each slot's first varying word
  __shared__ union {
    uint32_t hist[kFastBins];                  // the radix select's digit counts
    uint64_t cw[kFastCand * kFastWords];       // then the candidates' words from the first varying one on
  } u;
  __shared__ uint32_t cg[kFastCand];           // the candidates' slots
  __shared__ uint32_t red[kFastThreads / 64];
  __shared__ unsigned long long mm[2];
  __shared__ uint32_t s_cut, s_n;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint64_t n_groups = 0, n_out = 0;
  uint32_t fallback = 0;
  const uint32_t nc = P.n_comps;  // <= kFastWords (checked by the host)
  uint32_t occm = 0;
#pragma unroll
  for (uint32_t j = 0; j < kFastJ; j++) {  // every slot's count load in flight at once
    const uint32_t g = tid + j * kFastThreads;
    if (g < n_rec && *reinterpret_cast<const uint64_t*>(P.g_state + (uint64_t)g * P.g_rec + P.g_count_off) != 0)
      occm |= 1u << j;
  }
  const uint32_t occ_total = fast_block_sum((uint32_t)__popc(occm), red);
  n_groups = occ_total;
  // the first word whose min and max differ; its values stay in lw
  uint32_t first = nc;
  for (uint32_t c = 0; c < nc; c++) {
    if (tid == 0) { mm[0] = ~0ull; mm[1] = 0ull; }
    __syncthreads();
    const CompDesc cd = P.comps[c];
    uint64_t lo = ~0ull, hi = 0;
#pragma unroll 4
    for (uint32_t j = 0; j < kFastJ; j++) {
      if (!((occm >> j) & 1u)) continue;
      const uint32_t g = tid + j * kFastThreads;
      const uint64_t w = comp_word(P, cd, g);
      lw[g] = w;
      lo = w < lo ? w : lo;
      hi = w > hi ? w : hi;
    }
    for (int o = 32; o > 0; o >>= 1) {  // one LDS atomic per wavefront (atomics on one address serialise)
      const uint64_t l2 = shfl_xor_u64(lo, o), h2 = shfl_xor_u64(hi, o);
      lo = l2 < lo ? l2 : lo;
      hi = h2 > hi ? h2 : hi;
    }
    if (lane == 0) {
      atomicMin(&mm[0], (unsigned long long)lo);
      atomicMax(&mm[1], (unsigned long long)hi);
    }
    __syncthreads();
    const bool varies = mm[0] < mm[1];
    if (varies) { first = c; break; }
  }
  if (first == nc) {
    // every group ties on every word (or no ORDER BY): slot order is the
    // order; positions by wavefront ballots, round j covering slots
    // [j * kFastThreads, (j + 1) * kFastThreads)
    uint32_t base = 0;
    for (uint32_t j = 0; j * kFastThreads < n_rec; j++) {
      const bool occ = (occm >> j) & 1u;
      const uint64_t m = __ballot(occ);
      __syncthreads();
      if (lane == 0) red[wave] = (uint32_t)__popcll(m);
      __syncthreads();
      uint32_t before = 0, total = 0;
      for (uint32_t w = 0; w < kFastThreads / 64; w++) {
        if (w < wave) before += red[w];
        total += red[w];
      }
      const uint32_t pos = base + before + (uint32_t)__popcll(m & ((1ull << lane) - 1));
      if (occ && pos >= offset && pos < need) finalize_row(F, tid + j * kFastThreads, pos - offset, out);
      base += total;
    }
    n_out = base > offset ? (need < base ? need : base) - offset : 0;
  } else {
    // radix select: the region (w & mask) == val holds the want-th smallest
    // word; every group below the region is in
    const uint64_t lo0 = mm[0], hi0 = mm[1];
    int top = 63 - __clzll(lo0 ^ hi0);  // highest varying bit
    uint64_t mask = top < 63 ? ~((2ull << top) - 1) : 0ull;
    uint64_t val = lo0 & mask;
    const uint32_t want = need < occ_total ? need : occ_total;
    uint32_t in = 0, maybe = occ_total;
    while (top >= 0 && in + maybe > kFastCand) {
      const int sh = top >= 11 ? top - 11 : 0;
      const uint32_t bits = (uint32_t)(top - sh + 1);
      for (uint32_t b = tid; b < kFastBins; b += kFastThreads) u.hist[b] = 0;
      __syncthreads();
      // words cluster (most groups share the few top digits of a skewed
      // distribution): lanes with the digit of the wavefront's first active
      // lane add once through that lane, the others one by one
      for (uint32_t j = 0; j < kFastJ; j++) {
        const uint32_t g = tid + j * kFastThreads;
        const uint64_t w = (occm >> j) & 1u ? lw[g] : 0;
        const bool act = ((occm >> j) & 1u) && (w & mask) == val;
        const uint64_t am = __ballot(act);
        if (!am) continue;
        const uint32_t d = (uint32_t)(w >> sh) & ((1u << bits) - 1);
        const int lead = __ffsll((long long)am) - 1;
        const uint32_t d0 = __shfl(d, lead);
        const uint64_t same = __ballot(act && d == d0);
        if ((int)lane == lead) atomicAdd(&u.hist[d0], (uint32_t)__popcll(same));
        else if (act && d != d0) atomicAdd(&u.hist[d], 1u);
      }
      __syncthreads();
      // the bin where the count from the smallest end reaches `want`: each
      // thread sums its run of kFastBins / kFastThreads bins, a scan of the
      // runs places them, and the run holding the crossing names the bin
      {
        constexpr uint32_t per = kFastBins / kFastThreads;
        const uint32_t b0 = tid * per;
        uint32_t mine = 0;
        for (uint32_t q = 0; q < per; q++) mine += u.hist[b0 + q];
        uint32_t incl = mine;  // inclusive scan within the wavefront
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t t = __shfl_up(incl, o);
          if ((int)lane >= o) incl += t;
        }
        if (lane == 63) red[wave] = incl;
        __syncthreads();
        uint32_t before = 0;
        for (uint32_t w = 0; w < wave; w++) before += red[w];
        uint32_t cum = in + before + incl - mine;  // groups below this thread's run
        if (cum < want && cum + mine >= want) {
          uint32_t q = 0;
          for (; q < per; q++) {
            if (cum + u.hist[b0 + q] >= want) break;
            cum += u.hist[b0 + q];
          }
          s_cut = b0 + q;
          s_n = cum;  // groups below the cut bin
        }
        __syncthreads();
      }
      const uint32_t b = s_cut;
      maybe = u.hist[b];
      in = s_n;
      mask |= ((1ull << bits) - 1) << sh;
      val |= (uint64_t)b << sh;
      top = sh - 1;
      __syncthreads();
    }
    if (in + maybe > kFastCand) {
      fallback = 1;  // more than kFastCand groups tie at the cut
    } else {
      // candidates: every group whose word is at most the region's top
      const uint64_t region_max = val | ~mask;
      const uint32_t nw = nc - first;
      if (tid == 0) s_n = 0;
      __syncthreads();
      for (uint32_t j = 0; j < kFastJ; j++) {
        const uint32_t g = tid + j * kFastThreads;
        const uint64_t w = (occm >> j) & 1u ? lw[g] : 0;
        const bool take = ((occm >> j) & 1u) && w <= region_max;
        const uint64_t m = __ballot(take);  // one LDS atomic per wavefront and slot round
        if (!m) continue;
        const int lead = __ffsll((long long)m) - 1;
        uint32_t at0 = 0;
        if ((int)lane == lead) at0 = atomicAdd(&s_n, (uint32_t)__popcll(m));
        at0 = __shfl(at0, lead);
        if (!take) continue;
        const uint32_t at = at0 + (uint32_t)__popcll(m & ((1ull << lane) - 1));
        cg[at] = g;
        u.cw[at * kFastWords] = w;
        for (uint32_t c = 1; c < nw; c++) u.cw[at * kFastWords + c] = comp_word(P, P.comps[first + c], g);
      }
      __syncthreads();
      const uint32_t n = s_n;
      const uint32_t end = want < n ? want : n;
      n_out = end > offset ? end - offset : 0;
      if (n <= kFastRank) {
        // rank = the candidates ordering before this one by (words, slot),
        // distinct since slots are; ranks [offset, end) are the rows out.
        // q neighbouring lanes share a candidate, each comparing every q-th
        // other one, and add their counts by shuffles
        uint32_t q = 1;
        while (q < 16 && n * (q << 1) <= kFastThreads) q <<= 1;
        const uint32_t a = tid / q, part = tid & (q - 1);
        uint32_t r = 0, ga = 0;
        if (a < n) {
          ga = cg[a];
          uint64_t xa[kFastWords];
#pragma unroll
          for (uint32_t c = 0; c < kFastWords; c++) xa[c] = c < nw ? u.cw[a * kFastWords + c] : 0;
#pragma unroll 2
          for (uint32_t i = part; i < n; i += q) {
            bool before = cg[i] < ga, decided = false;
#pragma unroll
            for (uint32_t c = 0; c < kFastWords; c++) {
              if (c >= nw) break;
              const uint64_t xb = u.cw[i * kFastWords + c];
              if (!decided && xb != xa[c]) { before = xb < xa[c]; decided = true; }
            }
            r += before;
          }
        }
        for (uint32_t o = 1; o < q; o <<= 1) r += __shfl_xor(r, (int)o);
        if (a < n && part == 0 && r >= offset && r < end) finalize_row(F, ga, r - offset, out);
      } else {
        uint32_t m = 1;
        while (m < n) m <<= 1;
        for (uint32_t i = n + tid; i < m; i += kFastThreads) cg[i] = 0xffffffffu;  // after every group
        __syncthreads();
        // bitonic sort by (words, slot)
        for (uint32_t k = 2; k <= m; k <<= 1) {
          for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < m; i += kFastThreads) {
              const uint32_t l = i ^ j;
              if (l <= i) continue;
              const uint32_t ga = cg[i], gb = cg[l];
              bool l_less;  // element l orders before element i
              if (gb == 0xffffffffu) l_less = false;
              else if (ga == 0xffffffffu) l_less = true;
              else {
                l_less = gb < ga;
                for (uint32_t c = 0; c < nw; c++) {
                  const uint64_t xa = u.cw[i * kFastWords + c], xb = u.cw[l * kFastWords + c];
                  if (xa != xb) { l_less = xb < xa; break; }
                }
              }
              const bool up = (i & k) == 0;
              if (l_less == up) {
                cg[i] = gb;
                cg[l] = ga;
                for (uint32_t c = 0; c < nw; c++) {
                  const uint64_t t = u.cw[i * kFastWords + c];
                  u.cw[i * kFastWords + c] = u.cw[l * kFastWords + c];
                  u.cw[l * kFastWords + c] = t;
                }
              }
            }
            __syncthreads();
          }
        }
        for (uint32_t i = offset + tid; i < end; i += kFastThreads) finalize_row(F, cg[i], i - offset, out);
      }
    }
  }
  if (tid == 0) {
    hdr->n_groups = n_groups;
    hdr->n_out = fallback ? 0 : n_out;
    hdr->err = err[0];
    hdr->fallback = fallback;
  }
}

hipError_t launch_dense_topn(const OrderParams& op, const FinalParams& fp, uint32_t n_rec, uint32_t offset,
                             uint32_t need, const uint32_t* err, uint8_t* out, FastHeader* hdr, hipStream_t s) {
  if (n_rec == 0 || n_rec > kFastRec) return hipErrorInvalidValue;
  hipLaunchKernelGGL(dense_topn_kernel, dim3(1), dim3(kFastThreads), 0, s, op, fp, n_rec, offset, need, err, out,
                     hdr);
  return hipGetLastError();
}

}  // namespace tdbg
