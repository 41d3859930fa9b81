// batch_device.h — the batched persistent kernel's loop (batch.hip runs it
// with the generic strip; jit.cpp with strips generated for a batch's plans).
//
// A batch is a set of work items (BatchItem) over the plans' tiles, radix
// leaves and result records; every workgroup of the grid takes the next item
// until every item was taken, so tables of very different sizes share the
// GPU without a launch (or a drain) per plan.  Per item the workgroup runs
// the same routines as the single-plan kernels:
//   BI_SCAN        scan/filter/aggregate tiles [a0, a1) into an LDS table
//                  (direct: one key range `part`; hash) or HBM atomics
//   BI_RX_HIST     radix level-1 digit counts of the filtered rows of tiles
//                  [a0, a1); the plan's last histogram item turns every
//                  item's counts into its write offsets and the bucket ranges
//   BI_RX_SCATTER  radix scatter of tiles [a0, a1) at its item's offsets (no
//                  device atomics)
//   BI_RX_AGG      radix leaves [a0, a1) into the plan's dense record array
//   BI_EMIT        records [a0, a1): the groups among them (occupied,
//                  passing HAVING) written as typed result rows straight
//                  into the output buffer in host memory, in record order
// Items are taken when they are ready, never before, so no workgroup waits
// for another's work: histogram and scan items sit in a static queue; the
// workgroup that completes a plan's last item of a stage appends the plan's
// next-stage items (scatters, leaves, emits) to that stage's ready queue.
// Deeper stages are taken first, so a plan runs to its end soon after it
// starts and its result rows cross PCIe while other plans compute.  At most
// `emit_cap` emit items run at once while other work is ready: a few dozen
// workgroups keep PCIe writes at their peak, and the rest keep computing.
//
// Device-scope release and acquire are L2 writebacks and invalidations on
// this chip (each XCD has its own L2), so they are spent once per item, not
// per access, and only where plain loads must see other workgroups' plain
// stores: an item whose output other workgroups read (histogram counts,
// partition rows, leaf records) ends with one release by one thread before
// it counts itself done; a leaf item starts with one acquire (it reads the
// partition rows).  A scatter item reads its write offsets, and an emit item
// its records, with device-coherent loads that bypass a stale L2 line
// instead, so neither invalidates the L2 other workgroups of its XCD are
// reusing (an emit item over HAVING, or over records too wide to stage,
// still acquires).  Queue words, slots and emit counts are relaxed atomics.
// Scan items update records only through device atomics.
//
// An emit item counts its rows and publishes the count before it waits for
// anything; its first output row is the sum of the counts of the plan's
// earlier emit items, which were taken before it (a queue hands out its
// slots in order) by running workgroups.  Every wait therefore ends; one that
// exceeds spin_limit polls raises ERR_BATCH_WAIT instead of hanging.
#pragma once

#include "radix_device.h"

namespace tdbg {

constexpr uint32_t kNoItem = 0xffffffffu;  // nothing ready
constexpr uint32_t kExit = 0xfffffffeu;    // every item was taken

// Takes the next slot of ready queue k, if one was appended (thread 0).
// The producer reserved the slot before writing it: its item index arrives
// after a short wait.  *live is set while the queue still has slots to hand
// out (appended or not yet appended).
__device__ __forceinline__ uint32_t rq_take(const BatchArgs& A, uint32_t k, bool* live) {
  if (A.q_items[k] == 0) return kNoItem;  // nothing is ever appended
  BatchSched* S = A.sched;
  uint32_t t = __hip_atomic_load(&S->tail[k].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (t == 0) {  // nothing appended yet
    *live = true;
    return kNoItem;
  }
  uint32_t h = __hip_atomic_load(&S->head[k].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (;;) {
    if (h >= t) {
      if (h < A.q_items[k]) *live = true;
      return kNoItem;
    }
    if (__hip_atomic_compare_exchange_strong(&S->head[k].v, &h, h + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      break;
    t = __hip_atomic_load(&S->tail[k].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  uint32_t v, spins = 0;
  while ((v = __hip_atomic_load(&A.slots[k][h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0) {
    __builtin_amdgcn_s_sleep(1);
    if (++spins >= A.spin_limit) {
      atomicOr(A.err, ERR_BATCH_WAIT);
      return kExit;
    }
  }
  return v - 1;
}

// The next item for this workgroup (thread 0): a ready emit item while
// fewer than emit_cap run, then ready leaves, scatters, the static queue,
// then emit items beyond the cap; kExit once every item was taken (every
// queue handed out all its slots: no counter of taken items, which every
// pick would update).
__device__ __forceinline__ uint32_t batch_pick(const BatchArgs& A) {
  BatchSched* S = A.sched;
  for (uint32_t spins = 0;; spins++) {
    uint32_t it = kNoItem;
    bool leaf = false;  // a leaf item reads partition rows other workgroups wrote
    bool live = false;  // a ready queue still has slots to hand out
    const bool room = __hip_atomic_load(&S->emitters.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < A.emit_cap;
    if (room && (it = rq_take(A, BQ_EMIT, &live)) < kExit) atomicAdd(&S->emitters.v, 1u);
    if (it == kNoItem) leaf = (it = rq_take(A, BQ_AGG, &live)) < kExit;
    if (it == kNoItem) it = rq_take(A, BQ_SCATTER, &live);
    if (it == kNoItem && __hip_atomic_load(&S->head0.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < A.n_static) {
      const uint32_t h = atomicAdd(&S->head0.v, 1u);
      if (h < A.n_static) it = h;
    }
    if (it == kNoItem && !room && (it = rq_take(A, BQ_EMIT, &live)) < kExit) atomicAdd(&S->emitters.v, 1u);
    if (it == kExit) return kExit;
    if (it != kNoItem) {
      // the scatters released their writes before the item was appended
      if (leaf) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      return it;
    }
    if (!live) return kExit;  // every queue was examined above; the static one is taken
    if (spins >= A.spin_limit) {
      atomicOr(A.err, ERR_BATCH_WAIT);
      return kExit;
    }
    // idle: a dependency is still running; back off (to a few microseconds
    // between polls) so idle workgroups do not crowd the queue words
    if (spins < 2) __builtin_amdgcn_s_sleep(8);
    else if (spins < 8) __builtin_amdgcn_s_sleep(32);
    else __builtin_amdgcn_s_sleep(127);
  }
}

// Counts a completed item of a plan's stage (*done of `total`); the
// workgroup that completes the last one returns true (uniform).  RELEASE:
// the item's plain stores are read by other workgroups, so they are written
// back first (one release, after the barrier has every wave's stores in
// L2); ACQUIRE: the last workgroup reads what the others released.
template <bool RELEASE, bool ACQUIRE>
__device__ __forceinline__ bool stage_done(uint32_t* done, uint32_t total, uint32_t* s_flag) {
  __syncthreads();
  if (threadIdx.x == 0) {
    if (RELEASE) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const bool last = atomicAdd(done, 1u) + 1 == total;
    if (ACQUIRE && last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *s_flag = last ? 1u : 0u;
  }
  __syncthreads();
  return *s_flag != 0;
}

// Appends items [first, first + n) to ready queue k (whole workgroup).
__device__ __forceinline__ void rq_push(const BatchArgs& A, uint32_t k, uint32_t first, uint32_t n, uint32_t* s_base) {
  if (threadIdx.x == 0) *s_base = atomicAdd(&A.sched->tail[k].v, n);
  __syncthreads();
  const uint32_t base = *s_base;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
    __hip_atomic_store(&A.slots[k][base + i], first + i + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
}

// Emit item `e` of plan q over records [r0, r1) (r1 - r0 <= kEmitChunk):
// record r0 + s * blockDim.x + t is thread t's in slab s.  First every
// thread requests its records' count words at once and the item publishes
// how many rows it writes; then, slab by slab, the slab's records are copied
// into LDS with coalesced 8-byte loads and each kept row is written from
// there (a record read straight from HBM per output column would stall on
// one memory round trip per column), as consecutive rows of every column.
__device__ __forceinline__ void batch_emit(const BatchArgs& A, const BatchQuery& q, uint32_t e, uint64_t r0,
                                           uint64_t r1, uint8_t* lds) {
  constexpr uint32_t kWaves = kBatchThreads / 64;
  constexpr uint32_t kPerLane = kEmitPer * kWaves / 64;  // (slab, wavefront) counts per lane of the scan
  static_assert(kEmitPer * kWaves == 64 * kPerLane, "one wavefront scans the (slab, wavefront) counts");
  // LDS: kept rows per (slab, wavefront), then their offsets; the item's
  // count and the wavefronts' partial sums; the keep masks; staged records
  constexpr uint32_t kRedOff = 4 * kEmitPer * kWaves, kKeepOff = kRedOff + 128;
  static_assert(8 * (kWaves + 1) <= 128 && kKeepOff + 8 * kEmitPer * kWaves <= kEmitStageOff, "emit LDS layout");
  uint32_t* cnt = reinterpret_cast<uint32_t*>(lds);
  uint64_t* red = reinterpret_cast<uint64_t*>(lds + kRedOff);
  uint64_t* keep = reinterpret_cast<uint64_t*>(lds + kKeepOff);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const FinalParams& F = q.fp;
  const uint32_t g_rec = F.g_rec;
  const bool staged = (g_rec & 7u) == 0 && kEmitStageOff + (uint64_t)kBatchThreads * g_rec <= A.lds_bytes;
  if (q.n_having || !staged) {  // records read with plain loads below: see the producers' writes
    if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __syncthreads();
  }
  constexpr uint32_t kLoads = 4;  // count words in flight per thread
  for (uint32_t s0 = 0; s0 < kEmitPer; s0 += kLoads) {
    uint64_t cw[kLoads];
#pragma unroll
    for (uint32_t j = 0; j < kLoads; j++) {
      const uint64_t r = r0 + (uint64_t)(s0 + j) * kBatchThreads + threadIdx.x;
      cw[j] = r < r1 ? __hip_atomic_load(reinterpret_cast<const uint64_t*>(F.g_state + r * g_rec + F.g_count_off),
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                     : 0ull;
    }
#pragma unroll
    for (uint32_t j = 0; j < kLoads; j++) {
      const uint32_t s = s0 + j;
      const uint64_t r = r0 + (uint64_t)s * kBatchThreads + threadIdx.x;
      const bool k = cw[j] != 0 && (q.n_having == 0 || having_pass(F, q.having, q.n_having, (uint32_t)r));
      const uint64_t m = __ballot(k);
      if (lane == 0) {
        cnt[s * kWaves + wave] = (uint32_t)__popcll(m);
        keep[s * kWaves + wave] = m;
      }
    }
  }
  __syncthreads();
  if (wave == 0) {  // (slab, wavefront)-major exclusive offsets; the item's row count
    uint32_t c[kPerLane], v = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPerLane; j++) {
      c[j] = cnt[lane * kPerLane + j];
      v += c[j];
    }
    uint32_t x = v;
    for (uint32_t o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    uint32_t run = x - v;
#pragma unroll
    for (uint32_t j = 0; j < kPerLane; j++) {
      cnt[lane * kPerLane + j] = run;
      run += c[j];
    }
    if (lane == 63) {
      red[0] = x;
      __hip_atomic_store(&q.emit_status[e], kEmitReady | x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
  const uint64_t mine = red[0];
  // the rows of the plan's earlier emit items (each publishes its count
  // before it waits for anything)
  uint64_t before = 0;
  for (uint32_t k = threadIdx.x; k < e; k += blockDim.x) {
    uint64_t v;
    uint32_t spins = 0;
    while (!((v = __hip_atomic_load(&q.emit_status[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) & kEmitReady)) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins >= A.spin_limit) {
        atomicOr(A.err, ERR_BATCH_WAIT);
        v = kEmitReady;
        break;
      }
    }
    before += v & ~kEmitReady;
  }
  for (int o = 32; o > 0; o >>= 1) before += __shfl_xor(before, o);
  __syncthreads();
  if (lane == 0) red[1 + wave] = before;
  __syncthreads();
  uint64_t first = 0;
  for (uint32_t w = 0; w < kWaves; w++) first += red[1 + w];
  uint8_t* out = A.out;
  uint8_t* stage = lds + kEmitStageOff;
  for (uint32_t s = 0; s < kEmitPer; s++) {  // uniform
    const uint64_t base = r0 + (uint64_t)s * kBatchThreads;
    if (base >= r1) break;
    const uint32_t slab_end = s + 1 < kEmitPer ? cnt[(s + 1) * kWaves] : (uint32_t)mine;
    if (slab_end == cnt[s * kWaves]) continue;  // no row in this slab
    const uint64_t n = r1 - base < kBatchThreads ? r1 - base : kBatchThreads;
    if (staged) {
      const uint64_t* src = reinterpret_cast<const uint64_t*>(F.g_state + base * g_rec);
      uint64_t* dst = reinterpret_cast<uint64_t*>(stage);
      const uint32_t words = (uint32_t)(n * g_rec / 8);
      for (uint32_t w = threadIdx.x; w < words; w += kBatchThreads)
        dst[w] = __hip_atomic_load(src + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
    }
    const uint64_t m = keep[s * kWaves + wave];
    if ((m >> lane) & 1ull) {
      const uint64_t r = base + threadIdx.x;
      const uint8_t* rec = staged ? stage + (uint64_t)threadIdx.x * g_rec : F.g_state + r * g_rec;
      const uint64_t row = first + cnt[s * kWaves + wave] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
      finalize_rec(F, rec, (uint32_t)r, row, out);
    }
    if (staged) __syncthreads();  // the next slab reuses the stage
  }
  if (e + 1 == q.n_emit && threadIdx.x == 0) *q.n_out = first + mine;
}

// The persistent loop of a batch kernel; SCAN::run(B, P, lds, it) runs a
// scan item (the generic strip ahead of time, or strips generated for the
// batch's plans at run time).  Kernels pass Q and items as restrict
// parameters: the plans' descriptors are read-only, so the scan paths keep
// them in registers across stores.
template <class SCAN>
__device__ __forceinline__ void batch_loop(const BatchQuery* __restrict__ Q, const BatchItem* __restrict__ items,
                                           const BatchArgs* __restrict__ Ap) {
  const BatchArgs& A = *Ap;
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ uint32_t s_item;
  __shared__ uint32_t s_flag;  // stage_done's verdict
  __shared__ uint32_t s_base;  // rq_push's first slot
  for (;;) {
    if (threadIdx.x == 0) s_item = batch_pick(A);
    __syncthreads();
    const uint32_t idx = __builtin_amdgcn_readfirstlane(s_item);
    if (idx == kExit) break;  // uniform: every workgroup leaves once every item was taken
    const BatchItem it = items[idx];
    const BatchQuery& B = Q[it.q];
    const ScanParams& P = B.sp;
    switch (it.kind) {
      case BI_SCAN: {
        SCAN::run(B, P, lds, it);
        // records change only through device atomics: nothing to release
        if (stage_done<false, false>(&B.sync[2], B.n_comp, &s_flag)) rq_push(A, BQ_EMIT, B.em0, B.n_emit, &s_base);
        break;
      }
      case BI_RX_HIST: {
        const uint32_t nb = 1u << B.rx.bits1;
        rx_item_hist(P, B.rx, it.a0, it.a1, reinterpret_cast<uint32_t*>(lds), B.item_hist + (uint64_t)it.part * nb);
        if (stage_done<true, true>(&B.sync[0], B.n_hist, &s_flag)) {
          // the plan's last histogram item turns counts into offsets
          rx_item_offsets(B.rx, B.item_hist, B.n_hist, reinterpret_cast<uint64_t*>(lds));
          __syncthreads();
          if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
          rq_push(A, BQ_SCATTER, B.sc0, B.n_scatter, &s_base);
        }
        break;
      }
      case BI_RX_SCATTER: {
        const uint32_t nb = 1u << B.rx.bits1;
        rx_scatter_range(P, B.rx, it.a0, it.a1, lds, B.item_hist + (uint64_t)it.part * nb);
        if (stage_done<true, false>(&B.sync[1], B.n_scatter, &s_flag)) rq_push(A, BQ_AGG, B.ag0, B.n_comp, &s_base);
        break;
      }
      case BI_RX_AGG: {
        // a batched plan has at most 2^10 leaves (one partition level), so
        // its leaves hold thousands of rows: 16 rows' keys per thread per
        // memory round trip
        for (uint32_t leaf = it.a0; leaf < it.a1; leaf++) rx_agg_leaf<true, false, 16>(P, B.rx, leaf, lds, nullptr);
        if (stage_done<true, false>(&B.sync[2], B.n_comp, &s_flag)) rq_push(A, BQ_EMIT, B.em0, B.n_emit, &s_base);
        break;
      }
      default: {  // BI_EMIT
        batch_emit(A, B, it.part, it.a0, it.a1, lds);
        __syncthreads();
        if (threadIdx.x == 0) atomicSub(&A.sched->emitters.v, 1u);
        break;
      }
    }
    __syncthreads();  // s_item and the LDS are reused by the next item
  }
}

}  // namespace tdbg
