// decode.hip — expansion of encoded column blocks into a table's HBM arena.
//
// Replaces the reference's per-block decode: decodeColumnPayload /
// decodeForColumn / decodeRleColumn (src/storage/segment_reader.zig:1727-1760,
// :1989-2075, :2133-2247).  thinDB's block cache keeps blocks decompressed but
// encoded and expands them on every scan; here the encoded body crosses PCIe
// once and is expanded once, into the native-width arena the scan kernels
// read (the columns stay resident in HBM).
//
//   FOR  out[i] = base + code[i], codes `width` bytes little-endian, the add
//        wrapping at the element width (expandWidth, :2058-2075)
//   RLE  out[i] = value of the run covering row i (fillRuns, :2227-2247);
//        rows past the last run (a block whose lengths sum below its row
//        count) get 0
#include "engine.h"

namespace tdbg {

template <typename T, typename C>
__global__ void for_expand_kernel(const C* __restrict__ codes, int64_t base, uint64_t n, T* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (T)((uint64_t)base + (uint64_t)codes[i]);
}

// starts[r] = first row of run r (ascending, starts[0] = 0); runs cover rows
// [0, covered).
template <typename T>
__global__ void rle_expand_kernel(const T* __restrict__ values, const uint32_t* __restrict__ starts, uint32_t n_runs,
                                  uint32_t covered, uint64_t n, T* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (i >= covered || n_runs == 0) {
      out[i] = (T)0;
      continue;
    }
    uint32_t lo = 0, hi = n_runs - 1;  // the last run starting at or before i
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) / 2;
      if (starts[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    out[i] = values[lo];
  }
}

static uint32_t grid_of(uint64_t n) {
  const uint64_t g = (n + 255) / 256;
  return (uint32_t)(g < 4096 ? (g ? g : 1) : 4096);
}

template <typename T, typename C>
static void for_launch(const void* codes, int64_t base, uint64_t n, void* out, hipStream_t s) {
  hipLaunchKernelGGL((for_expand_kernel<T, C>), dim3(grid_of(n)), dim3(256), 0, s, static_cast<const C*>(codes), base,
                     n, static_cast<T*>(out));
}

template <typename T>
static hipError_t for_dispatch(const void* codes, uint32_t width, int64_t base, uint64_t n, void* out, hipStream_t s) {
  switch (width) {
    case 1: for_launch<T, uint8_t>(codes, base, n, out, s); break;
    case 2: for_launch<T, uint16_t>(codes, base, n, out, s); break;
    case 4: for_launch<T, uint32_t>(codes, base, n, out, s); break;
    case 8: for_launch<T, uint64_t>(codes, base, n, out, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_for_expand(const void* codes, uint32_t width, int64_t base, uint32_t elem, uint64_t n, void* out,
                             hipStream_t s) {
  switch (elem) {
    case 1: return for_dispatch<uint8_t>(codes, width, base, n, out, s);
    case 2: return for_dispatch<uint16_t>(codes, width, base, n, out, s);
    case 4: return for_dispatch<uint32_t>(codes, width, base, n, out, s);
    case 8: return for_dispatch<uint64_t>(codes, width, base, n, out, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_rle_expand(const void* values, const uint32_t* starts, uint32_t n_runs, uint32_t covered,
                             uint32_t elem, uint64_t n, void* out, hipStream_t s) {
  const dim3 g(grid_of(n)), b(256);
  switch (elem) {
    case 1:
      hipLaunchKernelGGL(rle_expand_kernel<uint8_t>, g, b, 0, s, static_cast<const uint8_t*>(values), starts, n_runs,
                         covered, n, static_cast<uint8_t*>(out));
      break;
    case 2:
      hipLaunchKernelGGL(rle_expand_kernel<uint16_t>, g, b, 0, s, static_cast<const uint16_t*>(values), starts, n_runs,
                         covered, n, static_cast<uint16_t*>(out));
      break;
    case 4:
      hipLaunchKernelGGL(rle_expand_kernel<uint32_t>, g, b, 0, s, static_cast<const uint32_t*>(values), starts, n_runs,
                         covered, n, static_cast<uint32_t*>(out));
      break;
    case 8:
      hipLaunchKernelGGL(rle_expand_kernel<uint64_t>, g, b, 0, s, static_cast<const uint64_t*>(values), starts, n_runs,
                         covered, n, static_cast<uint64_t*>(out));
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace tdbg
