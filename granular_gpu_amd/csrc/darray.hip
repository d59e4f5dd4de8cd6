// Generic device-array primitives behind the host DArray class: stable
// compaction by a flag array and a 32-bit fill.
//
// Reference: DArray<T>::compact (DArray.hpp:62-122) keeps the elements whose
// flag != 1, in order, via thrust count_if + copy_if into a temporary vector
// and a copy back.  Here: three passes over 1024-element tiles (count per
// tile, scan of tile counts, ordered scatter with a block-local ballot
// prefix), straight into the caller's destination, scratch passed in.
#include <hip/hip_runtime.h>

#include "gg_device.hpp"
#include "gg_internal.hpp"

namespace gg {

constexpr int kCompactTile = 1024;  // elements per tile: 256 threads x 4

__device__ __forceinline__ bool keep_flag(const int *flags, long long i, long long n) {
  return i < n && flags[i] != 1;
}

__global__ void __launch_bounds__(256)
    k_compact_count(const int *__restrict__ flags, long long n, int *__restrict__ tile_count) {
  const long long base = (long long)blockIdx.x * kCompactTile;
  int c = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) c += keep_flag(flags, base + threadIdx.x + 256 * q, n) ? 1 : 0;
  __shared__ int red[4];
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) tile_count[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// One block: exclusive scan of the tile counts in place; total -> out.
__global__ void __launch_bounds__(1024)
    k_compact_scan(int *__restrict__ tile_count, int ntiles, int *__restrict__ total_out) {
  __shared__ int wsum[16];
  __shared__ int carry_s;
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int start = 0; start < ntiles; start += 1024) {
    const int i = start + threadIdx.x;
    const int v = i < ntiles ? tile_count[i] : 0;
    // inclusive scan inside the wave
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      int acc = 0;
      for (int w = 0; w < 16; ++w) {
        const int t = wsum[w];
        wsum[w] = acc;
        acc += t;
      }
    }
    __syncthreads();
    const int carry = carry_s;
    if (i < ntiles) tile_count[i] = carry + wsum[wid] + x - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry_s = carry + wsum[wid] + x;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total_out = carry_s;
}

// Ordered scatter: element i of tile t goes to tile_off[t] + (number of kept
// elements before i inside the tile).  Elements are elem_words 32-bit words.
__global__ void __launch_bounds__(256)
    k_compact_scatter(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                      const int *__restrict__ flags, long long n, int elem_words,
                      const int *__restrict__ tile_off) {
  __shared__ int wcount[16];
  const long long base = (long long)blockIdx.x * kCompactTile;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // sub-tile q covers elements base + 256 q .. +255; wave w of sub-tile q is
  // element block (q, w); blocks are ranked in (q, w) order = index order
  bool k[4];
  unsigned long long m[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    k[q] = keep_flag(flags, base + threadIdx.x + 256 * q, n);
    m[q] = __ballot(k[q]);
    if (lane == 0) wcount[q * 4 + wid] = __popcll(m[q]);
  }
  __syncthreads();
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int t0 = tile_off[blockIdx.x];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (!k[q]) continue;
    int before = 0;
    for (int b = 0; b < q * 4 + wid; ++b) before += wcount[b];
    const long long d = (long long)t0 + before + __popcll(m[q] & below);
    const long long s = base + threadIdx.x + 256 * q;
    for (int w = 0; w < elem_words; ++w)
      dst[d * elem_words + w] = src[s * elem_words + w];
  }
}

__global__ void k_fill_words(uint32_t *__restrict__ p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace gg

using namespace gg;

extern "C" size_t gg_compact_scratch_bytes(long long n) {
  if (n < 0) return 0;
  const long long ntiles = (n + kCompactTile - 1) / kCompactTile;
  return (size_t)(ntiles + 1) * sizeof(int);
}

extern "C" int gg_compact(const void *src, void *dst, long long n, int elem_words,
                          const int *flags, int *count_out, void *scratch, size_t scratch_bytes,
                          hipStream_t stream) {
  if (n < 0 || elem_words <= 0 || !count_out) return GG_ERR_INVALID;
  if (n > 0x7fffffffLL) return GG_ERR_CAPACITY;
  if (n > 0 && (!src || !dst || !flags || !scratch)) return GG_ERR_INVALID;
  if (scratch_bytes < gg_compact_scratch_bytes(n)) return GG_ERR_WORKSPACE;
  if (src == dst && n > 0) return GG_ERR_INVALID;
  int *tiles = static_cast<int *>(scratch);
  const int ntiles = (int)((n + kCompactTile - 1) / kCompactTile);
  if (ntiles > 0) {
    k_compact_count<<<ntiles, 256, 0, stream>>>(flags, n, tiles);
    GG_LAUNCH_CHECK();
  }
  k_compact_scan<<<1, 1024, 0, stream>>>(tiles, ntiles, count_out);
  GG_LAUNCH_CHECK();
  if (ntiles > 0) {
    k_compact_scatter<<<ntiles, 256, 0, stream>>>(static_cast<const uint32_t *>(src),
                                                  static_cast<uint32_t *>(dst), flags, n,
                                                  elem_words, tiles);
    GG_LAUNCH_CHECK();
  }
  return GG_OK;
}

extern "C" int gg_fill_u32(void *dst, size_t words, uint32_t value, hipStream_t stream) {
  if (words == 0) return GG_OK;
  if (!dst) return GG_ERR_INVALID;
  size_t blocks = (words + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_fill_words<<<(unsigned)blocks, 256, 0, stream>>>(static_cast<uint32_t *>(dst), words, value);
  GG_LAUNCH_CHECK();
  return GG_OK;
}
