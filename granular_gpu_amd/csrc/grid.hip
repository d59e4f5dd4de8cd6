// Uniform-grid neighbour search: cell hash, stable counting sort by cell id,
// cell table (cellStart) and permutation of the particle payload.
//
// Reference: GranularSystem.cu:65-280 (neighbor_search_granular/boundary),
// CUDAFunctions.cuh:54-107 (mapParticles2Cells_CUDA, countingInCell_CUDA),
// thrust::sort_by_key / gather / exclusive_scan.  The sort here is a
// counting sort: integer atomics give each particle a slot in its cell, the
// cell table is the exclusive scan of the histogram, and a per-cell sort by
// old index then restores stability, so the permutation equals a stable
// sort by cell id (what thrust's radix sort_by_key returns) on every run.
#include <hip/hip_runtime.h>

#include "gg_device.hpp"
#include "gg_internal.hpp"
#include "grid_kernels.hpp"

namespace gg {

// key_i = hash(x_i) (or keys_in[i] when given: out-of-range keys go to the
// sentinel cell); slot_i = arrival order inside the cell.
__global__ void __launch_bounds__(kBlock) k_hash_count(const float *__restrict__ pos,
                                                       const int *__restrict__ keys_in, int n,
                                                       DevGrid g, int *__restrict__ key,
                                                       int *__restrict__ slot,
                                                       int *__restrict__ count,
                                                       int *__restrict__ ctr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    ctr[0] = 0;                // big-cell counter of this build
    ctr[kHugeCounter] = 0;     // cells for the multi-block path
  }
  // All lanes take part in the wave operations; lanes past n carry key -1.
  int c = -1;
  if (i < n) {
    if (keys_in) {
      const int C = g.cx * g.cy * g.cz;
      c = keys_in[i];
      if (c < 0 || c > C) c = C;
    } else {
      const Packed3 p3 = ld3(pos, i);
      c = hash_cell(p3.x, p3.y, p3.z, g);
    }
    key[i] = c;
  }
  // Wave-aggregated slot reservation: each run of equal keys in adjacent
  // lanes (the input is the previous step's cell order, so runs are long)
  // reserves its slots with one atomic; lanes take consecutive slots.
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(c, 1, 64);
  const bool head = (lane == 0) || (prev != c);
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = heads & ((2ull << lane) - 1ull);  // heads at lanes <= lane
  const int hl = 63 - __clzll((long long)upto);                       // my run's head lane
  const unsigned long long after = heads & ~((2ull << lane) - 1ull);  // heads above lane
  const int nl = after ? __ffsll((long long)after) - 1 : 64;          // next run's head lane
  int basev = 0;
  if (head && c >= 0) {
    // run length = lanes up to the next head (lanes past n are key -1 runs)
    basev = atomicAdd(&count[c], nl - lane);
  }
  basev = __shfl(basev, hl, 64);
  if (i < n) slot[i] = basev + (lane - hl);
}

// Block sums of kScanTile-element tiles of count[0..m).
__global__ void __launch_bounds__(kBlock) k_scan_reduce(const int *__restrict__ count, int m,
                                                        int *__restrict__ part) {
  __shared__ int wsum[kBlock / 64];
  const int base = blockIdx.x * kScanTile + threadIdx.x * 8;
  int s = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int idx = base + k;
    if (idx < m) s += count[idx];
  }
  s = wave_sum(s);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) wsum[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += wsum[w];
    part[blockIdx.x] = t;
  }
}

// Exclusive scan of the (<= kScanMaxParts) partials in one block.
__global__ void __launch_bounds__(1024) k_scan_parts(int *__restrict__ part, int np) {
  __shared__ int buf[kScanMaxParts];
  __shared__ int wtot[16];
  const int t = threadIdx.x;
  for (int k = t; k < np; k += 1024) buf[k] = part[k];
  __syncthreads();
  // each thread owns 8 consecutive entries
  const int b0 = t * 8;
  int loc[8];
  int s = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    loc[k] = (b0 + k < np) ? buf[b0 + k] : 0;
    s += loc[k];
  }
  // inclusive wave scan of per-thread sums
  const int lane = t & 63, wid = t >> 6;
  int inc = s;
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(inc, o, 64);
    if (lane >= o) inc += v;
  }
  if (lane == 63) wtot[wid] = inc;
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int w = 0; w < 16; ++w) {
      const int v = wtot[w];
      wtot[w] = run;
      run += v;
    }
  }
  __syncthreads();
  int run = wtot[wid] + inc - s;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (b0 + k < np) buf[b0 + k] = run;
    run += loc[k];
  }
  __syncthreads();
  for (int k = t; k < np; k += 1024) part[k] = buf[k];
}

// cell_start[c] = exclusive prefix of count; cell_start[m] = total.  Clears
// count for the next build.
__global__ void __launch_bounds__(kBlock) k_scan_apply(int *__restrict__ count, int m,
                                                       const int *__restrict__ part,
                                                       int *__restrict__ cell_start) {
  __shared__ int wtot[kBlock / 64];
  const int base = blockIdx.x * kScanTile + threadIdx.x * 8;
  int loc[8];
  int s = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int idx = base + k;
    loc[k] = idx < m ? count[idx] : 0;
    s += loc[k];
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int inc = s;
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(inc, o, 64);
    if (lane >= o) inc += v;
  }
  if (lane == 63) wtot[wid] = inc;
  __syncthreads();
  int woff = 0;
  for (int w = 0; w < wid; ++w) woff += wtot[w];
  int run = part[blockIdx.x] + woff + inc - s;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int idx = base + k;
    if (idx < m) {
      cell_start[idx] = run;
      count[idx] = 0;
      if (idx == m - 1) cell_start[m] = run + loc[k];
    }
    run += loc[k];
  }
}

__global__ void __launch_bounds__(kBlock) k_scatter(const int *__restrict__ key,
                                                    const int *__restrict__ slot, int n,
                                                    const int *__restrict__ cell_start,
                                                    int *__restrict__ perm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int d = cell_start[key[i]] + slot[i];
  if (d >= 0 && d < n) perm[d] = i;
}

// One thread per sorted slot d: raw[d] (the scatter's arrival order) goes
// to cell_start[c] + (number of entries of its cell c with a smaller old
// index), i.e. ascending old index within the cell.  Cells longer than
// kSmallCell are copied as they are and queued (by their first slot) for
// k_sort_big.
__global__ void __launch_bounds__(kBlock) k_rank_small(const int *__restrict__ raw, int n,
                                                       const int *__restrict__ key,
                                                       const int *__restrict__ cell_start,
                                                       int *__restrict__ perm,
                                                       int *__restrict__ big,
                                                       int *__restrict__ ctr) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= n) return;
  const int v = raw[d];
  const int c = key[v];
  const int s = cell_start[c], e = cell_start[c + 1];
  if (e - s > kSmallCell) {
    perm[d] = v;
    if (d == s) big[atomicAdd(&ctr[0], 1)] = c;
    return;
  }
  int rank = 0;
  for (int k = s; k < e; ++k) rank += raw[k] < v ? 1 : 0;
  perm[s + rank] = v;
}

// One block per queued cell (grid-stride over the queue).  Cells up to
// kBigLds entries: bitonic sort of the old indices in LDS.  Larger cells
// (e.g. the sentinel cell of particles outside the grid) go to the
// multi-block path: the block records the cell and the range [lo, hi) of its
// members' old indices in huge[] (k_huge_count / k_huge_write then compact
// {i in [lo, hi) : key[i] == c} in index order over many blocks).
__global__ void __launch_bounds__(kBlock) k_sort_big(const int *__restrict__ cell_start,
                                                     int *__restrict__ perm,
                                                     const int *__restrict__ big,
                                                     int *__restrict__ ctr,
                                                     int *__restrict__ huge) {
  __shared__ int buf[kBigLds];
  __shared__ int red[2][kBlock / 64];
  const int nbig = ctr[0];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int b = blockIdx.x; b < nbig; b += gridDim.x) {
    const int c = big[b];
    const int s = cell_start[c], e = cell_start[c + 1];
    const int len = e - s;
    if (len <= kBigLds) {
      int P = 1;
      while (P < len) P <<= 1;
      for (int k = threadIdx.x; k < P; k += blockDim.x) buf[k] = k < len ? perm[s + k] : 0x7fffffff;
      __syncthreads();
      for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
          for (int t = threadIdx.x; t < P; t += blockDim.x) {
            const int u = t ^ j;
            if (u > t) {
              const int x = buf[t], y = buf[u];
              const bool up = (t & k) == 0;
              if ((x > y) == up) {
                buf[t] = y;
                buf[u] = x;
              }
            }
          }
          __syncthreads();
        }
      }
      for (int k = threadIdx.x; k < len; k += blockDim.x) perm[s + k] = buf[k];
      __syncthreads();
    } else {
      int lo = 0x7fffffff, hi = -1;
      for (int k = s + threadIdx.x; k < e; k += blockDim.x) {
        const int v = perm[k];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
      }
      for (int o = 32; o > 0; o >>= 1) {
        const int a = __shfl_xor(lo, o, 64), z = __shfl_xor(hi, o, 64);
        lo = a < lo ? a : lo;
        hi = z > hi ? z : hi;
      }
      if (lane == 0) {
        red[0][wid] = lo;
        red[1][wid] = hi;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) {
          lo = red[0][w] < lo ? red[0][w] : lo;
          hi = red[1][w] > hi ? red[1][w] : hi;
        }
        const int h = atomicAdd(&ctr[kHugeCounter], 1);
        huge[3 * h] = c;
        huge[3 * h + 1] = lo;
        huge[3 * h + 2] = hi + 1;
      }
      __syncthreads();
    }
  }
}

// Chunk p of kHugeChunks equal parts of [lo, hi).
__device__ __forceinline__ void huge_chunk(int lo, int hi, int p, int &a, int &b) {
  const long long w = (long long)hi - lo;
  a = lo + (int)(w * p / kHugeChunks);
  b = lo + (int)(w * (p + 1) / kHugeChunks);
}

// Block p counts the members of each recorded cell in its chunk of the
// cell's range: cnt[h * kHugeChunks + p].
__global__ void __launch_bounds__(kBlock) k_huge_count(const int *__restrict__ key,
                                                       const int *__restrict__ ctr,
                                                       const int *__restrict__ huge,
                                                       int *__restrict__ cnt) {
  __shared__ int red[kBlock / 64];
  const int nh = ctr[kHugeCounter];
  const int p = blockIdx.x, lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int h = 0; h < nh; ++h) {
    const int c = huge[3 * h];
    int a = 0, b = 0;
    huge_chunk(huge[3 * h + 1], huge[3 * h + 2], p, a, b);
    int m = 0;
    for (int i = a + threadIdx.x; i < b; i += blockDim.x) m += key[i] == c ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) m += __shfl_xor(m, o, 64);
    if (lane == 0) red[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int w = 0; w < kBlock / 64; ++w) t += red[w];
      cnt[h * kHugeChunks + p] = t;
    }
    __syncthreads();
  }
}

// Block p writes the members of each recorded cell found in its chunk, in
// index order, after those of the chunks before it: a stable compaction of
// {i in [lo, hi) : key[i] == c} into the cell's slots.
__global__ void __launch_bounds__(kBlock) k_huge_write(const int *__restrict__ key,
                                                       const int *__restrict__ cell_start,
                                                       const int *__restrict__ ctr,
                                                       const int *__restrict__ huge,
                                                       const int *__restrict__ cnt,
                                                       int *__restrict__ perm) {
  __shared__ int wcnt[kBlock / 64];
  const int nh = ctr[kHugeCounter];
  const int p = blockIdx.x, lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int h = 0; h < nh; ++h) {
    const int c = huge[3 * h];
    const int s = cell_start[c], len = cell_start[c + 1] - s;
    int a = 0, b = 0;
    huge_chunk(huge[3 * h + 1], huge[3 * h + 2], p, a, b);
    // members in the chunks before this one
    int before = 0;
    for (int q = threadIdx.x; q < p; q += blockDim.x) before += cnt[h * kHugeChunks + q];
    for (int o = 32; o > 0; o >>= 1) before += __shfl_xor(before, o, 64);
    if (lane == 0) wcnt[wid] = before;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < kBlock / 64; ++w) base += wcnt[w];
    __syncthreads();
    for (int t0 = a; t0 < b; t0 += 4 * kBlock) {
      // thread owns keys t0 + 4*tid .. +3 (index order within the tile)
      const int i0 = t0 + 4 * threadIdx.x;
      int m[4];
      int cn = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        m[q] = (i0 + q < b && key[i0 + q] == c) ? 1 : 0;
        cn += m[q];
      }
      int inc = cn;
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
      }
      if (lane == 63) wcnt[wid] = inc;
      __syncthreads();
      int woff = 0, tot = 0;
      for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wid) woff += wcnt[w];
        tot += wcnt[w];
      }
      int at = base + woff + inc - cn;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (m[q]) {
          if (at < len) perm[s + at] = i0 + q;
          ++at;
        }
      }
      base += tot;
      __syncthreads();
    }
  }
}

// dst[k] = src[perm[k]] for records of `words` 32-bit words.
__global__ void __launch_bounds__(kBlock) k_gather_words(const uint32_t *__restrict__ src,
                                                         uint32_t *__restrict__ dst,
                                                         const int *__restrict__ perm, int n,
                                                         int words) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int o = perm[k];
  for (int w = 0; w < words; ++w) dst[(size_t)k * words + w] = src[(size_t)o * words + w];
}

__global__ void __launch_bounds__(kBlock) k_gather_keys(const int *__restrict__ key,
                                                        const int *__restrict__ perm, int n,
                                                        int *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  out[k] = key[perm[k]];
}

__global__ void __launch_bounds__(kBlock) k_pack_boundary(const float *__restrict__ pos, int n,
                                                          float rb, float4 *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  out[k] = make_float4(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2], rb);
}

static inline int blocks_for(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// Hash (or the given keys), counting sort, cell table and perm for `n`
// particles.
int sort_by_cell_keys(const float *pos, const int *keys_in, int n, int *cell_start,
                      const gg_workspace *ws, hipStream_t st) {
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                         ws->bmax);
  const DevGrid g = dev_grid(ws->grid);
  const int m = (int)(grid_cells(ws->grid) + 1);  // C+1 bins incl. sentinel
  int *count = at<int>(ws, L.count);
  int *key = at<int>(ws, L.key);
  int *slot = at<int>(ws, L.slot);
  int *perm = at<int>(ws, L.perm);
  int *part = at<int>(ws, L.part);
  int *big = at<int>(ws, L.big);
  int *ctr = at<int>(ws, L.ctr);
  const int nb = n > 0 ? blocks_for(n) : 1;
  k_hash_count<<<nb, kBlock, 0, st>>>(pos, keys_in, n, g, key, slot, count, ctr);
  GG_LAUNCH_CHECK();
  const int np = (m + kScanTile - 1) / kScanTile;
  k_scan_reduce<<<np, kBlock, 0, st>>>(count, m, part);
  GG_LAUNCH_CHECK();
  k_scan_parts<<<1, 1024, 0, st>>>(part, np);
  GG_LAUNCH_CHECK();
  k_scan_apply<<<np, kBlock, 0, st>>>(count, m, part, cell_start);
  GG_LAUNCH_CHECK();
  if (n > 0) {
    int *raw = at<int>(ws, L.itmp);
    k_scatter<<<blocks_for(n), kBlock, 0, st>>>(key, slot, n, cell_start, raw);
    GG_LAUNCH_CHECK();
    k_rank_small<<<blocks_for(n), kBlock, 0, st>>>(raw, n, key, cell_start, perm, big, ctr);
    GG_LAUNCH_CHECK();
  }
  // the huge-cell records and per-chunk counts live in the arrival-slot
  // array, which nothing reads after the scatter: (3 + kHugeChunks) ints per
  // cell of more than kBigLds particles (at most n / kBigLds of them)
  int *huge = slot;
  int *hcnt = slot + 3 * (n / kBigLds + 1);
  k_sort_big<<<256, kBlock, 0, st>>>(cell_start, perm, big, ctr, huge);
  GG_LAUNCH_CHECK();
  if (n > kBigLds) {
    k_huge_count<<<kHugeChunks, kBlock, 0, st>>>(key, ctr, huge, hcnt);
    GG_LAUNCH_CHECK();
    k_huge_write<<<kHugeChunks, kBlock, 0, st>>>(key, cell_start, ctr, huge, hcnt, perm);
    GG_LAUNCH_CHECK();
  }
  return GG_OK;
}

int sort_by_cell(const float *pos, int n, int *cell_start, const gg_workspace *ws,
                 hipStream_t st) {
  return sort_by_cell_keys(pos, nullptr, n, cell_start, ws, st);
}

// out[k] = sum of in[0..k) for k in [0, m], in[] cleared.  m <= kScanTile *
// kScanMaxParts (the workspace's partials array holds the tile sums).
int scan_exclusive(int *in, int m, int *out, const gg_workspace *ws, hipStream_t st) {
  if (m <= 0) return GG_OK;
  if ((long long)m > (long long)kScanTile * kScanMaxParts) return GG_ERR_CAPACITY;
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                         ws->bmax);
  int *part = at<int>(ws, L.part);
  const int np = (m + kScanTile - 1) / kScanTile;
  k_scan_reduce<<<np, kBlock, 0, st>>>(in, m, part);
  GG_LAUNCH_CHECK();
  k_scan_parts<<<1, 1024, 0, st>>>(part, np);
  GG_LAUNCH_CHECK();
  k_scan_apply<<<np, kBlock, 0, st>>>(in, m, part, out);
  GG_LAUNCH_CHECK();
  return GG_OK;
}

int gather_inplace(void *arr, int words, const int *perm, int n, void *tmp, hipStream_t st) {
  if (!arr || n <= 0) return GG_OK;
  k_gather_words<<<blocks_for(n), kBlock, 0, st>>>(static_cast<const uint32_t *>(arr),
                                                   static_cast<uint32_t *>(tmp), perm, n, words);
  GG_LAUNCH_CHECK();
  return hip_status(
      hipMemcpyAsync(arr, tmp, (size_t)n * words * 4, hipMemcpyDeviceToDevice, st));
}

int grid_build_general(gg_particles *set, gg_adaptive *ad, int *cell_start, int *perm_out,
                       const gg_workspace *ws, hipStream_t st, int capacity_limit,
                       const int *keys) {
  if (!set || !cell_start) return GG_ERR_INVALID;
  GG_TRY(check_ws(ws));
  const int n = set->n;
  if (n < 0) return GG_ERR_INVALID;
  if (n > capacity_limit) return GG_ERR_CAPACITY;
  if (n > 0 && !set->pos) return GG_ERR_INVALID;
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                         ws->bmax);
  GG_TRY(sort_by_cell_keys(reinterpret_cast<const float *>(set->pos), keys, n, cell_start, ws,
                           st));
  const int *perm = at<int>(ws, L.perm);
  const int *key = at<int>(ws, L.key);
  if (n == 0) return GG_OK;
  void *t3 = at<void>(ws, L.v3tmp0);
  void *t1 = at<void>(ws, L.itmp0);
  GG_TRY(gather_inplace(set->pos, 3, perm, n, t3, st));
  GG_TRY(gather_inplace(set->vel, 3, perm, n, t3, st));
  GG_TRY(gather_inplace(set->mass, 1, perm, n, t1, st));
  GG_TRY(gather_inplace(set->is_animated, 1, perm, n, t1, st));
  if (ad) {
    GG_TRY(gather_inplace(ad->merge, 1, perm, n, t1, st));
    GG_TRY(gather_inplace(ad->merge_count, 1, perm, n, t1, st));
    GG_TRY(gather_inplace(ad->remove, 1, perm, n, t1, st));
  }
  if (set->particle_2_cell) {
    k_gather_keys<<<blocks_for(n), kBlock, 0, st>>>(key, perm, n, set->particle_2_cell);
    GG_LAUNCH_CHECK();
  }
  if (perm_out)
    GG_TRY(hip_status(
        hipMemcpyAsync(perm_out, perm, (size_t)n * 4, hipMemcpyDeviceToDevice, st)));
  return GG_OK;
}

}  // namespace gg

using namespace gg;

extern "C" int gg_grid_build(gg_particles *set, gg_adaptive *ad, int *cell_start, int *perm_out,
                             const gg_workspace *ws, hipStream_t stream) {
  if (!ws) return GG_ERR_INVALID;
  return grid_build_general(set, ad, cell_start, perm_out, ws, stream, ws->capacity);
}

namespace gg {
int pack_boundary(const gg_particles *boundary, const gg_workspace *ws, hipStream_t st) {
  if (boundary->n <= 0) return GG_OK;
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                         ws->bmax);
  gg_params dp;
  gg_params_default(&dp);
  k_pack_boundary<<<blocks_for(boundary->n), kBlock, 0, st>>>(
      reinterpret_cast<const float *>(boundary->pos), boundary->n, dp.boundary_radius,
      at<float4>(ws, L.bnd4));
  GG_LAUNCH_CHECK();
  return GG_OK;
}
}  // namespace gg

extern "C" int gg_boundary_build(gg_particles *boundary, int *bcell_start,
                                 const gg_workspace *ws, hipStream_t stream) {
  if (!boundary || !bcell_start || !ws) return GG_ERR_INVALID;
  GG_TRY(grid_build_general(boundary, nullptr, bcell_start, nullptr, ws, stream,
                            ws->boundary_capacity));
  return pack_boundary(boundary, ws, stream);
}
