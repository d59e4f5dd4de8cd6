// Learned upsampler on gfx950: kNN-16 features, the one-step teacher (the
// plain upsampler applied to a 2x2x2 template), BF16 MFMA training kernels
// with FP32 accumulation, and the fused inference kernel.
//
// Definitions shared with the oracle (features, flags, teacher, network,
// mixed precision, loss) are stated in the public header.  No reference
// code: the learned upsampler is an addition beside update_upsampled_cuda
// (Solver.cu:1244-1387), whose update is the teacher.
//
// MFMA: v_mfma_f32_16x16x16_bf16 (one wave, 64 lanes).  Lane l = c + 16 q
// (c = l % 16, q = l / 16) holds A[c][4q..4q+3], B[4q..4q+3][c] and
// D[4q..4q+3][c].  The inference kernel computes each layer transposed,
// Z^T = W H^T (weights as A, activations as B): the D tile of output
// features 16t..16t+15 is then exactly the B operand of k-step t of the
// next layer, so activations never leave registers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "gg_device.hpp"
#include "gg_internal.hpp"
#include "profiler.hpp"
#include "upsample_dev.hpp"

namespace gg {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 96, kHid = 128, kOut = 24, kOutPad = 32, kKnn = 16;
constexpr int kWOff1 = 0, kWOff2 = 12288, kWOff3 = 28672, kWOff4 = 45056;
constexpr int kBOff1 = 0, kBOff2 = 128, kBOff3 = 256, kBOff4 = 384;

// Round to the nearest bfloat16 (ties to even); NaN stays NaN.
__device__ __forceinline__ uint16_t to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u | 0x00400000u) >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2v __attribute__((ext_vector_type(2)));

// Two floats rounded to bfloat16 by the hardware (v_cvt_pk_bf16_f32), lo in
// the low half.  Equal to to_bf16 for every non-NaN input:
// gg_bf16_round_check compares the two over all 2^32 bit patterns.
__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
  const f32x2v v = {lo, hi};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2v));
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));

// v_mfma_f32_16x16x32_bf16: lane l = c + 16 q holds 8 elements of row c of
// A and of column c of B at k-slots owned by q, and D[4q..4q+3][c].
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8v, a),
                                                 __builtin_bit_cast(bf16x8v, b), c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 cat8(bf16x4 lo, bf16x4 hi) {
  return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

__device__ __forceinline__ float templ(int k, int axis) {
  return ((k >> (2 - axis)) & 1) ? 0.5f : -0.5f;
}

struct MlpConst {
  float radius, vscale, guard_boundary, guard_floor, guard_side;
};

// ---- fallback flags -----------------------------------------------------------
// flag bit 1: a boundary particle of the stencil within guard_boundary;
// bit 2: below guard_floor or near the x = 0 / z = 0 planes.  Flagged
// particles are appended to a compacted list (order is irrelevant: each one
// writes only its own outputs).
__device__ __forceinline__ int guard_flag(const Packed3 &xi, const float *__restrict__ bpos,
                                          const int *__restrict__ bcs, const DevGrid &g,
                                          const MlpConst &c) {
  int flag = 0;
  const float gb2 = c.guard_boundary * c.guard_boundary;
  int bx, by, bz;
  if (floor_cell(xi.x, xi.y, xi.z, g, bx, by, bz)) {
    stencil_rows(bx, by, bz, g, bcs, [&](int k0, int k1) {
      int k = k0;
      for (; k + 4 <= k1; k += 4) {
        const Packed3 q0 = ld3(bpos, k), q1 = ld3(bpos, k + 1), q2 = ld3(bpos, k + 2),
                      q3 = ld3(bpos, k + 3);
        auto near = [&](const Packed3 &q) {
          const float dx = q.x - xi.x, dy = q.y - xi.y, dz = q.z - xi.z;
          return dx * dx + dy * dy + dz * dz < gb2;
        };
        if (near(q0) || near(q1) || near(q2) || near(q3)) flag |= 1;
      }
      for (; k < k1; ++k) {
        const Packed3 q = ld3(bpos, k);
        const float dx = q.x - xi.x, dy = q.y - xi.y, dz = q.z - xi.z;
        if (dx * dx + dy * dy + dz * dz < gb2) flag |= 1;
      }
    });
  }
  if (xi.y < c.guard_floor || xi.x < c.guard_side || xi.z < c.guard_side) flag |= 2;
  return flag;
}

// Appends the wave's flagged particles to the compacted list.
__device__ __forceinline__ void list_flagged(int i, int flag, int *__restrict__ fb_list,
                                             int *__restrict__ fb_count) {
  const unsigned long long m = __ballot(flag != 0);
  if (flag) {
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    int slot = 0;
    if (lane == leader) slot = atomicAdd(fb_count, __popcll(m));
    slot = __shfl(slot, leader, 64) + __popcll(m & ((1ull << lane) - 1ull));
    fb_list[slot] = i;
  }
}

__global__ void __launch_bounds__(256)
    k_mlp_flags(int n, const float *__restrict__ pos, const float *__restrict__ bpos,
                const int *__restrict__ bcs, DevGrid g, MlpConst c, int *__restrict__ flags,
                int *__restrict__ fb_list, int *__restrict__ fb_count) {
  const int i = xcd_block(blockIdx.x, gridDim.x, 1) * blockDim.x + threadIdx.x;
  int flag = 0;
  if (i < n) {
    flag = guard_flag(ld3(pos, i), bpos, bcs, g, c);
    flags[i] = flag;
  }
  if (fb_list) list_flagged(i, flag, fb_list, fb_count);
}

// ---- features ---------------------------------------------------------------
// Stencil rows (the three cells dz = -1..1 of one (dx, dy), one contiguous
// index range) nearest-first: centre, the four faces, the four corners, as
// r = 3 (dx + 1) + (dy + 1).  The 16 smallest keys are a set, so the
// visiting order does not change them; near rows first let later chunks be
// skipped when no lane can improve.
__constant__ int kRowNearFirst[9] = {4, 1, 3, 5, 7, 0, 2, 6, 8};

// Compare-exchange of 64-bit keys: a <- min, b <- max.
__device__ __forceinline__ void kv_ce(unsigned long long &a, unsigned long long &b) {
  const bool sw = b < a;
  const unsigned long long lo = sw ? b : a, hi = sw ? a : b;
  a = lo;
  b = hi;
}

// Batcher's odd-even merge sort of 16 keys, ascending: 63 compare-exchanges
// (pairs (a, b), a < b, the minimum to a), where a bitonic network needs 80.
// A test checks it on all 2^16 zero-one inputs (0-1 principle).
constexpr unsigned char kOddEvenMerge16[126] = {
    0,  1,  2,  3,  0,  2,  1,  3,  1,  2,  4,  5,  6,  7,  4,  6,  5,  7,  5,  6,  0,
    4,  2,  6,  2,  4,  1,  5,  3,  7,  3,  5,  1,  2,  3,  4,  5,  6,  8,  9,  10, 11,
    8,  10, 9,  11, 9,  10, 12, 13, 14, 15, 12, 14, 13, 15, 13, 14, 8,  12, 10, 14, 10,
    12, 9,  13, 11, 15, 11, 13, 9,  10, 11, 12, 13, 14, 0,  8,  4,  12, 4,  8,  2,  10,
    6,  14, 6,  10, 2,  4,  6,  8,  10, 12, 1,  9,  5,  13, 5,  9,  3,  11, 7,  15, 7,
    11, 3,  5,  7,  9,  11, 13, 1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14};

__device__ __forceinline__ void kv_sort16(unsigned long long (&v)[16]) {
#pragma unroll
  for (int k = 0; k < 63; ++k) kv_ce(v[kOddEvenMerge16[2 * k]], v[kOddEvenMerge16[2 * k + 1]]);
}

// Batcher's network for 8 keys (19 compare-exchanges), checked the same way
// on all 2^8 zero-one inputs.
constexpr unsigned char kOddEvenMerge8[38] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 2, 1, 3, 4,
                                              6, 5, 7, 1, 2, 5, 6, 0, 4, 1, 5, 2, 6,
                                              3, 7, 2, 4, 3, 5, 1, 2, 3, 4, 5, 6};

__device__ __forceinline__ void kv_sort8(unsigned long long (&v)[8]) {
#pragma unroll
  for (int k = 0; k < 19; ++k) kv_ce(v[kOddEvenMerge8[2 * k]], v[kOddEvenMerge8[2 * k + 1]]);
}

// top <- the 16 smallest keys of top and c (both ascending), ascending:
// min(top[i], c[15 - i]) holds them as a bitonic sequence, which a
// bitonic merger sorts (16 minima + 32 compare-exchanges).
__device__ __forceinline__ void kv_merge16(unsigned long long (&top)[16],
                                           const unsigned long long (&c)[16]) {
#pragma unroll
  for (int i = 0; i < 16; ++i) top[i] = c[15 - i] < top[i] ? c[15 - i] : top[i];
#pragma unroll
  for (int j = 8; j > 0; j >>= 1)
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int l = i ^ j;
      if (l > i) kv_ce(top[i], top[l]);
    }
}

// The same with 8 new keys (c padded with the largest key: the first 8
// minima keep top[0..7]).
__device__ __forceinline__ void kv_merge8(unsigned long long (&top)[16],
                                          const unsigned long long (&c)[8]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) top[8 + i] = c[7 - i] < top[8 + i] ? c[7 - i] : top[8 + i];
#pragma unroll
  for (int j = 8; j > 0; j >>= 1)
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int l = i ^ j;
      if (l > i) kv_ce(top[i], top[l]);
    }
}

// One thread per coarse particle: the 16 smallest (d2, j) of the stencil as
// 64-bit keys (d2 bits << 32 | j; d2 >= 0 orders like its bits).  The
// candidates come in chunks of 16; a chunk is sorted by a sorting network
// and merged into the kept 16, the same work on every lane (a per-candidate
// insertion runs whenever any lane of the wave inserts).
// GROUPED = 0: feat[i][96] (sample-major, the training layout);
// GROUPED = 1: feat[g][i][4] for the 24 groups g of 4 consecutive features
// (the inference layout: a wave's stores and the network's B-operand loads
// are contiguous).
// CHUNK: candidates per selection step (16, or 8 with fewer registers:
// 96 instead of 114, 5 waves per SIMD instead of 4).
template <int GROUPED, int CHUNK>
__global__ void __launch_bounds__(256)
    k_mlp_features(int n, const float *__restrict__ pos, const float *__restrict__ vel,
                   const int *__restrict__ cs, DevGrid g, MlpConst c,
                   uint16_t *__restrict__ feat, int *__restrict__ knn) {
  // consecutive blocks stay on one XCD (spatial neighbours share its L2)
  const int i = xcd_block(blockIdx.x, gridDim.x, 1) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Packed3 xi = ld3(pos, i);
  const float xx = xi.x, xy = xi.y, xz = xi.z;
  constexpr unsigned long long kNone = ~0ull;
  static_assert(kKnn == 16, "the selection network is built for 16 keys");
  unsigned long long key[kKnn];
#pragma unroll
  for (int s = 0; s < kKnn; ++s) key[s] = kNone;
  int bx = 0, by = 0, bz = 0;
  const bool ok = floor_cell(xx, xy, xz, g, bx, by, bz);
  const int z0 = bz - 1 > 0 ? bz - 1 : 0;
  const int z1 = bz + 1 < g.cz - 1 ? bz + 1 : g.cz - 1;
  // index range of row r of kRowNearFirst (empty outside the grid)
  auto row_range = [&](int r, int &a, int &b) {
    a = 0;
    b = 0;
    if (r >= 9) return;
    const int rr = kRowNearFirst[r];
    const int x = bx + rr / 3 - 1, y = by + rr % 3 - 1;
    if (x >= 0 && x < g.cx && y >= 0 && y < g.cy) {
      const int c0 = (x * g.cy + y) * g.cz;
      a = cs[c0 + z0];
      b = cs[c0 + z1 + 1];
    }
  };
  int r = ok && z0 <= z1 ? 0 : 9;  // row whose range is in (na, nb)
  int na = 0, nb = 0;              // loaded one row ahead of use
  row_range(r, na, nb);
  int j = 0, j1 = 0;               // rest of the current row
  for (bool first = true;; first = false) {
    // the next CHUNK candidates (index -1 past the last row)
    int jj[CHUNK];
#pragma unroll
    for (int s = 0; s < CHUNK; ++s) {
      while (j >= j1 && r < 9) {
        j = na;
        j1 = nb;
        row_range(++r, na, nb);
      }
      jj[s] = j < j1 ? j++ : -1;
    }
    unsigned long long c[CHUNK];
    bool need = false;
#pragma unroll
    for (int s = 0; s < CHUNK; ++s) {
      const int js = jj[s];
      const Packed3 pj = ld3(pos, js >= 0 ? js : i);
      const float dx = pj.x - xx, dy = pj.y - xy, dz = pj.z - xz;
      const float d2 = dx * dx + dy * dy + dz * dz;
      // finite d2 only (NaN and inf never enter), and not the particle itself
      const bool take = js >= 0 && js != i && d2 < __builtin_inff();
      c[s] = take ? ((unsigned long long)__float_as_uint(d2) << 32) | (unsigned)js : kNone;
      need |= c[s] < key[kKnn - 1];
    }
    if (first) {
      // nothing kept yet: the sorted chunk is the merge's result
      if constexpr (CHUNK == 16) kv_sort16(c); else kv_sort8(c);
#pragma unroll
      for (int s = 0; s < CHUNK; ++s) key[s] = c[s];
    } else if (__ballot(need)) {
      if constexpr (CHUNK == 16) {
        kv_sort16(c);
        kv_merge16(key, c);
      } else {
        kv_sort8(c);
        kv_merge8(key, c);
      }
    }
    if (!__ballot(j < j1 || r < 9)) break;
  }
  int kj[kKnn];
#pragma unroll
  for (int s = 0; s < kKnn; ++s) kj[s] = key[s] == kNone ? -1 : (int)(unsigned)(key[s] & 0xffffffffu);
  const Packed3 vi = ld3(vel, i);
  const float vx = vi.x, vy = vi.y, vz = vi.z;
  const float inv_r = 1.0f / c.radius, inv_v = 1.0f / c.vscale;
  // Neighbours in groups of kGather: a group's positions and velocities are
  // loaded together (an empty slot loads the particle itself and writes
  // zeros), so that the features wait for kKnn / kGather rounds of loads,
  // not 2 kKnn.
  constexpr int kGather = 4;
#pragma unroll
  for (int h = 0; h < kKnn; h += kGather) {
    Packed3 pj[kGather], vj[kGather];
#pragma unroll
    for (int u = 0; u < kGather; ++u) {
      const int j = kj[h + u];
      const int jl = j >= 0 ? j : i;
      pj[u] = ld3(pos, jl);
      vj[u] = ld3(vel, jl);
    }
    // two neighbour slots = 12 features = 3 groups of 4
#pragma unroll
    for (int s = h; s < h + kGather; s += 2) {
      float f[12];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int j = kj[s + u];
        if (knn) knn[(size_t)i * kKnn + s + u] = j;
        const Packed3 &q = pj[s - h + u], &w = vj[s - h + u];
        const bool has = j >= 0;
        f[6 * u + 0] = has ? (q.x - xx) * inv_r : 0.0f;
        f[6 * u + 1] = has ? (q.y - xy) * inv_r : 0.0f;
        f[6 * u + 2] = has ? (q.z - xz) * inv_r : 0.0f;
        f[6 * u + 3] = has ? (w.x - vx) * inv_v : 0.0f;
        f[6 * u + 4] = has ? (w.y - vy) * inv_v : 0.0f;
        f[6 * u + 5] = has ? (w.z - vz) * inv_v : 0.0f;
      }
      // hardware round-to-nearest-even (the defined rounding on every non-NaN
      // value; a NaN stays a NaN)
#pragma unroll
      for (int gq = 0; gq < 3; ++gq) {
        const uint2 wv = make_uint2(pack_bf16x2(f[4 * gq], f[4 * gq + 1]),
                                    pack_bf16x2(f[4 * gq + 2], f[4 * gq + 3]));
        const int grp = 3 * (s / 2) + gq;  // features 4 grp .. 4 grp + 3
        if (GROUPED) {
          reinterpret_cast<uint2 *>(feat)[(size_t)grp * n + i] = wv;
        } else {
          reinterpret_cast<uint2 *>(feat + (size_t)i * kIn)[grp] = wv;
        }
      }
    }
  }
}

// ---- teacher ------------------------------------------------------------------
// One thread per (coarse particle, template slot).  only_flagged: skip the
// particles whose flag is 0 (the inference fallback).
__device__ __forceinline__ void teacher_slot(int i, int k, const float *__restrict__ pos,
                                             const float *__restrict__ vel,
                                             const int *__restrict__ cs,
                                             const float *__restrict__ bpos,
                                             const int *__restrict__ bcs, const DevGrid &g,
                                             const UpConst &u, float radius,
                                             float *__restrict__ y, float *__restrict__ fine_pos,
                                             float *__restrict__ fine_vel) {
  const float vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
  const float px = pos[3 * i] + radius * templ(k, 0);
  const float py = pos[3 * i + 1] + radius * templ(k, 1);
  const float pz = pos[3 * i + 2] + radius * templ(k, 2);
  float nx, ny, nz, wx, wy, wz;
  upsample_one(px, py, pz, vx, vy, vz, pos, vel, cs, bpos, bcs, g, u, nx, ny, nz, wx, wy, wz);
  const size_t o = (size_t)i * kOut + 3 * k;
  if (y) {
    y[o] = wx - vx;
    y[o + 1] = wy - vy;
    y[o + 2] = wz - vz;
  }
  if (fine_pos) {
    fine_pos[o] = nx;
    fine_pos[o + 1] = ny;
    fine_pos[o + 2] = nz;
  }
  if (fine_vel) {
    fine_vel[o] = wx;
    fine_vel[o + 1] = wy;
    fine_vel[o + 2] = wz;
  }
}

// Fallback: one thread per (listed particle, slot), grid-stride over the
// device-side count of the compacted flagged list.
__global__ void __launch_bounds__(256)
    k_mlp_fallback(const int *__restrict__ list, const int *__restrict__ count,
                   const float *__restrict__ pos, const float *__restrict__ vel,
                   const int *__restrict__ cs, const float *__restrict__ bpos,
                   const int *__restrict__ bcs, DevGrid g, UpConst u, float radius,
                   float *__restrict__ fine_pos, float *__restrict__ fine_vel) {
  const long long total = 8LL * (*count);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x)
    teacher_slot(list[t >> 3], (int)(t & 7), pos, vel, cs, bpos, bcs, g, u, radius, nullptr,
                 fine_pos, fine_vel);
}

// Fine-run targets: one thread per (coarse particle, slot).  The fine
// particle nearest to the template point among the 27 fine cells around it,
// ties to the lower index (so the visiting order does not matter).
__global__ void __launch_bounds__(256)
    k_mlp_fine_targets(int n, const float *__restrict__ pos, const float *__restrict__ vel,
                       const float *__restrict__ fpos, const float *__restrict__ fvel,
                       const int *__restrict__ fcs, DevGrid fg, float radius,
                       float *__restrict__ y, float *__restrict__ d2, int *__restrict__ jm) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 8LL * n) return;
  const int i = (int)(t >> 3), k = (int)(t & 7);
  const float qx = pos[3 * i] + radius * templ(k, 0);
  const float qy = pos[3 * i + 1] + radius * templ(k, 1);
  const float qz = pos[3 * i + 2] + radius * templ(k, 2);
  float best = INFINITY;
  int bj = -1;
  int bx, by, bz;
  if (floor_cell(qx, qy, qz, fg, bx, by, bz)) {
    stencil_rows(bx, by, bz, fg, fcs, [&](int j0, int j1) {
      for (int j = j0; j < j1; ++j) {
        const Packed3 f = ld3(fpos, j);
        const float dx = f.x - qx, dy = f.y - qy, dz = f.z - qz;
        const float dd = dx * dx + dy * dy + dz * dz;
        if (dd < best || (dd == best && j < bj)) {
          best = dd;
          bj = j;
        }
      }
    });
  }
  const size_t o = (size_t)i * kOut + 3 * k;
  if (bj < 0) {
    y[o] = y[o + 1] = y[o + 2] = 0.0f;
  } else {
    const Packed3 w = ld3(fvel, bj);
    y[o] = w.x - vel[3 * i];
    y[o + 1] = w.y - vel[3 * i + 1];
    y[o + 2] = w.z - vel[3 * i + 2];
  }
  if (d2) d2[t] = best;
  if (jm) jm[t] = bj;
}

__global__ void k_zero_int(int *p) { *p = 0; }
__global__ void k_copy_int(const int *src, int *dst) { *dst = *src; }

__global__ void __launch_bounds__(256)
    k_mlp_teacher(int n, const float *__restrict__ pos, const float *__restrict__ vel,
                  const int *__restrict__ cs, const float *__restrict__ bpos,
                  const int *__restrict__ bcs, DevGrid g, UpConst u, float radius,
                  const int *__restrict__ flags, float *__restrict__ y,
                  float *__restrict__ fine_pos, float *__restrict__ fine_vel) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 8LL * n) return;
  const int i = (int)(t >> 3), k = (int)(t & 7);
  if (flags && flags[i] == 0) return;
  teacher_slot(i, k, pos, vel, cs, bpos, bcs, g, u, radius, y, fine_pos, fine_vel);
}

// ---- training: generic bf16 GEMM (fp32 out) ---------------------------------
// C[m][n] = sum_k A(m,k) B(k,n), A(m,k) = A[m sam + k sak],
// B(k,n) = B[k sbk + n sbn]; one wave per 16x16 tile of C, 4 waves a block.
// Bounds are checked per element (zero fill), so any M, N, K work.
struct GemmArgs {
  const uint16_t *A;
  const uint16_t *B;
  float *C;
  int M, N, K;
  long long sam, sak, sbk, sbn;
  int ldc;
};

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs a) {
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  const int tiles_n = (a.N + 15) / 16;
  const int tiles = ((a.M + 15) / 16) * tiles_n;
  if (wave >= tiles) return;  // whole waves exit together
  const int m0 = (wave / tiles_n) * 16, n0 = (wave % tiles_n) * 16;
  const int c = lane & 15, q = lane >> 4;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < a.K; k0 += 16) {
    bf16x4 av, bv;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = k0 + 4 * q + e;
      const int m = m0 + c, nn = n0 + c;
      av[e] = (m < a.M && k < a.K) ? (short)a.A[m * a.sam + k * a.sak] : (short)0;
      bv[e] = (nn < a.N && k < a.K) ? (short)a.B[k * a.sbk + nn * a.sbn] : (short)0;
    }
    acc = mfma16(av, bv, acc);
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int m = m0 + 4 * q + e, nn = n0 + c;
    if (m < a.M && nn < a.N) a.C[(size_t)m * a.ldc + nn] = acc[e];
  }
}

// Z = C + bias (fp32, kept for the backward mask); H = bf16(relu(Z)).
// last: only Z (the network output y), N = 24 of the padded 32 columns.
__global__ void k_bias_act(int B, int N, int ldc, const float *__restrict__ C,
                           const float *__restrict__ bias, float *__restrict__ Z,
                           uint16_t *__restrict__ H, int last) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)B * N) return;
  const int b = (int)(t / N), nn = (int)(t % N);
  const float z = C[(size_t)b * ldc + nn] + bias[nn];
  Z[(size_t)b * N + nn] = z;
  if (!last) H[(size_t)b * N + nn] = to_bf16(z > 0.0f ? z : 0.0f);
}

// g = 2 (y - t) / (24 B) on the 24 outputs, 0 on the 8 padding columns;
// G = bf16(g); sq[b] = sum of the sample's squared errors.
__global__ void k_loss_grad(int B, const float *__restrict__ y, const float *__restrict__ t,
                            float *__restrict__ g, uint16_t *__restrict__ G,
                            float *__restrict__ sq) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float scale = 2.0f / (float)(kOut * B);
  float s = 0.0f;
  for (int nn = 0; nn < kOutPad; ++nn) {
    float gv = 0.0f;
    if (nn < kOut) {
      const float e = y[(size_t)b * kOut + nn] - t[(size_t)b * kOut + nn];
      s = s + e * e;
      gv = e * scale;
    }
    g[(size_t)b * kOutPad + nn] = gv;
    G[(size_t)b * kOutPad + nn] = to_bf16(gv);
  }
  if (sq) sq[b] = s;
}

// g_prev = D * [Z > 0]; G_prev = bf16(g_prev).
__global__ void k_mask(long long count, const float *__restrict__ D, const float *__restrict__ Z,
                       float *__restrict__ g, uint16_t *__restrict__ G) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const float v = Z[t] > 0.0f ? D[t] : 0.0f;
  g[t] = v;
  G[t] = to_bf16(v);
}

// db[n] = sum over b of g[b][n]: one block per column; thread t sums
// samples t, t + 256, ... in order, then a fixed tree in LDS (the same
// order on every run).
__global__ void __launch_bounds__(256) k_colsum(int B, int N, const float *__restrict__ g,
                                                float *__restrict__ db) {
  __shared__ float part[256];
  const int nn = blockIdx.x;
  float s = 0.0f;
  for (int b = threadIdx.x; b < B; b += 256) s = s + g[(size_t)b * N + nn];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] = part[threadIdx.x] + part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[nn] = part[0];
}

__global__ void k_adam(int count, float *__restrict__ w, float *__restrict__ m,
                       float *__restrict__ v, const float *__restrict__ g,
                       uint16_t *__restrict__ wb, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
  m[i] = mi;
  v[i] = vi;
  const float mh = mi * c1, vh = vi * c2;
  const float wi = w[i] - (lr * mh) / (sqrtf(vh) + eps);
  w[i] = wi;
  if (wb) wb[i] = to_bf16(wi);
}

__global__ void k_to_bf16(long long count, const float *__restrict__ x, uint16_t *__restrict__ h) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) h[t] = to_bf16(x[t]);
}

// ---- inference ----------------------------------------------------------------
// Persistent blocks of WAVES waves; the weights (bfloat16, 96 KB) and biases
// live in LDS, shared by the waves.  A wave takes 16 TILES samples per
// iteration: TILES MFMA column tiles that share every weight fragment.

// The inference kernel uses v_mfma_f32_16x16x32_bf16.  A k-step covers 32
// features, and element e of lane (c, q) stands for feature
//   f(ks, q, e) = 32 ks + 16 (e / 4) + 4 q + e % 4,
// which is exactly where the previous layer's output tiles 2ks (e < 4) and
// 2ks + 1 (e >= 4) leave that feature in lane (c, q): activations go from
// one layer to the next without any cross-lane move.  The weights use the
// same k order (a sum over all k does not depend on it).
//
// Weights in LDS are stored in fragment order: the A fragment of output
// tile t and k-step ks for lane l is the 8 bfloat16 at
// ((t * KSTEPS + ks) * 64 + l) * 8, so a wave's fragment read is one
// contiguous 1 KB block (no bank conflicts).
__device__ __forceinline__ const bf16x8 &frag(const uint16_t *__restrict__ W, int ksteps, int t,
                                              int ks, int lane) {
  return *reinterpret_cast<const bf16x8 *>(W + ((t * ksteps + ks) * 64 + lane) * 8);
}

// The output layer's rows in the inference kernel: row 16 t + 4 q + e holds
// output 3 (4 t + q) + e (slot 4 t + q, axis e) for e < 3 and zeros for
// e = 3, so that lane (c, q) of output tile t ends with the three velocity
// offsets of one fine particle.  -1: a zero row.
__device__ __forceinline__ int out_row_source(int row) {
  const int e = row & 3, slot = 4 * (row >> 4) + ((row >> 2) & 3);
  return e < 3 ? 3 * slot + e : -1;
}

// Copies one layer's row-major weights (rows x 32*ksteps) into fragment order;
// OUT_ROWS: the rows of the output layer reordered by out_row_source.
template <bool OUT_ROWS = false>
__device__ __forceinline__ void stage_layer(uint16_t *__restrict__ dst,
                                            const uint16_t *__restrict__ src, int rows,
                                            int ksteps) {
  const int frags = (rows / 16) * ksteps * 64;  // 8-element fragments
  for (int f = threadIdx.x; f < frags; f += blockDim.x) {
    const int lane = f & 63, fk = f >> 6;
    const int t = fk / ksteps, ks = fk % ksteps;
    const int o = 16 * t + (lane & 15), k = 32 * ks + 4 * (lane >> 4);
    const int so = OUT_ROWS ? out_row_source(o) : o;
    uint2 lo = make_uint2(0u, 0u), hi = lo;
    if (so >= 0) {
      const uint16_t *row = src + (size_t)so * (32 * ksteps);
      lo = *reinterpret_cast<const uint2 *>(row + k);
      hi = *reinterpret_cast<const uint2 *>(row + k + 16);
    }
    reinterpret_cast<uint4 *>(dst)[f] = make_uint4(lo.x, lo.y, hi.x, hi.y);
  }
}

// One layer, transposed: output tile t (features 16t..16t+15) of every
// sample tile from KSTEPS 32-wide k-steps, accumulated onto the bias; ReLU
// and hardware bfloat16 rounding give the next layer's operand halves
// (out[s][t] = features 16t + 4q .. + 3 of lane (c, q)'s sample).
// ReLU (no NaN survives it), then hardware round-to-nearest-even.
__device__ __forceinline__ bf16x4 relu_bf16x4(const f32x4 &z) {
  const u32x2v h = {pack_bf16x2(z[0] > 0.0f ? z[0] : 0.0f, z[1] > 0.0f ? z[1] : 0.0f),
                    pack_bf16x2(z[2] > 0.0f ? z[2] : 0.0f, z[3] > 0.0f ? z[3] : 0.0f)};
  return __builtin_bit_cast(bf16x4, h);
}

// Software-pipelined over the 8 output tiles: each tile's weight fragments
// and bias are read from LDS one tile ahead, and its ReLU runs after the
// next tile's MFMAs are issued, so that neither the LDS latency nor the
// MFMA result latency stalls the wave.
template <int TILES, int KSTEPS>
__device__ __forceinline__ void hidden_layer(const uint16_t *__restrict__ W,
                                             const float *__restrict__ bias,
                                             const bf16x8 (&in)[TILES][4],
                                             bf16x4 (&out)[TILES][8], int lane) {
  const int q = lane >> 4;
  bf16x8 w[KSTEPS];
  f32x4 b = *reinterpret_cast<const f32x4 *>(bias + 4 * q);
#pragma unroll
  for (int ks = 0; ks < KSTEPS; ++ks) w[ks] = frag(W, KSTEPS, 0, ks, lane);
  f32x4 prev[TILES];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    bf16x8 wn[KSTEPS];
    f32x4 bn = b;
    if (t + 1 < 8) {
      bn = *reinterpret_cast<const f32x4 *>(bias + 16 * (t + 1) + 4 * q);
#pragma unroll
      for (int ks = 0; ks < KSTEPS; ++ks) wn[ks] = frag(W, KSTEPS, t + 1, ks, lane);
    }
    f32x4 a[TILES];
#pragma unroll
    for (int s = 0; s < TILES; ++s) a[s] = b;
#pragma unroll
    for (int ks = 0; ks < KSTEPS; ++ks)
#pragma unroll
      for (int s = 0; s < TILES; ++s) a[s] = mfma32(w[ks], in[s][ks], a[s]);
    if (t > 0) {
#pragma unroll
      for (int s = 0; s < TILES; ++s) out[s][t - 1] = relu_bf16x4(prev[s]);
    }
#pragma unroll
    for (int s = 0; s < TILES; ++s) prev[s] = a[s];
    if (t + 1 < 8) {
      b = bn;
#pragma unroll
      for (int ks = 0; ks < KSTEPS; ++ks) w[ks] = wn[ks];
    }
    // one tile in flight at a time (bounds live registers)
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int s = 0; s < TILES; ++s) out[s][7] = relu_bf16x4(prev[s]);
}

// The next layer's k-step operands from a hidden layer's output tiles.
template <int TILES>
__device__ __forceinline__ void operands(const bf16x4 (&h)[TILES][8], bf16x8 (&in)[TILES][4]) {
#pragma unroll
  for (int s = 0; s < TILES; ++s)
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) in[s][ks] = cat8(h[s][2 * ks], h[s][2 * ks + 1]);
}

// Output layer: 2 tiles (24 outputs + 8 zero rows, in out_row_source's
// order) onto the bias, no activation.
template <int TILES>
__device__ __forceinline__ void output_layer(const uint16_t *__restrict__ W,
                                             const float *__restrict__ bias,
                                             const bf16x8 (&in)[TILES][4],
                                             f32x4 (&y)[TILES][2], int lane) {
  const int q = lane >> 4;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const f32x4 b = *reinterpret_cast<const f32x4 *>(bias + 16 * t + 4 * q);
#pragma unroll
    for (int s = 0; s < TILES; ++s) y[s][t] = b;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const bf16x8 w = frag(W, 4, t, ks, lane);
#pragma unroll
      for (int s = 0; s < TILES; ++s) y[s][t] = mfma32(w, in[s][ks], y[s][t]);
    }
  }
}

// Layer-1 operands of sample group grp: k-step ks of lane (c, q) = feature
// groups 8 ks + q and 8 ks + 4 + q of sample grp * 16 TILES + 16 s + c.
template <int TILES>
__device__ __forceinline__ void load_features(int n, const uint2 *__restrict__ feat4, int grp,
                                              int groups, int c, int q, uint4 (&f)[TILES][3]) {
#pragma unroll
  for (int s = 0; s < TILES; ++s) {
    const int smp = grp * 16 * TILES + 16 * s + c;
#pragma unroll
    for (int ks = 0; ks < 3; ++ks) {
      uint2 lo = make_uint2(0u, 0u), hi = make_uint2(0u, 0u);
      if (grp < groups && smp < n) {
        lo = feat4[(size_t)(8 * ks + q) * n + smp];
        hi = feat4[(size_t)(8 * ks + 4 + q) * n + smp];
      }
      f[s][ks] = make_uint4(lo.x, lo.y, hi.x, hi.y);
    }
  }
}

// Flags, positions and velocities of sample group grp (index clamped to the
// last sample: the loads need no branch).
template <int TILES>
__device__ __forceinline__ void load_samples(int n, const int *__restrict__ flags,
                                             const float *__restrict__ pos,
                                             const float *__restrict__ vel, int grp, int c,
                                             int (&fl)[TILES], Packed3 (&ps)[TILES],
                                             Packed3 (&vs)[TILES]) {
#pragma unroll
  for (int s = 0; s < TILES; ++s) {
    const int smp = grp * 16 * TILES + 16 * s + c;
    const int ld = smp < n ? smp : n - 1;
    fl[s] = flags[ld];
    ps[s] = ld3(pos, ld);
    vs[s] = ld3(vel, ld);
  }
}

// feat4: grouped features, feat4[g * n + i] = features 4g..4g+3 of sample i.
// Each wave loads the next group's features, flags, positions and
// velocities before computing the current group, so that their latency
// hides behind the network (loads complete in issue order: the current
// group's inputs, loaded one iteration earlier, are waited for without
// waiting for the next group's).  With the output
// rows reordered, lane (c, q) of output tile t holds the velocity offsets of
// fine particle 8 i + 4 t + q (i its sample), stored as one position and one
// velocity triple.
template <int WAVES, int TILES>
__global__ void __launch_bounds__(64 * WAVES)
    k_mlp_infer(int n, const uint2 *__restrict__ feat4, const int *__restrict__ flags,
                const float *__restrict__ pos, const float *__restrict__ vel,
                const uint16_t *__restrict__ wts, const float *__restrict__ bias_g, float radius,
                float dt, float *__restrict__ fine_pos, float *__restrict__ fine_vel) {
  __shared__ __attribute__((aligned(16))) uint16_t W[49152];
  __shared__ __attribute__((aligned(16))) float bias[416];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = lane & 15, q = lane >> 4;
  const int per_iter = 16 * TILES;
  const int groups = (n + per_iter - 1) / per_iter;
  const int stride = gridDim.x * WAVES;
  int grp = blockIdx.x * WAVES + wave;
  uint4 cur[TILES][3];
  int fl[TILES];
  Packed3 ps[TILES], vs[TILES];
  load_features<TILES>(n, feat4, grp, groups, c, q, cur);
  load_samples<TILES>(n, flags, pos, vel, grp, c, fl, ps, vs);
  stage_layer(W + kWOff1, wts + kWOff1, kHid, kIn / 32);
  stage_layer(W + kWOff2, wts + kWOff2, kHid, kHid / 32);
  stage_layer(W + kWOff3, wts + kWOff3, kHid, kHid / 32);
  stage_layer<true>(W + kWOff4, wts + kWOff4, kOutPad, kHid / 32);
  for (int i = threadIdx.x; i < 416; i += blockDim.x) {
    const int so = i < kBOff4 ? i : out_row_source(i - kBOff4);
    bias[i] = so < 0 ? 0.0f : (i < kBOff4 ? bias_g[i] : bias_g[kBOff4 + so]);
  }
  // the template offsets radius * T of this lane's two slots q and 4 + q
  float rt[2][3];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) rt[t][ax] = radius * templ(4 * t + q, ax);
  __syncthreads();
  for (; grp < groups; grp += stride) {
    const int base = grp * per_iter;
    uint4 nxt[TILES][3];
    int fl_n[TILES];
    Packed3 ps_n[TILES], vs_n[TILES];
    load_features<TILES>(n, feat4, grp + stride, groups, c, q, nxt);
    load_samples<TILES>(n, flags, pos, vel, grp + stride, c, fl_n, ps_n, vs_n);
    __builtin_amdgcn_sched_barrier(0);
    bf16x8 in[TILES][4];
    bf16x4 h[TILES][8];
    f32x4 y[TILES][2];
#pragma unroll
    for (int s = 0; s < TILES; ++s) {
#pragma unroll
      for (int ks = 0; ks < 3; ++ks) in[s][ks] = __builtin_bit_cast(bf16x8, cur[s][ks]);
      in[s][3] = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
    }
    hidden_layer<TILES, 3>(W + kWOff1, bias + kBOff1, in, h, lane);
    operands<TILES>(h, in);
    hidden_layer<TILES, 4>(W + kWOff2, bias + kBOff2, in, h, lane);
    operands<TILES>(h, in);
    hidden_layer<TILES, 4>(W + kWOff3, bias + kBOff3, in, h, lane);
    operands<TILES>(h, in);
    output_layer<TILES>(W + kWOff4, bias + kBOff4, in, y, lane);
    // fine particle 8 smp + 4 t + q of sample smp = base + 16 s + c (computed
    // on every lane, stored for the samples that exist and are not flagged)
#pragma unroll
    for (int s = 0; s < TILES; ++s) {
      const int smp = base + 16 * s + c;
      const bool store = smp < n && fl[s] == 0;
      const Packed3 p = ps[s], v = vs[s];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const float fvx = v.x + y[s][t][0];
        const float fvy = v.y + y[s][t][1];
        const float fvz = v.z + y[s][t][2];
        const float fpx = (p.x + rt[t][0]) + dt * fvx;
        const float fpy = (p.y + rt[t][1]) + dt * fvy;
        const float fpz = (p.z + rt[t][2]) + dt * fvz;
        if (store) {
          const long long fi = 8ll * smp + 4 * t + q;
          st3(fine_pos, fi, fpx, fpy, fpz);
          if (fine_vel) st3(fine_vel, fi, fvx, fvy, fvz);
        }
      }
    }
#pragma unroll
    for (int s = 0; s < TILES; ++s) {
#pragma unroll
      for (int ks = 0; ks < 3; ++ks) cur[s][ks] = nxt[s][ks];
      fl[s] = fl_n[s];
      ps[s] = ps_n[s];
      vs[s] = vs_n[s];
    }
  }
}

// Layout probe: D = A B for one 16x16x16 tile (A, B row-major bf16).
__global__ void k_mfma_probe(const uint16_t *__restrict__ A, const uint16_t *__restrict__ B,
                             float *__restrict__ D) {
  const int lane = threadIdx.x;
  if (lane >= 64) return;
  const int c = lane & 15, q = lane >> 4;
  bf16x4 a, b;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    a[e] = (short)A[c * 16 + 4 * q + e];
    b[e] = (short)B[(4 * q + e) * 16 + c];
  }
  f32x4 d = mfma16(a, b, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
  for (int e = 0; e < 4; ++e) D[(4 * q + e) * 16 + c] = d[e];
}

static inline int nblocks(long long n, int bs) { return (int)((n + bs - 1) / bs); }

MlpConst mlp_const(const gg_mlp_config *c) {
  MlpConst m;
  m.radius = c->radius;
  m.vscale = c->vscale;
  m.guard_boundary = c->guard_boundary;
  m.guard_floor = c->guard_floor;
  m.guard_side = c->guard_side;
  return m;
}

int gemm(const uint16_t *A, long long sam, long long sak, const uint16_t *B, long long sbk,
         long long sbn, float *C, int ldc, int M, int N, int K, hipStream_t st) {
  if (M <= 0 || N <= 0) return GG_OK;
  GemmArgs a{A, B, C, M, N, K, sam, sak, sbk, sbn, ldc};
  const long long tiles = (long long)((M + 15) / 16) * ((N + 15) / 16);
  k_gemm_bf16<<<nblocks(tiles, 4), 256, 0, st>>>(a);
  GG_LAUNCH_CHECK();
  return GG_OK;
}

// Scratch of the training path for batch B.
struct TrainScratch {
  float *C, *Z[4], *g0, *g1, *dwork;
  uint16_t *H[4], *G;
};

size_t train_scratch_layout(int B, TrainScratch *t, char *base) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *p = base ? base + off : nullptr;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  };
  const size_t b = (size_t)(B > 0 ? B : 0);
  TrainScratch s{};
  s.C = reinterpret_cast<float *>(take(b * kHid * 4));
  for (int l = 0; l < 4; ++l) s.Z[l] = reinterpret_cast<float *>(take(b * kHid * 4));
  s.g0 = reinterpret_cast<float *>(take(b * kHid * 4));
  s.g1 = reinterpret_cast<float *>(take(b * kHid * 4));
  s.dwork = reinterpret_cast<float *>(take(b * kHid * 4));
  for (int l = 0; l < 4; ++l) s.H[l] = reinterpret_cast<uint16_t *>(take(b * kHid * 2));
  s.G = reinterpret_cast<uint16_t *>(take(b * kHid * 2));
  if (t) *t = s;
  return off;
}

constexpr int kDimIn[4] = {kIn, kHid, kHid, kHid};
constexpr int kDimOut[4] = {kHid, kHid, kHid, kOutPad};
constexpr int kWOffs[4] = {kWOff1, kWOff2, kWOff3, kWOff4};
constexpr int kBOffs[4] = {kBOff1, kBOff2, kBOff3, kBOff4};

// Training-path forward: Z_l, H_l kept in scratch; y = Z_4 (B x 24).
int train_forward(const uint16_t *wb, const float *bias, const uint16_t *x, int B,
                  TrainScratch &s, hipStream_t st) {
  const uint16_t *h = x;
  int K = kIn;
  for (int l = 0; l < 4; ++l) {
    const int N = kDimOut[l];
    // C[b][n] = sum_k h[b][k] W_l[n][k]
    GG_TRY(gemm(h, K, 1, wb + kWOffs[l], 1, K, s.C, kHid, B, N, K, st));
    const int Nz = (l == 3) ? kOut : N;
    k_bias_act<<<nblocks((long long)B * Nz, 256), 256, 0, st>>>(B, Nz, kHid, s.C,
                                                               bias + kBOffs[l], s.Z[l], s.H[l],
                                                               l == 3 ? 1 : 0);
    GG_LAUNCH_CHECK();
    h = s.H[l];
    K = N;
  }
  return GG_OK;
}

}  // namespace gg

using namespace gg;

extern "C" void gg_mlp_config_default(gg_mlp_config *c) {
  if (!c) return;
  c->radius = 0.02f;
  c->vscale = 1.0f;
  c->guard_boundary = 0.05f;
  c->guard_floor = 0.08f;
  c->guard_side = 0.05f;
  c->infer_waves = 0;
  c->infer_tiles = 0;
  c->select_chunk = 0;
}

extern "C" int gg_mlp_features(const gg_particles *coarse, const int *cs,
                               const gg_particles *boundary, const int *bcs, gg_grid grid,
                               const gg_mlp_config *c, uint16_t *feat, int *flags, int *knn,
                               hipStream_t stream) {
  if (!coarse || !cs || !boundary || !bcs || !c || !feat || !flags) return GG_ERR_INVALID;
  if (coarse->n < 0 || grid.cx <= 0 || grid.cy <= 0 || grid.cz <= 0) return GG_ERR_INVALID;
  if (!(c->radius > 0.0f) || !(c->vscale > 0.0f)) return GG_ERR_INVALID;
  if (coarse->n == 0) return GG_OK;
  if (!coarse->pos || !coarse->vel || (boundary->n > 0 && !boundary->pos)) return GG_ERR_INVALID;
  k_mlp_flags<<<nblocks(coarse->n, 256), 256, 0, stream>>>(
      coarse->n, reinterpret_cast<const float *>(coarse->pos),
      reinterpret_cast<const float *>(boundary->pos), bcs, dev_grid(grid), mlp_const(c), flags,
      nullptr, nullptr);
  GG_LAUNCH_CHECK();
  k_mlp_features<0, 16><<<nblocks(coarse->n, 256), 256, 0, stream>>>(
      coarse->n, reinterpret_cast<const float *>(coarse->pos),
      reinterpret_cast<const float *>(coarse->vel), cs, dev_grid(grid), mlp_const(c), feat, knn);
  GG_LAUNCH_CHECK();
  return GG_OK;
}

extern "C" int gg_mlp_teacher(const gg_particles *coarse, const int *cs,
                              const gg_particles *boundary, const int *bcs, gg_grid grid,
                              const gg_upsample_params *u, const gg_mlp_config *c,
                              const int *flags, float *y, float *fine_pos, float *fine_vel,
                              hipStream_t stream) {
  if (!coarse || !cs || !boundary || !bcs || !u || !c) return GG_ERR_INVALID;
  if (coarse->n < 0 || grid.cx <= 0 || grid.cy <= 0 || grid.cz <= 0) return GG_ERR_INVALID;
  if (coarse->n == 0) return GG_OK;
  if (!coarse->pos || !coarse->vel || (boundary->n > 0 && !boundary->pos)) return GG_ERR_INVALID;
  k_mlp_teacher<<<nblocks(8LL * coarse->n, 256), 256, 0, stream>>>(
      coarse->n, reinterpret_cast<const float *>(coarse->pos),
      reinterpret_cast<const float *>(coarse->vel), cs,
      reinterpret_cast<const float *>(boundary->pos), bcs, dev_grid(grid), up_const(u),
      c->radius, flags, y, fine_pos, fine_vel);
  GG_LAUNCH_CHECK();
  return GG_OK;
}

extern "C" int gg_mlp_fine_targets(const gg_particles *coarse, const gg_particles *fine,
                                   const int *fine_cell_start, gg_grid fine_grid,
                                   const gg_mlp_config *c, float *y, float *d2, int *match,
                                   hipStream_t stream) {
  if (!coarse || !fine || !fine_cell_start || !c || !y) return GG_ERR_INVALID;
  if (coarse->n < 0 || fine->n < 0) return GG_ERR_INVALID;
  if (fine_grid.cx <= 0 || fine_grid.cy <= 0 || fine_grid.cz <= 0 ||
      !(fine_grid.cell_length > 0.0f) || !(c->radius > 0.0f))
    return GG_ERR_INVALID;
  if (coarse->n == 0) return GG_OK;
  if (!coarse->pos || !coarse->vel || (fine->n > 0 && (!fine->pos || !fine->vel)))
    return GG_ERR_INVALID;
  k_mlp_fine_targets<<<nblocks(8LL * coarse->n, 256), 256, 0, stream>>>(
      coarse->n, reinterpret_cast<const float *>(coarse->pos),
      reinterpret_cast<const float *>(coarse->vel), reinterpret_cast<const float *>(fine->pos),
      reinterpret_cast<const float *>(fine->vel), fine_cell_start, dev_grid(fine_grid),
      c->radius, y, d2, match);
  GG_LAUNCH_CHECK();
  return GG_OK;
}

extern "C" size_t gg_mlp_train_scratch_bytes(int batch) {
  if (batch <= 0) return 0;
  return train_scratch_layout(batch, nullptr, nullptr);
}

extern "C" int gg_mlp_forward(const gg_mlp_weights *w, const uint16_t *x, int B, float *y,
                              void *scratch, size_t scratch_bytes, hipStream_t stream) {
  if (!w || !w->w_bf16 || !w->bias || !x || !y || B <= 0 || !scratch) return GG_ERR_INVALID;
  if (scratch_bytes < gg_mlp_train_scratch_bytes(B)) return GG_ERR_WORKSPACE;
  TrainScratch s;
  train_scratch_layout(B, &s, static_cast<char *>(scratch));
  GG_TRY(train_forward(w->w_bf16, w->bias, x, B, s, stream));
  return hip_status(hipMemcpyAsync(y, s.Z[3], sizeof(float) * (size_t)B * kOut,
                                   hipMemcpyDeviceToDevice, stream));
}

extern "C" int gg_mlp_grad(const gg_mlp_weights *w, const uint16_t *x, const float *target,
                           int B, float *dW, float *db, float *sq_err, void *scratch,
                           size_t scratch_bytes, hipStream_t stream) {
  if (!w || !w->w_bf16 || !w->bias || !x || !target || !dW || !db || B <= 0 || !scratch)
    return GG_ERR_INVALID;
  if (scratch_bytes < gg_mlp_train_scratch_bytes(B)) return GG_ERR_WORKSPACE;
  TrainScratch s;
  train_scratch_layout(B, &s, static_cast<char *>(scratch));
  GG_TRY(train_forward(w->w_bf16, w->bias, x, B, s, stream));
  // output layer: g (B x 32), G = bf16(g)
  k_loss_grad<<<nblocks(B, 256), 256, 0, stream>>>(B, s.Z[3], target, s.g0, s.G, sq_err);
  GG_LAUNCH_CHECK();
  float *g = s.g0, *gp = s.g1;
  for (int l = 3; l >= 0; --l) {
    const int K = kDimIn[l], N = kDimOut[l];
    const uint16_t *hin = (l == 0) ? x : s.H[l - 1];
    // dW_l[n][k] = sum_b G[b][n] h[b][k]
    GG_TRY(gemm(s.G, 1, N, hin, K, 1, dW + kWOffs[l], K, N, K, B, stream));
    k_colsum<<<N, 256, 0, stream>>>(B, N, g, db + kBOffs[l]);
    GG_LAUNCH_CHECK();
    if (l == 0) break;
    // D[b][k] = sum_n G[b][n] W_l[n][k]; g_prev = D * [Z_{l-1} > 0]
    GG_TRY(gemm(s.G, N, 1, w->w_bf16 + kWOffs[l], K, 1, s.dwork, K, B, K, N, stream));
    k_mask<<<nblocks((long long)B * K, 256), 256, 0, stream>>>((long long)B * K, s.dwork,
                                                               s.Z[l - 1], gp, s.G);
    GG_LAUNCH_CHECK();
    float *tmp = g;
    g = gp;
    gp = tmp;
  }
  return GG_OK;
}

extern "C" int gg_mlp_adam(gg_mlp_weights *w, const float *dW, const float *db, float lr,
                           float beta1, float beta2, float eps, int step, hipStream_t stream) {
  if (!w || !w->w_master || !w->m_w || !w->v_w || !w->b_master || !w->m_b || !w->v_b ||
      !w->w_bf16 || !w->bias || !dW || !db || step < 1)
    return GG_ERR_INVALID;
  const float c1 = (float)(1.0 / (1.0 - std::pow((double)beta1, step)));
  const float c2 = (float)(1.0 / (1.0 - std::pow((double)beta2, step)));
  k_adam<<<nblocks(GG_MLP_WEIGHTS, 256), 256, 0, stream>>>(GG_MLP_WEIGHTS, w->w_master, w->m_w,
                                                          w->v_w, dW, w->w_bf16, lr, beta1,
                                                          beta2, eps, c1, c2);
  GG_LAUNCH_CHECK();
  k_adam<<<nblocks(GG_MLP_BIASES, 256), 256, 0, stream>>>(GG_MLP_BIASES, w->b_master, w->m_b,
                                                         w->v_b, db, nullptr, lr, beta1, beta2,
                                                         eps, c1, c2);
  GG_LAUNCH_CHECK();
  return hip_status(hipMemcpyAsync(w->bias, w->b_master, sizeof(float) * GG_MLP_BIASES,
                                   hipMemcpyDeviceToDevice, stream));
}

extern "C" int gg_mlp_sync_weights(gg_mlp_weights *w, hipStream_t stream) {
  if (!w || !w->w_master || !w->w_bf16 || !w->b_master || !w->bias) return GG_ERR_INVALID;
  k_to_bf16<<<nblocks(GG_MLP_WEIGHTS, 256), 256, 0, stream>>>(GG_MLP_WEIGHTS, w->w_master,
                                                             w->w_bf16);
  GG_LAUNCH_CHECK();
  return hip_status(hipMemcpyAsync(w->bias, w->b_master, sizeof(float) * GG_MLP_BIASES,
                                   hipMemcpyDeviceToDevice, stream));
}

extern "C" size_t gg_upsample_mlp_scratch_bytes(int n) {
  if (n < 0) return 0;
  const size_t feat = ((size_t)n * kIn * 2 + 255) & ~(size_t)255;
  const size_t fl = ((size_t)n * 4 + 255) & ~(size_t)255;
  return feat + 2 * fl + 256;
}

template <int WAVES, int TILES>
static int launch_infer(int n, const uint2 *feat4, const int *flags, const float *pos,
                        const float *vel, const uint16_t *w, const float *bias, float radius,
                        float dt, float *fine_pos, float *fine_vel, int cus, hipStream_t st) {
  const int groups = (n + 16 * TILES - 1) / (16 * TILES);
  int blocks = (groups + WAVES - 1) / WAVES;
  if (blocks > cus) blocks = cus;
  k_mlp_infer<WAVES, TILES><<<blocks, 64 * WAVES, 0, st>>>(
      n, feat4, flags, pos, vel, w, bias, radius, dt, fine_pos, fine_vel);
  GG_LAUNCH_CHECK();
  return GG_OK;
}

// A second stream per device (default priority) on which the teacher
// fallback runs beside the network kernel, with the two events that fork it
// from and join it back into the caller's stream.  Created on first use,
// kept for the life of the process.
struct SideStream {
  hipStream_t st = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, flags = nullptr;
};
static int side_stream(SideStream *&out) {
  static std::mutex mu;
  static SideStream per_dev[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return GG_ERR_HIP;
  std::lock_guard<std::mutex> lock(mu);
  SideStream &s = per_dev[dev];
  if (!s.st) {
    if (hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking) != hipSuccess) return GG_ERR_HIP;
    if (hipEventCreateWithFlags(&s.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s.join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s.flags, hipEventDisableTiming) != hipSuccess)
      return GG_ERR_HIP;
  }
  out = &s;
  return GG_OK;
}

extern "C" int gg_upsample_mlp(const gg_particles *coarse, const int *cs,
                               const gg_particles *boundary, const int *bcs, gg_grid grid,
                               const gg_upsample_params *u, const gg_mlp_config *c,
                               const uint16_t *w_bf16, const float *bias, float *fine_pos,
                               float *fine_vel, void *scratch, size_t scratch_bytes,
                               int *fallback_count, hipStream_t stream) {
  if (!coarse || !cs || !boundary || !bcs || !u || !c || !w_bf16 || !bias || !fine_pos ||
      !scratch)
    return GG_ERR_INVALID;
  const int n = coarse->n;
  if (n < 0 || grid.cx <= 0 || grid.cy <= 0 || grid.cz <= 0) return GG_ERR_INVALID;
  if (!(c->radius > 0.0f) || !(c->vscale > 0.0f)) return GG_ERR_INVALID;
  // default 16 waves of one tile: at 250k and 1M coarse particles 4-6%
  // faster than 8 waves of 4 tiles (more waves hide the feature loads)
  const int waves = c->infer_waves == 0 ? 16 : c->infer_waves;
  const int tiles = c->infer_tiles == 0 ? 1 : c->infer_tiles;
  // (16 waves, 2 or 4 tiles) would spill registers: not offered
  if (!(((waves == 4 || waves == 8) && (tiles == 1 || tiles == 2 || tiles == 4)) ||
        (waves == 16 && tiles == 1)))
    return GG_ERR_INVALID;
  if (scratch_bytes < gg_upsample_mlp_scratch_bytes(n)) return GG_ERR_WORKSPACE;
  if (n == 0) return GG_OK;
  if (!coarse->pos || !coarse->vel || (boundary->n > 0 && !boundary->pos)) return GG_ERR_INVALID;
  uint16_t *feat = static_cast<uint16_t *>(scratch);
  const size_t fl_bytes = ((size_t)n * 4 + 255) & ~(size_t)255;
  int *flags = reinterpret_cast<int *>(static_cast<char *>(scratch) +
                                       (((size_t)n * kIn * 2 + 255) & ~(size_t)255));
  int *fb_list = reinterpret_cast<int *>(reinterpret_cast<char *>(flags) + fl_bytes);
  int *fb_count = reinterpret_cast<int *>(reinterpret_cast<char *>(fb_list) + fl_bytes);
  const float *pos = reinterpret_cast<const float *>(coarse->pos);
  const float *vel = reinterpret_cast<const float *>(coarse->vel);
  const DevGrid g = dev_grid(grid);
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const uint2 *feat4 = reinterpret_cast<const uint2 *>(feat);
  // The side stream computes the flags and the compacted list of flagged
  // particles, then the teacher on them (the plain upsampler on the
  // template), beside the features pass and the network on the caller's
  // stream (disjoint outputs); the network waits for the flags only.  Fused
  // into the features pass instead, the fallback would start only after it.
  SideStream *side = nullptr;
  GG_TRY(side_stream(side));
  if (hipEventRecord(side->fork, stream) != hipSuccess ||
      hipStreamWaitEvent(side->st, side->fork, 0) != hipSuccess)
    return GG_ERR_HIP;
  prof_begin(GG_STAGE_MLP_FALLBACK, side->st);
  k_zero_int<<<1, 1, 0, side->st>>>(fb_count);
  k_mlp_flags<<<nblocks(n, 256), 256, 0, side->st>>>(
      n, pos, reinterpret_cast<const float *>(boundary->pos), bcs, g, mlp_const(c), flags,
      fb_list, fb_count);
  if (fallback_count) k_copy_int<<<1, 1, 0, side->st>>>(fb_count, fallback_count);
  const bool flags_ok =
      hipGetLastError() == hipSuccess && hipEventRecord(side->flags, side->st) == hipSuccess;
  if (flags_ok)
    k_mlp_fallback<<<cus * 8, 256, 0, side->st>>>(
        fb_list, fb_count, pos, vel, cs, reinterpret_cast<const float *>(boundary->pos), bcs,
        g, up_const(u), c->radius, fine_pos, fine_vel);
  prof_end(GG_STAGE_MLP_FALLBACK, side->st);
  const bool side_ok = flags_ok && hipGetLastError() == hipSuccess;
  // the caller's stream always joins the side stream before returning
  if (hipEventRecord(side->join, side->st) != hipSuccess) return GG_ERR_HIP;
  if (!side_ok) {
    (void)hipStreamWaitEvent(stream, side->join, 0);
    return GG_ERR_HIP;
  }
  prof_begin(GG_STAGE_MLP_FEATURES, stream);
  int frc = GG_OK;
  if (c->select_chunk != 16)
    k_mlp_features<1, 8><<<nblocks(n, 256), 256, 0, stream>>>(n, pos, vel, cs, g, mlp_const(c),
                                                              feat, nullptr);
  else
    k_mlp_features<1, 16><<<nblocks(n, 256), 256, 0, stream>>>(n, pos, vel, cs, g,
                                                               mlp_const(c), feat, nullptr);
  if (hipGetLastError() != hipSuccess) frc = GG_ERR_HIP;
  if (frc == GG_OK && hipStreamWaitEvent(stream, side->flags, 0) != hipSuccess) frc = GG_ERR_HIP;
  if (frc != GG_OK) {
    (void)hipStreamWaitEvent(stream, side->join, 0);
    return frc;
  }
  prof_end(GG_STAGE_MLP_FEATURES, stream);
  prof_begin(GG_STAGE_MLP, stream);
  int rc;
#define GG_INFER(W_, T_)                                                                     \
  rc = launch_infer<W_, T_>(n, feat4, flags, pos, vel, w_bf16, bias, c->radius, u->dt,       \
                            fine_pos, fine_vel, cus, stream)
  if (waves == 4 && tiles == 1) GG_INFER(4, 1);
  else if (waves == 4 && tiles == 2) GG_INFER(4, 2);
  else if (waves == 4 && tiles == 4) GG_INFER(4, 4);
  else if (waves == 8 && tiles == 1) GG_INFER(8, 1);
  else if (waves == 8 && tiles == 2) GG_INFER(8, 2);
  else if (waves == 8 && tiles == 4) GG_INFER(8, 4);
  else GG_INFER(16, 1);
#undef GG_INFER
  if (rc != GG_OK) {
    (void)hipStreamWaitEvent(stream, side->join, 0);
    return rc;
  }
  prof_end(GG_STAGE_MLP, stream);
  if (hipStreamWaitEvent(stream, side->join, 0) != hipSuccess) return GG_ERR_HIP;
  return GG_OK;
}

// Counts the non-NaN float bit patterns first .. first + count - 1 (mod
// 2^32) on which pack_bf16x2 and to_bf16 disagree (either half).
__global__ void __launch_bounds__(256)
    k_bf16_round_check(uint64_t first, uint64_t count, unsigned long long *__restrict__ bad) {
  unsigned long long local = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
    const uint32_t u = (uint32_t)(first + k);
    if ((u & 0x7fffffffu) > 0x7f800000u) continue;
    const float x = __uint_as_float(u);
    const float y = __uint_as_float(u ^ 0x80000000u);
    const uint32_t hw = pack_bf16x2(x, y);
    if ((hw & 0xffffu) != to_bf16(x) || (hw >> 16) != to_bf16(y)) ++local;
  }
  for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o, 64);
  if ((threadIdx.x & 63) == 0 && local) atomicAdd(bad, local);
}

extern "C" int gg_bf16_round_check(uint64_t first, uint64_t count, unsigned long long *mismatches,
                                   hipStream_t stream) {
  if (!mismatches) return GG_ERR_INVALID;
  if (count == 0) return GG_OK;
  k_bf16_round_check<<<2048, 256, 0, stream>>>(first, count, mismatches);
  GG_LAUNCH_CHECK();
  return GG_OK;
}

extern "C" int gg_mfma_probe(const uint16_t *a, const uint16_t *b, float *d, hipStream_t stream) {
  if (!a || !b || !d) return GG_ERR_INVALID;
  k_mfma_probe<<<1, 64, 0, stream>>>(a, b, d);
  GG_LAUNCH_CHECK();
  return GG_OK;
}
