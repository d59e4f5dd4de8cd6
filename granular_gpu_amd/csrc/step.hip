// Plain PBD step: permute + predict, neighbour candidate lists, contact
// sweeps (stabilisation + solver), finalize.
//
// Reference: Solver::step (Solver.cu:82-109) -> add_external_force,
// update_particle_positions, apply_mass_scaling (Solver.cu:125-191),
// project (Solver.cu:416-513) with compute_delta_pos / boundary_constraint /
// particles_constraint (Solver.cu:272-401), final_update (Solver.cu:193-270).
//
// Deterministic semantics (the reference scatters into delta_pos[j] and
// n[j] without atomics): each sweep is a gather.  For particle i,
//   boundary k in contact (d < R_b + max(r_i, rmin)): A += ((rc-d)/d) p, n += 1
//   granular j in contact (d <= r_i + r_j)         : A += ((rc-d)/d) p, n += 2
// which is the reference's own term plus the mirror term thread j scatters
// into i (their sum is independent of the mass scaling).  Then
// x* += A / max(n, 1).  Positions are carried as base + offset: the offset
// D = x* - x stays small, so differences and the final x = base + D are
// rounded once.
//
// Neighbour candidate lists: built once per step with a skin; a sweep uses
// particle i's list only when (a) the stencil base of x*_i is the one the
// list was built for and (b) i's accumulated displacement plus the largest
// displacement of any particle is below 0.9*skin, which guarantees every
// pair outside the list is still out of contact.  Otherwise the particle
// scans the full 27-cell stencil, in its own lane or, when few lanes of its
// wave need to, queued and scanned by a whole wave in a second pass.  All paths
// visit contacts in the same order, so the result is identical to always
// scanning the stencil.
#include <hip/hip_runtime.h>

#include "gg_device.hpp"
#include "gg_internal.hpp"
#include "grid_kernels.hpp"
#include "profiler.hpp"

namespace gg {

__global__ void __launch_bounds__(kBlock)
    k_permute_predict(int n, const int *__restrict__ perm, const int *__restrict__ key,
                      const float *__restrict__ pos, const float *__restrict__ vel,
                      const float *__restrict__ mass, float4 *__restrict__ x0,
                      float4 *__restrict__ off0, float4 *__restrict__ d0,
                      float4 *__restrict__ xs,
                      float *__restrict__ mass_tmp, float *__restrict__ ms_out,
                      int *__restrict__ p2c, const int *__restrict__ rem_in,
                      const float *__restrict__ mer_in, const int *__restrict__ mc_in,
                      int *__restrict__ rem_tmp, float *__restrict__ mer_tmp,
                      int *__restrict__ mc_tmp, float dt, float dvx, float dvy, float dvz,
                      double denom, int *__restrict__ dirty, int ndirty,
                      int *__restrict__ fbcount, int *__restrict__ fbq_len,
                      unsigned int *__restrict__ bpart) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = i; k < ndirty; k += gridDim.x * blockDim.x) dirty[k] = 0;
  if (i < kFallbackSlots) fbcount[i] = 0;
  if (i < 2) fbq_len[i] = 0;
  float dmax = 0.0f, rmax = 0.0f;
  if (i < n) {
    const int o = perm ? perm[i] : i;
    const Packed3 p3 = ld3(pos, o), v3 = ld3(vel, o);
    const float px = p3.x, py = p3.y, pz = p3.z;
    float vx = v3.x, vy = v3.y, vz = v3.z;
    vx = vx + dvx;
    vy = vy + dvy;
    vz = vz + dvz;
    const float Dx = dt * vx, Dy = dt * vy, Dz = dt * vz;
    const float m = mass[o];
    const float r = radius_contact(m, denom);
    x0[i] = make_float4(px, py, pz, r);
    off0[i] = make_float4(Dx, Dy, Dz, 0.0f);
    d0[i] = make_float4(Dx, Dy, Dz, 0.0f);
    xs[i] = make_float4(px + Dx, py + Dy, pz + Dz, r);
    mass_tmp[i] = m;
    if (ms_out) ms_out[i] = m * exp_f(-py);
    if (perm && p2c) p2c[i] = key[o];
    if (rem_in) rem_tmp[i] = rem_in[o];
    if (mer_in) mer_tmp[i] = mer_in[o];
    if (mc_in) mc_tmp[i] = mc_in[o];
    dmax = fmaxf(fabsf(Dx), fmaxf(fabsf(Dy), fabsf(Dz)));
    if (!(Dx == Dx && Dy == Dy && Dz == Dz)) dmax = __builtin_inff();
    rmax = r == r ? r : __builtin_inff();
  }
  // per-block largest |D| component and radius; k_reduce_bounds folds the
  // blocks
  __shared__ float red[2][kBlock / 64];
  const float wd = wave_max(dmax);
  const float wr = wave_max(rmax);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = wd;
    red[1][threadIdx.x >> 6] = wr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float a = red[0][0], b = red[1][0];
    for (int w = 1; w < kBlock / 64; ++w) {
      a = fmaxf(a, red[0][w]);
      b = fmaxf(b, red[1][w]);
    }
    bpart[2 * blockIdx.x] = __float_as_uint(a);
    bpart[2 * blockIdx.x + 1] = __float_as_uint(b);
  }
}

// bounds[0..1] = maxima of the nb per-block values (non-negative floats
// order like their bit patterns).  One block.
__global__ void __launch_bounds__(1024) k_reduce_bounds(const unsigned int *__restrict__ bpart,
                                                        int nb, unsigned int *__restrict__ bounds) {
  __shared__ unsigned int red[2][16];
  unsigned int a = 0u, b = 0u;
  for (int k = threadIdx.x; k < nb; k += 1024) {
    a = max(a, bpart[2 * k]);
    b = max(b, bpart[2 * k + 1]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    a = max(a, (unsigned int)__shfl_xor((int)a, o, 64));
    b = max(b, (unsigned int)__shfl_xor((int)b, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < 16; ++w) {
      a = max(a, red[0][w]);
      b = max(b, red[1][w]);
    }
    bounds[0] = a;
    bounds[1] = b;
  }
}

// Gate of a mid-step list refresh: the refresh runs when the full-stencil
// count of the sweep just done (*ck) exceeds the first sweep's (*c0) by more
// than threshold.  ck null: always run (the build of the step's lists).
struct RefreshGate {
  const int *__restrict__ ck;
  const int *__restrict__ c0;
  int threshold;
};
__device__ __forceinline__ bool gate_open(const RefreshGate &gt) {
  return !gt.ck || *gt.ck - *gt.c0 > gt.threshold;
}

// Candidate lists: particles of the stencil of trunc(x*/L) with
// |x*_i - x*_j| < r_i + r_j + skin (boundary: |x*_i - q| < R_b + max(r_i,
// rmin) + skin), in scan order (cells ascending, index ascending).  x* is
// the predicted position fl(base + offset) with the contact radius in w.
// SCENES: the particle's scene gives the grid and the range of the cell
// tables (meta keeps the scene-local stencil id and sorted cell).
template <int BS, bool SCENES>
__global__ void __launch_bounds__(BS)
    k_nlist(int n, const float4 *__restrict__ x, const float4 *__restrict__ xs,
            const float4 *__restrict__ bnd4, const int *__restrict__ cs_all,
            const int *__restrict__ bcs_all, DevGrid g0, SweepConst sc, float skin,
            int4 *__restrict__ meta, int *__restrict__ glist, int *__restrict__ blist,
            int remap, SceneView sv, RefreshGate gate) {
  if (!gate_open(gate)) return;
  const int i = xcd_block(blockIdx.x, gridDim.x, remap) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const LocalGrid lg = local_grid<SCENES>(sv, i, g0, sc);
  const DevGrid &g = lg.g;
  const int *__restrict__ cs = cs_all + lg.base;
  const int *__restrict__ bcs = bcs_all + lg.base;
  const float4 si = xs[i];
  int bx = 0, by = 0, bz = 0;
  const bool ok = trunc_cell(si.x, si.y, si.z, g, bx, by, bz);
  const int ext = ext_base_id(ok, bx, by, bz, g);
  const float rbi = fmaxf(si.w, sc.rmin);
  const float blim = (sc.rb + rbi) + skin;
  int gc = 0, bc = 0;
  if (ext >= 0) {
    stencil_rows(bx, by, bz, g, bcs, [&](int k0, int k1) {
      for (int k = k0; k < k1; ++k) {
        const float4 q = bnd4[k];
        const float px = si.x - q.x, py = si.y - q.y, pz = si.z - q.z;
        const float d2 = px * px + py * py + pz * pz;
        if (d2 < blim * blim) {
          if (bc < sc.bmax) blist[(size_t)bc * sc.cap + i] = k;
          ++bc;
        }
      }
    });
    // 1 if j (at x* sj) is a candidate of i (self excluded below)
    auto near = [&](const float4 &sj) -> unsigned int {
      const float px = si.x - sj.x, py = si.y - sj.y, pz = si.z - sj.z;
      const float d2 = px * px + py * py + pz * pz;
      const float lim = (si.w + sj.w) + skin;
      return d2 < lim * lim ? 1u : 0u;
    };
    // Rows in chunks of 32 candidates: the tests only set bits of a mask;
    // the (few) candidates found are appended afterwards in index order.
    stencil_rows(bx, by, bz, g, cs, [&](int j0, int j1) {
      for (int c0 = j0; c0 < j1; c0 += 32) {
        const int cn = j1 - c0 < 32 ? j1 - c0 : 32;
        unsigned int mask = 0u;
        int t = 0;
        for (; t + 4 <= cn; t += 4) {
          const float4 s0 = xs[c0 + t], s1 = xs[c0 + t + 1], s2 = xs[c0 + t + 2],
                       s3 = xs[c0 + t + 3];
          mask |= (near(s0) | (near(s1) << 1) | (near(s2) << 2) | (near(s3) << 3)) << t;
        }
        for (; t < cn; ++t) mask |= near(xs[c0 + t]) << t;
        if (i >= c0 && i < c0 + cn) mask &= ~(1u << (i - c0));
        while (mask) {
          const int b = __ffs((int)mask) - 1;
          mask &= mask - 1u;
          if (gc < sc.gmax) glist[(size_t)gc * sc.cap + i] = c0 + b;
          ++gc;
        }
      }
    });
  }
  const float4 bi = x[i];
  meta[i] = make_int4(gc, bc, ext, hash_cell(bi.x, bi.y, bi.z, g));
}

struct SweepArgs {
  int n;
  const float4 *__restrict__ base;
  const float4 *__restrict__ off_in;
  float4 *__restrict__ off_out;
  const int4 *__restrict__ meta;
  const int *__restrict__ glist;
  const int *__restrict__ blist;
  const float4 *__restrict__ bnd4;
  const int *__restrict__ cs;
  const int *__restrict__ bcs;
  int *__restrict__ dirty;
  const float4 *__restrict__ d0;
  float4 *__restrict__ xb_out;
  int *__restrict__ stats;    // full-stencil particle count of this sweep
  int *__restrict__ fbq;      // particles queued for a whole-wave stencil scan
  int *__restrict__ fbq_count;  // queue length of this sweep
  int *__restrict__ fbq_next;   // queue length of the next sweep (cleared by the queue pass)
  int fbq_global;  // 1: queue in fbq for k_sweep_queue, 0: in the block's LDS queue
  DevGrid g;
  SweepConst sc;
  int use_lists;
  int remap;
  int fb_wave_max;  // a wave hands at most this many fallbacks to whole waves
  const unsigned int *__restrict__ bounds;  // largest |D| component, contact radius
  float near_limit;  // a list stays usable one cell over while its reach is below this
  SceneView sv;      // scenes sharing the set (read when SCENES)
};

// Shared epilogue: new offset (and for the last stabilisation sweep the
// position base).  A particle whose accumulated displacement has reached
// half the valid skin raises the displacement mark of the 27 cells around
// its sorted cell to that displacement (every sweep while it moves), so that
// particles scanning those cells can tell whether their lists still hold.
template <bool LAST_STAB>
__device__ __forceinline__ void sweep_store(const SweepArgs &a, int i, const float4 &bi,
                                            const float4 &oi, float3 ab, float3 ag, int cnt,
                                            int cell, const LocalGrid &lg) {
  const float nf = (float)(cnt > 1 ? cnt : 1);
  const float ix = (ab.x + ag.x) / nf;
  const float iy = (ab.y + ag.y) / nf;
  const float iz = (ab.z + ag.z) / nf;
  const float ox = oi.x + ix, oy = oi.y + iy, oz = oi.z + iz;
  const float cum_new = oi.w + sqrtf(ix * ix + iy * iy + iz * iz);
  if (a.use_lists && !(cum_new < a.sc.half_valid) && cum_new != oi.w) {
    const DevGrid &g = lg.g;
    if (cell >= 0 && cell < g.cx * g.cy * g.cz) {
      const int cz0 = cell % g.cz, cy0 = (cell / g.cz) % g.cy, cx0 = cell / (g.cz * g.cy);
      // non-negative floats (and NaN above them) order like their bits
      const unsigned int mark = __float_as_uint(cum_new);
      unsigned int *marks = reinterpret_cast<unsigned int *>(a.dirty) + lg.base;
      for (int m = 0; m < 27; ++m) {
        const int cc = stencil_cell(cx0, cy0, cz0, m, g);
        if (cc >= 0) atomicMax(marks + cc, mark);
      }
    }
  }
  if (LAST_STAB) {
    const float4 d = a.d0[i];
    a.xb_out[i] = make_float4(bi.x + (ox - d.x), bi.y + (oy - d.y), bi.z + (oz - d.z), bi.w);
    a.off_out[i] = make_float4(d.x, d.y, d.z, cum_new);
  } else {
    a.off_out[i] = make_float4(ox, oy, oz, cum_new);
  }
}

// Full 27-cell stencil evaluation for particle i (row order = scan order).
template <bool FRIC>
__device__ __forceinline__ void full_stencil(const SweepArgs &a, const SweepConst &sc,
                                             const LocalGrid &lg, int i, const float4 &bi,
                                             const float4 &oi, int bx, int by, int bz, float rcb,
                                             float preb, float3 &ab, float3 &ag, int &nb,
                                             int &ng) {
  stencil_rows(bx, by, bz, lg.g, a.bcs + lg.base, [&](int k0, int k1) {
    int k = k0;
    for (; k + 4 <= k1; k += 4) {
      const float4 q0 = a.bnd4[k], q1 = a.bnd4[k + 1], q2 = a.bnd4[k + 2], q3 = a.bnd4[k + 3];
      boundary_contact<FRIC>(bi, oi, q0, rcb, preb, sc, ab, nb);
      boundary_contact<FRIC>(bi, oi, q1, rcb, preb, sc, ab, nb);
      boundary_contact<FRIC>(bi, oi, q2, rcb, preb, sc, ab, nb);
      boundary_contact<FRIC>(bi, oi, q3, rcb, preb, sc, ab, nb);
    }
    for (; k < k1; ++k) boundary_contact<FRIC>(bi, oi, a.bnd4[k], rcb, preb, sc, ab, nb);
  });
  stencil_rows(bx, by, bz, lg.g, a.cs + lg.base, [&](int j0, int j1) {
    int j = j0;
    for (; j + 4 <= j1; j += 4) {
      const float4 b0 = a.base[j], b1 = a.base[j + 1], b2 = a.base[j + 2], b3 = a.base[j + 3];
      const float4 o0 = a.off_in[j], o1 = a.off_in[j + 1], o2 = a.off_in[j + 2],
                   o3 = a.off_in[j + 3];
      if (j != i) granular_contact<FRIC>(bi, oi, b0, o0, sc, ag, ng);
      if (j + 1 != i) granular_contact<FRIC>(bi, oi, b1, o1, sc, ag, ng);
      if (j + 2 != i) granular_contact<FRIC>(bi, oi, b2, o2, sc, ag, ng);
      if (j + 3 != i) granular_contact<FRIC>(bi, oi, b3, o3, sc, ag, ng);
    }
    for (; j < j1; ++j)
      if (j != i) granular_contact<FRIC>(bi, oi, a.base[j], a.off_in[j], sc, ag, ng);
  });
}

// Rows of the 27-cell stencil of (bx, by, bz) over cell table cs, held
// wave-uniform: row r (dx = r/3 - 1, dy = r%3 - 1, the three z cells) is the
// index range [start[r], start[r] + len[r]).  Rows in scan order.
struct StencilRows {
  int start[9], len[9], total;
};

// The row tables of one stencil over the boundary (bcs) and granular (cs)
// cell tables; lanes 0..8 read row r = lane of both, all loads in flight at
// once.
__device__ __forceinline__ void stencil_row_tables(int bx, int by, int bz, const DevGrid &g,
                                                   const int *__restrict__ bcs,
                                                   const int *__restrict__ cs, int lane,
                                                   StencilRows &rb, StencilRows &rg) {
  int sb = 0, eb = 0, sg = 0, eg = 0;
  const int z0 = bz - 1 > 0 ? bz - 1 : 0;
  const int z1 = bz + 1 < g.cz - 1 ? bz + 1 : g.cz - 1;
  if (lane < 9 && z0 <= z1) {
    const int x = bx + lane / 3 - 1, y = by + lane % 3 - 1;
    if (x >= 0 && x < g.cx && y >= 0 && y < g.cy) {
      const int c = (x * g.cy + y) * g.cz;
      sb = bcs[c + z0];
      eb = bcs[c + z1 + 1];
      sg = cs[c + z0];
      eg = cs[c + z1 + 1];
    }
  }
  const int lb = eb - sb, lg = eg - sg;
  rb.total = 0;
  rg.total = 0;
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    rb.start[r] = __builtin_amdgcn_readlane(sb, r);
    rb.len[r] = __builtin_amdgcn_readlane(lb, r);
    rb.total += rb.len[r];
    rg.start[r] = __builtin_amdgcn_readlane(sg, r);
    rg.len[r] = __builtin_amdgcn_readlane(lg, r);
    rg.total += rg.len[r];
  }
}

// Index of the t-th candidate of the row table in scan order (-1 past the end).
__device__ __forceinline__ int stencil_row_index(const StencilRows &rt, int t) {
  int j = -1, cum = 0;
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int d = t - cum;
    if (d >= 0 && d < rt.len[r]) j = rt.start[r] + d;
    cum += rt.len[r];
  }
  return j;
}

__device__ __forceinline__ float readlane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Adds the terms of the contacts flagged in mask, lowest lane first.
template <bool FRIC>
__device__ __forceinline__ void add_lanes(unsigned long long mask, const float3 &c, const float3 &f,
                                          float3 &acc) {
  while (mask) {
    const int l = __ffsll((long long)mask) - 1;
    mask &= mask - 1;
    const float3 cl = make_float3(readlane_f(c.x, l), readlane_f(c.y, l), readlane_f(c.z, l));
    float3 fl = make_float3(0.0f, 0.0f, 0.0f);
    if (FRIC) fl = make_float3(readlane_f(f.x, l), readlane_f(f.y, l), readlane_f(f.z, l));
    add_terms<FRIC>(acc, cl, fl);
  }
}

// full_stencil for one particle by a whole wave: the lanes evaluate the
// stencil's candidates 64 at a time in scan order, and the contacts found
// are added one after another in that order, so the sums equal
// full_stencil's bit for bit.  Every lane of the wave calls it with the
// same particle; the accumulators come back wave-uniform.
template <bool FRIC>
__device__ __forceinline__ void wave_stencil(const SweepArgs &a, const SweepConst &sc,
                                             const LocalGrid &lg, int i, const float4 &bi,
                                             const float4 &oi, int bx, int by, int bz, float rcb,
                                             float preb, float3 &ab, float3 &ag, int &nb,
                                             int &ng) {
  const int lane = threadIdx.x & 63;
  StencilRows rb, rg;
  stencil_row_tables(bx, by, bz, lg.g, a.bcs + lg.base, a.cs + lg.base, lane, rb, rg);
  // The first 2 x 64 boundary and 4 x 64 granular candidates are loaded
  // before any is used (one memory latency instead of one per 64).
  const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  int kq[2], jq[4];
  float4 q[2], gb[4], go[4];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    kq[u] = stencil_row_index(rb, 64 * u + lane);
    q[u] = kq[u] >= 0 ? a.bnd4[kq[u]] : z4;
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    jq[u] = stencil_row_index(rg, 64 * u + lane);
    gb[u] = jq[u] >= 0 ? a.base[jq[u]] : z4;
    go[u] = jq[u] >= 0 ? a.off_in[jq[u]] : z4;
  }
  for (int t0 = 0; t0 < rb.total; t0 += 64) {
    const int u = t0 >> 6;
    const int k = u < 2 ? kq[u & 1] : stencil_row_index(rb, t0 + lane);
    float3 c = make_float3(0.0f, 0.0f, 0.0f), f = c;
    bool hit = false;
    if (k >= 0) hit = boundary_terms<FRIC>(bi, oi, u < 2 ? q[u & 1] : a.bnd4[k], rcb, preb, sc, c, f);
    const unsigned long long m = __ballot(hit);
    nb += __popcll(m);
    add_lanes<FRIC>(m, c, f, ab);
  }
  for (int t0 = 0; t0 < rg.total; t0 += 64) {
    const int u = t0 >> 6;
    const int j = u < 4 ? jq[u & 3] : stencil_row_index(rg, t0 + lane);
    float3 c = make_float3(0.0f, 0.0f, 0.0f), f = c;
    bool hit = false;
    if (j >= 0 && j != i)
      hit = granular_terms<FRIC>(bi, oi, u < 4 ? gb[u & 3] : a.base[j],
                                 u < 4 ? go[u & 3] : a.off_in[j], sc, c, f);
    const unsigned long long m = __ballot(hit);
    ng += 2 * __popcll(m);
    add_lanes<FRIC>(m, c, f, ag);
  }
}

// wave_stencil without the hoisted loads (each group of 64 candidates is
// loaded when it is used): the variant k_sweep inlines, where registers set
// the occupancy of the list work.
template <bool FRIC>
__device__ __forceinline__ void wave_stencil_seq(const SweepArgs &a, const SweepConst &sc,
                                                 const LocalGrid &lg, int i, const float4 &bi,
                                                 const float4 &oi, int bx, int by, int bz,
                                                 float rcb, float preb, float3 &ab, float3 &ag,
                                                 int &nb, int &ng) {
  const int lane = threadIdx.x & 63;
  StencilRows rb, rg;
  stencil_row_tables(bx, by, bz, lg.g, a.bcs + lg.base, a.cs + lg.base, lane, rb, rg);
  for (int t0 = 0; t0 < rb.total; t0 += 64) {
    const int k = stencil_row_index(rb, t0 + lane);
    float3 c = make_float3(0.0f, 0.0f, 0.0f), f = c;
    bool hit = false;
    if (k >= 0) hit = boundary_terms<FRIC>(bi, oi, a.bnd4[k], rcb, preb, sc, c, f);
    const unsigned long long m = __ballot(hit);
    nb += __popcll(m);
    add_lanes<FRIC>(m, c, f, ab);
  }
  for (int t0 = 0; t0 < rg.total; t0 += 64) {
    const int j = stencil_row_index(rg, t0 + lane);
    float3 c = make_float3(0.0f, 0.0f, 0.0f), f = c;
    bool hit = false;
    if (j >= 0 && j != i) hit = granular_terms<FRIC>(bi, oi, a.base[j], a.off_in[j], sc, c, f);
    const unsigned long long m = __ballot(hit);
    ng += 2 * __popcll(m);
    add_lanes<FRIC>(m, c, f, ag);
  }
}


// One sweep.  Particles whose candidate list is provably exact for this
// sweep use it; the others scan the full stencil, either in their own lane
// or (when few lanes of the wave need it) queued for k_sweep_queue, which
// scans each with a whole wave.  All paths visit contacts in
// the same order, so the result does not depend on which path ran.
template <bool FRIC, bool LAST_STAB, int BS, bool SCENES>
__global__ void __launch_bounds__(BS, 6) k_sweep(SweepArgs a) {
  __shared__ int lqueue[BS];
  __shared__ int lqueued;
  const int lane = threadIdx.x & 63;
  const bool local_q = a.use_lists && !a.fbq_global;
  if (local_q) {
    if (threadIdx.x == 0) lqueued = 0;
    __syncthreads();
  }
  const int i = xcd_block(blockIdx.x, gridDim.x, a.remap) * blockDim.x + threadIdx.x;
  // the particle's grid, cell-table offset and friction (its scene's)
  LocalGrid lg{a.g, 0, a.sc.mu_s, a.sc.mu_k};
  if (SCENES && i < a.n) lg = local_grid<SCENES>(a.sv, i, a.g, a.sc);
  const DevGrid &g = lg.g;
  SweepConst sc = a.sc;
  sc.mu_s = lg.mu_s;
  sc.mu_k = lg.mu_k;
  // 0 idle lane, 1 list, 2 stencil scan, 3 stencil outside the grid
  int mode = 0;
  float4 bi = make_float4(0.0f, 0.0f, 0.0f, 0.0f), oi = bi;
  int bx = 0, by = 0, bz = 0;
  int4 mt = make_int4(0, 0, -1, -1);
  if (i < a.n) {
    bi = a.base[i];
    oi = a.off_in[i];
    const bool ok = trunc_cell(bi.x + oi.x, bi.y + oi.y, bi.z + oi.z, g, bx, by, bz);
    const int ext = ext_base_id(ok, bx, by, bz, g);
    bool list_ok = false;
    if (a.use_lists) {
      mt = a.meta[i];
      const int bcell = base_cell_id(ok, bx, by, bz, g);
      bool base_ok = ext == mt.z;
      if (!base_ok && ext >= 0) {
        // The stencil base moved one cell (x* moved less than half the
        // skin).  Scan order is ascending index for every base, and no pair
        // entering or leaving the stencil can be in contact while
        // r_i + r_max + |D|_max + 2 * half_valid < L, so the list still
        // holds exactly the contacts of the new stencil.
        const float dmax = __uint_as_float(a.bounds[0]), rmax = __uint_as_float(a.bounds[1]);
        const float reach = fmaxf((bi.w + rmax) + dmax, sc.rb + fmaxf(bi.w, sc.rmin));
        base_ok = reach + 2.0f * sc.half_valid < a.near_limit;
      }
      // A pair outside the list stays out of contact while
      // cum_i + cum_j < 2 * half_valid (0.9 skin); cum_j < half_valid
      // unless j marked the cell.
      bool moved_ok = false;
      if (bcell >= 0 && oi.w < sc.half_valid) {
        const unsigned int mark =
            reinterpret_cast<const unsigned int *>(a.dirty)[lg.base + bcell];
        moved_ok = mark == 0u || oi.w + __uint_as_float(mark) < 2.0f * sc.half_valid;
      }
      list_ok = base_ok && moved_ok && mt.x <= sc.gmax && mt.y <= sc.bmax;
    }
    mode = list_ok ? 1 : (ext >= 0 ? 2 : 3);
  }
  const unsigned long long fb = __ballot(mode >= 2);
  if (a.stats && fb && lane == __ffsll((long long)fb) - 1) {
    // count particles that scanned the full stencil (one atomic per wave)
    atomicAdd(a.stats, (int)__popcll(fb));
  }
  // Stencil scans of a few lanes are queued (one wave each): scanned in
  // their own lane they would hold this wave several times longer than its
  // list work.  They queue in the block (LDS, at most BS entries), scanned
  // by the block's waves at the end, or with GG_FLAG_FALLBACK_WAVE
  // sweep-wide for k_sweep_queue, which spreads them over the grid.  A
  // sweep-wide slot
  // past the capacity (a queue that was not drained) scans in its lane.
  const unsigned long long scan = __ballot(mode == 2);
  bool queued = false;
  if (a.use_lists && scan && __popcll(scan) <= a.fb_wave_max) {
    const int leader = __ffsll((long long)scan) - 1;
    const int rank = __popcll(scan & ((1ull << lane) - 1ull));
    int at0 = 0;
    if (a.fbq_global) {
      if (lane == leader) at0 = atomicAdd(a.fbq_count, (int)__popcll(scan));
      at0 = __shfl(at0, leader, 64);
      if (mode == 2 && at0 + rank < sc.cap) {
        a.fbq[at0 + rank] = i;
        queued = true;
      }
    } else {
      if (lane == leader) at0 = atomicAdd(&lqueued, (int)__popcll(scan));
      at0 = __shfl(at0, leader, 64);
      if (mode == 2) {
        lqueue[at0 + rank] = i;
        queued = true;
      }
    }
  }
  const float rcb = sc.rb + fmaxf(bi.w, sc.rmin);
  const float preb = contact_prefilter(rcb);
  float3 ab = make_float3(0.0f, 0.0f, 0.0f), ag = make_float3(0.0f, 0.0f, 0.0f);
  int nb = 0, ng = 0;
  if (mode == 1) {
    // boundary candidates in list order, gathered four at a time (a particle
    // beside a dense boundary has dozens)
    {
      const int *bl = a.blist + i;
      const size_t bstride = (size_t)sc.cap;
      int t = 0;
      for (; t + 4 <= mt.y; t += 4) {
        const int k0 = bl[(t + 0) * bstride], k1 = bl[(t + 1) * bstride];
        const int k2 = bl[(t + 2) * bstride], k3 = bl[(t + 3) * bstride];
        const float4 q0 = a.bnd4[k0], q1 = a.bnd4[k1], q2 = a.bnd4[k2], q3 = a.bnd4[k3];
        boundary_contact<FRIC>(bi, oi, q0, rcb, preb, sc, ab, nb);
        boundary_contact<FRIC>(bi, oi, q1, rcb, preb, sc, ab, nb);
        boundary_contact<FRIC>(bi, oi, q2, rcb, preb, sc, ab, nb);
        boundary_contact<FRIC>(bi, oi, q3, rcb, preb, sc, ab, nb);
      }
      for (; t < mt.y; ++t)
        boundary_contact<FRIC>(bi, oi, a.bnd4[bl[t * bstride]], rcb, preb, sc, ab, nb);
    }
    // candidates in list order, gathered four at a time
    const int *gl = a.glist + i;
    const size_t stride = (size_t)sc.cap;
    int t = 0;
    for (; t + 4 <= mt.x; t += 4) {
      const int j0 = gl[(t + 0) * stride], j1 = gl[(t + 1) * stride];
      const int j2 = gl[(t + 2) * stride], j3 = gl[(t + 3) * stride];
      const float4 b0 = a.base[j0], b1 = a.base[j1], b2 = a.base[j2], b3 = a.base[j3];
      const float4 o0 = a.off_in[j0], o1 = a.off_in[j1], o2 = a.off_in[j2], o3 = a.off_in[j3];
      granular_contact<FRIC>(bi, oi, b0, o0, sc, ag, ng);
      granular_contact<FRIC>(bi, oi, b1, o1, sc, ag, ng);
      granular_contact<FRIC>(bi, oi, b2, o2, sc, ag, ng);
      granular_contact<FRIC>(bi, oi, b3, o3, sc, ag, ng);
    }
    for (; t < mt.x; ++t) {
      const int j = gl[t * stride];
      granular_contact<FRIC>(bi, oi, a.base[j], a.off_in[j], sc, ag, ng);
    }
  } else if (mode == 2 && !queued) {
    full_stencil<FRIC>(a, sc, lg, i, bi, oi, bx, by, bz, rcb, preb, ab, ag, nb, ng);
  }
  if (mode == 1 || mode == 3 || (mode == 2 && !queued))
    sweep_store<LAST_STAB>(a, i, bi, oi, ab, ag, nb + ng, mt.w, lg);
  if (!local_q) return;
  // the block's queued particles: one wave each
  __syncthreads();
  const int nq = lqueued;
  for (int q = threadIdx.x >> 6; q < nq; q += BS / 64) {
    const int iq = lqueue[q];
    const LocalGrid lq = local_grid<SCENES>(a.sv, iq, a.g, a.sc);
    SweepConst scq = a.sc;
    scq.mu_s = lq.mu_s;
    scq.mu_k = lq.mu_k;
    const float4 bq = a.base[iq];
    const float4 oq = a.off_in[iq];
    int qx = 0, qy = 0, qz = 0;
    trunc_cell(bq.x + oq.x, bq.y + oq.y, bq.z + oq.z, lq.g, qx, qy, qz);
    const float rq = scq.rb + fmaxf(bq.w, scq.rmin);
    float3 qb = make_float3(0.0f, 0.0f, 0.0f), qg = qb;
    int nqb = 0, nqg = 0;
    wave_stencil_seq<FRIC>(a, scq, lq, iq, bq, oq, qx, qy, qz, rq, contact_prefilter(rq), qb,
                           qg, nqb, nqg);
    if (lane == 0)
      sweep_store<LAST_STAB>(a, iq, bq, oq, qb, qg, nqb + nqg, a.meta[iq].w, lq);
  }
}

// The particles k_sweep queued: one wave per particle, the waves of the
// whole grid taking them in turn, so that the few particles whose lists
// lapsed cost the sweep about one stencil scan of latency.  Same inputs and
// outputs as k_sweep (a Jacobi sweep reads off_in only).  Sweeps alternate
// between two queue lengths; this pass clears the next sweep's, which no
// kernel reads or writes meanwhile (predict clears both at the start of a
// step).  A queue that was not cleared (a sweep repeated without its
// neighbours) only makes particles scan again: the full-stencil result is
// every particle's exact result, so rewriting it changes nothing.
template <bool FRIC, bool LAST_STAB, bool SCENES>
__global__ void __launch_bounds__(256) k_sweep_queue(SweepArgs a) {
  const int lane = threadIdx.x & 63;
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.fbq_next = 0;
  const int len = __atomic_load_n(a.fbq_count, __ATOMIC_RELAXED);
  const int nq = len < a.sc.cap ? len : a.sc.cap;
  const int waves = gridDim.x * (blockDim.x >> 6);
  for (int q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); q < nq; q += waves) {
    const int iq = a.fbq[q];
    if (iq < 0 || iq >= a.n) continue;
    const LocalGrid lq = local_grid<SCENES>(a.sv, iq, a.g, a.sc);
    SweepConst sc = a.sc;
    sc.mu_s = lq.mu_s;
    sc.mu_k = lq.mu_k;
    const float4 bq = a.base[iq];
    const float4 oq = a.off_in[iq];
    int qx = 0, qy = 0, qz = 0;
    // (only a stale entry can fail this; k_sweep stored its result)
    if (!trunc_cell(bq.x + oq.x, bq.y + oq.y, bq.z + oq.z, lq.g, qx, qy, qz)) continue;
    const float rq = sc.rb + fmaxf(bq.w, sc.rmin);
    float3 qb = make_float3(0.0f, 0.0f, 0.0f), qg = qb;
    int nqb = 0, nqg = 0;
    wave_stencil<FRIC>(a, sc, lq, iq, bq, oq, qx, qy, qz, rq, contact_prefilter(rq), qb, qg, nqb,
                       nqg);
    if (lane == 0)
      sweep_store<LAST_STAB>(a, iq, bq, oq, qb, qg, nqb + nqg, a.meta[iq].w, lq);
  }
}

// Mid-step list refresh, first pass (when the gate is open): the current
// predicted positions x* = base + offset into xs, every accumulated
// displacement reset to zero, every displacement mark cleared, and the bound
// on |x* - x0| (x0: the sorted start-of-step position, whose cell placed the
// particle in the cell tables) raised to cover the new x*.  The list build
// then runs on xs, so that each list again holds exactly the candidates
// within the skin of x*, and the sweeps' validity tests restart from zero.
__global__ void __launch_bounds__(kBlock)
    k_refresh_prep(int n, const float4 *__restrict__ base, float4 *__restrict__ off,
                   const float4 *__restrict__ x0, float4 *__restrict__ xs,
                   int *__restrict__ dirty, int ndirty, unsigned int *__restrict__ bounds,
                   RefreshGate gate) {
  if (!gate_open(gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = i; k < ndirty; k += gridDim.x * blockDim.x) dirty[k] = 0;
  float dmax = 0.0f;
  if (i < n) {
    const float4 b = base[i];
    const float4 o = off[i];
    const float4 s0 = x0[i];
    const float px = b.x + o.x, py = b.y + o.y, pz = b.z + o.z;
    xs[i] = make_float4(px, py, pz, b.w);
    off[i] = make_float4(o.x, o.y, o.z, 0.0f);
    const float dx = px - s0.x, dy = py - s0.y, dz = pz - s0.z;
    dmax = fmaxf(fabsf(dx), fmaxf(fabsf(dy), fabsf(dz)));
    if (!(dmax == dmax)) dmax = __builtin_inff();
  }
  // the block's maximum, then one atomic only where it raises the bound
  // (non-negative floats order like their bits; most blocks stay below the
  // bound of the step's start, so the atomics do not pile up on one address)
  __shared__ float red[kBlock / 64];
  const float wd = wave_max(dmax);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wd;
  __syncthreads();
  if (threadIdx.x == 0) {
    float bd = red[0];
    for (int w = 1; w < kBlock / 64; ++w) bd = fmaxf(bd, red[w]);
    const unsigned int bits = __float_as_uint(bd);
    if (bits > __atomic_load_n(bounds, __ATOMIC_RELAXED)) atomicMax(bounds, bits);
  }
}

struct FinalConst {
  float inv_dt, sleep, damp_speed, damping, adaptive_damping;
};

__global__ void __launch_bounds__(kBlock)
    k_finalize(int n, const float4 *__restrict__ base, const float4 *__restrict__ off,
               float *__restrict__ pos, float *__restrict__ vel, int *__restrict__ adl,
               const float *__restrict__ mass_tmp, float *__restrict__ mass,
               const int *__restrict__ rem_tmp, int *__restrict__ rem,
               const float *__restrict__ mer_tmp, float *__restrict__ mer,
               const int *__restrict__ mc_tmp, int *__restrict__ mc, FinalConst fc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 b = base[i];
  const float4 o = off[i];
  float vx = fc.inv_dt * o.x, vy = fc.inv_dt * o.y, vz = fc.inv_dt * o.z;
  const float speed = sqrtf(vx * vx + vy * vy + vz * vz);
  if (speed < fc.sleep) {
    vx = 0.0f;
    vy = 0.0f;
    vz = 0.0f;
  } else if (speed > fc.damp_speed) {
    const float e = exp_f(-speed * 0.1f);
    const float f = fc.damping + (1.0f - fc.damping) * e;
    vx = vx * f;
    vy = vy * f;
    vz = vz * f;
    if (adl && adl[i] == 1) {
      vx = vx * fc.adaptive_damping;
      vy = vy * fc.adaptive_damping;
      vz = vz * fc.adaptive_damping;
    }
  }
  st3(vel, i, vx, vy, vz);
  st3(pos, i, b.x + o.x, b.y + o.y, b.z + o.z);
  if (adl) adl[i] = 0;
  if (mass) mass[i] = mass_tmp[i];
  if (rem) rem[i] = rem_tmp[i];
  if (mer) mer[i] = mer_tmp[i];
  if (mc) mc[i] = mc_tmp[i];
}

static inline int blocks_for(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// Threads per block of the list build and the sweeps: 0 (the default, 64) or
// one of these.
static inline bool block_ok(int b) { return b == 0 || b == 64 || b == 128 || b == 256 || b == 512; }

static int check_params(const gg_params *p, const gg_workspace *ws) {
  if (!p) return GG_ERR_INVALID;
  if (p->stab_iters < 0 || p->solver_iters < 0) return GG_ERR_INVALID;
  if (p->stab_iters + p->solver_iters > ws->max_sweeps) return GG_ERR_CAPACITY;
  if (!(p->dt > 0.0f)) return GG_ERR_INVALID;
  if (!block_ok(p->sweep_block) || !block_ok(p->nlist_block)) return GG_ERR_INVALID;
  return GG_OK;
}

static bool lists_on(const gg_params *p, const gg_workspace *ws) {
  return p->skin > 0.0f && !(p->flags & GG_FLAG_NO_NLIST) && ws->gmax > 0;
}

static SweepConst sweep_const(const gg_params *p, const gg_workspace *ws) {
  SweepConst sc;
  sc.rb = p->boundary_radius;
  sc.rmin = p->contact_radius_min;
  sc.mu_s = p->mu_s;
  sc.mu_k = p->mu_k;
  sc.half_valid = 0.5f * kListValidFrac * p->skin;
  sc.gmax = ws->gmax;
  sc.bmax = ws->bmax;
  sc.cap = ws->capacity + 1;
  return sc;
}

template <int BS, bool SCENES>
static int launch_nlist_bs(int n, const float4 *x, const float4 *xs, const float4 *bnd4,
                           const int *cs, const int *bcs, DevGrid g, SweepConst sc, float skin,
                           int4 *meta, int *glist, int *blist, int remap, SceneView sv,
                           RefreshGate gate, hipStream_t st) {
  k_nlist<BS, SCENES><<<(n + BS - 1) / BS, BS, 0, st>>>(n, x, xs, bnd4, cs, bcs, g, sc, skin,
                                                         meta, glist, blist, remap, sv, gate);
  GG_LAUNCH_CHECK();
  return GG_OK;
}

template <bool SCENES>
static int launch_nlist_sc(int bs, int n, const float4 *x, const float4 *xs, const float4 *bnd4,
                           const int *cs, const int *bcs, DevGrid g, SweepConst sc, float skin,
                           int4 *meta, int *glist, int *blist, int remap, SceneView sv,
                           RefreshGate gate, hipStream_t st) {
  switch (bs) {
    case 64: return launch_nlist_bs<64, SCENES>(n, x, xs, bnd4, cs, bcs, g, sc, skin, meta, glist, blist, remap, sv, gate, st);
    case 128: return launch_nlist_bs<128, SCENES>(n, x, xs, bnd4, cs, bcs, g, sc, skin, meta, glist, blist, remap, sv, gate, st);
    case 512: return launch_nlist_bs<512, SCENES>(n, x, xs, bnd4, cs, bcs, g, sc, skin, meta, glist, blist, remap, sv, gate, st);
    default: return launch_nlist_bs<256, SCENES>(n, x, xs, bnd4, cs, bcs, g, sc, skin, meta, glist, blist, remap, sv, gate, st);
  }
}

static int launch_nlist(int bs, int n, const float4 *x, const float4 *xs, const float4 *bnd4,
                        const int *cs, const int *bcs, DevGrid g, SweepConst sc, float skin,
                        int4 *meta, int *glist, int *blist, int remap, const SceneView *sv,
                        hipStream_t st, RefreshGate gate = RefreshGate{nullptr, nullptr, 0}) {
  if (sv)
    return launch_nlist_sc<true>(bs, n, x, xs, bnd4, cs, bcs, g, sc, skin, meta, glist, blist,
                                 remap, *sv, gate, st);
  return launch_nlist_sc<false>(bs, n, x, xs, bnd4, cs, bcs, g, sc, skin, meta, glist, blist,
                                remap, SceneView{nullptr, nullptr, 0}, gate, st);
}

// permute (perm may be null = identity) + predict + lists.  sv: scenes
// sharing the set (null: one scene).
static int run_predict(gg_particles *set, gg_adaptive *ad, const int *perm, const int *key,
                       const int *cs, const int *bcs, int nbnd, const gg_params *p,
                       const gg_workspace *ws, hipStream_t st, const SceneView *sv = nullptr) {
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                         ws->bmax);
  const int n = set->n;
  const int ndirty = (int)(grid_cells(ws->grid) + 1);
  const float dvx = p->dt * p->gx, dvy = p->dt * p->gy, dvz = p->dt * p->gz;
  const double denom = 4.0 * kPiD * (double)(int)p->density;
  const int nblk = n > 0 ? blocks_for(n) : 1;
  prof_begin(GG_STAGE_PREDICT, st);
  k_permute_predict<<<nblk, kBlock, 0, st>>>(
      n, perm, key, reinterpret_cast<const float *>(set->pos),
      reinterpret_cast<const float *>(set->vel), set->mass, at<float4>(ws, L.x0),
      at<float4>(ws, L.off0), at<float4>(ws, L.d0), at<float4>(ws, L.off1), at<float>(ws, L.ftmp0),
      set->scaled_mass,
      set->particle_2_cell, ad ? ad->remove : nullptr, ad ? ad->merge : nullptr,
      ad ? ad->merge_count : nullptr, at<int>(ws, L.itmp0), at<float>(ws, L.ftmp1),
      at<int>(ws, L.itmp1), p->dt, dvx, dvy, dvz, denom, at<int>(ws, L.dirty), ndirty,
      at<int>(ws, L.ctr) + 1, at<int>(ws, L.fbqc), at<unsigned int>(ws, L.bpart));
  GG_LAUNCH_CHECK();
  k_reduce_bounds<<<1, 1024, 0, st>>>(at<unsigned int>(ws, L.bpart), nblk,
                                      at<unsigned int>(ws, L.bounds));
  GG_LAUNCH_CHECK();
  prof_end(GG_STAGE_PREDICT, st);
  if (lists_on(p, ws) && n > 0) {
    prof_begin(GG_STAGE_NLIST, st);
    GG_TRY(launch_nlist(p->nlist_block == 0 ? 64 : p->nlist_block, n, at<float4>(ws, L.x0), at<float4>(ws, L.off1),
                        at<float4>(ws, L.bnd4), cs, bcs, dev_grid(ws->grid), sweep_const(p, ws),
                        p->skin, at<int4>(ws, L.meta), at<int>(ws, L.glist),
                        at<int>(ws, L.blist), (p->flags & GG_FLAG_NO_XCD_REMAP) ? 0 : 1, sv,
                        st));
    prof_end(GG_STAGE_NLIST, st);
  }
  (void)nbnd;
  return GG_OK;
}

template <bool FRIC, bool LAST, bool SCENES>
static int launch_sweep_sc(const SweepArgs &a, int bs, hipStream_t st) {
  const int n = a.n;
  switch (bs) {
    case 64: k_sweep<FRIC, LAST, 64, SCENES><<<(n + 63) / 64, 64, 0, st>>>(a); break;
    case 128: k_sweep<FRIC, LAST, 128, SCENES><<<(n + 127) / 128, 128, 0, st>>>(a); break;
    case 512: k_sweep<FRIC, LAST, 512, SCENES><<<(n + 511) / 512, 512, 0, st>>>(a); break;
    default: k_sweep<FRIC, LAST, 256, SCENES><<<(n + 255) / 256, 256, 0, st>>>(a); break;
  }
  GG_LAUNCH_CHECK();
  return GG_OK;
}

template <bool FRIC, bool LAST>
static int launch_sweep(const SweepArgs &a, int bs, hipStream_t st) {
  return a.sv.scene_of ? launch_sweep_sc<FRIC, LAST, true>(a, bs, st)
                       : launch_sweep_sc<FRIC, LAST, false>(a, bs, st);
}

template <bool FRIC, bool LAST>
static int launch_sweep_queue(const SweepArgs &a, hipStream_t st) {
  // up to 2048 waves take the queued particles in turn
  const int qblocks = a.n / 256 + 1 < 512 ? a.n / 256 + 1 : 512;
  if (a.sv.scene_of)
    k_sweep_queue<FRIC, LAST, true><<<qblocks, 256, 0, st>>>(a);
  else
    k_sweep_queue<FRIC, LAST, false><<<qblocks, 256, 0, st>>>(a);
  GG_LAUNCH_CHECK();
  return GG_OK;
}

// fric_scenes: with scenes, whether their sweeps apply friction (each with
// its own coefficients).
static int run_sweep(const gg_particles *set, const int *cs, const int *bcs, const gg_params *p,
                     int k, const gg_workspace *ws, hipStream_t st,
                     const SceneView *sv = nullptr, bool fric_scenes = false) {
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                         ws->bmax);
  const int n = set->n;
  if (n <= 0) return GG_OK;
  const int S = p->stab_iters;
  const bool last_stab = (k == S - 1);
  const bool fric =
      (k >= S) && (sv ? fric_scenes : (p->mu_s > 0.0f || p->mu_k > 0.0f));
  SweepArgs a;
  a.n = n;
  a.base = (k < S || S == 0) ? at<float4>(ws, L.x0) : at<float4>(ws, L.xb);
  a.off_in = (k % 2 == 0) ? at<float4>(ws, L.off0) : at<float4>(ws, L.off1);
  a.off_out = (k % 2 == 0) ? at<float4>(ws, L.off1) : at<float4>(ws, L.off0);
  a.meta = at<int4>(ws, L.meta);
  a.glist = at<int>(ws, L.glist);
  a.blist = at<int>(ws, L.blist);
  a.bnd4 = at<float4>(ws, L.bnd4);
  a.cs = cs;
  a.bcs = bcs;
  a.dirty = at<int>(ws, L.dirty);
  a.d0 = at<float4>(ws, L.d0);
  a.xb_out = at<float4>(ws, L.xb);
  a.stats = at<int>(ws, L.ctr) + 1 + (k % kFallbackSlots);
  a.fbq = at<int>(ws, L.fbq);
  a.fbq_count = at<int>(ws, L.fbqc) + (k & 1);
  a.fbq_next = at<int>(ws, L.fbqc) + ((k + 1) & 1);
  a.g = dev_grid(ws->grid);
  a.sc = sweep_const(p, ws);
  a.use_lists = lists_on(p, ws) ? 1 : 0;
  a.remap = (p->flags & GG_FLAG_NO_XCD_REMAP) ? 0 : 1;
  a.bounds = at<unsigned int>(ws, L.bounds);
  a.near_limit = 0.999f * ws->grid.cell_length;
  a.sv = sv ? *sv : SceneView{nullptr, nullptr, 0};
  // GG_FLAG_FALLBACK_WAVE queues every stencil scan sweep-wide, whatever the
  // size of the set
  a.fbq_global = (p->flags & GG_FLAG_FALLBACK_WAVE) ? 1 : 0;
  a.fb_wave_max = (p->flags & GG_FLAG_FALLBACK_LANE)   ? 0
                  : (p->flags & GG_FLAG_FALLBACK_WAVE) ? 64
                  : a.fbq_global                       ? kFallbackWaveMax
                                                       : kFallbackWaveMaxLocal;
  const int bs = p->sweep_block == 0 ? 64 : p->sweep_block;
  const int stage = k < S ? GG_STAGE_STAB : GG_STAGE_SWEEP;
  prof_begin(stage, st);
  int rc;
  if (last_stab)
    rc = launch_sweep<false, true>(a, bs, st);
  else if (fric)
    rc = launch_sweep<true, false>(a, bs, st);
  else
    rc = launch_sweep<false, false>(a, bs, st);
  if (rc != GG_OK) return rc;
  prof_end(stage, st);
  // nothing is queued sweep-wide without lists, per block, or in lanes only
  if (!a.use_lists || !a.fbq_global || a.fb_wave_max == 0) return GG_OK;
  prof_begin(GG_STAGE_SWEEP_QUEUE, st);
  if (last_stab)
    rc = launch_sweep_queue<false, true>(a, st);
  else if (fric)
    rc = launch_sweep_queue<true, false>(a, st);
  else
    rc = launch_sweep_queue<false, false>(a, st);
  if (rc != GG_OK) return rc;
  prof_end(GG_STAGE_SWEEP_QUEUE, st);
  return GG_OK;
}

// Whether a list refresh may follow sweep k: after every sweep but the first
// and the last (the gate on the device decides whether it runs).
static bool refresh_point(const gg_params *p, const gg_workspace *ws, int k) {
  if (!(p->flags & GG_FLAG_LIST_REFRESH) || !lists_on(p, ws)) return false;
  const int K = p->stab_iters + p->solver_iters;
  return k >= 1 && k <= K - 2;
}

// The refresh after sweep k, gated on the device: it runs when sweep k
// scanned the stencil for more particles than sweep 0 (whose lists were
// fresh) by more than n / (kRefreshCost * R), R = the sweeps left in the
// step, i.e. when the scans it can save over the rest of the step outweigh
// one list build.  Positions, marks and bounds by k_refresh_prep, then the
// list build on the refreshed positions.  The positions go to a buffer that
// no later sweep of the step reads before rewriting it: the stabilisation
// base (written by the last stabilisation sweep) while stabilisation
// sweeps remain, dt*v (read only by the last stabilisation sweep) after.
static int run_refresh(const gg_particles *set, const int *cs, const int *bcs,
                       const gg_params *p, int k, const gg_workspace *ws, hipStream_t st,
                       const SceneView *sv = nullptr) {
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                         ws->bmax);
  const int n = set->n;
  if (n <= 0) return GG_OK;
  const int S = p->stab_iters, K = p->stab_iters + p->solver_iters;
  const int left = K - 1 - k;
  const int *ctr = at<int>(ws, L.ctr) + 1;
  const RefreshGate gate{ctr + (k % kFallbackSlots), ctr,
                         (int)((long long)n / ((long long)kRefreshCost * left))};
  // the offsets the next sweep reads and its position base
  float4 *off = ((k + 1) % 2 == 0) ? at<float4>(ws, L.off0) : at<float4>(ws, L.off1);
  const bool stab_next = k + 1 < S;
  const float4 *base = (S == 0 || stab_next) ? at<float4>(ws, L.x0) : at<float4>(ws, L.xb);
  float4 *xs = stab_next ? at<float4>(ws, L.xb) : at<float4>(ws, L.d0);
  const int ndirty = (int)(grid_cells(ws->grid) + 1);
  prof_begin(GG_STAGE_NLIST, st);
  k_refresh_prep<<<blocks_for(n), kBlock, 0, st>>>(n, base, off, at<float4>(ws, L.x0), xs,
                                                    at<int>(ws, L.dirty), ndirty,
                                                    at<unsigned int>(ws, L.bounds), gate);
  GG_LAUNCH_CHECK();
  GG_TRY(launch_nlist(p->nlist_block == 0 ? 64 : p->nlist_block, n, at<float4>(ws, L.x0), xs,
                      at<float4>(ws, L.bnd4), cs, bcs, dev_grid(ws->grid), sweep_const(p, ws),
                      p->skin, at<int4>(ws, L.meta), at<int>(ws, L.glist), at<int>(ws, L.blist),
                      (p->flags & GG_FLAG_NO_XCD_REMAP) ? 0 : 1, sv, st, gate));
  prof_end(GG_STAGE_NLIST, st);
  return GG_OK;
}

static int run_finalize(gg_particles *set, gg_adaptive *ad, const gg_params *p,
                        const gg_workspace *ws, hipStream_t st) {
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                         ws->bmax);
  const int n = set->n;
  if (n <= 0) return GG_OK;
  const int S = p->stab_iters, I = p->solver_iters;
  const float4 *base = S > 0 ? at<float4>(ws, L.xb) : at<float4>(ws, L.x0);
  const float4 *off = ((S + I) % 2 == 0) ? at<float4>(ws, L.off0) : at<float4>(ws, L.off1);
  FinalConst fc;
  fc.inv_dt = 1.0f / p->dt;
  fc.sleep = p->sleep_speed;
  fc.damp_speed = p->damp_speed;
  fc.damping = p->damping;
  fc.adaptive_damping = p->adaptive_damping;
  prof_begin(GG_STAGE_FINALIZE, st);
  k_finalize<<<blocks_for(n), kBlock, 0, st>>>(
      n, base, off, reinterpret_cast<float *>(set->pos), reinterpret_cast<float *>(set->vel),
      set->adaptive_last_step, at<float>(ws, L.ftmp0), set->mass, at<int>(ws, L.itmp0),
      ad ? ad->remove : nullptr, at<float>(ws, L.ftmp1), ad ? ad->merge : nullptr,
      at<int>(ws, L.itmp1), ad ? ad->merge_count : nullptr, fc);
  GG_LAUNCH_CHECK();
  prof_end(GG_STAGE_FINALIZE, st);
  return GG_OK;
}

static int check_set(const gg_particles *set, const gg_workspace *ws) {
  if (!set) return GG_ERR_INVALID;
  if (set->n < 0) return GG_ERR_INVALID;
  if (set->n > ws->capacity) return GG_ERR_CAPACITY;
  if (set->n > 0 && (!set->pos || !set->vel || !set->mass)) return GG_ERR_INVALID;
  return GG_OK;
}

static int check_boundary(const gg_particles *b, const int *bcs, const gg_workspace *ws) {
  if (!b || !bcs) return GG_ERR_INVALID;
  if (b->n < 0) return GG_ERR_INVALID;
  if (b->n > ws->boundary_capacity) return GG_ERR_CAPACITY;
  return GG_OK;
}

}  // namespace gg

using namespace gg;

extern "C" int gg_step(gg_particles *set, gg_adaptive *ad, int *cell_start,
                       const gg_particles *boundary, const int *bcell_start,
                       const gg_params *params, const gg_workspace *ws, hipStream_t stream) {
  GG_TRY(check_ws(ws));
  GG_TRY(check_params(params, ws));
  GG_TRY(check_set(set, ws));
  GG_TRY(check_boundary(boundary, bcell_start, ws));
  if (!cell_start) return GG_ERR_INVALID;
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                         ws->bmax);
  prof_begin(GG_STAGE_SORT, stream);
  GG_TRY(sort_by_cell(reinterpret_cast<const float *>(set->pos), set->n, cell_start, ws, stream));
  prof_end(GG_STAGE_SORT, stream);
  GG_TRY(run_predict(set, ad, at<int>(ws, L.perm), at<int>(ws, L.key), cell_start, bcell_start,
                     boundary->n, params, ws, stream));
  const int K = params->stab_iters + params->solver_iters;
  for (int k = 0; k < K; ++k) {
    GG_TRY(run_sweep(set, cell_start, bcell_start, params, k, ws, stream));
    if (refresh_point(params, ws, k))
      GG_TRY(run_refresh(set, cell_start, bcell_start, params, k, ws, stream));
  }
  return run_finalize(set, ad, params, ws, stream);
}

extern "C" int gg_step_scenes(gg_particles *set, int *cell_start, const gg_particles *boundary,
                              const int *bcell_start, const gg_scenes *scenes,
                              const gg_params *params, const gg_workspace *ws,
                              hipStream_t stream) {
  GG_TRY(check_ws(ws));
  GG_TRY(check_params(params, ws));
  GG_TRY(check_set(set, ws));
  GG_TRY(check_boundary(boundary, bcell_start, ws));
  if (!cell_start) return GG_ERR_INVALID;
  GG_TRY(check_scenes(scenes, ws));
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                         ws->bmax);
  const SceneView sv{scenes->scene_of, scenes->desc, scenes->count};
  // cell keys over the concatenated tables, then the ordinary sort
  int *keys = at<int>(ws, L.itmp2);
  prof_begin(GG_STAGE_SORT, stream);
  GG_TRY(scene_keys(reinterpret_cast<const float *>(set->pos), set->n, scenes->scene_of, scenes,
                    ws->grid.cell_length, keys, stream));
  GG_TRY(sort_by_cell_keys(reinterpret_cast<const float *>(set->pos), keys, set->n, cell_start,
                           ws, stream));
  prof_end(GG_STAGE_SORT, stream);
  GG_TRY(run_predict(set, nullptr, at<int>(ws, L.perm), at<int>(ws, L.key), cell_start,
                     bcell_start, boundary->n, params, ws, stream, &sv));
  const int K = params->stab_iters + params->solver_iters;
  for (int k = 0; k < K; ++k) {
    GG_TRY(run_sweep(set, cell_start, bcell_start, params, k, ws, stream, &sv,
                     scenes->friction != 0));
    if (refresh_point(params, ws, k))
      GG_TRY(run_refresh(set, cell_start, bcell_start, params, k, ws, stream, &sv));
  }
  return run_finalize(set, nullptr, params, ws, stream);
}

extern "C" int gg_predict(gg_particles *set, const int *cell_start, const gg_particles *boundary,
                          const int *bcell_start, const gg_params *params,
                          const gg_workspace *ws, hipStream_t stream) {
  GG_TRY(check_ws(ws));
  GG_TRY(check_params(params, ws));
  GG_TRY(check_set(set, ws));
  GG_TRY(check_boundary(boundary, bcell_start, ws));
  if (!cell_start) return GG_ERR_INVALID;
  return run_predict(set, nullptr, nullptr, nullptr, cell_start, bcell_start, boundary->n, params,
                     ws, stream);
}

extern "C" int gg_contact_sweep(gg_particles *set, const int *cell_start,
                                const gg_particles *boundary, const int *bcell_start,
                                const gg_params *params, int sweep, const gg_workspace *ws,
                                hipStream_t stream) {
  GG_TRY(check_ws(ws));
  GG_TRY(check_params(params, ws));
  GG_TRY(check_set(set, ws));
  GG_TRY(check_boundary(boundary, bcell_start, ws));
  if (!cell_start) return GG_ERR_INVALID;
  if (sweep < 0 || sweep >= params->stab_iters + params->solver_iters) return GG_ERR_INVALID;
  return run_sweep(set, cell_start, bcell_start, params, sweep, ws, stream);
}

extern "C" int gg_finalize(gg_particles *set, const gg_params *params, const gg_workspace *ws,
                           hipStream_t stream) {
  GG_TRY(check_ws(ws));
  GG_TRY(check_params(params, ws));
  GG_TRY(check_set(set, ws));
  return run_finalize(set, nullptr, params, ws, stream);
}

extern "C" const void *gg_workspace_view(const gg_workspace *ws, int which) {
  if (check_ws(ws) != GG_OK) return nullptr;
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                         ws->bmax);
  switch (which) {
    case 0: return at<void>(ws, L.x0);
    case 1: return at<void>(ws, L.off0);
    case 2: return at<void>(ws, L.meta);
    case 3: return at<void>(ws, L.glist);
    case 4: return at<void>(ws, L.blist);
    case 5: return at<void>(ws, L.off1);
    case 6: return at<void>(ws, L.xb);
    case 7: return at<void>(ws, L.d0);
    case 8: return at<void>(ws, L.ctr);
    default: return nullptr;
  }
}
