// Adaptive merge/split: surface flags, distance to the surface, merge
// pairing and blending, split and the compaction that removes merged donors
// and appends split children.
//
// Reference: GranularSystem.cu:506-699 (find_cover_vector, find_surface,
// compute_distance_to_surface) and Solver.cu:526-1242 (merge_mark_gpu,
// merge_count_gpu, split_gpu, extract_split_particles_kernel,
// adaptive_sampling's compactions and appends).
//
// Deterministic semantics: the reference pairs merges with atomicCAS in
// scheduling order and hands out split slots with atomicAdd.  Here merge
// candidates are computed from the pre-kernel snapshot of remove[] and the
// claims are resolved exactly as if the threads ran in ascending index
// order; split children are appended in ascending parent order.
#include <hip/hip_runtime.h>

#include <climits>

#include "gg_device.hpp"
#include "gg_internal.hpp"
#include "grid_kernels.hpp"
#include "profiler.hpp"

namespace gg {

enum MergeState : int {
  M_NONE = 0,
  M_UND = 1,          // undecided
  M_CLAIMED = 2,      // taken as a target by an earlier claimer
  M_SELF_LOCK = 3,    // own CAS succeeded, own conditions failed: stays -1
  M_RELEASE = 4,      // partner already taken: released
  M_SUCCESS = 5,      // merged with its candidate
  M_PARTNER_LOCK = 6  // partner locked (-1) but not merged, self released
};

// Approximate 1/sqrt(x): Lomont's bit-level estimate ("Fast Inverse Square
// Root", 2003) and two Newton-Raphson steps, relative error below 5e-6.
// The reference normalizes with the device's approximate rsqrtf; an
// approximation (rather than a correctly rounded 1/sqrt) keeps the cover
// vector of a regular-lattice interior particle nonzero, as there.
__device__ __forceinline__ float rsqrt_approx(float x) {
  float y = __uint_as_float(0x5f3759dfu - (__float_as_uint(x) >> 1));
  const float hx = 0.5f * x;
  y = y * (1.5f - hx * y * y);
  y = y * (1.5f - hx * y * y);
  return y;
}

// normalize(v) = v * rsqrt(v.v) (helper_math.h:1325-1329); 0 for v = 0.
__device__ __forceinline__ float3 nrm(float x, float y, float z) {
  const float inv = rsqrt_approx(x * x + y * y + z * z);
  return make_float3(x * inv, y * inv, z * inv);
}

// f(j) for every particle j of the stencil of trunc(p/L), in scan order.
// False only when fl(sqrt(d2)) <= lim is impossible (the prefilter of the
// contact terms, whose bound holds while lim^2 is a normal float; below
// that, and for NaN, the exact test decides alone).
__device__ __forceinline__ bool maybe_within(float d2, float lim) {
  return !(lim >= 0x1p-60f) || d2 <= contact_prefilter(lim);
}

// Calls f(j, x_j, a[j]) for the particles of the stencil in scan order, with
// the positions and a[j] of four candidates loaded ahead of their calls, so
// that a candidate's test does not wait for its own loads one at a time.
template <class T, class F>
__device__ __forceinline__ void for_stencil_loaded(float px, float py, float pz,
                                                   const DevGrid &g, const int *__restrict__ cs,
                                                   const float *__restrict__ pos,
                                                   const T *__restrict__ a, F &&f) {
  int bx, by, bz;
  if (!trunc_cell(px, py, pz, g, bx, by, bz)) return;
  stencil_rows(bx, by, bz, g, cs, [&](int j0, int j1) {
    int j = j0;
    for (; j + 4 <= j1; j += 4) {
      Packed3 q[4];
      T v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        q[u] = ld3(pos, j + u);
        v[u] = a[j + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) f(j + u, q[u], v[u]);
    }
    for (; j < j1; ++j) f(j, ld3(pos, j), a[j]);
  });
}

template <class F>
__device__ __forceinline__ void for_stencil(float px, float py, float pz, const DevGrid &g,
                                            const int *__restrict__ cs, F &&f) {
  int bx, by, bz;
  if (!trunc_cell(px, py, pz, g, bx, by, bz)) return;
  stencil_rows(bx, by, bz, g, cs, [&](int j0, int j1) {
    for (int j = j0; j < j1; ++j) f(j);
  });
}

// Surface radius (GranularSystem.cu:518): cbrtf(3m / (4*PI_d*density)).
__global__ void __launch_bounds__(kBlock) k_radius_surface(int n, const float *__restrict__ mass,
                                                           double denom, float *__restrict__ rs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float three_m = 3.0f * mass[i];
  rs[i] = cbrt_f((float)((double)three_m / denom));
}

// K5: cover vector = sum of normalize(x_i - x_j) over neighbours with
// |x_i - x_j| <= 2 (r_i + r_j).
__global__ void __launch_bounds__(kBlock)
    k_cover(int n, const float *__restrict__ pos, const float *__restrict__ rs,
            const int *__restrict__ cs, DevGrid g, float4 *__restrict__ cover) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
  const float ri = rs[i];
  float cx = 0.0f, cy = 0.0f, cz = 0.0f;
  for_stencil_loaded(px, py, pz, g, cs, pos, rs, [&](int j, const Packed3 &q, float rj) {
    if (j == i) return;
    const float bx = px - q.x, by = py - q.y, bz = pz - q.z;
    const float d2 = bx * bx + by * by + bz * bz;
    const float lim = 2.0f * (ri + rj);
    // (a candidate failing the prefilter fails the test: no square root)
    if (!maybe_within(d2, lim)) return;
    const float dis = sqrtf(d2);
    if (!(dis <= lim)) return;
    const float3 u = nrm(bx, by, bz);
    cx = cx + u.x;
    cy = cy + u.y;
    cz = cz + u.z;
  });
  cover[i] = make_float4(cx, cy, cz, 0.0f);
}

// K6: surface <=> n <= 5 or no neighbour inside the pi/3 cone around the
// normalised cover vector (acos(d) <= pi/3 tested as d >= 0.5).
__global__ void __launch_bounds__(kBlock)
    k_surface(int n, const float *__restrict__ pos, const float *__restrict__ rs,
              const float4 *__restrict__ cover, const int *__restrict__ cs, DevGrid g,
              int *__restrict__ surf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
  const float ri = rs[i];
  const float4 c = cover[i];
  const float3 cv = nrm(c.x, c.y, c.z);
  int nn = 0;
  bool vote = false;
  for_stencil_loaded(px, py, pz, g, cs, pos, rs, [&](int j, const Packed3 &q, float rj) {
    if (j == i) return;
    const float bx = q.x - px, by = q.y - py, bz = q.z - pz;
    const float d2 = bx * bx + by * by + bz * bz;
    const float lim = 2.0f * (ri + rj);
    if (!maybe_within(d2, lim)) return;
    const float dis = sqrtf(d2);
    if (!(dis <= lim)) return;
    const float3 u = nrm(bx, by, bz);
    if (u.x * cv.x + u.y * cv.y + u.z * cv.z >= 0.5f) vote = true;
    ++nn;
  });
  surf[i] = (nn <= 5 || !vote) ? 1 : 0;
}

// Merge / split rules.  Reference rules (Solver.cu:526-653, 780-827):
// "deep" = not surface and no surface particle in the stencil (sd > 100:
// sd is a squared distance or the sentinel 1e5), partners within 0.035 m,
// merging needs 1 <= m < max_mass, splitting a surface particle with m > 2.
// Depth rules (radius levels): r = the surface radius of the mass; deep =
// not surface and sd > (merge_depth r)^2, partners within merge_dist
// (r_i + r_j), splitting needs sd < (split_depth r)^2 and m > split_mass.
struct MergeRules {
  int depth;
  float min_mass, split_mass, merge_dist, merge_depth, split_depth;
};

inline MergeRules rules_of(const gg_params *p) {
  MergeRules R;
  R.depth = p->adaptive_rules;
  R.min_mass = p->min_mass;
  R.split_mass = p->split_mass;
  R.merge_dist = p->merge_dist;
  R.merge_depth = p->merge_depth;
  R.split_depth = p->split_depth;
  return R;
}

__device__ __forceinline__ bool deep(const MergeRules &R, int surf, float sd, float r) {
  if (!R.depth) return surf == 0 && sd > 100.0f;
  const float t = R.merge_depth * r;
  return surf == 0 && sd > t * t;
}

__device__ __forceinline__ float partner_dist(const MergeRules &R, float ri, float rj) {
  return R.depth ? R.merge_dist * (ri + rj) : R.merge_dist;
}

// K7: squared distance to the nearest surface particle of the stencil.
__global__ void __launch_bounds__(kBlock)
    k_surface_distance(int n, const float *__restrict__ pos, const int *__restrict__ surf,
                       const int *__restrict__ cs, DevGrid g, float *__restrict__ sd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (surf[i] == 1) {
    sd[i] = 0.0f;
    return;
  }
  const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
  float best = 100000.0f;
  for_stencil_loaded(px, py, pz, g, cs, pos, surf, [&](int j, const Packed3 &q, int sj) {
    if (j == i || sj == 0) return;
    const float dx = px - q.x, dy = py - q.y, dz = pz - q.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < best) best = d2;
  });
  sd[i] = best;
}

// K8, first part: merge candidate and local density from the snapshot.
__global__ void __launch_bounds__(kBlock)
    k_merge_cand(int n, const float *__restrict__ pos, const float *__restrict__ mass,
                 const int *__restrict__ surf, const float *__restrict__ sd,
                 const int *__restrict__ remove, const int *__restrict__ cs, DevGrid g,
                 float max_mass, MergeRules R, const float *__restrict__ rs,
                 int *__restrict__ cand, float *__restrict__ dens, int *__restrict__ state,
                 int *__restrict__ owner, int *__restrict__ und0, int *__restrict__ und1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  owner[i] = INT_MAX;
  und0[i] = INT_MAX;
  und1[i] = INT_MAX;
  cand[i] = -1;
  state[i] = M_NONE;
  const float mi = mass[i];
  const float ri = R.depth ? rs[i] : 0.0f;
  if (surf[i] != 0 || remove[i] != 0 || mi < R.min_mass || mi >= max_mass ||
      !deep(R, surf[i], sd[i], ri))
    return;
  const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
  float closest = 1000.0f;
  int ci = -1;
  float ld = 0.0f;
  int lc = 0;
  // (a = the radii under the depth rules, else any array of n floats)
  const float *a = R.depth ? rs : mass;
  for_stencil_loaded(px, py, pz, g, cs, pos, a, [&](int j, const Packed3 &q, float aj) {
    if (j >= n) return;
    const float dx = px - q.x, dy = py - q.y, dz = pz - q.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float lim = partner_dist(R, ri, R.depth ? aj : 0.0f);
    // everything below needs dis < lim, which a candidate failing the
    // prefilter cannot meet
    if (!maybe_within(d2, lim)) return;
    const float dis = sqrtf(d2);
    if (dis < lim) {
      ld = ld + mass[j];
      ++lc;
    }
    if (j <= i || remove[j] != 0 || surf[j] != 0) return;
    if (mi + mass[j] > max_mass) return;
    if (dis < lim && dis < closest) {
      closest = dis;
      ci = j;
    }
  });
  if (lc > 0) {
    cand[i] = ci;
    dens[i] = ld / (float)lc;
    if (ci >= 0) state[i] = M_UND;
  }
}

// Round phase A: per target, lowest undecided claimer and lowest claimer
// that took it (integer minima: order-independent).
__global__ void __launch_bounds__(kBlock)
    k_merge_fold(int n, const int *__restrict__ cand, const int *__restrict__ state,
                 int *__restrict__ owner, int *__restrict__ und) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int st = state[i];
  if (st == M_UND)
    atomicMin(&und[cand[i]], i);
  else if (st == M_SUCCESS || st == M_PARTNER_LOCK)
    atomicMin(&owner[cand[i]], i);
}

// Round phase B: a candidate decides once every claimer below it that can
// affect it has decided.  Resets the other und buffer for the next round.
__global__ void __launch_bounds__(kBlock)
    k_merge_decide(int n, const int *__restrict__ cand, int *__restrict__ state,
                   const int *__restrict__ owner, const int *__restrict__ und,
                   int *__restrict__ und_next, const int *__restrict__ surf,
                   const float *__restrict__ sd, MergeRules R, const float *__restrict__ rs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  und_next[i] = INT_MAX;
  if (state[i] != M_UND) return;
  if (owner[i] != INT_MAX) {  // every claimer of i has a lower index
    state[i] = M_CLAIMED;
    return;
  }
  if (und[i] != INT_MAX) return;
  if (!deep(R, surf[i], sd[i], R.depth ? rs[i] : 0.0f)) {
    state[i] = M_SELF_LOCK;
    return;
  }
  const int c = cand[i];
  if (owner[c] < i) {
    state[i] = M_RELEASE;
    return;
  }
  if (und[c] < i) return;
  state[i] = deep(R, surf[c], sd[c], R.depth ? rs[c] : 0.0f) ? M_SUCCESS : M_PARTNER_LOCK;
}

// Ordered list of the still-undecided candidates (one block per 256
// particles; block offsets from a scan of per-block counts).
__global__ void __launch_bounds__(kBlock) k_count_und(int n, const int *__restrict__ state,
                                                      int *__restrict__ bsum) {
  __shared__ int ws[kBlock / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool f = i < n && state[i] == M_UND;
  const int c = (int)__popcll(__ballot(f));
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += ws[w];
    bsum[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kBlock)
    k_list_und(int n, const int *__restrict__ state, const int *__restrict__ boff,
               int *__restrict__ list) {
  __shared__ int ws[kBlock / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool f = i < n && state[i] == M_UND;
  const unsigned long long b = __ballot(f);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) ws[w] = (int)__popcll(b);
  __syncthreads();
  int off = boff[blockIdx.x];
  for (int k = 0; k < w; ++k) off += ws[k];
  if (f) list[off + (int)__popcll(b & ((1ull << lane) - 1ull))] = i;
}

// Sequential finisher for whatever the rounds left undecided, in index order.
__global__ void k_merge_finish(const int *__restrict__ list, const int *__restrict__ count,
                               const int *__restrict__ cand, int *__restrict__ state,
                               int *__restrict__ owner, const int *__restrict__ surf,
                               const float *__restrict__ sd, MergeRules R,
                               const float *__restrict__ rs) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int m = *count;
  for (int k = 0; k < m; ++k) {
    const int i = list[k];
    const int c = cand[i];
    int st;
    if (owner[i] != INT_MAX)
      st = M_CLAIMED;
    else if (!deep(R, surf[i], sd[i], R.depth ? rs[i] : 0.0f))
      st = M_SELF_LOCK;
    else if (owner[c] < i)
      st = M_RELEASE;
    else
      st = deep(R, surf[c], sd[c], R.depth ? rs[c] : 0.0f) ? M_SUCCESS : M_PARTNER_LOCK;
    state[i] = st;
    if ((st == M_SUCCESS || st == M_PARTNER_LOCK) && i < owner[c]) owner[c] = i;
  }
}

// K8, final writes.
__global__ void __launch_bounds__(kBlock)
    k_merge_apply(int n, const int *__restrict__ cand, const int *__restrict__ state,
                  const float *__restrict__ dens, const float *__restrict__ mass,
                  const float *__restrict__ vel, int *__restrict__ remove,
                  float *__restrict__ merge, float *__restrict__ mv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int st = state[i];
  if (st == M_SELF_LOCK) {
    remove[i] = -1;
  } else if (st == M_PARTNER_LOCK) {
    remove[cand[i]] = -1;
  } else if (st == M_SUCCESS) {
    const int c = cand[i];
    remove[i] = -1;
    remove[c] = -1;
    const float mi = mass[i], mc = mass[c];
    merge[c] = mi;
    merge[i] = -mi;
    const float mt = 1.0f / (mi + mc);
    const float vx = mt * (mi * vel[3 * i] + mc * vel[3 * c]);
    const float vy = mt * (mi * vel[3 * i + 1] + mc * vel[3 * c + 1]);
    const float vz = mt * (mi * vel[3 * i + 2] + mc * vel[3 * c + 2]);
    mv[3 * i] = -vx;
    mv[3 * i + 1] = -vy;
    mv[3 * i + 2] = -vz;
    const float df = fminf(1.0f, dens[i] / 100.0f);
    const float boost = 1.0f + (1.0f - df) * 0.1f;
    mv[3 * c] = vx * boost;
    mv[3 * c + 1] = vy * boost;
    mv[3 * c + 2] = vz * boost;
  }
}

// K9: 40-step blend.
__global__ void __launch_bounds__(kBlock)
    k_merge_blend(int n, float *__restrict__ merge, float *__restrict__ mass,
                  float *__restrict__ vel, int *__restrict__ mcount, int *__restrict__ remove,
                  const float *__restrict__ mv, int blend, int *__restrict__ adl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (remove[i] != -1) return;
  const int old = mcount[i];
  mcount[i] = old + 1;
  const float bf = (float)blend;
  const float dm = merge[i] / bf;
  const float dvx = mv[3 * i] / bf, dvy = mv[3 * i + 1] / bf, dvz = mv[3 * i + 2] / bf;
  mass[i] = mass[i] + dm;
  adl[i] = 1;
  vel[3 * i] = vel[3 * i] + dvx;
  vel[3 * i + 1] = vel[3 * i + 1] + dvy;
  vel[3 * i + 2] = vel[3 * i + 2] + dvz;
  if (old + 1 == blend) {
    if (merge[i] < 0.0f) {
      remove[i] = 1;
      adl[i] = 0;
    } else {
      remove[i] = 0;
    }
    mcount[i] = 0;
    merge[i] = 0.0f;
  }
}

// K10 check: child at the first empty non-centre stencil cell, no boundary
// particle within 2r.  Also counts survivors (remove != 1) and candidates
// per block for the ordered compaction.
__global__ void __launch_bounds__(kBlock)
    k_split_check(int n, const float *__restrict__ pos, const float *__restrict__ mass,
                  const int *__restrict__ remove, const int *__restrict__ surf,
                  const float *__restrict__ sd, const float *__restrict__ rs, MergeRules R,
                  const float4 *__restrict__ bnd4, const int *__restrict__ cs,
                  const int *__restrict__ bcs, DevGrid g, float denom_f,
                  int *__restrict__ ok, float4 *__restrict__ child, int *__restrict__ bsurv,
                  int *__restrict__ bsplit) {
  __shared__ int w1[kBlock / 64], w2[kBlock / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool split = false, surv = false;
  if (i < n) {
    surv = remove[i] != 1;
    const float m = mass[i];
    bool want = remove[i] == 0 && m > R.split_mass;
    if (want && R.depth) {
      const float t = R.split_depth * rs[i];
      want = sd[i] < t * t;
    } else if (want) {
      want = surf[i] == 1;
    }
    if (want) {
      const float r = cbrt_f((3.0f * m) / denom_f);
      const float maxd = 2.0f * r;
      const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
      int bx, by, bz;
      if (trunc_cell(px, py, pz, g, bx, by, bz)) {
        bool found = false, blocked = false;
        float4 cc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (int mm = 0; mm < 27 && !blocked; ++mm) {
          const int cell = stencil_cell(bx, by, bz, mm, g);
          if (cell < 0) continue;
          if (!found && mm != 13 && cs[cell] == cs[cell + 1] && bcs[cell] == bcs[cell + 1]) {
            const int x = bx + mm / 9 - 1, y = by + (mm % 9) / 3 - 1, z = bz + mm % 3 - 1;
            cc = make_float4(((float)x + 0.5f) * g.L, ((float)y + 0.5f) * g.L,
                             ((float)z + 0.5f) * g.L, 0.0f);
            found = true;
          }
          for (int k = bcs[cell]; k < bcs[cell + 1]; ++k) {
            const float4 q = bnd4[k];
            const float dx = px - q.x, dy = py - q.y, dz = pz - q.z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (maybe_within(d2, maxd) && sqrtf(d2) < maxd) {
              blocked = true;
              break;
            }
          }
        }
        split = found && !blocked;
        child[i] = cc;
      }
    }
    ok[i] = split ? 1 : 0;
  }
  const int c1 = (int)__popcll(__ballot(surv));
  const int c2 = (int)__popcll(__ballot(split));
  if ((threadIdx.x & 63) == 0) {
    w1[threadIdx.x >> 6] = c1;
    w2[threadIdx.x >> 6] = c2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int a = 0, b = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      a += w1[w];
      b += w2[w];
    }
    bsurv[blockIdx.x] = a;
    bsplit[blockIdx.x] = b;
  }
}

// Compaction + split: survivors keep their order, children follow in parent
// order, at most cap - survivors children (later candidates do not split).
// Writes into the staging arrays; counts[0..3] = survivors, children made,
// candidates refused, new n.
__global__ void __launch_bounds__(kBlock)
    k_compact(int n, int cap, const float *__restrict__ pos, const float *__restrict__ vel,
              const float *__restrict__ mass, const int *__restrict__ adl,
              const int *__restrict__ remove, const float *__restrict__ merge,
              const int *__restrict__ mcount, const int *__restrict__ ok,
              const float4 *__restrict__ child, const int *__restrict__ osurv,
              const int *__restrict__ osplit, int nblocks, float *__restrict__ t_pos,
              float *__restrict__ t_vel, float *__restrict__ t_mass, int *__restrict__ t_adl,
              int *__restrict__ t_rem, float *__restrict__ t_mer, int *__restrict__ t_mc,
              int *__restrict__ counts) {
  __shared__ int w1[kBlock / 64], w2[kBlock / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nsurv = osurv[nblocks], nsplit = osplit[nblocks];
  const int room = cap - nsurv > 0 ? cap - nsurv : 0;
  const int made = nsplit < room ? nsplit : room;
  if (i == 0) {
    counts[0] = nsurv;
    counts[1] = made;
    counts[2] = nsplit - made;
    counts[3] = nsurv + made;
  }
  const bool surv = i < n && remove[i] != 1;
  const bool split = i < n && ok[i] != 0;
  const unsigned long long b1 = __ballot(surv), b2 = __ballot(split);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    w1[w] = (int)__popcll(b1);
    w2[w] = (int)__popcll(b2);
  }
  __syncthreads();
  if (i >= n) return;
  int s1 = osurv[blockIdx.x], s2 = osplit[blockIdx.x];
  for (int k = 0; k < w; ++k) {
    s1 += w1[k];
    s2 += w2[k];
  }
  const unsigned long long below = (1ull << lane) - 1ull;
  s1 += (int)__popcll(b1 & below);
  s2 += (int)__popcll(b2 & below);
  const bool makes = split && s2 < made;
  float m = mass[i];
  if (makes) m = m / 2.0f;
  if (surv) {
    t_pos[3 * s1] = pos[3 * i];
    t_pos[3 * s1 + 1] = pos[3 * i + 1];
    t_pos[3 * s1 + 2] = pos[3 * i + 2];
    t_vel[3 * s1] = vel[3 * i];
    t_vel[3 * s1 + 1] = vel[3 * i + 1];
    t_vel[3 * s1 + 2] = vel[3 * i + 2];
    t_mass[s1] = m;
    t_adl[s1] = makes ? 1 : adl[i];
    t_rem[s1] = remove[i];
    t_mer[s1] = merge[i];
    t_mc[s1] = mcount[i];
  }
  if (makes) {
    const int d = nsurv + s2;
    const float4 c = child[i];
    t_pos[3 * d] = c.x;
    t_pos[3 * d + 1] = c.y;
    t_pos[3 * d + 2] = c.z;
    t_vel[3 * d] = vel[3 * i];
    t_vel[3 * d + 1] = vel[3 * i + 1];
    t_vel[3 * d + 2] = vel[3 * i + 2];
    t_mass[d] = m;
    t_adl[d] = 1;
    t_rem[d] = 0;
    t_mer[d] = 0.0f;
    t_mc[d] = 0;
  }
}

// Copies the first counts[3] staged records back into the caller's arrays.
__global__ void __launch_bounds__(kBlock)
    k_copy_back(int cap, const int *__restrict__ counts, const float *__restrict__ t_pos,
                const float *__restrict__ t_vel, const float *__restrict__ t_mass,
                const int *__restrict__ t_adl, const int *__restrict__ t_rem,
                const float *__restrict__ t_mer, const int *__restrict__ t_mc,
                float *__restrict__ pos, float *__restrict__ vel, float *__restrict__ mass,
                int *__restrict__ adl, int *__restrict__ rem, float *__restrict__ mer,
                int *__restrict__ mc) {
  const int m = counts[3];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m && k < cap;
       k += gridDim.x * blockDim.x) {
    pos[3 * k] = t_pos[3 * k];
    pos[3 * k + 1] = t_pos[3 * k + 1];
    pos[3 * k + 2] = t_pos[3 * k + 2];
    vel[3 * k] = t_vel[3 * k];
    vel[3 * k + 1] = t_vel[3 * k + 1];
    vel[3 * k + 2] = t_vel[3 * k + 2];
    mass[k] = t_mass[k];
    adl[k] = t_adl[k];
    rem[k] = t_rem[k];
    mer[k] = t_mer[k];
    mc[k] = t_mc[k];
  }
}

static inline int blocks_for(long long n) { return (int)((n + kBlock - 1) / kBlock); }

static Layout layout_of(const gg_workspace *ws) {
  return make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                     ws->bmax);
}

static int check_adaptive_set(const gg_particles *set, const gg_workspace *ws) {
  GG_TRY(check_ws(ws));
  if (!set || set->n < 0) return GG_ERR_INVALID;
  if (set->n > ws->capacity) return GG_ERR_CAPACITY;
  if (set->n > 0 && (!set->pos || !set->mass)) return GG_ERR_INVALID;
  return GG_OK;
}

}  // namespace gg

using namespace gg;

extern "C" int gg_surface(gg_particles *set, const int *cell_start, const gg_params *params,
                          const gg_workspace *ws, hipStream_t stream) {
  GG_TRY(check_adaptive_set(set, ws));
  if (!cell_start || !params || !set->surface) return GG_ERR_INVALID;
  const int n = set->n;
  if (n == 0) return GG_OK;
  Layout L = layout_of(ws);
  const DevGrid g = dev_grid(ws->grid);
  const double denom = 4.0 * kPiD * (double)params->density;
  float *rs = at<float>(ws, L.dens);
  prof_begin(GG_STAGE_ADAPTIVE, stream);
  k_radius_surface<<<blocks_for(n), kBlock, 0, stream>>>(n, set->mass, denom, rs);
  GG_LAUNCH_CHECK();
  const float *pos = reinterpret_cast<const float *>(set->pos);
  k_cover<<<blocks_for(n), kBlock, 0, stream>>>(n, pos, rs, cell_start, g,
                                                at<float4>(ws, L.cover));
  GG_LAUNCH_CHECK();
  k_surface<<<blocks_for(n), kBlock, 0, stream>>>(n, pos, rs, at<float4>(ws, L.cover),
                                                  cell_start, g, set->surface);
  GG_LAUNCH_CHECK();
  prof_end(GG_STAGE_ADAPTIVE, stream);
  return GG_OK;
}

extern "C" int gg_surface_distance(gg_particles *set, const int *cell_start,
                                   const gg_workspace *ws, hipStream_t stream) {
  GG_TRY(check_adaptive_set(set, ws));
  if (!cell_start || !set->surface || !set->surface_distance) return GG_ERR_INVALID;
  const int n = set->n;
  if (n == 0) return GG_OK;
  prof_begin(GG_STAGE_ADAPTIVE, stream);
  k_surface_distance<<<blocks_for(n), kBlock, 0, stream>>>(
      n, reinterpret_cast<const float *>(set->pos), set->surface, cell_start,
      dev_grid(ws->grid), set->surface_distance);
  GG_LAUNCH_CHECK();
  prof_end(GG_STAGE_ADAPTIVE, stream);
  return GG_OK;
}

static int scan_blocks(int *bsum, int nblocks, int *out, const gg_workspace *ws,
                       hipStream_t st) {
  return scan_exclusive(bsum, nblocks, out, ws, st);
}

// Surface radii of the set (the depth rules' r) into the workspace.
static int merge_radii(const gg_particles *set, float density, const gg_workspace *ws,
                       hipStream_t stream, const float **out) {
  Layout L = layout_of(ws);
  float *rs = at<float>(ws, L.mrad);
  const double denom = 4.0 * kPiD * (double)density;
  k_radius_surface<<<blocks_for(set->n), kBlock, 0, stream>>>(set->n, set->mass, denom, rs);
  GG_LAUNCH_CHECK();
  *out = rs;
  return GG_OK;
}

static int merge_mark_rules(gg_particles *set, gg_adaptive *ad, const int *cell_start,
                            float max_mass, const MergeRules &R, const float *rs,
                            const gg_workspace *ws, hipStream_t stream) {
  GG_TRY(check_adaptive_set(set, ws));
  if (!ad || !ad->remove || !ad->merge || !ad->merge_velocity || !cell_start || !set->vel ||
      !set->surface || !set->surface_distance)
    return GG_ERR_INVALID;
  const int n = set->n;
  if (n == 0) return GG_OK;
  Layout L = layout_of(ws);
  const DevGrid g = dev_grid(ws->grid);
  const int nb = blocks_for(n);
  int *cand = at<int>(ws, L.cand), *state = at<int>(ws, L.mstate), *owner = at<int>(ws, L.owner);
  int *und0 = at<int>(ws, L.und0), *und1 = at<int>(ws, L.und1);
  float *dens = at<float>(ws, L.dens);
  const float *pos = reinterpret_cast<const float *>(set->pos);
  prof_begin(GG_STAGE_ADAPTIVE, stream);
  k_merge_cand<<<nb, kBlock, 0, stream>>>(n, pos, set->mass, set->surface, set->surface_distance,
                                          ad->remove, cell_start, g, max_mass, R, rs, cand, dens,
                                          state, owner, und0, und1);
  GG_LAUNCH_CHECK();
  for (int r = 0; r < kMergeRounds; ++r) {
    int *u = (r % 2 == 0) ? und0 : und1;
    int *un = (r % 2 == 0) ? und1 : und0;
    k_merge_fold<<<nb, kBlock, 0, stream>>>(n, cand, state, owner, u);
    GG_LAUNCH_CHECK();
    k_merge_decide<<<nb, kBlock, 0, stream>>>(n, cand, state, owner, u, un, set->surface,
                                              set->surface_distance, R, rs);
    GG_LAUNCH_CHECK();
  }
  // fold the last round's decisions, then finish the rest in index order
  int *ufin = (kMergeRounds % 2 == 0) ? und0 : und1;
  k_merge_fold<<<nb, kBlock, 0, stream>>>(n, cand, state, owner, ufin);
  GG_LAUNCH_CHECK();
  int *bsum = at<int>(ws, L.bsum), *boff = at<int>(ws, L.bsum2);
  int *list = at<int>(ws, L.flags);
  k_count_und<<<nb, kBlock, 0, stream>>>(n, state, bsum);
  GG_LAUNCH_CHECK();
  GG_TRY(scan_blocks(bsum, nb, boff, ws, stream));
  k_list_und<<<nb, kBlock, 0, stream>>>(n, state, boff, list);
  GG_LAUNCH_CHECK();
  k_merge_finish<<<1, 64, 0, stream>>>(list, boff + nb, cand, state, owner, set->surface,
                                       set->surface_distance, R, rs);
  GG_LAUNCH_CHECK();
  k_merge_apply<<<nb, kBlock, 0, stream>>>(n, cand, state, dens, set->mass,
                                           reinterpret_cast<const float *>(set->vel), ad->remove,
                                           ad->merge,
                                           reinterpret_cast<float *>(ad->merge_velocity));
  GG_LAUNCH_CHECK();
  prof_end(GG_STAGE_ADAPTIVE, stream);
  return GG_OK;
}

extern "C" int gg_merge_mark(gg_particles *set, gg_adaptive *ad, const int *cell_start,
                             float max_mass, const gg_workspace *ws, hipStream_t stream) {
  gg_params ref;
  gg_params_default(&ref);
  return merge_mark_rules(set, ad, cell_start, max_mass, rules_of(&ref), nullptr, ws, stream);
}

extern "C" int gg_merge_blend(gg_particles *set, gg_adaptive *ad, int blend_steps,
                              const gg_workspace *ws, hipStream_t stream) {
  GG_TRY(check_adaptive_set(set, ws));
  if (!ad || !ad->remove || !ad->merge || !ad->merge_velocity || !ad->merge_count ||
      !set->vel || !set->adaptive_last_step || blend_steps <= 0)
    return GG_ERR_INVALID;
  const int n = set->n;
  if (n == 0) return GG_OK;
  prof_begin(GG_STAGE_ADAPTIVE, stream);
  k_merge_blend<<<blocks_for(n), kBlock, 0, stream>>>(
      n, ad->merge, set->mass, reinterpret_cast<float *>(set->vel), ad->merge_count, ad->remove,
      reinterpret_cast<const float *>(ad->merge_velocity), blend_steps, set->adaptive_last_step);
  GG_LAUNCH_CHECK();
  prof_end(GG_STAGE_ADAPTIVE, stream);
  return GG_OK;
}

static int split_compact_rules(gg_particles *set, gg_adaptive *ad, const int *cell_start,
                               const gg_particles *boundary, const int *bcell_start,
                               const gg_params *params, const float *rs,
                               const gg_workspace *ws, hipStream_t stream, int *stats) {
  GG_TRY(check_adaptive_set(set, ws));
  if (!ad || !ad->remove || !ad->merge || !ad->merge_count || !cell_start || !boundary ||
      !bcell_start || !params || !set->vel || !set->surface || !set->adaptive_last_step)
    return GG_ERR_INVALID;
  const MergeRules R = rules_of(params);
  if (R.depth && (!set->surface_distance || !rs)) return GG_ERR_INVALID;
  const int n = set->n;
  const int cap = set->capacity < ws->capacity ? set->capacity : ws->capacity;
  Layout L = layout_of(ws);
  int *counts = at<int>(ws, L.ctr2);
  int host[4] = {0, 0, 0, 0};
  if (n > 0) {
    const int nb = blocks_for(n);
    const DevGrid g = dev_grid(ws->grid);
    const float denom_f = (4.0f * kPiF) * params->density;
    int *ok = at<int>(ws, L.flags);
    int *bsurv = at<int>(ws, L.bsum), *bsplit = at<int>(ws, L.bsum2);
    int *osurv = at<int>(ws, L.und0), *osplit = at<int>(ws, L.und1);
    prof_begin(GG_STAGE_ADAPTIVE, stream);
    k_split_check<<<nb, kBlock, 0, stream>>>(
        n, reinterpret_cast<const float *>(set->pos), set->mass, ad->remove, set->surface,
        set->surface_distance, rs, R, at<float4>(ws, L.bnd4), cell_start, bcell_start, g, denom_f,
        ok, at<float4>(ws, L.child), bsurv, bsplit);
    GG_LAUNCH_CHECK();
    GG_TRY(scan_exclusive(bsurv, nb, osurv, ws, stream));
    GG_TRY(scan_exclusive(bsplit, nb, osplit, ws, stream));
    float *tp = at<float>(ws, L.v3tmp0), *tv = at<float>(ws, L.v3tmp1);
    float *tm = at<float>(ws, L.ftmp0), *tme = at<float>(ws, L.ftmp1);
    int *ta = at<int>(ws, L.itmp0), *tr = at<int>(ws, L.itmp1), *tc = at<int>(ws, L.itmp2);
    k_compact<<<nb, kBlock, 0, stream>>>(
        n, cap, reinterpret_cast<const float *>(set->pos),
        reinterpret_cast<const float *>(set->vel), set->mass, set->adaptive_last_step, ad->remove,
        ad->merge, ad->merge_count, ok, at<float4>(ws, L.child), osurv, osplit, nb, tp, tv, tm,
        ta, tr, tme, tc, counts);
    GG_LAUNCH_CHECK();
    const int cb = blocks_for(cap) < 4096 ? blocks_for(cap) : 4096;
    k_copy_back<<<cb > 0 ? cb : 1, kBlock, 0, stream>>>(
        cap, counts, tp, tv, tm, ta, tr, tme, tc, reinterpret_cast<float *>(set->pos),
        reinterpret_cast<float *>(set->vel), set->mass, set->adaptive_last_step, ad->remove,
        ad->merge, ad->merge_count);
    GG_LAUNCH_CHECK();
    prof_end(GG_STAGE_ADAPTIVE, stream);
    GG_TRY(hip_status(hipMemcpyAsync(host, counts, sizeof(host), hipMemcpyDeviceToHost, stream)));
    GG_TRY(hip_status(hipStreamSynchronize(stream)));
  }
  if (stats) {
    stats[0] = n - host[0];  // removed
    stats[1] = host[1];      // split
    stats[2] = host[2];      // refused for capacity
  }
  set->n = n > 0 ? host[3] : 0;
  return GG_OK;
}

extern "C" int gg_split_compact(gg_particles *set, gg_adaptive *ad, const int *cell_start,
                                const gg_particles *boundary, const int *bcell_start,
                                const gg_params *params, const gg_workspace *ws,
                                hipStream_t stream, int *stats) {
  if (!params || !set) return GG_ERR_INVALID;
  const float *rs = nullptr;
  if (params->adaptive_rules && set->n > 0) {
    GG_TRY(check_adaptive_set(set, ws));
    GG_TRY(merge_radii(set, params->density, ws, stream, &rs));
  }
  return split_compact_rules(set, ad, cell_start, boundary, bcell_start, params, rs, ws, stream,
                             stats);
}

extern "C" int gg_adaptive_sampling(gg_particles *set, gg_adaptive *ad, const int *cell_start,
                                    const gg_particles *boundary, const int *bcell_start,
                                    const gg_params *params, float max_mass, int do_merge,
                                    int blend_steps, const gg_workspace *ws, hipStream_t stream,
                                    int *stats) {
  if (!params || !set) return GG_ERR_INVALID;
  const MergeRules R = rules_of(params);
  const float *rs = nullptr;
  if (R.depth && set->n > 0) {
    GG_TRY(check_adaptive_set(set, ws));
    GG_TRY(merge_radii(set, params->density, ws, stream, &rs));
  }
  if (do_merge) GG_TRY(merge_mark_rules(set, ad, cell_start, max_mass, R, rs, ws, stream));
  GG_TRY(gg_merge_blend(set, ad, blend_steps, ws, stream));
  // the split rule reads the radii of the masses after the blend
  if (R.depth && set->n > 0) GG_TRY(merge_radii(set, params->density, ws, stream, &rs));
  return split_compact_rules(set, ad, cell_start, boundary, bcell_start, params, rs, ws, stream,
                             stats);
}
