// Device-side building blocks shared by the kernels.  Every floating-point
// expression here is evaluated exactly as written (the library is compiled
// with -ffp-contract=off and IEEE division / square root), which is what
// makes the FP32 results reproducible bit for bit on the host.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gg_internal.hpp"

namespace gg {

struct DevGrid {
  float L;
  int cx, cy, cz;
};

inline DevGrid dev_grid(const gg_grid &g) { return DevGrid{g.cell_length, g.cx, g.cy, g.cz}; }

// Cell id of a position: floor(p/L) per axis, x-major, sentinel C outside
// the grid (CUDAFunctions.cuh:71-107).  NaN / huge coordinates -> sentinel.
__device__ __forceinline__ int hash_cell(float px, float py, float pz, const DevGrid &g) {
  const float fx = floorf(px / g.L);
  const float fy = floorf(py / g.L);
  const float fz = floorf(pz / g.L);
  if (!(fx >= 0.0f && fx < (float)g.cx && fy >= 0.0f && fy < (float)g.cy && fz >= 0.0f &&
        fz < (float)g.cz))
    return g.cx * g.cy * g.cz;
  return ((int)fx * g.cy + (int)fy) * g.cz + (int)fz;
}

// Stencil base: make_int3(p / L), truncation toward zero
// (helper_math.h:176-179 as used in Solver.cu:382).
__device__ __forceinline__ bool trunc_cell(float px, float py, float pz, const DevGrid &g, int &bx,
                                           int &by, int &bz) {
  const float qx = px / g.L, qy = py / g.L, qz = pz / g.L;
  const float lim = 1073741824.0f;
  if (!(qx > -lim && qx < lim && qy > -lim && qy < lim && qz > -lim && qz < lim)) return false;
  bx = (int)qx;
  by = (int)qy;
  bz = (int)qz;
  return true;
}

// Upsampler base: floorf(p / L) (Solver.cu:1271).
__device__ __forceinline__ bool floor_cell(float px, float py, float pz, const DevGrid &g, int &bx,
                                           int &by, int &bz) {
  const float qx = floorf(px / g.L), qy = floorf(py / g.L), qz = floorf(pz / g.L);
  const float lim = 1073741824.0f;
  if (!(qx > -lim && qx < lim && qy > -lim && qy < lim && qz > -lim && qz < lim)) return false;
  bx = (int)qx;
  by = (int)qy;
  bz = (int)qz;
  return true;
}

// Identifies a stencil base inside the grid padded by one cell on every side;
// -1 when the whole 27-cell stencil lies outside the grid.  Two bases with
// equal ids have identical stencils.
__device__ __forceinline__ int ext_base_id(bool ok, int bx, int by, int bz, const DevGrid &g) {
  if (!ok) return -1;
  if (bx < -1 || bx > g.cx || by < -1 || by > g.cy || bz < -1 || bz > g.cz) return -1;
  return ((bx + 1) * (g.cy + 2) + (by + 1)) * (g.cz + 2) + (bz + 1);
}

// Stencil cell m in 0..26, offsets (m/9-1, (m%9)/3-1, m%3-1) in ascending
// cell-id order (Solver.cu:381-386); -1 if outside the grid.
__device__ __forceinline__ int stencil_cell(int bx, int by, int bz, int m, const DevGrid &g) {
  const int x = bx + m / 9 - 1;
  const int y = by + (m % 9) / 3 - 1;
  const int z = bz + m % 3 - 1;
  if (x < 0 || x >= g.cx || y < 0 || y >= g.cy || z < 0 || z >= g.cz) return -1;
  return (x * g.cy + y) * g.cz + z;
}

// The 27-cell stencil of base (bx,by,bz) as 9 rows: for each (dx,dy), the
// three cells dz = -1..1 have consecutive ids, so their particles form one
// contiguous index range [cs[first], cs[last+1]).  Rows come in ascending
// cell-id order, so visiting rows in order and indices ascending inside a row
// is exactly the cell-by-cell scan order.  f(begin, end) is called per row.
template <class F>
__device__ __forceinline__ void stencil_rows(int bx, int by, int bz, const DevGrid &g,
                                             const int *__restrict__ cs, F &&f) {
  const int z0 = bz - 1 > 0 ? bz - 1 : 0;
  const int z1 = bz + 1 < g.cz - 1 ? bz + 1 : g.cz - 1;
  if (z0 > z1) return;
  for (int dx = -1; dx <= 1; ++dx) {
    const int x = bx + dx;
    if (x < 0 || x >= g.cx) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int y = by + dy;
      if (y < 0 || y >= g.cy) continue;
      const int c = (x * g.cy + y) * g.cz;
      f(cs[c + z0], cs[c + z1 + 1]);
    }
  }
}

// Element i of a packed float3 array (x, y, z at 3i), read with one 12-byte
// load per lane instead of three scalar ones.
struct Packed3 {
  float x, y, z;
};
__device__ __forceinline__ Packed3 ld3(const float *__restrict__ p, long long i) {
  return reinterpret_cast<const Packed3 *>(p)[i];
}
__device__ __forceinline__ void st3(float *__restrict__ p, long long i, float x, float y, float z) {
  reinterpret_cast<Packed3 *>(p)[i] = Packed3{x, y, z};
}

// exp / cbrt evaluated in double and rounded once to float.
__device__ __forceinline__ float exp_f(float x) { return (float)exp((double)x); }
__device__ __forceinline__ float cbrt_f(float x) { return (float)cbrt((double)x); }

// Contact radius (Solver.cu:281, 320): cbrtf(3m / (4*pi*int(density))) with
// the denominator formed in double on the host.
__device__ __forceinline__ float radius_contact(float m, double denom) {
  const float three_m = 3.0f * m;
  const float arg = (float)((double)three_m / denom);
  return cbrt_f(arg);
}

struct SweepConst {
  float rb;        // boundary particle radius
  float rmin;      // clamp of r_i against the boundary
  float mu_s, mu_k;
  float half_valid;  // list usable while every displacement < half_valid
  int gmax, bmax, cap;
};

// Scenes sharing one set (gg_step_scenes): the scene of each particle and
// the scenes' grids, cell offsets and friction.  With SCENES false (one
// scene) the launch's own grid at cell offset 0 and its own friction.
struct SceneView {
  const int *__restrict__ scene_of;
  const gg_scene_desc *__restrict__ desc;
  int count;  // scene indices are clamped to [0, count)
};
struct LocalGrid {
  DevGrid g;
  int base;  // offset of the scene's cells in the cell tables
  float mu_s, mu_k;
};
template <bool SCENES>
__device__ __forceinline__ LocalGrid local_grid(const SceneView &sv, int i, const DevGrid &g0,
                                                const SweepConst &sc) {
  if (!SCENES) return LocalGrid{g0, 0, sc.mu_s, sc.mu_k};
  const int k = sv.scene_of[i];
  const gg_scene_desc d = sv.desc[k < 0 ? 0 : (k < sv.count ? k : sv.count - 1)];
  return LocalGrid{DevGrid{g0.L, d.cx, d.cy, d.cz}, d.cell_base, d.mu_s, d.mu_k};
}

// Correctly rounded square root, reciprocal and quotient on restricted
// domains.  The compiler's IEEE sequences (what sqrtf and '/' compile to
// under -fhip-fp32-correctly-rounded-divide-sqrt) wrap a core in steps that
// are identities on these domains: for sqrt, the input scaling below 2^-96
// and the +-0 / +inf class check; for division, v_div_scale (no scaling when
// numerator and denominator are normal, the numerator's biased exponent
// exceeds 23, the exponent difference is below 96 and neither 1/d nor n/d
// is denormal), v_div_fmas (a plain fma when not scaled) and v_div_fixup
// (pass-through for finite normal quotients).  The cores below are those
// sequences without the identity steps, so their results equal sqrtf and
// '/' bit for bit on the domains (checked over every input by
// gg_fast_math_check); callers fall back to sqrtf and '/' outside them.
constexpr float kSqrtFastMin = 0x1p-96f;  // sqrt_core domain [2^-96, max float]

__device__ __forceinline__ float sqrt_core(float x) {
  float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __int_as_float(__float_as_int(s) - 1);
  const float sp = __int_as_float(__float_as_int(s) + 1);
  const float rm = __builtin_fmaf(-sm, s, x);
  const float rp = __builtin_fmaf(-sp, s, x);
  s = rm <= 0.0f ? sm : s;
  s = rp > 0.0f ? sp : s;
  return s;
}

// n / d for normal n, d with the conditions above.
__device__ __forceinline__ float div_core(float n, float d) {
  const float r0 = __builtin_amdgcn_rcpf(d);
  const float e0 = __builtin_fmaf(-d, r0, 1.0f);
  const float r1 = __builtin_fmaf(e0, r0, r0);
  const float q0 = n * r1;
  const float e1 = __builtin_fmaf(-d, q0, n);
  const float q1 = __builtin_fmaf(e1, r1, q0);
  const float e2 = __builtin_fmaf(-d, q1, n);
  return __builtin_fmaf(e2, r1, q1);
}

// 1 / d (the numerator is exactly 1, so n * r1 = r1).
__device__ __forceinline__ float rcp_core(float d) {
  const float r0 = __builtin_amdgcn_rcpf(d);
  const float e0 = __builtin_fmaf(-d, r0, 1.0f);
  const float r1 = __builtin_fmaf(e0, r0, r0);
  const float e1 = __builtin_fmaf(-d, r1, 1.0f);
  const float q1 = __builtin_fmaf(e1, r1, r1);
  const float e2 = __builtin_fmaf(-d, q1, 1.0f);
  return __builtin_fmaf(e2, r1, q1);
}

// d = sqrt(d2) and inv = 1/d, as sqrtf and 1.0f/d give them.  For d2 in
// [2^-96, max float], d lies in [2^-48, 2^64), where 1/d meets the
// division conditions.
__device__ __forceinline__ void sqrt_rcp(float d2, float &d, float &inv) {
  if (d2 >= kSqrtFastMin && d2 <= 3.40282347e+38f) {
    d = sqrt_core(d2);
    inv = rcp_core(d);
  } else {
    d = sqrtf(d2);
    inv = 1.0f / d;
  }
}

// PBD friction (Macklin et al. 2014, sec. 6.1) on the relative displacement;
// inv = 1/d of the contact.  Static when |t|^2 < (mu_s delta)^2 (squares:
// the static case needs no square root); otherwise t scaled by
// min(mu_k delta / |t|, 1).  In the kinetic case |t| = sqrt(tl2) with tl2
// in [2^-96, 2^80] and mu_k delta in [2^-80, 2^40] keep the quotient inside
// the division domain (quotient in [2^-120, 2^88]).
__device__ __forceinline__ float3 friction(float px, float py, float pz, float inv, float delta,
                                           float rx, float ry, float rz, float mu_s, float mu_k) {
  const float nx = px * inv, ny = py * inv, nz = pz * inv;
  const float rn = rx * nx + ry * ny + rz * nz;
  float tx = rx - rn * nx;
  float ty = ry - rn * ny;
  float tz = rz - rn * nz;
  const float tl2 = tx * tx + ty * ty + tz * tz;
  const float lim = mu_s * delta;
  if (tl2 < lim * lim) return make_float3(tx, ty, tz);
  const float num = mu_k * delta;
  float c;
  if (tl2 >= kSqrtFastMin && tl2 <= 0x1p80f && num >= 0x1p-80f && num <= 0x1p40f)
    c = div_core(num, sqrt_core(tl2));
  else
    c = num / sqrtf(tl2);
  c = fminf(c, 1.0f);
  return make_float3(tx * c, ty * c, tz * c);
}

// Necessary condition for fl(sqrt(d2)) <= rc with a correctly rounded sqrt:
// d2 <= rc^2 (1 + 2^-23) < fl(fl(rc*rc) * 1.000001f).  Skipping the square
// root when it fails cannot change any result.
__device__ __forceinline__ float contact_prefilter(float rc) { return rc * rc * 1.000001f; }

// Boundary contact terms of particle (bi, oi) with static boundary particle
// q: the correction c = s p and the friction f.  False when not in contact.
template <bool FRIC>
__device__ __forceinline__ bool boundary_terms(const float4 &bi, const float4 &oi, const float4 &q,
                                               float rc, float pre, const SweepConst &sc,
                                               float3 &c, float3 &f) {
  const float px = (bi.x - q.x) + oi.x;
  const float py = (bi.y - q.y) + oi.y;
  const float pz = (bi.z - q.z) + oi.z;
  const float d2 = px * px + py * py + pz * pz;
  if (!(d2 <= pre)) return false;
  float d, inv;
  sqrt_rcp(d2, d, inv);
  if (!(d < rc && d > 0.0f)) return false;
  const float s = (rc - d) * inv;
  c = make_float3(s * px, s * py, s * pz);
  if (FRIC) f = friction(px, py, pz, inv, rc - d, oi.x, oi.y, oi.z, sc.mu_s, sc.mu_k);
  return true;
}

// Granular contact terms of particle i with particle j (symmetric gather).
template <bool FRIC>
__device__ __forceinline__ bool granular_terms(const float4 &bi, const float4 &oi, const float4 &bj,
                                               const float4 &oj, const SweepConst &sc, float3 &c,
                                               float3 &f) {
  const float px = (bi.x - bj.x) + (oi.x - oj.x);
  const float py = (bi.y - bj.y) + (oi.y - oj.y);
  const float pz = (bi.z - bj.z) + (oi.z - oj.z);
  const float d2 = px * px + py * py + pz * pz;
  const float rc = bi.w + bj.w;
  if (!(d2 <= contact_prefilter(rc))) return false;
  float d, inv;
  sqrt_rcp(d2, d, inv);
  if (!(d <= rc && d > 0.0f)) return false;
  const float s = (rc - d) * inv;
  c = make_float3(s * px, s * py, s * pz);
  if (FRIC) {
    const float rx = oi.x - oj.x;
    const float ry = oi.y - oj.y;
    const float rz = oi.z - oj.z;
    f = friction(px, py, pz, inv, rc - d, rx, ry, rz, sc.mu_s, sc.mu_k);
  }
  return true;
}

// Adds one contact's terms: a + c, then - f (the order every path uses).
template <bool FRIC>
__device__ __forceinline__ void add_terms(float3 &a, const float3 &c, const float3 &f) {
  a.x = a.x + c.x;
  a.y = a.y + c.y;
  a.z = a.z + c.z;
  if (FRIC) {
    a.x = a.x - f.x;
    a.y = a.y - f.y;
    a.z = a.z - f.z;
  }
}

// Boundary contact of particle (bi, oi) with static boundary particle q.
template <bool FRIC>
__device__ __forceinline__ void boundary_contact(const float4 &bi, const float4 &oi, const float4 &q,
                                                 float rc, float pre, const SweepConst &sc,
                                                 float3 &ab, int &nb) {
  float3 c, f = make_float3(0.0f, 0.0f, 0.0f);
  if (!boundary_terms<FRIC>(bi, oi, q, rc, pre, sc, c, f)) return;
  add_terms<FRIC>(ab, c, f);
  nb += 1;
}

// Granular contact of particle i with particle j (symmetric gather).
template <bool FRIC>
__device__ __forceinline__ void granular_contact(const float4 &bi, const float4 &oi, const float4 &bj,
                                                 const float4 &oj, const SweepConst &sc, float3 &ag,
                                                 int &ng) {
  float3 c, f = make_float3(0.0f, 0.0f, 0.0f);
  if (!granular_terms<FRIC>(bi, oi, bj, oj, sc, c, f)) return;
  add_terms<FRIC>(ag, c, f);
  ng += 2;
}

__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Remapping hardware block b so that each XCD receives a contiguous
// range of logical blocks keeps a block's spatial neighbours (nearby
// particle indices) in the same L2.  Bijective on [0, nb).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int nb, int remap) {
  if (!remap) return b;
  const int x = b % kXcds, idx = b / kXcds;
  const int q = nb / kXcds, r = nb % kXcds;
  return x * q + (x < r ? x : r) + idx;
}

// In-grid cell id of stencil base (bx,by,bz), or -1.
__device__ __forceinline__ int base_cell_id(bool ok, int bx, int by, int bz, const DevGrid &g) {
  if (!ok || bx < 0 || bx >= g.cx || by < 0 || by >= g.cy || bz < 0 || bz >= g.cz) return -1;
  return (bx * g.cy + by) * g.cz + bz;
}

__device__ __forceinline__ int wave_sum(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace gg
