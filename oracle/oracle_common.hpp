// Shared helpers of the CPU oracle -- TEST INFRASTRUCTURE ONLY.
#pragma once
#include "oracle.h"

#include <cmath>
#include <cstdint>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

template <class T> struct V3 {
  T x, y, z;
};

template <class T> inline V3<T> add(V3<T> a, V3<T> b) {
  return V3<T>{a.x + b.x, a.y + b.y, a.z + b.z};
}

struct Grid {
  float L;
  int cx, cy, cz;
};

// Reference: particlePos2cellIdx + mapParticles2Cells_CUDA
// (CUDAFunctions.cuh:71-107): floorf(p/L) per axis with IEEE division,
// x-major id, sentinel C when outside [0,c).  Non-finite -> sentinel.
template <class T> inline int hash_cell(V3<T> p, T L, const Grid &g) {
  const T fx = std::floor(p.x / L);
  const T fy = std::floor(p.y / L);
  const T fz = std::floor(p.z / L);
  if (!(fx >= (T)0 && fx < (T)g.cx && fy >= (T)0 && fy < (T)g.cy &&
        fz >= (T)0 && fz < (T)g.cz))
    return g.cx * g.cy * g.cz;
  return ((int)fx * g.cy + (int)fy) * g.cz + (int)fz;
}

// Stencil base of the sweeps / surface / merge / split kernels:
// make_int3(pos / cell_length) truncates toward zero (helper_math.h:176-179).
// Quotients outside +-2^30 (or NaN) have no in-grid stencil cell.
template <class T>
inline bool trunc_cell(V3<T> p, T L, int &bx, int &by, int &bz) {
  const T qx = p.x / L, qy = p.y / L, qz = p.z / L;
  const T lim = (T)1073741824.0;
  if (!(qx > -lim && qx < lim && qy > -lim && qy < lim && qz > -lim &&
        qz < lim))
    return false;
  bx = (int)qx;
  by = (int)qy;
  bz = (int)qz;
  return true;
}

// Upsampler base: floorf(pos / cell_length) (Solver.cu:1271).
template <class T>
inline bool floor_cell(V3<T> p, T L, int &bx, int &by, int &bz) {
  const T qx = std::floor(p.x / L), qy = std::floor(p.y / L),
          qz = std::floor(p.z / L);
  const T lim = (T)1073741824.0;
  if (!(qx > -lim && qx < lim && qy > -lim && qy < lim && qz > -lim &&
        qz < lim))
    return false;
  bx = (int)qx;
  by = (int)qy;
  bz = (int)qz;
  return true;
}

inline int cell_id_unchecked(int x, int y, int z, const Grid &g) {
  if (x < 0 || x >= g.cx || y < 0 || y >= g.cy || z < 0 || z >= g.cz)
    return g.cx * g.cy * g.cz;
  return (x * g.cy + y) * g.cz + z;
}

// Identifier of a stencil base in the grid padded by one cell per side; -1
// when the whole stencil lies outside the grid.
inline int ext_base_id(int bx, int by, int bz, const Grid &g) {
  if (bx < -1 || bx > g.cx || by < -1 || by > g.cy || bz < -1 || bz > g.cz)
    return -1;
  return ((bx + 1) * (g.cy + 2) + (by + 1)) * (g.cz + 2) + (bz + 1);
}

// Stencil cell m in 0..26: offsets (m/9-1, (m%9)/3-1, m%3-1)
// (Solver.cu:381-386).  Returns -1 for out-of-grid cells (skipped).
inline int stencil_cell(int bx, int by, int bz, int m, const Grid &g) {
  const int x = bx + m / 9 - 1;
  const int y = by + (m % 9) / 3 - 1;
  const int z = bz + m % 3 - 1;
  if (x < 0 || x >= g.cx || y < 0 || y >= g.cy || z < 0 || z >= g.cz)
    return -1;
  return (x * g.cy + y) * g.cz + z;
}

// FP32 exp / cbrt: evaluated in double and rounded once to float, so the
// result is the correctly rounded float value except in astronomically rare
// near-tie cases.  FP64: plain double functions.
template <class T> inline T exp_t(T x);
template <> inline float exp_t<float>(float x) {
  return (float)std::exp((double)x);
}
template <> inline double exp_t<double>(double x) { return std::exp(x); }

static const double kPiD = 3.14159265358979323846;
static const float kPiF = 3.14159265358979323846f;

// Contact radius (Solver.cu:281, 320-321): cbrtf((3*m)/(4*pi*density)) with
// pi a double constant and density an int kernel parameter.
template <class T> inline T radius_contact_t(T m, float density);
template <> inline float radius_contact_t<float>(float m, float density) {
  const float three_m = 3.0f * m;
  const double denom = 4.0 * kPiD * (double)(int)density;
  const float arg = (float)((double)three_m / denom);
  return (float)std::cbrt((double)arg);
}
template <> inline double radius_contact_t<double>(double m, float density) {
  return std::cbrt((3.0 * m) / (4.0 * kPiD * (double)(int)density));
}

// Surface radius (GranularSystem.cu:518, 536): double PI_d, float density.
inline float radius_surface_f(float m, float density) {
  const float three_m = 3.0f * m;
  const double denom = 4.0 * kPiD * (double)density;
  const float arg = (float)((double)three_m / denom);
  return (float)std::cbrt((double)arg);
}

// Split radius (Solver.cu:799): all-float PI and density.
inline float radius_split_f(float m, float density) {
  const float denom = (4.0f * kPiF) * density;
  const float arg = (3.0f * m) / denom;
  return (float)std::cbrt((double)arg);
}

template <class T> struct SweepParams {
  int own_only;
  T rb, rmin, mu_s, mu_k;
  T sleep, damp_speed, damping, adaptive_damping;
};

template <class T> inline SweepParams<T> make_sweep_params(const oracle_params *p) {
  SweepParams<T> s;
  s.own_only = p->own_only;
  s.rb = (T)p->boundary_radius;
  s.rmin = (T)p->contact_radius_min;
  s.mu_s = (T)p->mu_s;
  s.mu_k = (T)p->mu_k;
  s.sleep = (T)p->sleep_speed;
  s.damp_speed = (T)p->damp_speed;
  s.damping = (T)p->damping;
  s.adaptive_damping = (T)p->adaptive_damping;
  return s;
}

// PBD friction (Macklin et al. 2014, "Unified Particle Physics for
// Real-Time Applications", sec. 6.1): tangential part t of the relative
// displacement rel; static if |t|^2 < (mu_s*delta)^2 (squares, so that the
// static case needs no square root), else t*min(mu_k*delta/|t|,1).
// inv = 1/d of the contact.
template <class T>
inline V3<T> friction(V3<T> p, T inv, T delta, V3<T> rel, const SweepParams<T> &sp) {
  const T nx = p.x * inv, ny = p.y * inv, nz = p.z * inv;
  const T rn = rel.x * nx + rel.y * ny + rel.z * nz;
  V3<T> t;
  t.x = rel.x - rn * nx;
  t.y = rel.y - rn * ny;
  t.z = rel.z - rn * nz;
  const T tl2 = t.x * t.x + t.y * t.y + t.z * t.z;
  const T lim = sp.mu_s * delta;
  if (tl2 < lim * lim) return t;
  const T tl = std::sqrt(tl2);
  T c = (sp.mu_k * delta) / tl;
  c = std::fmin(c, (T)1);
  return V3<T>{t.x * c, t.y * c, t.z * c};
}

}  // namespace orc
