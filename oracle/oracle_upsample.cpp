// CPU oracle of the plain (kernel-weighted) upsampler -- TEST INFRASTRUCTURE
// ONLY.  Restates update_upsampled_cuda (Solver.cu:1244-1369): each fine
// particle blends its own ballistic velocity with the poly6-style weighted
// velocity of the coarse particles in its 27-cell stencil (base floor(x/L)),
// is repelled by boundary particles closer than 0.015, clamped, advected and
// kept above the floor.  FP32, operations in the order written.
#include <cmath>

#include "oracle.h"
#include "oracle_internal.hpp"

using namespace orc;

extern "C" void oracle_upsample_params_default(oracle_upsample_params *u) {
  u->dt = 0.002f;                // Solver.cu:1260
  u->gx = 0.0f;                  // Solver.cu:1259
  u->gy = -9.8f;
  u->gz = 0.0f;
  u->r9 = 493.8271605f;          // Solver.cu:29 (1/h^2, h = 0.045)
  u->boundary_radius = 0.015f;   // Solver.cu:1263
  u->repulsion = 3.0f;           // Solver.cu:1264
  u->min_weight = 1e-6f;         // Solver.cu:1313
  u->near_damping = 0.8f;        // Solver.cu:1325
  u->ground_y = 0.1f;            // Solver.cu:1329
  u->ground_damping = 0.5f;      // Solver.cu:1330
  u->max_velocity = 7.0f;        // Solver.cu:1339
  u->floor_y = 0.04f;            // Solver.cu:1354
  u->floor_friction = 0.95f;     // Solver.cu:1357
}

// One fine particle's update (Solver.cu:1252-1368): position pi, velocity
// vi in; new position and velocity out.  Shared with the learned
// upsampler's teacher (the same update applied to a template point).
namespace orc {
void upsample_one(V3<float> pi, V3<float> vi, const V3<float> *cpos, const V3<float> *cvel,
                  const int *cs, const V3<float> *bpos, const int *bcs, const Grid &g,
                  const oracle_upsample_params *u, V3<float> &np_out, V3<float> &v_out) {
  const V3<float> gdt{u->gx * u->dt, u->gy * u->dt, u->gz * u->dt};
  V3<float> wv{0.0f, 0.0f, 0.0f};
  float gw = 0.0f, maxw = 0.0f;
  V3<float> rep{0.0f, 0.0f, 0.0f};
  bool near = false;
  int bx, by, bz;
  if (floor_cell<float>(pi, g.L, bx, by, bz)) {
    for (int m = 0; m < 27; ++m) {
      const int cell = stencil_cell(bx, by, bz, m, g);
      if (cell < 0) continue;
      for (int j = cs[cell]; j < cs[cell + 1]; ++j) {
        const float dx = pi.x - cpos[j].x, dy = pi.y - cpos[j].y, dz = pi.z - cpos[j].z;
        const float dis = std::sqrt(dx * dx + dy * dy + dz * dz);
        const float t1 = 1.0f - dis * dis * u->r9;
        const float w = std::fmax(0.0f, t1 * t1 * t1);
        gw = gw + w;
        wv.x = wv.x + w * cvel[j].x;
        wv.y = wv.y + w * cvel[j].y;
        wv.z = wv.z + w * cvel[j].z;
        maxw = std::fmax(maxw, w);
      }
      for (int k = bcs[cell]; k < bcs[cell + 1]; ++k) {
        const float tx = pi.x - bpos[k].x, ty = pi.y - bpos[k].y, tz = pi.z - bpos[k].z;
        const float dis = std::sqrt(tx * tx + ty * ty + tz * tz);
        if (dis < u->boundary_radius) {
          near = true;
          const float den = dis + 1e-6f;
          const float force = (u->repulsion * (u->boundary_radius - dis)) / u->boundary_radius;
          rep.x = rep.x + (tx / den) * force;
          rep.y = rep.y + (ty / den) * force;
          rep.z = rep.z + (tz / den) * force;
        }
      }
    }
  }
  V3<float> v;
  if (gw > u->min_weight) {
    wv.x = wv.x / gw;
    wv.y = wv.y / gw;
    wv.z = wv.z / gw;
    const float a = std::fmax(0.0f, 1.0f - maxw);
    v.x = a * (vi.x + gdt.x) + (1.0f - a) * wv.x;
    v.y = a * (vi.y + gdt.y) + (1.0f - a) * wv.y;
    v.z = a * (vi.z + gdt.z) + (1.0f - a) * wv.z;
  } else {
    v.x = vi.x + gdt.x;
    v.y = vi.y + gdt.y;
    v.z = vi.z + gdt.z;
    if (near) {
      v.x = v.x * u->near_damping;
      v.y = v.y * u->near_damping;
      v.z = v.z * u->near_damping;
      if (pi.y < u->ground_y) v.y = v.y * u->ground_damping;
    }
  }
  v.x = v.x + rep.x;
  v.y = v.y + rep.y;
  v.z = v.z + rep.z;
  const float vm = u->max_velocity;
  v.x = std::fmax(-vm, std::fmin(v.x, vm));
  v.y = std::fmax(-vm, std::fmin(v.y, vm));
  v.z = std::fmax(-vm, std::fmin(v.z, vm));
  // fmin/fmax return the non-NaN operand, so a NaN component leaves the
  // clamp as +max_velocity and this test (kept from the reference) is dead
  if (std::isnan(v.x) || std::isnan(v.y) || std::isnan(v.z)) v = V3<float>{0.0f, 0.0f, 0.0f};
  V3<float> np;
  np.x = pi.x + v.x * u->dt;
  np.y = pi.y + v.y * u->dt;
  np.z = pi.z + v.z * u->dt;
  if (np.y < u->floor_y) {
    np.y = u->floor_y;
    v.y = 0.0f;
    v.x = v.x * u->floor_friction;
    v.z = v.z * u->floor_friction;
  }
  if (np.x < 0.0f) {
    np.x = 0.0f;
    v.x = 0.0f;
  }
  if (np.z < 0.0f) {
    np.z = 0.0f;
    v.z = 0.0f;
  }
  np_out = np;
  v_out = v;
}
}  // namespace orc

extern "C" int oracle_upsample_plain(int nf, float *fpos_, float *fvel_, int n,
                                     const float *cpos_, const float *cvel_, const int *cs,
                                     int nb, const float *bpos_, const int *bcs,
                                     const oracle_params *p, const oracle_upsample_params *u) {
  (void)n;
  (void)nb;
  Grid g{p->cell_length, p->cx, p->cy, p->cz};
  V3<float> *fpos = reinterpret_cast<V3<float> *>(fpos_);
  V3<float> *fvel = reinterpret_cast<V3<float> *>(fvel_);
  const V3<float> *cpos = reinterpret_cast<const V3<float> *>(cpos_);
  const V3<float> *cvel = reinterpret_cast<const V3<float> *>(cvel_);
  const V3<float> *bpos = reinterpret_cast<const V3<float> *>(bpos_);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < nf; ++i) {
    V3<float> np, v;
    upsample_one(fpos[i], fvel[i], cpos, cvel, cs, bpos, bcs, g, u, np, v);
    fpos[i] = np;
    fvel[i] = v;
  }
  return 0;
}
