// Per-particle update of the plain upsampler (update_upsampled_cuda,
// Solver.cu:1252-1368), shared by the upsampler pass and the learned
// upsampler's teacher.  FP32, operations in the order of the reference.
#pragma once
#include <hip/hip_runtime.h>

#include "gg_device.hpp"
#include "gg_internal.hpp"

namespace gg {

struct UpConst {
  float dt, gdx, gdy, gdz, r9, brad, rep, minw, neard, groundy, groundd, vmax, floory, floorf;
};

__device__ __forceinline__ void coarse_term(float px, float py, float pz, float cx, float cy,
                                            float cz, float vx, float vy, float vz, float r9,
                                            float &gw, float &wx, float &wy, float &wz,
                                            float &maxw) {
  const float dx = px - cx, dy = py - cy, dz = pz - cz;
  const float dis = sqrtf(dx * dx + dy * dy + dz * dz);
  const float t1 = 1.0f - dis * dis * r9;
  const float w = fmaxf(0.0f, t1 * t1 * t1);
  gw = gw + w;
  wx = wx + w * vx;
  wy = wy + w * vy;
  wz = wz + w * vz;
  maxw = fmaxf(maxw, w);
}

// Position (px,py,pz) and velocity (vix,viy,viz) in; new position and
// velocity out.
__device__ __forceinline__ void upsample_one(
    float px, float py, float pz, float vix, float viy, float viz,
    const float *__restrict__ cpos, const float *__restrict__ cvel, const int *__restrict__ cs,
    const float *__restrict__ bpos, const int *__restrict__ bcs, const DevGrid &g,
    const UpConst &u, float &nx_out, float &ny_out, float &nz_out, float &vx_out, float &vy_out,
    float &vz_out) {
  float wx = 0.0f, wy = 0.0f, wz = 0.0f, gw = 0.0f, maxw = 0.0f;
  float rx = 0.0f, ry = 0.0f, rz = 0.0f;
  bool near = false;
  int bx, by, bz;
  if (floor_cell(px, py, pz, g, bx, by, bz)) {
    stencil_rows(bx, by, bz, g, cs, [&](int j0, int j1) {
      int j = j0;
      for (; j + 2 <= j1; j += 2) {
        const Packed3 a = ld3(cpos, j), b = ld3(cpos, j + 1);
        const Packed3 v = ld3(cvel, j), w = ld3(cvel, j + 1);
        coarse_term(px, py, pz, a.x, a.y, a.z, v.x, v.y, v.z, u.r9, gw, wx, wy, wz, maxw);
        coarse_term(px, py, pz, b.x, b.y, b.z, w.x, w.y, w.z, u.r9, gw, wx, wy, wz, maxw);
      }
      for (; j < j1; ++j) {
        const Packed3 a = ld3(cpos, j), v = ld3(cvel, j);
        coarse_term(px, py, pz, a.x, a.y, a.z, v.x, v.y, v.z, u.r9, gw, wx, wy, wz, maxw);
      }
    });
    auto wall = [&](const Packed3 &q) {
      const float tx = px - q.x, ty = py - q.y, tz = pz - q.z;
      const float dis = sqrtf(tx * tx + ty * ty + tz * tz);
      if (dis < u.brad) {
        near = true;
        const float den = dis + 1e-6f;
        const float force = (u.rep * (u.brad - dis)) / u.brad;
        rx = rx + (tx / den) * force;
        ry = ry + (ty / den) * force;
        rz = rz + (tz / den) * force;
      }
    };
    stencil_rows(bx, by, bz, g, bcs, [&](int k0, int k1) {
      int k = k0;
      for (; k + 4 <= k1; k += 4) {
        const Packed3 q0 = ld3(bpos, k), q1 = ld3(bpos, k + 1), q2 = ld3(bpos, k + 2),
                      q3 = ld3(bpos, k + 3);
        wall(q0);
        wall(q1);
        wall(q2);
        wall(q3);
      }
      for (; k < k1; ++k) wall(ld3(bpos, k));
    });
  }
  float vx, vy, vz;
  if (gw > u.minw) {
    wx = wx / gw;
    wy = wy / gw;
    wz = wz / gw;
    const float a = fmaxf(0.0f, 1.0f - maxw);
    vx = a * (vix + u.gdx) + (1.0f - a) * wx;
    vy = a * (viy + u.gdy) + (1.0f - a) * wy;
    vz = a * (viz + u.gdz) + (1.0f - a) * wz;
  } else {
    vx = vix + u.gdx;
    vy = viy + u.gdy;
    vz = viz + u.gdz;
    if (near) {
      vx = vx * u.neard;
      vy = vy * u.neard;
      vz = vz * u.neard;
      if (py < u.groundy) vy = vy * u.groundd;
    }
  }
  vx = vx + rx;
  vy = vy + ry;
  vz = vz + rz;
  vx = fmaxf(-u.vmax, fminf(vx, u.vmax));
  vy = fmaxf(-u.vmax, fminf(vy, u.vmax));
  vz = fmaxf(-u.vmax, fminf(vz, u.vmax));
  // v_min/v_max return the non-NaN operand: a NaN leaves the clamp as
  // +vmax, so (as in the reference) this test never fires
  if (isnan(vx) || isnan(vy) || isnan(vz)) {
    vx = 0.0f;
    vy = 0.0f;
    vz = 0.0f;
  }
  float nx = px + vx * u.dt, ny = py + vy * u.dt, nz = pz + vz * u.dt;
  if (ny < u.floory) {
    ny = u.floory;
    vy = 0.0f;
    vx = vx * u.floorf;
    vz = vz * u.floorf;
  }
  if (nx < 0.0f) {
    nx = 0.0f;
    vx = 0.0f;
  }
  if (nz < 0.0f) {
    nz = 0.0f;
    vz = 0.0f;
  }
  nx_out = nx;
  ny_out = ny;
  nz_out = nz;
  vx_out = vx;
  vy_out = vy;
  vz_out = vz;
}

inline UpConst up_const(const gg_upsample_params *up) {
  UpConst u;
  u.dt = up->dt;
  u.gdx = up->gx * up->dt;
  u.gdy = up->gy * up->dt;
  u.gdz = up->gz * up->dt;
  u.r9 = up->r9;
  u.brad = up->boundary_radius;
  u.rep = up->repulsion;
  u.minw = up->min_weight;
  u.neard = up->near_damping;
  u.groundy = up->ground_y;
  u.groundd = up->ground_damping;
  u.vmax = up->max_velocity;
  u.floory = up->floor_y;
  u.floorf = up->floor_friction;
  return u;
}

}  // namespace gg
