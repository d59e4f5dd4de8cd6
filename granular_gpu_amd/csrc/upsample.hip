// Plain upsampler: fine particles advected by the kernel-weighted velocity
// of the coarse particles around them, and the fine-particle cell sort.
//
// Reference: update_upsampled_cuda (Solver.cu:1244-1369) and
// neighbor_search_upsampled (GranularSystem.cu:282-395).  The reference
// sorts the fine positions and velocities with two separate stable sorts
// keyed by the same unsorted keys; here one stable sort and one gather give
// the same permutation.  particle_2_cell keeps the pre-sort keys and the
// fine cell table is built but not used by the upsampler, as in the
// reference.
#include <hip/hip_runtime.h>

#include "gg_device.hpp"
#include "gg_internal.hpp"
#include "grid_kernels.hpp"
#include "profiler.hpp"

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

__global__ void __launch_bounds__(kBlock)
    k_upsample(int nf, float *__restrict__ fpos, float *__restrict__ fvel,
               const float *__restrict__ cpos, const float *__restrict__ cvel,
               const int *__restrict__ cs, const float *__restrict__ bpos,
               const int *__restrict__ bcs, DevGrid g, UpConst u, int remap) {
  const int i = xcd_block(blockIdx.x, gridDim.x, remap) * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  const float px = fpos[3 * i], py = fpos[3 * i + 1], pz = fpos[3 * i + 2];
  float wx = 0.0f, wy = 0.0f, wz = 0.0f, gw = 0.0f, maxw = 0.0f;
  float rx = 0.0f, ry = 0.0f, rz = 0.0f;
  bool near = false;
  int bx, by, bz;
  if (floor_cell(px, py, pz, g, bx, by, bz)) {
    stencil_rows(bx, by, bz, g, cs, [&](int j0, int j1) {
      int j = j0;
      for (; j + 2 <= j1; j += 2) {
        const float a0 = cpos[3 * j], a1 = cpos[3 * j + 1], a2 = cpos[3 * j + 2];
        const float b0 = cpos[3 * j + 3], b1 = cpos[3 * j + 4], b2 = cpos[3 * j + 5];
        const float v0 = cvel[3 * j], v1 = cvel[3 * j + 1], v2 = cvel[3 * j + 2];
        const float w0 = cvel[3 * j + 3], w1 = cvel[3 * j + 4], w2 = cvel[3 * j + 5];
        coarse_term(px, py, pz, a0, a1, a2, v0, v1, v2, u.r9, gw, wx, wy, wz, maxw);
        coarse_term(px, py, pz, b0, b1, b2, w0, w1, w2, u.r9, gw, wx, wy, wz, maxw);
      }
      for (; j < j1; ++j)
        coarse_term(px, py, pz, cpos[3 * j], cpos[3 * j + 1], cpos[3 * j + 2], cvel[3 * j],
                    cvel[3 * j + 1], cvel[3 * j + 2], u.r9, gw, wx, wy, wz, maxw);
    });
    stencil_rows(bx, by, bz, g, bcs, [&](int k0, int k1) {
      for (int k = k0; k < k1; ++k) {
        const float tx = px - bpos[3 * k], ty = py - bpos[3 * k + 1], tz = pz - bpos[3 * k + 2];
        const float dis = sqrtf(tx * tx + ty * ty + tz * tz);
        if (dis < u.brad) {
          near = true;
          const float den = dis + 1e-6f;
          const float force = (u.rep * (u.brad - dis)) / u.brad;
          rx = rx + (tx / den) * force;
          ry = ry + (ty / den) * force;
          rz = rz + (tz / den) * force;
        }
      }
    });
  }
  const float vix = fvel[3 * i], viy = fvel[3 * i + 1], viz = fvel[3 * i + 2];
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
  fpos[3 * i] = nx;
  fpos[3 * i + 1] = ny;
  fpos[3 * i + 2] = nz;
  fvel[3 * i] = vx;
  fvel[3 * i + 1] = vy;
  fvel[3 * i + 2] = vz;
}

__global__ void __launch_bounds__(kBlock) k_copy_int(const int *__restrict__ src,
                                                     int *__restrict__ dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

static inline int blocks_for(long long n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace gg

using namespace gg;

extern "C" void gg_upsample_params_default(gg_upsample_params *u) {
  if (!u) return;
  u->dt = 0.002f;               // Solver.cu:1260
  u->gx = 0.0f;                 // Solver.cu:1259
  u->gy = -9.8f;
  u->gz = 0.0f;
  u->r9 = 493.8271605f;         // Solver.cu:29
  u->boundary_radius = 0.015f;  // Solver.cu:1263
  u->repulsion = 3.0f;          // Solver.cu:1264
  u->min_weight = 1e-6f;        // Solver.cu:1313
  u->near_damping = 0.8f;       // Solver.cu:1325
  u->ground_y = 0.1f;           // Solver.cu:1329
  u->ground_damping = 0.5f;     // Solver.cu:1330
  u->max_velocity = 7.0f;       // Solver.cu:1339
  u->floor_y = 0.04f;           // Solver.cu:1354
  u->floor_friction = 0.95f;    // Solver.cu:1357
}

extern "C" int gg_upsampled_sort(gg_particles *fine, int *fine_cell_start,
                                 const gg_workspace *fine_ws, hipStream_t stream) {
  if (!fine || !fine_cell_start || !fine_ws) return GG_ERR_INVALID;
  GG_TRY(check_ws(fine_ws));
  if (fine->n < 0) return GG_ERR_INVALID;
  if (fine->n > fine_ws->capacity) return GG_ERR_CAPACITY;
  prof_begin(GG_STAGE_UPSAMPLE_SORT, stream);
  // keep the pre-sort keys for particle_2_cell (GranularSystem.cu:338-345)
  int *p2c = fine->particle_2_cell;
  gg_particles tmp = *fine;
  tmp.particle_2_cell = nullptr;
  tmp.mass = nullptr;
  tmp.is_animated = nullptr;
  GG_TRY(grid_build_general(&tmp, nullptr, fine_cell_start, nullptr, fine_ws, stream,
                            fine_ws->capacity));
  if (p2c && fine->n > 0) {
    Layout L = make_layout(fine_ws->capacity, fine_ws->boundary_capacity, fine_ws->grid,
                           fine_ws->max_sweeps, fine_ws->gmax, fine_ws->bmax);
    k_copy_int<<<blocks_for(fine->n), kBlock, 0, stream>>>(at<int>(fine_ws, L.key), p2c,
                                                           fine->n);
    GG_LAUNCH_CHECK();
  }
  prof_end(GG_STAGE_UPSAMPLE_SORT, stream);
  return GG_OK;
}

extern "C" int gg_upsample_plain(gg_particles *fine, const gg_particles *coarse,
                                 const int *cell_start, const gg_particles *boundary,
                                 const int *bcell_start, gg_grid grid,
                                 const gg_upsample_params *up, int flags, hipStream_t stream) {
  if (!fine || !coarse || !cell_start || !boundary || !bcell_start || !up) return GG_ERR_INVALID;
  if (fine->n < 0 || coarse->n < 0 || boundary->n < 0) return GG_ERR_INVALID;
  if (grid.cx <= 0 || grid.cy <= 0 || grid.cz <= 0 || !(grid.cell_length > 0.0f))
    return GG_ERR_INVALID;
  if (fine->n == 0) return GG_OK;
  if (!fine->pos || !fine->vel || (coarse->n > 0 && (!coarse->pos || !coarse->vel)) ||
      (boundary->n > 0 && !boundary->pos))
    return GG_ERR_INVALID;
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
  prof_begin(GG_STAGE_UPSAMPLE, stream);
  k_upsample<<<blocks_for(fine->n), kBlock, 0, stream>>>(
      fine->n, reinterpret_cast<float *>(fine->pos), reinterpret_cast<float *>(fine->vel),
      reinterpret_cast<const float *>(coarse->pos), reinterpret_cast<const float *>(coarse->vel),
      cell_start, reinterpret_cast<const float *>(boundary->pos), bcell_start, dev_grid(grid), u,
      (flags & GG_FLAG_NO_XCD_REMAP) ? 0 : 1);
  GG_LAUNCH_CHECK();
  prof_end(GG_STAGE_UPSAMPLE, stream);
  return GG_OK;
}
