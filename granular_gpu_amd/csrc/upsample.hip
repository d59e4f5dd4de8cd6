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
#include "upsample_dev.hpp"

namespace gg {

__global__ void __launch_bounds__(kBlock)
    k_upsample(int nf, float *__restrict__ fpos, float *__restrict__ fvel,
               const float *__restrict__ cpos, const float *__restrict__ cvel,
               const int *__restrict__ cs, const float *__restrict__ bpos,
               const int *__restrict__ bcs, DevGrid g, UpConst u, int remap) {
  const int i = xcd_block(blockIdx.x, gridDim.x, remap) * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  float nx, ny, nz, vx, vy, vz;
  upsample_one(fpos[3 * i], fpos[3 * i + 1], fpos[3 * i + 2], fvel[3 * i], fvel[3 * i + 1],
               fvel[3 * i + 2], cpos, cvel, cs, bpos, bcs, g, u, nx, ny, nz, vx, vy, vz);
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
  const UpConst u = up_const(up);
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
