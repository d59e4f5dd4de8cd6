// The reference Solver's separate per-particle stages, as entry points of
// their own.  gg_step and gg_predict run them fused into one pass
// (k_permute_predict); these exist for callers that drive the stages one
// by one, as the reference's public Solver methods do.
//
// Reference: Solver::add_external_force (Solver.cu:125-131),
// update_particle_positions / predict_position_functor (:133-160),
// apply_mass_scaling / mass_scaling_functor (:162-191),
// update_neighborhood (:112-123).
#include <hip/hip_runtime.h>

#include "gg_device.hpp"
#include "gg_internal.hpp"
#include "grid_kernels.hpp"

namespace gg {

// v += dv per particle (dv = dt * G formed by the caller, as the reference
// forms it on the host).
__global__ void __launch_bounds__(kBlock) k_add_velocity(float *__restrict__ vel, int n, float dvx,
                                                         float dvy, float dvz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Packed3 v = ld3(vel, i);
  st3(vel, i, v.x + dvx, v.y + dvy, v.z + dvz);
}

// out = x + dt * v per component (no contraction: -ffp-contract=off).
__global__ void __launch_bounds__(kBlock) k_predict_positions(const float *__restrict__ pos,
                                                              const float *__restrict__ vel, int n,
                                                              float dt, float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Packed3 p = ld3(pos, i), v = ld3(vel, i);
  st3(out, i, p.x + dt * v.x, p.y + dt * v.y, p.z + dt * v.z);
}

// m_s = m * exp(-y) of the current position, exp in double rounded once.
__global__ void __launch_bounds__(kBlock) k_mass_scaling(const float *__restrict__ pos,
                                                         const float *__restrict__ mass, int n,
                                                         float *__restrict__ ms) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ms[i] = mass[i] * exp_f(-pos[3 * (size_t)i + 1]);
}

static inline int blocks_of(int n) { return (n + kBlock - 1) / kBlock; }

}  // namespace gg

using namespace gg;

extern "C" int gg_add_external_force(gg_particles *set, gg_float3 dv, hipStream_t stream) {
  if (!set || set->n < 0) return GG_ERR_INVALID;
  if (set->n == 0) return GG_OK;
  if (!set->vel) return GG_ERR_INVALID;
  k_add_velocity<<<blocks_of(set->n), kBlock, 0, stream>>>(reinterpret_cast<float *>(set->vel),
                                                           set->n, dv.x, dv.y, dv.z);
  GG_LAUNCH_CHECK();
  return GG_OK;
}

extern "C" int gg_predict_positions(const gg_particles *set, float dt, gg_float3 *out,
                                    hipStream_t stream) {
  if (!set || set->n < 0) return GG_ERR_INVALID;
  if (set->n == 0) return GG_OK;
  if (!set->pos || !set->vel || !out) return GG_ERR_INVALID;
  k_predict_positions<<<blocks_of(set->n), kBlock, 0, stream>>>(
      reinterpret_cast<const float *>(set->pos), reinterpret_cast<const float *>(set->vel),
      set->n, dt, reinterpret_cast<float *>(out));
  GG_LAUNCH_CHECK();
  return GG_OK;
}

extern "C" int gg_mass_scaling(gg_particles *set, hipStream_t stream) {
  if (!set || set->n < 0) return GG_ERR_INVALID;
  if (set->n == 0) return GG_OK;
  if (!set->pos || !set->mass || !set->scaled_mass) return GG_ERR_INVALID;
  k_mass_scaling<<<blocks_of(set->n), kBlock, 0, stream>>>(
      reinterpret_cast<const float *>(set->pos), set->mass, set->n, set->scaled_mass);
  GG_LAUNCH_CHECK();
  return GG_OK;
}

extern "C" int gg_sort_positions_by_key(gg_particles *set, int *cell_start,
                                        const gg_workspace *ws, hipStream_t stream) {
  GG_TRY(check_ws(ws));
  if (!set || !cell_start || set->n < 0) return GG_ERR_INVALID;
  if (set->n > ws->capacity) return GG_ERR_CAPACITY;
  if (set->n == 0) return GG_OK;
  if (!set->pos || !set->particle_2_cell) return GG_ERR_INVALID;
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                         ws->bmax);
  GG_TRY(sort_by_cell_keys(nullptr, set->particle_2_cell, set->n, cell_start, ws, stream));
  return gather_inplace(set->pos, 3, at<int>(ws, L.perm), set->n, at<void>(ws, L.v3tmp0), stream);
}
