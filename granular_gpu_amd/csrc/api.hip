// Library-level entry points: version, status strings, device check,
// default parameters and workspace sizing / initialisation.
#include <hip/hip_runtime.h>

#include <cstring>

#include "gg_device.hpp"
#include "gg_internal.hpp"

namespace gg {
__global__ void k_fill_u32(uint32_t *__restrict__ p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}
}  // namespace gg

using namespace gg;

extern "C" const char *gg_version(void) { return "granular_gpu_amd 0.1.0 (gfx950)"; }

extern "C" const char *gg_status_string(int s) {
  switch (s) {
    case GG_OK: return "ok";
    case GG_ERR_INVALID: return "invalid argument";
    case GG_ERR_CAPACITY: return "capacity exceeded";
    case GG_ERR_HIP: return "HIP runtime error";
    case GG_ERR_WORKSPACE: return "workspace too small or not initialised";
    case GG_ERR_NO_DEVICE: return "no gfx950 device";
    default: return "unknown status";
  }
}

extern "C" int gg_device_check(void) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return GG_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return GG_ERR_NO_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return GG_ERR_NO_DEVICE;
  return GG_OK;
}

extern "C" void gg_params_default(gg_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->dt = 0.002f;  // scenes/*.json "dt"
  p->gx = 0.0f;
  p->gy = -9.8f;
  p->gz = 0.0f;
  p->density = 238732.4146f;  // main.cu:47: m = 1 <=> r = 0.01
  p->stab_iters = 3;          // Solver.cu:428
  p->solver_iters = 5;        // Solver.cu:99
  p->mu_s = 0.0f;
  p->mu_k = 0.0f;
  p->boundary_radius = 0.01f;     // Solver.cu:283
  p->contact_radius_min = 0.01f;  // Solver.cu:281
  p->sleep_speed = 0.01f;         // Solver.cu:223
  p->damp_speed = 1.0f;           // Solver.cu:202
  p->damping = 0.99f;             // Solver.cu:204-205
  p->adaptive_damping = 0.98f;    // Solver.cu:242
  p->skin = 0.0075f;
  p->flags = 0;
  p->adaptive_rules = 0;  // reference merge / split rules
  p->min_mass = 1.0f;     // Solver.cu:542
  p->split_mass = 2.0f;   // Solver.cu:795
  p->merge_dist = 0.035f; // Solver.cu:572
  p->merge_depth = 3.0f;
  p->split_depth = 2.0f;
}

extern "C" size_t gg_workspace_bytes(int capacity, int boundary_capacity, gg_grid grid,
                                     int max_sweeps, int gmax, int bmax) {
  if (capacity < 0 || boundary_capacity < 0 || max_sweeps < 0 || gmax < 0 || bmax < 0) return 0;
  if (grid.cx <= 0 || grid.cy <= 0 || grid.cz <= 0) return 0;
  return make_layout(capacity, boundary_capacity, grid, max_sweeps, gmax, bmax).total;
}

extern "C" int gg_workspace_init(const gg_workspace *ws, hipStream_t stream) {
  GG_TRY(check_ws(ws));
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                         ws->bmax);
  const size_t words = L.total / 4;
  k_fill_u32<<<1024, 256, 0, stream>>>(static_cast<uint32_t *>(ws->base), words, 0u);
  GG_LAUNCH_CHECK();
  return GG_OK;
}
