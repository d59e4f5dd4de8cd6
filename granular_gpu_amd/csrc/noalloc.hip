// gg_step_noalloc: the plain step for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  It runs gg_step with its workspace placed in a
// file-scope device array, `scratch`, instead of a caller-owned buffer.
//
// - Size: GG_NOALLOC_SCRATCH_BYTES (32 MiB) per device, held from the
//   library's first use on that device until the process ends.  A request
//   whose workspace (gg_workspace_bytes at the set's n, the boundary's n,
//   the step's sweeps, list widths 48 / 16) does not fit is refused with
//   GG_ERR_CAPACITY before any device work.
// - Host code never passes `scratch` itself (its host-side name is not a
//   device address): on the first call on each device, with that device
//   current, hipGetSymbolAddress gives the device's own copy's address,
//   kept per device and used for every kernel argument and fill.
// - Every call clears the counters and the cell histogram the step expects
//   at zero (fill kernels, no memset nodes, so the call can be captured) and
//   re-packs the boundary positions into the array; everything else the
//   step reads it writes first in the same call.
// - The array keeps the last call's contents and is shared by every call on
//   its device: calls that overlap on one device (two streams, two host
//   threads) race on it and must be ordered by the caller.
#include <hip/hip_runtime.h>

#include <mutex>

#include "gg_internal.hpp"
#include "grid_kernels.hpp"

namespace {

constexpr size_t kScratchBytes = GG_NOALLOC_SCRATCH_BYTES;
constexpr int kListWidth = 48;      // granular candidates per neighbour list
constexpr int kBoundaryWidth = 16;  // boundary candidates per neighbour list
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_lock;
void *g_scratch_addr[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (looked up once).
int scratch_address(void **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return GG_ERR_HIP;
  if (dev < 0 || dev >= kMaxDevices) return GG_ERR_NO_DEVICE;
  std::lock_guard<std::mutex> guard(g_lock);
  if (!g_scratch_addr[dev]) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return GG_ERR_HIP;
    g_scratch_addr[dev] = ptr;
  }
  *out = g_scratch_addr[dev];
  return GG_OK;
}

}  // namespace

using namespace gg;

extern "C" int gg_step_noalloc(gg_particles *set, gg_adaptive *ad, int *cell_start,
                               const gg_particles *boundary, const int *bcell_start, gg_grid grid,
                               const gg_params *params, hipStream_t stream) {
  if (!set || !boundary || !params || !cell_start || !bcell_start) return GG_ERR_INVALID;
  if (set->n < 0 || boundary->n < 0 || params->stab_iters < 0 || params->solver_iters < 0)
    return GG_ERR_INVALID;
  if (grid.cx <= 0 || grid.cy <= 0 || grid.cz <= 0 || !(grid.cell_length > 0.0f))
    return GG_ERR_INVALID;
  if (boundary->n > 0 && !boundary->pos) return GG_ERR_INVALID;
  gg_workspace ws{};
  ws.capacity = set->n > 0 ? set->n : 1;
  ws.boundary_capacity = boundary->n > 0 ? boundary->n : 1;
  ws.grid = grid;
  ws.max_sweeps = params->stab_iters + params->solver_iters;
  ws.gmax = kListWidth;
  ws.bmax = kBoundaryWidth;
  const size_t need = gg_workspace_bytes(ws.capacity, ws.boundary_capacity, grid, ws.max_sweeps,
                                         ws.gmax, ws.bmax);
  if (need == 0) return GG_ERR_INVALID;
  if (need > kScratchBytes) return GG_ERR_CAPACITY;
  void *base = nullptr;
  GG_TRY(scratch_address(&base));
  ws.base = base;
  ws.bytes = kScratchBytes;
  GG_TRY(check_ws(&ws));
  // zero what the step expects at zero: the cell histogram and the counters
  const Layout L = make_layout(ws.capacity, ws.boundary_capacity, grid, ws.max_sweeps, ws.gmax,
                               ws.bmax);
  GG_TRY(gg_fill_u32(at<void>(&ws, L.count), (size_t)grid_cells(grid) + 1, 0u, stream));
  GG_TRY(gg_fill_u32(at<void>(&ws, L.ctr), 64, 0u, stream));
  GG_TRY(gg_fill_u32(at<void>(&ws, L.ctr2), 64, 0u, stream));
  GG_TRY(gg_fill_u32(at<void>(&ws, L.fbqc), 2, 0u, stream));
  GG_TRY(pack_boundary(boundary, &ws, stream));
  return gg_step(set, ad, cell_start, boundary, bcell_start, params, &ws, stream);
}
