// Scenes sharing one set: cell keys over the concatenated cell tables, the
// boundary build, and the captured multi-scene step with its block-size
// tuner.  gg_step_scenes itself (the pipeline of gg_step with per-scene
// grids) is in step.hip.
//
// No reference counterpart: the specification's batched configuration
// (many independent scenes advanced by one launch per stage).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "gg_device.hpp"
#include "gg_internal.hpp"
#include "grid_kernels.hpp"

namespace gg {

static inline int blocks_for(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// key = cell_base of the particle's scene + its cell in the scene's grid
// (the scene's own sentinel outside it).
__global__ void __launch_bounds__(kBlock)
    k_scene_keys(const float *__restrict__ pos, int n, const int *__restrict__ scene_of,
                 const gg_scene_desc *__restrict__ desc, int count, float L,
                 int *__restrict__ keys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int k = scene_of[i];
  k = k < 0 ? 0 : (k < count ? k : count - 1);
  const gg_scene_desc d = desc[k];
  const Packed3 p = ld3(pos, i);
  keys[i] = d.cell_base + hash_cell(p.x, p.y, p.z, DevGrid{L, d.cx, d.cy, d.cz});
}

int check_scenes(const gg_scenes *sc, const gg_workspace *ws) {
  if (!sc || sc->count < 1 || !sc->desc || !sc->host_desc || !sc->scene_of)
    return GG_ERR_INVALID;
  long long next = 0;
  for (int s = 0; s < sc->count; ++s) {
    const gg_scene_desc &d = sc->host_desc[s];
    if (d.cx <= 0 || d.cy <= 0 || d.cz <= 0 || d.cell_base != next) return GG_ERR_INVALID;
    if (!(d.mu_s >= 0.0f) || !(d.mu_k >= 0.0f)) return GG_ERR_INVALID;
    next += (long long)d.cx * d.cy * d.cz + 1;
  }
  // the workspace's grid holds the concatenated tables: T - 1 cells + sentinel
  if (next != grid_cells(ws->grid) + 1) return GG_ERR_INVALID;
  return GG_OK;
}

int scene_keys(const float *pos, int n, const int *scene_of, const gg_scenes *sc, float L,
               int *keys, hipStream_t st) {
  if (n <= 0) return GG_OK;
  k_scene_keys<<<blocks_for(n), kBlock, 0, st>>>(pos, n, scene_of, sc->desc, sc->count, L, keys);
  GG_LAUNCH_CHECK();
  return GG_OK;
}

}  // namespace gg

using namespace gg;

extern "C" int gg_scenes_boundary_build(gg_particles *boundary, int *bcell_start,
                                        const gg_scenes *scenes, const gg_workspace *ws,
                                        hipStream_t stream) {
  if (!boundary || !bcell_start || !ws) return GG_ERR_INVALID;
  GG_TRY(check_ws(ws));
  GG_TRY(check_scenes(scenes, ws));
  if (boundary->n < 0 || boundary->n > ws->boundary_capacity) return GG_ERR_INVALID;
  if (boundary->n > 0 && (!boundary->pos || !scenes->bscene_of)) return GG_ERR_INVALID;
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid, ws->max_sweeps, ws->gmax,
                         ws->bmax);
  int *keys = at<int>(ws, L.itmp2);
  GG_TRY(scene_keys(reinterpret_cast<const float *>(boundary->pos), boundary->n,
                    scenes->bscene_of, scenes, ws->grid.cell_length, keys, stream));
  GG_TRY(grid_build_general(boundary, nullptr, bcell_start, nullptr, ws, stream,
                            ws->boundary_capacity, keys));
  return pack_boundary(boundary, ws, stream);
}

// ---- captured multi-scene step ------------------------------------------------
struct gg_scenes_graph {
  gg_particles set, boundary;
  int *cell_start = nullptr;
  const int *bcell_start = nullptr;
  gg_scenes scenes{};
  std::vector<gg_scene_desc> host_desc;
  gg_params params{};
  gg_workspace ws{};
  int steps = 1;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void release(gg_scenes_graph *g) {
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->exec = nullptr;
  g->graph = nullptr;
}

int capture(gg_scenes_graph *g) {
  release(g);
  if (!g->cap) GG_TRY(hip_status(hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking)));
  GG_TRY(hip_status(hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal)));
  int rc = GG_OK;
  for (int t = 0; t < g->steps && rc == GG_OK; ++t)
    rc = gg_step_scenes(&g->set, g->cell_start, &g->boundary, g->bcell_start, &g->scenes,
                        &g->params, &g->ws, g->cap);
  hipGraph_t graph = nullptr;
  const int rc_end = hip_status(hipStreamEndCapture(g->cap, &graph));
  if (rc != GG_OK || rc_end != GG_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc != GG_OK ? rc : rc_end;
  }
  g->graph = graph;
  return hip_status(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
}

bool block_ok(int b) { return b == 0 || b == 64 || b == 128 || b == 256 || b == 512; }

}  // namespace

extern "C" int gg_scenes_graph_create(gg_particles *set, int *cell_start,
                                      const gg_particles *boundary, const int *bcell_start,
                                      const gg_scenes *scenes, const gg_params *params,
                                      const gg_workspace *ws, int steps_per_launch,
                                      gg_scenes_graph **out) {
  if (!out) return GG_ERR_INVALID;
  *out = nullptr;
  if (!set || !cell_start || !boundary || !bcell_start || !scenes || !params || !ws)
    return GG_ERR_INVALID;
  if (steps_per_launch < 1 || steps_per_launch > 1000) return GG_ERR_INVALID;
  GG_TRY(check_ws(ws));
  GG_TRY(check_scenes(scenes, ws));
  gg_scenes_graph *g = new (std::nothrow) gg_scenes_graph;
  if (!g) return GG_ERR_HIP;
  g->set = *set;
  g->boundary = *boundary;
  g->cell_start = cell_start;
  g->bcell_start = bcell_start;
  g->host_desc.assign(scenes->host_desc, scenes->host_desc + scenes->count);
  g->scenes = *scenes;
  g->scenes.host_desc = g->host_desc.data();
  g->params = *params;
  g->ws = *ws;
  g->steps = steps_per_launch;
  const int rc = capture(g);
  if (rc != GG_OK) {
    gg_scenes_graph_destroy(g);
    return rc;
  }
  *out = g;
  return GG_OK;
}

extern "C" int gg_scenes_graph_launch(gg_scenes_graph *g, hipStream_t stream) {
  if (!g || !g->exec) return GG_ERR_INVALID;
  return hip_status(hipGraphLaunch(g->exec, stream));
}

extern "C" int gg_scenes_graph_autotune(gg_scenes_graph *g, int trials, hipStream_t stream,
                                        int *sweep_block, int *nlist_block, float *best_ms) {
  if (!g || trials < 1 || trials > 100) return GG_ERR_INVALID;
  const int n = g->set.n;
  // pos, vel, mass (and adaptive_last_step when present) saved and restored
  // around every candidate, so tuning leaves the scenes as it found them
  void *snap = nullptr;
  const size_t bytes = (size_t)std::max(n, 1) * 32;
  GG_TRY(hip_status(hipMalloc(&snap, bytes)));
  auto copy_state = [&](bool restore) -> int {
    char *b = static_cast<char *>(snap);
    void *part[4] = {g->set.pos, g->set.vel, g->set.mass, g->set.adaptive_last_step};
    const size_t sz[4] = {12 * (size_t)n, 12 * (size_t)n, 4 * (size_t)n, 4 * (size_t)n};
    size_t o = 0;
    for (int a = 0; a < 4; ++a) {
      if (part[a] && n > 0) {
        void *dst = restore ? part[a] : b + o;
        const void *src = restore ? (const void *)(b + o) : part[a];
        GG_TRY(hip_status(hipMemcpyAsync(dst, src, sz[a], hipMemcpyDeviceToDevice, stream)));
      }
      o += sz[a];
    }
    return GG_OK;
  };
  hipEvent_t t0 = nullptr, t1 = nullptr;
  int rc = copy_state(false);
  if (rc == GG_OK) rc = hip_status(hipEventCreate(&t0));
  if (rc == GG_OK) rc = hip_status(hipEventCreate(&t1));
  const gg_params keep = g->params;
  gg_params best = keep;
  float best_t = -1.0f;
  const int sw_cand[] = {64, 128, 256, 512};
  const int nl_cand[] = {128, 256};
  for (int sw : sw_cand) {
    for (int nl : nl_cand) {
      if (rc != GG_OK) break;
      g->params.sweep_block = sw;
      g->params.nlist_block = nl;
      rc = capture(g);
      if (rc == GG_OK) rc = hip_status(hipGraphLaunch(g->exec, stream));  // warm-up
      if (rc == GG_OK) rc = hip_status(hipEventRecord(t0, stream));
      for (int t = 0; t < trials && rc == GG_OK; ++t)
        rc = hip_status(hipGraphLaunch(g->exec, stream));
      if (rc == GG_OK) rc = hip_status(hipEventRecord(t1, stream));
      if (rc == GG_OK) rc = hip_status(hipEventSynchronize(t1));
      float ms = 0.0f;
      if (rc == GG_OK) rc = hip_status(hipEventElapsedTime(&ms, t0, t1));
      if (rc == GG_OK) rc = copy_state(true);
      if (rc != GG_OK) break;
      ms /= (float)trials;
      if (best_t < 0.0f || ms < best_t) {
        best_t = ms;
        best = g->params;
      }
    }
  }
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  const int rc_sync = hip_status(hipStreamSynchronize(stream));
  (void)hipFree(snap);
  g->params = rc == GG_OK ? best : keep;
  const int rc_cap = capture(g);
  if (rc != GG_OK) return rc;
  if (rc_sync != GG_OK) return rc_sync;
  if (rc_cap != GG_OK) return rc_cap;
  if (!block_ok(g->params.sweep_block) || !block_ok(g->params.nlist_block)) return GG_ERR_INVALID;
  if (sweep_block) *sweep_block = g->params.sweep_block;
  if (nlist_block) *nlist_block = g->params.nlist_block;
  if (best_ms) *best_ms = best_t;
  return GG_OK;
}

extern "C" void gg_scenes_graph_destroy(gg_scenes_graph *g) {
  if (!g) return;
  release(g);
  if (g->cap) (void)hipStreamDestroy(g->cap);
  delete g;
}
