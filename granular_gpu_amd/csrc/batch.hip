// Batched independent scenes: many plain scenes of uneven size advanced by
// one hipGraph launch, and the autotuner that picks the graph's shape and
// the kernels' block sizes.
//
// The graph holds `steps_per_launch` steps of every scene.  Scenes are
// spread over `branches` parallel branches (streams forked from and joined
// back into the capture stream) by longest-first greedy packing on particle
// count, so large and small scenes overlap on the GPU; inside a branch the
// order is step-major.  Block size and branch count do not change results:
// every kernel computes each particle independently of the launch shape.
//
// No reference counterpart (the reference steps one scene at a time with
// host syncs between launches); this is the batched path of the
// specification's batched-scenes configuration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "gg_internal.hpp"

struct gg_batch {
  std::vector<gg_batch_scene> sc;
  gg_batch_config cfg{};
  std::vector<std::vector<int>> branch;  // scene indices per branch
  std::vector<hipStream_t> streams;
  hipStream_t cap = nullptr;
  std::vector<hipEvent_t> ev;  // [0] fork, [1 + b] join of branch b
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace gg {
namespace {

bool block_valid(int b) { return b == 0 || b == 64 || b == 128 || b == 256 || b == 512; }

void release_graph(gg_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

void release_streams(gg_batch *b) {
  for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
  b->streams.clear();
  for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
  b->ev.clear();
  if (b->cap) (void)hipStreamDestroy(b->cap);
  b->cap = nullptr;
}

// Longest-first greedy packing of scenes onto branches.
void pack(gg_batch *b) {
  const int nb = b->cfg.branches;
  b->branch.assign(nb, {});
  std::vector<int> order(b->sc.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return b->sc[x].set.n > b->sc[y].set.n; });
  std::vector<long long> load(nb, 0);
  for (int s : order) {
    const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    b->branch[k].push_back(s);
    load[k] += b->sc[s].set.n;
  }
  for (auto &v : b->branch) std::sort(v.begin(), v.end());
}

int scene_params(const gg_batch *b, int s, gg_params *p) {
  *p = b->sc[s].params;
  if (b->cfg.sweep_block) p->sweep_block = b->cfg.sweep_block;
  if (b->cfg.nlist_block) p->nlist_block = b->cfg.nlist_block;
  return GG_OK;
}

// Enqueues one step of every scene of branch k on stream st.
int enqueue_branch(gg_batch *b, int k, hipStream_t st) {
  for (int s : b->branch[k]) {
    gg_batch_scene &e = b->sc[s];
    gg_params p;
    scene_params(b, s, &p);
    GG_TRY(gg_step(&e.set, nullptr, e.cell_start, &e.boundary, e.bcell_start, &p, &e.ws, st));
  }
  return GG_OK;
}

int build_graph(gg_batch *b) {
  release_graph(b);
  release_streams(b);
  pack(b);
  const int nb = b->cfg.branches;
  GG_TRY(hip_status(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking)));
  b->streams.resize(nb, nullptr);
  for (int k = 0; k < nb; ++k)
    GG_TRY(hip_status(hipStreamCreateWithFlags(&b->streams[k], hipStreamNonBlocking)));
  b->ev.resize(nb + 1, nullptr);
  for (auto &e : b->ev) GG_TRY(hip_status(hipEventCreateWithFlags(&e, hipEventDisableTiming)));
  GG_TRY(hip_status(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal)));
  int rc = hip_status(hipEventRecord(b->ev[0], b->cap));
  for (int k = 0; k < nb && rc == GG_OK; ++k)
    rc = hip_status(hipStreamWaitEvent(b->streams[k], b->ev[0], 0));
  for (int t = 0; t < b->cfg.steps_per_launch && rc == GG_OK; ++t)
    for (int k = 0; k < nb && rc == GG_OK; ++k) rc = enqueue_branch(b, k, b->streams[k]);
  for (int k = 0; k < nb && rc == GG_OK; ++k) {
    rc = hip_status(hipEventRecord(b->ev[1 + k], b->streams[k]));
    if (rc == GG_OK) rc = hip_status(hipStreamWaitEvent(b->cap, b->ev[1 + k], 0));
  }
  hipGraph_t g = nullptr;
  const int rc_end = hip_status(hipStreamEndCapture(b->cap, &g));
  if (rc != GG_OK || rc_end != GG_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc != GG_OK ? rc : rc_end;
  }
  b->graph = g;
  GG_TRY(hip_status(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0)));
  return GG_OK;
}

int check_config(const gg_batch_config *c, int nscenes) {
  if (!c) return GG_ERR_INVALID;
  if (c->branches < 1 || c->branches > 16 || c->branches > std::max(nscenes, 1))
    return GG_ERR_INVALID;
  if (!block_valid(c->sweep_block) || !block_valid(c->nlist_block)) return GG_ERR_INVALID;
  if (c->steps_per_launch < 1 || c->steps_per_launch > 1000) return GG_ERR_INVALID;
  return GG_OK;
}

// Device copies of every scene's pos, vel, mass and adaptive_last_step, so
// that tuning runs leave the scenes as they found them.
struct Snapshot {
  void *mem = nullptr;
  std::vector<size_t> off;
  ~Snapshot() {
    if (mem) (void)hipFree(mem);
  }
};

int snapshot(gg_batch *b, Snapshot &sn, bool restore, hipStream_t st) {
  if (!restore) {
    size_t total = 0;
    sn.off.clear();
    for (auto &e : b->sc) {
      sn.off.push_back(total);
      total += (size_t)e.set.n * 32 + 256;
    }
    GG_TRY(hip_status(hipMalloc(&sn.mem, std::max<size_t>(total, 256))));
  }
  for (size_t s = 0; s < b->sc.size(); ++s) {
    const gg_particles &p = b->sc[s].set;
    const size_t n = (size_t)p.n;
    if (!n) continue;
    char *base = static_cast<char *>(sn.mem) + sn.off[s];
    void *part[4] = {p.pos, p.vel, p.mass, p.adaptive_last_step};
    const size_t bytes[4] = {12 * n, 12 * n, 4 * n, 4 * n};
    size_t o = 0;
    for (int a = 0; a < 4; ++a) {
      if (part[a]) {
        void *dst = restore ? part[a] : base + o;
        const void *src = restore ? (const void *)(base + o) : part[a];
        GG_TRY(hip_status(hipMemcpyAsync(dst, src, bytes[a], hipMemcpyDeviceToDevice, st)));
      }
      o += bytes[a];
    }
  }
  return GG_OK;
}

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_batch_create(const gg_batch_scene *scenes, int nscenes,
                               const gg_batch_config *cfg, gg_batch **out) {
  if (!out || nscenes < 1 || !scenes) return GG_ERR_INVALID;
  *out = nullptr;
  gg_batch_config c;
  if (cfg) {
    c = *cfg;
  } else {
    c.branches = std::min(nscenes, 4);
    c.sweep_block = 0;
    c.nlist_block = 0;
    c.steps_per_launch = 1;
  }
  GG_TRY(check_config(&c, nscenes));
  for (int s = 0; s < nscenes; ++s) {
    const gg_batch_scene &e = scenes[s];
    if (!e.cell_start || !e.bcell_start || e.set.n < 0 || e.set.n > e.ws.capacity ||
        e.boundary.n > e.ws.boundary_capacity || !e.ws.base)
      return GG_ERR_INVALID;
    if (e.params.stab_iters + e.params.solver_iters > e.ws.max_sweeps) return GG_ERR_CAPACITY;
  }
  gg_batch *b = new (std::nothrow) gg_batch;
  if (!b) return GG_ERR_HIP;
  b->sc.assign(scenes, scenes + nscenes);
  b->cfg = c;
  const int rc = build_graph(b);
  if (rc != GG_OK) {
    release_graph(b);
    release_streams(b);
    delete b;
    return rc;
  }
  *out = b;
  return GG_OK;
}

extern "C" int gg_batch_step(gg_batch *b, hipStream_t stream) {
  if (!b || !b->exec) return GG_ERR_INVALID;
  return hip_status(hipGraphLaunch(b->exec, stream));
}

extern "C" int gg_batch_step_eager(gg_batch *b, hipStream_t stream) {
  if (!b) return GG_ERR_INVALID;
  for (int t = 0; t < b->cfg.steps_per_launch; ++t)
    for (size_t s = 0; s < b->sc.size(); ++s) {
      gg_batch_scene &e = b->sc[s];
      gg_params p;
      scene_params(b, (int)s, &p);
      GG_TRY(gg_step(&e.set, nullptr, e.cell_start, &e.boundary, e.bcell_start, &p, &e.ws,
                     stream));
    }
  return GG_OK;
}

extern "C" int gg_batch_get_config(const gg_batch *b, gg_batch_config *cfg) {
  if (!b || !cfg) return GG_ERR_INVALID;
  *cfg = b->cfg;
  return GG_OK;
}

extern "C" int gg_batch_set_config(gg_batch *b, const gg_batch_config *cfg) {
  if (!b) return GG_ERR_INVALID;
  GG_TRY(check_config(cfg, (int)b->sc.size()));
  b->cfg = *cfg;
  return build_graph(b);
}

extern "C" int gg_batch_autotune(gg_batch *b, int trials, hipStream_t stream,
                                 gg_batch_config *best_out, float *best_ms_out) {
  if (!b || trials < 1 || trials > 100) return GG_ERR_INVALID;
  const gg_batch_config keep = b->cfg;
  Snapshot sn;
  GG_TRY(snapshot(b, sn, false, stream));
  hipEvent_t t0 = nullptr, t1 = nullptr;
  GG_TRY(hip_status(hipEventCreate(&t0)));
  if (hipEventCreate(&t1) != hipSuccess) {
    (void)hipEventDestroy(t0);
    return GG_ERR_HIP;
  }
  const int nsc = (int)b->sc.size();
  const int br_cand[] = {1, 2, 4, 8};
  const int sw_cand[] = {64, 128, 256, 512};
  const int nl_cand[] = {128, 256};
  gg_batch_config best = keep;
  float best_ms = -1.0f;
  int rc = GG_OK;
  for (int br : br_cand) {
    if (br > nsc) continue;
    for (int sw : sw_cand)
      for (int nl : nl_cand) {
        gg_batch_config c = keep;
        c.branches = br;
        c.sweep_block = sw;
        c.nlist_block = nl;
        b->cfg = c;
        rc = build_graph(b);
        if (rc == GG_OK) rc = hip_status(hipGraphLaunch(b->exec, stream));  // warm-up
        if (rc == GG_OK) rc = hip_status(hipEventRecord(t0, stream));
        for (int t = 0; t < trials && rc == GG_OK; ++t)
          rc = hip_status(hipGraphLaunch(b->exec, stream));
        if (rc == GG_OK) rc = hip_status(hipEventRecord(t1, stream));
        if (rc == GG_OK) rc = hip_status(hipEventSynchronize(t1));
        float ms = 0.0f;
        if (rc == GG_OK) rc = hip_status(hipEventElapsedTime(&ms, t0, t1));
        if (rc == GG_OK) rc = snapshot(b, sn, true, stream);
        if (rc != GG_OK) break;
        ms /= (float)trials;
        if (best_ms < 0.0f || ms < best_ms) {
          best_ms = ms;
          best = c;
        }
      }
    if (rc != GG_OK) break;
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  const int rc_sync = hip_status(hipStreamSynchronize(stream));
  b->cfg = (rc == GG_OK) ? best : keep;
  const int rc_build = build_graph(b);
  if (rc != GG_OK) return rc;
  if (rc_sync != GG_OK) return rc_sync;
  if (rc_build != GG_OK) return rc_build;
  if (best_out) *best_out = best;
  if (best_ms_out) *best_ms_out = best_ms;
  return GG_OK;
}

extern "C" void gg_batch_destroy(gg_batch *b) {
  if (!b) return;
  release_graph(b);
  release_streams(b);
  delete b;
}

// ---- captured step -----------------------------------------------------------
struct gg_step_graph {
  gg_batch *b;
};

extern "C" int gg_step_graph_create(const gg_particles *set, int *cell_start,
                                    const gg_particles *boundary, const int *bcell_start,
                                    const gg_params *params, const gg_workspace *ws,
                                    int steps_per_launch, gg_step_graph **out) {
  if (!out) return GG_ERR_INVALID;
  *out = nullptr;
  if (!set || !cell_start || !boundary || !bcell_start || !params || !ws) return GG_ERR_INVALID;
  gg_batch_scene sc;
  sc.set = *set;
  sc.cell_start = cell_start;
  sc.boundary = *boundary;
  sc.bcell_start = bcell_start;
  sc.params = *params;
  sc.ws = *ws;
  gg_batch_config c;
  c.branches = 1;
  c.sweep_block = 0;
  c.nlist_block = 0;
  c.steps_per_launch = steps_per_launch;
  gg_batch *b = nullptr;
  GG_TRY(gg_batch_create(&sc, 1, &c, &b));
  gg_step_graph *g = new (std::nothrow) gg_step_graph{b};
  if (!g) {
    gg_batch_destroy(b);
    return GG_ERR_HIP;
  }
  *out = g;
  return GG_OK;
}

extern "C" int gg_step_graph_launch(gg_step_graph *g, hipStream_t stream) {
  if (!g) return GG_ERR_INVALID;
  return gg_batch_step(g->b, stream);
}

extern "C" void gg_step_graph_destroy(gg_step_graph *g) {
  if (!g) return;
  gg_batch_destroy(g->b);
  delete g;
}
