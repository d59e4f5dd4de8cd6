// PBD solver of the host API (Solver.hpp:8-118 of the reference) over the
// library's stage entry points.
//
// Differences a caller can see, each a design decision:
// - step() and the system's step run add_external_force,
//   update_particle_positions and apply_mass_scaling as one fused pass,
//   predict() (gg_predict); separate passes would add two full read/write
//   sweeps of the particle arrays.  The reference's separate methods are
//   kept as entry points of their own for callers that drive the stages.
// - final_update also clears adaptive_last_step (the reference clears it in
//   step() right after final_update).
// - The merge-skip toggle is per instance (the reference's t_merge_iter is a
//   process global, Solver.cu:36); with one system per process they agree.
// - The workspace (cell-sort, neighbour-list and sweep scratch) is owned
//   here and sized from the particle capacity, the boundary size and the
//   grid; it is rebuilt, with the boundary pack, when any of them grows.
#pragma once
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "GranularParticles.hpp"

class Solver {
 public:
  explicit Solver(const std::shared_ptr<GranularParticles> &particles);
  Solver(const Solver &) = delete;
  Solver &operator=(const Solver &) = delete;
  ~Solver() {}

  // predict -> project(5 solver sweeps) -> final_update, on a set that is
  // already sorted by cell (Solver.cu:82-109).
  void step(std::shared_ptr<GranularParticles> &particles,
            const std::shared_ptr<GranularParticles> &boundary,
            const DArray<int> &cell_start_particle, const DArray<int> &cell_start_boundary,
            float3 space_size, int3 cell_size, float cell_length, float dt, float3 G,
            const float density);

  // Cell sort of the granular set (neighbor_search_granular) fused with
  // step(): the system's per-step path (gg_step).
  void sort_and_step(std::shared_ptr<GranularParticles> &particles,
                     const std::shared_ptr<GranularParticles> &boundary,
                     DArray<int> &cell_start_particle, const DArray<int> &cell_start_boundary,
                     float3 space_size, int3 cell_size, float cell_length, float dt, float3 G,
                     const float density);

  // Cell sort of the granular set and its merge buffers (gg_grid_build).
  void sort_particles(std::shared_ptr<GranularParticles> &particles,
                      const std::shared_ptr<GranularParticles> &boundary,
                      DArray<int> &cell_start_particle, int3 cell_size, float cell_length);

  // Cell sort + pack of the boundary set (gg_boundary_build).
  void sort_boundary(const std::shared_ptr<GranularParticles> &boundary,
                     DArray<int> &cell_start_boundary, int3 cell_size, float cell_length);

  // External force, prediction and mass scaling, fused (gg_predict).
  void predict(std::shared_ptr<GranularParticles> &particles,
               const std::shared_ptr<GranularParticles> &boundary,
               const DArray<int> &cell_start_particle, const DArray<int> &cell_start_boundary,
               int3 cell_size, float cell_length, float dt, float3 G, const float density);

  // 3 stabilisation + max_iter solver contact sweeps (Solver.cu:403-513).
  void project(std::shared_ptr<GranularParticles> &particles,
               const std::shared_ptr<GranularParticles> &boundaries,
               const DArray<int> &cell_start_granular, const DArray<int> &cell_start_boundary,
               int3 cell_size, float3 space_size, float cell_length, int max_iter,
               const float density);

  void final_update(std::shared_ptr<GranularParticles> &particles, float dt);

  // The reference's separate stages (Solver.cu:112-191), one pass each:
  // v += dt G; predicted positions x + dt v into the solver's _pos_t
  // buffer (get_pos_t_ptr); scaled mass m exp(-y); and the stable sort of
  // the positions alone by particle_2_cell (the keys stay as they are).
  void update_neighborhood(const std::shared_ptr<GranularParticles> &particles);
  void add_external_force(std::shared_ptr<GranularParticles> &particles, float dt, float3 G);
  void update_particle_positions(std::shared_ptr<GranularParticles> &particles, float dt);
  void apply_mass_scaling(std::shared_ptr<GranularParticles> &particles);
  float3 *get_pos_t_ptr() const { return _pos_t ? _pos_t->addr() : nullptr; }

  // Merge (skipped on the first call), blend, split, compaction
  // (Solver.cu:856-1242); N changes and the stats tuple is recorded.
  void adaptive_sampling(std::shared_ptr<GranularParticles> &particles,
                         const std::shared_ptr<GranularParticles> &boundaries,
                         const DArray<int> &cell_start_granular,
                         const DArray<int> &cell_start_boundary, const float max_mass,
                         int3 cell_size, float3 space_size, float cell_length,
                         const float density);

  // Plain upsampler pass (Solver.cu:1370-1387).
  void upsampled_update(std::shared_ptr<GranularParticles> &particles,
                        const std::shared_ptr<GranularParticles> &boundaries,
                        std::shared_ptr<GranularParticles> &upsampled,
                        const DArray<int> &cell_start_upsampled,
                        const DArray<int> &cell_start_granular,
                        const DArray<int> &cell_start_boundary, int3 cell_size,
                        float3 space_size, float cell_length, const float density);

  // Surface flags and distance to the surface (GranularSystem.cu:613-699).
  void surface(std::shared_ptr<GranularParticles> &particles,
               const DArray<int> &cell_start_granular, int3 cell_size, float cell_length,
               const float density);
  void surface_distance(std::shared_ptr<GranularParticles> &particles,
                        const DArray<int> &cell_start_granular, int3 cell_size,
                        float cell_length);

  int *get_buffer_merge_count_ptr() const { return _buffer_merge_count.addr(); }
  float *get_buffer_merge_ptr() const { return _buffer_merge.addr(); }
  int *get_buffer_remove_ptr() const { return _buffer_remove.addr(); }
  float3 *get_buffer_merge_velocity_ptr() const { return _buffer_merge_velocity.addr(); }

  void resize(const int size) {
    _buffer_merge_count.resize((unsigned int)size);
    _buffer_merge.resize((unsigned int)size);
    _buffer_remove.resize((unsigned int)size);
  }

  // Solver::trigger_explosion (Solver.cu:1411-1442): radial impulse
  // force / (1 + d^2) / m away from the centre of mass (gg_explosion; the
  // centre is summed in double on the device).  The centre of the last
  // explosion is kept for get_explosion_center().
  void trigger_explosion(std::shared_ptr<GranularParticles> &particles, float explosion_force);
  float3 get_explosion_center() const { return _explosion_center; }

  void save_particle_stats(const std::string &filename);
  // Particles that scanned their whole 27-cell stencil (instead of their
  // neighbour list) in each sweep of the last step, sweep 0 first (at most
  // max counts; returns the number of sweeps).  A diagnostic of the lists:
  // results never depend on it.
  int sweep_scan_counts(int *out, int max) const;
  const std::vector<std::tuple<int, int, int>> &get_particle_stats() const {
    return _particle_stats;
  }

  // Grows the merge buffers (and the workspace on next use) to cap particles.
  void reserve(unsigned int cap);

  // Options the reference fixes in code: iteration counts (3 + 5), friction
  // (none), library flags (GG_FLAG_*).
  void set_iterations(int stab, int solver);
  void set_friction(float mu_s, float mu_k);
  void set_flags(int flags);
  // Merge/split rules (gg_params adaptive_rules, min_mass, split_mass,
  // merge_dist, merge_depth, split_depth); the reference rules until set.
  void set_merge_rules(int depth_rules, float min_mass, float split_mass, float merge_dist,
                       float merge_depth, float split_depth);
  // Neighbour lists: granular and boundary candidates per particle and the
  // skin (m).  Defaults 48, 16, 0.0075 fit r = 0.01 with 0.01 boundary
  // particles; smaller particles near dense boundaries need more boundary
  // slots (a particle whose list overflows scans its whole stencil every
  // sweep, with the same results).
  void set_neighbor_lists(int gmax, int bmax, float skin);

 private:
  gg_params params(float dt, float3 G, float density) const;
  gg_adaptive adaptive_view();
  gg_grid grid(int3 cell_size, float cell_length) const;
  void ensure_workspace(int cap, int bcap, gg_grid g);
  const gg_workspace &workspace(const std::shared_ptr<GranularParticles> &particles,
                                const std::shared_ptr<GranularParticles> &boundary,
                                const DArray<int> *cell_start_boundary, int3 cell_size,
                                float cell_length);

  const int _max_iter;
  const int _blend_factor;
  DArray<int> _buffer_remove;
  DArray<float> _buffer_merge;
  DArray<float3> _buffer_merge_velocity;
  DArray<int> _buffer_merge_count;
  std::unique_ptr<DArray<unsigned int>> _ws_mem;
  gg_workspace _ws;
  bool _ws_ready = false;
  // boundary set and cell table last sorted into the workspace (the packed
  // boundary is rebuilt from them when the workspace is reallocated)
  std::shared_ptr<GranularParticles> _bnd;
  const DArray<int> *_bcs = nullptr;
  int _gmax = 48, _bmax = 16;
  float _skin = 0.0075f;
  int _rules = 0;
  float _min_mass = 1.0f, _split_mass = 2.0f, _merge_dist = 0.035f, _merge_depth = 3.0f,
        _split_depth = 2.0f;
  int _stab = 3;
  int _solver_iters = 5;
  float _mu_s = 0.0f, _mu_k = 0.0f;
  int _flags = 0;
  int _t_merge_iter = 0;
  // dt, gravity and density of the last predict (final_update and the
  // adaptive passes take them from there, as the reference's kernels do
  // from the system)
  float _last_dt = 0.002f;
  float3 _last_G = make_float3(0.0f, -9.8f, 0.0f);
  float _last_density = 238732.4146f;
  std::vector<std::tuple<int, int, int>> _particle_stats;
  // predicted positions of update_particle_positions, and the cell table
  // that update_neighborhood's sort fills
  std::unique_ptr<DArray<float3>> _pos_t;
  std::unique_ptr<DArray<int>> _key_cells;
  float3 _explosion_center = make_float3(0.0f, 0.0f, 0.0f);
};
