// Granular system of the host API (GranularSystem.hpp:6-107 of the
// reference): owns the granular, boundary and fine particle sets (moved in
// from the caller, whose pointers become null), their cell tables and the
// solver, and runs one step() in the constructor.
//
// step(): boundary re-sort when the boundary moves, cell sort + PBD step of
// the granular set, fine-particle sort + plain upsampler, then (adaptive)
// surface flags, distance to the surface and merge/split.  With the learned
// upsampler set, the fine set is rewritten by the network instead.  Returns 1 and
// appends the frame time in ms to get_frame_times(), printing it as the
// reference does (set_verbose(false) silences the print).  Exceptions are
// caught and printed, as in the reference.
#pragma once
#include <memory>
#include <vector>

#include "GranularParticles.hpp"
#include "Solver.hpp"

class GranularSystem {
 public:
  GranularSystem(std::shared_ptr<GranularParticles> &particles,
                 std::shared_ptr<GranularParticles> &boundary_particles,
                 std::shared_ptr<GranularParticles> &upsampled_particles, float3 space_size,
                 float cell_length, float dt, float3 g, int3 cell_size, const float density,
                 const float upsampled_radius, const bool is_move_boundary,
                 const int is_adaptive);

  GranularSystem(const GranularSystem &) = delete;
  GranularSystem &operator=(const GranularSystem &) = delete;

  float step();

  int size() const { return (*_particles).size(); }
  int upsampled_size() const { return (*_upsampled).size(); }
  int boundary_size() const { return (*_boundaries).size(); }
  int total_size() const {
    return (*_particles).size() + (*_boundaries).size() + (*_upsampled).size();
  }

  auto get_particles() const {
    return static_cast<const std::shared_ptr<GranularParticles>>(_particles);
  }
  auto get_particles_non_const() const {
    return static_cast<std::shared_ptr<GranularParticles>>(_particles);
  }
  auto get_boundaries() const {
    return static_cast<const std::shared_ptr<GranularParticles>>(_boundaries);
  }
  auto get_upsampled() const {
    return static_cast<const std::shared_ptr<GranularParticles>>(_upsampled);
  }

  float get_max_mass() const { return _merge_max_mass; }
  float get_min_mass() const { return (float)_min_mass; }
  float get_upsampled_radius() const { return _upsampled_radius; }

  const std::vector<float> &get_frame_times() const { return frame_times; }

  ~GranularSystem() { (void)hipDeviceSynchronize(); }

  Solver &get_solver() { return _solver; }

  void set_verbose(bool v) { _verbose = v; }

  // Merge/split rules of the adaptive passes (gg_params' adaptive_rules,
  // min_mass .. split_depth) and the largest mass a merge may start from;
  // the reference rules with max mass 4 until set.
  void set_merge_rules(int depth_rules, float min_mass, float split_mass, float merge_dist,
                       float merge_depth, float split_depth, float max_mass);

  // Learned output: from the next step on, the fine set holds 8 particles per
  // granular particle (fine index 8 i + k, slot k of the 2x2x2 template of
  // granular particle i) written by gg_upsample_mlp instead of the plain
  // upsampler's update.  w_bf16 (GG_MLP_WEIGHTS bfloat16) and bias
  // (GG_MLP_BIASES) are host arrays, copied to the device.
  void set_learned_upsampler(const uint16_t *w_bf16, const float *bias,
                             const gg_mlp_config &cfg);
  bool learned_upsampler() const { return _learned; }

 private:
  std::shared_ptr<GranularParticles> _particles;
  const std::shared_ptr<GranularParticles> _boundaries;
  std::shared_ptr<GranularParticles> _upsampled;
  const float _cell_length;
  const float3 _space_size;
  const float _dt;
  const float3 _g;
  const float _density;
  const int _max_mass;
  const int _min_mass;
  const int3 _cell_size;
  const float _upsampled_radius;
  // cell tables: C + 2 entries (cell starts, sentinel, end)
  DArray<int> _cell_start_particle;
  DArray<int> _cell_start_boundary;
  DArray<int> _cell_start_upsampled;
  std::unique_ptr<DArray<unsigned int>> _fine_ws_mem;
  gg_workspace _fine_ws;
  std::vector<float> frame_times;
  const bool _is_move_boundary;
  const int _is_adaptive;
  bool _verbose = true;
  float _merge_max_mass = 4.0f;
  bool _learned = false;
  std::unique_ptr<DArray<unsigned int>> _mlp_w;  // bfloat16 pairs
  std::unique_ptr<DArray<float>> _mlp_b;
  std::unique_ptr<DArray<unsigned int>> _mlp_scratch;
  gg_mlp_config _mlp_cfg{};
  Solver _solver;

  void learned_update();

  void ensure_capacity();
  void neighbor_search_granular(const std::shared_ptr<GranularParticles> &particles,
                                DArray<int> &cell_start);
  void neighbor_search_boundary(const std::shared_ptr<GranularParticles> &particles,
                                DArray<int> &cell_start);
  void neighbor_search_upsampled(const std::shared_ptr<GranularParticles> &particles,
                                 DArray<int> &cell_start);
  void set_surface_particles(const std::shared_ptr<GranularParticles> &particles,
                             DArray<int> &cell_start);
  void find_distance_to_surface(const std::shared_ptr<GranularParticles> &particles,
                                DArray<int> &cell_start);
};
