// Granular system of the host API (GranularSystem.hpp:6-107 of the
// reference): owns the granular, boundary and fine particle sets (moved in
// from the caller, whose pointers become null), their cell tables and the
// solver, and runs one step() in the constructor.
//
// step(): boundary re-sort when the boundary moves, cell sort + PBD step of
// the granular set, fine-particle sort + plain upsampler, then (adaptive)
// surface flags, distance to the surface and merge/split.  Returns 1 and
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

  float get_max_mass() const { return (float)_max_mass; }
  float get_min_mass() const { return (float)_min_mass; }
  float get_upsampled_radius() const { return _upsampled_radius; }

  const std::vector<float> &get_frame_times() const { return frame_times; }

  ~GranularSystem() { (void)hipDeviceSynchronize(); }

  Solver &get_solver() { return _solver; }

  void set_verbose(bool v) { _verbose = v; }

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
  Solver _solver;

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
