// Granular system of the host API over the library's C-ABI.
#include "granular/GranularSystem.hpp"

#include <iostream>
#include <stdexcept>

namespace {

void check(int rc, const char *what) {
  if (rc != GG_OK)
    throw std::runtime_error(std::string(what) + ": " + gg_status_string(rc));
}

unsigned int cells_plus2(int3 c) {
  const long long n = (long long)c.x * c.y * c.z;
  if (c.x <= 0 || c.y <= 0 || c.z <= 0 || n + 2 > 0x7fffffffLL)
    throw std::invalid_argument("invalid cell_size");
  return (unsigned int)(n + 2);
}

const std::shared_ptr<GranularParticles> &nonnull(const std::shared_ptr<GranularParticles> &p,
                                                  const char *what) {
  if (!p) throw std::invalid_argument(std::string("GranularSystem: null ") + what);
  return p;
}

}  // namespace

GranularSystem::GranularSystem(std::shared_ptr<GranularParticles> &granular_particles,
                               std::shared_ptr<GranularParticles> &boundary_particles,
                               std::shared_ptr<GranularParticles> &upsampled_particles,
                               const float3 space_size, const float cell_length, const float dt,
                               const float3 g, int3 cell_size, const float density,
                               const float upsampled_radius, const bool is_move_boundary,
                               const int is_adaptive)
    : _particles(std::move(granular_particles)),
      _boundaries(std::move(boundary_particles)),
      _upsampled(std::move(upsampled_particles)),
      _cell_length(cell_length),
      _space_size(space_size),
      _dt(dt),
      _g(g),
      _density(density),
      _max_mass(4),
      _min_mass(1),
      _cell_size(cell_size),
      _upsampled_radius(upsampled_radius),
      _cell_start_particle(cells_plus2(cell_size)),
      _cell_start_boundary(cells_plus2(cell_size)),
      _cell_start_upsampled(cells_plus2(cell_size)),
      _is_move_boundary(is_move_boundary),
      _is_adaptive(is_adaptive),
      _solver(nonnull(_particles, "particles")) {
  nonnull(_boundaries, "boundary particles");
  nonnull(_upsampled, "upsampled particles");
  // fine-particle sort workspace: cell sort only
  {
    gg_grid gr{cell_length, cell_size.x, cell_size.y, cell_size.z};
    const int nf = (int)std::max(_upsampled->size(), 1u);
    const size_t bytes = gg_workspace_bytes(nf, 0, gr, 0, 0, 0);
    if (bytes == 0) throw std::runtime_error("invalid fine workspace request");
    _fine_ws_mem.reset(new DArray<unsigned int>((unsigned int)((bytes + 3) / 4)));
    _fine_ws = gg_workspace{};
    _fine_ws.base = _fine_ws_mem->addr();
    _fine_ws.bytes = bytes;
    _fine_ws.capacity = nf;
    _fine_ws.boundary_capacity = 0;
    _fine_ws.grid = gr;
    check(gg_workspace_init(&_fine_ws, nullptr), "gg_workspace_init");
  }
  neighbor_search_boundary(_boundaries, _cell_start_boundary);
  // every granular particle starts with mass 1 (GranularSystem.cu:55-57)
  check(gg_fill_u32(_particles->get_mass_ptr(), _particles->size(), 0x3f800000u, nullptr),
        "gg_fill_u32");
  ensure_capacity();
  neighbor_search_granular(_particles, _cell_start_particle);
  neighbor_search_upsampled(_upsampled, _cell_start_upsampled);
  step();
}

void GranularSystem::ensure_capacity() {
  // merge/split work in place: room for every particle to split once
  const unsigned int n = _particles->size();
  const unsigned int want = _is_adaptive == 1 ? 2 * std::max(n, 1u) : std::max(n, 1u);
  if (_particles->device_capacity() < want) _particles->reserve(want);
  _solver.reserve(want);
}

void GranularSystem::neighbor_search_granular(const std::shared_ptr<GranularParticles> &particles,
                                              DArray<int> &cell_start) {
  auto p = particles;
  _solver.sort_particles(p, _boundaries, cell_start, _cell_size, _cell_length);
}

void GranularSystem::neighbor_search_boundary(const std::shared_ptr<GranularParticles> &particles,
                                              DArray<int> &cell_start) {
  _solver.sort_boundary(particles, cell_start, _cell_size, _cell_length);
}

void GranularSystem::neighbor_search_upsampled(const std::shared_ptr<GranularParticles> &particles,
                                               DArray<int> &cell_start) {
  gg_particles f = particles->view();
  check(gg_upsampled_sort(&f, cell_start.addr(), &_fine_ws, nullptr), "gg_upsampled_sort");
}

void GranularSystem::set_surface_particles(const std::shared_ptr<GranularParticles> &particles,
                                           DArray<int> &cell_start) {
  auto p = particles;
  _solver.surface(p, cell_start, _cell_size, _cell_length, _density);
}

void GranularSystem::find_distance_to_surface(
    const std::shared_ptr<GranularParticles> &particles, DArray<int> &cell_start) {
  auto p = particles;
  _solver.surface_distance(p, cell_start, _cell_size, _cell_length);
}

float GranularSystem::step() {
  hipEvent_t start = nullptr, stop = nullptr;
  HIP_CALL(hipEventCreate(&start));
  HIP_CALL(hipEventCreate(&stop));
  HIP_CALL(hipEventRecord(start, nullptr));
  try {
    ensure_capacity();
    if (_is_move_boundary) neighbor_search_boundary(_boundaries, _cell_start_boundary);
    // neighbor_search_granular + Solver::step, fused
    _solver.sort_and_step(_particles, _boundaries, _cell_start_particle, _cell_start_boundary,
                          _space_size, _cell_size, _cell_length, _dt, _g, _density);
    if (_learned) {
      learned_update();
    } else {
      neighbor_search_upsampled(_upsampled, _cell_start_upsampled);
      _solver.upsampled_update(_particles, _boundaries, _upsampled, _cell_start_upsampled,
                               _cell_start_particle, _cell_start_boundary, _cell_size,
                               _space_size, _cell_length, _density);
    }
    if (_is_adaptive == 1) {
      set_surface_particles(_particles, _cell_start_particle);
      find_distance_to_surface(_particles, _cell_start_particle);
      _solver.adaptive_sampling(_particles, _boundaries, _cell_start_particle,
                                _cell_start_boundary, _merge_max_mass, _cell_size, _space_size,
                                _cell_length, _density);
    }
  } catch (const char *s) {
    std::cout << s << "\n";
  } catch (const std::exception &e) {
    std::cout << "Exception at " << __FILE__ << ": " << e.what() << "\n";
  } catch (...) {
    std::cout << "Unknown Exception at " << __FILE__ << ": line" << __LINE__ << "\n";
  }
  float milliseconds = 0.0f;
  HIP_CALL(hipEventRecord(stop, nullptr));
  HIP_CALL(hipEventSynchronize(stop));
  HIP_CALL(hipEventElapsedTime(&milliseconds, start, stop));
  HIP_CALL(hipEventDestroy(start));
  HIP_CALL(hipEventDestroy(stop));
  frame_times.push_back(milliseconds);
  if (_verbose) std::cout << "Frame time : " << milliseconds << "\n";
  return 1;
}

void GranularSystem::set_merge_rules(int depth_rules, float min_mass, float split_mass,
                                     float merge_dist, float merge_depth, float split_depth,
                                     float max_mass) {
  if (!(min_mass > 0.0f) || !(split_mass > 0.0f) || !(merge_dist > 0.0f) ||
      !(max_mass > min_mass))
    throw std::invalid_argument("set_merge_rules: masses and distances must be positive");
  _solver.set_merge_rules(depth_rules, min_mass, split_mass, merge_dist, merge_depth,
                          split_depth);
  _merge_max_mass = max_mass;
}

void GranularSystem::set_learned_upsampler(const uint16_t *w_bf16, const float *bias,
                                           const gg_mlp_config &cfg) {
  if (!w_bf16 || !bias) throw std::invalid_argument("set_learned_upsampler: null weights");
  if (!(cfg.radius > 0.0f) || !(cfg.vscale > 0.0f))
    throw std::invalid_argument("set_learned_upsampler: radius and vscale must be positive");
  if (!_mlp_w) _mlp_w.reset(new DArray<unsigned int>(GG_MLP_WEIGHTS / 2));
  if (!_mlp_b) _mlp_b.reset(new DArray<float>(GG_MLP_BIASES));
  HIP_CALL(hipMemcpy(_mlp_w->addr(), w_bf16, sizeof(uint16_t) * GG_MLP_WEIGHTS,
                     hipMemcpyHostToDevice));
  HIP_CALL(hipMemcpy(_mlp_b->addr(), bias, sizeof(float) * GG_MLP_BIASES,
                     hipMemcpyHostToDevice));
  _mlp_cfg = cfg;
  _learned = true;
}

// 8 fine particles per granular particle from the network (features over
// the step's granular cell table; teacher fallback near boundaries).
void GranularSystem::learned_update() {
  const unsigned int n = _particles->size();
  const unsigned int want = 8u * std::max(_particles->device_capacity(), 1u);
  if (_upsampled->device_capacity() < want) _upsampled->reserve(want);
  _upsampled->set_size(8u * n);
  if (n == 0) return;
  const size_t need = gg_upsample_mlp_scratch_bytes((int)n);
  const size_t words = (need + 3) / 4;
  if (words > 0xffffffffu) throw std::runtime_error("learned_update: scratch too large");
  if (!_mlp_scratch || _mlp_scratch->length() < words)
    _mlp_scratch.reset(new DArray<unsigned int>((unsigned int)words));
  gg_particles c = _particles->view();
  gg_particles b = _boundaries->view();
  gg_upsample_params u;
  gg_upsample_params_default(&u);
  const gg_grid gr{_cell_length, _cell_size.x, _cell_size.y, _cell_size.z};
  check(gg_upsample_mlp(&c, _cell_start_particle.addr(), &b, _cell_start_boundary.addr(), gr, &u,
                        &_mlp_cfg, reinterpret_cast<const uint16_t *>(_mlp_w->addr()),
                        _mlp_b->addr(),
                        reinterpret_cast<float *>(_upsampled->get_pos_ptr()),
                        reinterpret_cast<float *>(_upsampled->get_vel_ptr()),
                        _mlp_scratch->addr(), 4 * (size_t)_mlp_scratch->length(), nullptr,
                        nullptr),
        "gg_upsample_mlp");
}
