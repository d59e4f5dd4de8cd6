// Solver of the host API over the library's C-ABI.
#include "granular/Solver.hpp"

#include <fstream>
#include <iostream>
#include <stdexcept>

namespace {

void check(int rc, const char *what) {
  if (rc != GG_OK)
    throw std::runtime_error(std::string(what) + ": " + gg_status_string(rc));
}

// gg_workspace_view(8): counter 1 + k % 62 holds sweep k's stencil-scan count
constexpr int kScanSlots = 62;

}  // namespace

Solver::Solver(const std::shared_ptr<GranularParticles> &particles)
    : _max_iter(5),
      _blend_factor(40),
      _buffer_remove(particles->size()),
      _buffer_merge(particles->size()),
      _buffer_merge_velocity(particles->size()),
      _buffer_merge_count(particles->size()) {
  _ws = gg_workspace{};
}

void Solver::set_iterations(int stab, int solver) {
  if (stab < 0 || solver < 0) throw std::invalid_argument("iteration counts must be >= 0");
  _stab = stab;
  _solver_iters = solver;
}

void Solver::set_friction(float mu_s, float mu_k) {
  _mu_s = mu_s;
  _mu_k = mu_k;
}

void Solver::set_neighbor_lists(int gmax, int bmax, float skin) {
  if (gmax < 0 || bmax < 0 || !(skin >= 0.0f))
    throw std::invalid_argument("set_neighbor_lists: negative capacity or skin");
  _gmax = gmax;
  _bmax = bmax;
  _skin = skin;
}

void Solver::set_merge_rules(int depth_rules, float min_mass, float split_mass,
                             float merge_dist, float merge_depth, float split_depth) {
  _rules = depth_rules ? 1 : 0;
  _min_mass = min_mass;
  _split_mass = split_mass;
  _merge_dist = merge_dist;
  _merge_depth = merge_depth;
  _split_depth = split_depth;
}

void Solver::set_flags(int flags) { _flags = flags; }

void Solver::reserve(unsigned int cap) {
  _buffer_remove.reserve(cap);
  _buffer_merge.reserve(cap);
  _buffer_merge_velocity.reserve(cap);
  _buffer_merge_count.reserve(cap);
}

gg_params Solver::params(float dt, float3 G, float density) const {
  gg_params p;
  gg_params_default(&p);
  p.dt = dt;
  p.gx = G.x;
  p.gy = G.y;
  p.gz = G.z;
  p.density = density;
  p.stab_iters = _stab;
  p.solver_iters = _solver_iters;
  p.mu_s = _mu_s;
  p.mu_k = _mu_k;
  p.flags = _flags;
  p.skin = _skin;
  p.adaptive_rules = _rules;
  p.min_mass = _min_mass;
  p.split_mass = _split_mass;
  p.merge_dist = _merge_dist;
  p.merge_depth = _merge_depth;
  p.split_depth = _split_depth;
  return p;
}

gg_adaptive Solver::adaptive_view() {
  gg_adaptive a;
  a.remove = _buffer_remove.addr();
  a.merge = _buffer_merge.addr();
  a.merge_velocity = reinterpret_cast<gg_float3 *>(_buffer_merge_velocity.addr());
  a.merge_count = _buffer_merge_count.addr();
  return a;
}

gg_grid Solver::grid(int3 cell_size, float cell_length) const {
  gg_grid g;
  g.cell_length = cell_length;
  g.cx = cell_size.x;
  g.cy = cell_size.y;
  g.cz = cell_size.z;
  return g;
}

void Solver::ensure_workspace(int cap, int bcap, gg_grid g) {
  const int sweeps = _stab + _solver_iters;
  const bool same_grid = _ws_ready && _ws.grid.cx == g.cx && _ws.grid.cy == g.cy &&
                         _ws.grid.cz == g.cz && _ws.grid.cell_length == g.cell_length;
  if (same_grid && _ws.capacity >= cap && _ws.boundary_capacity >= bcap &&
      _ws.max_sweeps >= sweeps && _ws.gmax == _gmax && _ws.bmax == _bmax)
    return;
  gg_workspace w{};
  w.capacity = same_grid ? std::max(cap, _ws.capacity) : cap;
  w.boundary_capacity = same_grid ? std::max(bcap, _ws.boundary_capacity) : bcap;
  w.grid = g;
  w.max_sweeps = same_grid ? std::max(sweeps, _ws.max_sweeps) : sweeps;
  w.gmax = _gmax;
  w.bmax = _bmax;
  const size_t bytes = gg_workspace_bytes(w.capacity, w.boundary_capacity, g, w.max_sweeps,
                                          w.gmax, w.bmax);
  if (bytes == 0) throw std::runtime_error("invalid workspace request");
  _ws_mem.reset();
  _ws_mem.reset(new DArray<unsigned int>((unsigned int)((bytes + 3) / 4)));
  w.base = _ws_mem->addr();
  w.bytes = bytes;
  check(gg_workspace_init(&w, nullptr), "gg_workspace_init");
  _ws = w;
  _ws_ready = true;
  // the packed boundary lives in the workspace: rebuild it
  if (_bnd && _bcs) {
    gg_particles b = _bnd->view();
    check(gg_boundary_build(&b, _bcs->addr(), &_ws, nullptr), "gg_boundary_build");
  }
}

const gg_workspace &Solver::workspace(const std::shared_ptr<GranularParticles> &particles,
                                      const std::shared_ptr<GranularParticles> &boundary,
                                      const DArray<int> *cell_start_boundary, int3 cell_size,
                                      float cell_length) {
  if (boundary && cell_start_boundary) {
    _bnd = boundary;
    _bcs = cell_start_boundary;
  }
  int cap = (int)std::min({_buffer_remove.capacity(), _buffer_merge.capacity(),
                           _buffer_merge_velocity.capacity(), _buffer_merge_count.capacity()});
  if (particles) cap = std::min(cap, (int)particles->device_capacity());
  const int bcap = _bnd ? (int)std::max(_bnd->size(), 1u) : 1;
  ensure_workspace(std::max(cap, 1), bcap, grid(cell_size, cell_length));
  return _ws;
}

void Solver::sort_boundary(const std::shared_ptr<GranularParticles> &boundary,
                           DArray<int> &cell_start_boundary, int3 cell_size, float cell_length) {
  _bnd = boundary;
  _bcs = &cell_start_boundary;
  const int cap = (int)std::max(std::min({_buffer_remove.capacity(), _buffer_merge.capacity(),
                                          _buffer_merge_velocity.capacity(),
                                          _buffer_merge_count.capacity()}),
                                1u);
  ensure_workspace(cap, (int)std::max(boundary->size(), 1u), grid(cell_size, cell_length));
  gg_particles b = boundary->view();
  check(gg_boundary_build(&b, cell_start_boundary.addr(), &_ws, nullptr), "gg_boundary_build");
}

void Solver::sort_particles(std::shared_ptr<GranularParticles> &particles,
                            const std::shared_ptr<GranularParticles> &boundary,
                            DArray<int> &cell_start_particle, int3 cell_size, float cell_length) {
  const gg_workspace &ws = workspace(particles, boundary, nullptr, cell_size, cell_length);
  gg_particles s = particles->view();
  gg_adaptive a = adaptive_view();
  check(gg_grid_build(&s, &a, cell_start_particle.addr(), nullptr, &ws, nullptr), "gg_grid_build");
}

void Solver::predict(std::shared_ptr<GranularParticles> &particles,
                     const std::shared_ptr<GranularParticles> &boundary,
                     const DArray<int> &cell_start_particle,
                     const DArray<int> &cell_start_boundary, int3 cell_size, float cell_length,
                     float dt, float3 G, const float density) {
  const gg_workspace &ws =
      workspace(particles, boundary, &cell_start_boundary, cell_size, cell_length);
  gg_particles s = particles->view();
  gg_particles b = boundary->view();
  const gg_params p = params(dt, G, density);
  check(gg_predict(&s, cell_start_particle.addr(), &b, cell_start_boundary.addr(), &p, &ws,
                   nullptr),
        "gg_predict");
  _last_dt = dt;
  _last_G = G;
  _last_density = density;
}

void Solver::project(std::shared_ptr<GranularParticles> &particles,
                     const std::shared_ptr<GranularParticles> &boundaries,
                     const DArray<int> &cell_start_granular,
                     const DArray<int> &cell_start_boundary, int3 cell_size, float3 space_size,
                     float cell_length, int max_iter, const float density) {
  (void)space_size;
  const int saved = _solver_iters;
  _solver_iters = max_iter;
  try {
    const gg_workspace &ws =
        workspace(particles, boundaries, &cell_start_boundary, cell_size, cell_length);
    gg_particles s = particles->view();
    gg_particles b = boundaries->view();
    const gg_params p = params(_last_dt, _last_G, density);
    for (int k = 0; k < _stab + max_iter; ++k)
      check(gg_contact_sweep(&s, cell_start_granular.addr(), &b, cell_start_boundary.addr(), &p,
                             k, &ws, nullptr),
            "gg_contact_sweep");
  } catch (...) {
    _solver_iters = saved;
    throw;
  }
  _solver_iters = saved;
}

void Solver::final_update(std::shared_ptr<GranularParticles> &particles, float dt) {
  if (!_ws_ready) throw std::runtime_error("final_update before predict");
  gg_particles s = particles->view();
  const gg_params p = params(dt, _last_G, _last_density);
  check(gg_finalize(&s, &p, &_ws, nullptr), "gg_finalize");
}

void Solver::step(std::shared_ptr<GranularParticles> &particles,
                  const std::shared_ptr<GranularParticles> &boundary,
                  const DArray<int> &cell_start_particle, const DArray<int> &cell_start_boundary,
                  float3 space_size, int3 cell_size, float cell_length, float dt, float3 G,
                  const float density) {
  predict(particles, boundary, cell_start_particle, cell_start_boundary, cell_size, cell_length,
          dt, G, density);
  project(particles, boundary, cell_start_particle, cell_start_boundary, cell_size, space_size,
          cell_length, _solver_iters, density);
  final_update(particles, dt);
}

void Solver::sort_and_step(std::shared_ptr<GranularParticles> &particles,
                           const std::shared_ptr<GranularParticles> &boundary,
                           DArray<int> &cell_start_particle,
                           const DArray<int> &cell_start_boundary, float3 space_size,
                           int3 cell_size, float cell_length, float dt, float3 G,
                           const float density) {
  (void)space_size;
  const gg_workspace &ws =
      workspace(particles, boundary, &cell_start_boundary, cell_size, cell_length);
  gg_particles s = particles->view();
  gg_particles b = boundary->view();
  gg_adaptive a = adaptive_view();
  const gg_params p = params(dt, G, density);
  check(gg_step(&s, &a, cell_start_particle.addr(), &b, cell_start_boundary.addr(), &p, &ws,
                nullptr),
        "gg_step");
  _last_dt = dt;
  _last_G = G;
  _last_density = density;
}

void Solver::surface(std::shared_ptr<GranularParticles> &particles,
                     const DArray<int> &cell_start_granular, int3 cell_size, float cell_length,
                     const float density) {
  const gg_workspace &ws = workspace(particles, nullptr, nullptr, cell_size, cell_length);
  gg_particles s = particles->view();
  const gg_params p = params(_last_dt, _last_G, density);
  check(gg_surface(&s, cell_start_granular.addr(), &p, &ws, nullptr), "gg_surface");
}

void Solver::surface_distance(std::shared_ptr<GranularParticles> &particles,
                              const DArray<int> &cell_start_granular, int3 cell_size,
                              float cell_length) {
  const gg_workspace &ws = workspace(particles, nullptr, nullptr, cell_size, cell_length);
  gg_particles s = particles->view();
  check(gg_surface_distance(&s, cell_start_granular.addr(), &ws, nullptr), "gg_surface_distance");
}

void Solver::adaptive_sampling(std::shared_ptr<GranularParticles> &particles,
                               const std::shared_ptr<GranularParticles> &boundaries,
                               const DArray<int> &cell_start_granular,
                               const DArray<int> &cell_start_boundary, const float max_mass,
                               int3 cell_size, float3 space_size, float cell_length,
                               const float density) {
  (void)space_size;
  const int num = (int)particles->size();
  if (num == 0) return;
  const gg_workspace &ws =
      workspace(particles, boundaries, &cell_start_boundary, cell_size, cell_length);
  // merge is skipped on the first call (Solver.cu:910-922)
  const int do_merge = (_t_merge_iter == 1) ? 1 : 0;
  if (do_merge) _t_merge_iter = 0;
  _t_merge_iter++;
  gg_particles s = particles->view();
  gg_particles b = boundaries->view();
  gg_adaptive a = adaptive_view();
  const gg_params p = params(_last_dt, _last_G, density);
  int stats[3] = {0, 0, 0};
  check(gg_adaptive_sampling(&s, &a, cell_start_granular.addr(), &b, cell_start_boundary.addr(),
                             &p, max_mass, do_merge, _blend_factor, &ws, nullptr, stats),
        "gg_adaptive_sampling");
  particles->set_size((unsigned int)s.n);
  _buffer_remove.set_length((unsigned int)s.n);
  _buffer_merge.set_length((unsigned int)s.n);
  _buffer_merge_count.set_length((unsigned int)s.n);
  _particle_stats.push_back(std::make_tuple(stats[0], stats[1], num));
}

void Solver::upsampled_update(std::shared_ptr<GranularParticles> &particles,
                              const std::shared_ptr<GranularParticles> &boundaries,
                              std::shared_ptr<GranularParticles> &upsampled,
                              const DArray<int> &cell_start_upsampled,
                              const DArray<int> &cell_start_granular,
                              const DArray<int> &cell_start_boundary, int3 cell_size,
                              float3 space_size, float cell_length, const float density) {
  (void)cell_start_upsampled;
  (void)space_size;
  (void)density;
  gg_particles f = upsampled->view();
  gg_particles c = particles->view();
  gg_particles b = boundaries->view();
  gg_upsample_params u;
  gg_upsample_params_default(&u);
  check(gg_upsample_plain(&f, &c, cell_start_granular.addr(), &b, cell_start_boundary.addr(),
                          grid(cell_size, cell_length), &u, _flags, nullptr),
        "gg_upsample_plain");
}

void Solver::trigger_explosion(std::shared_ptr<GranularParticles> &particles,
                               float explosion_force) {
  _explosion_center = make_float3(0.0f, 0.0f, 0.0f);
  const int n = (int)particles->size();
  if (n == 0) return;
  DArray<unsigned int> scratch((unsigned int)((gg_explosion_scratch_bytes(n) + 3) / 4));
  gg_particles s = particles->view();
  gg_float3 c{0.0f, 0.0f, 0.0f};
  check(gg_explosion(&s, explosion_force, scratch.addr(), 4 * (size_t)scratch.length(), &c,
                     nullptr),
        "gg_explosion");
  _explosion_center = make_float3(c.x, c.y, c.z);
}

void Solver::add_external_force(std::shared_ptr<GranularParticles> &particles, float dt,
                                float3 G) {
  gg_particles s = particles->view();
  const gg_float3 dv{dt * G.x, dt * G.y, dt * G.z};
  check(gg_add_external_force(&s, dv, nullptr), "gg_add_external_force");
}

void Solver::update_particle_positions(std::shared_ptr<GranularParticles> &particles, float dt) {
  const unsigned int n = particles->size();
  if (!_pos_t || _pos_t->capacity() < n) _pos_t.reset(new DArray<float3>(n));
  _pos_t->set_length(n);
  gg_particles s = particles->view();
  check(gg_predict_positions(&s, dt, reinterpret_cast<gg_float3 *>(_pos_t->addr()), nullptr),
        "gg_predict_positions");
}

void Solver::apply_mass_scaling(std::shared_ptr<GranularParticles> &particles) {
  gg_particles s = particles->view();
  check(gg_mass_scaling(&s, nullptr), "gg_mass_scaling");
}

void Solver::update_neighborhood(const std::shared_ptr<GranularParticles> &particles) {
  if (!_ws_ready)
    throw std::runtime_error("update_neighborhood: the grid is set by the first sort or step");
  if ((int)particles->size() > _ws.capacity)
    ensure_workspace((int)particles->size(), _ws.boundary_capacity, _ws.grid);
  const unsigned int cells = (unsigned int)((long long)_ws.grid.cx * _ws.grid.cy * _ws.grid.cz + 2);
  if (!_key_cells || _key_cells->length() != cells) _key_cells.reset(new DArray<int>(cells));
  gg_particles s = particles->view();
  check(gg_sort_positions_by_key(&s, _key_cells->addr(), &_ws, nullptr),
        "gg_sort_positions_by_key");
}

int Solver::sweep_scan_counts(int *out, int max) const {
  const int sweeps = _stab + _solver_iters;
  if (!_ws_ready || !out || max <= 0) return sweeps;
  const int *ctr = static_cast<const int *>(gg_workspace_view(&_ws, 8));
  if (!ctr) throw std::runtime_error("sweep_scan_counts: no workspace");
  const int k = sweeps < max ? sweeps : max;
  // ctr[1 + k] holds sweep k's count (sweeps beyond the slot count wrap)
  const int m = k < kScanSlots ? k : kScanSlots;
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(out, ctr + 1, sizeof(int) * (size_t)m, hipMemcpyDeviceToHost) != hipSuccess)
    throw std::runtime_error("sweep_scan_counts: copy failed");
  for (int i = m; i < k; ++i) out[i] = out[i % kScanSlots];
  return sweeps;
}

void Solver::save_particle_stats(const std::string &filename) {
  std::ofstream outfile(filename);
  if (!outfile.is_open()) {
    std::cerr << "Failed to open file for particle statistics: " << filename << std::endl;
    return;
  }
  outfile << "Frame,MergedParticles,SplitParticles,TotalParticles" << std::endl;
  for (size_t i = 0; i < _particle_stats.size(); i++) {
    const auto &st = _particle_stats[i];
    outfile << i << "," << std::get<0>(st) << "," << std::get<1>(st) << "," << std::get<2>(st)
            << std::endl;
  }
  outfile.close();
  std::cout << "Particle statistics saved to " << filename << std::endl;
}
