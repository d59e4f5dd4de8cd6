// C-ABI over the host C++ API (handles for ctypes and other FFI callers).
#include "granular_host.h"

#include <algorithm>
#include <cstring>
#include <exception>
#include <iostream>
#include <sstream>
#include <memory>
#include <string>
#include <vector>

#include "granular/Animation.hpp"
#include "granular/GranularSystem.hpp"
#include "granular/Output.hpp"
#include "granular/SceneConfig.hpp"

struct ggh_system {
  std::unique_ptr<GranularSystem> sys;
};

namespace {

thread_local std::string g_err;

int fail(const std::exception &e) {
  g_err = e.what();
  return -1;
}

void copy_str(char *dst, size_t cap, const std::string &s) {
  std::strncpy(dst, s.c_str(), cap - 1);
  dst[cap - 1] = '\0';
}

void to_c(const SceneConfig &c, ggh_scene *o) {
  std::memset(o, 0, sizeof(*o));
  o->space_size[0] = c.space_size.x;
  o->space_size[1] = c.space_size.y;
  o->space_size[2] = c.space_size.z;
  o->dt = c.dt;
  o->gravity[0] = c.G.x;
  o->gravity[1] = c.G.y;
  o->gravity[2] = c.G.z;
  o->sph_spacing = c.sphSpacing;
  o->init_spacing = c.initSpacing;
  o->smoothing_radius = c.smoothing_radius;
  o->cell_length = c.cell_length;
  o->cell_size[0] = c.cell_size.x;
  o->cell_size[1] = c.cell_size.y;
  o->cell_size[2] = c.cell_size.z;
  o->density = c.density;
  o->box_size[0] = c.box_size.x;
  o->box_size[1] = c.box_size.y;
  o->box_size[2] = c.box_size.z;
  o->boundary_translation[0] = c.boundary_translation.x;
  o->boundary_translation[1] = c.boundary_translation.y;
  o->boundary_translation[2] = c.boundary_translation.z;
  o->particle_translation[0] = c.particle_translation.x;
  o->particle_translation[1] = c.particle_translation.y;
  o->particle_translation[2] = c.particle_translation.z;
  o->scene = c.scene;
  o->is_adaptive = c.is_adaptive;
  copy_str(o->obj_file, sizeof(o->obj_file), c.obj_file);
  copy_str(o->save_file, sizeof(o->save_file), c.save_file);
}

SceneConfig from_c(const ggh_scene *s) {
  SceneConfig c;
  c.space_size = make_float3(s->space_size[0], s->space_size[1], s->space_size[2]);
  c.dt = s->dt;
  c.G = make_float3(s->gravity[0], s->gravity[1], s->gravity[2]);
  c.sphSpacing = s->sph_spacing;
  c.initSpacing = s->init_spacing;
  c.smoothing_radius = s->smoothing_radius;
  c.cell_length = s->cell_length;
  c.cell_size = make_int3(s->cell_size[0], s->cell_size[1], s->cell_size[2]);
  c.density = s->density;
  c.box_size = make_float3(s->box_size[0], s->box_size[1], s->box_size[2]);
  c.boundary_translation =
      make_float3(s->boundary_translation[0], s->boundary_translation[1], s->boundary_translation[2]);
  c.particle_translation =
      make_float3(s->particle_translation[0], s->particle_translation[1], s->particle_translation[2]);
  c.scene = s->scene;
  c.obj_file = s->obj_file;
  c.is_adaptive = s->is_adaptive;
  c.save_file = s->save_file;
  return c;
}

void to_c(const AnimationState &a, ggh_animation *o) {
  o->is_animating = a.is_animating ? 1 : 0;
  o->animation_time = a.animation_time;
  o->start_index = a.start_index;
  o->num_particles = a.num_particles;
  o->translation_direction[0] = a.translation_direction.x;
  o->translation_direction[1] = a.translation_direction.y;
  o->translation_direction[2] = a.translation_direction.z;
  o->translation_speed = a.translation_speed;
  o->translation_complete = a.translation_complete ? 1 : 0;
  o->rotation_angle = a.rotation_angle;
  o->rotation_speed = a.rotation_speed;
  o->rotation_center[0] = a.rotation_center.x;
  o->rotation_center[1] = a.rotation_center.y;
  o->rotation_center[2] = a.rotation_center.z;
  o->translation_time = a.translation_time;
  o->rotation_limit = a.rotation_limit;
  o->center_from_marked = a.center_from_marked ? 1 : 0;
}

AnimationState from_c(const ggh_animation *o) {
  AnimationState a;
  a.is_animating = o->is_animating != 0;
  a.animation_time = o->animation_time;
  a.start_index = o->start_index;
  a.num_particles = o->num_particles;
  a.translation_direction = make_float3(o->translation_direction[0], o->translation_direction[1],
                                        o->translation_direction[2]);
  a.translation_speed = o->translation_speed;
  a.translation_complete = o->translation_complete != 0;
  a.rotation_angle = o->rotation_angle;
  a.rotation_speed = o->rotation_speed;
  a.rotation_center =
      make_float3(o->rotation_center[0], o->rotation_center[1], o->rotation_center[2]);
  a.translation_time = o->translation_time;
  a.rotation_limit = o->rotation_limit;
  a.center_from_marked = o->center_from_marked != 0;
  return a;
}

int emit(const std::vector<float3> &v, float *out, int max) {
  if (out) {
    const size_t k = std::min(v.size(), (size_t)(max > 0 ? max : 0));
    std::memcpy(out, v.data(), k * sizeof(float3));
  }
  return (int)v.size();
}

std::vector<float3> triples(const float *p, int n) {
  std::vector<float3> v((size_t)(n > 0 ? n : 0));
  if (n > 0) std::memcpy(v.data(), p, (size_t)n * sizeof(float3));
  return v;
}

std::shared_ptr<GranularParticles> set_of(const ggh_system *h, int which) {
  switch (which) {
    case 0: return h->sys->get_particles_non_const();
    case 1: return h->sys->get_boundaries();
    case 2: return h->sys->get_upsampled();
    default: throw std::invalid_argument("set index must be 0, 1 or 2");
  }
}

}  // namespace

extern "C" {

const char *ggh_last_error(void) { return g_err.c_str(); }

int ggh_parse_scene(const char *json_text, ggh_scene *out) {
  try {
    if (!json_text || !out) throw std::invalid_argument("null argument");
    to_c(parseSceneConfig(json_text), out);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_load_scene(const char *path, ggh_scene *out) {
  try {
    if (!path || !out) throw std::invalid_argument("null argument");
    to_c(loadSceneConfig(path), out);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_scene_lattice(const ggh_scene *s, float *out, int max) {
  try {
    if (!s) throw std::invalid_argument("null scene");
    return emit(granular_lattice(from_c(s)), out, max);
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_scene_walls(const ggh_scene *s, float *out, int max) {
  try {
    if (!s) throw std::invalid_argument("null scene");
    return emit(boundary_walls(from_c(s)), out, max);
  } catch (const std::exception &e) {
    return fail(e);
  }
}

void ggh_animation_default(ggh_animation *a) {
  if (a) to_c(AnimationState{}, a);
}

int ggh_scene_boundary(const ggh_scene *s, float *out, int *animated, int max,
                       ggh_animation *anim) {
  try {
    if (!s) throw std::invalid_argument("null scene");
    const SceneBoundary b = scene_boundary(from_c(s));
    emit(b.pos, out, max);
    if (animated) {
      const size_t k = std::min(b.animated.size(), (size_t)(max > 0 ? max : 0));
      std::copy(b.animated.begin(), b.animated.begin() + (long)k, animated);
    }
    if (anim) to_c(b.anim, anim);
    return (int)b.pos.size();
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_upsampled_cloud(const float *coarse, int n, float radius, int per, float *out, int max) {
  try {
    if (n < 0 || per < 0 || (n > 0 && !coarse)) throw std::invalid_argument("bad arguments");
    return emit(upsampled_cloud(triples(coarse, n), radius, per), out, max);
  } catch (const std::exception &e) {
    return fail(e);
  }
}

ggh_system *ggh_system_create(const float *pos, int n, const float *boundary, int nb,
                              const int *boundary_animated, const float *fine, int nf,
                              const float space_size[3], float cell_length, float dt,
                              const float gravity[3], const int cell_size[3], float density,
                              float upsampled_radius, int is_move_boundary, int is_adaptive,
                              int verbose) {
  try {
    if (n < 0 || nb < 0 || nf < 0 || !space_size || !gravity || !cell_size)
      throw std::invalid_argument("bad arguments");
    auto gp = std::make_shared<GranularParticles>(triples(pos, n));
    auto bp = std::make_shared<GranularParticles>(triples(boundary, nb));
    auto fp = std::make_shared<GranularParticles>(triples(fine, nf));
    if (boundary_animated && nb > 0)
      HIP_CALL(hipMemcpy(bp->get_is_animated_ptr(), boundary_animated, sizeof(int) * (size_t)nb,
                         hipMemcpyHostToDevice));
    std::unique_ptr<ggh_system> h(new ggh_system);
    // the constructor runs one step; print nothing for FFI callers unless asked
    std::streambuf *saved = nullptr;
    std::ostringstream sink;
    if (!verbose) saved = std::cout.rdbuf(sink.rdbuf());
    try {
      h->sys.reset(new GranularSystem(
          gp, bp, fp, make_float3(space_size[0], space_size[1], space_size[2]), cell_length, dt,
          make_float3(gravity[0], gravity[1], gravity[2]),
          make_int3(cell_size[0], cell_size[1], cell_size[2]), density, upsampled_radius,
          is_move_boundary != 0, is_adaptive));
    } catch (...) {
      if (saved) std::cout.rdbuf(saved);
      throw;
    }
    if (saved) std::cout.rdbuf(saved);
    h->sys->set_verbose(verbose != 0);
    return h.release();
  } catch (const std::exception &e) {
    fail(e);
    return nullptr;
  }
}

void ggh_system_destroy(ggh_system *h) { delete h; }

int ggh_system_step(ggh_system *h, int nsteps) {
  try {
    if (!h || nsteps < 0) throw std::invalid_argument("bad arguments");
    int r = 0;
    for (int k = 0; k < nsteps; ++k) r = (int)h->sys->step();
    return r;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_size(const ggh_system *h, int which) {
  if (!h) return -1;
  switch (which) {
    case 0: return h->sys->size();
    case 1: return h->sys->boundary_size();
    case 2: return h->sys->upsampled_size();
    case 3: return h->sys->total_size();
    default: return -1;
  }
}

int ggh_system_read(const ggh_system *h, int which, int field, void *out, int count) {
  try {
    if (!h || !out || count < 0) throw std::invalid_argument("bad arguments");
    auto s = set_of(h, which);
    if ((unsigned int)count > s->size()) throw std::invalid_argument("count exceeds set size");
    const void *src = nullptr;
    size_t esz = 4;
    switch (field) {
      case 0: src = s->get_pos_ptr(); esz = 12; break;
      case 1: src = s->get_vel_ptr(); esz = 12; break;
      case 2: src = s->get_mass_ptr(); break;
      case 3: src = s->get_particle_2_cell(); break;
      case 4: src = s->get_surface_ptr(); break;
      case 5: src = s->get_surface_distance_ptr(); break;
      case 6: src = s->get_adaptive_last_step_ptr(); break;
      case 7: src = s->get_is_animated_ptr(); break;
      default: throw std::invalid_argument("unknown field");
    }
    if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("device error");
    if (count > 0 && hipMemcpy(out, src, esz * (size_t)count, hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("copy failed");
    return count;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_translate_animated(ggh_system *h, float dx, float dy, float dz) {
  try {
    if (!h) throw std::invalid_argument("null handle");
    auto b = h->sys->get_boundaries();
    const unsigned int nb = b->size();
    if (nb == 0) return 0;
    std::vector<int> a(nb);
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(a.data(), b->get_is_animated_ptr(), 4 * (size_t)nb, hipMemcpyDeviceToHost) !=
            hipSuccess)
      throw std::runtime_error("copy failed");
    granular_animation::animate_boundary(*h->sys, make_float3(dx, dy, dz), 0.0f,
                                         make_float3(0.0f, 0.0f, 0.0f));
    return (int)std::count(a.begin(), a.end(), 1);
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_animate(ggh_system *h, const float translation[3], float angle_deg,
                       const float center[3]) {
  try {
    if (!h || !translation || !center) throw std::invalid_argument("null argument");
    granular_animation::animate_boundary(*h->sys,
                                         make_float3(translation[0], translation[1], translation[2]),
                                         angle_deg, make_float3(center[0], center[1], center[2]));
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_explode(ggh_system *h, float force, float com_out[3]) {
  try {
    if (!h) throw std::invalid_argument("null handle");
    auto p = h->sys->get_particles_non_const();
    h->sys->get_solver().trigger_explosion(p, force);
    const float3 c = h->sys->get_solver().get_explosion_center();
    if (com_out) {
      com_out[0] = c.x;
      com_out[1] = c.y;
      com_out[2] = c.z;
    }
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_pick(const ggh_system *h, const float center[3], float radius, int *indices,
                    int max) {
  try {
    if (!h || !center) throw std::invalid_argument("null argument");
    auto p = h->sys->get_particles_non_const();
    const int n = (int)p->size();
    if (n == 0) return 0;
    DArray<int> idx((unsigned int)n), count(1u);
    DArray<unsigned int> scratch((unsigned int)((gg_pick_scratch_bytes(n) + 3) / 4));
    const gg_particles s = p->view();
    const int rc = gg_pick_sphere(&s, gg_float3{center[0], center[1], center[2]}, radius,
                                  idx.addr(), count.addr(), scratch.addr(),
                                  4 * (size_t)scratch.length(), nullptr);
    if (rc != GG_OK) throw std::runtime_error(std::string("gg_pick_sphere: ") + gg_status_string(rc));
    int c = 0;
    if (hipMemcpy(&c, count.addr(), sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("copy failed");
    if (indices && max > 0 && c > 0 &&
        hipMemcpy(indices, idx.addr(), sizeof(int) * (size_t)std::min(c, max),
                  hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("copy failed");
    return c;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_move_picked(ggh_system *h, const int *indices, int count, const float delta[3],
                           const float lo[3], const float hi[3]) {
  try {
    if (!h || !delta || !lo || !hi || count < 0 || (count > 0 && !indices))
      throw std::invalid_argument("bad arguments");
    auto p = h->sys->get_particles_non_const();
    const int n = (int)p->size();
    if (count > n) throw std::invalid_argument("more picked indices than particles");
    for (int k = 0; k < count; ++k)
      if (indices[k] < 0 || indices[k] >= n) throw std::out_of_range("picked index out of range");
    if (count == 0) return 0;
    DArray<int> idx((unsigned int)count), dc(1u);
    if (hipMemcpy(idx.addr(), indices, sizeof(int) * (size_t)count, hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(dc.addr(), &count, sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      throw std::runtime_error("copy failed");
    gg_particles s = p->view();
    const int rc = gg_move_picked(&s, idx.addr(), dc.addr(), count,
                                  gg_float3{delta[0], delta[1], delta[2]},
                                  gg_float3{lo[0], lo[1], lo[2]}, gg_float3{hi[0], hi[1], hi[2]},
                                  nullptr);
    if (rc != GG_OK || hipDeviceSynchronize() != hipSuccess)
      throw std::runtime_error(std::string("gg_move_picked: ") + gg_status_string(rc));
    return count;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_update_animation(ggh_system *h, ggh_animation *a, float dt) {
  try {
    if (!h || !a) throw std::invalid_argument("null argument");
    AnimationState s = from_c(a);
    const bool moved = granular_animation::update_animation(*h->sys, s, dt);
    to_c(s, a);
    return moved ? 1 : 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_frame_times(const ggh_system *h, float *out, int max) {
  if (!h) return -1;
  const auto &ft = h->sys->get_frame_times();
  if (out)
    for (int i = 0; i < (int)ft.size() && i < max; ++i) out[i] = ft[(size_t)i];
  return (int)ft.size();
}

int ggh_system_particle_stats(const ggh_system *h, int *out, int max_rows) {
  if (!h) return -1;
  const auto &st = h->sys->get_solver().get_particle_stats();
  if (out)
    for (int i = 0; i < (int)st.size() && i < max_rows; ++i) {
      out[3 * i] = std::get<0>(st[(size_t)i]);
      out[3 * i + 1] = std::get<1>(st[(size_t)i]);
      out[3 * i + 2] = std::get<2>(st[(size_t)i]);
    }
  return (int)st.size();
}

int ggh_system_sweep_scans(const ggh_system *h, int *out, int max) {
  try {
    if (!h) throw std::invalid_argument("bad arguments");
    return h->sys->get_solver().sweep_scan_counts(out, max);
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_save_stats(ggh_system *h, const char *path) {
  try {
    if (!h || !path) throw std::invalid_argument("bad arguments");
    h->sys->get_solver().save_particle_stats(path);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_set_solver(ggh_system *h, int stab, int solver, float mu_s, float mu_k,
                          int flags) {
  try {
    if (!h) throw std::invalid_argument("null handle");
    Solver &s = h->sys->get_solver();
    s.set_iterations(stab, solver);
    s.set_friction(mu_s, mu_k);
    s.set_flags(flags);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_set_lists(ggh_system *h, int gmax, int bmax, float skin) {
  try {
    if (!h) throw std::invalid_argument("null handle");
    h->sys->get_solver().set_neighbor_lists(gmax, bmax, skin);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_set_merge_rules(ggh_system *h, int depth_rules, float min_mass, float split_mass,
                               float merge_dist, float merge_depth, float split_depth,
                               float max_mass) {
  try {
    if (!h) throw std::invalid_argument("null handle");
    h->sys->set_merge_rules(depth_rules, min_mass, split_mass, merge_dist, merge_depth,
                            split_depth, max_mass);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_set_learned(ggh_system *h, const unsigned short *w_bf16, const float *bias,
                           float radius, float vscale, float guard_boundary, float guard_floor,
                           float guard_side) {
  try {
    if (!h) throw std::invalid_argument("null handle");
    gg_mlp_config c;
    gg_mlp_config_default(&c);
    c.radius = radius;
    c.vscale = vscale;
    c.guard_boundary = guard_boundary;
    c.guard_floor = guard_floor;
    c.guard_side = guard_side;
    h->sys->set_learned_upsampler(reinterpret_cast<const uint16_t *>(w_bf16), bias, c);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_write_vtk(const char *path, const char *title, const float *xyz, int n) {
  try {
    if (!path || !title || n < 0 || (n > 0 && !xyz)) throw std::invalid_argument("bad arguments");
    return granular_output::write_vtk_points(path, title, triples(xyz, n));
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_save_upsampled_vtk(const ggh_system *h, const char *save_file, int is_adaptive,
                                  int frame) {
  try {
    if (!h || !save_file) throw std::invalid_argument("bad arguments");
    return granular_output::save_upsampled_vtk(*h->sys->get_upsampled(), save_file,
                                               is_adaptive != 0, frame);
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_save_boundary_vtk(const ggh_system *h, const char *save_file, int frame,
                                 int centred) {
  try {
    if (!h || !save_file) throw std::invalid_argument("bad arguments");
    const auto fine = h->sys->get_upsampled();
    return granular_output::save_animated_boundary_vtk(*h->sys->get_boundaries(),
                                                       centred ? fine.get() : nullptr,
                                                       save_file, frame);
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_save_mass_vtk(const ggh_system *h, const char *save_file, int frame,
                             int counts[3]) {
  try {
    if (!h || !save_file || !counts) throw std::invalid_argument("bad arguments");
    granular_output::save_particles_by_mass_vtk(*h->sys->get_particles(), save_file, frame,
                                                counts);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_checkpoint_write(const char *path, const float *pos, const float *vel,
                         const float *mass, const int *adaptive_last_step, int n,
                         const float *boundary, const int *boundary_animated, int nb,
                         const float *fine_pos, const float *fine_vel, int nf) {
  try {
    if (!path || n < 0 || nb < 0 || nf < 0) throw std::invalid_argument("bad arguments");
    if ((n > 0 && (!pos || !vel || !mass || !adaptive_last_step)) ||
        (nb > 0 && (!boundary || !boundary_animated)) || (nf > 0 && (!fine_pos || !fine_vel)))
      throw std::invalid_argument("null array");
    granular_output::Checkpoint c;
    c.pos = triples(pos, n);
    c.vel = triples(vel, n);
    c.mass.assign(mass, mass + n);
    c.adaptive_last_step.assign(adaptive_last_step, adaptive_last_step + n);
    c.boundary_pos = triples(boundary, nb);
    c.boundary_animated.assign(boundary_animated, boundary_animated + nb);
    c.fine_pos = triples(fine_pos, nf);
    c.fine_vel = triples(fine_vel, nf);
    granular_output::write_checkpoint(path, c);
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_system_save_checkpoint(const ggh_system *h, const char *path) {
  try {
    if (!h || !path) throw std::invalid_argument("bad arguments");
    granular_output::write_checkpoint(path, granular_output::checkpoint_of(*h->sys));
    return 0;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

int ggh_checkpoint_read(const char *path, int which, int field, void *out, int max) {
  try {
    if (!path || max < 0) throw std::invalid_argument("bad arguments");
    const granular_output::Checkpoint c = granular_output::read_checkpoint(path);
    const void *src = nullptr;
    size_t count = 0, esz = 4;
    auto pick = [&](const void *p, size_t n, size_t e) {
      src = p;
      count = n;
      esz = e;
    };
    if (which == 0 && field == 0) pick(c.pos.data(), c.pos.size(), 12);
    else if (which == 0 && field == 1) pick(c.vel.data(), c.vel.size(), 12);
    else if (which == 0 && field == 2) pick(c.mass.data(), c.mass.size(), 4);
    else if (which == 0 && field == 6)
      pick(c.adaptive_last_step.data(), c.adaptive_last_step.size(), 4);
    else if (which == 1 && field == 0) pick(c.boundary_pos.data(), c.boundary_pos.size(), 12);
    else if (which == 1 && field == 7)
      pick(c.boundary_animated.data(), c.boundary_animated.size(), 4);
    else if (which == 2 && field == 0) pick(c.fine_pos.data(), c.fine_pos.size(), 12);
    else if (which == 2 && field == 1) pick(c.fine_vel.data(), c.fine_vel.size(), 12);
    else throw std::invalid_argument("unknown set or field");
    if (out && count > 0) std::memcpy(out, src, esz * std::min(count, (size_t)max));
    return (int)count;
  } catch (const std::exception &e) {
    return fail(e);
  }
}

}  // extern "C"
