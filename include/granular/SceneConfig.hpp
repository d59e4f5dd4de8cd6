// Scene JSON reader and scene builders of the host API.
//
// SceneConfig / loadSceneConfig follow main.cu:81-98 and 123-167 of the
// reference: the same keys (space_size, dt, gravity, box_size, sph_spacing,
// init_spacing, smoothing_radius, cell_length, density, scene,
// boundary_translation, particle_translation, obj_file, is_adaptive,
// save_file), JSON numbers read as double and rounded to float,
// cell_size = ceil(space_size / cell_length) per axis, std::runtime_error on
// an unreadable file or a missing key.  The parser is this library's own
// (the reference uses nlohmann_json).
#pragma once
#include <string>
#include <vector>

#include "Global.hpp"

struct SceneConfig {
  float3 space_size;
  float dt;
  float3 G;
  float sphSpacing;
  float initSpacing;
  float smoothing_radius;
  float cell_length;
  int3 cell_size;
  float density;
  float3 box_size;
  float3 boundary_translation;
  float3 particle_translation;
  int scene;
  std::string obj_file;
  int is_adaptive;
  std::string save_file;
};

enum Scene { PILING = 0, BOX = 1, EXCAVATOR = 2, FUNNEL = 3, CORNER = 4, BOX_OPENING = 5 };

SceneConfig loadSceneConfig(const std::string &config_file);
SceneConfig parseSceneConfig(const std::string &json_text);

// Whitespace-separated x y z triples (main.cu:169-188); start_index = count.
std::vector<float3> readBoundaryParticlesFromFile(const std::string &filename, int &start_index);

// Granular lattice of init_granular_system (main.cu:747-760): box_size.x x
// box_size.y x box_size.z particles, (0.27 + s j, 0.13 + s i, 0.17 + s k) +
// particle_translation with s = init_spacing, loops i (x count), j, k.
std::vector<float3> granular_lattice(const SceneConfig &c);

// Wall particles of the six recipes (main.cu:822-1167) in the reference's
// insertion order.  PILING, BOX, CORNER and BOX_OPENING are translated by
// boundary_translation; EXCAVATOR and FUNNEL build the box without its top
// (y = max) wall and without the translation (main.cu:889-925, 985-1021).
std::vector<float3> boundary_walls(const SceneConfig &c);

// Kinematic motion of the animated boundary particles.  The fields and their
// defaults are the reference's AnimationState (main.cu:100-111);
// translation_time and rotation_limit are the reference's hard-coded 2.5 s
// and 45 degrees of updateAnimation (main.cu:1219, 1255-1262), made
// settable for scripted moves.  The rotation centre is recomputed, when the
// translation ends, over the index range [start_index, start_index +
// num_particles) as the reference does; the boundary is cell-sorted by then,
// so that range no longer holds the cloud.  center_from_marked = true takes
// the centroid of the particles marked animated instead.
struct AnimationState {
  bool is_animating = false;
  float animation_time = 0.0f;
  int start_index = 0;
  int num_particles = 0;
  float3 translation_direction = make_float3(-0.1f, 0.0f, 0.0f);
  float translation_speed = 5.0f;
  bool translation_complete = false;
  float rotation_angle = 0.0f;
  float rotation_speed = 90.0f;
  float3 rotation_center = make_float3(0.0f, 0.0f, 0.0f);
  float translation_time = 2.5f;
  float rotation_limit = 45.0f;
  bool center_from_marked = false;
};

// The boundary set of a scene: walls, then (EXCAVATOR, FUNNEL) the point
// cloud of obj_file if that file exists (main.cu:927-968, 1023-1075):
//  - EXCAVATOR: cloud + (2.5, 0, 0.5); rotation_center = centroid of the
//    cloud as read;
//  - FUNNEL: cloud rotated 90 degrees about the x axis through its
//    centroid (cosf/sinf of the float-rounded angle), then + (1.25, 0.3, 1.5).
// Cloud particles are marked animated (1), walls 0; `anim` receives
// start_index = wall count, num_particles = cloud size, is_animating = true.
// A missing obj_file leaves walls only (the reference skips it silently).
struct SceneBoundary {
  std::vector<float3> pos;
  std::vector<int> animated;
  AnimationState anim;
};
SceneBoundary scene_boundary(const SceneConfig &c);

// `per` fine particles per coarse particle, offsets drawn per axis from
// std::uniform_real_distribution<float>(-radius, radius) on one
// std::default_random_engine in x, y, z order (main.cu:804-816).
std::vector<float3> upsampled_cloud(const std::vector<float3> &coarse, float radius, int per = 25);
