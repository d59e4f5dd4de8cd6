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

// Wall particles of the PILING, BOX, CORNER and BOX_OPENING recipes
// (main.cu:822-1167) in the reference's insertion order, translated by
// boundary_translation.  EXCAVATOR and FUNNEL need external geometry and
// throw std::runtime_error.
std::vector<float3> boundary_walls(const SceneConfig &c);

// `per` fine particles per coarse particle, offsets drawn per axis from
// std::uniform_real_distribution<float>(-radius, radius) on one
// std::default_random_engine in x, y, z order (main.cu:804-816).
std::vector<float3> upsampled_cloud(const std::vector<float3> &coarse, float radius, int per = 25);
