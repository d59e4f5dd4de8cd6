// Headless scene runner written against the host C++ API the way the
// reference's main.cu uses its classes: load a scene JSON, build the
// granular lattice, the walls and 25 jittered fine particles per granular
// particle, construct the GranularSystem (which runs one step), step it,
// then save the particle statistics and the frame times.  EXCAVATOR and
// FUNNEL scenes add the obj_file point cloud as an animated boundary; with
// "animate" it moves by updateAnimation before every step (the reference
// ships with that call switched off).
//
//   granular_demo <scene.json> [frames] [out_dir] [animate]
#include <hip/hip_runtime_api.h>

#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "granular/Animation.hpp"
#include "granular/GranularSystem.hpp"
#include "granular/SceneConfig.hpp"

static const float upsampled_particle_radius = 0.005f;  // main.cu:42
static const float particle_radius = 0.01f;             // main.cu:46

int main(int argc, char **argv) {
  if (argc < 2) {
    std::cerr << "usage: " << argv[0] << " <scene.json> [frames] [out_dir] [animate]\n";
    return 2;
  }
  const int frames = argc > 2 ? std::stoi(argv[2]) : 100;
  const std::string out_dir = argc > 3 ? argv[3] : ".";
  const bool animate = argc > 4 && std::string(argv[4]) == "animate";
  try {
    SceneConfig cfg = loadSceneConfig(argv[1]);
    std::vector<float3> pos = granular_lattice(cfg);
    auto granular_particles = std::make_shared<GranularParticles>(pos);
    auto upsampled_particles =
        std::make_shared<GranularParticles>(upsampled_cloud(pos, particle_radius, 25));
    SceneBoundary b = scene_boundary(cfg);
    auto boundary_particles = std::make_shared<GranularParticles>(b.pos);
    if (!b.animated.empty() &&
        hipMemcpy(boundary_particles->get_is_animated_ptr(), b.animated.data(),
                  sizeof(int) * b.animated.size(), hipMemcpyHostToDevice) != hipSuccess)
      throw std::runtime_error("copy of the animated markers failed");
    // the reference sets is_move for EXCAVATOR only (main.cu:983); a boundary
    // moved by "animate" is re-sorted every step as well
    const bool is_move = cfg.scene == EXCAVATOR || animate;
    std::cout << "granular " << granular_particles->size() << ", boundary "
              << boundary_particles->size() << " (animated " << b.anim.num_particles
              << "), fine " << upsampled_particles->size()
              << ", cells " << cfg.cell_size.x << "x" << cfg.cell_size.y << "x"
              << cfg.cell_size.z << "\n";
    auto p_system = std::make_shared<GranularSystem>(
        granular_particles, boundary_particles, upsampled_particles, cfg.space_size,
        cfg.cell_length, cfg.dt, cfg.G, cfg.cell_size, cfg.density, upsampled_particle_radius,
        is_move, cfg.is_adaptive);
    if (granular_particles || boundary_particles || upsampled_particles) {
      std::cerr << "ownership was not moved into the system\n";
      return 1;
    }
    for (int f = 1; f < frames; ++f) {
      if (animate) granular_animation::update_animation(*p_system, b.anim, cfg.dt);
      p_system->step();
    }
    p_system->get_solver().save_particle_stats(out_dir + "/stats.csv");
    std::ofstream ft(out_dir + "/frame_times.txt");
    for (float t : p_system->get_frame_times()) ft << t << "\n";
    std::cout << "final granular count " << p_system->size() << "\n";
    return 0;
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
}
