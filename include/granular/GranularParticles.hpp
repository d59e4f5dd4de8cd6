// Granular particle set (GranularParticles.hpp:5-110 of the reference):
// positions and velocities plus mass, scaled mass, surface flag, distance
// to the surface, surface-neighbour count, cell key, animation flag and the
// adaptive_last_step flag.
//
// remove_elements compacts mass, adaptive_last_step, pos and vel only, and
// add_elements appends mass, adaptive_last_step = 1, pos and vel, as in the
// reference; the other arrays keep their constructed length.
#pragma once
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "DArray.hpp"
#include "Particles.hpp"

class GranularParticles final : public Particles {
 public:
  explicit GranularParticles(const std::vector<float3> &p)
      : Particles(p),
        _mass((unsigned int)p.size()),
        _scaled_mass((unsigned int)p.size()),
        _surface((unsigned int)p.size()),
        _surface_distance((unsigned int)p.size()),
        _num_surface((unsigned int)p.size()),
        _particle_2_cell((unsigned int)p.size()),
        _is_animated((unsigned int)p.size()),
        _adaptive_last_step((unsigned int)p.size()) {}

  GranularParticles(const GranularParticles &) = delete;
  GranularParticles &operator=(const GranularParticles &) = delete;

  int *get_particle_2_cell() const { return _particle_2_cell.addr(); }
  float *get_mass_ptr() const { return _mass.addr(); }
  float *get_scaled_mass_ptr() const { return _scaled_mass.addr(); }
  int *get_surface_ptr() const { return _surface.addr(); }
  float *get_surface_distance_ptr() const { return _surface_distance.addr(); }
  int *get_num_surface_ptr() const { return _num_surface.addr(); }
  int *get_adaptive_last_step_ptr() const { return _adaptive_last_step.addr(); }
  int *get_is_animated_ptr() { return _is_animated.addr(); }
  const DArray<float> &get_mass() const { return _mass; }

  // mass and adaptive_last_step report a flag-length mismatch on stderr and
  // stay as they are; pos and vel (Particles::remove_elements) throw it.
  void remove_elements(const DArray<int> &removal_flags) {
    compact_or_report(_mass, removal_flags, "mass");
    compact_or_report(_adaptive_last_step, removal_flags, "adaptive_last_step");
    Particles::remove_elements(removal_flags);
  }

  void add_elements(const DArray<float> &mass, const DArray<float3> &pos,
                    const DArray<float3> &vel, const int num) {
    if ((int)mass.length() != num || (int)pos.length() != num || (int)vel.length() != num)
      throw std::runtime_error("GranularParticles::add_elements: count " + std::to_string(num) +
                               " but " + std::to_string(mass.length()) + " masses, " +
                               std::to_string(pos.length()) + " positions, " +
                               std::to_string(vel.length()) + " velocities");
    _mass.reserve(_mass.length() + (unsigned int)num);
    _mass.append(mass);
    DArray<int> ones((unsigned int)num);
    if (num > 0) {
      const int rc = gg_fill_u32(ones.addr(), (size_t)num, 1u, nullptr);
      if (rc != GG_OK)
        throw std::runtime_error(std::string("GranularParticles::add_elements: flag fill: ") +
                                 gg_status_string(rc));
    }
    _adaptive_last_step.append(ones);
    Particles::add_elements(pos, vel);
  }

  // Grows every array's capacity to at least cap (lengths unchanged), so
  // that device passes may work in place on up to cap particles.
  void reserve(unsigned int cap) {
    _pos.reserve(cap);
    _vel.reserve(cap);
    _mass.reserve(cap);
    _scaled_mass.reserve(cap);
    _surface.reserve(cap);
    _surface_distance.reserve(cap);
    _num_surface.reserve(cap);
    _particle_2_cell.reserve(cap);
    _is_animated.reserve(cap);
    _adaptive_last_step.reserve(cap);
  }

  // Smallest capacity over the arrays that device passes index by particle.
  unsigned int device_capacity() const {
    unsigned int c = _pos.capacity();
    for (unsigned int v : {_vel.capacity(), _mass.capacity(), _scaled_mass.capacity(),
                           _surface.capacity(), _surface_distance.capacity(),
                           _num_surface.capacity(), _particle_2_cell.capacity(),
                           _is_animated.capacity(), _adaptive_last_step.capacity()})
      c = std::min(c, v);
    return c;
  }

  // After an in-place device pass changed N (merge/split compaction): the
  // arrays the reference compacts and appends take the new length.
  void set_size(unsigned int n) {
    _pos.set_length(n);
    _vel.set_length(n);
    _mass.set_length(n);
    _adaptive_last_step.set_length(n);
  }

  // The library's view of this set.
  gg_particles view() {
    gg_particles s;
    s.n = (int)size();
    s.capacity = (int)device_capacity();
    s.pos = reinterpret_cast<gg_float3 *>(_pos.addr());
    s.vel = reinterpret_cast<gg_float3 *>(_vel.addr());
    s.mass = _mass.addr();
    s.scaled_mass = _scaled_mass.addr();
    s.surface = _surface.addr();
    s.surface_distance = _surface_distance.addr();
    s.num_surface = _num_surface.addr();
    s.particle_2_cell = _particle_2_cell.addr();
    s.is_animated = _is_animated.addr();
    s.adaptive_last_step = _adaptive_last_step.addr();
    return s;
  }

  virtual ~GranularParticles() noexcept {}

 protected:
  template <typename T>
  static void compact_or_report(DArray<T> &a, const DArray<int> &flags, const char *name) {
    try {
      a.compact(flags);
    } catch (const std::exception &e) {
      std::cerr << "GranularParticles::remove_elements left " << name << " unchanged: "
                << e.what() << std::endl;
    }
  }

  DArray<float> _mass;
  DArray<float> _scaled_mass;
  DArray<int> _surface;
  DArray<float> _surface_distance;
  DArray<int> _num_surface;
  DArray<int> _particle_2_cell;
  DArray<int> _is_animated;
  DArray<int> _adaptive_last_step;
};
