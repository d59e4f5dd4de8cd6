// Particle set with positions and velocities (Particles.hpp:7-62 of the
// reference): built from host positions, velocities start at zero.
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "DArray.hpp"
#include "Global.hpp"

class Particles {
 public:
  explicit Particles(const std::vector<float3> &p)
      : _pos((unsigned int)p.size()), _vel((unsigned int)p.size()) {
    if (!p.empty())
      HIP_CALL(hipMemcpy(_pos.addr(), p.data(), sizeof(float3) * p.size(),
                         hipMemcpyHostToDevice));
  }

  Particles(const Particles &) = delete;
  Particles &operator=(const Particles &) = delete;

  unsigned int size() const { return _pos.length(); }
  float3 *get_pos_ptr() const { return _pos.addr(); }
  float3 *get_vel_ptr() const { return _vel.addr(); }
  const DArray<float3> &get_pos() const { return _pos; }

  void remove_elements(const DArray<int> &removal_flags) {
    _pos.compact(removal_flags);
    _vel.compact(removal_flags);
  }

  void add_elements(const DArray<float3> &pos, const DArray<float3> &vel) {
    if (pos.length() != vel.length())
      throw std::runtime_error("Particles::add_elements: " + std::to_string(pos.length()) +
                               " positions but " + std::to_string(vel.length()) + " velocities");
    _pos.reserve(_pos.length() + pos.length());
    _vel.reserve(_vel.length() + vel.length());
    _pos.append(pos);
    _vel.append(vel);
    if (_pos.length() != _vel.length())
      throw std::runtime_error("Particles::add_elements: position and velocity lengths differ "
                               "after the append");
  }

  virtual ~Particles() noexcept {}

 public:
  DArray<float3> _pos;
  DArray<float3> _vel;
};
