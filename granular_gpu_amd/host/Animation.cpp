// Kinematic boundary animation (main.cu:1175-1277 of the reference).
#include "granular/Animation.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "granular/GranularParticles.hpp"
#include "granular/GranularSystem.hpp"
#include "granular_gpu.h"

namespace granular_animation {

void animate_boundary(GranularSystem &sys, float3 translation, float angle_deg, float3 center) {
  gg_particles b = sys.get_boundaries()->view();
  const gg_float3 t{translation.x, translation.y, translation.z};
  const gg_float3 c{center.x, center.y, center.z};
  const int rc = gg_boundary_animate(&b, t, angle_deg, c, nullptr);
  if (rc != GG_OK)
    throw std::runtime_error(std::string("gg_boundary_animate: ") + gg_status_string(rc));
  if (hipStreamSynchronize(nullptr) != hipSuccess)
    throw std::runtime_error("gg_boundary_animate: device error");
}

bool update_animation(GranularSystem &sys, AnimationState &a, float dt) {
  if (!a.is_animating) return false;
  a.animation_time += dt;
  float3 translation = make_float3(0.0f, 0.0f, 0.0f);
  float rotation = 0.0f;
  if (a.animation_time <= a.translation_time && !a.translation_complete) {
    const float3 d = a.translation_direction;
    const float k = a.translation_speed;
    translation = make_float3((d.x * k) * dt, (d.y * k) * dt, (d.z * k) * dt);
  } else if (!a.translation_complete) {
    a.translation_complete = true;
    a.animation_time = 0.0f;
    const auto b = sys.get_boundaries();
    long long lo = a.start_index, hi = (long long)a.start_index + a.num_particles;
    if (a.center_from_marked) {
      lo = 0;
      hi = b->size();
    } else if (lo < 0 || hi > (long long)b->size()) {
      throw std::out_of_range("update_animation: animated range outside the boundary set");
    }
    if (hi > lo) {
      std::vector<float3> p((size_t)(hi - lo));
      std::vector<int> m(a.center_from_marked ? p.size() : 0);
      if (hipDeviceSynchronize() != hipSuccess ||
          hipMemcpy(p.data(), b->get_pos_ptr() + lo, sizeof(float3) * p.size(),
                    hipMemcpyDeviceToHost) != hipSuccess ||
          (!m.empty() && hipMemcpy(m.data(), b->get_is_animated_ptr(), sizeof(int) * m.size(),
                                   hipMemcpyDeviceToHost) != hipSuccess))
        throw std::runtime_error("update_animation: device-to-host copy failed");
      float3 s = make_float3(0.0f, 0.0f, 0.0f);
      int count = 0;
      for (size_t i = 0; i < p.size(); ++i) {
        if (!m.empty() && m[i] != 1) continue;
        s = make_float3(s.x + p[i].x, s.y + p[i].y, s.z + p[i].z);
        ++count;
      }
      if (count > 0) {
        const float n = (float)count;
        a.rotation_center = make_float3(s.x / n, s.y / n, s.z / n);
      }
    }
  }
  if (a.translation_complete && a.rotation_angle < a.rotation_limit) {
    const float delta = -a.rotation_speed * dt;
    a.rotation_angle += std::fabs(delta);
    rotation = delta;
    if (a.rotation_angle >= a.rotation_limit) a.is_animating = false;
  }
  if (!a.is_animating) return false;
  animate_boundary(sys, translation, rotation, a.rotation_center);
  return true;
}

}  // namespace granular_animation
