// CPU oracle -- TEST INFRASTRUCTURE ONLY.  Kinematic boundary motion:
// restates animateParticles (main.cu:1175-1208) and updateAnimation
// (main.cu:1210-1277) of the reference on host arrays.
#include <cmath>

#include "oracle.h"

extern "C" {

// angle_rad = angle * M_PI / 180.0f is formed in double and rounded to float
// (main.cu:1194); cosf/sinf of it (main.cu:1195-1196).
void oracle_rotation_cs(float angle_deg, float *c, float *s) {
  const float rad = (float)((double)angle_deg * M_PI / 180.0);
  *c = cosf(rad);
  *s = sinf(rad);
}

// main.cu:1183-1206: pos += translation; if the angle is non-zero, shift by
// the centre in x and y, rotate, shift back.  Each operation rounds to float.
void oracle_animate(int n, float *pos, const int *animated, float tx, float ty, float tz,
                    float angle_deg, float cx, float cy) {
  float c = 1.0f, s = 0.0f;
  oracle_rotation_cs(angle_deg, &c, &s);
  for (int i = 0; i < n; ++i) {
    if (animated[i] != 1) continue;
    float *p = pos + 3 * (long long)i;
    float x = p[0] + tx, y = p[1] + ty, z = p[2] + tz;
    if (angle_deg != 0.0f) {
      x -= cx;
      y -= cy;
      const float nx = x * c - y * s;
      const float ny = x * s + y * c;
      x = nx + cx;
      y = ny + cy;
    }
    p[0] = x;
    p[1] = y;
    p[2] = z;
  }
}

// main.cu:1210-1277 with the time and angle limits taken from the state.
int oracle_update_animation(oracle_animation *a, int n, float *pos, const int *animated,
                            float dt) {
  if (!a->is_animating) return 0;
  a->animation_time += dt;
  float t[3] = {0.0f, 0.0f, 0.0f};
  float rotation = 0.0f;
  if (a->animation_time <= a->translation_time && !a->translation_complete) {
    for (int k = 0; k < 3; ++k) t[k] = a->translation_direction[k] * a->translation_speed * dt;
  } else if (!a->translation_complete) {
    a->translation_complete = 1;
    a->animation_time = 0.0f;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    int count = 0;
    const int lo = a->center_from_marked ? 0 : a->start_index;
    const int hi = a->center_from_marked ? n : a->start_index + a->num_particles;
    for (int i = lo; i < hi; ++i) {
      if (a->center_from_marked && animated[i] != 1) continue;
      sx += pos[3 * i];
      sy += pos[3 * i + 1];
      sz += pos[3 * i + 2];
      count++;
    }
    if (count > 0) {
      a->rotation_center[0] = sx / (float)count;
      a->rotation_center[1] = sy / (float)count;
      a->rotation_center[2] = sz / (float)count;
    }
  }
  if (a->translation_complete && a->rotation_angle < a->rotation_limit) {
    const float delta = -a->rotation_speed * dt;
    a->rotation_angle += fabsf(delta);
    rotation = delta;
    if (a->rotation_angle >= a->rotation_limit) a->is_animating = 0;
  }
  if (!a->is_animating) return 0;
  oracle_animate(n, pos, animated, t[0], t[1], t[2], rotation, a->rotation_center[0],
                 a->rotation_center[1]);
  return 1;
}

}  // extern "C"
