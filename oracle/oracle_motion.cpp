// CPU oracle -- TEST INFRASTRUCTURE ONLY.  Kinematic boundary motion:
// restates animateParticles (main.cu:1175-1208) and updateAnimation
// (main.cu:1210-1277) of the reference on host arrays; and the interactive
// extras (explosion impulse, sphere picking and the clamped move).
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

// ---- interactive extras (Solver.cu:1389-1442, main.cu:1375-1426) ----------

extern "C" {

// Centre of mass sum(m p) / sum(m), accumulated in double in index order and
// rounded to float; 0 when the total mass is not positive.
void oracle_center_of_mass(int n, const float *pos, const float *mass, float *com) {
  double x = 0.0, y = 0.0, z = 0.0, m = 0.0;
  for (int i = 0; i < n; ++i) {
    x += (double)pos[3 * i] * mass[i];
    y += (double)pos[3 * i + 1] * mass[i];
    z += (double)pos[3 * i + 2] * mass[i];
    m += mass[i];
  }
  if (m > 0.0) {
    com[0] = (float)(x / m);
    com[1] = (float)(y / m);
    com[2] = (float)(z / m);
  } else {
    com[0] = com[1] = com[2] = 0.0f;
  }
}

// apply_explosion_force (Solver.cu:1389-1409) about a given centre.
void oracle_explode(int n, const float *pos, float *vel, const float *mass, const float *com,
                    float force) {
  for (int i = 0; i < n; ++i) {
    const float dx = pos[3 * i] - com[0], dy = pos[3 * i + 1] - com[1],
                dz = pos[3 * i + 2] - com[2];
    const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
    if (dist < 1e-4f) continue;
    const float ux = dx / dist, uy = dy / dist, uz = dz / dist;
    const float mag = force / (1.0f + dist * dist);
    vel[3 * i] += (ux * mag) / mass[i];
    vel[3 * i + 1] += (uy * mag) / mass[i];
    vel[3 * i + 2] += (uz * mag) / mass[i];
  }
}

// markParticlesInSphere (main.cu:1375-1400) with the indices in ascending
// order; returns their number.
int oracle_pick_sphere(int n, const float *pos, const float *center, float radius, int *out) {
  int c = 0;
  for (int i = 0; i < n; ++i) {
    const float dx = pos[3 * i] - center[0], dy = pos[3 * i + 1] - center[1],
                dz = pos[3 * i + 2] - center[2];
    if (sqrtf(dx * dx + dy * dy + dz * dz) <= radius) out[c++] = i;
  }
  return c;
}

// updatePickedParticlesPositions (main.cu:1402-1426) with bounds lo, hi.
void oracle_move_picked(float *pos, const int *idx, int count, const float *delta,
                        const float *lo, const float *hi) {
  for (int k = 0; k < count; ++k)
    for (int a = 0; a < 3; ++a) {
      float *p = pos + 3 * (long long)idx[k] + a;
      *p = fmaxf(lo[a], fminf(*p + delta[a], hi[a]));
    }
}

}  // extern "C"
