// Kinematic boundary animation of the host API: the reference's
// updateAnimation / animateParticles (main.cu:1175-1277) over a
// GranularSystem's boundary set.  The system must be built with
// is_move_boundary so that its next step re-sorts the moved boundary
// (GranularSystem.cu:402-404).
#pragma once
#include "SceneConfig.hpp"

class GranularSystem;

namespace granular_animation {

// Moves the system's animated boundary particles: translation, then a
// rotation by angle_deg about the z axis through (center.x, center.y)
// (gg_boundary_animate).  Throws std::runtime_error on a device error.
void animate_boundary(GranularSystem &sys, float3 translation, float angle_deg, float3 center);

// One call of updateAnimation(dt) (main.cu:1210-1277):
//  - animation_time += dt;
//  - while animation_time <= translation_time and the translation is not
//    complete: move by translation_direction * translation_speed * dt;
//  - on the first call past translation_time: mark the translation complete,
//    reset animation_time, and set rotation_center to the float centroid of
//    boundary particles [start_index, start_index + num_particles) (or of
//    the marked particles, with center_from_marked), in set order;
//  - then, while rotation_angle < rotation_limit: rotation_angle +=
//    rotation_speed * dt and rotate by -rotation_speed * dt degrees; the
//    call that reaches the limit ends the animation and moves nothing.
// Returns true if the boundary was moved.  Does nothing once is_animating
// is false.
bool update_animation(GranularSystem &sys, AnimationState &a, float dt);

}  // namespace granular_animation
