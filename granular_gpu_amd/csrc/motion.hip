// Kinematic boundary motion: translate, then rotate about the z axis, the
// boundary particles marked animated.  Replaces animateParticles
// (main.cu:1175-1208 of the reference).
#include <hip/hip_runtime.h>

#include <cmath>

#include "gg_device.hpp"
#include "gg_internal.hpp"

namespace gg {
namespace {

// One thread per particle; unmarked particles are not written.  With
// -ffp-contract=off every product and sum rounds separately, so the result
// is the float arithmetic of the formulas as written.
__global__ __launch_bounds__(256) void k_animate(float *__restrict__ pos,
                                                 const int *__restrict__ animated, int n,
                                                 float tx, float ty, float tz, int rotate, float c,
                                                 float s, float cx, float cy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || animated[i] != 1) return;
  const Packed3 p = ld3(pos, i);
  float x = p.x + tx, y = p.y + ty;
  const float z = p.z + tz;
  if (rotate) {
    const float u = x - cx, v = y - cy;
    x = (u * c - v * s) + cx;
    y = (u * s + v * c) + cy;
  }
  st3(pos, i, x, y, z);
}

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_boundary_animate(gg_particles *boundary, gg_float3 translation, float angle_deg,
                                   gg_float3 center, hipStream_t stream) {
  if (!boundary || boundary->n < 0) return GG_ERR_INVALID;
  if (boundary->n == 0) return GG_OK;
  if (!boundary->pos || !boundary->is_animated) return GG_ERR_INVALID;
  // The reference forms angle_deg * M_PI / 180 in double and rounds it to
  // float before cosf/sinf; here both are evaluated once, on the host.
  const float rad = (float)((double)angle_deg * M_PI / 180.0);
  const float c = std::cos(rad), s = std::sin(rad);
  const int blocks = (boundary->n + 255) / 256;
  k_animate<<<blocks, 256, 0, stream>>>(reinterpret_cast<float *>(boundary->pos),
                                        boundary->is_animated, boundary->n, translation.x,
                                        translation.y, translation.z, angle_deg != 0.0f, c, s,
                                        center.x, center.y);
  GG_LAUNCH_CHECK();
  return GG_OK;
}
