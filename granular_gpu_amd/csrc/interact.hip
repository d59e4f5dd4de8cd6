// Interactive extras: the explosion impulse (Solver::trigger_explosion +
// apply_explosion_force, Solver.cu:1389-1442 of the reference) and sphere
// picking with a clamped move (markParticlesInSphere +
// updatePickedParticlesPositions, main.cu:1375-1426).
#include <hip/hip_runtime.h>

#include "gg_device.hpp"
#include "gg_internal.hpp"

namespace gg {
namespace {

constexpr int kComBlock = 256;
constexpr int kComMaxBlocks = 1024;

int com_blocks(int n) {
  const int b = (n + kComBlock - 1) / kComBlock;
  return b < 1 ? 1 : (b > kComMaxBlocks ? kComMaxBlocks : b);
}

// Per-block sums (x m, y m, z m, m) in double over a grid-stride range,
// reduced in LDS in a fixed tree order: the result depends only on n and
// the data, not on scheduling.
__global__ __launch_bounds__(kComBlock) void k_com_partials(const float *__restrict__ pos,
                                                            const float *__restrict__ mass,
                                                            int n, double *__restrict__ part) {
  __shared__ double s[4][kComBlock];
  double ax = 0.0, ay = 0.0, az = 0.0, am = 0.0;
  for (int i = blockIdx.x * kComBlock + threadIdx.x; i < n; i += gridDim.x * kComBlock) {
    const Packed3 p = ld3(pos, i);
    const double m = mass[i];
    ax += (double)p.x * m;
    ay += (double)p.y * m;
    az += (double)p.z * m;
    am += m;
  }
  const int t = threadIdx.x;
  s[0][t] = ax;
  s[1][t] = ay;
  s[2][t] = az;
  s[3][t] = am;
  __syncthreads();
  for (int w = kComBlock / 2; w > 0; w >>= 1) {
    if (t < w)
      for (int k = 0; k < 4; ++k) s[k][t] += s[k][t + w];
    __syncthreads();
  }
  if (t < 4) part[4 * blockIdx.x + t] = s[t][0];
}

// One block: partials in block order, then the centre of mass rounded to
// float (0 if the total mass is not positive), stored in com[0..2].
__global__ __launch_bounds__(64) void k_com_final(const double *__restrict__ part, int nparts,
                                                  float *__restrict__ com) {
  if (threadIdx.x != 0) return;
  double x = 0.0, y = 0.0, z = 0.0, m = 0.0;
  for (int b = 0; b < nparts; ++b) {
    x += part[4 * b];
    y += part[4 * b + 1];
    z += part[4 * b + 2];
    m += part[4 * b + 3];
  }
  if (m > 0.0) {
    com[0] = (float)(x / m);
    com[1] = (float)(y / m);
    com[2] = (float)(z / m);
  } else {
    com[0] = com[1] = com[2] = 0.0f;
  }
}

// apply_explosion_force: v += (d / |d|) F / (1 + |d|^2) / m, skipped within
// 1e-4 of the centre; float arithmetic in the reference's order.
__global__ __launch_bounds__(256) void k_explode(const float *__restrict__ pos,
                                                 float *__restrict__ vel,
                                                 const float *__restrict__ mass, int n,
                                                 const float *__restrict__ com, float force) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Packed3 p = ld3(pos, i);
  const float dx = p.x - com[0], dy = p.y - com[1], dz = p.z - com[2];
  const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
  if (dist < 1e-4f) return;
  const float ux = dx / dist, uy = dy / dist, uz = dz / dist;
  const float mag = force / (1.0f + dist * dist);
  const float m = mass[i];
  const Packed3 v = ld3(vel, i);
  st3(vel, i, v.x + (ux * mag) / m, v.y + (uy * mag) / m, v.z + (uz * mag) / m);
}

// flags[i] = 0 for particles within `radius` of c (kept by the compaction),
// 1 otherwise; iota[i] = i.
__global__ __launch_bounds__(256) void k_pick_flags(const float *__restrict__ pos, int n,
                                                    float cx, float cy, float cz, float radius,
                                                    int *__restrict__ flags,
                                                    int *__restrict__ iota) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Packed3 p = ld3(pos, i);
  const float dx = p.x - cx, dy = p.y - cy, dz = p.z - cz;
  const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
  flags[i] = dist <= radius ? 0 : 1;
  iota[i] = i;
}

__global__ __launch_bounds__(256) void k_move_picked(float *__restrict__ pos,
                                                     const int *__restrict__ idx,
                                                     const int *__restrict__ count, int max_count,
                                                     float dx, float dy, float dz, float lx,
                                                     float ly, float lz, float hx, float hy,
                                                     float hz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = *count;
  if (k >= max_count || k >= c) return;
  const int i = idx[k];
  const Packed3 p = ld3(pos, i);
  st3(pos, i, fmaxf(lx, fminf(p.x + dx, hx)), fmaxf(ly, fminf(p.y + dy, hy)),
      fmaxf(lz, fminf(p.z + dz, hz)));
}

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" size_t gg_explosion_scratch_bytes(int n) {
  return align256(sizeof(double) * 4 * (size_t)com_blocks(n)) + 256;
}

extern "C" int gg_explosion(gg_particles *set, float force, void *scratch, size_t scratch_bytes,
                            gg_float3 *com_out, hipStream_t stream) {
  if (!set || set->n < 0) return GG_ERR_INVALID;
  if (set->n == 0) return GG_OK;
  if (!set->pos || !set->vel || !set->mass || !scratch) return GG_ERR_INVALID;
  if (scratch_bytes < gg_explosion_scratch_bytes(set->n)) return GG_ERR_WORKSPACE;
  const int nb = com_blocks(set->n);
  double *part = static_cast<double *>(scratch);
  float *com = reinterpret_cast<float *>(static_cast<char *>(scratch) +
                                         align256(sizeof(double) * 4 * (size_t)nb));
  const float *pos = reinterpret_cast<const float *>(set->pos);
  k_com_partials<<<nb, kComBlock, 0, stream>>>(pos, set->mass, set->n, part);
  GG_LAUNCH_CHECK();
  k_com_final<<<1, 64, 0, stream>>>(part, nb, com);
  GG_LAUNCH_CHECK();
  k_explode<<<(set->n + 255) / 256, 256, 0, stream>>>(
      pos, reinterpret_cast<float *>(set->vel), set->mass, set->n, com, force);
  GG_LAUNCH_CHECK();
  if (com_out) {
    if (hipMemcpyAsync(com_out, com, 3 * sizeof(float), hipMemcpyDeviceToHost, stream) !=
            hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
      return GG_ERR_HIP;
  }
  return GG_OK;
}

extern "C" size_t gg_pick_scratch_bytes(int n) {
  const size_t m = n > 0 ? (size_t)n : 0;
  return 2 * align256(sizeof(int) * m) + align256(gg_compact_scratch_bytes((long long)m));
}

extern "C" int gg_pick_sphere(const gg_particles *set, gg_float3 center, float radius,
                              int *indices, int *count, void *scratch, size_t scratch_bytes,
                              hipStream_t stream) {
  if (!set || set->n < 0 || !count) return GG_ERR_INVALID;
  if (set->n == 0) return gg_fill_u32(count, 1, 0u, stream);
  if (!set->pos || !indices || !scratch) return GG_ERR_INVALID;
  if (scratch_bytes < gg_pick_scratch_bytes(set->n)) return GG_ERR_WORKSPACE;
  const size_t words = align256(sizeof(int) * (size_t)set->n);
  int *flags = static_cast<int *>(scratch);
  int *iota = reinterpret_cast<int *>(static_cast<char *>(scratch) + words);
  void *cs = static_cast<char *>(scratch) + 2 * words;
  k_pick_flags<<<(set->n + 255) / 256, 256, 0, stream>>>(
      reinterpret_cast<const float *>(set->pos), set->n, center.x, center.y, center.z, radius,
      flags, iota);
  GG_LAUNCH_CHECK();
  return gg_compact(iota, indices, set->n, 1, flags, count, cs,
                    align256(gg_compact_scratch_bytes(set->n)), stream);
}

extern "C" int gg_move_picked(gg_particles *set, const int *indices, const int *count,
                              int max_count, gg_float3 delta, gg_float3 lo, gg_float3 hi,
                              hipStream_t stream) {
  if (!set || max_count < 0 || max_count > set->n) return GG_ERR_INVALID;
  if (max_count == 0) return GG_OK;
  if (!set->pos || !indices || !count) return GG_ERR_INVALID;
  k_move_picked<<<(max_count + 255) / 256, 256, 0, stream>>>(
      reinterpret_cast<float *>(set->pos), indices, count, max_count, delta.x, delta.y, delta.z,
      lo.x, lo.y, lo.z, hi.x, hi.y, hi.z);
  GG_LAUNCH_CHECK();
  return GG_OK;
}
