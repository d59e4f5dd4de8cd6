// gg_fast_math_check: the short correctly rounded square root, reciprocal
// and quotient of the contact kernels (sqrt_core, rcp_core, div_core in
// gg_device.hpp) against the compiler's full IEEE sequences (sqrtf, '/'),
// input by input on the device.
//   mode 0: x = every float bit pattern first .. first+count-1 that lies in
//           [2^-96, max float]: sqrt_core(x) vs sqrtf(x);
//   mode 1: d = every pattern in [2^-48, 2^64): rcp_core(d) vs 1.0f / d;
//   mode 2: pattern k gives the numerator n = 2^e (1 + m) with e, m from k's
//           bits spread over [2^-80, 2^40], and a SplitMix64 hash of k the
//           denominator d over [2^-48, 2^40] (the domains the friction
//           quotient uses): div_core(n, d) vs n / d.
// Mismatches are added to *mismatches (device).
#include <hip/hip_runtime.h>

#include "gg_device.hpp"
#include "gg_internal.hpp"

namespace gg {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// A positive normal float with biased exponent in [elo, ehi] and the
// mantissa from the low 23 bits of h.
__device__ __forceinline__ float in_range(uint64_t h, int elo, int ehi) {
  const int e = elo + (int)((h >> 23) % (uint64_t)(ehi - elo + 1));
  return __uint_as_float(((uint32_t)e << 23) | (uint32_t)(h & 0x7FFFFFu));
}

__global__ void __launch_bounds__(256) k_fast_math_check(int mode, uint64_t first, uint64_t count,
                                                         unsigned long long *mismatches) {
  unsigned long long bad = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < count; k += stride) {
    const uint32_t bits = (uint32_t)(first + k);
    if (mode == 0) {
      const float x = __uint_as_float(bits);
      if (!(x >= kSqrtFastMin && x <= 3.40282347e+38f)) continue;
      bad += __float_as_uint(sqrt_core(x)) != __float_as_uint(sqrtf(x));
    } else if (mode == 1) {
      const float d = __uint_as_float(bits);
      if (!(d >= 0x1p-48f && d < 0x1p64f)) continue;
      bad += __float_as_uint(rcp_core(d)) != __float_as_uint(1.0f / d);
    } else {
      const uint64_t h = mix64(first + k);
      const float n = in_range((uint64_t)bits * 0x9E3779B1ull, 127 - 80, 127 + 40);
      const float d = in_range(h, 127 - 48, 127 + 40);
      bad += __float_as_uint(div_core(n, d)) != __float_as_uint(n / d);
    }
  }
  for (int o = 32; o > 0; o >>= 1) bad += __shfl_xor(bad, o, 64);
  if ((threadIdx.x & 63) == 0 && bad) atomicAdd(mismatches, bad);
}

}  // namespace gg

using namespace gg;

extern "C" int gg_fast_math_check(int mode, uint64_t first, uint64_t count,
                                  unsigned long long *mismatches, hipStream_t stream) {
  if (!mismatches || mode < 0 || mode > 2) return GG_ERR_INVALID;
  if (count == 0) return GG_OK;
  k_fast_math_check<<<2048, 256, 0, stream>>>(mode, first, count, mismatches);
  GG_LAUNCH_CHECK();
  return GG_OK;
}
