// Shared constants and the HIP error-report macro of the host API.
// Mirrors Global.hpp of the reference (block_size, EPSILON, PI, CUDA_CALL
// that prints and continues), with HIP types.
#pragma once
#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include <cstdio>

const int block_size = 256;  // Global.hpp:5
#ifndef EPSILON
#define EPSILON (1e-6f)
#endif
#ifndef PI
#define PI (3.14159265358979323846f)
#endif
#define MAX_A (1000.0f)

// Prints and continues, as CUDA_CALL does (Global.hpp:8-15).
#define HIP_CALL(x)                                                                \
  do {                                                                             \
    hipError_t _gg_e = (x);                                                        \
    if (_gg_e != hipSuccess)                                                       \
      printf("HIP Error at %s:%d\t Error code = %d\n", __FILE__, __LINE__, (int)_gg_e); \
  } while (0)
