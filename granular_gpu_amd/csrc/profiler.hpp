// Stage profiler hooks used by the entry points (no-ops unless enabled).
#pragma once
#include <hip/hip_runtime.h>

namespace gg {
void prof_begin(int stage, hipStream_t st);
void prof_end(int stage, hipStream_t st);
}  // namespace gg
