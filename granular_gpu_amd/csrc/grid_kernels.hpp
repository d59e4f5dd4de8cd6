// Host-side entry points shared between the library's translation units.
#pragma once
#include "granular_gpu.h"

namespace gg {
// Hash + stable counting sort of n positions: fills cell_start (C+2) and
// the workspace's key / perm arrays.
int sort_by_cell(const float *pos, int n, int *cell_start, const gg_workspace *ws,
                 hipStream_t st);
// sort_by_cell + in-place permutation of the set's payload arrays.
int grid_build_general(gg_particles *set, gg_adaptive *ad, int *cell_start, int *perm_out,
                       const gg_workspace *ws, hipStream_t st, int capacity_limit);
// Exclusive scan of m ints: out[0..m] (out[m] = total); in[] is cleared.
int scan_exclusive(int *in, int m, int *out, const gg_workspace *ws, hipStream_t st);
}  // namespace gg
