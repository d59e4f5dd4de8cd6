// Host-side entry points shared between the library's translation units.
#pragma once
#include "granular_gpu.h"

namespace gg {
// Hash + stable counting sort of n positions: fills cell_start (C+2) and
// the workspace's key / perm arrays.
int sort_by_cell(const float *pos, int n, int *cell_start, const gg_workspace *ws,
                 hipStream_t st);
// The same sort by given cell keys (keys outside [0, C] count as C).
int sort_by_cell_keys(const float *pos, const int *keys_in, int n, int *cell_start,
                      const gg_workspace *ws, hipStream_t st);
// arr[k] = arr[perm[k]] for n records of `words` 32-bit words, through tmp.
int gather_inplace(void *arr, int words, const int *perm, int n, void *tmp, hipStream_t st);
// sort_by_cell (or by the given keys) + in-place permutation of the set's
// payload arrays.
int grid_build_general(gg_particles *set, gg_adaptive *ad, int *cell_start, int *perm_out,
                       const gg_workspace *ws, hipStream_t st, int capacity_limit,
                       const int *keys = nullptr);
// The packed float4 copy of the (already sorted) boundary positions that
// the contact kernels read, into the workspace.
int pack_boundary(const gg_particles *boundary, const gg_workspace *ws, hipStream_t st);
// Exclusive scan of m ints: out[0..m] (out[m] = total); in[] is cleared.
int scan_exclusive(int *in, int m, int *out, const gg_workspace *ws, hipStream_t st);
// Scenes sharing one set: checks the scene table against the workspace's
// cell count, and writes the cell ids of n particles (scene cell_base +
// cell in the scene's grid) to keys.
int check_scenes(const gg_scenes *sc, const gg_workspace *ws);
int scene_keys(const float *pos, int n, const int *scene_of, const gg_scenes *sc, float L,
               int *keys, hipStream_t st);
}  // namespace gg
