// Declarations shared between the oracle's translation units -- TEST
// INFRASTRUCTURE ONLY.
#pragma once
#include "oracle.h"
#include "oracle_common.hpp"

namespace orc {
void surface(int n, const V3<float> *pos, const float *mass, const int *cs, const Grid &g,
             float density, V3<float> *cover, int *surf);
void surface_distance(int n, const V3<float> *pos, const int *surf, const int *cs, const Grid &g,
                      float *sd);
void merge_mark(int n, const V3<float> *pos, const float *mass, const V3<float> *vel,
                const int *surf, const float *sd, int *remove, float *merge, V3<float> *mv,
                const int *cs, const Grid &g, float max_mass, int *cand_out, float *dens_out);
void merge_blend(int n, float *merge, float *mass, V3<float> *vel, int *mcount, int *remove,
                 const V3<float> *mv, int blend, int *adl);
void upsample_one(V3<float> pi, V3<float> vi, const V3<float> *cpos, const V3<float> *cvel,
                  const int *cs, const V3<float> *bpos, const int *bcs, const Grid &g,
                  const oracle_upsample_params *u, V3<float> &np_out, V3<float> &v_out);
}  // namespace orc

extern "C" int oracle_split_compact(int n, int cap, float *pos, float *vel, float *mass,
                                    int *adl, int *remove, float *merge, int *mcount,
                                    const int *surf, const float *bpos, const int *cs,
                                    const int *bcs, const oracle_params *p, int *stats);
