// Declarations shared between the oracle's translation units -- TEST
// INFRASTRUCTURE ONLY.
#pragma once
#include "oracle.h"
#include "oracle_common.hpp"

namespace orc {
// Merge / split rules.  Reference rules (Solver.cu:526-653, 780-827):
// "deep" = not surface and no surface particle in the stencil (sd >= 100:
// the sentinel 1e5 of compute_distance_to_surface), partners within 0.035,
// merging needs 1 <= m < max_mass, splitting needs a surface particle with
// m > 2.  Depth rules (radius levels): r = the surface radius of the mass;
// deep = not surface and sd > (merge_depth r)^2 (sd is a squared distance),
// partners within merge_dist (r_i + r_j), splitting needs sd <
// (split_depth r)^2 and m > split_mass.
struct Rules {
  bool depth;
  float min_mass, split_mass, merge_dist, merge_depth, split_depth, density;
};

inline Rules rules_of(const oracle_params *p) {
  return Rules{p->depth_rules != 0, p->min_mass, p->split_mass, p->merge_dist,
               p->merge_depth, p->split_depth, p->density};
}

// Deep enough to merge (the reference's eligibility test sd >= 100 and its
// partner test sd > 100 agree: sd is a squared distance below 100 or the
// sentinel 1e5).
inline bool deep(const Rules &R, int surf, float sd, float r) {
  if (!R.depth) return surf == 0 && sd > 100.0f;
  const float t = R.merge_depth * r;
  return surf == 0 && sd > t * t;
}

inline float partner_dist(const Rules &R, float ri, float rj) {
  return R.depth ? R.merge_dist * (ri + rj) : R.merge_dist;
}

void surface(int n, const V3<float> *pos, const float *mass, const int *cs, const Grid &g,
             float density, V3<float> *cover, int *surf);
void surface_distance(int n, const V3<float> *pos, const int *surf, const int *cs, const Grid &g,
                      float *sd);
void merge_mark(int n, const V3<float> *pos, const float *mass, const V3<float> *vel,
                const int *surf, const float *sd, int *remove, float *merge, V3<float> *mv,
                const int *cs, const Grid &g, float max_mass, const Rules &R, int *cand_out,
                float *dens_out);
void merge_blend(int n, float *merge, float *mass, V3<float> *vel, int *mcount, int *remove,
                 const V3<float> *mv, int blend, int *adl);
void upsample_one(V3<float> pi, V3<float> vi, const V3<float> *cpos, const V3<float> *cvel,
                  const int *cs, const V3<float> *bpos, const int *bcs, const Grid &g,
                  const oracle_upsample_params *u, V3<float> &np_out, V3<float> &v_out);
}  // namespace orc

extern "C" int oracle_split_compact(int n, int cap, float *pos, float *vel, float *mass,
                                    int *adl, int *remove, float *merge, int *mcount,
                                    const int *surf, const float *sd, const float *bpos,
                                    const int *cs, const int *bcs, const oracle_params *p,
                                    int *stats);
