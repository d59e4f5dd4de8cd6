// CPU oracle of the adaptive merge/split path -- TEST INFRASTRUCTURE ONLY.
// Restates GranularSystem.cu:506-699 (surface, surface distance) and
// Solver.cu:526-1242 (merge mark / blend, split, compaction, append) with
// the deterministic semantics documented in oracle.h.  FP32, operations in
// the order written (-ffp-contract=off).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdint>
#include <vector>

#include "oracle.h"
#include "oracle_internal.hpp"

namespace orc {

static inline V3<float> sub3(V3<float> a, V3<float> b) {
  return V3<float>{a.x - b.x, a.y - b.y, a.z - b.z};
}
static inline float dot3(V3<float> a, V3<float> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float len3(V3<float> a) { return std::sqrt(dot3(a, a)); }
// Approximate 1/sqrt(x): the bit-level initial estimate of Lomont, "Fast
// Inverse Square Root" (2003), refined by two Newton-Raphson steps
// y <- y (1.5 - (0.5 x) y y); relative error below 5e-6.  Basic IEEE
// operations only, so the device computes the same bits.
static inline float rsqrt_approx(float x) {
  uint32_t i;
  std::memcpy(&i, &x, 4);
  i = 0x5f3759dfu - (i >> 1);
  float y;
  std::memcpy(&y, &i, 4);
  const float hx = 0.5f * x;
  y = y * (1.5f - hx * y * y);
  y = y * (1.5f - hx * y * y);
  return y;
}

// normalize(v) = v * rsqrt(v.v) (helper_math.h:1325-1329).  The reference's
// surface kernels run on the device, where rsqrtf is an approximation; its
// error makes normalize(+d e) + normalize(-d' e) nonzero for d != d', so the
// interior of a regular lattice gets a tiny nonzero cover vector and is not
// surface.  A correctly rounded 1/sqrt returns exactly +-1 for axis vectors
// and flags every lattice interior as surface.  rsqrt_approx keeps the
// approximate behaviour deterministically.  v = 0 gives 0 (the vote then
// fails exactly as with the NaN of an exact 1/sqrt).
static inline V3<float> nrm3(V3<float> v) {
  const float inv = rsqrt_approx(dot3(v, v));
  return V3<float>{v.x * inv, v.y * inv, v.z * inv};
}

template <class F>
static void for_stencil(V3<float> p, const Grid &g, const int *cs, F &&f) {
  int bx, by, bz;
  if (!trunc_cell<float>(p, g.L, bx, by, bz)) return;
  for (int m = 0; m < 27; ++m) {
    const int cell = stencil_cell(bx, by, bz, m, g);
    if (cell < 0) continue;
    for (int j = cs[cell]; j < cs[cell + 1]; ++j) f(j);
  }
}

// K5 + K6 (GranularSystem.cu:506-611).  Neighbours: stencil particles j != i
// with |x_i - x_j| <= 2 (r_i + r_j), r from the float-density variant.
// cover c = sum normalize(x_i - x_j); surface <=> n <= 5 or no neighbour
// with normalize(x_j - x_i) . normalize(c) >= 0.5 (acos(.) <= pi/3).
void surface(int n, const V3<float> *pos, const float *mass, const int *cs, const Grid &g,
             float density, V3<float> *cover, int *surf) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    const float ri = radius_surface_f(mass[i], density);
    V3<float> c{0.0f, 0.0f, 0.0f};
    for_stencil(pos[i], g, cs, [&](int j) {
      if (j == i) return;
      const float rj = radius_surface_f(mass[j], density);
      const V3<float> b = sub3(pos[i], pos[j]);
      const float dis = len3(b);
      if (!(dis <= 2.0f * (ri + rj))) return;
      const V3<float> u = nrm3(b);
      c.x = c.x + u.x;
      c.y = c.y + u.y;
      c.z = c.z + u.z;
    });
    cover[i] = c;
  }
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    const float ri = radius_surface_f(mass[i], density);
    const V3<float> cv = nrm3(cover[i]);
    int nn = 0;
    bool vote = false;
    for_stencil(pos[i], g, cs, [&](int j) {
      if (j == i) return;
      const float rj = radius_surface_f(mass[j], density);
      const V3<float> b = sub3(pos[j], pos[i]);
      const float dis = len3(b);
      if (!(dis <= 2.0f * (ri + rj))) return;
      const V3<float> u = nrm3(b);
      if (dot3(u, cv) >= 0.5f) vote = true;
      ++nn;
    });
    surf[i] = (nn <= 5 || !vote) ? 1 : 0;
  }
}

// K7 (GranularSystem.cu:639-688): 0 on the surface, else the smallest
// squared distance to a surface particle of the stencil, 1e5 if none.
void surface_distance(int n, const V3<float> *pos, const int *surf, const int *cs, const Grid &g,
                      float *sd) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    if (surf[i] == 1) {
      sd[i] = 0.0f;
      continue;
    }
    float best = 100000.0f;
    for_stencil(pos[i], g, cs, [&](int j) {
      if (j == i || surf[j] == 0) return;
      const V3<float> d = sub3(pos[i], pos[j]);
      const float d2 = dot3(d, d);
      if (d2 < best) best = d2;
    });
    sd[i] = best;
  }
}

// K8 (Solver.cu:526-653) with claims resolved in ascending i.
void merge_mark(int n, const V3<float> *pos, const float *mass, const V3<float> *vel,
                const int *surf, const float *sd, int *remove, float *merge, V3<float> *mv,
                const int *cs, const Grid &g, float max_mass, const Rules &R, int *cand_out,
                float *dens_out) {
  std::vector<int> cand(n, -1);
  std::vector<float> dens(n, 0.0f);
  std::vector<float> rs(R.depth ? n : 0);
  if (R.depth)
    for (int i = 0; i < n; ++i) rs[i] = radius_surface_f(mass[i], R.density);
  auto rad = [&](int i) { return R.depth ? rs[i] : 0.0f; };
  // candidates from the pre-kernel snapshot of remove[]
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    const float mi = mass[i];
    if (surf[i] != 0 || remove[i] != 0 || mi < R.min_mass || mi >= max_mass ||
        !deep(R, surf[i], sd[i], rad(i)))
      continue;
    float closest = 1000.0f;
    int ci = -1;
    float ld = 0.0f;
    int lc = 0;
    for_stencil(pos[i], g, cs, [&](int j) {
      if (j >= n) return;
      const float dis = len3(sub3(pos[i], pos[j]));
      const float lim = partner_dist(R, rad(i), rad(j));
      if (dis < lim) {
        ld = ld + mass[j];
        ++lc;
      }
      if (j <= i || remove[j] != 0 || surf[j] != 0) return;
      if (mi + mass[j] > max_mass) return;
      if (dis < lim && dis < closest) {
        closest = dis;
        ci = j;
      }
    });
    if (lc > 0) {
      cand[i] = ci;
      dens[i] = ld / (float)lc;
    }
  }
  // sequential claim resolution
  std::vector<int> rem(remove, remove + n);
  for (int i = 0; i < n; ++i) {
    const int c = cand[i];
    if (c < 0) continue;
    if (rem[i] != 0) continue;  // claimed as a target earlier
    rem[i] = -1;
    if (!deep(R, surf[i], sd[i], rad(i))) continue;  // stays locked, no partner
    if (rem[c] != 0) {
      rem[i] = 0;
      continue;
    }
    rem[c] = -1;
    if (deep(R, surf[c], sd[c], rad(c))) {
      const float mi = mass[i], mc = mass[c];
      merge[c] = mi;
      merge[i] = -mi;
      const float mt = 1.0f / (mi + mc);
      V3<float> vt;
      vt.x = mt * (mi * vel[i].x + mc * vel[c].x);
      vt.y = mt * (mi * vel[i].y + mc * vel[c].y);
      vt.z = mt * (mi * vel[i].z + mc * vel[c].z);
      mv[i] = V3<float>{-vt.x, -vt.y, -vt.z};
      const float df = std::fmin(1.0f, dens[i] / 100.0f);
      const float boost = 1.0f + (1.0f - df) * 0.1f;
      mv[c] = V3<float>{vt.x * boost, vt.y * boost, vt.z * boost};
    } else {
      rem[i] = 0;  // partner stays locked (reference quirk)
    }
  }
  for (int i = 0; i < n; ++i) remove[i] = rem[i];
  if (cand_out)
    for (int i = 0; i < n; ++i) cand_out[i] = cand[i];
  if (dens_out)
    for (int i = 0; i < n; ++i) dens_out[i] = dens[i];
}

// K9 (Solver.cu:658-699).
void merge_blend(int n, float *merge, float *mass, V3<float> *vel, int *mcount, int *remove,
                 const V3<float> *mv, int blend, int *adl) {
  for (int i = 0; i < n; ++i) {
    if (remove[i] != -1) continue;
    const int old = mcount[i];
    mcount[i] = old + 1;
    const float bf = (float)blend;
    const float dm = merge[i] / bf;
    V3<float> dv{mv[i].x / bf, mv[i].y / bf, mv[i].z / bf};
    mass[i] = mass[i] + dm;
    adl[i] = 1;
    vel[i].x = vel[i].x + dv.x;
    vel[i].y = vel[i].y + dv.y;
    vel[i].z = vel[i].z + dv.z;
    if (old + 1 == blend) {
      if (merge[i] < 0.0f) {
        remove[i] = 1;
        adl[i] = 0;
      } else {
        remove[i] = 0;
      }
      mcount[i] = 0;
      merge[i] = 0.0f;
    }
  }
}

// K10 check (Solver.cu:708-827): surface particles with remove == 0 and
// m > 2 (depth rules: within split_depth r of the surface and m >
// split_mass); child at the centre of the first empty (granular and boundary)
// non-centre stencil cell; no boundary particle within 2 r.
bool split_check(int i, const V3<float> *pos, const float *mass, const int *remove,
                 const int *surf, const float *sd, const V3<float> *bpos, const int *cs,
                 const int *bcs, const Grid &g, const Rules &R, V3<float> &child) {
  if (remove[i] != 0 || !(mass[i] > R.split_mass)) return false;
  if (R.depth) {
    const float t = R.split_depth * radius_surface_f(mass[i], R.density);
    if (!(sd[i] < t * t)) return false;
  } else if (surf[i] != 1) {
    return false;
  }
  const float r = radius_split_f(mass[i], R.density);
  const float maxd = 2.0f * r;
  int bx, by, bz;
  if (!trunc_cell<float>(pos[i], g.L, bx, by, bz)) return false;
  bool found = false;
  for (int m = 0; m < 27; ++m) {
    const int x = bx + m / 9 - 1, y = by + (m % 9) / 3 - 1, z = bz + m % 3 - 1;
    const int cell = stencil_cell(bx, by, bz, m, g);
    if (cell < 0) continue;
    if (!found && m != 13 && cs[cell] == cs[cell + 1] && bcs[cell] == bcs[cell + 1]) {
      child = V3<float>{((float)x + 0.5f) * g.L, ((float)y + 0.5f) * g.L, ((float)z + 0.5f) * g.L};
      found = true;
    }
    for (int k = bcs[cell]; k < bcs[cell + 1]; ++k)
      if (len3(sub3(pos[i], bpos[k])) < maxd) return false;
  }
  return found;
}

}  // namespace orc

using namespace orc;

extern "C" int oracle_surface(int n, const float *pos, const float *mass, const int *cs,
                              const oracle_params *p, float *cover, int *surf) {
  Grid g{p->cell_length, p->cx, p->cy, p->cz};
  surface(n, reinterpret_cast<const V3<float> *>(pos), mass, cs, g, p->density,
          reinterpret_cast<V3<float> *>(cover), surf);
  return 0;
}

extern "C" int oracle_surface_distance(int n, const float *pos, const int *surf, const int *cs,
                                       const oracle_params *p, float *sd) {
  Grid g{p->cell_length, p->cx, p->cy, p->cz};
  surface_distance(n, reinterpret_cast<const V3<float> *>(pos), surf, cs, g, sd);
  return 0;
}

extern "C" int oracle_merge_mark(int n, const float *pos, const float *mass, const float *vel,
                                 const int *surf, const float *sd, int *remove, float *merge,
                                 float *mv, const int *cs, const oracle_params *p,
                                 float max_mass, int *cand, float *dens) {
  Grid g{p->cell_length, p->cx, p->cy, p->cz};
  merge_mark(n, reinterpret_cast<const V3<float> *>(pos), mass,
             reinterpret_cast<const V3<float> *>(vel), surf, sd, remove, merge,
             reinterpret_cast<V3<float> *>(mv), cs, g, max_mass, rules_of(p), cand, dens);
  return 0;
}

extern "C" int oracle_merge_blend(int n, float *merge, float *mass, float *vel, int *mcount,
                                  int *remove, const float *mv, int blend, int *adl) {
  merge_blend(n, merge, mass, reinterpret_cast<V3<float> *>(vel), mcount, remove,
              reinterpret_cast<const V3<float> *>(mv), blend, adl);
  return 0;
}

// Split + compaction + append (Solver.cu:971-1221; DArray.hpp:62-155;
// GranularParticles.hpp:30-95).  Arrays are in/out with room for `cap`
// particles.  Survivors (remove != 1) keep their order; children follow in
// ascending parent index; at most cap - survivors children are made (the
// later candidates do not split).  Returns the new count; stats[0] =
// removed, stats[1] = split, stats[2] = candidates refused for capacity.
extern "C" int oracle_split_compact(int n, int cap, float *pos, float *vel, float *mass,
                                    int *adl, int *remove, float *merge, int *mcount,
                                    const int *surf, const float *sd, const float *bpos,
                                    const int *cs, const int *bcs, const oracle_params *p,
                                    int *stats) {
  const Rules R = rules_of(p);
  if (R.depth && !sd) return -1;
  Grid g{p->cell_length, p->cx, p->cy, p->cz};
  V3<float> *P = reinterpret_cast<V3<float> *>(pos);
  V3<float> *V = reinterpret_cast<V3<float> *>(vel);
  const V3<float> *BP = reinterpret_cast<const V3<float> *>(bpos);
  std::vector<int> ok(n, 0);
  std::vector<V3<float>> child(n);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i)
    ok[i] = split_check(i, P, mass, remove, surf, sd, BP, cs, bcs, g, R, child[i]) ? 1 : 0;
  int survivors = 0, removed = 0;
  for (int i = 0; i < n; ++i) {
    if (remove[i] != 1) ++survivors;
    else ++removed;
  }
  const int room = cap - survivors;
  int made = 0, refused = 0;
  std::vector<V3<float>> cpos, cvel;
  std::vector<float> cmass;
  for (int i = 0; i < n; ++i) {
    if (!ok[i]) continue;
    if (made >= room) {
      ++refused;
      continue;
    }
    const float half = mass[i] / 2.0f;
    cpos.push_back(child[i]);
    cvel.push_back(V[i]);
    cmass.push_back(half);
    mass[i] = half;
    adl[i] = 1;
    ++made;
  }
  int k = 0;
  for (int i = 0; i < n; ++i) {
    if (remove[i] == 1) continue;
    P[k] = P[i];
    V[k] = V[i];
    mass[k] = mass[i];
    adl[k] = adl[i];
    merge[k] = merge[i];
    mcount[k] = mcount[i];
    remove[k] = remove[i];
    ++k;
  }
  for (int c = 0; c < made; ++c) {
    P[k] = cpos[c];
    V[k] = cvel[c];
    mass[k] = cmass[c];
    adl[k] = 1;
    merge[k] = 0.0f;
    mcount[k] = 0;
    remove[k] = 0;
    ++k;
  }
  if (stats) {
    stats[0] = removed;
    stats[1] = made;
    stats[2] = refused;
  }
  return k;
}
