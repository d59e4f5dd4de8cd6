// CPU oracle for the plain granular PBD step -- TEST INFRASTRUCTURE ONLY.
// Restates the reference algorithm (stingyemperor/granular_gpu) with the
// deterministic semantics documented in oracle.h.  Compiled with
// -ffp-contract=off and without -ffast-math: every FP32 operation below is an
// individually rounded IEEE operation, in the order written.
#include "oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// Grid build.  Hash = CUDAFunctions.cuh:71-107 (floorf of p/L per axis,
// x-major id (ix*cy+iy)*cz+iz, out-of-grid -> sentinel C).  Sort = stable by
// cell id (GranularSystem.cu:85-88 thrust::sort_by_key on int keys).  Cell
// table = exclusive scan of the per-cell histogram (CUDAFunctions.cuh:54-61,
// GranularSystem.cu:149-166), extended by one entry (C+1 -> n).
// ---------------------------------------------------------------------------
template <class T>
void grid_build(const V3<T> *pos, int n, const Grid &g, int *keys_sorted,
                int *perm, int *cell_start) {
  const int C = g.cx * g.cy * g.cz;
  std::vector<int> key(n);
  const T L = (T)g.L;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) key[i] = hash_cell<T>(pos[i], L, g);
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(),
                   [&](int a, int b) { return key[a] < key[b]; });
  std::vector<int> count(C + 1, 0);
  for (int i = 0; i < n; ++i) count[key[i]]++;
  int run = 0;
  for (int c = 0; c <= C; ++c) {
    cell_start[c] = run;
    run += count[c];
  }
  cell_start[C + 1] = run;
  for (int k = 0; k < n; ++k) {
    perm[k] = idx[k];
    keys_sorted[k] = key[idx[k]];
  }
}

// ---------------------------------------------------------------------------
// One contact sweep for particle i (Solver.cu:361-401 with the
// boundary_constraint / particles_constraint bodies, Solver.cu:272-359),
// evaluated as a symmetric gather:
//   boundary contact k : d < R_b + max(r_i, rmin)   -> ab += ((rc-d)/d) p
//   granular contact j : d <= r_i + r_j              -> ag += ((rc-d)/d) p
//                                                       n  += 2
// (own_ij + mirror_ji of the reference's scatter; mass scaling cancels).
// Optional PBD friction in the solver phase removes the relative tangential
// displacement (static if |t|^2 < (mu_s*delta)^2, else scaled by
// mu_k*delta/|t|).
// p = (b_i - b_j) + (o_i - o_j): positions are base + offset, never summed
// before differencing, so relative geometry keeps full precision.
//
// MODE: FREE evaluates every decision in precision T; RECORD also writes
// the discrete decisions (stencil base, contact, friction branch) to a tape;
// REPLAY takes them from the tape (decision-shadowed FP64 run).
// ---------------------------------------------------------------------------
enum { FREE = 0, RECORD = 1, REPLAY = 2 };

struct Tape {
  std::vector<uint8_t> bits;
  size_t pos = 0;
  bool okb = false;
  int bx = 0, by = 0, bz = 0;
  void clear() {
    bits.clear();
    pos = 0;
  }
};

// Friction with recorded / replayed branch (bit1 static, bit2 clamped).
template <class T, int MODE>
inline V3<T> friction_m(V3<T> p, T inv, T delta, V3<T> rel, const SweepParams<T> &sp,
                        uint8_t &bits) {
  const T nx = p.x * inv, ny = p.y * inv, nz = p.z * inv;
  const T rn = rel.x * nx + rel.y * ny + rel.z * nz;
  V3<T> t;
  t.x = rel.x - rn * nx;
  t.y = rel.y - rn * ny;
  t.z = rel.z - rn * nz;
  const T tl2 = t.x * t.x + t.y * t.y + t.z * t.z;
  const T lim = sp.mu_s * delta;
  bool is_static;
  if (MODE == REPLAY) is_static = bits & 2;
  else is_static = tl2 < lim * lim;
  if (MODE == RECORD && is_static) bits |= 2;
  if (is_static) return t;
  const T tl = std::sqrt(tl2);
  T c = (sp.mu_k * delta) / tl;
  bool clamp;
  if (MODE == REPLAY) clamp = bits & 4;
  else clamp = !(c < (T)1);
  if (MODE == RECORD && clamp) bits |= 4;
  if (clamp) c = (T)1;
  return V3<T>{t.x * c, t.y * c, t.z * c};
}

template <class T, int MODE>
V3<T> sweep_one(int i, const V3<T> *base, const T *rad, const V3<T> *off,
                const V3<T> *bpos, const int *cs, const int *bcs,
                const Grid &g, const SweepParams<T> &sp, bool solver_phase,
                Tape *tape, const T *ims) {
  const V3<T> bi = base[i];
  const V3<T> oi = off[i];
  const T ri = rad[i];
  const T rbi = ri > sp.rmin ? ri : sp.rmin;
  int bx = 0, by = 0, bz = 0;
  bool okb;
  if (MODE == REPLAY) {
    okb = tape->okb;
    bx = tape->bx;
    by = tape->by;
    bz = tape->bz;
  } else {
    const V3<T> xs = add(bi, oi);
    okb = trunc_cell<T>(xs, (T)g.L, bx, by, bz);
    if (MODE == RECORD) {
      tape->okb = okb;
      tape->bx = bx;
      tape->by = by;
      tape->bz = bz;
    }
  }
  V3<T> ab = {0, 0, 0}, ag = {0, 0, 0};
  int nb = 0, ng = 0;
  const bool fric = solver_phase && (sp.mu_s > 0 || sp.mu_k > 0);
  if (okb) {
    for (int m = 0; m < 27; ++m) {
      const int cell = stencil_cell(bx, by, bz, m, g);
      if (cell < 0) continue;
      for (int k = bcs[cell]; k < bcs[cell + 1]; ++k) {
        const V3<T> q = bpos[k];
        V3<T> p;
        p.x = (bi.x - q.x) + oi.x;
        p.y = (bi.y - q.y) + oi.y;
        p.z = (bi.z - q.z) + oi.z;
        const T d = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
        const T rc = sp.rb + rbi;
        uint8_t bits = 0;
        bool c;
        if (MODE == REPLAY) {
          bits = tape->bits[tape->pos++];
          c = bits & 1;
        } else {
          c = d < rc && d > 0;
          bits = c ? 1 : 0;
        }
        if (c) {
          const T inv = (T)1 / d;
          const T s = (rc - d) * inv;
          ab.x = ab.x + s * p.x;
          ab.y = ab.y + s * p.y;
          ab.z = ab.z + s * p.z;
          if (fric) {
            const V3<T> f = friction_m<T, MODE>(p, inv, rc - d, oi, sp, bits);
            ab.x = ab.x - f.x;
            ab.y = ab.y - f.y;
            ab.z = ab.z - f.z;
          }
          nb += 1;
        }
        if (MODE == RECORD) tape->bits.push_back(bits);
      }
      for (int j = cs[cell]; j < cs[cell + 1]; ++j) {
        if (j == i) continue;
        const V3<T> bj = base[j];
        const V3<T> oj = off[j];
        V3<T> p;
        p.x = (bi.x - bj.x) + (oi.x - oj.x);
        p.y = (bi.y - bj.y) + (oi.y - oj.y);
        p.z = (bi.z - bj.z) + (oi.z - oj.z);
        const T d = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
        const T rc = ri + rad[j];
        uint8_t bits = 0;
        bool c;
        if (MODE == REPLAY) {
          bits = tape->bits[tape->pos++];
          c = bits & 1;
        } else {
          c = d <= rc && d > 0;
          bits = c ? 1 : 0;
        }
        if (c) {
          const T inv = (T)1 / d;
          T s = (rc - d) * inv;
          // own-only: the reference's own term w_i*(mag/d)*p with
          // w_i = (1/ms_i) / (1/ms_i + 1/ms_j) (Solver.cu:317-345)
          const T w = sp.own_only ? ims[i] / (ims[i] + ims[j]) : (T)1;
          if (sp.own_only) s = w * s;
          ag.x = ag.x + s * p.x;
          ag.y = ag.y + s * p.y;
          ag.z = ag.z + s * p.z;
          if (fric) {
            V3<T> rel;
            rel.x = oi.x - oj.x;
            rel.y = oi.y - oj.y;
            rel.z = oi.z - oj.z;
            V3<T> f = friction_m<T, MODE>(p, inv, rc - d, rel, sp, bits);
            if (sp.own_only) {
              f.x = w * f.x;
              f.y = w * f.y;
              f.z = w * f.z;
            }
            ag.x = ag.x - f.x;
            ag.y = ag.y - f.y;
            ag.z = ag.z - f.z;
          }
          ng += 2;
        }
        if (MODE == RECORD) tape->bits.push_back(bits);
      }
    }
  }
  const int n = nb + ng;
  const T nf = (T)(n > 1 ? n : 1);
  V3<T> o;
  o.x = oi.x + (ab.x + ag.x) / nf;
  o.y = oi.y + (ab.y + ag.y) / nf;
  o.z = oi.z + (ab.z + ag.z) / nf;
  return o;
}

// Predict (Solver.cu:125-191): v += dt*g (dv = dt*G formed first, on the
// host in the reference), D0 = dt*v, scaled mass = m*exp(-y), radius.
template <class T>
void predict(int n, V3<T> *vel, const V3<T> *pos, const T *mass,
             V3<T> *d0, T *ms, T *rad, const SweepParams<T> &sp, T dt,
             V3<T> gvec, float density) {
  V3<T> dv;
  dv.x = dt * gvec.x;
  dv.y = dt * gvec.y;
  dv.z = dt * gvec.z;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    V3<T> v = vel[i];
    v.x = v.x + dv.x;
    v.y = v.y + dv.y;
    v.z = v.z + dv.z;
    vel[i] = v;
    d0[i].x = dt * v.x;
    d0[i].y = dt * v.y;
    d0[i].z = dt * v.z;
    ms[i] = mass[i] * exp_t<T>(-pos[i].y);
    rad[i] = radius_contact_t<T>(mass[i], density);
  }
  (void)sp;
}

// Finalize (Solver.cu:193-270) for particle i: v = D/dt with sleep below
// sleep_speed, damping above damp_speed and the adaptive x0.98; x = base + D.
// Decision bits: bit0 sleep, bit1 damped.
template <class T, int MODE>
void finalize_one(int i, V3<T> *pos, V3<T> *vel, const V3<T> *xb, const V3<T> *D,
                  const int *adaptive_last, const SweepParams<T> &sp, T dt,
                  uint8_t &bits) {
  const T inv_dt = (T)1 / dt;
  V3<T> v;
  v.x = inv_dt * D[i].x;
  v.y = inv_dt * D[i].y;
  v.z = inv_dt * D[i].z;
  const T speed = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  bool sleep, damp;
  if (MODE == REPLAY) {
    sleep = bits & 1;
    damp = bits & 2;
  } else {
    sleep = speed < sp.sleep;
    damp = !sleep && speed > sp.damp_speed;
    bits = (sleep ? 1 : 0) | (damp ? 2 : 0);
  }
  if (sleep) {
    v.x = 0;
    v.y = 0;
    v.z = 0;
  } else if (damp) {
    const T e = exp_t<T>(-speed * (T)0.1f);
    const T f = sp.damping + ((T)1 - sp.damping) * e;
    v.x = v.x * f;
    v.y = v.y * f;
    v.z = v.z * f;
    if (adaptive_last && adaptive_last[i] == 1) {
      v.x = v.x * sp.adaptive_damping;
      v.y = v.y * sp.adaptive_damping;
      v.z = v.z * sp.adaptive_damping;
    }
  }
  vel[i] = v;
  pos[i].x = xb[i].x + D[i].x;
  pos[i].y = xb[i].y + D[i].y;
  pos[i].z = xb[i].z + D[i].z;
}

// Per-run scratch of one precision.
template <class T> struct Sim {
  int n = 0;
  V3<T> *pos = nullptr, *vel = nullptr;
  T *mass = nullptr;
  const V3<T> *bpos = nullptr;
  std::vector<V3<T>> tp, tv, d0, xb, a, b;
  std::vector<T> tm, ms, rad, ims;
  SweepParams<T> sp;
  void init(int n_, T *p, T *v, T *m, const T *bp, const oracle_params *prm) {
    n = n_;
    pos = reinterpret_cast<V3<T> *>(p);
    vel = reinterpret_cast<V3<T> *>(v);
    mass = m;
    bpos = reinterpret_cast<const V3<T> *>(bp);
    tp.resize(n);
    tv.resize(n);
    d0.resize(n);
    xb.resize(n);
    a.resize(n);
    b.resize(n);
    tm.resize(n);
    ms.resize(n);
    rad.resize(n);
    ims.resize(n);
    sp = make_sweep_params<T>(prm);
  }
  void permute(const std::vector<int> &idx) {
    for (int k = 0; k < n; ++k) {
      tp[k] = pos[idx[k]];
      tv[k] = vel[idx[k]];
      tm[k] = mass[idx[k]];
    }
    for (int k = 0; k < n; ++k) {
      pos[k] = tp[k];
      vel[k] = tv[k];
      mass[k] = tm[k];
    }
  }
  // Predict (Solver.cu:125-191): v += dt*g (dv = dt*G formed first, on the
  // host in the reference), D0 = dt*v, scaled mass m*exp(-y), radius.
  void predict(const oracle_params *p) {
    const T dt = (T)p->dt;
    V3<T> dv;
    dv.x = dt * (T)p->gx;
    dv.y = dt * (T)p->gy;
    dv.z = dt * (T)p->gz;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      V3<T> v = vel[i];
      v.x = v.x + dv.x;
      v.y = v.y + dv.y;
      v.z = v.z + dv.z;
      vel[i] = v;
      d0[i].x = dt * v.x;
      d0[i].y = dt * v.y;
      d0[i].z = dt * v.z;
      ms[i] = mass[i] * exp_t<T>(-pos[i].y);
      ims[i] = (T)1 / ms[i];
      rad[i] = radius_contact_t<T>(mass[i], p->density);
      a[i] = d0[i];
    }
  }
  // After the S stabilisation sweeps (Solver.cu:428-476 moved x and x* by
  // the same amounts): x = x0 + (P - D0), and the solver offset is D0.
  void end_stabilisation(int S) {
    for (int i = 0; i < n; ++i) {
      if (S > 0) {
        V3<T> e;
        e.x = a[i].x - d0[i].x;
        e.y = a[i].y - d0[i].y;
        e.z = a[i].z - d0[i].z;
        xb[i].x = pos[i].x + e.x;
        xb[i].y = pos[i].y + e.y;
        xb[i].z = pos[i].z + e.z;
      } else {
        xb[i] = pos[i];
      }
      a[i] = d0[i];
    }
  }
};

static void sort_perm(const V3<float> *pos, int n, const Grid &g, std::vector<int> &idx,
                      std::vector<int> &keys_sorted, int *cs) {
  const int C = g.cx * g.cy * g.cz;
  std::vector<int> key(n);
  for (int i = 0; i < n; ++i) key[i] = hash_cell<float>(pos[i], g.L, g);
  idx.resize(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return key[x] < key[y]; });
  std::vector<int> count(C + 1, 0);
  for (int i = 0; i < n; ++i) count[key[i]]++;
  int run = 0;
  for (int c = 0; c <= C; ++c) {
    cs[c] = run;
    run += count[c];
  }
  cs[C + 1] = run;
  keys_sorted.resize(n);
  for (int k = 0; k < n; ++k) keys_sorted[k] = key[idx[k]];
}

template <class T>
static void sort_perm_t(const V3<T> *pos, int n, const Grid &g, std::vector<int> &idx,
                        std::vector<int> &keys_sorted, int *cs) {
  if constexpr (std::is_same<T, float>::value) {
    sort_perm(pos, n, g, idx, keys_sorted, cs);
  } else {
    const int C = g.cx * g.cy * g.cz;
    std::vector<int> key(n);
    for (int i = 0; i < n; ++i) key[i] = hash_cell<T>(pos[i], (T)g.L, g);
    idx.resize(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return key[x] < key[y]; });
    std::vector<int> count(C + 1, 0);
    for (int i = 0; i < n; ++i) count[key[i]]++;
    int run = 0;
    for (int c = 0; c <= C; ++c) {
      cs[c] = run;
      run += count[c];
    }
    cs[C + 1] = run;
    keys_sorted.resize(n);
    for (int k = 0; k < n; ++k) keys_sorted[k] = key[idx[k]];
  }
}

// Free run in precision T: every decision taken in T.
template <class T>
int plain_steps(int n, T *pos_, T *vel_, T *mass, int nb, const T *bpos_, const int *bcs,
                const oracle_params *p, int nsteps, float *ms_out, int *p2c_out,
                int *adaptive_last, int *ids) {
  Grid g{p->cell_length, p->cx, p->cy, p->cz};
  Sim<T> s;
  s.init(n, pos_, vel_, mass, bpos_, p);
  const int C = g.cx * g.cy * g.cz;
  std::vector<int> cs(C + 2), idx, keys;
  std::vector<int> idtmp(n);
  const int S = p->stab_iters, I = p->solver_iters;
  for (int step = 0; step < nsteps; ++step) {
    sort_perm_t<T>(s.pos, n, g, idx, keys, cs.data());
    s.permute(idx);
    if (ids) {
      for (int k = 0; k < n; ++k) idtmp[k] = ids[idx[k]];
      for (int k = 0; k < n; ++k) ids[k] = idtmp[k];
    }
    s.predict(p);
    for (int k = 0; k < S; ++k) {
#pragma omp parallel for schedule(dynamic, 256)
      for (int i = 0; i < n; ++i)
        s.b[i] = sweep_one<T, FREE>(i, s.pos, s.rad.data(), s.a.data(), s.bpos, cs.data(), bcs,
                                    g, s.sp, false, nullptr, s.ims.data());
      std::swap(s.a, s.b);
    }
    s.end_stabilisation(S);
    for (int k = 0; k < I; ++k) {
#pragma omp parallel for schedule(dynamic, 256)
      for (int i = 0; i < n; ++i)
        s.b[i] = sweep_one<T, FREE>(i, s.xb.data(), s.rad.data(), s.a.data(), s.bpos, cs.data(),
                                    bcs, g, s.sp, true, nullptr, s.ims.data());
      std::swap(s.a, s.b);
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      uint8_t bits = 0;
      finalize_one<T, FREE>(i, s.pos, s.vel, s.xb.data(), s.a.data(), adaptive_last, s.sp,
                            (T)p->dt, bits);
    }
    if (adaptive_last)
      for (int i = 0; i < n; ++i) adaptive_last[i] = 0;
  }
  if (ms_out)
    for (int i = 0; i < n; ++i) ms_out[i] = (float)s.ms[i];
  if (p2c_out)
    for (int i = 0; i < n; ++i) p2c_out[i] = keys.empty() ? 0 : keys[i];
  return 0;
}

// Lockstep FP32 + FP64 run: the FP64 state takes every discrete decision
// (sort order, stencil base, contact, friction branch, sleep / damping) from
// the FP32 state, so it follows the same discrete path and measures only
// the rounding error of the FP32 arithmetic.
int shadow_steps(int n, float *pf, float *vf, float *mf, double *pd, double *vd, double *md,
                 int nb, const float *bpf, const double *bpd, const int *bcs,
                 const oracle_params *p, int nsteps, int *adaptive_last) {
  Grid g{p->cell_length, p->cx, p->cy, p->cz};
  Sim<float> f;
  Sim<double> d;
  f.init(n, pf, vf, mf, bpf, p);
  d.init(n, pd, vd, md, bpd, p);
  const int C = g.cx * g.cy * g.cz;
  std::vector<int> cs(C + 2), idx, keys;
  const int S = p->stab_iters, I = p->solver_iters;
  for (int step = 0; step < nsteps; ++step) {
    sort_perm(f.pos, n, g, idx, keys, cs.data());
    f.permute(idx);
    d.permute(idx);
    f.predict(p);
    d.predict(p);
    for (int k = 0; k < S + I; ++k) {
      const bool solver = k >= S;
      if (k == S) {
        f.end_stabilisation(S);
        d.end_stabilisation(S);
      }
      const V3<float> *bf = solver ? f.xb.data() : f.pos;
      const V3<double> *bd = solver ? d.xb.data() : d.pos;
#pragma omp parallel for schedule(dynamic, 256)
      for (int i = 0; i < n; ++i) {
        Tape t;
        f.b[i] = sweep_one<float, RECORD>(i, bf, f.rad.data(), f.a.data(), f.bpos, cs.data(),
                                          bcs, g, f.sp, solver, &t, f.ims.data());
        d.b[i] = sweep_one<double, REPLAY>(i, bd, d.rad.data(), d.a.data(), d.bpos, cs.data(),
                                           bcs, g, d.sp, solver, &t, d.ims.data());
      }
      std::swap(f.a, f.b);
      std::swap(d.a, d.b);
    }
    if (S + I == S) {
      f.end_stabilisation(S);
      d.end_stabilisation(S);
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      uint8_t bits = 0;
      finalize_one<float, RECORD>(i, f.pos, f.vel, f.xb.data(), f.a.data(), adaptive_last, f.sp,
                                  (float)p->dt, bits);
      finalize_one<double, REPLAY>(i, d.pos, d.vel, d.xb.data(), d.a.data(), adaptive_last,
                                   d.sp, (double)p->dt, bits);
    }
    if (adaptive_last)
      for (int i = 0; i < n; ++i) adaptive_last[i] = 0;
  }
  return 0;
}

}  // namespace orc

using namespace orc;

extern "C" int oracle_hash(const float *pos, int n, float L, int cx, int cy,
                           int cz, int *keys) {
  Grid g{L, cx, cy, cz};
  const V3<float> *p = reinterpret_cast<const V3<float> *>(pos);
  for (int i = 0; i < n; ++i) keys[i] = hash_cell<float>(p[i], L, g);
  return 0;
}

extern "C" int oracle_grid_build(const float *pos, int n, float L, int cx,
                                 int cy, int cz, int *keys_sorted, int *perm,
                                 int *cell_start) {
  if (n < 0) return -1;
  Grid g{L, cx, cy, cz};
  grid_build<float>(reinterpret_cast<const V3<float> *>(pos), n, g,
                    keys_sorted, perm, cell_start);
  return 0;
}

extern "C" int oracle_plain_steps_f32(int n, float *pos, float *vel,
                                      float *mass, int nb, const float *bpos,
                                      const int *bcs, const oracle_params *p,
                                      int nsteps, float *ms_out, int *p2c_out,
                                      int *adaptive_last) {
  if (n < 0 || nb < 0 || !p) return -1;
  return plain_steps<float>(n, pos, vel, mass, nb, bpos, bcs, p, nsteps,
                            ms_out, p2c_out, adaptive_last, nullptr);
}

extern "C" int oracle_plain_steps_f64(int n, double *pos, double *vel,
                                      double *mass, int nb, const double *bpos,
                                      const int *bcs, const oracle_params *p,
                                      int nsteps, int *adaptive_last,
                                      int *ids) {
  if (n < 0 || nb < 0 || !p) return -1;
  return plain_steps<double>(n, pos, vel, mass, nb, bpos, bcs, p, nsteps,
                             nullptr, nullptr, adaptive_last, ids);
}

extern "C" int oracle_sweep_f32(int n, const float *base4, const float *off4,
                                float *off4_out, int nb, const float *bpos,
                                const int *cs, const int *bcs,
                                const oracle_params *p, int solver_phase) {
  (void)nb;
  if (p->own_only) return -2;
  Grid g{p->cell_length, p->cx, p->cy, p->cz};
  SweepParams<float> sp = make_sweep_params<float>(p);
  std::vector<V3<float>> base(n), off(n);
  std::vector<float> rad(n);
  for (int i = 0; i < n; ++i) {
    base[i] = {base4[4 * i], base4[4 * i + 1], base4[4 * i + 2]};
    rad[i] = base4[4 * i + 3];
    off[i] = {off4[4 * i], off4[4 * i + 1], off4[4 * i + 2]};
  }
  const V3<float> *bp = reinterpret_cast<const V3<float> *>(bpos);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    V3<float> o = sweep_one<float, FREE>(i, base.data(), rad.data(), off.data(),
                                         bp, cs, bcs, g, sp, solver_phase != 0,
                                         nullptr, nullptr);
    off4_out[4 * i] = o.x;
    off4_out[4 * i + 1] = o.y;
    off4_out[4 * i + 2] = o.z;
    off4_out[4 * i + 3] = off4[4 * i + 3];
  }
  return 0;
}

extern "C" int oracle_nlist_f32(int n, const float *base4, const float *off4,
                                int nb, const float *bpos, const int *cs,
                                const int *bcs, const oracle_params *p,
                                int gmax, int bmax, int *gcount, int *glist,
                                int *bcount, int *blist, int *base_cell) {
  (void)nb;
  Grid g{p->cell_length, p->cx, p->cy, p->cz};
  SweepParams<float> sp = make_sweep_params<float>(p);
  const V3<float> *bp = reinterpret_cast<const V3<float> *>(bpos);
  const float skin = p->skin;
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    V3<float> bi{base4[4 * i], base4[4 * i + 1], base4[4 * i + 2]};
    V3<float> oi{off4[4 * i], off4[4 * i + 1], off4[4 * i + 2]};
    const float ri = base4[4 * i + 3];
    const float rbi = ri > sp.rmin ? ri : sp.rmin;
    V3<float> xs = add(bi, oi);
    int bx, by, bz;
    const bool okb = trunc_cell<float>(xs, g.L, bx, by, bz);
    base_cell[i] = okb ? ext_base_id(bx, by, bz, g) : -1;
    int gc = 0, bc = 0;
    if (okb) {
      for (int m = 0; m < 27; ++m) {
        const int cell = stencil_cell(bx, by, bz, m, g);
        if (cell < 0) continue;
        for (int k = bcs[cell]; k < bcs[cell + 1]; ++k) {
          V3<float> p3;
          p3.x = xs.x - bp[k].x;
          p3.y = xs.y - bp[k].y;
          p3.z = xs.z - bp[k].z;
          const float d2 = p3.x * p3.x + p3.y * p3.y + p3.z * p3.z;
          const float lim = (sp.rb + rbi) + skin;
          if (d2 < lim * lim) {
            if (bc < bmax) blist[(size_t)i * bmax + bc] = k;
            bc++;
          }
        }
        for (int j = cs[cell]; j < cs[cell + 1]; ++j) {
          if (j == i) continue;
          const V3<float> bj{base4[4 * j], base4[4 * j + 1], base4[4 * j + 2]};
          const V3<float> oj{off4[4 * j], off4[4 * j + 1], off4[4 * j + 2]};
          const V3<float> xsj = add(bj, oj);
          V3<float> p3;
          p3.x = xs.x - xsj.x;
          p3.y = xs.y - xsj.y;
          p3.z = xs.z - xsj.z;
          const float d2 = p3.x * p3.x + p3.y * p3.y + p3.z * p3.z;
          const float lim = (ri + base4[4 * j + 3]) + skin;
          if (d2 < lim * lim) {
            if (gc < gmax) glist[(size_t)i * gmax + gc] = j;
            gc++;
          }
        }
      }
    }
    gcount[i] = gc;
    bcount[i] = bc;
  }
  return 0;
}

extern "C" float oracle_radius_contact(float m, float density) {
  return radius_contact_t<float>(m, density);
}
extern "C" float oracle_radius_surface(float m, float density) {
  return radius_surface_f(m, density);
}
extern "C" float oracle_radius_split(float m, float density) {
  return radius_split_f(m, density);
}
extern "C" float oracle_expf(float x) { return exp_t<float>(x); }

extern "C" void oracle_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

extern "C" int oracle_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

extern "C" int oracle_shadow_steps(int n, float *pf, float *vf, float *mf,
                                   double *pd, double *vd, double *md, int nb,
                                   const float *bpf, const double *bpd,
                                   const int *bcs, const oracle_params *p,
                                   int nsteps, int *adaptive_last) {
  if (n < 0 || nb < 0 || !p) return -1;
  return shadow_steps(n, pf, vf, mf, pd, vd, md, nb, bpf, bpd, bcs, p, nsteps,
                      adaptive_last);
}

// Full adaptive steps: the plain step (the sort also permutes merge,
// merge_count and remove, GranularSystem.cu:99-135; adaptive_last_step and
// merge_velocity are not permuted), then set_surface_particles,
// find_distance_to_surface and Solver::adaptive_sampling
// (GranularSystem.cu:420-432, Solver.cu:856-1242) against the cell table of
// the start of the step.  merge_first = 0 skips the merge mark of the first
// step (the reference's t_merge_iter toggle, Solver.cu:910-922).
extern "C" int oracle_adaptive_steps(int *n_io, int cap, float *pos_, float *vel_, float *mass,
                                     int *adl, int *remove, float *merge, int *mcount,
                                     float *mv_, int *surf, float *sd, int nb,
                                     const float *bpos_, const int *bcs,
                                     const oracle_params *p, int nsteps, int merge_first,
                                     float max_mass, int *stats) {
  Grid g{p->cell_length, p->cx, p->cy, p->cz};
  const int C = g.cx * g.cy * g.cz;
  std::vector<int> cs(C + 2), idx, keys;
  V3<float> *pos = reinterpret_cast<V3<float> *>(pos_);
  V3<float> *vel = reinterpret_cast<V3<float> *>(vel_);
  V3<float> *mv = reinterpret_cast<V3<float> *>(mv_);
  const int S = p->stab_iters, I = p->solver_iters;
  (void)nb;
  for (int step = 0; step < nsteps; ++step) {
    const int n = *n_io;
    Sim<float> s;
    s.init(n, pos_, vel_, mass, bpos_, p);
    sort_perm(pos, n, g, idx, keys, cs.data());
    s.permute(idx);
    {
      std::vector<float> tm(n);
      std::vector<int> ti(n), tr(n);
      for (int k = 0; k < n; ++k) {
        tm[k] = merge[idx[k]];
        ti[k] = mcount[idx[k]];
        tr[k] = remove[idx[k]];
      }
      for (int k = 0; k < n; ++k) {
        merge[k] = tm[k];
        mcount[k] = ti[k];
        remove[k] = tr[k];
      }
    }
    s.predict(p);
    for (int k = 0; k < S; ++k) {
#pragma omp parallel for schedule(dynamic, 256)
      for (int i = 0; i < n; ++i)
        s.b[i] = sweep_one<float, FREE>(i, s.pos, s.rad.data(), s.a.data(), s.bpos, cs.data(),
                                        bcs, g, s.sp, false, nullptr, s.ims.data());
      std::swap(s.a, s.b);
    }
    s.end_stabilisation(S);
    for (int k = 0; k < I; ++k) {
#pragma omp parallel for schedule(dynamic, 256)
      for (int i = 0; i < n; ++i)
        s.b[i] = sweep_one<float, FREE>(i, s.xb.data(), s.rad.data(), s.a.data(), s.bpos,
                                        cs.data(), bcs, g, s.sp, true, nullptr, s.ims.data());
      std::swap(s.a, s.b);
    }
    for (int i = 0; i < n; ++i) {
      uint8_t bits = 0;
      finalize_one<float, FREE>(i, s.pos, s.vel, s.xb.data(), s.a.data(), adl, s.sp, p->dt,
                                bits);
    }
    for (int i = 0; i < n; ++i) adl[i] = 0;
    std::vector<V3<float>> cover(n);
    surface(n, pos, mass, cs.data(), g, p->density, cover.data(), surf);
    surface_distance(n, pos, surf, cs.data(), g, sd);
    if (step > 0 || merge_first)
      merge_mark(n, pos, mass, vel, surf, sd, remove, merge, mv, cs.data(), g, max_mass,
                 rules_of(p), nullptr, nullptr);
    merge_blend(n, merge, mass, vel, mcount, remove, mv, 40, adl);
    int st[3] = {0, 0, 0};
    *n_io = oracle_split_compact(n, cap, pos_, vel_, mass, adl, remove, merge, mcount, surf,
                                 sd, bpos_, cs.data(), bcs, p, st);
    if (stats) {
      stats[3 * step] = st[0];
      stats[3 * step + 1] = st[1];
      stats[3 * step + 2] = n;
    }
  }
  return 0;
}
