// CPU oracle of the learned upsampler -- TEST INFRASTRUCTURE ONLY.
//
// Definitions (no reference code: the learned upsampler is a specification
// addition beside update_upsampled_cuda, Solver.cu:1244-1387):
//
// Features of coarse particle i: the 16 coarse particles j != i of the
// 27-cell stencil of floor(x_i/L) with the smallest (|x_j - x_i|^2, j),
// ascending; slot s holds ((x_j - x_i) fl(1/R), (v_j - v_i) fl(1/V)) in the
// particle's co-moving frame (R = coarse radius, V = velocity scale, one
// rounded reciprocal each); empty slots are zero; every feature is rounded
// to bfloat16.
//
// Fallback flags: bit 0 = a boundary particle of the stencil closer than
// guard_boundary; bit 1 = y_i < guard_floor, x_i < guard_side or
// z_i < guard_side.  Flagged particles take the teacher's output.
//
// Teacher: the plain upsampler's update (upsample_one) applied to the 8
// template points x_f = x_i + R T_k, T_k = (+-1/2, +-1/2, +-1/2) with
// k = 4 sx + 2 sy + sz, starting at the parent's velocity v_i.  Target
// y_k = v_f' - v_i (the fine velocity relative to the parent).
//
// Network: 96 -> 128 -> 128 -> 128 -> 24, ReLU; weights of layer l stored
// [n_out][k_in] (rows of layer 4 padded to 32 with zeros) at offsets
// 0, 12288, 28672, 45056; biases at 0, 128, 256, 384.  Mixed precision as
// on the GPU: layer inputs and weights in bfloat16, products accumulated in
// FP32 (here: double, rounded to float at the end of the sum), biases FP32.
//   h0 = x; z_l = h_{l-1} W_l^T + b_l; h_l = bf16(relu(z_l)); y = z_4.
// Loss L = mean over B x 24 of (y - t)^2.  Backward:
//   g4 = 2 (y - t) / (24 B); G_l = bf16(g_l);
//   dW_l[n][k] = sum_b G_l[b][n] h_{l-1}[b][k]; db_l[n] = sum_b g_l[b][n];
//   g_{l-1} = (G_l W_l)[.][k] * [z_{l-1} > 0].
// exact = 1 switches every bfloat16 rounding off (finite-difference checks).
// Fine position of slot k: fl(fl(x_i + R T_k) + dt (v_i + y_k)).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"
#include "oracle_internal.hpp"

using namespace orc;

namespace {

constexpr int kIn = 96, kH = 128, kOut = 24, kOutPad = 32, kKnn = 16;
constexpr int kWOff[4] = {0, 12288, 28672, 45056};
constexpr int kBOff[4] = {0, 128, 256, 384};
constexpr int kDimIn[4] = {kIn, kH, kH, kH};
constexpr int kDimOut[4] = {kH, kH, kH, kOutPad};

// Round to the nearest bfloat16 (ties to even); NaN stays NaN.
float bf16r(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) {
    u |= 0x00400000u;
  } else {
    u += 0x7fffu + ((u >> 16) & 1u);
  }
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

inline float rnd(float x, int exact) { return exact ? x : bf16r(x); }

V3<float> templ(int k) {
  return V3<float>{(k & 4) ? 0.5f : -0.5f, (k & 2) ? 0.5f : -0.5f, (k & 1) ? 0.5f : -0.5f};
}

struct Net {
  // forward of one sample; keeps pre-activations and layer inputs
  double z[4][kOutPad > kH ? kOutPad : kH];
  float h[4][kH];  // h[l] = input of layer l (h[0] = features, 96 used)
};

void forward_one(const float *x, const float *W, const float *b, int exact, Net &t, float *y) {
  for (int k = 0; k < kIn; ++k) t.h[0][k] = rnd(x[k], exact);
  for (int l = 0; l < 4; ++l) {
    const int K = kDimIn[l], N = (l == 3) ? kOut : kDimOut[l];
    for (int n = 0; n < N; ++n) {
      double acc = 0.0;
      const float *w = W + kWOff[l] + (size_t)n * K;
      for (int k = 0; k < K; ++k) acc += (double)t.h[l][k] * (double)rnd(w[k], exact);
      const float zf = (float)acc + b[kBOff[l] + n];
      t.z[l][n] = zf;
      if (l < 3) {
        t.h[l + 1][n] = rnd(zf > 0.0f ? zf : 0.0f, exact);
      } else {
        y[n] = zf;
      }
    }
  }
}

}  // namespace

extern "C" float oracle_bf16_round(float x) { return bf16r(x); }

extern "C" int oracle_mlp_features(int n, const float *pos_, const float *vel_, const int *cs,
                                   int nb, const float *bpos_, const int *bcs,
                                   const oracle_params *p, const oracle_mlp_cfg *c, float *feat,
                                   int *flags, int *knn) {
  (void)nb;
  Grid g{p->cell_length, p->cx, p->cy, p->cz};
  const V3<float> *pos = reinterpret_cast<const V3<float> *>(pos_);
  const V3<float> *vel = reinterpret_cast<const V3<float> *>(vel_);
  const V3<float> *bpos = reinterpret_cast<const V3<float> *>(bpos_);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    const V3<float> xi = pos[i], vi = vel[i];
    float kd[kKnn];
    int kj[kKnn];
    for (int s = 0; s < kKnn; ++s) {
      kd[s] = INFINITY;
      kj[s] = -1;
    }
    int flag = 0;
    const float gb2 = c->guard_boundary * c->guard_boundary;
    int bx, by, bz;
    if (floor_cell<float>(xi, g.L, bx, by, bz)) {
      for (int m = 0; m < 27; ++m) {
        const int cell = stencil_cell(bx, by, bz, m, g);
        if (cell < 0) continue;
        for (int j = cs[cell]; j < cs[cell + 1]; ++j) {
          if (j == i) continue;
          const float dx = pos[j].x - xi.x, dy = pos[j].y - xi.y, dz = pos[j].z - xi.z;
          const float d2 = dx * dx + dy * dy + dz * dz;
          // insert (d2, j) if it beats the current 16th
          if (d2 < kd[kKnn - 1] || (d2 == kd[kKnn - 1] && j < kj[kKnn - 1])) {
            int s = kKnn - 1;
            while (s > 0 && (d2 < kd[s - 1] || (d2 == kd[s - 1] && j < kj[s - 1]))) {
              kd[s] = kd[s - 1];
              kj[s] = kj[s - 1];
              --s;
            }
            kd[s] = d2;
            kj[s] = j;
          }
        }
        for (int k = bcs[cell]; k < bcs[cell + 1]; ++k) {
          const float dx = bpos[k].x - xi.x, dy = bpos[k].y - xi.y, dz = bpos[k].z - xi.z;
          if (dx * dx + dy * dy + dz * dz < gb2) flag |= 1;
        }
      }
    }
    if (xi.y < c->guard_floor || xi.x < c->guard_side || xi.z < c->guard_side) flag |= 2;
    float *f = feat + (size_t)i * kIn;
    const float inv_r = 1.0f / c->radius, inv_v = 1.0f / c->vscale;
    for (int s = 0; s < kKnn; ++s) {
      const int j = kj[s];
      if (knn) knn[(size_t)i * kKnn + s] = j;
      if (j < 0) {
        for (int q = 0; q < 6; ++q) f[6 * s + q] = 0.0f;
        continue;
      }
      f[6 * s + 0] = bf16r((pos[j].x - xi.x) * inv_r);
      f[6 * s + 1] = bf16r((pos[j].y - xi.y) * inv_r);
      f[6 * s + 2] = bf16r((pos[j].z - xi.z) * inv_r);
      f[6 * s + 3] = bf16r((vel[j].x - vi.x) * inv_v);
      f[6 * s + 4] = bf16r((vel[j].y - vi.y) * inv_v);
      f[6 * s + 5] = bf16r((vel[j].z - vi.z) * inv_v);
    }
    flags[i] = flag;
  }
  return 0;
}

extern "C" int oracle_mlp_teacher(int n, const float *pos_, const float *vel_, const int *cs,
                                  const float *bpos_, const int *bcs, const oracle_params *p,
                                  const oracle_upsample_params *u, const oracle_mlp_cfg *c,
                                  float *y, float *fine) {
  Grid g{p->cell_length, p->cx, p->cy, p->cz};
  const V3<float> *pos = reinterpret_cast<const V3<float> *>(pos_);
  const V3<float> *vel = reinterpret_cast<const V3<float> *>(vel_);
  const V3<float> *bpos = reinterpret_cast<const V3<float> *>(bpos_);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    const V3<float> xi = pos[i], vi = vel[i];
    for (int k = 0; k < 8; ++k) {
      const V3<float> t = templ(k);
      const V3<float> xf{xi.x + c->radius * t.x, xi.y + c->radius * t.y, xi.z + c->radius * t.z};
      V3<float> np, v;
      upsample_one(xf, vi, pos, vel, cs, bpos, bcs, g, u, np, v);
      if (y) {
        y[(size_t)i * kOut + 3 * k + 0] = v.x - vi.x;
        y[(size_t)i * kOut + 3 * k + 1] = v.y - vi.y;
        y[(size_t)i * kOut + 3 * k + 2] = v.z - vi.z;
      }
      if (fine) {
        fine[(size_t)i * kOut + 3 * k + 0] = np.x;
        fine[(size_t)i * kOut + 3 * k + 1] = np.y;
        fine[(size_t)i * kOut + 3 * k + 2] = np.z;
      }
    }
  }
  return 0;
}

// Fine-run targets (the learned upsampler's training pairs from a coarse and
// a fine plain run, no reference counterpart): for coarse particle i and
// template slot k the fine particle j nearest to x_i + radius T_k among the
// fine particles of the 27 fine cells around floor(point / L_f), ties to the
// lower index; y = v_fine(j) - v_i, d2 = its squared distance, jm = j; no
// candidate: y = 0, d2 = +inf, jm = -1.  fcs is the fine cell table over
// the fine grid (fg_* its cell length and counts).
extern "C" int oracle_mlp_fine_targets(int n, const float *pos_, const float *vel_,
                                       const float *fpos_, const float *fvel_, const int *fcs,
                                       float fL, int fcx, int fcy, int fcz, float radius,
                                       float *y, float *d2, int *jm) {
  Grid g{fL, fcx, fcy, fcz};
  const V3<float> *pos = reinterpret_cast<const V3<float> *>(pos_);
  const V3<float> *vel = reinterpret_cast<const V3<float> *>(vel_);
  const V3<float> *fpos = reinterpret_cast<const V3<float> *>(fpos_);
  const V3<float> *fvel = reinterpret_cast<const V3<float> *>(fvel_);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    const V3<float> xi = pos[i], vi = vel[i];
    for (int k = 0; k < 8; ++k) {
      const V3<float> t = templ(k);
      const V3<float> q{xi.x + radius * t.x, xi.y + radius * t.y, xi.z + radius * t.z};
      float best = INFINITY;
      int bj = -1;
      int bx, by, bz;
      if (floor_cell<float>(q, g.L, bx, by, bz)) {
        for (int m = 0; m < 27; ++m) {
          const int cell = stencil_cell(bx, by, bz, m, g);
          if (cell < 0) continue;
          for (int j = fcs[cell]; j < fcs[cell + 1]; ++j) {
            const float dx = fpos[j].x - q.x, dy = fpos[j].y - q.y, dz = fpos[j].z - q.z;
            const float dd = dx * dx + dy * dy + dz * dz;
            if (dd < best || (dd == best && j < bj)) {
              best = dd;
              bj = j;
            }
          }
        }
      }
      float *o = y + (size_t)i * kOut + 3 * k;
      if (bj < 0) {
        o[0] = o[1] = o[2] = 0.0f;
      } else {
        o[0] = fvel[bj].x - vi.x;
        o[1] = fvel[bj].y - vi.y;
        o[2] = fvel[bj].z - vi.z;
      }
      if (d2) d2[(size_t)i * 8 + k] = best;
      if (jm) jm[(size_t)i * 8 + k] = bj;
    }
  }
  return 0;
}

extern "C" int oracle_mlp_positions(int n, const float *pos_, const float *vel_, const float *y,
                                    const oracle_mlp_cfg *c, float dt, float *fine) {
  const V3<float> *pos = reinterpret_cast<const V3<float> *>(pos_);
  const V3<float> *vel = reinterpret_cast<const V3<float> *>(vel_);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 8; ++k) {
      const V3<float> t = templ(k);
      const float *yk = y + (size_t)i * kOut + 3 * k;
      float *o = fine + (size_t)i * kOut + 3 * k;
      o[0] = (pos[i].x + c->radius * t.x) + dt * (vel[i].x + yk[0]);
      o[1] = (pos[i].y + c->radius * t.y) + dt * (vel[i].y + yk[1]);
      o[2] = (pos[i].z + c->radius * t.z) + dt * (vel[i].z + yk[2]);
    }
  return 0;
}

extern "C" int oracle_mlp_forward(int B, const float *x, const float *W, const float *b, float *y,
                                  int exact) {
#pragma omp parallel for schedule(static)
  for (int s = 0; s < B; ++s) {
    Net t;
    forward_one(x + (size_t)s * kIn, W, b, exact, t, y + (size_t)s * kOut);
  }
  return 0;
}

extern "C" double oracle_mlp_backward(int B, const float *x, const float *W, const float *b,
                                      const float *target, float *dW, float *db, int exact) {
  std::vector<Net> nets((size_t)B);
  std::vector<float> y((size_t)B * kOut);
  double loss = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : loss)
  for (int s = 0; s < B; ++s) {
    forward_one(x + (size_t)s * kIn, W, b, exact, nets[(size_t)s], y.data() + (size_t)s * kOut);
    for (int n = 0; n < kOut; ++n) {
      const double e = (double)y[(size_t)s * kOut + n] - (double)target[(size_t)s * kOut + n];
      loss += e * e;
    }
  }
  loss /= (double)B * kOut;
  // g[b][n] per layer, fp32 as on the GPU
  std::vector<float> g((size_t)B * kH), gp((size_t)B * kH);
  const float scale = 2.0f / (float)(kOut * B);
  for (int s = 0; s < B; ++s)
    for (int n = 0; n < kOutPad; ++n)
      g[(size_t)s * kH + n] =
          n < kOut ? (y[(size_t)s * kOut + n] - target[(size_t)s * kOut + n]) * scale : 0.0f;
  std::memset(dW, 0, sizeof(float) * 49152);
  std::memset(db, 0, sizeof(float) * 416);
  for (int l = 3; l >= 0; --l) {
    const int K = kDimIn[l], N = kDimOut[l];
    // dW, db
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
      double dbs = 0.0;
      std::vector<double> acc((size_t)K, 0.0);
      for (int s = 0; s < B; ++s) {
        const float gv = g[(size_t)s * kH + n];
        dbs += gv;
        const double G = rnd(gv, exact);
        if (G == 0.0) continue;
        for (int k = 0; k < K; ++k) acc[(size_t)k] += G * (double)nets[(size_t)s].h[l][k];
      }
      for (int k = 0; k < K; ++k) dW[kWOff[l] + (size_t)n * K + k] = (float)acc[(size_t)k];
      db[kBOff[l] + n] = (float)dbs;
    }
    if (l == 0) break;
    // g_{l-1} = (G_l W_l) * [z_{l-1} > 0]
#pragma omp parallel for schedule(static)
    for (int s = 0; s < B; ++s)
      for (int k = 0; k < K; ++k) {
        double acc = 0.0;
        for (int n = 0; n < N; ++n)
          acc += (double)rnd(g[(size_t)s * kH + n], exact) *
                 (double)rnd(W[kWOff[l] + (size_t)n * K + k], exact);
        gp[(size_t)s * kH + k] = nets[(size_t)s].z[l - 1][k] > 0.0 ? (float)acc : 0.0f;
      }
    std::swap(g, gp);
  }
  return loss;
}

extern "C" int oracle_mlp_adam(int count, float *w, float *m, float *v, const float *g, float lr,
                               float beta1, float beta2, float eps, float c1, float c2) {
  for (int i = 0; i < count; ++i) {
    m[i] = beta1 * m[i] + (1.0f - beta1) * g[i];
    v[i] = beta2 * v[i] + (1.0f - beta2) * (g[i] * g[i]);
    const float mh = m[i] * c1, vh = v[i] * c2;
    w[i] = w[i] - (lr * mh) / (std::sqrt(vh) + eps);
  }
  return 0;
}
