/*
 * CPU oracle for the granular PBD step -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a from-scratch C++/OpenMP restatement of the reference solver's
 * algorithm (stingyemperor/granular_gpu, src/GranularSystem.cu, src/Solver.cu,
 * src/CUDAFunctions.cuh), written to check the HIP implementation.  Only the
 * test suite, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product library never links, loads or calls anything here.
 *
 * Two precisions:
 *   *_f32 follows the exact FP32 operation order of the deterministic
 *         semantics (no FMA contraction, IEEE division and square root), so a
 *         conforming GPU implementation matches it bit for bit;
 *   *_f64 runs the same algorithm in double precision and is the accuracy
 *         reference for positions (1e-5*r after 1 step, 1e-3*r after 10).
 *
 * Array layout: float3 / double3 quantities are packed AoS (x,y,z per
 * particle), exactly as the reference's float3 arrays (Particles.hpp:19-20).
 * Cell tables have C+2 entries: cell_start[c] for c in [0, C] is the
 * reference's cellStart (C = sentinel cell of out-of-grid particles) and
 * cell_start[C+1] = n.
 */
#ifndef GRANULAR_ORACLE_H
#define GRANULAR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_params {
  float dt;
  float gx, gy, gz;
  float cell_length;
  int cx, cy, cz;
  float density;            /* scene density; contacts use (int)density     */
  int stab_iters;           /* reference: 3  (Solver.cu:428)                */
  int solver_iters;         /* reference: 5  (Solver.cu:99)                 */
  float mu_s, mu_k;         /* PBD friction; 0/0 = reference (inert clamp)  */
  float boundary_radius;    /* 0.01 (Solver.cu:283)                          */
  float contact_radius_min; /* 0.01 clamp of r_i vs boundary (Solver.cu:281) */
  float sleep_speed;        /* 0.01 (Solver.cu:223)                          */
  float damp_speed;         /* 1.0  (Solver.cu:202)                          */
  float damping;            /* 0.99 (Solver.cu:204-205)                      */
  float adaptive_damping;   /* 0.98 (Solver.cu:242)                          */
  float skin;               /* neighbour-list skin (only for oracle_nlist)   */
  int own_only;             /* 0: symmetric gather; 1: own term only        */
  /* merge / split rules (oracle_merge_mark, oracle_split_compact) */
  int depth_rules;          /* 0: reference rules; 1: depth rules             */
  float min_mass;           /* merging needs m >= min_mass (reference 1)      */
  float split_mass;         /* splitting needs m > split_mass (reference 2)   */
  float merge_dist;         /* partners within merge_dist (reference 0.035 m);
                               depth rules: within merge_dist (r_i + r_j)     */
  float merge_depth;        /* depth rules: merge when sd > (merge_depth r)^2 */
  float split_depth;        /* depth rules: split when sd < (split_depth r)^2 */
} oracle_params;

/* Grid build: hash, stable sort by cell id, cell_start (C+2 entries).
 * perm[k] = old index of the particle placed at sorted slot k.            */
int oracle_grid_build(const float *pos, int n, float cell_length, int cx,
                      int cy, int cz, int *keys_sorted, int *perm,
                      int *cell_start);
int oracle_hash(const float *pos, int n, float cell_length, int cx, int cy,
                int cz, int *keys);

/* Full plain step(s) (grid build, predict, S+I contact sweeps, finalize) on
 * host arrays, in place.  Boundary arrays must already be cell-sorted with
 * their cell_start (C+2).  Optional outputs may be NULL.                   */
int oracle_plain_steps_f32(int n, float *pos, float *vel, float *mass,
                           int nb, const float *bpos, const int *bcell_start,
                           const oracle_params *p, int nsteps,
                           float *scaled_mass_out, int *p2c_out,
                           int *adaptive_last_step);
int oracle_plain_steps_f64(int n, double *pos, double *vel, double *mass,
                           int nb, const double *bpos, const int *bcell_start,
                           const oracle_params *p, int nsteps,
                           int *adaptive_last_step, int *ids);

/* Lockstep FP32 + FP64 run.  The FP64 state takes every discrete decision
 * (sort order, stencil base, contact yes/no, friction branch, sleep and
 * damping branch) from the FP32 state, so both follow the same discrete path
 * and the difference is the rounding error of the FP32 arithmetic.  (An
 * independent FP64 run can take different discrete decisions: pairs resolved
 * to exactly d = r_i + r_j sit on the contact-count threshold.)            */
int oracle_shadow_steps(int n, float *pos32, float *vel32, float *mass32,
                        double *pos64, double *vel64, double *mass64, int nb,
                        const float *bpos32, const double *bpos64,
                        const int *bcell_start, const oracle_params *p,
                        int nsteps, int *adaptive_last_step);

/* One contact sweep in FP32 on explicit base/offset arrays (stage test).
 * base4/off4 are float4 records (x,y,z,w): base w = radius.               */
int oracle_sweep_f32(int n, const float *base4, const float *off4,
                     float *off4_out, int nb, const float *bpos,
                     const int *cell_start, const int *bcell_start,
                     const oracle_params *p, int solver_phase);

/* Neighbour candidate lists (granular + boundary) in scan order, as the GPU
 * builds them: candidate iff in the stencil of trunc(xstar / L) and
 * |p|^2 < lim^2 with p = xstar_i - xstar_j (xstar = fl(base + off)) or
 * xstar_i - q, lim = r_i + r_j + skin (granular) or R_b + max(r_i,rmin) +
 * skin (boundary).  base_cell = padded stencil-base id.
 * Lists are written row-major [i][k]; counts may exceed gmax/bmax (then
 * only the first gmax/bmax entries are stored).                           */
int oracle_nlist_f32(int n, const float *base4, const float *off4, int nb,
                     const float *bpos, const int *cell_start,
                     const int *bcell_start, const oracle_params *p, int gmax,
                     int bmax, int *gcount, int *glist, int *bcount,
                     int *blist, int *base_cell);

/* ---- adaptive merge/split (GranularSystem.cu:506-699, Solver.cu:526-1242) */
/* Surface flags (K5+K6): cover vector and surface = n <= 5 or no neighbour
 * inside the pi/3 cone around the cover direction (tested as dot >= 0.5). */
int oracle_surface(int n, const float *pos, const float *mass, const int *cs,
                   const oracle_params *p, float *cover, int *surface);
/* K7: squared distance to the nearest surface particle of the stencil.    */
int oracle_surface_distance(int n, const float *pos, const int *surface,
                            const int *cs, const oracle_params *p, float *sd);
/* K8 with the pairing resolved as if threads ran in ascending index:
 * candidates from the pre-kernel snapshot of remove[]; i claims itself
 * unless already claimed, then its candidate unless already claimed.
 * cand/dens (optional) receive each particle's candidate and density.     */
int oracle_merge_mark(int n, const float *pos, const float *mass,
                      const float *vel, const int *surface, const float *sd,
                      int *remove, float *merge, float *merge_velocity,
                      const int *cs, const oracle_params *p, float max_mass,
                      int *cand, float *dens);
/* K9: 40-step mass/velocity blend.                                         */
int oracle_merge_blend(int n, float *merge, float *mass, float *vel,
                       int *merge_count, int *remove,
                       const float *merge_velocity, int blend, int *adl);
/* K10/K11 + compaction + append; returns the new particle count.  sd (the
 * squared surface distances) is read by the depth rules only (may be NULL
 * with the reference rules).                                              */
int oracle_split_compact(int n, int cap, float *pos, float *vel, float *mass,
                         int *adl, int *remove, float *merge,
                         int *merge_count, const int *surface, const float *sd,
                         const float *bpos, const int *cs, const int *bcs,
                         const oracle_params *p, int *stats);
/* Whole adaptive steps (plain step + the passes above).  stats[3k..3k+2] =
 * (removed, split, count before) of step k.                               */
int oracle_adaptive_steps(int *n, int cap, float *pos, float *vel, float *mass,
                          int *adl, int *remove, float *merge,
                          int *merge_count, float *merge_velocity,
                          int *surface, float *sd, int nb, const float *bpos,
                          const int *bcell_start, const oracle_params *p,
                          int nsteps, int merge_first, float max_mass,
                          int *stats);

/* ---- plain upsampler (update_upsampled_cuda, Solver.cu:1244-1369) ------ */
typedef struct oracle_upsample_params {
  float dt, gx, gy, gz;    /* 0.002, (0,-9.8,0): hard-coded in the reference */
  float r9;                /* 1/h^2 = 493.8271605                             */
  float boundary_radius;   /* 0.015 */
  float repulsion;         /* 3.0   */
  float min_weight;        /* 1e-6  */
  float near_damping;      /* 0.8   */
  float ground_y;          /* 0.1   */
  float ground_damping;    /* 0.5   */
  float max_velocity;      /* 7.0   */
  float floor_y;           /* 0.04  */
  float floor_friction;    /* 0.95  */
} oracle_upsample_params;
void oracle_upsample_params_default(oracle_upsample_params *u);
/* One upsampler pass over nf fine particles (in place) against the coarse
 * positions/velocities and their cell table (stencil base floor(x/L)).    */
int oracle_upsample_plain(int nf, float *fpos, float *fvel, int n,
                          const float *cpos, const float *cvel, const int *cs,
                          int nb, const float *bpos, const int *bcs,
                          const oracle_params *p, const oracle_upsample_params *u);

/* Radius helpers (three reference variants), exposed for unit tests.     */
float oracle_radius_contact(float m, float density);
float oracle_radius_surface(float m, float density);
float oracle_radius_split(float m, float density);
float oracle_expf(float x);

/* ---- kinematic boundary (oracle_motion.cpp) ------------------------------ */
/* Same layout as ggh_animation.                                           */
typedef struct oracle_animation {
  int is_animating;
  float animation_time;
  int start_index;
  int num_particles;
  float translation_direction[3];
  float translation_speed;
  int translation_complete;
  float rotation_angle;
  float rotation_speed;
  float rotation_center[3];
  float translation_time;
  float rotation_limit;
  int center_from_marked;
} oracle_animation;
void oracle_rotation_cs(float angle_deg, float *c, float *s);
/* animateParticles on n AoS positions: marked (== 1) particles translate,
 * then rotate by angle_deg about z through (cx, cy) if angle_deg != 0.    */
void oracle_animate(int n, float *pos, const int *animated, float tx, float ty, float tz,
                    float angle_deg, float cx, float cy);
/* One updateAnimation(dt) call; returns 1 if the positions moved.         */
int oracle_update_animation(oracle_animation *a, int n, float *pos, const int *animated,
                            float dt);

/* ---- interactive extras (oracle_motion.cpp) ------------------------------ */
void oracle_center_of_mass(int n, const float *pos, const float *mass, float *com);
void oracle_explode(int n, const float *pos, float *vel, const float *mass, const float *com,
                    float force);
int oracle_pick_sphere(int n, const float *pos, const float *center, float radius, int *out);
void oracle_move_picked(float *pos, const int *idx, int count, const float *delta,
                        const float *lo, const float *hi);

/* Number of OpenMP threads the oracle uses (0 = default).                */
void oracle_set_threads(int t);
int oracle_get_threads(void);

/* ---- learned upsampler (see oracle_mlp.cpp for the definitions) ---------- */
typedef struct oracle_mlp_cfg {
  float radius;          /* coarse radius R: position feature scale, template size */
  float vscale;          /* velocity feature scale V (m/s)                          */
  float guard_boundary;  /* boundary particle closer than this: fallback            */
  float guard_floor;     /* y below this: fallback                                  */
  float guard_side;      /* x or z below this: fallback                             */
} oracle_mlp_cfg;
float oracle_bf16_round(float x);
int oracle_mlp_features(int n, const float *pos, const float *vel, const int *cs, int nb,
                        const float *bpos, const int *bcs, const oracle_params *p,
                        const oracle_mlp_cfg *c, float *feat, int *flags, int *knn);
int oracle_mlp_teacher(int n, const float *pos, const float *vel, const int *cs,
                       const float *bpos, const int *bcs, const oracle_params *p,
                       const oracle_upsample_params *u, const oracle_mlp_cfg *c, float *y,
                       float *fine);
int oracle_mlp_fine_targets(int n, const float *pos, const float *vel, const float *fpos,
                            const float *fvel, const int *fcs, float fL, int fcx, int fcy,
                            int fcz, float radius, float *y, float *d2, int *jm);
int oracle_mlp_positions(int n, const float *pos, const float *vel, const float *y,
                         const oracle_mlp_cfg *c, float dt, float *fine);
int oracle_mlp_forward(int B, const float *x, const float *W, const float *b, float *y,
                       int exact);
double oracle_mlp_backward(int B, const float *x, const float *W, const float *b,
                           const float *target, float *dW, float *db, int exact);
int oracle_mlp_adam(int count, float *w, float *m, float *v, const float *g, float lr,
                    float beta1, float beta2, float eps, float c1, float c2);

#ifdef __cplusplus
}
#endif
#endif
