/*
 * granular_gpu.h -- C-ABI of the MI355X-native granular PBD solver.
 *
 * Every entry point is extern "C", takes plain device pointers, sizes and
 * small parameter structs, returns GG_OK (0) or a negative GG_ERR_* code, and
 * never throws.  Nothing here allocates device memory: scratch space is a
 * caller-owned workspace sized by gg_workspace_bytes().  All work is enqueued
 * on the given HIP stream (NULL = the null stream) and is asynchronous unless
 * stated otherwise.
 *
 * The reference (stingyemperor/granular_gpu) has no C-ABI of its own: its
 * boundary is the C++ class API (GranularSystem / Solver / GranularParticles /
 * DArray) whose methods launch raw-pointer CUDA kernels.  Each entry point
 * below names the reference functions and kernels it stands in for.
 *
 * Array conventions (same as the reference):
 *   - positions / velocities are packed float3 (12 bytes, AoS);
 *   - a cell table has C+2 ints for a grid of C = cx*cy*cz cells:
 *     cell_start[c] for c in [0, C] is the reference's cellStart array
 *     (entry C is the sentinel cell of particles outside the grid) and
 *     cell_start[C+1] = n;
 *   - cell ids are x-major: ((ix*cy)+iy)*cz+iz with ix = floor(x/L).
 */
#ifndef GRANULAR_GPU_H
#define GRANULAR_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same type as hip_runtime_api.h's hipStream_t. */
typedef struct ihipStream_t *hipStream_t;

enum {
  GG_OK = 0,
  GG_ERR_INVALID = -1,   /* bad argument (NULL pointer, negative size ...) */
  GG_ERR_CAPACITY = -2,  /* n exceeds the capacity the workspace was sized for */
  GG_ERR_HIP = -3,       /* a HIP runtime call failed */
  GG_ERR_WORKSPACE = -4, /* workspace too small or not initialised */
  GG_ERR_NO_DEVICE = -5  /* no usable gfx950 device */
};

typedef struct gg_float3 {
  float x, y, z;
} gg_float3;

/* Uniform grid: cell length L and cell counts.  Reference: _cell_length,
 * _cell_size (GranularSystem.hpp:247-253), cell_size = ceil(space/L)
 * (main.cu:158-160). */
typedef struct gg_grid {
  float cell_length;
  int cx, cy, cz;
} gg_grid;

/* Step parameters.  gg_params_default() fills the reference's constants. */
typedef struct gg_params {
  float dt;                 /* time step (scene JSON "dt")                 */
  float gx, gy, gz;         /* gravity (scene JSON "gravity")              */
  float density;            /* scene JSON "density"                         */
  int stab_iters;           /* stabilisation sweeps, reference 3            */
  int solver_iters;         /* solver sweeps, reference 5                   */
  float mu_s, mu_k;         /* PBD friction; 0,0 reproduces the reference   */
  float boundary_radius;    /* 0.01  (Solver.cu:283)                        */
  float contact_radius_min; /* 0.01  (Solver.cu:281)                        */
  float sleep_speed;        /* 0.01  (Solver.cu:223)                        */
  float damp_speed;         /* 1.0   (Solver.cu:202)                        */
  float damping;            /* 0.99  (Solver.cu:204-205)                    */
  float adaptive_damping;   /* 0.98  (Solver.cu:242)                        */
  float skin;               /* neighbour-list skin; <= 0 disables lists     */
  int flags;                /* GG_FLAG_*                                    */
  int sweep_block;          /* threads per block of the contact sweeps:
                               64, 128, 256 or 512; 0 = 64 (autotuned by
                               gg_batch_autotune)                           */
  int nlist_block;          /* same for the neighbour-list build; 0 = 64    */
  /* Merge / split rules of gg_adaptive_sampling and gg_split_compact
   * (gg_merge_mark applies the reference rules).  Reference rules
   * (Solver.cu:526-653, 780-827): merge partners deep (not surface, no
   * surface particle in the stencil) within merge_dist = 0.035 m, merging
   * needs min_mass <= m < max_mass, splitting a surface particle with
   * m > split_mass = 2.  Depth rules (radius levels, r the radius of the
   * mass): merge when sd > (merge_depth r)^2 with partners within
   * merge_dist (r_i + r_j), split when sd < (split_depth r)^2 and
   * m > split_mass; sd is the squared distance to the surface. */
  int adaptive_rules;       /* 0 reference rules, 1 depth rules             */
  float min_mass;           /* 1.0                                          */
  float split_mass;         /* 2.0                                          */
  float merge_dist;         /* 0.035 m (depth rules: a factor, e.g. 1.75)   */
  float merge_depth;        /* 3.0 (depth rules)                            */
  float split_depth;        /* 2.0 (depth rules)                            */
} gg_params;

#define GG_FLAG_NO_NLIST 1     /* always scan the 27-cell stencil */
#define GG_FLAG_NO_XCD_REMAP 2 /* keep hardware block order (no XCD grouping) */
/* Particles whose neighbour list is not valid for a sweep scan the 27-cell
 * stencil instead: one wave per particle when few lanes of a wave need it
 * (queued in the block and scanned by its waves after their list work),
 * one lane per particle otherwise.  LANE forces every such particle to scan
 * in its lane; WAVE queues every one sweep-wide for a second pass after the
 * sweep.  Results are the same bit for bit. */
#define GG_FLAG_FALLBACK_LANE 4
#define GG_FLAG_FALLBACK_WAVE 8
/* Lists that lapse inside a step (particles driven far by their contacts,
 * e.g. columns pushed through their floors) are rebuilt from the current
 * positions after a sweep (any but the first and the last) that scanned the
 * stencil for more particles than the first sweep by more than n/(8 R), R
 * the sweeps left in the step: when the scans saved outweigh one list
 * build (decided on the device, so the step stays one fixed sequence of
 * launches).  Results are the same bit for bit; each refresh point adds two
 * launches, which exit at once when the lists still hold. */
#define GG_FLAG_LIST_REFRESH 16

/* Device arrays of one particle set.  Mirrors GranularParticles
 * (GranularParticles.hpp:5-110) + Particles (Particles.hpp:7-62).
 * Arrays that a stage does not use may be NULL. */
typedef struct gg_particles {
  int n;
  int capacity;
  gg_float3 *pos;
  gg_float3 *vel;
  float *mass;
  float *scaled_mass;
  int *surface;
  float *surface_distance;
  int *num_surface;
  int *particle_2_cell;
  int *is_animated;
  int *adaptive_last_step;
} gg_particles;

/* Merge/split bookkeeping.  Mirrors Solver's _buffer_remove, _buffer_merge,
 * _buffer_merge_velocity, _buffer_merge_count (Solver.hpp:108-112). */
typedef struct gg_adaptive {
  int *remove;
  float *merge;
  gg_float3 *merge_velocity;
  int *merge_count;
} gg_adaptive;

/* Caller-owned scratch.  Size with gg_workspace_bytes(), allocate device
 * memory of that size, fill the struct, call gg_workspace_init() once. */
typedef struct gg_workspace {
  void *base;            /* device pointer, 256-byte aligned */
  size_t bytes;
  int capacity;          /* max particles of the granular set */
  int boundary_capacity; /* max particles of the boundary set */
  gg_grid grid;
  int max_sweeps;        /* >= stab_iters + solver_iters */
  int gmax, bmax;        /* neighbour-list widths */
} gg_workspace;

/* ---- library ------------------------------------------------------------ */
const char *gg_version(void);
const char *gg_status_string(int status);
/* Current device is a gfx950 (MI355X).  Returns GG_OK or GG_ERR_NO_DEVICE. */
int gg_device_check(void);
void gg_params_default(gg_params *p);

/* ---- workspace ---------------------------------------------------------- */
size_t gg_workspace_bytes(int capacity, int boundary_capacity, gg_grid grid,
                          int max_sweeps, int gmax, int bmax);
/* Clears the counters the kernels rely on.  Stream-ordered. */
int gg_workspace_init(const gg_workspace *ws, hipStream_t stream);

/* ---- neighbour search --------------------------------------------------- */
/* Hash + stable sort by cell id + cell table + in-place permutation of the
 * set's pos, vel, mass, is_animated and (if ad != NULL) the merge/remove
 * buffers.  Sorted keys go to particle_2_cell.  perm_out (optional, n ints)
 * receives the old index of each sorted slot.
 * Replaces: neighbor_search_granular / neighbor_search_boundary
 * (GranularSystem.cu:65-280), mapParticles2Cells_CUDA + countingInCell_CUDA
 * (CUDAFunctions.cuh:54-107), thrust sort_by_key / gather / exclusive_scan. */
int gg_grid_build(gg_particles *set, gg_adaptive *ad, int *cell_start,
                  int *perm_out, const gg_workspace *ws, hipStream_t stream);

/* Boundary preparation: gg_grid_build on the boundary set, plus the packed
 * copy the contact kernels read.  Call once per boundary change.
 * Replaces: the GranularSystem ctor's neighbor_search_boundary
 * (GranularSystem.cu:55) and the moving-boundary re-search (:402-404). */
int gg_boundary_build(gg_particles *boundary, int *bcell_start,
                      const gg_workspace *ws, hipStream_t stream);

/* Kinematic boundary motion.  Every particle i with is_animated[i] == 1
 * moves by `translation`; then, if angle_deg != 0, it is rotated by
 * angle_deg degrees about the z axis through (center.x, center.y):
 *   u = x - cx, v = y - cy, x' = (u c - v s) + cx, y' = (u s + v c) + cy,
 * with c, s = cosf, sinf of (float)(angle_deg * pi / 180) evaluated once on
 * the host.  Other particles are untouched.  Call gg_boundary_build before
 * the next step (the moving-boundary re-sort).
 * Replaces: animateParticles (main.cu:1175-1208), which evaluates cosf/sinf
 * per thread on the device. */
int gg_boundary_animate(gg_particles *boundary, gg_float3 translation, float angle_deg,
                        gg_float3 center, hipStream_t stream);

/* ---- plain PBD step ----------------------------------------------------- */
/* Grid build + predict + stab/solver contact sweeps + finalize for the
 * granular set, against a boundary prepared by gg_boundary_build.
 * Replaces: GranularSystem::step's neighbor_search_granular and
 * Solver::step (Solver.cu:82-109: add_external_force,
 * update_particle_positions, apply_mass_scaling, project, final_update,
 * adaptive_last_step reset). */
int gg_step(gg_particles *set, gg_adaptive *ad, int *cell_start,
            const gg_particles *boundary, const int *bcell_start,
            const gg_params *params, const gg_workspace *ws,
            hipStream_t stream);

/* No counterpart in the reference.  gg_step for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the workspace is a library-owned
 * static device array of GG_NOALLOC_SCRATCH_BYTES per device, held until the
 * process ends, instead of a caller buffer.  grid is the grid the cell
 * tables are built on; boundary->pos must be in the cell order bcell_start
 * describes (as gg_boundary_build leaves it).  Requests whose workspace
 * (gg_workspace_bytes(n, boundary n, grid, stab + solver sweeps, 48, 16))
 * exceeds the array return GG_ERR_CAPACITY before any device work, e.g.
 * the 32,768-particle box with 14,408 walls needs 15.4 MiB.  Each call
 * clears the counters it relies on and re-packs the boundary; the array
 * keeps the last call's contents and is shared by all calls on its device,
 * so calls that overlap on one device race and must be ordered by the
 * caller.  Results equal gg_step's bit for bit. */
#define GG_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int gg_step_noalloc(gg_particles *set, gg_adaptive *ad, int *cell_start,
                    const gg_particles *boundary, const int *bcell_start, gg_grid grid,
                    const gg_params *params, hipStream_t stream);

/* ---- adaptive merge/split ----------------------------------------------- */
/* Surface flags into set->surface (K5 find_cover_vector + K6 find_surface,
 * GranularSystem.cu:506-637).  Uses the cell table of the start of the step
 * with the current positions, as the reference does. */
int gg_surface(gg_particles *set, const int *cell_start, const gg_params *params,
               const gg_workspace *ws, hipStream_t stream);
/* Squared distance to the nearest surface particle into
 * set->surface_distance (K7 compute_distance_to_surface,
 * GranularSystem.cu:639-699). */
int gg_surface_distance(gg_particles *set, const int *cell_start,
                        const gg_workspace *ws, hipStream_t stream);
/* Merge pairing (K8 merge_mark_gpu, Solver.cu:526-653): candidates from the
 * pre-kernel snapshot of remove[], claims resolved as if threads ran in
 * ascending index order.  Writes remove, merge, merge_velocity. */
int gg_merge_mark(gg_particles *set, gg_adaptive *ad, const int *cell_start,
                  float max_mass, const gg_workspace *ws, hipStream_t stream);
/* Mass/velocity blend over blend_steps steps (K9 merge_count_gpu,
 * Solver.cu:658-699; the reference uses 40). */
int gg_merge_blend(gg_particles *set, gg_adaptive *ad, int blend_steps,
                   const gg_workspace *ws, hipStream_t stream);
/* Split (K10/K11, Solver.cu:708-843) + compaction of pos, vel, mass,
 * adaptive_last_step, remove, merge, merge_count (survivors keep their order)
 * + append of the children in ascending parent order (Solver.cu:1037-1221).
 * At most capacity - survivors children are made.  Synchronises the stream
 * and writes the new count to set->n; stats (optional, 3 ints) = removed,
 * split, candidates refused for capacity. */
int gg_split_compact(gg_particles *set, gg_adaptive *ad, const int *cell_start,
                     const gg_particles *boundary, const int *bcell_start,
                     const gg_params *params, const gg_workspace *ws,
                     hipStream_t stream, int *stats);
/* Solver::adaptive_sampling (Solver.cu:856-1242): merge mark (if do_merge),
 * blend, split + compaction. */
int gg_adaptive_sampling(gg_particles *set, gg_adaptive *ad, const int *cell_start,
                         const gg_particles *boundary, const int *bcell_start,
                         const gg_params *params, float max_mass, int do_merge,
                         int blend_steps, const gg_workspace *ws,
                         hipStream_t stream, int *stats);

/* ---- upsampling ----------------------------------------------------------- */
/* Constants of update_upsampled_cuda (hard-coded in the reference:
 * Solver.cu:29, 1259-1264, 1313-1359); gg_upsample_params_default fills them. */
typedef struct gg_upsample_params {
  float dt, gx, gy, gz;   /* 0.002, (0,-9.8,0) */
  float r9;               /* 1/h^2 = 493.8271605 (h = 0.045) */
  float boundary_radius;  /* 0.015 */
  float repulsion;        /* 3.0 */
  float min_weight;       /* 1e-6 */
  float near_damping;     /* 0.8 */
  float ground_y;         /* 0.1 */
  float ground_damping;   /* 0.5 */
  float max_velocity;     /* 7.0 */
  float floor_y;          /* 0.04 */
  float floor_friction;   /* 0.95 */
} gg_upsample_params;
void gg_upsample_params_default(gg_upsample_params *u);

/* Fine-particle cell sort (neighbor_search_upsampled,
 * GranularSystem.cu:282-395): pos and vel permuted by one stable sort;
 * particle_2_cell keeps the pre-sort keys; fine_cell_start (C+2) built.
 * fine_ws is a workspace whose capacity covers the fine set. */
int gg_upsampled_sort(gg_particles *fine, int *fine_cell_start,
                      const gg_workspace *fine_ws, hipStream_t stream);

/* One pass of update_upsampled_cuda (Solver.cu:1244-1387): fine pos/vel
 * updated in place from the coarse pos/vel, their cell table (start of the
 * step, as in the reference) and the boundary set.  No workspace needed. */
int gg_upsample_plain(gg_particles *fine, const gg_particles *coarse,
                      const int *cell_start, const gg_particles *boundary,
                      const int *bcell_start, gg_grid grid,
                      const gg_upsample_params *up, int flags, hipStream_t stream);

/* Stage entry points of gg_step, in order (for stage-level checking).
 * gg_predict: T7-T9 (Solver.cu:125-191) on the sorted set; builds the
 *   neighbour lists unless GG_FLAG_NO_NLIST.
 * gg_contact_sweep: one compute_delta_pos + change_position pass
 *   (Solver.cu:361-513); sweep in [0, stab_iters+solver_iters).
 * gg_finalize: final_update + adaptive_last_step reset (Solver.cu:193-270,
 *   103-108). */
int gg_predict(gg_particles *set, const int *cell_start,
               const gg_particles *boundary, const int *bcell_start,
               const gg_params *params, const gg_workspace *ws,
               hipStream_t stream);
int gg_contact_sweep(gg_particles *set, const int *cell_start,
                     const gg_particles *boundary, const int *bcell_start,
                     const gg_params *params, int sweep,
                     const gg_workspace *ws, hipStream_t stream);
int gg_finalize(gg_particles *set, const gg_params *params,
                const gg_workspace *ws, hipStream_t stream);

/* The reference Solver's separate per-particle stages (gg_step / gg_predict
 * run them fused into one pass):
 * gg_add_external_force: v += dv for the set's n particles; dv = dt * G
 *   formed by the caller (Solver::add_external_force, Solver.cu:125-131).
 * gg_predict_positions: out[i] = x_i + dt * v_i per component, into a
 *   caller array of n float3 (update_particle_positions, Solver.cu:133-160,
 *   which writes the solver's _pos_t).
 * gg_mass_scaling: scaled_mass[i] = mass[i] * exp(-x_i.y) of the current
 *   positions, exp in double rounded once (apply_mass_scaling,
 *   Solver.cu:162-191).
 * gg_sort_positions_by_key: stable sort of pos alone by particle_2_cell
 *   (the keys themselves unchanged; keys outside [0, C] sort as C); the
 *   cell table of those keys goes to cell_start (C+2)
 *   (update_neighborhood, Solver.cu:112-123). */
int gg_add_external_force(gg_particles *set, gg_float3 dv, hipStream_t stream);
int gg_predict_positions(const gg_particles *set, float dt, gg_float3 *out, hipStream_t stream);
int gg_mass_scaling(gg_particles *set, hipStream_t stream);
int gg_sort_positions_by_key(gg_particles *set, int *cell_start, const gg_workspace *ws,
                             hipStream_t stream);

/* ---- stage profiler ------------------------------------------------------ */
/* When enabled, every stage launched by gg_step and the other entry points
 * is bracketed by hipEvents on its stream (the reference times only the whole
 * step, GranularSystem.cu:398-449).  gg_profiler_collect synchronises the events,
 * adds each stage's elapsed milliseconds to total_ms[stage] and its launch
 * count to count[stage] for stage < nstages, and clears the record list. */
enum {
  GG_STAGE_SORT = 0,     /* hash + counting sort + cell table */
  GG_STAGE_PREDICT = 1,  /* permute + predict */
  GG_STAGE_NLIST = 2,    /* neighbour candidate lists */
  GG_STAGE_SWEEP = 3,    /* one solver contact sweep */
  GG_STAGE_FINALIZE = 4, /* final update */
  GG_STAGE_STAB = 5,     /* one stabilisation contact sweep */
  GG_STAGE_ADAPTIVE = 6, /* surface / distance / merge / split passes */
  GG_STAGE_UPSAMPLE = 7, /* plain upsampler pass */
  GG_STAGE_UPSAMPLE_SORT = 8, /* fine-particle cell sort */
  GG_STAGE_MLP = 9,      /* learned upsampler: fused MFMA network */
  GG_STAGE_MLP_FEATURES = 10, /* learned upsampler: kNN-16 features  */
  GG_STAGE_MLP_FALLBACK = 11, /* learned upsampler: teacher fallback */
  GG_STAGE_SWEEP_QUEUE = 12,  /* one pass over the particles a contact sweep
                                 queued for whole-wave stencil scans */
  GG_STAGE_COUNT = 13
};
int gg_profiler_enable(int max_records); /* 0 disables */
/* Nonzero stops recording new stages (records so far are kept), 0 resumes;
 * lets a caller time a sample of the steps of a longer run. */
int gg_profiler_pause(int paused);
int gg_profiler_collect(double *total_ms, int *count, int nstages);

/* Read-only views into the workspace for checking (device pointers):
 * which = 0 base float4 (x,y,z,r) of the current phase, 1 offset float4
 * (dx,dy,dz,cum) after the last sweep, 2 neighbour meta int4
 * (gcount,bcount,base_cell,cell), 3 granular list [gmax][capacity],
 * 4 boundary list [bmax][capacity], 5 offset pong, 6 base after
 * stabilisation, 7 dt*v after gravity, 8 int counters: [1 + k % 62] =
 * particles of sweep k that scanned the full stencil. */
const void *gg_workspace_view(const gg_workspace *ws, int which);

/* ---- learned upsampler ------------------------------------------------------ */
/* No reference counterpart (a specification addition beside
 * update_upsampled_cuda, Solver.cu:1244-1387).  Definitions:
 * - features of coarse particle i: its 16 nearest coarse particles j != i
 *   in the 27-cell stencil of floor(x_i/L), ordered by (|x_j - x_i|^2, j);
 *   slot s = ((x_j - x_i) fl(1/radius), (v_j - v_i) fl(1/vscale)), zeros
 *   when fewer than 16; 96 bfloat16 values;
 * - flags: bit 0 = a boundary particle closer than guard_boundary; bit 1 =
 *   y < guard_floor or x, z < guard_side.  Flagged particles take the
 *   teacher's result (the fallback);
 * - teacher: the plain upsampler's update applied to the 8 template points
 *   x_i + radius (+-1/2, +-1/2, +-1/2) (slot k = 4 sx + 2 sy + sz) with the
 *   parent's velocity; target y_k = v_fine - v_i;
 * - network 96 -> 128 -> 128 -> 128 -> 24, ReLU; bfloat16 weights and layer
 *   inputs, FP32 accumulation and biases; weights [n_out][k_in] per layer at
 *   offsets 0, 12288, 28672, 45056 (layer 4 padded to 32 rows), biases at
 *   0, 128, 256, 384;
 * - fine position of slot k: (x_i + radius T_k) + dt (v_i + y_k); fine
 *   velocity v_i + y_k (dt from gg_upsample_params). */
#define GG_MLP_WEIGHTS 49152
#define GG_MLP_BIASES 416
typedef struct gg_mlp_config {
  float radius;          /* coarse radius (0.02)                 */
  float vscale;          /* velocity scale, m/s (1.0)            */
  float guard_boundary;  /* 0.05                                  */
  float guard_floor;     /* 0.08                                  */
  float guard_side;      /* 0.05                                  */
  int infer_waves;       /* inference: waves per block, 4, 8 or 16 (0 = 16) */
  int infer_tiles;       /* inference: 16-sample tiles per wave, 1, 2 or 4
                            (0 = 1); 16 waves only with 1 tile             */
  int select_chunk;      /* inference features: candidates sorted and merged
                            per step of the kNN selection, 8 or 16 (0 = 8);
                            the selected set does not depend on it         */
} gg_mlp_config;
typedef struct gg_mlp_weights {
  uint16_t *w_bf16;  /* GG_MLP_WEIGHTS bfloat16 (inference and training)  */
  float *bias;       /* GG_MLP_BIASES                                      */
  float *w_master;   /* FP32 master weights (training)                     */
  float *m_w, *v_w;  /* Adam moments of the weights                        */
  float *b_master;   /* FP32 master biases                                 */
  float *m_b, *v_b;  /* Adam moments of the biases                         */
} gg_mlp_weights;
void gg_mlp_config_default(gg_mlp_config *c);
/* feat: n x 96 bfloat16; flags: n; knn (optional): n x 16 indices, -1 pad. */
int gg_mlp_features(const gg_particles *coarse, const int *cell_start,
                    const gg_particles *boundary, const int *bcell_start, gg_grid grid,
                    const gg_mlp_config *c, uint16_t *feat, int *flags, int *knn,
                    hipStream_t stream);
/* Teacher for every particle (flags NULL) or only the flagged ones; each
 * output (n x 24) optional. */
int gg_mlp_teacher(const gg_particles *coarse, const int *cell_start,
                   const gg_particles *boundary, const int *bcell_start, gg_grid grid,
                   const gg_upsample_params *u, const gg_mlp_config *c, const int *flags,
                   float *y, float *fine_pos, float *fine_vel, hipStream_t stream);
/* Training targets from a fine plain run beside the coarse one (the two
 * seeded identically, stepped side by side; no reference counterpart).  For
 * coarse particle i and slot k: the fine particle j nearest to the template
 * point x_i + radius T_k among the fine particles of the 27 fine cells
 * around floor(point / L_f), ties to the lower index; y (n x 24) = v_j - v_i,
 * d2 (n x 8, optional) its squared distance, match (n x 8, optional) = j; no
 * candidate: y = 0, d2 = +inf, match = -1.  The fine set must be sorted by
 * fine_cell_start (gg_grid_build over fine_grid). */
int gg_mlp_fine_targets(const gg_particles *coarse, const gg_particles *fine,
                        const int *fine_cell_start, gg_grid fine_grid, const gg_mlp_config *c,
                        float *y, float *d2, int *match, hipStream_t stream);
/* Training path (MFMA GEMMs, scratch of gg_mlp_train_scratch_bytes(B)). */
size_t gg_mlp_train_scratch_bytes(int batch);
int gg_mlp_forward(const gg_mlp_weights *w, const uint16_t *x, int batch, float *y,
                   void *scratch, size_t scratch_bytes, hipStream_t stream);
/* Loss gradients of mean((y - t)^2) over batch x 24: dW (GG_MLP_WEIGHTS),
 * db (GG_MLP_BIASES), per-sample squared error sum sq_err (optional). */
int gg_mlp_grad(const gg_mlp_weights *w, const uint16_t *x, const float *target, int batch,
                float *dW, float *db, float *sq_err, void *scratch, size_t scratch_bytes,
                hipStream_t stream);
/* Adam on the FP32 masters (bias-corrected, step >= 1); refreshes the
 * bfloat16 weights and the biases. */
int gg_mlp_adam(gg_mlp_weights *w, const float *dW, const float *db, float lr, float beta1,
                float beta2, float eps, int step, hipStream_t stream);
/* bfloat16 weights and biases from the FP32 masters. */
int gg_mlp_sync_weights(gg_mlp_weights *w, hipStream_t stream);
/* Inference: features, fused register-resident MFMA network, teacher on
 * the flagged particles.  fine_pos / fine_vel: n x 24 (8 fine per coarse,
 * fine_vel optional); scratch: gg_upsample_mlp_scratch_bytes(n);
 * fallback_count (optional, one int of device memory) receives the number of
 * flagged particles, whose outputs the teacher wrote.  The teacher
 * runs on a library-owned second stream (one per device, default priority)
 * forked from and joined back into `stream` by events, so the call stays
 * ordered on `stream`; calls that overlap on one device share that stream. */
size_t gg_upsample_mlp_scratch_bytes(int n);
int gg_upsample_mlp(const gg_particles *coarse, const int *cell_start,
                    const gg_particles *boundary, const int *bcell_start, gg_grid grid,
                    const gg_upsample_params *u, const gg_mlp_config *c, const uint16_t *w_bf16,
                    const float *bias, float *fine_pos, float *fine_vel, void *scratch,
                    size_t scratch_bytes, int *fallback_count, hipStream_t stream);
/* One 16x16x16 bfloat16 MFMA, D = A B (row-major): checks the lane layout. */
int gg_mfma_probe(const uint16_t *a, const uint16_t *b, float *d, hipStream_t stream);
/* Adds to *mismatches (device counter) the number of non-NaN float bit
 * patterns first .. first+count-1 (mod 2^32) whose hardware bfloat16 rounding
 * (used by the inference kernel's hidden layers) differs from the
 * round-to-nearest-even the header defines. */
int gg_bf16_round_check(uint64_t first, uint64_t count, unsigned long long *mismatches,
                        hipStream_t stream);

/* Adds to *mismatches (device counter) the number of inputs on which the
 * contact kernels' short correctly rounded square root, reciprocal and
 * quotient differ from the full IEEE sequences (sqrtf and '/'): mode 0,
 * sqrt of every float bit pattern first .. first+count-1 (mod 2^32) in
 * [2^-96, max float]; mode 1, 1/d of every pattern in [2^-48, 2^64);
 * mode 2, n/d for count pseudo-random pairs derived from the patterns,
 * n in [2^-80, 2^40], d in [2^-48, 2^40]. */
int gg_fast_math_check(int mode, uint64_t first, uint64_t count, unsigned long long *mismatches,
                       hipStream_t stream);

/* ---- scenes sharing one set ------------------------------------------------ */
/* Independent plain scenes stored back to back in one particle set and one
 * boundary set, advanced by one launch per stage (no reference counterpart:
 * the specification's batched configuration).  Scene s owns a contiguous
 * range of the set and of the boundary, its own grid (cx, cy, cz; the cell
 * length is the workspace's), and the cells [cell_base, cell_base + cx cy cz]
 * of the concatenated cell tables: its grid cells, then its own sentinel
 * (cell_base of scene s + 1 = cell_base of s + cx cy cz + 1).  A particle's
 * cell id is its scene's cell_base plus its cell in its scene's grid, so
 * the stable sort keeps every scene contiguous and in the order its own
 * sort would give; every stencil is read through the scene's range of the
 * tables, so each scene's result equals gg_step on that scene alone, bit
 * for bit.  The workspace's grid is (L, T - 1, 1, 1) with T = the total of
 * the scenes' cx cy cz + 1; its capacity covers the whole set. */
typedef struct gg_scene_desc {
  int cx, cy, cz;     /* the scene's grid                                   */
  int cell_base;      /* its first cell in the concatenated cell tables     */
  float mu_s, mu_k;   /* friction of its solver sweeps                      */
} gg_scene_desc;
typedef struct gg_scenes {
  int count;
  const gg_scene_desc *desc;       /* device array [count]                  */
  const gg_scene_desc *host_desc;  /* the same on the host (the calls check
                                      the cell ranges against it)           */
  const int *scene_of;   /* device: scene of each granular particle, in
                            scene order (non-decreasing)                    */
  const int *bscene_of;  /* device: scene of each boundary particle, the
                            same way                                        */
  int friction;               /* nonzero: the solver sweeps apply friction
                                 with each scene's mu_s, mu_k (scenes with
                                 zero coefficients take the same arithmetic) */
} gg_scenes;
/* Sorts the boundary set (scenes back to back) into bcell_start and packs
 * it into the workspace; gg_boundary_build for scenes. */
int gg_scenes_boundary_build(gg_particles *boundary, int *bcell_start, const gg_scenes *scenes,
                             const gg_workspace *ws, hipStream_t stream);
/* One plain step of every scene: gg_step with per-scene grids and friction
 * (params' own mu_s, mu_k are not read). */
int gg_step_scenes(gg_particles *set, int *cell_start, const gg_particles *boundary,
                   const int *bcell_start, const gg_scenes *scenes, const gg_params *params,
                   const gg_workspace *ws, hipStream_t stream);
/* gg_step_scenes captured into a hipGraph: steps_per_launch steps per
 * launch, the list build and sweeps at the given block sizes (0 = 64);
 * arguments are captured by value, device arrays by address. */
typedef struct gg_scenes_graph gg_scenes_graph;
int gg_scenes_graph_create(gg_particles *set, int *cell_start, const gg_particles *boundary,
                           const int *bcell_start, const gg_scenes *scenes,
                           const gg_params *params, const gg_workspace *ws, int steps_per_launch,
                           gg_scenes_graph **out);
int gg_scenes_graph_launch(gg_scenes_graph *g, hipStream_t stream);
/* Times every sweep / list block size pair (64..512 / 128, 256) over
 * `trials` launches, restoring pos, vel and mass after each, and recaptures
 * with the fastest; best_ms (optional) = its time per launch. */
int gg_scenes_graph_autotune(gg_scenes_graph *g, int trials, hipStream_t stream,
                             int *sweep_block, int *nlist_block, float *best_ms);
void gg_scenes_graph_destroy(gg_scenes_graph *g);

/* ---- batched scenes -------------------------------------------------------- */
/* Many independent plain scenes advanced together by one hipGraph launch
 * (no reference counterpart: the specification's batched-scenes
 * configuration).  Each scene brings its own arrays, cell tables, params
 * (per-scene friction) and workspace; N is fixed (no merge/split).  A batch
 * of one scene is the captured fixed-N step.  The batch object (graph,
 * streams, events) is allocated by the library; gg_batch_autotune also
 * allocates a temporary snapshot of the scenes. */
typedef struct gg_batch gg_batch;
typedef struct gg_batch_scene {
  gg_particles set;
  int *cell_start;          /* C+2 */
  gg_particles boundary;    /* built by gg_boundary_build into ws */
  const int *bcell_start;
  gg_params params;
  gg_workspace ws;
} gg_batch_scene;
typedef struct gg_batch_config {
  int branches;          /* parallel graph branches (1..16, <= scenes)     */
  int sweep_block;       /* 0 = each scene's own params value, else 64..512 */
  int nlist_block;       /* same for the neighbour-list build               */
  int steps_per_launch;  /* steps of every scene per graph launch (1..1000) */
} gg_batch_config;
/* cfg NULL: min(4, nscenes) branches, scene block sizes, 1 step per launch. */
int gg_batch_create(const gg_batch_scene *scenes, int nscenes, const gg_batch_config *cfg,
                    gg_batch **out);
/* One graph launch on `stream`: steps_per_launch steps of every scene. */
int gg_batch_step(gg_batch *b, hipStream_t stream);
/* The same steps launched kernel by kernel, scene after scene. */
int gg_batch_step_eager(gg_batch *b, hipStream_t stream);
int gg_batch_get_config(const gg_batch *b, gg_batch_config *cfg);
int gg_batch_set_config(gg_batch *b, const gg_batch_config *cfg);
/* Times `trials` launches of every candidate (branches x sweep block x list
 * block), restores the scenes after each, keeps the fastest.  Scene state
 * is unchanged on return; results never depend on the configuration. */
int gg_batch_autotune(gg_batch *b, int trials, hipStream_t stream, gg_batch_config *best,
                      float *best_ms);
void gg_batch_destroy(gg_batch *b);

/* ---- captured step -------------------------------------------------------- */
/* gg_step of one fixed-N plain scene (no merge/split) captured into a
 * hipGraph: steps_per_launch (1..1000) steps per launch, replayed without
 * per-kernel launch overhead.  The set's arrays and n, the cell tables, the
 * params and the workspace are captured at creation; after changing any of
 * them, destroy and create again.  A one-scene gg_batch underneath. */
typedef struct gg_step_graph gg_step_graph;
int gg_step_graph_create(const gg_particles *set, int *cell_start, const gg_particles *boundary,
                         const int *bcell_start, const gg_params *params,
                         const gg_workspace *ws, int steps_per_launch, gg_step_graph **out);
int gg_step_graph_launch(gg_step_graph *g, hipStream_t stream);
void gg_step_graph_destroy(gg_step_graph *g);

/* ---- device-array primitives (host DArray class) --------------------------- */
/* ---- interactive extras ------------------------------------------------- */
/* Explosion impulse.  Centre of mass c = sum(m p) / sum(m) accumulated in
 * double (fixed-order per-block sums, then the blocks in order; 0 if the
 * total mass is not positive), rounded to float; then for every particle
 * with d = p - c, |d| >= 1e-4:  v += (d / |d|) * (force / (1 + |d|^2)) / m.
 * scratch: gg_explosion_scratch_bytes(n) bytes of device memory.  com_out
 * (host, optional) receives c; asking for it synchronizes the stream.
 * Replaces: Solver::trigger_explosion + apply_explosion_force
 * (Solver.cu:1389-1442), which sums the centre of mass sequentially in
 * float on the host. */
size_t gg_explosion_scratch_bytes(int n);
int gg_explosion(gg_particles *set, float force, void *scratch, size_t scratch_bytes,
                 gg_float3 *com_out, hipStream_t stream);

/* Sphere picking: the indices i with sqrtf(|p_i - center|^2) <= radius,
 * ascending, to indices (device, room for n) and their number to *count
 * (device).  scratch: gg_pick_scratch_bytes(n) bytes.
 * Replaces: markParticlesInSphere (main.cu:1375-1400), whose atomic append
 * gives the indices in no fixed order. */
size_t gg_pick_scratch_bytes(int n);
int gg_pick_sphere(const gg_particles *set, gg_float3 center, float radius, int *indices,
                   int *count, void *scratch, size_t scratch_bytes, hipStream_t stream);
/* Moves the first min(*count, max_count) picked particles (indices in
 * [0, n), device) by delta and clamps each coordinate to [lo, hi] with
 * fmaxf(lo, fminf(p + delta, hi)).  max_count <= n sizes the launch.
 * Replaces: updatePickedParticlesPositions (main.cu:1402-1426), whose bounds
 * are fixed at 0.01 and 0.99. */
int gg_move_picked(gg_particles *set, const int *indices, const int *count, int max_count,
                   gg_float3 delta, gg_float3 lo, gg_float3 hi, hipStream_t stream);

/* Stable compaction (DArray<T>::compact, DArray.hpp:62-122): copies the
 * elements i of src with flags[i] != 1, in order, to dst (no aliasing);
 * elements are elem_words 32-bit words; the kept count is written to the
 * device int *count_out.  scratch: gg_compact_scratch_bytes(n) bytes. */
size_t gg_compact_scratch_bytes(long long n);
int gg_compact(const void *src, void *dst, long long n, int elem_words, const int *flags,
               int *count_out, void *scratch, size_t scratch_bytes, hipStream_t stream);
/* words x 32-bit fill (thrust::fill on the reference's arrays). */
int gg_fill_u32(void *dst, size_t words, uint32_t value, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* GRANULAR_GPU_H */
