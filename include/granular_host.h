/* C-ABI of the host C++ API (GranularSystem / Solver / GranularParticles /
 * DArray / SceneConfig), for callers that cannot use C++ classes (ctypes,
 * cgo, JNI).  Handles own one GranularSystem each.  Functions return 0 (or a
 * count) on success and a negative value on error; ggh_last_error() gives
 * the message of the last failure on the calling thread.
 *
 * Reference interface replaced: the GranularSystem class and its accessors
 * (GranularSystem.hpp:8-56), Solver::save_particle_stats (Solver.cu:1444),
 * loadSceneConfig (main.cu:123-167), the scene builders of main.cu
 * (747-816, 822-1167) and its boundary animation (1175-1277). */
#ifndef GRANULAR_HOST_H
#define GRANULAR_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ggh_system ggh_system;

typedef struct ggh_scene {
  float space_size[3];
  float dt;
  float gravity[3];
  float sph_spacing;
  float init_spacing;
  float smoothing_radius;
  float cell_length;
  int cell_size[3];
  float density;
  float box_size[3];
  float boundary_translation[3];
  float particle_translation[3];
  int scene;
  int is_adaptive;
  char obj_file[256];
  char save_file[256];
} ggh_scene;

const char *ggh_last_error(void);

/* Scene JSON (CPU only). */
int ggh_load_scene(const char *path, ggh_scene *out);
int ggh_parse_scene(const char *json_text, ggh_scene *out);
/* Scene builders (CPU only): write up to max xyz triples to out (out may be
 * NULL to count) and return the total count. */
int ggh_scene_lattice(const ggh_scene *s, float *out, int max);
int ggh_scene_walls(const ggh_scene *s, float *out, int max);
int ggh_upsampled_cloud(const float *coarse, int n, float radius, int per, float *out, int max);

/* Kinematic animation state.  Fields and defaults of the reference's
 * AnimationState (main.cu:100-111); translation_time (2.5 s) and
 * rotation_limit (45 degrees) are constants of its updateAnimation
 * (main.cu:1219, 1255-1262). */
typedef struct ggh_animation {
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
  int center_from_marked; /* 1: centre of the marked particles, not the range */
} ggh_animation;
void ggh_animation_default(ggh_animation *a);

/* Boundary set of a scene (CPU only; main.cu:822-1167): the walls, then for
 * EXCAVATOR / FUNNEL the xyz point cloud of obj_file when that file exists,
 * transformed as the reference does (bucket + (2.5, 0, 0.5); funnel rotated
 * 90 degrees about x through its centroid, + (1.25, 0.3, 1.5)).  Writes up
 * to max xyz triples to out and animated markers (0 walls, 1 cloud) to
 * animated (either may be NULL), fills *anim if not NULL, and returns the
 * total count. */
int ggh_scene_boundary(const ggh_scene *s, float *out, int *animated, int max,
                       ggh_animation *anim);

/* System: host arrays are copied to the device; the constructor runs one
 * step (as the reference's does). */
ggh_system *ggh_system_create(const float *pos, int n, const float *boundary, int nb,
                              const int *boundary_animated, const float *fine, int nf,
                              const float space_size[3], float cell_length, float dt,
                              const float gravity[3], const int cell_size[3], float density,
                              float upsampled_radius, int is_move_boundary, int is_adaptive,
                              int verbose);
void ggh_system_destroy(ggh_system *h);
/* Runs nsteps steps; returns the last step()'s value (1). */
int ggh_system_step(ggh_system *h, int nsteps);
/* which: 0 granular, 1 boundary, 2 fine, 3 total. */
int ggh_system_size(const ggh_system *h, int which);
/* Copies count elements of a field of set `which` (0 granular, 1 boundary,
 * 2 fine) to host memory.  field: 0 pos (3 floats), 1 vel (3 floats),
 * 2 mass, 3 particle_2_cell (int), 4 surface (int), 5 surface_distance,
 * 6 adaptive_last_step (int), 7 is_animated (int). */
int ggh_system_read(const ggh_system *h, int which, int field, void *out, int count);
/* Moves the boundary particles flagged as animated by (dx,dy,dz) (a
 * kinematic boundary; the next step re-sorts a moving boundary); returns
 * the number of animated particles. */
int ggh_system_translate_animated(ggh_system *h, float dx, float dy, float dz);
/* animateParticles (main.cu:1175-1208) on the boundary set: translation,
 * then a rotation by angle_deg about the z axis through center (x, y);
 * see gg_boundary_animate. */
int ggh_system_animate(ggh_system *h, const float translation[3], float angle_deg,
                       const float center[3]);
/* One updateAnimation(dt) call (main.cu:1210-1277) advancing *a; returns 1
 * if the boundary moved, 0 if not, negative on error. */
int ggh_system_update_animation(ggh_system *h, ggh_animation *a, float dt);
/* Interactive extras on the granular set.  Explosion: Solver::
 * trigger_explosion (Solver.cu:1411), the 'e' key of main.cu:1905 uses
 * force 50; com_out (optional) receives the centre of mass.  Picking
 * (main.cu:1375-1426): ggh_system_pick writes up to max indices within
 * radius of center, ascending, and returns their number;
 * ggh_system_move_picked moves those particles by delta, clamped to
 * [lo, hi] per coordinate (the reference clamps to [0.01, 0.99]). */
int ggh_system_explode(ggh_system *h, float force, float com_out[3]);
int ggh_system_pick(const ggh_system *h, const float center[3], float radius, int *indices,
                    int max);
int ggh_system_move_picked(ggh_system *h, const int *indices, int count, const float delta[3],
                           const float lo[3], const float hi[3]);
int ggh_system_frame_times(const ggh_system *h, float *out, int max);
/* Rows (merged, split, N before) of the adaptive passes; returns the count. */
int ggh_system_particle_stats(const ggh_system *h, int *out, int max_rows);
/* Solver::sweep_scan_counts: particles that scanned their whole stencil in
 * each sweep of the last step (up to max); returns the sweep count, -1 on
 * error.  No counterpart in the reference (a diagnostic of the lists). */
int ggh_system_sweep_scans(const ggh_system *h, int *out, int max);
int ggh_system_save_stats(ggh_system *h, const char *path);

/* Output formats (include/granular/Output.hpp).  VTK: the reference's
 * legacy-ASCII point files (main.cu:188-372), NaN points dropped, y and z
 * swapped.  ggh_write_vtk (CPU only) writes n xyz points and returns the
 * count written; the system versions write the fine set to
 * <save_file>_adaptive|_normal/<frame>.vtk (returns the count) or the
 * granular set by mass class to <save_file>_mass_light|_medium|_heavy/
 * <frame>.vtk (counts[3]). */
int ggh_write_vtk(const char *path, const char *title, const float *xyz, int n);
int ggh_system_save_upsampled_vtk(const ggh_system *h, const char *save_file, int is_adaptive,
                                  int frame);
int ggh_system_save_mass_vtk(const ggh_system *h, const char *save_file, int frame,
                             int counts[3]);
/* The animated boundary particles: centred == 0 writes them as they are to
 * <save_file>/<frame>.vtk (saveFunnelBoundaryParticlesToVTK, main.cu:370);
 * centred != 0 centres and lifts them by the fine set's centroid and lowest
 * point into <save_file>_boundary_adaptive/<frame>.vtk
 * (saveBoundaryParticlesToVTK, main.cu:434).  Returns the count. */
int ggh_system_save_boundary_vtk(const ggh_system *h, const char *save_file, int frame,
                                 int centred);
/* Binary checkpoint of the particle state (no reference counterpart): the
 * granular set's pos, vel, mass, adaptive_last_step, the boundary's pos and
 * is_animated, the fine set's pos and vel.  ggh_checkpoint_write (CPU only)
 * writes host arrays; ggh_checkpoint_read copies up to max elements of one
 * array (which: 0 granular, 1 boundary, 2 fine; field numbers as
 * ggh_system_read: 0 pos, 1 vel, 2 mass, 6 adaptive_last_step, 7
 * is_animated) and returns its length. */
int ggh_checkpoint_write(const char *path, const float *pos, const float *vel,
                         const float *mass, const int *adaptive_last_step, int n,
                         const float *boundary, const int *boundary_animated, int nb,
                         const float *fine_pos, const float *fine_vel, int nf);
int ggh_system_save_checkpoint(const ggh_system *h, const char *path);
int ggh_checkpoint_read(const char *path, int which, int field, void *out, int max);
int ggh_system_set_solver(ggh_system *h, int stab, int solver, float mu_s, float mu_k,
                          int flags);
/* Solver::set_neighbor_lists: granular and boundary candidates per
 * neighbour list and the list skin (m); defaults 48, 16, 0.0075. */
int ggh_system_set_lists(ggh_system *h, int gmax, int bmax, float skin);
/* GranularSystem::set_merge_rules: merge/split rules of the adaptive passes
 * (0 reference rules, 1 depth rules; the fields of gg_params) and the
 * largest mass a merge may start from. */
int ggh_system_set_merge_rules(ggh_system *h, int depth_rules, float min_mass, float split_mass,
                               float merge_dist, float merge_depth, float split_depth,
                               float max_mass);
/* GranularSystem::set_learned_upsampler: from the next step on the fine set
 * holds 8 particles per granular particle written by gg_upsample_mlp.
 * w_bf16: GG_MLP_WEIGHTS bfloat16 bit patterns, bias: GG_MLP_BIASES floats
 * (host arrays); the rest are gg_mlp_config's fields. */
int ggh_system_set_learned(ggh_system *h, const unsigned short *w_bf16, const float *bias,
                           float radius, float vscale, float guard_boundary, float guard_floor,
                           float guard_side);

/* Device self-test of DArray (compact, append, remove, resize,
 * verifyArrayOrdering) and GranularParticles (remove_elements,
 * add_elements) on n elements; 0 when every check passes. */
int ggh_darray_selftest(int n, char *msg, int msg_len);
/* Device self-test of the Solver's separate stages (add_external_force,
 * update_particle_positions, apply_mass_scaling, update_neighborhood,
 * trigger_explosion; Solver.cu:112-191, 1411-1442) on n particles against
 * host-computed values; 0 and "ok" in msg on success. */
int ggh_solver_stages_selftest(int n, char *msg, int msg_len);

#ifdef __cplusplus
}
#endif
#endif
