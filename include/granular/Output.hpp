// Output formats of the host API: the reference's legacy-ASCII VTK point
// files and a binary checkpoint of the particle state.
//
// VTK (main.cu:188-253 saveFunnelUpsampledPositionsToVTK, main.cu:256-372
// saveParticlesByMass): "# vtk DataFile Version 3.0", a title line, ASCII,
// UNSTRUCTURED_GRID; points written as "x z y" (y and z swapped, for a
// z-up viewer) with the stream's default float formatting; one VTK_VERTEX
// cell per point.  Points with a NaN coordinate are dropped.
//
// Checkpoint (no reference counterpart; for replay and test fixtures):
// little-endian "GGCK", format version 1, then three sets (granular,
// boundary, fine), each as a count followed by its arrays.
#pragma once
#include <hip/hip_vector_types.h>

#include <cstdint>
#include <string>
#include <vector>

class GranularParticles;
class GranularSystem;

namespace granular_output {

// Writes the points (NaN ones dropped) as a VTK file; returns the number of
// points written.  As in the reference, a file whose points are all NaN is
// created empty.  Throws std::runtime_error if the file cannot be opened.
int write_vtk_points(const std::string &path, const std::string &title,
                     const std::vector<float3> &points);

// saveFunnelUpsampledPositionsToVTK: <save_file>_adaptive or _normal /
// <frame>.vtk, title "Upsampled Granular Particles".  Returns the count.
int save_upsampled_vtk(const GranularParticles &fine, const std::string &save_file,
                       bool is_adaptive, int frame);

// saveParticlesByMass: points centred on their centroid, lifted so that the
// lowest is at y = 0 if it was below, split by mass (< 2 light, [2, 3]
// medium, > 3 heavy) into <save_file>_mass_light / _medium / _heavy /
// <frame>.vtk (title "Particles by Mass"); an empty class writes no file.
// Returns the three counts in out[0..2].
void save_particles_by_mass_vtk(const GranularParticles &particles,
                                const std::string &save_file, int frame, int out[3]);

// The animated (is_animated == 1) boundary particles, title "Animated
// Boundary Particles", NaN points kept.  fine == nullptr:
// saveFunnelBoundaryParticlesToVTK (main.cu:370-432), <save_file>/<frame>.vtk
// with the positions as they are.  Otherwise saveBoundaryParticlesToVTK
// (main.cu:434-525), <save_file>_boundary_adaptive/<frame>.vtk with the
// positions centred on the centroid of the finite fine particles and lifted
// by the same amount as the fine set's lowest point.  With no animated
// particle the file is created empty.  Returns the count written.
int save_animated_boundary_vtk(const GranularParticles &boundary, const GranularParticles *fine,
                               const std::string &save_file, int frame);

struct Checkpoint {
  std::vector<float3> pos, vel;  // granular set
  std::vector<float> mass;
  std::vector<int> adaptive_last_step;
  std::vector<float3> boundary_pos;
  std::vector<int> boundary_animated;
  std::vector<float3> fine_pos, fine_vel;
};

// Writes / reads the checkpoint file.  Reading checks the magic, version,
// counts (at most INT32_MAX, and never more than the bytes left in the
// file) and length; throws std::runtime_error on any mismatch.  The file
// holds the state a plain step starts from; the merge/split bookkeeping of
// an adaptive run (remove flags, merge masses and velocities, blend counts)
// is not in it, so an adaptive run resumes equivalently only from a step
// where no merge blend is in flight.
void write_checkpoint(const std::string &path, const Checkpoint &c);
Checkpoint read_checkpoint(const std::string &path);

// The current state of a system (device arrays copied to the host).
Checkpoint checkpoint_of(const GranularSystem &sys);

}  // namespace granular_output
