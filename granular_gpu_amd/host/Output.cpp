// VTK point files (main.cu:188-372 of the reference) and the binary
// checkpoint of the host API.
#include "granular/Output.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <stdexcept>

#include "granular/GranularParticles.hpp"
#include "granular/GranularSystem.hpp"

namespace granular_output {
namespace {

template <class T>
std::vector<T> download(const T *src, size_t n) {
  std::vector<T> out(n);
  if (n == 0) return out;
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(out.data(), src, sizeof(T) * n, hipMemcpyDeviceToHost) != hipSuccess)
    throw std::runtime_error("device-to-host copy failed");
  return out;
}

bool finite_point(const float3 &p) {
  return !std::isnan(p.x) && !std::isnan(p.y) && !std::isnan(p.z);
}

void write_vtk(std::ofstream &f, const std::string &title, const std::vector<float3> &pts) {
  f << "# vtk DataFile Version 3.0\n";
  f << title << "\n";
  f << "ASCII\n";
  f << "DATASET UNSTRUCTURED_GRID\n";
  f << "POINTS " << pts.size() << " float\n";
  for (const auto &p : pts) f << p.x << " " << p.z << " " << p.y << "\n";
  f << "CELLS " << pts.size() << " " << pts.size() * 2 << "\n";
  for (size_t i = 0; i < pts.size(); i++) f << "1 " << i << "\n";
  f << "CELL_TYPES " << pts.size() << "\n";
  for (size_t i = 0; i < pts.size(); i++) f << "1\n";
}

const char kMagic[4] = {'G', 'G', 'C', 'K'};
constexpr uint32_t kVersion = 1;

template <class T>
void put(std::ofstream &f, const std::vector<T> &v) {
  if (!v.empty()) f.write(reinterpret_cast<const char *>(v.data()), sizeof(T) * v.size());
}

// Reads n elements after checking that the file still holds them (a
// corrupt count must not size the vector beyond the file).
template <class T>
void get(std::ifstream &f, std::vector<T> &v, uint64_t n, uint64_t &left) {
  if (n > left / sizeof(T)) throw std::runtime_error("checkpoint: truncated file");
  v.resize(n);
  if (n && !f.read(reinterpret_cast<char *>(v.data()), sizeof(T) * n))
    throw std::runtime_error("checkpoint: truncated file");
  left -= sizeof(T) * n;
}

// A count field: at most INT32_MAX (counts cross the C-ABI as int).
uint64_t get_count(std::ifstream &f, uint64_t &left) {
  uint64_t n = 0;
  if (left < 8 || !f.read(reinterpret_cast<char *>(&n), 8) || n > 0x7fffffffull)
    throw std::runtime_error("checkpoint: bad count");
  left -= 8;
  return n;
}

}  // namespace

int write_vtk_points(const std::string &path, const std::string &title,
                     const std::vector<float3> &points) {
  std::ofstream f(path);
  if (!f.is_open()) throw std::runtime_error("cannot open " + path);
  std::vector<float3> valid;
  valid.reserve(points.size());
  for (const auto &p : points)
    if (finite_point(p)) valid.push_back(p);
  if (valid.empty()) return 0;
  write_vtk(f, title, valid);
  if (!f) throw std::runtime_error("write failed: " + path);
  return (int)valid.size();
}

int save_upsampled_vtk(const GranularParticles &fine, const std::string &save_file,
                       bool is_adaptive, int frame) {
  const std::string dir = save_file + (is_adaptive ? "_adaptive" : "_normal");
  std::filesystem::create_directories(dir);
  return write_vtk_points(dir + "/" + std::to_string(frame) + ".vtk",
                          "Upsampled Granular Particles",
                          download(fine.get_pos_ptr(), fine.size()));
}

void save_particles_by_mass_vtk(const GranularParticles &particles,
                                const std::string &save_file, int frame, int out[3]) {
  const std::vector<float3> pos = download(particles.get_pos_ptr(), particles.size());
  const std::vector<float> mass = download(particles.get_mass_ptr(), particles.size());
  float3 c = make_float3(0.0f, 0.0f, 0.0f);
  int valid = 0;
  for (const auto &p : pos)
    if (finite_point(p)) {
      c.x += p.x;
      c.y += p.y;
      c.z += p.z;
      valid++;
    }
  if (valid > 0) c = make_float3(c.x / (float)valid, c.y / (float)valid, c.z / (float)valid);
  float min_y = 1000000.f;
  for (const auto &p : pos)
    if (finite_point(p)) min_y = std::fmin(min_y, p.y - c.y);
  const float lift = min_y < 0.0f ? -min_y : 0.0f;
  std::vector<float3> cls[3];
  for (size_t i = 0; i < pos.size(); i++) {
    if (!finite_point(pos[i])) continue;
    const float3 q = make_float3(pos[i].x - c.x, pos[i].y - c.y + lift, pos[i].z - c.z);
    const int k = mass[i] < 2.0f ? 0 : (mass[i] <= 3.0f ? 1 : 2);
    cls[k].push_back(q);
  }
  const char *suffix[3] = {"_mass_light", "_mass_medium", "_mass_heavy"};
  for (int k = 0; k < 3; ++k) {
    const std::string dir = save_file + suffix[k];
    std::filesystem::create_directories(dir);
    out[k] = (int)cls[k].size();
    if (cls[k].empty()) continue;
    std::ofstream f(dir + "/" + std::to_string(frame) + ".vtk");
    if (!f.is_open()) throw std::runtime_error("cannot open " + dir);
    write_vtk(f, "Particles by Mass", cls[k]);
  }
}

int save_animated_boundary_vtk(const GranularParticles &boundary, const GranularParticles *fine,
                               const std::string &save_file, int frame) {
  const std::string dir = save_file + (fine ? "_boundary_adaptive" : "");
  std::filesystem::create_directories(dir);
  const std::string path = dir + "/" + std::to_string(frame) + ".vtk";
  std::ofstream f(path);
  if (!f.is_open()) throw std::runtime_error("cannot open " + path);
  const std::vector<float3> pos = download(boundary.get_pos_ptr(), boundary.size());
  const std::vector<int> marks =
      download(const_cast<GranularParticles &>(boundary).get_is_animated_ptr(), boundary.size());
  std::vector<float3> moving;
  for (size_t i = 0; i < pos.size(); ++i)
    if (marks[i] == 1) moving.push_back(pos[i]);
  if (moving.empty()) return 0;
  if (fine) {
    std::vector<float3> valid;
    for (const auto &p : download(fine->get_pos_ptr(), fine->size()))
      if (finite_point(p)) valid.push_back(p);
    float3 c = make_float3(0.0f, 0.0f, 0.0f);
    for (const auto &p : valid) c = make_float3(c.x + p.x, c.y + p.y, c.z + p.z);
    const float n = (float)valid.size();
    c = make_float3(c.x / n, c.y / n, c.z / n);
    float min_y = 1000000.f;
    for (const auto &p : valid) min_y = std::fmin(min_y, p.y - c.y);
    const float lift = min_y < 0.0f ? -min_y : 0.0f;
    for (auto &p : moving) p = make_float3(p.x - c.x, p.y - c.y + lift, p.z - c.z);
  }
  write_vtk(f, "Animated Boundary Particles", moving);
  if (!f) throw std::runtime_error("write failed: " + path);
  return (int)moving.size();
}

void write_checkpoint(const std::string &path, const Checkpoint &c) {
  const uint64_t n = c.pos.size(), nb = c.boundary_pos.size(), nf = c.fine_pos.size();
  if (c.vel.size() != n || c.mass.size() != n || c.adaptive_last_step.size() != n ||
      c.boundary_animated.size() != nb || c.fine_vel.size() != nf)
    throw std::invalid_argument("checkpoint: array lengths differ");
  std::ofstream f(path, std::ios::binary);
  if (!f.is_open()) throw std::runtime_error("cannot open " + path);
  f.write(kMagic, 4);
  f.write(reinterpret_cast<const char *>(&kVersion), 4);
  f.write(reinterpret_cast<const char *>(&n), 8);
  put(f, c.pos);
  put(f, c.vel);
  put(f, c.mass);
  put(f, c.adaptive_last_step);
  f.write(reinterpret_cast<const char *>(&nb), 8);
  put(f, c.boundary_pos);
  put(f, c.boundary_animated);
  f.write(reinterpret_cast<const char *>(&nf), 8);
  put(f, c.fine_pos);
  put(f, c.fine_vel);
  if (!f) throw std::runtime_error("write failed: " + path);
}

Checkpoint read_checkpoint(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) throw std::runtime_error("cannot open " + path);
  char magic[4];
  uint32_t version = 0;
  if (!f.read(magic, 4) || std::memcmp(magic, kMagic, 4) != 0)
    throw std::runtime_error("checkpoint: bad magic");
  if (!f.read(reinterpret_cast<char *>(&version), 4) || version != kVersion)
    throw std::runtime_error("checkpoint: unsupported version");
  Checkpoint c;
  f.seekg(0, std::ios::end);
  const uint64_t size = (uint64_t)f.tellg();
  f.seekg(8, std::ios::beg);
  uint64_t left = size - 8;
  const uint64_t n = get_count(f, left);
  get(f, c.pos, n, left);
  get(f, c.vel, n, left);
  get(f, c.mass, n, left);
  get(f, c.adaptive_last_step, n, left);
  const uint64_t nb = get_count(f, left);
  get(f, c.boundary_pos, nb, left);
  get(f, c.boundary_animated, nb, left);
  const uint64_t nf = get_count(f, left);
  get(f, c.fine_pos, nf, left);
  get(f, c.fine_vel, nf, left);
  char extra;
  if (f.read(&extra, 1)) throw std::runtime_error("checkpoint: trailing bytes");
  return c;
}

Checkpoint checkpoint_of(const GranularSystem &sys) {
  Checkpoint c;
  const auto p = sys.get_particles();
  const auto b = sys.get_boundaries();
  const auto u = sys.get_upsampled();
  c.pos = download(p->get_pos_ptr(), p->size());
  c.vel = download(p->get_vel_ptr(), p->size());
  c.mass = download(p->get_mass_ptr(), p->size());
  c.adaptive_last_step = download(p->get_adaptive_last_step_ptr(), p->size());
  c.boundary_pos = download(b->get_pos_ptr(), b->size());
  c.boundary_animated = download(const_cast<GranularParticles &>(*b).get_is_animated_ptr(),
                                 b->size());
  c.fine_pos = download(u->get_pos_ptr(), u->size());
  c.fine_vel = download(u->get_vel_ptr(), u->size());
  return c;
}

}  // namespace granular_output
