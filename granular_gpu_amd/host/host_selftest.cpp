// Device self-tests of the host API: the containers (DArray,
// GranularParticles) through ggh_darray_selftest and the Solver's separate
// stages through ggh_solver_stages_selftest.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "granular/GranularParticles.hpp"
#include "granular/Solver.hpp"
#include "granular_host.h"

namespace {

template <typename T>
std::vector<T> host(const DArray<T> &a) {
  std::vector<T> v(a.length());
  if (!v.empty() &&
      hipMemcpy(v.data(), a.addr(), sizeof(T) * v.size(), hipMemcpyDeviceToHost) != hipSuccess)
    throw std::runtime_error("copy to host failed");
  return v;
}

template <typename T>
void upload(DArray<T> &a, const std::vector<T> &v) {
  if (v.size() > a.length()) throw std::runtime_error("upload too large");
  if (!v.empty() &&
      hipMemcpy(a.addr(), v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
    throw std::runtime_error("copy to device failed");
}

void expect(bool ok, const char *what) {
  if (!ok) throw std::runtime_error(what);
}

// flag pattern with 0, 1 and -1 (only 1 removes), deterministic
int flag_of(unsigned int i) {
  const unsigned int h = (i * 2654435761u) >> 7;
  return (h % 5 == 0) ? 1 : ((h % 7 == 0) ? -1 : 0);
}

void run(int n) {
  const unsigned int N = (unsigned int)n;
  // compact keeps flag != 1 in order
  DArray<float3> a(N);
  std::vector<float3> av(N);
  for (unsigned int i = 0; i < N; ++i) av[i] = make_float3((float)i, 2.0f * i, 3.0f * i);
  upload(a, av);
  DArray<int> f(N);
  std::vector<int> fv(N);
  for (unsigned int i = 0; i < N; ++i) fv[i] = flag_of(i);
  upload(f, fv);
  std::vector<float3> want;
  for (unsigned int i = 0; i < N; ++i)
    if (fv[i] != 1) want.push_back(av[i]);
  expect(a.compact(f) == want.size(), "compact: kept count");
  const auto got = host(a);
  expect(got.size() == want.size() &&
             (want.empty() || std::memcmp(got.data(), want.data(), 12 * want.size()) == 0),
         "compact: kept elements or order");
  // length mismatch throws
  bool threw = false;
  try {
    DArray<int> g(N + 1);
    a.compact(g);
  } catch (const std::runtime_error &) {
    threw = true;
  }
  expect(threw, "compact: flags length mismatch must throw");
  // a second compaction by the same flags (the array now lives in the
  // other of its two blocks) removes nothing more from the survivors
  {
    DArray<int> f2((unsigned int)want.size());
    expect(a.compact(f2) == want.size(), "compact twice: kept count");
    auto again = host(a);
    expect(std::memcmp(again.data(), want.data(), want.size() * sizeof(float3)) == 0,
           "compact twice: contents");
  }
  // append grows to max(new length, capacity + capacity / 2 + 64)
  DArray<float> x(5), y(7);
  upload(x, std::vector<float>{0, 1, 2, 3, 4});
  upload(y, std::vector<float>{5, 6, 7, 8, 9, 10, 11});
  x.append(y);
  expect(x.length() == 12 && x.capacity() == 71, "append: length/capacity");
  auto xv = host(x);
  for (int i = 0; i < 12; ++i) expect(xv[(size_t)i] == (float)i, "append: values");
  x.append(y);
  expect(x.length() == 19 && x.capacity() == 71, "append: fits the capacity");
  // remove(start, count)
  x.remove(2, 3);
  xv = host(x);
  expect(x.length() == 16 && xv[1] == 1.0f && xv[2] == 5.0f && xv[8] == 11.0f,
         "remove: closes the gap");
  threw = false;
  try {
    x.remove(10, 10);
  } catch (const std::runtime_error &) {
    threw = true;
  }
  expect(threw, "remove: out of range must throw");
  // resize keeps min(length, capacity)
  x.resize(4);
  xv = host(x);
  expect(x.length() == 4 && x.capacity() == 4 && xv[2] == 5.0f && xv[3] == 6.0f, "resize: shrink");
  // verifyArrayOrdering over two arrays
  DArray<float> p(N), q(N);
  std::vector<float> pv(N), qv(N);
  for (unsigned int i = 0; i < N; ++i) {
    pv[i] = (float)i;
    qv[i] = -(float)i;
  }
  upload(p, pv);
  upload(q, qv);
  const char *names[] = {"p", "q"};
  DArray<float>::verifyArrayOrdering({&p, &q}, f, names);
  expect(p.length() == want.size() && q.length() == want.size(), "verifyArrayOrdering: lengths");
  // GranularParticles: remove_elements compacts mass, adaptive_last_step,
  // pos, vel; add_elements appends them with adaptive_last_step = 1
  std::vector<float3> gp(N);
  for (unsigned int i = 0; i < N; ++i) gp[i] = make_float3((float)i, 0.5f, 0.25f);
  GranularParticles g(gp);
  std::vector<float> gm(N);
  for (unsigned int i = 0; i < N; ++i) gm[i] = 1.0f + (float)(i % 3);
  if (N) HIP_CALL(hipMemcpy(g.get_mass_ptr(), gm.data(), 4 * (size_t)N, hipMemcpyHostToDevice));
  g.remove_elements(f);
  expect(g.size() == want.size(), "remove_elements: size");
  std::vector<float> gmm(g.size());
  if (!gmm.empty())
    HIP_CALL(hipMemcpy(gmm.data(), g.get_mass_ptr(), 4 * gmm.size(), hipMemcpyDeviceToHost));
  for (size_t k = 0, i = 0; i < N; ++i)
    if (fv[i] != 1) expect(gmm[k++] == gm[i], "remove_elements: mass order");
  DArray<float> nm(3);
  DArray<float3> np(3), nv(3);
  upload(nm, std::vector<float>{7, 8, 9});
  upload(np, std::vector<float3>{make_float3(1, 1, 1), make_float3(2, 2, 2), make_float3(3, 3, 3)});
  const unsigned int before = g.size();
  g.add_elements(nm, np, nv, 3);
  expect(g.size() == before + 3, "add_elements: size");
  std::vector<int> adl(g.size());
  HIP_CALL(hipMemcpy(adl.data(), g.get_adaptive_last_step_ptr(), 4 * adl.size(),
                     hipMemcpyDeviceToHost));
  for (unsigned int i = 0; i < g.size(); ++i)
    expect(adl[i] == (i >= before ? 1 : 0), "add_elements: adaptive_last_step");
  std::vector<float3> pp(g.size());
  HIP_CALL(hipMemcpy(pp.data(), g.get_pos_ptr(), 12 * pp.size(), hipMemcpyDeviceToHost));
  expect(pp[before + 2].x == 3.0f, "add_elements: positions appended at the end");
}

// The Solver's separate stages (add_external_force,
// update_particle_positions, apply_mass_scaling, update_neighborhood,
// trigger_explosion) against values computed here on the host.
void run_solver_stages(int n) {
  std::vector<float3> p0((size_t)n), v0((size_t)n);
  std::vector<float> m0((size_t)n);
  for (int i = 0; i < n; ++i) {
    p0[(size_t)i] = make_float3(0.013f * (float)(i % 17), 0.5f + 0.007f * (float)(i % 29),
                                0.011f * (float)(i % 23));
    v0[(size_t)i] = make_float3(0.1f * (float)(i % 5) - 0.2f, 0.03f * (float)(i % 7),
                                -0.05f * (float)(i % 3));
    m0[(size_t)i] = 0.5f + 0.25f * (float)(i % 4);
  }
  auto parts = std::make_shared<GranularParticles>(p0);
  upload(parts->_vel, v0);
  DArray<float> &mass = const_cast<DArray<float> &>(parts->get_mass());
  upload(mass, m0);
  auto bnd = std::make_shared<GranularParticles>(std::vector<float3>{make_float3(0.1f, 0.1f, 0.1f)});
  Solver solver(parts);
  const float dt = 0.002f;
  const float3 G = make_float3(0.25f, -9.8f, 0.0f);
  solver.add_external_force(parts, dt, G);
  const float dvx = dt * G.x, dvy = dt * G.y, dvz = dt * G.z;
  std::vector<float3> v1 = host(parts->_vel);
  for (int i = 0; i < n; ++i) {
    const float3 a = v0[(size_t)i], b = v1[(size_t)i];
    expect(b.x == a.x + dvx && b.y == a.y + dvy && b.z == a.z + dvz,
           "add_external_force: v + dt G");
  }
  solver.update_particle_positions(parts, dt);
  std::vector<float3> pt((size_t)n);
  if (n > 0)
    HIP_CALL(hipMemcpy(pt.data(), solver.get_pos_t_ptr(), 12 * (size_t)n, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i) {
    const float3 x = p0[(size_t)i], v = v1[(size_t)i], q = pt[(size_t)i];
    expect(q.x == x.x + dt * v.x && q.y == x.y + dt * v.y && q.z == x.z + dt * v.z,
           "update_particle_positions: x + dt v");
  }
  expect(host(parts->_pos).size() == (size_t)n, "update_particle_positions: positions kept");
  solver.apply_mass_scaling(parts);
  std::vector<float> ms((size_t)n);
  if (n > 0)
    HIP_CALL(hipMemcpy(ms.data(), parts->get_scaled_mass_ptr(), 4 * (size_t)n,
                       hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i)
    expect(ms[(size_t)i] == m0[(size_t)i] * (float)std::exp(-(double)p0[(size_t)i].y),
           "apply_mass_scaling: m exp(-y)");
  // update_neighborhood: stable sort of pos alone by particle_2_cell on a
  // 4^3 grid (keys outside [0, 64] sort as the sentinel 64)
  DArray<int> bcs(4 * 4 * 4 + 2);
  solver.sort_boundary(bnd, bcs, make_int3(4, 4, 4), 0.25f);
  std::vector<int> keys((size_t)n);
  for (int i = 0; i < n; ++i) keys[(size_t)i] = (i % 11 == 3) ? -5 : (i % 13 == 7 ? 99 : (i * 37) % 65);
  if (n > 0)
    HIP_CALL(hipMemcpy(parts->get_particle_2_cell(), keys.data(), 4 * (size_t)n,
                       hipMemcpyHostToDevice));
  solver.update_neighborhood(parts);
  std::vector<int> order((size_t)n);
  for (int i = 0; i < n; ++i) order[(size_t)i] = i;
  auto eff = [&](int i) { const int k = keys[(size_t)i]; return (k < 0 || k > 64) ? 64 : k; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return eff(a) < eff(b); });
  const std::vector<float3> ps = host(parts->_pos);
  for (int k = 0; k < n; ++k)
    expect(std::memcmp(&ps[(size_t)k], &p0[(size_t)order[(size_t)k]], 12) == 0,
           "update_neighborhood: positions in stable key order");
  std::vector<int> keys_after((size_t)n);
  if (n > 0)
    HIP_CALL(hipMemcpy(keys_after.data(), parts->get_particle_2_cell(), 4 * (size_t)n,
                       hipMemcpyDeviceToHost));
  expect(keys_after == keys, "update_neighborhood: keys unchanged");
  // trigger_explosion keeps the centre of mass it pushed from
  solver.trigger_explosion(parts, 1.0f);
  // (update_neighborhood moved the positions only: the centre pairs the
  // device arrays as they are)
  std::vector<float> mnow = host(mass);
  double cx = 0.0, cm = 0.0;
  for (int i = 0; i < n; ++i) {
    cx += (double)mnow[(size_t)i] * (double)ps[(size_t)i].x;
    cm += (double)mnow[(size_t)i];
  }
  if (n > 0)
    expect(std::fabs(solver.get_explosion_center().x - (float)(cx / cm)) <= 1e-6f,
           "trigger_explosion: centre of mass");
}

}  // namespace

extern "C" int ggh_solver_stages_selftest(int n, char *msg, int msg_len) {
  try {
    if (n < 0) throw std::invalid_argument("n must be >= 0");
    run_solver_stages(n);
    if (msg && msg_len > 0) std::snprintf(msg, (size_t)msg_len, "ok");
    return 0;
  } catch (const std::exception &e) {
    if (msg && msg_len > 0) std::snprintf(msg, (size_t)msg_len, "%s", e.what());
    return -1;
  }
}

extern "C" int ggh_darray_selftest(int n, char *msg, int msg_len) {
  try {
    if (n < 0) throw std::invalid_argument("n must be >= 0");
    run(n);
    if (msg && msg_len > 0) std::snprintf(msg, (size_t)msg_len, "ok");
    return 0;
  } catch (const std::exception &e) {
    if (msg && msg_len > 0) std::snprintf(msg, (size_t)msg_len, "%s", e.what());
    return -1;
  }
}
