// Device self-test of the host containers (DArray, GranularParticles),
// exported for the test suite through ggh_darray_selftest.
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "granular/GranularParticles.hpp"
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
  // append grows to max(new length, 2 x capacity)
  DArray<float> x(5), y(7);
  upload(x, std::vector<float>{0, 1, 2, 3, 4});
  upload(y, std::vector<float>{5, 6, 7, 8, 9, 10, 11});
  x.append(y);
  expect(x.length() == 12 && x.capacity() == 12, "append: length/capacity");
  auto xv = host(x);
  for (int i = 0; i < 12; ++i) expect(xv[(size_t)i] == (float)i, "append: values");
  x.append(y);
  expect(x.length() == 19 && x.capacity() == 24, "append: doubling");
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

}  // namespace

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
