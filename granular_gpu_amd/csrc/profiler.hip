// Stage profiler: hipEvents bracketing each stage on the stream it runs on.
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "gg_internal.hpp"
#include "profiler.hpp"

namespace gg {
namespace {
constexpr int kMaxStages = 64;
struct Prof {
  bool on = false;
  bool paused = false;
  std::vector<hipEvent_t> ev;
  std::vector<std::pair<int, int>> rec;  // (stage, first event index)
  int next = 0;
  int open[kMaxStages];  // first event index of each stage's open bracket, or -1
  Prof() {
    for (int &o : open) o = -1;
  }
};
Prof &prof() {
  static Prof p;
  return p;
}
}  // namespace

// Stages enqueued while a stream is being captured into a graph are not
// timed (an event record would become a graph node).
static bool capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// Each stage has its own open bracket, so brackets of different stages may
// interleave (e.g. one on a side stream inside another on the main stream).
void prof_begin(int stage, hipStream_t st) {
  Prof &p = prof();
  if (stage < 0 || stage >= kMaxStages) return;
  p.open[stage] = -1;
  if (!p.on || p.paused || p.next + 2 > (int)p.ev.size() || capturing(st)) return;
  if (hipEventRecord(p.ev[p.next], st) != hipSuccess) return;
  p.open[stage] = p.next;
  p.next += 2;
}

void prof_end(int stage, hipStream_t st) {
  Prof &p = prof();
  if (stage < 0 || stage >= kMaxStages) return;
  const int first = p.open[stage];
  p.open[stage] = -1;
  if (!p.on || first < 0) return;
  if (hipEventRecord(p.ev[first + 1], st) != hipSuccess) return;
  p.rec.emplace_back(stage, first);
}
}  // namespace gg

using namespace gg;

extern "C" int gg_profiler_enable(int max_records) {
  Prof &p = prof();
  for (hipEvent_t e : p.ev) (void)hipEventDestroy(e);
  p.ev.clear();
  p.rec.clear();
  p.next = 0;
  p.on = false;
  p.paused = false;
  for (int &o : p.open) o = -1;
  if (max_records <= 0) return GG_OK;
  p.ev.resize(2 * (size_t)max_records);
  for (auto &e : p.ev)
    if (hipEventCreate(&e) != hipSuccess) return GG_ERR_HIP;
  p.on = true;
  return GG_OK;
}

extern "C" int gg_profiler_pause(int paused) {
  prof().paused = paused != 0;
  return GG_OK;
}

extern "C" int gg_profiler_collect(double *total_ms, int *count, int nstages) {
  Prof &p = prof();
  if (!total_ms || !count || nstages < 0) return GG_ERR_INVALID;
  for (auto &r : p.rec) {
    if (hipEventSynchronize(p.ev[r.second + 1]) != hipSuccess) return GG_ERR_HIP;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.ev[r.second], p.ev[r.second + 1]) != hipSuccess)
      return GG_ERR_HIP;
    if (r.first >= 0 && r.first < nstages) {
      total_ms[r.first] += ms;
      count[r.first] += 1;
    }
  }
  p.rec.clear();
  p.next = 0;
  return GG_OK;
}
