// Internal helpers of the granular_gpu library: workspace layout, launch
// checks, shared constants.  Host and device code include this file.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "granular_gpu.h"

namespace gg {

constexpr int kBlock = 256;        // threads per block for particle kernels
constexpr int kScanTile = 2048;    // elements per block in the cell-table scan
constexpr int kScanMaxParts = 8192;
constexpr int kSmallCell = 32;     // cells up to this size are ranked one thread per entry
constexpr int kBigLds = 4096;      // cells up to this size are sorted in LDS
constexpr int kHugeChunks = 256;   // blocks compacting each larger cell
constexpr int kHugeCounter = 63;   // ctr[63]: cells larger than kBigLds in this build
constexpr double kPiD = 3.14159265358979323846;
constexpr float kPiF = 3.14159265358979323846f;
// A particle's list stays exact while its own accumulated correction and
// that of every particle of its stencil stay below half of this fraction of
// the skin.
constexpr float kListValidFrac = 0.9f;
// A wave with at most this many particles that must scan the full stencil
// queues them for one wave each (else each scans in its own lane), in its
// block's LDS queue, scanned by the block's waves after their list work.
// GG_FLAG_FALLBACK_WAVE instead queues them sweep-wide for a second pass
// after the sweep (at most kFallbackWaveMax per wave).  Measured on the
// 1M-particle pile, the block queue is 6% faster at steps 110..115 and 2%
// at 480..485 (no second launch per sweep), 3% slower at 300..305.
constexpr int kFallbackWaveMax = 32;
constexpr int kFallbackWaveMaxLocal = 6;
constexpr int kFallbackSlots = 62;  // ctr[1 + k % 62] = full-stencil count of sweep k
// GG_FLAG_LIST_REFRESH: a list build costs about as much as kRefreshCost
// stencil scans per particle rebuilt, measured: a scan by a lane whose wave
// holds lists is divergent and waits on its dependent loads (about 1.3 ns per
// scan against 0.15 ns per particle of a build at 4M particles)
constexpr int kRefreshCost = 8;
// ctr[63..] are not used by the sweeps; the adaptive passes use ctr2 (own region)
constexpr int kMergeRounds = 12;    // parallel claim-resolution rounds before the finisher

// Byte offsets of every buffer inside the caller's workspace.  Every region
// starts on a 256-byte boundary.
struct Layout {
  size_t count;    // int[C+1]       per-cell histogram, zero between builds
  size_t key;      // int[cmax]      cell id per particle (input order)
  size_t slot;     // int[cmax]      arrival slot inside the cell
  size_t perm;     // int[cmax]      sorted slot -> old index
  size_t itmp;     // int[cmax]      cell order as scattered (before the stable ranking)
  size_t part;     // int[parts+1]   scan partials
  size_t big;      // int[C+1]       list of cells needing the big sort
  size_t ctr;      // int[64]        device counters ([0] = big-cell count)
  size_t ctr2;     // int[64]        adaptive counters
  size_t dirty;    // int[C+1]       cells near a particle past half the skin
  size_t bounds;   // uint[4]        largest |D| component and contact radius of the step (float bits)
  size_t bpart;    // uint[2*blocks] per-block maxima reduced into bounds
  size_t x0;       // float4[cap]    start-of-step position + radius
  size_t off0;     // float4[cap]    offset ping
  size_t off1;     // float4[cap]    offset pong
  size_t d0;       // float4[cap]    dt*v after gravity
  size_t xb;       // float4[cap]    position after stabilisation + radius
  size_t meta;     // int4[cap]      neighbour list meta
  size_t glist;    // int[gmax*cap]  granular candidates, [k][i]
  size_t blist;    // int[bmax*cap]  boundary candidates, [k][i]
  size_t fbq;      // int[cap]       particles a sweep queued for a whole-wave stencil scan
  size_t fbqc;     // int[2]         queue lengths of the even / odd sweeps
  size_t ftmp0;    // float[cmax]    payload staging
  size_t itmp0;    // int[cmax]
  size_t itmp1;    // int[cmax]
  size_t ftmp1;    // float[cmax]
  size_t v3tmp0;   // float3[cmax]
  size_t v3tmp1;   // float3[cmax]
  size_t bnd4;     // float4[bcap]   packed sorted boundary
  // adaptive merge/split
  size_t cover;    // float4[cap]    cover vectors
  size_t cand;     // int[cap]       merge candidate
  size_t dens;     // float[cap]     local density
  size_t mstate;   // int[cap]       merge claim state
  size_t owner;    // int[cap]       lowest successful claimer per target
  size_t und0;     // int[cap]       lowest undecided claimer per target (ping)
  size_t und1;     // int[cap]       (pong)
  size_t flags;    // int[cap]       generic predicate / position scratch
  size_t bsum;     // int[blocks+1]  per-block counts for ordered compaction
  size_t bsum2;    // int[blocks+1]
  size_t child;    // float4[cap]    split child positions
  size_t mrad;     // float[cap]     surface radii for the depth merge / split rules
  size_t itmp2;    // int[cmax]
  size_t itmp3;    // int[cmax]
  size_t total;
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

inline long long grid_cells(const gg_grid &g) {
  return (long long)g.cx * (long long)g.cy * (long long)g.cz;
}

inline Layout make_layout(int capacity, int bcap, gg_grid g, int max_sweeps,
                          int gmax, int bmax) {
  Layout L{};
  const size_t C1 = (size_t)grid_cells(g) + 1;
  const size_t cmax = (size_t)(capacity > bcap ? capacity : bcap) + 1;
  const size_t cap = (size_t)capacity + 1;
  const size_t nblk0 = (cap + kBlock - 1) / kBlock + 2;
  const size_t parts = ((C1 > nblk0 ? C1 : nblk0) + kScanTile - 1) / kScanTile + 1;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    size_t at = o;
    o = align256(o + bytes);
    return at;
  };
  L.count = take(C1 * 4);
  L.key = take(cmax * 4);
  L.slot = take(cmax * 4);
  L.perm = take(cmax * 4);
  L.itmp = take(cmax * 4);
  L.part = take(parts * 4);
  L.big = take(C1 * 4);
  L.ctr = take(64 * 4);
  L.ctr2 = take(64 * 4);
  L.dirty = take(C1 * 4);
  L.bounds = take(16);
  L.bpart = take(nblk0 * 8);
  L.x0 = take(cap * 16);
  L.off0 = take(cap * 16);
  L.off1 = take(cap * 16);
  L.d0 = take(cap * 16);
  L.xb = take(cap * 16);
  L.meta = take(cap * 16);
  L.glist = take((size_t)gmax * cap * 4);
  L.blist = take((size_t)bmax * cap * 4);
  L.fbq = take(cap * 4);
  L.fbqc = take(2 * 4);
  L.ftmp0 = take(cmax * 4);
  L.itmp0 = take(cmax * 4);
  L.itmp1 = take(cmax * 4);
  L.ftmp1 = take(cmax * 4);
  L.v3tmp0 = take(cmax * 12);
  L.v3tmp1 = take(cmax * 12);
  L.bnd4 = take(((size_t)bcap + 1) * 16);
  const size_t nblk = (cap + kBlock - 1) / kBlock + 2;
  L.cover = take(cap * 16);
  L.cand = take(cap * 4);
  L.dens = take(cap * 4);
  L.mstate = take(cap * 4);
  L.owner = take(cap * 4);
  L.und0 = take(cap * 4);
  L.und1 = take(cap * 4);
  L.flags = take(cap * 4);
  L.bsum = take(nblk * 4);
  L.bsum2 = take(nblk * 4);
  L.child = take(cap * 16);
  L.mrad = take(cap * 4);
  L.itmp2 = take(cmax * 4);
  L.itmp3 = take(cmax * 4);
  L.total = o;
  return L;
}

template <class T> inline T *at(const gg_workspace *ws, size_t off) {
  return reinterpret_cast<T *>(static_cast<char *>(ws->base) + off);
}

inline int check_ws(const gg_workspace *ws) {
  if (!ws || !ws->base) return GG_ERR_INVALID;
  if (ws->capacity < 0 || ws->boundary_capacity < 0 || ws->max_sweeps < 0 ||
      ws->gmax < 0 || ws->bmax < 0)
    return GG_ERR_INVALID;
  if (ws->grid.cx <= 0 || ws->grid.cy <= 0 || ws->grid.cz <= 0 ||
      !(ws->grid.cell_length > 0.0f))
    return GG_ERR_INVALID;
  const long long C1 = grid_cells(ws->grid) + 1;
  if (C1 > (long long)kScanTile * kScanMaxParts) return GG_ERR_INVALID;
  Layout L = make_layout(ws->capacity, ws->boundary_capacity, ws->grid,
                         ws->max_sweeps, ws->gmax, ws->bmax);
  if (ws->bytes < L.total) return GG_ERR_WORKSPACE;
  return GG_OK;
}

inline int hip_status(hipError_t e) { return e == hipSuccess ? GG_OK : GG_ERR_HIP; }

}  // namespace gg

#define GG_LAUNCH_CHECK()                                  \
  do {                                                     \
    hipError_t _e = hipGetLastError();                     \
    if (_e != hipSuccess) return GG_ERR_HIP;               \
  } while (0)

#define GG_TRY(expr)                                       \
  do {                                                     \
    int _s = (expr);                                       \
    if (_s != GG_OK) return _s;                            \
  } while (0)
