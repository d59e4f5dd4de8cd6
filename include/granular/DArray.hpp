// Owning device array of the host API.  Keeps the member functions and
// signatures of the reference's DArray<T> (DArray.hpp:12-257: length(),
// capacity(), addr(), compact(flags), append(other), remove(start, count),
// resize(capacity), verifyArrayOrdering(...)) so that code written against it
// compiles unchanged; the storage scheme is this library's own:
//
// - Every array owns two device blocks of `capacity` elements, the live one
//   and a spare.  compact() and remove() write the surviving elements into
//   the spare (the library's ordered compaction, gg_compact) and then swap
//   the two, so neither allocates nor copies device-to-device per call.
// - compact() also keeps a scratch block for the compaction's per-block
//   counts, grown with the capacity; the kept count is read back once.
// - append() grows the capacity geometrically (by half, plus a 64-element
//   floor) when the new length does not fit.
// - Device errors throw std::runtime_error naming the operation and the HIP
//   error string.
// Elements must be a whole number of 32-bit words (float, int, float3, ...).
#pragma once
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <typeinfo>
#include <utility>
#include <vector>

#include "../granular_gpu.h"
#include "Global.hpp"

namespace granular_detail {

inline void hip_ok(hipError_t e, const char *op) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string("DArray ") + op + ": " + hipGetErrorString(e));
}

// Zero-filled device block of `count` elements of `elem` bytes (at least one).
inline void *device_block(size_t count, size_t elem, const char *op) {
  const size_t n = std::max<size_t>(count, 1);
  if (n > std::numeric_limits<size_t>::max() / elem)
    throw std::runtime_error(std::string("DArray ") + op + ": size overflows");
  void *p = nullptr;
  hip_ok(hipMalloc(&p, n * elem), op);
  const hipError_t e = hipMemset(p, 0, n * elem);
  if (e != hipSuccess) {
    (void)hipFree(p);
    hip_ok(e, op);
  }
  return p;
}

inline void release(void *p) {
  if (p) (void)hipFree(p);
}

}  // namespace granular_detail

template <typename T>
class DArray {
  static_assert(sizeof(T) % 4 == 0, "DArray elements must be whole 32-bit words");

 public:
  explicit DArray(unsigned int size, unsigned int initial_capacity = 0)
      : _addr(nullptr), _length(size), _capacity(std::max(size, initial_capacity)) {
    _addr = static_cast<T *>(granular_detail::device_block(_capacity, sizeof(T), "allocate"));
  }

  DArray(const DArray &) = delete;
  DArray &operator=(const DArray &) = delete;

  ~DArray() {
    granular_detail::release(_addr);
    granular_detail::release(_spare);
    granular_detail::release(_scratch);
  }

  unsigned int length() const { return _length; }
  unsigned int capacity() const { return _capacity; }
  T *addr() const { return _addr; }

  // Keeps the elements whose flag is not 1, in their order; returns the new
  // length.  The flags must cover exactly the current length.
  unsigned int compact(const DArray<int> &removal_flags) {
    if (removal_flags.length() != _length)
      throw std::runtime_error(std::string("DArray<") + typeid(T).name() +
                               ">::compact: " + std::to_string(removal_flags.length()) +
                               " flags for " + std::to_string(_length) + " elements");
    if (_length == 0) return 0;
    ensure_spare();
    const size_t need = gg_compact_scratch_bytes(_length) + sizeof(int);
    if (need > _scratch_bytes) {
      granular_detail::release(_scratch);
      _scratch = nullptr;
      _scratch_bytes = 0;
      const size_t grow = gg_compact_scratch_bytes(_capacity) + sizeof(int);
      _scratch = granular_detail::device_block(std::max(need, grow), 1, "compact scratch");
      _scratch_bytes = std::max(need, grow);
    }
    int *kept_dev = reinterpret_cast<int *>(static_cast<char *>(_scratch) + _scratch_bytes -
                                            sizeof(int));
    const int rc = gg_compact(_addr, _spare, _length, (int)(sizeof(T) / 4), removal_flags.addr(),
                              kept_dev, _scratch, _scratch_bytes - sizeof(int), nullptr);
    if (rc != GG_OK) throw std::runtime_error(std::string("DArray compact: ") + gg_status_string(rc));
    int kept = 0;
    granular_detail::hip_ok(hipMemcpy(&kept, kept_dev, sizeof(int), hipMemcpyDeviceToHost),
                            "compact count");
    std::swap(_addr, _spare);
    _length = (unsigned int)kept;
    return _length;
  }

  // Appends the elements of other after the current ones.
  void append(const DArray<T> &other) {
    const unsigned int extra = other.length();
    if (extra == 0) return;
    const unsigned long long want = (unsigned long long)_length + extra;
    if (want > std::numeric_limits<unsigned int>::max())
      throw std::runtime_error("DArray append: length overflows");
    if (want > _capacity) {
      const unsigned long long grown = (unsigned long long)_capacity + _capacity / 2 + 64;
      resize_keep((unsigned int)std::min<unsigned long long>(
          std::max(want, grown), std::numeric_limits<unsigned int>::max()));
    }
    granular_detail::hip_ok(hipMemcpy(_addr + _length, other.addr(), sizeof(T) * (size_t)extra,
                                      hipMemcpyDeviceToDevice),
                            "append");
    _length = (unsigned int)want;
  }

  // Drops elements [start, start + count) and closes the gap.
  void remove(unsigned int start, unsigned int count) {
    if ((unsigned long long)start + count > _length)
      throw std::runtime_error("DArray remove: range [" + std::to_string(start) + ", " +
                               std::to_string((unsigned long long)start + count) +
                               ") is not inside the " + std::to_string(_length) + " elements");
    if (count == 0) return;
    const unsigned int tail = _length - start - count;
    ensure_spare();
    if (start > 0)
      granular_detail::hip_ok(
          hipMemcpy(_spare, _addr, sizeof(T) * (size_t)start, hipMemcpyDeviceToDevice), "remove");
    if (tail > 0)
      granular_detail::hip_ok(hipMemcpy(_spare + start, _addr + start + count,
                                        sizeof(T) * (size_t)tail, hipMemcpyDeviceToDevice),
                              "remove");
    std::swap(_addr, _spare);
    _length -= count;
  }

  // Compacts every array by the same flags and checks on the host that each
  // survivor i holds the value of the i-th unflagged element from before.
  // names (optional) label the arrays in the error text.
  static void verifyArrayOrdering(const std::vector<DArray<T> *> &arrays,
                                  const DArray<int> &removal_flags, const char *arrayNames[]) {
    if (arrays.empty()) return;
    auto label = [&](size_t a) {
      return arrayNames && arrayNames[a] ? std::string(arrayNames[a])
                                         : "array " + std::to_string(a);
    };
    const unsigned int n = arrays[0]->length();
    std::vector<int> flags(n);
    if (n > 0)
      granular_detail::hip_ok(hipMemcpy(flags.data(), removal_flags.addr(), sizeof(int) * (size_t)n,
                                        hipMemcpyDeviceToHost),
                              "verify flags");
    // expected survivors, computed on the host from the snapshots
    std::vector<std::vector<T>> expect(arrays.size());
    for (size_t a = 0; a < arrays.size(); ++a) {
      if (arrays[a]->length() != n)
        throw std::runtime_error("verifyArrayOrdering: " + label(a) + " has " +
                                 std::to_string(arrays[a]->length()) + " elements, " + label(0) +
                                 " has " + std::to_string(n));
      std::vector<T> snap(n);
      if (n > 0)
        granular_detail::hip_ok(hipMemcpy(snap.data(), arrays[a]->addr(), sizeof(T) * (size_t)n,
                                          hipMemcpyDeviceToHost),
                                "verify snapshot");
      for (unsigned int i = 0; i < n; ++i)
        if (flags[i] != 1) expect[a].push_back(snap[i]);
    }
    for (auto *arr : arrays) arr->compact(removal_flags);
    for (size_t a = 0; a < arrays.size(); ++a) {
      const unsigned int m = arrays[a]->length();
      if (m != expect[a].size())
        throw std::runtime_error("verifyArrayOrdering: " + label(a) + " kept " +
                                 std::to_string(m) + " elements, expected " +
                                 std::to_string(expect[a].size()));
      std::vector<T> got(m);
      if (m > 0)
        granular_detail::hip_ok(hipMemcpy(got.data(), arrays[a]->addr(), sizeof(T) * (size_t)m,
                                          hipMemcpyDeviceToHost),
                                "verify result");
      for (unsigned int i = 0; i < m; ++i)
        if (std::memcmp(&got[i], &expect[a][i], sizeof(T)) != 0)
          throw std::runtime_error("verifyArrayOrdering: " + label(a) + " survivor " +
                                   std::to_string(i) + " is not the " + std::to_string(i) +
                                   "-th unflagged element");
    }
  }

  // New capacity; the first min(length, new_capacity) elements are kept.
  void resize(unsigned int new_capacity) {
    const unsigned int keep = std::min(_length, new_capacity);
    _length = keep;
    resize_keep(new_capacity);
  }

  // Sets the length after a device pass changed the element count in place
  // (merge/split compaction works inside the capacity).
  void set_length(unsigned int n) {
    if (n > _capacity)
      throw std::runtime_error("DArray set_length: " + std::to_string(n) + " exceeds capacity " +
                               std::to_string(_capacity));
    _length = n;
  }

  // Grows the capacity to at least cap, keeping length and contents.
  void reserve(unsigned int cap) {
    if (cap > _capacity) resize_keep(cap);
  }

 private:
  // Reallocates both blocks at new_capacity (>= _length), keeping the
  // first _length elements.
  void resize_keep(unsigned int new_capacity) {
    T *fresh = static_cast<T *>(granular_detail::device_block(new_capacity, sizeof(T), "resize"));
    if (_length > 0) {
      const hipError_t e =
          hipMemcpy(fresh, _addr, sizeof(T) * (size_t)_length, hipMemcpyDeviceToDevice);
      if (e != hipSuccess) {
        granular_detail::release(fresh);
        granular_detail::hip_ok(e, "resize copy");
      }
    }
    granular_detail::release(_addr);
    granular_detail::release(_spare);
    _spare = nullptr;
    _addr = fresh;
    _capacity = new_capacity;
  }

  void ensure_spare() {
    if (!_spare)
      _spare = static_cast<T *>(granular_detail::device_block(_capacity, sizeof(T), "spare"));
  }

  T *_spare = nullptr;
  void *_scratch = nullptr;
  size_t _scratch_bytes = 0;

 public:
  T *_addr;
  unsigned int _length;
  unsigned int _capacity;
};
