// Owning device array: the host DArray<T> of the reference (DArray.hpp:12-257)
// over hipMalloc and the library's compaction primitive.
//
// Same members and semantics: length()/capacity()/addr(); compact(flags)
// keeps the elements with flag != 1 in order and throws on a length
// mismatch; append() grows to max(new length, 2 x capacity) when full;
// remove(start, count) closes the gap; resize(new_capacity) reallocates and
// keeps min(length, capacity) elements; verifyArrayOrdering() compacts a set
// of arrays and checks on the host that survivors kept their order.
// Elements must be a whole number of 32-bit words (float, int, float3, ...).
#pragma once
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <typeinfo>
#include <vector>

#include "../granular_gpu.h"
#include "Global.hpp"

template <typename T>
class DArray {
  static_assert(sizeof(T) % 4 == 0, "DArray elements must be whole 32-bit words");

 public:
  explicit DArray(unsigned int size, unsigned int initial_capacity = 0)
      : _addr(nullptr), _length(size), _capacity(std::max(size, initial_capacity)) {
    if (size > 1000000000u) throw std::runtime_error("Suspiciously large array size requested");
    const size_t bytes = sizeof(T) * (size_t)std::max(_capacity, 1u);
    size_t free_mem = 0, total_mem = 0;
    if (hipMemGetInfo(&free_mem, &total_mem) == hipSuccess && bytes > free_mem) {
      std::cerr << "Not enough GPU memory. Required: " << bytes << " Available: " << free_mem
                << std::endl;
      throw std::runtime_error("Insufficient GPU memory");
    }
    if (hipMalloc(reinterpret_cast<void **>(&_addr), bytes) != hipSuccess) {
      _addr = nullptr;
      throw std::runtime_error("HIP malloc failed");
    }
    if (hipMemset(_addr, 0, bytes) != hipSuccess) {
      (void)hipFree(_addr);
      _addr = nullptr;
      throw std::runtime_error("HIP memset failed");
    }
  }

  DArray(const DArray &) = delete;
  DArray &operator=(const DArray &) = delete;

  ~DArray() {
    if (_addr) {
      if (hipFree(_addr) != hipSuccess)
        std::cerr << "Failed to free HIP memory in destructor" << std::endl;
      _addr = nullptr;
    }
  }

  unsigned int length() const { return _length; }
  unsigned int capacity() const { return _capacity; }
  T *addr() const { return _addr; }

  // Stable removal of the elements whose flag is 1 (0 and -1 are kept).
  unsigned int compact(const DArray<int> &removal_flags) {
    if (removal_flags.length() != _length) {
      std::cerr << "Array type: " << typeid(T).name() << "\nRemoval flags length ("
                << removal_flags.length() << ") doesn't match array length (" << _length << ")"
                << std::endl;
      throw std::runtime_error("Removal flags length mismatch");
    }
    if (_length == 0) return 0;
    const size_t sbytes = gg_compact_scratch_bytes(_length);
    void *tmp = nullptr, *scratch = nullptr;
    if (hipMalloc(&tmp, sizeof(T) * (size_t)_length) != hipSuccess)
      throw std::runtime_error("HIP malloc failed in compact");
    if (hipMalloc(&scratch, sbytes + 256) != hipSuccess) {
      (void)hipFree(tmp);
      throw std::runtime_error("HIP malloc failed in compact");
    }
    int *d_count = reinterpret_cast<int *>(static_cast<char *>(scratch) + sbytes);
    const int rc = gg_compact(_addr, tmp, _length, (int)(sizeof(T) / 4), removal_flags.addr(),
                              d_count, scratch, sbytes, nullptr);
    int kept = 0;
    hipError_t e = hipSuccess;
    if (rc == GG_OK) e = hipMemcpy(&kept, d_count, sizeof(int), hipMemcpyDeviceToHost);
    if (rc == GG_OK && e == hipSuccess && kept > 0)
      e = hipMemcpy(_addr, tmp, sizeof(T) * (size_t)kept, hipMemcpyDeviceToDevice);
    (void)hipFree(tmp);
    (void)hipFree(scratch);
    if (rc != GG_OK) throw std::runtime_error(std::string("compact: ") + gg_status_string(rc));
    if (e != hipSuccess) throw std::runtime_error("Compact copy failed");
    _length = (unsigned int)kept;
    return _length;
  }

  void append(const DArray<T> &other) {
    const unsigned int other_length = other.length();
    if (other_length == 0) return;
    const unsigned int new_length = _length + other_length;
    if (new_length > _capacity) resize(std::max(new_length, _capacity * 2));
    if (hipMemcpy(_addr + _length, other.addr(), sizeof(T) * (size_t)other_length,
                  hipMemcpyDeviceToDevice) != hipSuccess)
      throw std::runtime_error("Failed to append array");
    _length = new_length;
    if (hipDeviceSynchronize() != hipSuccess)
      throw std::runtime_error("Error occurred during array append");
  }

  void remove(unsigned int start, unsigned int count) {
    if ((unsigned long long)start + count > _length)
      throw std::runtime_error("Remove range exceeds array bounds");
    const unsigned int tail = _length - start - count;
    if (count > 0 && tail > 0) {
      T *tmp = nullptr;
      if (hipMalloc(reinterpret_cast<void **>(&tmp), sizeof(T) * (size_t)tail) != hipSuccess)
        throw std::runtime_error("HIP malloc failed in remove");
      hipError_t e = hipMemcpy(tmp, _addr + start + count, sizeof(T) * (size_t)tail,
                               hipMemcpyDeviceToDevice);
      if (e == hipSuccess)
        e = hipMemcpy(_addr + start, tmp, sizeof(T) * (size_t)tail, hipMemcpyDeviceToDevice);
      (void)hipFree(tmp);
      if (e != hipSuccess) throw std::runtime_error("Remove copy failed");
    }
    _length -= count;
  }

  static void verifyArrayOrdering(const std::vector<DArray<T> *> &arrays,
                                  const DArray<int> &removal_flags, const char *arrayNames[]) {
    if (arrays.empty()) return;
    const unsigned int length = arrays[0]->length();
    for (size_t i = 1; i < arrays.size(); i++)
      if (arrays[i]->length() != length)
        throw std::runtime_error(std::string("Array length mismatch between ") + arrayNames[0] +
                                 " and " + arrayNames[i]);
    std::vector<std::vector<T>> before(arrays.size()), after(arrays.size());
    std::vector<int> flags(length);
    for (size_t i = 0; i < arrays.size(); i++) {
      before[i].resize(length);
      HIP_CALL(hipMemcpy(before[i].data(), arrays[i]->addr(), sizeof(T) * (size_t)length,
                         hipMemcpyDeviceToHost));
    }
    HIP_CALL(hipMemcpy(flags.data(), removal_flags.addr(), sizeof(int) * (size_t)length,
                       hipMemcpyDeviceToHost));
    for (auto *a : arrays) a->compact(removal_flags);
    const unsigned int new_length = arrays[0]->length();
    for (size_t i = 0; i < arrays.size(); i++) {
      if (arrays[i]->length() != new_length)
        throw std::runtime_error("Inconsistent lengths after compaction");
      after[i].resize(new_length);
      HIP_CALL(hipMemcpy(after[i].data(), arrays[i]->addr(), sizeof(T) * (size_t)new_length,
                         hipMemcpyDeviceToHost));
    }
    std::vector<unsigned int> survivors;
    for (unsigned int i = 0; i < length; i++)
      if (flags[i] != 1) survivors.push_back(i);
    if (survivors.size() != new_length)
      throw std::runtime_error("Array ordering not maintained during compaction");
    for (unsigned int i = 0; i < new_length; i++)
      for (size_t a = 0; a < arrays.size(); a++)
        if (std::memcmp(&after[a][i], &before[a][survivors[i]], sizeof(T)) != 0) {
          std::cerr << "Order mismatch in " << arrayNames[a] << " at index " << i
                    << " (original index " << survivors[i] << ")" << std::endl;
          throw std::runtime_error("Array ordering not maintained during compaction");
        }
  }

  void resize(unsigned int new_capacity) {
    T *new_addr = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&new_addr),
                  sizeof(T) * (size_t)std::max(new_capacity, 1u)) != hipSuccess)
      throw std::runtime_error("HIP malloc failed in resize");
    const unsigned int keep = std::min(_length, new_capacity);
    if (keep > 0)
      HIP_CALL(hipMemcpy(new_addr, _addr, sizeof(T) * (size_t)keep, hipMemcpyDeviceToDevice));
    (void)hipFree(_addr);
    _addr = new_addr;
    _capacity = new_capacity;
    _length = keep;
  }

  // Sets the length after a device pass changed the element count in place
  // (merge/split compaction works inside the capacity).
  void set_length(unsigned int n) {
    if (n > _capacity) throw std::runtime_error("set_length beyond capacity");
    _length = n;
  }

  // Grows the capacity to at least cap, keeping length and contents.
  void reserve(unsigned int cap) {
    if (cap > _capacity) {
      const unsigned int len = _length;
      resize(cap);
      _length = len;
    }
  }

 public:
  T *_addr;
  unsigned int _length;
  unsigned int _capacity;
};
