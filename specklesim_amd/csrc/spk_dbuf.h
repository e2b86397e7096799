// spk_dbuf.h -- a device array that owns its allocation (hipMalloc / hipFree),
// keeps it while later requests fit (so a handle set up again for a mesh of
// the same size allocates nothing) and is move-only.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace spk {

// device buffer of n elements; alloc keeps the allocation when the request
// fits its capacity, so the contents after alloc are undefined either way
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0, cap = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n), cap(o.cap) {
    o.p = nullptr;
    o.n = o.cap = 0;
  }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) {
      reset();
      p = o.p;
      n = o.n;
      cap = o.cap;
      o.p = nullptr;
      o.n = o.cap = 0;
    }
    return *this;
  }
  ~DBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = cap = 0;
  }
  hipError_t alloc(size_t count) {
    if (p && count <= cap) {
      n = count;
      return hipSuccess;
    }
    reset();
    if (count == 0) return hipSuccess;
    const hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    n = cap = count;
    return hipSuccess;
  }
  hipError_t upload(const std::vector<T> &v) {
    hipError_t e = alloc(v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  }
};

}  // namespace spk
