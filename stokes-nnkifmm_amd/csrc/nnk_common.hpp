// Shared host-side helpers: error type, HIP checks, device buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace nnk {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define NNK_HIP(x)                                                                    \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess)                                                             \
      throw ::nnk::Error(3, std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                                __FILE__ + ":" + std::to_string(__LINE__) + " (" #x ")"); \
  } while (0)

// Owning device allocation.
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t n) {
    release();
    if (n == 0) n = 16;
    NNK_HIP(hipMalloc(&p, n));
    bytes = n;
  }
  void zero(hipStream_t s = 0) { if (p) NNK_HIP(hipMemsetAsync(p, 0, bytes, s)); }
  template <class T> T *as() const { return static_cast<T *>(p); }
  template <class T> void upload(const std::vector<T> &v) {
    alloc(v.size() * sizeof(T));
    if (!v.empty()) NNK_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
};

}  // namespace nnk
