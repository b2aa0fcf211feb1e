// common.hpp -- shared host/device utilities for libnekrs_hip (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "nrs_hip.h"

namespace nrs {

// Error handling: every C entry point catches nrs::Error and returns its code;
// the message is kept per thread for nrs_last_error().
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define NRS_HIP(call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw ::nrs::Error(NRS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_) +   \
                                          " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define NRS_CHECK(cond, msg)                                                                \
  do {                                                                                      \
    if (!(cond)) throw ::nrs::Error(NRS_ERR_INVALID, std::string(msg));                     \
  } while (0)

#define NRS_LAUNCH_CHECK() NRS_HIP(hipGetLastError())

// Owning device buffer (setup-time allocations only; never on the solve path).
template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) NRS_HIP(hipMalloc(&p, count * sizeof(T)));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr; n = 0;
  }
  void upload(const T *h, size_t count) {
    if (count != n) alloc(count);
    if (count) NRS_HIP(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
  }
  void upload(const std::vector<T> &h) { upload(h.data(), h.size()); }
  std::vector<T> download() const {
    std::vector<T> h(n);
    if (n) NRS_HIP(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
  }
  T *get() const { return p; }
  size_t size() const { return n; }
};

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// Dtype tags used by the C ABI.
template <typename T> struct dtype_of;
template <> struct dtype_of<double> { static constexpr int value = NRS_F64; };
template <> struct dtype_of<float> { static constexpr int value = NRS_F32; };
template <> struct dtype_of<int> { static constexpr int value = NRS_I32; };

}  // namespace nrs
