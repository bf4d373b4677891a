// Shared host helpers: HIP error handling and RAII device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace cx {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};

#define CX_HIP(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::cx::HipError(std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                           __FILE__ + ":" + std::to_string(__LINE__));                   \
  } while (0)

#define CX_LAUNCH_CHECK() CX_HIP(hipGetLastError())

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    n = count;
    if (count) CX_HIP(hipMalloc(&p, count * sizeof(T)));
  }
  void zero(hipStream_t s) {
    if (n) CX_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
  void upload(const T* src, size_t count, hipStream_t s) {
    alloc(count);
    if (count) CX_HIP(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  void download(T* dst, size_t count, hipStream_t s) const {
    if (count) CX_HIP(hipMemcpyAsync(dst, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  operator T*() const { return p; }
};

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace cx
