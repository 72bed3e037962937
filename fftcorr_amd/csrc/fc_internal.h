// Shared runtime pieces of libfftcorr_hip: status codes, error capture,
// device buffers, contexts, and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fftcorr_hip.h"

namespace fc {

// Exceptions carry the C-ABI status they map to.
struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

inline void fail(int status, const std::string& msg) { throw Error(status, msg); }

void set_last_error(const std::string& msg);

// Records the device address of the no-allocation histogram scratch array on
// `device` (the current device); called once per device at context creation.
void noalloc_scratch_init(int device);

#define FC_HIP(expr)                                                              \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      ::fc::fail(_e == hipErrorOutOfMemory ? FC_ERR_NOMEM : FC_ERR_HIP,           \
                 std::string("HIP error: ") + hipGetErrorString(_e) + " at " +    \
                     __FILE__ + ":" + std::to_string(__LINE__) + " (" #expr ")"); \
    }                                                                             \
  } while (0)

#define FC_LAUNCH_CHECK() FC_HIP(hipGetLastError())

// Wraps a C-ABI body: converts exceptions into status codes + last error.
// The HIP runtime's per-thread error state is cleared first, so that the
// launch checks of this call report only this call's errors (and not one
// left by an earlier call whose status was not checked, e.g. in a release
// path).
template <typename F>
int guard(F&& f) {
  (void)hipGetLastError();
  try {
    f();
    return FC_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return FC_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return FC_ERR_RUNTIME;
  }
}

// Bytes of device memory the library holds per device (every DeviceBuffer
// allocation), and the peak since the last reset: the measured footprint the
// serving layer admits against.
constexpr int kMaxTrackedDevices = 64;
void note_device_bytes(int device, int64_t delta);

// Owning device allocation (hipMalloc / hipFree).
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t bytes) { allocate(bytes); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept
      : ptr_(o.ptr_), bytes_(o.bytes_), owned_(o.owned_), dev_(o.dev_) {
    o.ptr_ = nullptr;
    o.bytes_ = 0;
    o.owned_ = true;
  }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      ptr_ = o.ptr_;
      bytes_ = o.bytes_;
      owned_ = o.owned_;
      dev_ = o.dev_;
      o.ptr_ = nullptr;
      o.bytes_ = 0;
      o.owned_ = true;
    }
    return *this;
  }
  void allocate(size_t bytes) {
    release();
    if (bytes == 0) return;
    FC_HIP(hipGetDevice(&dev_));
    FC_HIP(hipMalloc(&ptr_, bytes));
    bytes_ = bytes;
    note_device_bytes(dev_, (int64_t)bytes);
  }
  // Grows (never shrinks); contents are not preserved.
  void reserve(size_t bytes) {
    if (bytes > bytes_) allocate(bytes);
  }
  void release() {
    if (ptr_ && owned_) {
      (void)hipFree(ptr_);
      note_device_bytes(dev_, -(int64_t)bytes_);
    }
    ptr_ = nullptr;
    bytes_ = 0;
    owned_ = true;
  }
  // Non-owning view of memory owned elsewhere (a slice of a batch buffer).
  void wrap(void* p, size_t bytes) {
    release();
    ptr_ = p;
    bytes_ = bytes;
    owned_ = false;
  }
  template <typename T = void>
  T* as() const {
    return static_cast<T*>(ptr_);
  }
  size_t bytes() const { return bytes_; }

 private:
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
  bool owned_ = true;
  int dev_ = 0;
};

}  // namespace fc

// One device + one stream.  Every handle created from a context issues its
// work on that stream, in call order.
struct fc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // Scratch reused across calls on this context (sort temp storage, staging).
  fc::DeviceBuffer scratch[12];  // per-context work buffers, grown on demand
  fc::DeviceBuffer wavelet;       // D12 sample table on this device
  uint64_t wavelet_ver = 0;       // version of the process-wide table it holds
  int arch_cus = 0;
};

namespace fc {

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    (void)hipGetDevice(&prev);
    if (prev != dev) FC_HIP(hipSetDevice(dev));
  }
  ~DeviceGuard() {
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (prev >= 0 && cur != prev) (void)hipSetDevice(prev);
  }
};

inline void check_ctx(const fc_ctx* c) {
  if (!c) fail(FC_ERR_INVALID, "null context");
}

inline unsigned grid1d(uint64_t n, unsigned block, unsigned cap = 1u << 20) {
  uint64_t g = (n + block - 1) / block;
  if (g == 0) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

// Host-side FP64 helpers that reproduce the reference's setup arithmetic
// (cc/array/array_ops.cc:7-13).
inline std::vector<double> sequence(double start, double step, int n) {
  std::vector<double> s(n);
  for (int i = 0; i < n; ++i) s[i] = start + i * step;
  return s;
}

}  // namespace fc
