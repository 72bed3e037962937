// Runtime entry points: error text, version, device count, contexts.
#include "fc_internal.h"

#include <algorithm>
#include <atomic>

namespace fc {
namespace {
thread_local std::string g_last_error;
std::atomic<int64_t> g_dev_bytes[kMaxTrackedDevices];
std::atomic<int64_t> g_dev_peak[kMaxTrackedDevices];
}  // namespace
void set_last_error(const std::string& msg) { g_last_error = msg; }

void note_device_bytes(int device, int64_t delta) {
  if (device < 0 || device >= kMaxTrackedDevices) return;
  const int64_t now = g_dev_bytes[device].fetch_add(delta) + delta;
  int64_t peak = g_dev_peak[device].load();
  while (now > peak && !g_dev_peak[device].compare_exchange_weak(peak, now)) {
  }
}
}  // namespace fc

extern "C" {

const char* fc_last_error(void) { return fc::g_last_error.c_str(); }

int fc_version(void) { return 1; }

int fc_device_count(int* count) {
  return fc::guard([&] {
    if (!count) fc::fail(FC_ERR_INVALID, "null count");
    int n = 0;
    FC_HIP(hipGetDeviceCount(&n));
    *count = n;
  });
}

int fc_device_memory(int device, size_t* free_bytes, size_t* total_bytes) {
  return fc::guard([&] {
    if (!free_bytes || !total_bytes) fc::fail(FC_ERR_INVALID, "null argument");
    int n = 0;
    FC_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) fc::fail(FC_ERR_RANGE, "no such device");
    fc::DeviceGuard dg(device);
    FC_HIP(hipMemGetInfo(free_bytes, total_bytes));
  });
}

int fc_device_bytes(int device, size_t* in_use, size_t* peak) {
  return fc::guard([&] {
    if (device < 0 || device >= fc::kMaxTrackedDevices)
      fc::fail(FC_ERR_RANGE, "no such device");
    if (in_use) *in_use = (size_t)std::max<int64_t>(0, fc::g_dev_bytes[device].load());
    if (peak) *peak = (size_t)std::max<int64_t>(0, fc::g_dev_peak[device].load());
  });
}

int fc_device_bytes_reset_peak(int device) {
  return fc::guard([&] {
    if (device < 0 || device >= fc::kMaxTrackedDevices)
      fc::fail(FC_ERR_RANGE, "no such device");
    fc::g_dev_peak[device].store(fc::g_dev_bytes[device].load());
  });
}

int fc_ctx_create(int device, fc_ctx** out) {
  return fc::guard([&] {
    if (!out) fc::fail(FC_ERR_INVALID, "null output handle");
    int n = 0;
    FC_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
      fc::fail(FC_ERR_INVALID, "device " + std::to_string(device) + " not present (" +
                                   std::to_string(n) + " visible)");
    fc::DeviceGuard g(device);
    std::unique_ptr<fc_ctx> c(new fc_ctx());
    c->device = device;
    // Default-priority stream (every stream in this library keeps the default).
    FC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    FC_HIP(hipGetDeviceProperties(&prop, device));
    c->arch_cus = prop.multiProcessorCount;
    fc::noalloc_scratch_init(device);  // once per device, with the device current
    *out = c.release();
  });
}

int fc_ctx_destroy(fc_ctx* ctx) {
  return fc::guard([&] {
    if (!ctx) return;
    fc::DeviceGuard g(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& b : ctx->scratch) b.release();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

int fc_ctx_footprint(fc_ctx* ctx, size_t* bytes) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    if (!bytes) fc::fail(FC_ERR_INVALID, "null argument");
    size_t b = ctx->wavelet.bytes();
    for (const auto& buf : ctx->scratch) b += buf.bytes();
    *bytes = b;
  });
}

int fc_ctx_release_scratch(fc_ctx* ctx) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    fc::DeviceGuard g(ctx->device);
    FC_HIP(hipStreamSynchronize(ctx->stream));
    for (auto& b : ctx->scratch) b.release();
  });
}

int fc_ctx_synchronize(fc_ctx* ctx) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    fc::DeviceGuard g(ctx->device);
    FC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
