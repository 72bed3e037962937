// Runtime entry points: error text, version, device count, contexts.
#include "fc_internal.h"

namespace fc {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
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

int fc_ctx_synchronize(fc_ctx* ctx) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    fc::DeviceGuard g(ctx->device);
    FC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
