// HistogramList::accumulate on the device (cc/histogram/histogram_list.h:33-45).
#include "hist_device.h"
#include "ylm_device.h"

namespace fc {
namespace {

struct PairVals {
  const double* x;
  const double* y;
  double minval, binsize;
  int nbins;
  __device__ __forceinline__ void operator()(int64_t idx, int* bin, double* v) const {
    *bin = hist_bin(x[idx], minval, binsize, nbins);
    v[0] = y[idx];
  }
};

}  // namespace
}  // namespace fc

extern "C" int fc_hist_accumulate(fc_ctx* ctx, double minval, double binsize, int nbins,
                                  const double* x, const double* y, int64_t n, int* counts,
                                  double* hist) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    if (nbins <= 0) fc::fail(FC_ERR_INVALID, "nbins must be positive");
    if (!(binsize > 0)) fc::fail(FC_ERR_INVALID, "binsize must be positive.");
    if (!counts || !hist) fc::fail(FC_ERR_INVALID, "null output row");
    if (n < 0) fc::fail(FC_ERR_INVALID, "negative length");
    if (n == 0) return;
    if (!x || !y) fc::fail(FC_ERR_INVALID, "null input");
    fc::DeviceGuard dg(ctx->device);
    hipStream_t s = ctx->stream;
    fc::DeviceBuffer dx(sizeof(double) * n), dy(sizeof(double) * n);
    fc::DeviceBuffer dh(sizeof(double) * nbins), dc(sizeof(int) * nbins);
    FC_HIP(hipMemcpyAsync(dx.as<void>(), x, sizeof(double) * n, hipMemcpyHostToDevice, s));
    FC_HIP(hipMemcpyAsync(dy.as<void>(), y, sizeof(double) * n, hipMemcpyHostToDevice, s));
    FC_HIP(hipMemsetAsync(dh.as<void>(), 0, sizeof(double) * nbins, s));
    FC_HIP(hipMemsetAsync(dc.as<void>(), 0, sizeof(int) * nbins, s));
    fc::HistScratch hs;
    fc::PairVals f{dx.as<double>(), dy.as<double>(), minval, binsize, nbins};
    fc::run_hist<1>(f, n, nbins, hs, dh.as<double>(), dc.as<int>(), 0, 0, 1, s);
    std::vector<double> h(nbins);
    std::vector<int> c(nbins);
    FC_HIP(hipMemcpyAsync(h.data(), dh.as<void>(), sizeof(double) * nbins, hipMemcpyDeviceToHost,
                          s));
    FC_HIP(hipMemcpyAsync(c.data(), dc.as<void>(), sizeof(int) * nbins, hipMemcpyDeviceToHost, s));
    FC_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < nbins; ++i) {
      hist[i] += h[i];
      counts[i] += c[i];
    }
  });
}
