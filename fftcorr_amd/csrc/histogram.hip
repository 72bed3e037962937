// HistogramList::accumulate on the device (cc/histogram/histogram_list.h:33-45).
#include "hist_device.h"
#include "ylm_device.h"

// Scratch of fc_hist_accumulate_device_noalloc: the per-block partial sums
// and counts of one histogram pass (kHistBlocks x nbins FP64 + int32), in a
// file-scope device array so that the entry point allocates nothing.  Every
// device the library runs on holds its own copy until the process ends.
// Host code never uses the array's name as an address: its device address
// on each device comes from hipGetSymbolAddress (noalloc_scratch_init).
constexpr size_t kNoAllocScratchBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace fc {

namespace {
std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_ptr;  // device -> address of `scratch` there
}  // namespace

// Looks up `scratch` on `device` once; the caller has made `device` current.
void noalloc_scratch_init(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr.count(device)) return;
  void* p = nullptr;
  FC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) fail(FC_ERR_HIP, "no device address for the histogram scratch array");
  g_scratch_ptr[device] = static_cast<unsigned char*>(p);
}

namespace {

unsigned char* noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it == g_scratch_ptr.end())
    fail(FC_ERR_RUNTIME, "histogram scratch not set up on device " + std::to_string(device));
  return it->second;
}

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

// Largest nbins whose partials fit the scratch array: kHistBlocks blocks of
// nbins FP64 sums and nbins int32 counts.
static constexpr int kNoAllocMaxBins =
    (int)(kNoAllocScratchBytes / ((size_t)fc::kHistBlocks * (sizeof(double) + sizeof(int))));

extern "C" int fc_hist_accumulate_device_noalloc(fc_ctx* ctx, double minval, double binsize,
                                                 int nbins, const double* x_dev,
                                                 const double* y_dev, int64_t n,
                                                 int* counts_dev, double* hist_dev) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    if (nbins <= 0) fc::fail(FC_ERR_INVALID, "nbins must be positive");
    if (nbins > kNoAllocMaxBins)
      fc::fail(FC_ERR_INVALID, "nbins " + std::to_string(nbins) +
                                   " exceeds the no-allocation limit of " +
                                   std::to_string(kNoAllocMaxBins) + " bins");
    if (!(binsize > 0)) fc::fail(FC_ERR_INVALID, "binsize must be positive.");
    if (!counts_dev || !hist_dev) fc::fail(FC_ERR_INVALID, "null output row");
    if (n < 0) fc::fail(FC_ERR_INVALID, "negative length");
    if (n == 0) return;
    if (!x_dev || !y_dev) fc::fail(FC_ERR_INVALID, "null input");
    const int thr = fc::hist_threads(1, nbins);
    if (thr == 0) fc::fail(FC_ERR_INVALID, "too many histogram bins");
    fc::DeviceGuard dg(ctx->device);
    unsigned char* base = fc::noalloc_scratch(ctx->device);
    // The same block count and kernels as fc_hist_accumulate: the sums come
    // out in the same order, so the two entry points agree bit for bit.
    int nb = (int)std::min<int64_t>(fc::kHistBlocks, (n + thr - 1) / thr);
    if (nb < 1) nb = 1;
    double* part_hist = reinterpret_cast<double*>(base);
    int* part_counts = reinterpret_cast<int*>(base + sizeof(double) * (size_t)nb * nbins);
    hipStream_t s = ctx->stream;
    fc::PairVals f{x_dev, y_dev, minval, binsize, nbins};
    // Every block < nb writes all of its partials: nothing in the scratch
    // needs clearing.
    hipLaunchKernelGGL((fc::k_hist<1, fc::PairVals>), dim3(nb), dim3(thr),
                       fc::hist_lds_bytes(1, nbins, thr), s, f, n, nbins, part_hist,
                       part_counts);
    FC_LAUNCH_CHECK();
    hipLaunchKernelGGL((fc::k_hist_final<1>), dim3(nbins), dim3(256), 0, s, part_hist,
                       part_counts, (const int*)nullptr, nb, nbins, hist_dev, counts_dev, 0, 0,
                       1);
    FC_LAUNCH_CHECK();
  });
}
