// Host construction of bin-sorted subgrid cells (see BinnedCells).
#include <algorithm>
#include <thread>

#include "hist_device.h"

namespace fc {

void build_binned_cells(BinnedCells& bc, const std::vector<double>& ax,
                        const std::vector<double>& ay, const std::vector<double>& az,
                        double binsize, int nbins, int64_t track_cell,
                        const std::function<uint64_t(uint32_t)>* order,
                        std::vector<int16_t>* cell_bins) {
  const int s0 = (int)ax.size(), s1 = (int)ay.size(), s2 = (int)az.size();
  if ((uint64_t)s0 * s1 * s2 > 0xffffffffULL) fail(FC_ERR_INVALID, "subgrid too large");
  if (cell_bins) {
    if (nbins >= 32768) fail(FC_ERR_RUNTIME, "internal: too many bins for 16-bit cell bins");
    cell_bins->assign((size_t)s0 * s1 * s2, (int16_t)-1);
  }
  // Bin of cell (i, j, k): floor((|r| - 0) / binsize) with
  // |r| = sqrt((x*x + y*y) + z*z) in FP64 (correlator.h:129-135, :180-186;
  // histogram_list.h:57).  Host x86-64 code without FMA, like the reference.
  auto bin_of = [&](int i, int j, int k) {
#pragma clang fp contract(off)
    const double norm = std::sqrt(ax[i] * ax[i] + ay[j] * ay[j] + az[k] * az[k]);
    const double f = std::floor((norm - 0.0) / binsize);
    return (f >= 0.0 && f < (double)nbins) ? (int)f : -1;
  };
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  nt = std::min<unsigned>(nt, (unsigned)s0);
  // Pass 1: per-thread counts over fixed i-slabs.
  std::vector<std::vector<int64_t>> cnt(nt, std::vector<int64_t>(nbins, 0));
  auto slab = [&](unsigned t, int* lo, int* hi) {
    *lo = (int)((int64_t)s0 * t / nt);
    *hi = (int)((int64_t)s0 * (t + 1) / nt);
  };
  {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
      th.emplace_back([&, t] {
        int lo, hi;
        slab(t, &lo, &hi);
        for (int i = lo; i < hi; ++i)
          for (int j = 0; j < s1; ++j)
            for (int k = 0; k < s2; ++k) {
              const int b = bin_of(i, j, k);
              if (b >= 0) ++cnt[t][b];
              if (cell_bins) (*cell_bins)[((size_t)i * s1 + j) * s2 + k] = (int16_t)b;
            }
      });
    for (auto& x : th) x.join();
  }
  // Offsets: bin-major, then slab order, so each bin keeps raster order.
  std::vector<int64_t> start(nbins + 1, 0);
  std::vector<std::vector<int64_t>> off(nt, std::vector<int64_t>(nbins, 0));
  int64_t pos = 0;
  bc.h_counts.assign(nbins, 0);
  for (int b = 0; b < nbins; ++b) {
    start[b] = pos;
    for (unsigned t = 0; t < nt; ++t) {
      off[t][b] = pos;
      pos += cnt[t][b];
    }
    bc.h_counts[b] = (int)(pos - start[b]);
  }
  start[nbins] = pos;
  if (pos > 0x7fffffffLL) fail(FC_ERR_INVALID, "subgrid too large");
  std::vector<uint32_t> cells(pos > 0 ? pos : 1);
  {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
      th.emplace_back([&, t] {
        int lo, hi;
        slab(t, &lo, &hi);
        std::vector<int64_t>& o = off[t];
        for (int i = lo; i < hi; ++i)
          for (int j = 0; j < s1; ++j)
            for (int k = 0; k < s2; ++k) {
              const int b = bin_of(i, j, k);
              if (b >= 0) {
                cells[o[b]++] = (uint32_t)(((int64_t)i * s1 + j) * s2 + k);
              }
            }
      });
    for (auto& x : th) x.join();
  }
  if (order) {
    // Each bin's segment sorted by (order key, cell); bins split over threads.
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
      th.emplace_back([&, t] {
        std::vector<std::pair<uint64_t, uint32_t>> tmp;
        for (int b = (int)t; b < nbins; b += (int)nt) {
          const int64_t lo = start[b], hi = start[b + 1];
          tmp.resize(hi - lo);
          for (int64_t q = lo; q < hi; ++q) tmp[q - lo] = {(*order)(cells[q]), cells[q]};
          std::sort(tmp.begin(), tmp.end());
          for (int64_t q = lo; q < hi; ++q) cells[q] = tmp[q - lo].second;
        }
      });
    for (auto& x : th) x.join();
  }
  bc.tracked_pos = -1;
  if (track_cell >= 0)
    for (int64_t q = 0; q < pos; ++q)
      if ((int64_t)cells[q] == track_cell) {
        bc.tracked_pos = q;
        break;
      }
  // Chunks of at most kSegChunk entries, never straddling a bin.
  std::vector<int4> chunks;
  std::vector<int> first(nbins + 1, 0);
  for (int b = 0; b < nbins; ++b) {
    first[b] = (int)chunks.size();
    for (int64_t q = start[b]; q < start[b + 1]; q += kSegChunk) {
      const int len = (int)std::min<int64_t>(kSegChunk, start[b + 1] - q);
      chunks.push_back(make_int4(b, (int)q, len, 0));
    }
  }
  first[nbins] = (int)chunks.size();
  std::vector<int> launch(chunks.size());
  for (size_t q = 0; q < launch.size(); ++q) launch[q] = (int)q;
  std::stable_sort(launch.begin(), launch.end(), [&](int a, int b) {
    return cells[chunks[a].y] < cells[chunks[b].y];
  });
  bc.n = pos;
  bc.nbins = nbins;
  bc.nchunks = (int)chunks.size();
  bc.cells.allocate(sizeof(uint32_t) * cells.size());
  bc.counts.allocate(sizeof(int) * nbins);
  bc.chunks.allocate(sizeof(int4) * std::max<size_t>(chunks.size(), 1));
  bc.chunk_first.allocate(sizeof(int) * first.size());
  bc.order.allocate(sizeof(int) * std::max<size_t>(launch.size(), 1));
  FC_HIP(hipMemcpy(bc.cells.as<void>(), cells.data(), sizeof(uint32_t) * cells.size(),
                   hipMemcpyHostToDevice));
  FC_HIP(hipMemcpy(bc.counts.as<void>(), bc.h_counts.data(), sizeof(int) * nbins,
                   hipMemcpyHostToDevice));
  if (!chunks.empty())
    FC_HIP(hipMemcpy(bc.chunks.as<void>(), chunks.data(), sizeof(int4) * chunks.size(),
                     hipMemcpyHostToDevice));
  FC_HIP(hipMemcpy(bc.chunk_first.as<void>(), first.data(), sizeof(int) * first.size(),
                   hipMemcpyHostToDevice));
  if (!launch.empty())
    FC_HIP(hipMemcpy(bc.order.as<void>(), launch.data(), sizeof(int) * launch.size(),
                     hipMemcpyHostToDevice));
}

}  // namespace fc
