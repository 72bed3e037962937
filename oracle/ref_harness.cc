// Checker-only C ABI over the FFTW-free reference headers.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile from the reference's own
// sources where they lie (cc/histogram, cc/correlate/spherical_harmonics.h,
// cc/correlate/window_correct.h, cc/grid/config_space_grid.h,
// cc/particle_mesh/*, cc/array/array_ops.cc, cc/catalog_reader.h) into
// oracle/_ref/libfftcorr_ref.so.  Only tests and the CPU-baseline leg of the
// benchmark load it, as a checker.  The correlator itself (cc/correlate/
// correlator.h) needs <fftw3.h>, which this image lacks, so it is not built.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>

#include "array/row_major_array.h"
#include "catalog_reader.h"
#include "correlate/spherical_harmonics.h"
#include "correlate/window_correct.h"
#include "grid/config_space_grid.h"
#include "histogram/histogram_list.h"
#include "particle_mesh/mass_assignor.h"

namespace {
bool g_wavelet_ready = false;
void ensure_wavelet() {
  if (!g_wavelet_ready) {
    setup_wavelet();
    g_wavelet_ready = true;
  }
}
}  // namespace

extern "C" {

int ref_sequence(double start, double step, int n, double* out) {
  Array1D<Float> s = array_ops::sequence(start, step, n);
  for (int i = 0; i < n; ++i) out[i] = s[i];
  return 0;
}

// make_ylm over the axes; mult (nullable) and out are [nx, ny, nz].
int ref_make_ylm(int ell, int m, const double* x, int nx, const double* y, int ny,
                 const double* z, int nz, double coeff, int exponent,
                 const double* mult, double* out) {
  try {
    Array1D<Float> ax(nx), ay(ny), az(nz);
    for (int i = 0; i < nx; ++i) ax[i] = x[i];
    for (int i = 0; i < ny; ++i) ay[i] = y[i];
    for (int i = 0; i < nz; ++i) az[i] = z[i];
    RowMajorArrayPtr<Float, 3> yp({nx, ny, nz}, out);
    if (mult) {
      RowMajorArrayPtr<Float, 3> mp({nx, ny, nz}, const_cast<double*>(mult));
      make_ylm(ell, m, ax, ay, az, coeff, exponent, &mp, &yp);
    } else {
      make_ylm(ell, m, ax, ay, az, coeff, exponent, NULL, &yp);
    }
  } catch (const std::exception&) {
    return 1;
  }
  return 0;
}

double ref_window(int kind, double k, double cell_size) {
  std::unique_ptr<WindowSquaredNorm> w =
      make_window_function(static_cast<WindowCorrection>(kind));
  return w->evaluate(k, cell_size);
}

// One HistogramList(n, minval, maxval, binsize); one accumulate(ih, x, y).
// Writes bins[nbins], counts[n*nbins], hist[n*nbins]; returns nbins or -1.
int ref_histogram(int n, double minval, double maxval, double binsize, int ih,
                  const double* x, const double* y, int64_t len, double* bins,
                  int* counts, double* hist) {
  try {
    HistogramList h(n, minval, maxval, binsize);
    if (len > 0) {
      RowMajorArrayPtr<Float, 1> xa({(int)len}, const_cast<double*>(x));
      RowMajorArrayPtr<Float, 1> ya({(int)len}, const_cast<double*>(y));
      h.accumulate(ih, xa, ya);
    }
    int nb = h.nbins();
    for (int i = 0; i < nb; ++i) bins[i] = h.bins()[i];
    for (int i = 0; i < n * nb; ++i) {
      counts[i] = h.counts().data()[i];
      hist[i] = h.hist_values().data()[i];
    }
    return nb;
  } catch (const std::exception&) {
    return -1;
  }
}

int ref_grid_coords(const int* shape, const double* posmin, double cell, int wrap,
                    const double* pos, double* out) {
  ConfigSpaceGrid g({shape[0], shape[1], shape[2]}, {posmin[0], posmin[1], posmin[2]},
                    cell);
  return g.get_grid_coords(pos, wrap != 0, out) ? 1 : 0;
}

// MassAssignor over posw [n,4]; grid_out [shape] gets the assigned density.
int ref_mass_assign(const int* shape, const double* posmin, double cell, int scheme,
                    int wrap, int64_t buffer_size, const double* posw, int64_t n,
                    double* grid_out, uint64_t* num_added, uint64_t* num_skipped,
                    double* totw, double* totwsq) {
  try {
    ensure_wavelet();
    ConfigSpaceGrid g({shape[0], shape[1], shape[2]}, {posmin[0], posmin[1], posmin[2]},
                      cell);
    MassAssignor ma(g, static_cast<DistributionScheme>(scheme), wrap != 0,
                    (uint64)buffer_size);
    if (n > 0) {
      RowMajorArrayPtr<Float, 2> p({(int)n, 4}, const_cast<double*>(posw));
      ma.add_particles_to_buffer(p);
    }
    ma.flush();
    std::memcpy(grid_out, g.data().data(), sizeof(double) * g.size());
    *num_added = ma.num_added();
    *num_skipped = ma.num_skipped();
    *totw = ma.totw();
    *totwsq = ma.totwsq();
  } catch (const std::exception&) {
    return 1;
  }
  return 0;
}

// The CLI's gridding steps (cc/fftcorr.cc:327-368): header, padding by
// sep/1.5 when not periodic, cell size, MassAssignor(batch 1e6), one or two
// catalogs.  Writes posmin_out[3], *cell_out and grid_out[ngrid^3]
// (grid_out may be NULL to query the geometry only).
int ref_cli_grid(const char* file1, const char* file2, int ngrid, double sep,
                 int periodic, int scheme, double* posmin_out, double* cell_out,
                 double* grid_out, uint64_t* num_added, double* totw,
                 double* totwsq) {
  try {
    ensure_wavelet();
    std::array<int, 3> shape = {ngrid, ngrid, ngrid};
    std::array<Float, 3> posmin, posmax;
    CatalogReader reader;
    reader.read_header(file1, posmin, posmax);
    if (!periodic) {
      Float pad = sep / 1.5;
      for (int i = 0; i < 3; i++) {
        posmin[i] -= pad;
        posmax[i] += pad;
      }
    }
    Float cell_size = -123.0;
    for (int i = 0; i < 3; ++i)
      cell_size = std::max(cell_size, (posmax[i] - posmin[i]) / shape[i]);
    for (int i = 0; i < 3; ++i) posmin_out[i] = posmin[i];
    *cell_out = cell_size;
    if (!grid_out) return 0;
    ConfigSpaceGrid grid(shape, posmin, cell_size);
    MassAssignor ma(grid, static_cast<DistributionScheme>(scheme), periodic != 0,
                    1000000);
    reader.read_galaxies(file1, &ma);
    if (file2 && file2[0]) reader.read_galaxies(file2, &ma);
    std::memcpy(grid_out, grid.data().data(), sizeof(double) * grid.size());
    *num_added = ma.num_added();
    *totw = ma.totw();
    *totwsq = ma.totwsq();
  } catch (const std::exception&) {
    return 1;
  }
  return 0;
}

// The D12 wavelet's sample table in the reference's own order
// (d12.h: 8 cells x 64 sub-cell samples).
int ref_wavelet_table(double* out) {
  for (int i = 0; i < WCELLS * WAVESAMPLE; ++i) out[i] = wavesample[i];
  return WCELLS * WAVESAMPLE;
}

}  // extern "C"
