/* libfftcorr_hip: MI355X (gfx950) implementation of fftcorr's FFT-based
 * anisotropic 2-point correlation and power-spectrum multipole path.
 *
 * Plain C ABI: opaque handles, plain pointers and sizes.  Host arrays passed
 * in are borrowed for the duration of the call; handles own device memory.
 * Every call returns an int status (FC_OK = 0); fc_last_error() returns the
 * calling thread's message for the last failure.  Each entry point names the
 * reference interface (file:line, relative to the cshallue/fftcorr root) it
 * replaces.
 *
 * Enum codes are the reference's ints, with new codes appended:
 *   distribution scheme  0 NGP, 1 TSC, 2 D12 wavelet (reference),
 *                        3 CIC (new)        cc/particle_mesh/distribution_functions.h:11-15
 *   window correction    0 none, 1 TSC, 2 TSC aliased (reference),
 *                        3 CIC, 4 CIC aliased (new)   cc/correlate/window_correct.h:8-12
 * Precision: FC_FP64 computes every stage in FP64 (the reference's Float);
 * FC_FP32 keeps grids, FFTs and products in FP32 with FP64 sums, Y_lm
 * normalisation and histogram accumulation; FC_MIXED stores grids and
 * spectra in FP32 like FC_FP32 but computes the z and y forward FFT passes in
 * FP64 (values rounded to FP32 only where they go to LDS or HBM) and the
 * (pruned) inverse FFT of the power/cross spectrum in FP64.
 */
#ifndef FFTCORR_HIP_H
#define FFTCORR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (mapped to Python exceptions by the host layer). */
#define FC_OK 0
#define FC_ERR_INVALID 1 /* invalid argument  -> ValueError   (std::invalid_argument) */
#define FC_ERR_RANGE 2   /* out of range      -> IndexError   (std::out_of_range)     */
#define FC_ERR_RUNTIME 3 /* runtime failure   -> RuntimeError (std::runtime_error)    */
#define FC_ERR_NOMEM 4   /* allocation failed -> MemoryError  (std::bad_alloc)        */
#define FC_ERR_HIP 5     /* HIP runtime error -> RuntimeError                          */

#define FC_FP64 0
#define FC_FP32 1
#define FC_MIXED 2

/* Correlation algorithm: FC_ALGO_FUSED (default) runs the fused x pass and
 * the pruned inverse; FC_ALGO_PLAIN runs the reference's stage sequence
 * (full forward FFT, conjugate product, k-subgrid extraction per multipole,
 * full inverse FFT, r-subgrid extraction per multipole;
 * cc/correlate/correlator.h:273-357, :405-491) on the device. */
#define FC_ALGO_FUSED 0
#define FC_ALGO_PLAIN 1

#define FC_PERIODIC 0 /* PeriodicCorrelator, cc/correlate/correlator.h:255-357 */
#define FC_YAMAMOTO 1 /* Correlator,         cc/correlate/correlator.h:360-505 */

typedef struct fc_ctx fc_ctx;
typedef struct fc_grid fc_grid;
typedef struct fc_corr fc_corr;
typedef struct fc_batch fc_batch;

/* ---- runtime ---------------------------------------------------------- */
const char* fc_last_error(void);
int fc_version(void);
int fc_device_count(int* count);
/* Free and total device memory in bytes (admission control of the serving
 * layer; no reference counterpart). */
int fc_device_memory(int device, size_t* free_bytes, size_t* total_bytes);
/* Device memory held by this library on `device` (every handle's and
 * context's buffers) and its peak since the last reset; the serving layer
 * admits jobs against it (no reference counterpart). */
int fc_device_bytes(int device, size_t* in_use, size_t* peak);
int fc_device_bytes_reset_peak(int device);
/* One device + one HIP stream; all handles made from it run on that stream. */
int fc_ctx_create(int device, fc_ctx** out);
int fc_ctx_destroy(fc_ctx* ctx);
int fc_ctx_synchronize(fc_ctx* ctx);
/* Device bytes held by the context's reused work buffers (mass assignment),
 * and their release (after the context's pending work; later calls grow them
 * again).  No reference counterpart. */
int fc_ctx_footprint(fc_ctx* ctx, size_t* bytes);
int fc_ctx_release_scratch(fc_ctx* ctx);

/* ---- ConfigSpaceGrid: cc/grid/config_space_grid.h:10-56,
 *      fftcorr/grid/config_space_grid.pyx:11-207 -------------------------- */
/* Device-resident [n0,n1,n2] density grid, zeroed (config_space_grid.h:12-16).
 * precision selects FP64 or FP32 storage. */
int fc_grid_create(fc_ctx* ctx, const int shape[3], const double posmin[3],
                   double cell_size, int precision, fc_grid** out);
int fc_grid_destroy(fc_grid* grid);
int fc_grid_clear(fc_grid* grid); /* config_space_grid.h:28 */
int fc_grid_upload(fc_grid* grid, const double* host);   /* [n0*n1*n2] float64 */
int fc_grid_download(fc_grid* grid, double* host);       /* [n0*n1*n2] float64 */
/* array_ops::sum / sumsq (cc/array/array_ops.cc:95-137), FP64, deterministic. */
int fc_grid_stats(fc_grid* grid, double* sum, double* sumsq);
/* array_ops::add_scalar / multiply_by (cc/array/array_ops.cc:55-93). */
int fc_grid_add_scalar(fc_grid* grid, double s);
int fc_grid_multiply_by(fc_grid* grid, double s);
/* Periodic delta normalisation (cc/fftcorr.cc:369-375): d = (d - mean) * (1/mean),
 * evaluated in FP64. */
int fc_grid_normalize(fc_grid* grid, double mean);
/* Raw device pointer of the grid data (FP64 or FP32 per precision). */
int fc_grid_device_ptr(fc_grid* grid, void** ptr);

/* ---- MassAssignor: cc/particle_mesh/mass_assignor.h:30-234,
 *      fftcorr/particle_mesh/mass_assignor.pyx:18-109 ---------------------- */
typedef struct fc_ma_stats {
  uint64_t num_added;   /* mass_assignor.h:55 */
  uint64_t num_skipped; /* mass_assignor.h:56 */
  double totw;          /* mass_assignor.h:57, FP64 */
  double totwsq;        /* mass_assignor.h:58, FP64 */
  double sort_time;     /* seconds, device sort (mass_assignor.h:59) */
  double window_time;   /* seconds, device assignment (mass_assignor.h:60) */
} fc_ma_stats;

/* Adds n particles to the grid (add_particles_to_buffer + flush,
 * mass_assignor.h:72-191).  pos is [n, ncol] float64 with ncol 3 or 4 (column 3
 * is the weight when ncol == 4); weights is [n] float64 or NULL (then
 * w_scalar is every particle's weight).  shift (grid units) displaces every
 * particle before assignment (0.5 builds an interlaced grid; new).  Stats are
 * accumulated into *stats.  pos/weights are host pointers. */
/* Scheme 2 (D12 wavelet, distribution_functions.h:162-217) uses the sample
 * table of cc/particle_mesh/d12.h (8 cells x 64 samples, `wavesample`
 * order), which the caller supplies (it is the reference's data, not part of
 * this library); NULL clears it.  Process-wide. */
int fc_set_wavelet_table(const double* table, int n);
int fc_mass_assign(fc_grid* grid, const double* pos, int64_t n, int ncol,
                   const double* weights, double w_scalar, int scheme,
                   int periodic_wrap, double shift, fc_ma_stats* stats);
/* Device memory a call of n particles on this grid needs beyond what its
 * context holds: *grow = growth of the context's reused work buffers (kept
 * after the call), *transient = staging of host inputs (host_input = 1, as
 * fc_mass_assign; freed on return).  No reference counterpart. */
int fc_mass_assign_footprint(fc_grid* grid, int64_t n, int ncol, int has_weights,
                             int host_input, size_t* grow, size_t* transient);
/* Same, with pos/weights already in device memory (HBM-resident catalogs). */
int fc_mass_assign_device(fc_grid* grid, const double* pos_dev, int64_t n, int ncol,
                          const double* weights_dev, double w_scalar, int scheme,
                          int periodic_wrap, double shift, fc_ma_stats* stats);

/* ---- Randoms and displacement fields: fftcorr/catalog/random_catalog.py:8-87,
 *      fftcorr/grid/utils.pyx:19-55 ------------------------------------------ */
/* apply_displacement_field: out[i] = pos[i] + disp[cell(pos[i])] for host
 * arrays pos/out [n][3] and disp [n0][n1][n2][3]; FC_ERR_INVALID (the
 * reference's ValueError) if a position is outside the grid without wrap. */
int fc_apply_displacement(fc_grid* grid, const double* pos, int64_t n, const double* disp,
                          int periodic_wrap, double* out);
/* add_random_particles with the randoms drawn on the device: n uniform points
 * in the grid's box from the counter-based stream `seed` (Philox4x32-10),
 * optionally displaced by disp (host [n0][n1][n2][3] or NULL), assigned with
 * weight `weight` in batches of `batch`, statistics accumulated into stats. */
int fc_add_random_particles(fc_grid* grid, int64_t n, double weight, int scheme,
                            int periodic_wrap, const double* disp, uint64_t seed, int64_t batch,
                            fc_ma_stats* stats);
/* Device pointers: pos [n][3] = particles first .. first+n-1 of stream seed,
 * uniform in [lo, hi) per axis; raw: out [n][4] Philox4x32-10 words of the
 * counters (first + i, w2, w3) under key seed (known-answer testing). */
int fc_random_uniform_device(fc_ctx* ctx, double* pos, int64_t n, const double lo[3],
                             const double hi[3], uint64_t seed, uint64_t first);
int fc_random_raw_device(fc_ctx* ctx, uint32_t* out, int64_t n, uint64_t seed, uint64_t first,
                         uint32_t w2, uint32_t w3);
/* The same into host arrays. */
int fc_random_uniform(fc_ctx* ctx, double* pos, int64_t n, const double lo[3], const double hi[3],
                      uint64_t seed, uint64_t first);
int fc_random_raw(fc_ctx* ctx, uint32_t* out, int64_t n, uint64_t seed, uint64_t first,
                  uint32_t w2, uint32_t w3);

/* ---- Correlators: cc/correlate/correlator.h:18-505,
 *      fftcorr/correlate/correlator.pyx:12-184 ----------------------------- */
/* kind FC_PERIODIC or FC_YAMAMOTO; posmin is used by FC_YAMAMOTO only (observer
 * at the survey origin, correlator.h:381-402).  Throws the reference's
 * invalid_argument cases (histogram_list.h:19-23) as FC_ERR_INVALID. */
int fc_corr_create(fc_ctx* ctx, int kind, const int shape[3], double cell_size,
                   const double posmin[3], int window_correct, double rmax, double dr,
                   double kmax, double dk, int maxell, int precision, fc_corr** out);
int fc_corr_destroy(fc_corr* corr);
/* Device bytes the correlator holds now (*held) and will hold once its
 * correlation path has run (*full; interlaced = 1 for
 * fc_corr_autocorrelate_interlaced).  No reference counterpart. */
int fc_corr_footprint(fc_corr* corr, int interlaced, size_t* held, size_t* full);
/* Geometry of the outputs: r/k bin counts, number of multipoles, subgrids. */
int fc_corr_geometry(fc_corr* corr, int* n_rbins, int* n_kbins, int* n_ell,
                     int rshape[3], int kshape[3]);
/* BaseCorrelator::set_grid2 (correlator.h:53-67): FFT of grid2 kept on device. */
int fc_corr_set_grid2(fc_corr* corr, fc_grid* grid2);
/* BaseCorrelator::set_grid2_fft (correlator.h:69-77): host complex128
 * [n0, n1, n2/2+1] (interleaved re, im). */
int fc_corr_set_grid2_fft(fc_corr* corr, const double* host_complex);
/* PeriodicCorrelator::autocorrelate (correlator.h:258-260) /
 * Correlator::autocorrelate (correlator.h:372-374). */
int fc_corr_autocorrelate(fc_corr* corr, fc_grid* grid);
/* cross_correlate(grid1) against the stored grid2 FFT (correlator.h:262-264,
 * :376-378). */
int fc_corr_cross_correlate(fc_corr* corr, fc_grid* grid1);
/* Interlaced periodic autocorrelation (new): grid_b holds the same catalogue
 * assigned with shift = 0.5 cells; F = (F_a + F_b e^{i k.(h/2)(1,1,1)}) / 2. */
int fc_corr_autocorrelate_interlaced(fc_corr* corr, fc_grid* grid_a, fc_grid* grid_b);
/* Output accessors (correlator.h:79-94).  Arrays sized per fc_corr_geometry;
 * counts/hists are [n_ell, nbins] row-major.  Any pointer may be NULL. */
int fc_corr_results(fc_corr* corr, double* r_mid, int* r_counts, double* r_hist,
                    double* k_mid, int* k_counts, double* k_hist, double* zerolag);
/* Device time (ms) of the last correlate call, from HIP events on the
 * correlator's stream. */
int fc_corr_last_time_ms(fc_corr* corr, double* ms);

/* Per-stage device timing (new; replaces the reference's Timer accessors of
 * correlator.h:96-106 and fft_grid.h:36-53).  With profiling on, every
 * correlate call brackets each stage's launches with HIP events on the
 * correlator's stream; fc_corr_stage_times returns the accumulated
 * milliseconds and launch counts per stage (arrays of FC_NSTAGES). */
#define FC_STAGE_FFT_Z_FWD 0 /* z-pass R2C (+ fused load)            */
#define FC_STAGE_FFT_Y_FWD 1 /* y-pass forward C2C                   */
#define FC_STAGE_FFT_X_FWD 2 /* x-pass forward C2C                   */
#define FC_STAGE_COMBINE 3   /* interlacing / conjugate products     */
#define FC_STAGE_HIST_K 4    /* k-space Y_lm weights + histogram     */
#define FC_STAGE_FFT_X_INV 5 /* x-pass inverse C2C                   */
#define FC_STAGE_FFT_Y_INV 6 /* y-pass inverse C2C                   */
#define FC_STAGE_FFT_Z_INV 7 /* z-pass C2R                           */
#define FC_STAGE_HIST_R 8    /* r-space Y_lm weights + histogram     */
#define FC_STAGE_SUBGRID 9   /* Yamamoto subgrid accumulation        */
#define FC_NSTAGES 10
int fc_corr_set_profiling(fc_corr* corr, int on);
/* Selects FC_ALGO_FUSED or FC_ALGO_PLAIN for later calls (no reference
 * counterpart: the reference has one algorithm).  FC_ALGO_PLAIN needs
 * FC_FP64 or FC_FP32. */
int fc_corr_set_algorithm(fc_corr* corr, int algo);
int fc_corr_stage_times(fc_corr* corr, double* ms, int64_t* launches);

/* ---- Batches of periodic autocorrelations (no reference counterpart: the
 *      reference correlates one grid per call, correlator.pyx:72-184; this
 *      runs many same-geometry mocks per launch) ---------------------------- */
/* nbatch member grids [shape] stored back to back on the device, one shared
 * PeriodicCorrelator geometry (arguments as fc_corr_create).  Every pass is
 * one launch over all members, except the x pass of lengths the fused x
 * kernel does not take (not a power of two from 8 to 512, e.g. 384), which
 * runs member by member. */
int fc_batch_create(fc_ctx* ctx, const int shape[3], double cell_size, int window_correct,
                    double rmax, double dr, double kmax, double dk, int maxell, int precision,
                    int nbatch, fc_batch** out);
int fc_batch_destroy(fc_batch* batch);
/* Member i's density grid (owned by the batch; fill it with fc_mass_assign,
 * fc_grid_upload, fc_grid_normalize ...; fc_grid_destroy on it fails). */
int fc_batch_grid(fc_batch* batch, int i, fc_grid** out);
/* 1: capture the launch sequence of fc_batch_autocorrelate as a HIP graph the
 * first time a member count is used and replay it afterwards. */
int fc_batch_set_graph(fc_batch* batch, int on);
/* Autocorrelates members 0 .. n-1 (PeriodicCorrelator::autocorrelate each). */
int fc_batch_autocorrelate(fc_batch* batch, int n);
/* Member i's histograms ([n_ell][nbins] each; bin centres and geometry from
 * fc_batch_correlator + fc_corr_results / fc_corr_geometry). */
int fc_batch_results(fc_batch* batch, int i, double* r_hist, int* r_counts, double* k_hist,
                     int* k_counts, double* zerolag);
/* The batch's shared correlator (geometry, bin centres; owned by the batch). */
int fc_batch_correlator(fc_batch* batch, fc_corr** out);
/* Device time of the last fc_batch_autocorrelate. */
int fc_batch_last_time_ms(fc_batch* batch, double* ms);

/* ---- HistogramList: cc/histogram/histogram_list.h:10-67,
 *      fftcorr/histogram/histogram_list.pyx:9-44 -------------------------- */
/* accumulate(ih, x, y) for one row: bin = floor((x - minval)/binsize), values
 * outside [0, nbins) dropped; counts (int32) and hist (float64) rows are
 * updated in place.  x, y are host float64 arrays of length n. */
int fc_hist_accumulate(fc_ctx* ctx, double minval, double binsize, int nbins,
                       const double* x, const double* y, int64_t n, int* counts,
                       double* hist);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): fc_hist_accumulate with x, y (length n) and the counts /
 * hist rows (nbins each, accumulated in place) in caller-owned device
 * memory.  It calls no hipMalloc, hipFree or allocating routine: its partial
 * sums live in a fixed 8 MiB device array per GPU, so nbins is limited to
 * 682 (FC_ERR_INVALID beyond).  Work is queued on ctx's stream and not waited
 * for (fc_ctx_synchronize).  Results equal fc_hist_accumulate's bit for bit.
 * All calls on one GPU share the array: calls that overlap in time on the
 * same GPU (several contexts or host threads) race and must be serialised by
 * the caller. */
int fc_hist_accumulate_device_noalloc(fc_ctx* ctx, double minval, double binsize, int nbins,
                                      const double* x_dev, const double* y_dev, int64_t n,
                                      int* counts_dev, double* hist_dev);

/* ---- 3D FFT on the device (FftGrid, cc/grid/fft_grid.cc:69-178) ---------- */
/* Unnormalised R2C of a real [n0,n1,n2] float64 host array into complex128
 * [n0,n1,n2/2+1] (forward sign -1), in the given precision. */
int fc_fft_r2c(fc_ctx* ctx, const int shape[3], int precision, const double* real_in,
               double* complex_out);
/* Unnormalised C2R (no 1/N), the inverse of fc_fft_r2c. */
int fc_fft_c2r(fc_ctx* ctx, const int shape[3], int precision,
               const double* complex_in, double* real_out);

/* ---- plan cache ("wisdom"): fftcorr/fftw/fftw.pyx:1-14 -------------------- */
int fc_plan_cache_export(const char* path);
int fc_plan_cache_import(const char* path);
int fc_plan_cache_forget(void);
/* Autotuner (the GPU analogue of FFTW_MEASURE planning, fft_grid.cc:69-115):
 * times the candidate radix plans and tiles of every pass of a [shape] plan
 * at `precision` on ctx's device, reps timed runs each, and records the
 * fastest in the plan cache (exported by fc_plan_cache_export).  FC_MIXED
 * times the mixed form's kernels (FP32 storage, FP64 arithmetic) and records
 * its choices for the FP32 plan.  best_ms (may be NULL) receives the best
 * times of the z pass, y pass, x pass and chunked z+y sequence.  Plans
 * created afterwards use the tuned choices; existing correlators keep
 * theirs. */
int fc_autotune(fc_ctx* ctx, const int shape[3], int precision, int reps, double best_ms[4]);
/* Tiles a new plan of this shape would use: {x columns, y columns, z rows,
 * z+y chunk KiB (0: unchunked)}. */
int fc_plan_tiles(int device, const int shape[3], int precision, int tiles[4]);
/* Radix plan variants a new plan of this shape would use: {x, y, z rows}
 * (0: first radix 8, 1: first radix 16; the z entry is the row FFT of
 * length n2/2).  The counterpart of the algorithm FFTW_MEASURE picks. */
int fc_plan_variants(int device, const int shape[3], int precision, int vars[3]);
/* The radices of plan variant `var` for an FFT of length L, first to last:
 * returns their count (at most 8) in *count, 0 when the register-chained
 * kernels have no such variant for L (other lengths use fixed plans). */
int fc_fft_radices(int L, int var, int radices[8], int* count);

#ifdef __cplusplus
}
#endif

#endif /* FFTCORR_HIP_H */
