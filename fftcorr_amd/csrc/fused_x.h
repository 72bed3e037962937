// Fused periodic x pass: forward x FFT of one or two grids (after their z and
// y passes), the interlacing combine, |F|^2 or F conj(G), the compact
// k-subgrid values, the inverse x FFT and the pruned store of the x lags the
// r subgrid reads -- one HBM read of each input and no write of the full
// power grid.  Replaces, for the x axis, the sequence execute_fft ->
// multiply_with_conjugation -> extract_fft_submatrix_c2r -> execute_ifft of
// cc/correlate/correlator.h:290-334 with cc/grid/fft_grid.cc:125-257.
#pragma once

#include "fft_kernels.h"

namespace fc {

constexpr int kModeAuto = 0, kModeCross = 1, kModeInterlaced = 2;

// ---- k-subgrid compact layout ---------------------------------------------
// pk[sx][sy][c] holds Re P at the element with signed offsets
// (sx - ox, sy - oy) along x, y and kz index c in [0, oz]: every element the
// k subgrid reads (directly or reflected, fft_grid.cc:215-257).  Slots along
// an axis of length n for index e: e + o (if e <= o) and e - n + o (if >= 0);
// both exist only when the subgrid was truncated to n + 1 cells.
struct KSlots {
  int ks0, ks1, ks2;
  __host__ __device__ __forceinline__ int oz() const { return ks2 / 2; }
  __host__ __device__ __forceinline__ int64_t at(int sx, int sy, int c) const {
    return ((int64_t)sx * ks1 + sy) * (ks2 / 2 + 1) + c;
  }
};

__device__ __forceinline__ int slot_a(int e, int o) { return e <= o ? e + o : -1; }
__device__ __forceinline__ int slot_b(int e, int n, int o) {
  return e - n + o >= 0 ? e - n + o : -1;
}

// Writes Re P of element (e, y, c) to every k-subgrid slot that reads it.
template <typename T>
__device__ __forceinline__ void put_pk(T* pk, const KSlots& ks, int n0, int n1, int e, int y,
                                       int c, T v) {
  if (c > ks.oz()) return;
  const int ox = ks.ks0 / 2, oy = ks.ks1 / 2;
  const int sx[2] = {slot_a(e, ox), slot_b(e, n0, ox)};
  const int sy[2] = {slot_a(y, oy), slot_b(y, n1, oy)};
  for (int a = 0; a < 2; ++a) {
    if (sx[a] < 0 || (a == 1 && sx[1] == sx[0])) continue;
    for (int b = 0; b < 2; ++b) {
      if (sy[b] < 0 || (b == 1 && sy[1] == sy[0])) continue;
      pk[ks.at(sx[a], sy[b], c)] = v;
    }
  }
}

// Combine of the fused x pass: slot s of v holds element e = t + out_m(s) L/E
// of the transformed tiles (v[0] grid a; v[NT - 1] grid b when interlaced;
// g the fully transformed grid 2 at the same elements in cross mode).  Writes
// P (in TI) to u in the inverse pass's input order and Re P (in T) to pkv[s]
// for x2_store_pk.  pyc: the y and z interlacing phases of this column
// (interlaced mode); ph0: the x phases exp(i pi f / L) of the signed
// frequencies f, loaded before the first use.
template <typename T, typename TI, int L, int MODE, int NT, int NG>
__device__ __forceinline__ void x2_combine(const cplx<T> (&v)[NT][FftCfg<L>::E],
                                           const cplx<T> (&g)[NG],
                                           cplx<TI> (&u)[1][FftCfg<L>::E],
                                           T (&pkv)[FftCfg<L>::E], const double2* ph0,
                                           double2 pyc, int t) {
  using CI = cplx<TI>;
  constexpr int E = FftCfg<L>::E;
  constexpr int TCOL = L / E;
  // The x phase of element e = t + m L/E is ph0[t] ph0[m L/E]: with
  // ph0[a] = exp(i pi f(a) / L) for the signed frequency f, f(e) = t + f(m L/E)
  // because L/2 is a multiple of L/E and t < L/E.  ph0[t] (times the column's
  // y and z phases) is one load per thread; the ph0[m L/E] are the same for
  // every lane (scalar loads).
  double2 pt = make_double2(1.0, 0.0);
  double2 pm[MODE == kModeInterlaced ? E : 1];
  if (MODE == kModeInterlaced) {
    pt = cmul(ph0[t], pyc);
#pragma unroll
    for (int s = 0; s < E; ++s) pm[s] = ph0[out_m<L>(s) * TCOL];
  }
  {
#pragma clang fp contract(off)
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const TI ax = (TI)v[0][s].x, ay = (TI)v[0][s].y;
      CI p;
      if (MODE == kModeInterlaced) {
        const double2 phd = cmul(pt, pm[s]);
        const TI px = (TI)phd.x, py = (TI)phd.y;
        const TI bx = (TI)v[NT - 1][s].x, by = (TI)v[NT - 1][s].y;
        const TI fx = TI(0.5) * (ax + (bx * px - by * py));
        const TI fy = TI(0.5) * (ay + (bx * py + by * px));
        p = CI{fx * fx + fy * fy, TI(0)};
      } else if (MODE == kModeCross) {
        const TI gx = (TI)g[s].x, gy = (TI)g[s].y;
        p = CI{ax * gx + ay * gy, ay * gx - ax * gy};
      } else {
        p = CI{ax * ax + ay * ay, TI(0)};
      }
      u[0][out_m<L>(s)] = p;
      pkv[s] = (T)p.x;
    }
  }
}

// A uniform value the optimiser cannot see through: offsets built as
// base + k * step from it stay additions (folding them back into
// (t + k) * stride would cost a quarter-rate integer multiply per element).
__device__ __forceinline__ uint32_t opaque_uniform(uint32_t x) {
  asm volatile("" : "+s"(x));
  return x;
}

// Re P of slot s (pkv, from x2_combine) to every compact k slot that reads
// element e = t + out_m(s) L/E of row y, column cg.  Issued after the inverse
// pass, so no load of the pass waits behind these stores.
template <typename T, int L>
__device__ __forceinline__ void x2_store_pk(const T (&pkv)[FftCfg<L>::E], T* pk,
                                            const KSlots& ks, int n1, int y, int cg, bool colok,
                                            int t) {
  constexpr int E = FftCfg<L>::E;
  constexpr int TCOL = L / E;
  const int oz = ks.oz();
  if (!(colok && cg <= oz)) return;
  const int ox = ks.ks0 / 2, oy = ks.ks1 / 2;
  int sy0 = slot_a(y, oy), sy1 = slot_b(y, n1, oy);
  if (sy1 == sy0) sy1 = -1;
  // Byte offsets (below 4 GB: the launcher checks the compact array's size)
  // from the uniform base, so the stores take a 32-bit offset each.
  constexpr uint32_t ts = sizeof(T);
  const uint32_t pstride = (uint32_t)ks.ks1 * (uint32_t)(oz + 1) * ts;
  const uint32_t prow0 = (uint32_t)(sy0 * (oz + 1) + cg) * ts;
  const uint32_t prow1 = (uint32_t)(sy1 * (oz + 1) + cg) * ts;
  // Slot rows (e + ox) pstride and (e - L + ox) pstride from e pstride =
  // t pstride + out_m(s) TCOL pstride: one multiply per thread.
  const uint32_t tp = (uint32_t)t * pstride;
  const uint32_t step = opaque_uniform((uint32_t)TCOL * pstride);
  const uint32_t r0 = opaque_uniform((uint32_t)ox * pstride);
  const uint32_t r1 = opaque_uniform((uint32_t)(ox - L) * pstride);
  char* pkb = reinterpret_cast<char*>(pk);
  auto put = [&](uint32_t off, T val) { *reinterpret_cast<T*>(pkb + off) = val; };
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int e = t + out_m<L>(s) * TCOL;
    int sx0 = slot_a(e, ox), sx1 = slot_b(e, L, ox);
    if (sx1 == sx0) sx1 = -1;
    const T val = pkv[s];
    const uint32_t ep = tp + (uint32_t)out_m<L>(s) * step;
    if (sx0 >= 0) {
      const uint32_t r = ep + r0;
      if (sy0 >= 0) put(r + prow0, val);
      if (sy1 >= 0) put(r + prow1, val);
    }
    if (sx1 >= 0) {
      const uint32_t r = ep + r1;
      if (sy0 >= 0) put(r + prow0, val);
      if (sy1 >= 0) put(r + prow1, val);
    }
  }
}

// Pruned store of the inverse x pass: q[sx][y][cg] for every sx < rs holds
// element (sx - off) mod L (offm = off mod L).
template <typename TI, int L>
__device__ __forceinline__ void x2_store_q(const cplx<TI> (&u)[1][FftCfg<L>::E], cplx<TI>* q,
                                           int n1, int y, int64_t ldq, int cg, int rs, int offm,
                                           int t) {
  using CI = cplx<TI>;
  constexpr int E = FftCfg<L>::E;
  constexpr int TCOL = L / E;
  // Byte offset sx A + (y ldq + cg) |CI| with A = n1 ldq |CI|: one multiply
  // per thread, then uniform steps (32-bit arithmetic throughout).
  const uint32_t A = (uint32_t)n1 * (uint32_t)ldq * (uint32_t)sizeof(CI);
  int base = t + offm;
  if (base >= L) base -= L;
  const uint32_t o0 =
      (uint32_t)base * A + ((uint32_t)y * (uint32_t)ldq + (uint32_t)cg) * (uint32_t)sizeof(CI);
  char* qb = reinterpret_cast<char*>(q);
  const uint32_t stepA = opaque_uniform((uint32_t)TCOL * A);
  const uint32_t wrapA = opaque_uniform((uint32_t)L * A);
#pragma unroll
  for (int s = 0; s < E; ++s) {
    int sx = base + out_m<L>(s) * TCOL;
    uint32_t off = o0 + (uint32_t)out_m<L>(s) * stepA;
    if (sx >= L) {
      sx -= L;
      off -= wrapA;
    }
    for (; sx < rs; sx += L, off += wrapA) st_stream(reinterpret_cast<CI*>(qb + off), u[0][s]);
  }
}

// ---- the fused kernel -------------------------------------------------------
// Columns per workgroup: 64 bytes of each row (8 FP32 or 4 FP64 complex).
template <typename T>
constexpr int px_tc() {
  return 64 / (int)sizeof(cplx<T>);
}
// Lengths the fused kernel takes: power-of-two plans whose first radix is E.
__host__ __device__ constexpr bool px_len_ok(int L) { return rc_len_ok(L); }
template <typename T, int L>
constexpr int px_threads() {
  return px_tc<T>() * (L / FftCfg<L>::E);
}
// Waves per SIMD the register budget is sized for: two workgroups per CU.
template <typename T>
constexpr int px_waves() {
  return sizeof(T) == 4 ? 4 : 2;
}
// The forward tiles hold the storage type T; the inverse tile holds TI.
template <typename T, typename TI, int L, int MODE>
constexpr size_t px_lds_bytes() {
  constexpr int tiles = MODE == kModeInterlaced ? 2 : 1;
  constexpr size_t fwd = tiles * (size_t)L * px_tc<T>() * sizeof(cplx<T>);
  constexpr size_t inv = (size_t)L * px_tc<T>() * sizeof(cplx<TI>);
  return fwd > inv ? fwd : inv;
}

// One workgroup per (batch member, y row, tile of TC kz columns) of the work
// grids [batch][n0][n1][ldc]; tiles are dealt so that each of the 8 XCDs
// (which receive consecutive workgroups in turn) runs a contiguous range of
// tiles and the two tiles sharing a 128-byte line meet in one L2.  Batch
// member b reads A, B at b * a_bstride and writes pk, q at b * pk_bstride,
// b * q_bstride (elements).  In cross mode B is the one shared grid-2
// spectrum (a_bstride does not apply to it).
//   A: grid a after its z and y passes (stored as T).
//   B: grid b likewise (interlaced) or the fully transformed grid 2 (cross).
//   pk: compact k-subgrid values (T); q[sx][y][c] (row stride ldq, stored
//   as TI): inverse x FFT at element (sx - off) mod L for every sx < rs
//   (offm = off mod L; rs may exceed L, then elements repeat).
// Loads, forward FFT (twiddles tw) and pk in T; combine, inverse and q in
// TI (FP64 for the mixed form), twiddles twi.
template <typename T, typename TI, int L, int MODE>
__global__ void __launch_bounds__((px_threads<T, L>()))
__attribute__((amdgpu_waves_per_eu(px_waves<T>())))
    k_periodic_x2(const cplx<T>* __restrict__ A, const cplx<T>* __restrict__ B, int n1, int n2h,
                  int64_t ldc, int tilesC, const cplx<T>* __restrict__ tw,
                  const cplx<TI>* __restrict__ twi, const double2* __restrict__ ph0,
                  const double2* __restrict__ ph1, const double2* __restrict__ ph2,
                  T* __restrict__ pk, KSlots ks, cplx<TI>* __restrict__ q, int64_t ldq, int rs,
                  int offm, int tiles_per_batch, int64_t a_bstride, int64_t pk_bstride,
                  int64_t q_bstride) {
  using C = cplx<T>;
  using CI = cplx<TI>;
  constexpr int E = FftCfg<L>::E;
  constexpr int TC = px_tc<T>();
  constexpr int TCOL = L / E;
  constexpr int NT = MODE == kModeInterlaced ? 2 : 1;
  using Rad = typename FftCfg<L>::radices;
  using Lay = LaySwz<TC>;
  static_assert(RadixInfo<Rad>::first == E, "first radix must equal E");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  CI* ldsi = reinterpret_cast<CI*>(smem_raw);
  // Stage twiddles of the forward (T) and inverse (TI) plans, one copy per
  // workgroup in LDS (one table when T and TI agree: the same x length and
  // precision give the same values).  13.8 KB for 512 points in the mixed
  // form, so two workgroups still fit a CU beside their 64 KB tiles.
  constexpr int SL = stage_twiddle_len<L, 0>();
  constexpr bool kSameTw = std::is_same<T, TI>::value;
  __shared__ __align__(16) C twf[SL > 0 ? SL : 1];
  __shared__ __align__(16) CI two[(SL > 0 && !kSameTw) ? SL : 1];
  const CI* twl = kSameTw ? reinterpret_cast<const CI*>(twf) : two;

  const int tile_all = xcd_tile(blockIdx.x, gridDim.x);
  const int bat = tile_all / tiles_per_batch;
  const int tile = tile_all - bat * tiles_per_batch;
  A += bat * a_bstride;
  if (MODE == kModeInterlaced) B += bat * a_bstride;
  pk += bat * pk_bstride;
  q += bat * q_bstride;
  const int y = tile / tilesC;
  const int c0 = (tile - y * tilesC) * TC;
  const int ncol = min(TC, n2h - c0);
  const int c = threadIdx.x % TC, t = threadIdx.x / TC;
  const int cg = c0 + c;
  const bool colok = c < ncol;
  // Byte offsets stay below 4 GB (checked by the launcher): 32-bit address
  // arithmetic on a uniform base.
  const uint32_t sL = (uint32_t)n1 * (uint32_t)ldc * (uint32_t)sizeof(C);
  // Columns past n2h read column c0 (discarded): the loads are unconditional
  // and issue back to back; row offsets are base + s TCOL sL.
  const uint32_t base =
      ((uint32_t)y * (uint32_t)ldc + (uint32_t)(colok ? cg : c0)) * (uint32_t)sizeof(C) +
      (uint32_t)t * sL;
  const uint32_t stepL = opaque_uniform((uint32_t)TCOL * sL);
  const char* Ab = reinterpret_cast<const char*>(A);
  const char* Bb = reinterpret_cast<const char*>(B);

  C v[NT][E];
  C g[MODE == kModeCross ? E : 1];
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const uint32_t o = base + (uint32_t)s * stepL;
    // Plain loads: the two tiles of a 128-byte line (64 bytes each) are read
    // by neighbouring workgroups of one XCD, the second from its L2.
    v[0][s] = *reinterpret_cast<const C*>(Ab + o);
    if (NT == 2) v[NT - 1][s] = *reinterpret_cast<const C*>(Bb + o);
    if (MODE == kModeCross)
      g[s] = *reinterpret_cast<const C*>(Bb + base + (uint32_t)out_m<L>(s) * stepL);
  }
  if (!colok) {
#pragma unroll
    for (int s = 0; s < E; ++s) {
      v[0][s] = C{T(0), T(0)};
      if (NT == 2) v[NT - 1][s] = C{T(0), T(0)};
      if (MODE == kModeCross) g[s] = C{T(0), T(0)};
    }
  }
  // The twiddle copies wait for their loads, which come after (and so wait
  // for) the input loads the first stage needs anyway; the first stage's
  // closing barrier orders them before any stage reads them.
  for (int i = threadIdx.x; i < SL; i += blockDim.x) {
    twf[i] = tw[L + i];
    if constexpr (!kSameTw) two[i] = twi[L + i];
  }
  rstages_<T, L, -1, NT, 1, 0, true, 0, NT, Lay, C, 0, BarWg>(Rad{}, v, lds, L * TC, twf, c, t,
                                                             Lay{});

  // Combine: slot s holds element e = t + out_m(s) L/E of both spectra.
  double2 pyc = make_double2(1.0, 0.0);
  if (MODE == kModeInterlaced) pyc = cmul(ph1[y], ph2[colok ? cg : 0]);
  CI u[1][E];
  T pkv[E];
  x2_combine<T, TI, L, MODE, NT, (MODE == kModeCross ? E : 1)>(v, g, u, pkv, ph0, pyc, t);
  __syncthreads();  // the forward pass's last LDS reads precede the inverse's stores
  rstages_<TI, L, +1, 1, 1, 0, true, 0, 1, Lay, CI, 0, BarWg>(Rad{}, u, ldsi, 0, twl, c, t, Lay{});
  x2_store_pk<T, L>(pkv, pk, ks, n1, y, cg, colok, t);
  if (colok) x2_store_q<TI, L>(u, q, n1, y, ldq, cg, rs, offm, t);
}

// ---- persistent cross-mode fused x pass with direct-to-LDS prefetch ---------
// The kernel above runs two workgroups per CU, each loading its column tiles,
// transforming and combining them and storing the pruned inverse in turn; its
// load and compute phases overlap only when the two workgroups happen to be
// in different phases.  This form of the cross mode (FP32 spectra, 512-point
// x: the Yamamoto correlator's per-(l, m) pass) runs one 8-wave workgroup per
// CU over the tiles k G + xcd_tile(b, G) (G = gridDim.x, so each XCD takes a
// contiguous range of every round and the two 64-byte tiles of a 128-byte
// line meet in one L2).  While it works on one tile, the next tile's rows of
// both spectra stream into its second LDS buffer with direct-to-LDS loads;
// the stage barriers wait for LDS only and the stage twiddles sit in LDS, so
// no global load queues behind the prefetch.  Arithmetic, combine and stores
// are those of k_periodic_x2.  At 512^3: 0.410 ms per pass against 0.480; the
// interlaced (two forward transforms per tile: 0.519 against 0.489) and auto
// (0.323 against 0.321) modes keep the two-workgroup kernel.
template <typename TI, int L>
__global__ void __launch_bounds__(8 * (L / 8))
    k_periodic_x2_cross_pipe(const cplx<float>* __restrict__ A, const cplx<float>* __restrict__ B,
                             int n1, int n2h, int64_t ldc, int tilesC,
                             const cplx<float>* __restrict__ tw, const cplx<TI>* __restrict__ twi,
                             float* __restrict__ pk, KSlots ks, cplx<TI>* __restrict__ q,
                             int64_t ldq, int rs, int offm, int ntiles) {
  using T = float;
  using C = cplx<T>;
  using CI = cplx<TI>;
  constexpr int E = FftCfg<L>::E;
  constexpr int TC = px_tc<T>();
  constexpr int TCOL = L / E;
  constexpr int SL = stage_twiddle_len<L, 0>();
  constexpr size_t RAW = 2 * (size_t)L * TC * sizeof(C);  // rows of A, then of B
  constexpr size_t INV = (size_t)L * TC * sizeof(CI);
  constexpr size_t BUF = RAW > INV ? RAW : INV;
  using Rad = typename FftCfg<L>::radices;
  using Lay = LaySwz<TC>;
  static_assert(E == 8 && TC * sizeof(C) == 64, "64-byte rows of 8 FP32 columns");
  __shared__ __align__(16) unsigned char bufA[BUF];
  __shared__ __align__(16) unsigned char bufB[BUF];
  __shared__ __align__(16) C twf[SL];
  __shared__ __align__(16) CI twl[SL];

  const int c = threadIdx.x % TC, t = threadIdx.x / TC;
  const int G = gridDim.x;
  const uint32_t sLb = (uint32_t)n1 * (uint32_t)ldc * (uint32_t)sizeof(C);  // x pitch
  auto fetch = [&](int tile, unsigned char* buf) {
    const int y = tile / tilesC;
    const int64_t off = (int64_t)y * ldc + (tile - y * tilesC) * TC;
    lds_fetch_rows<L, 64, TC * TCOL>(reinterpret_cast<const char*>(A + off), sLb, buf);
    lds_fetch_rows<L, 64, TC * TCOL>(reinterpret_cast<const char*>(B + off), sLb,
                                     buf + RAW / 2);
  };

  int tt = xcd_tile(blockIdx.x, G);
  if (tt >= ntiles) return;
  fetch(tt, bufA);
  for (int i = threadIdx.x; i < SL; i += blockDim.x) {
    twf[i] = tw[L + i];
    twl[i] = twi[L + i];
  }
  wait_vm_all();
  __syncthreads();
  if (tt + G < ntiles) fetch(tt + G, bufB);

  // One tile from cur; afterwards cur receives tile tt + 2 G.
  auto body = [&](unsigned char* cur) {
    const int y = tt / tilesC;
    const int cg = (tt - y * tilesC) * TC + c;
    const bool colok = cg < n2h;
    const C* rawA = reinterpret_cast<const C*>(cur);
    const C* rawB = rawA + L * TC;
    C v[1][E];
    C g[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
      v[0][s] = rawA[(t + s * TCOL) * TC + c];
      g[s] = rawB[(t + out_m<L>(s) * TCOL) * TC + c];
    }
    BarLds::sync();  // the raw tiles are in registers: cur takes the exchanges
    rstages_<T, L, -1, 1, 1, 0, true, 0, 1, Lay, C, 0, BarLds>(
        Rad{}, v, reinterpret_cast<C*>(cur), 0, twf, c, t, Lay{});
    CI u[1][E];
    T pkv[E];
    x2_combine<T, TI, L, kModeCross, 1, E>(v, g, u, pkv, nullptr, make_double2(1.0, 0.0), t);
    BarLds::sync();  // the forward pass's last LDS reads precede the inverse's stores
    rstages_<TI, L, +1, 1, 1, 0, true, 0, 1, Lay, CI, 0, BarLds>(
        Rad{}, u, reinterpret_cast<CI*>(cur), 0, twl, c, t, Lay{});
    wait_vm_all();   // the prefetch of tile tt + G and earlier stores are done
    BarLds::sync();  // in every wave; and every wave has read cur for the last time
    if (tt + 2 * G < ntiles) fetch(tt + 2 * G, cur);
    x2_store_pk<T, L>(pkv, pk, ks, n1, y, cg, colok, t);
    if (colok) x2_store_q<TI, L>(u, q, n1, y, ldq, cg, rs, offm, t);
  };
  for (;;) {
    body(bufA);
    tt += G;
    if (tt >= ntiles) break;
    body(bufB);
    tt += G;
    if (tt >= ntiles) break;
  }
}

}  // namespace fc
