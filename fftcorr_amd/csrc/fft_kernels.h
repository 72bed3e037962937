// LDS-staged Stockham FFT kernels for the 3D R2C / C2R transforms of the
// padded work grid [n0][n1][n2/2+1] (the layout of cc/grid/fft_grid.cc:10-50).
//
//  * z pass (rows, contiguous): a real row of n2 samples is packed as n2/2
//    complex values, transformed with a half-length complex FFT and
//    post-processed into the n2/2+1 Hermitian half (and the reverse for C2R).
//    Rows are loaded coalesced; the loader functor fuses the grid copy
//    (array_ops::copy_into_padded_array) and, for the Yamamoto correlator,
//    the Y_lm(x) weighting of the density into this load.
//  * y and x passes (columns, strided): a workgroup owns TC adjacent
//    columns (contiguous along z), so each global access is a run of TC
//    complex values; the TC transforms run in LDS with E complex values per
//    thread in registers between stages.
//
// Lengths 2^a (8..4096) and 3*2^a (24..3072) use compile-time radix plans;
// every other length falls back to a direct O(L^2) DFT per column/row (used
// for small or unusual test grids).  Forward transforms use sign -1 and no
// normalisation, the inverse sign +1 and no 1/N, matching FFTW's r2c/c2r as
// called by the reference (cc/grid/fft_grid.cc:81-84).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "xcd.h"

namespace fc {

template <typename T>
struct CplxOf;
template <>
struct CplxOf<float> {
  using type = float2;
};
template <>
struct CplxOf<double> {
  using type = double2;
};
template <typename T>
using cplx = typename CplxOf<T>::type;

template <typename C>
__host__ __device__ __forceinline__ C cadd(C a, C b) {
  return C{a.x + b.x, a.y + b.y};
}
template <typename C>
__host__ __device__ __forceinline__ C csub(C a, C b) {
  return C{a.x - b.x, a.y - b.y};
}
template <typename C>
__host__ __device__ __forceinline__ C cmul(C a, C b) {
  return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// Complex value in another precision (the storage type of a pass whose
// arithmetic runs in a wider type).
template <typename To, typename C>
__device__ __forceinline__ cplx<To> to_cplx(C a) {
  return cplx<To>{To(a.x), To(a.y)};
}
// Streaming global accesses for the whole-grid passes: every element is read
// once and written once per pass and the grids are far larger than the
// caches, so these loads and stores carry the nontemporal hint.
template <typename V>
struct NtVec;
template <>
struct NtVec<float2> {
  typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct NtVec<double2> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <typename V>
__device__ __forceinline__ V ld_stream(const V* p) {
  const auto e = __builtin_nontemporal_load(reinterpret_cast<const typename NtVec<V>::type*>(p));
  V v;
  v.x = e.x;
  v.y = e.y;
  return v;
}
template <typename V>
__device__ __forceinline__ void st_stream(V* p, V v) {
  typename NtVec<V>::type e;
  e.x = v.x;
  e.y = v.y;
  __builtin_nontemporal_store(e, reinterpret_cast<typename NtVec<V>::type*>(p));
}

template <typename C>
struct ScalarOf;
template <>
struct ScalarOf<float2> {
  using type = float;
};
template <>
struct ScalarOf<double2> {
  using type = double;
};
// a * (DIR * i): DIR = -1 multiplies by -i, DIR = +1 by +i.
template <int DIR, typename C>
__host__ __device__ __forceinline__ C mul_i(C a) {
  return DIR < 0 ? C{a.y, -a.x} : C{-a.y, a.x};
}

// ---- register butterflies (radix R, sign DIR) -----------------------------
template <int R, int DIR, typename T>
struct Bfly;

template <int DIR, typename T>
struct Bfly<2, DIR, T> {
  using C = cplx<T>;
  static __host__ __device__ __forceinline__ void run(C* v) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int DIR, typename T>
struct Bfly<4, DIR, T> {
  using C = cplx<T>;
  static __host__ __device__ __forceinline__ void run(C* v) {
    C t0 = cadd(v[0], v[2]);
    C t1 = csub(v[0], v[2]);
    C t2 = cadd(v[1], v[3]);
    C t3 = mul_i<DIR>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

// x * c for an irrational constant c.  In FP32 the constant is split as
// c = hi + lo (hi = c rounded to float) and the product formed as
// fma(hi, x, lo * x): the rounding of c alone would scale every product by
// the same relative error (-1.7e-8 for 1/sqrt(2)), a bias that does not
// average out in the multipole sums; with the split only the final rounding
// remains.  FP64 multiplies directly.
template <typename T>
struct ConstMul;
template <>
struct ConstMul<double> {
  double c;
  __host__ __device__ constexpr ConstMul(long double v) : c((double)v) {}
  __host__ __device__ __forceinline__ double operator()(double x) const { return c * x; }
};
template <>
struct ConstMul<float> {
  float hi, lo;
  __host__ __device__ constexpr ConstMul(long double v)
      : hi((float)v), lo((float)(v - (long double)(float)v)) {}
  __host__ __device__ __forceinline__ float operator()(float x) const {
    return __builtin_fmaf(hi, x, lo * x);
  }
};

template <int DIR, typename T>
struct Bfly<8, DIR, T> {
  using C = cplx<T>;
  static __host__ __device__ __forceinline__ void run(C* v) {
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    Bfly<4, DIR, T>::run(e);
    Bfly<4, DIR, T>::run(o);
    constexpr ConstMul<T> h(0.70710678118654752440084436210484903928L);
    // W8^1 = h (1 + DIR i), W8^2 = DIR i, W8^3 = h (-1 + DIR i)
    C o1 = C{h(o[1].x - DIR * o[1].y), h(o[1].y + DIR * o[1].x)};
    C o2 = mul_i<DIR>(o[2]);
    C o3 = C{h(-o[3].x - DIR * o[3].y), h(-o[3].y + DIR * o[3].x)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

// Radix 16 as 4 x 4: with n = 4 n1 + n2 and k = k1 + 4 k2,
// X[k1 + 4 k2] = sum_n2 W4^(n2 k2) W16^(n2 k1) sum_n1 W4^(n1 k1) x[4 n1 + n2]:
// four radix-4 butterflies over n1, the inner twiddles W16^(n2 k1), then four
// radix-4 butterflies over n2.
template <int DIR, typename T>
struct Bfly<16, DIR, T> {
  using C = cplx<T>;
  // x * (a + DIR i b) for the constant a + DIR i b.
  static __host__ __device__ __forceinline__ C rot(C x, ConstMul<T> a, ConstMul<T> b) {
    return C{a(x.x) - DIR * b(x.y), a(x.y) + DIR * b(x.x)};
  }
  static __host__ __device__ __forceinline__ void run(C* v) {
    constexpr ConstMul<T> h(0.70710678118654752440084436210484903928L);
    constexpr ConstMul<T> c1(0.92387953251128675612818318939678828682L);  // cos(pi/8)
    constexpr ConstMul<T> s1(0.38268343236508977172845998403039886676L);  // sin(pi/8)
    C y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      C a[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
      Bfly<4, DIR, T>::run(a);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = a[k1];
    }
    // W16^1, ^2, ^3 on row 1; ^2, ^4, ^6 on row 2; ^3, ^6, ^9 on row 3.
    y[1][1] = rot(y[1][1], c1, s1);
    y[1][2] = C{h(y[1][2].x - DIR * y[1][2].y), h(y[1][2].y + DIR * y[1][2].x)};
    y[1][3] = rot(y[1][3], s1, c1);
    y[2][1] = C{h(y[2][1].x - DIR * y[2][1].y), h(y[2][1].y + DIR * y[2][1].x)};
    y[2][2] = mul_i<DIR>(y[2][2]);
    y[2][3] = C{h(-y[2][3].x - DIR * y[2][3].y), h(-y[2][3].y + DIR * y[2][3].x)};
    y[3][1] = rot(y[3][1], s1, c1);
    y[3][2] = C{h(-y[3][2].x - DIR * y[3][2].y), h(-y[3][2].y + DIR * y[3][2].x)};
    {
      const C r = rot(y[3][3], c1, s1);  // W16^9 = -W16^1
      y[3][3] = C{-r.x, -r.y};
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      C b[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
      Bfly<4, DIR, T>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
    }
  }
};

template <int DIR, typename T>
struct Bfly<3, DIR, T> {
  using C = cplx<T>;
  static __host__ __device__ __forceinline__ void run(C* v) {
    constexpr ConstMul<T> s3(0.86602540378443864676372317075293618347L);  // sqrt(3)/2
    C t = cadd(v[1], v[2]);
    C d = mul_i<DIR>(csub(v[1], v[2]));
    C m = C{v[0].x - T(0.5) * t.x, v[0].y - T(0.5) * t.y};
    v[0] = cadd(v[0], t);
    const T sx = s3(d.x), sy = s3(d.y);
    v[1] = C{m.x + sx, m.y + sy};
    v[2] = C{m.x - sx, m.y - sy};
  }
};

template <int... Rs>
struct Radices {};

// Compile-time plans: E = complex values per thread (divisible by every
// radix), so a length-L column uses L/E threads.
template <int L>
struct FftCfg {
  static constexpr bool fast = false;
};
#define FC_FFT_CFG(Lv, Ev, ...)                 \
  template <>                                   \
  struct FftCfg<Lv> {                           \
    static constexpr bool fast = true;          \
    static constexpr int E = Ev;                \
    using radices = Radices<__VA_ARGS__>;       \
  };
FC_FFT_CFG(8, 8, 8)
FC_FFT_CFG(16, 8, 8, 2)
FC_FFT_CFG(32, 8, 8, 4)
FC_FFT_CFG(64, 8, 8, 8)
FC_FFT_CFG(128, 8, 8, 8, 2)
FC_FFT_CFG(256, 8, 8, 8, 4)
FC_FFT_CFG(512, 8, 8, 8, 8)
FC_FFT_CFG(1024, 8, 8, 8, 8, 2)
FC_FFT_CFG(2048, 8, 8, 8, 8, 4)
FC_FFT_CFG(4096, 8, 8, 8, 8, 8)
FC_FFT_CFG(24, 24, 8, 3)
FC_FFT_CFG(48, 24, 8, 2, 3)
FC_FFT_CFG(96, 24, 8, 4, 3)
FC_FFT_CFG(192, 24, 8, 8, 3)
FC_FFT_CFG(384, 24, 8, 8, 2, 3)
FC_FFT_CFG(768, 24, 8, 8, 4, 3)
FC_FFT_CFG(1536, 24, 8, 8, 8, 3)
FC_FFT_CFG(3072, 24, 8, 8, 8, 2, 3)
#undef FC_FFT_CFG

// Plan variants of the register-chained kernels, chosen per pass by the
// autotuner: 0 is the table above (first radix 8); 1 starts with radix 16
// (16 values per thread, half the threads per column or row, and for
// 256 = 16 x 16 one LDS exchange instead of two).
constexpr int kFftVariants = 2;
template <int L, int V>
struct FftCfgV : FftCfg<L> {};
template <int L>
struct FftCfgV<L, 1> {
  static constexpr bool fast = false;
};
#define FC_FFT_CFG16(Lv, ...)               \
  template <>                               \
  struct FftCfgV<Lv, 1> {                   \
    static constexpr bool fast = true;      \
    static constexpr int E = 16;            \
    using radices = Radices<__VA_ARGS__>;   \
  };
FC_FFT_CFG16(128, 16, 8)
FC_FFT_CFG16(256, 16, 16)
FC_FFT_CFG16(512, 16, 16, 2)
#undef FC_FFT_CFG16

// LDS addressing: element e of column c.
struct LayColFast {  // [e][TC]: columns contiguous (strided y/x passes)
  int tc;
  __device__ __forceinline__ int operator()(int c, int e) const { return e * tc + c; }
};
struct LayRow {  // [c][rs]: elements contiguous (row / z pass)
  int rs;
  __device__ __forceinline__ int operator()(int c, int e) const { return c * rs + e; }
};

// One Stockham stage of radix R on LDS column c, thread t of L/E.
// Twiddle table tw[m] = exp(-2 pi i m / L), m < L.
template <typename T, int L, int E, int R, int NS, int DIR, class Lay>
__device__ __forceinline__ void stockham_stage(cplx<T>* lds, const cplx<T>* __restrict__ tw,
                                               int c, int t, Lay lay) {
  using C = cplx<T>;
  constexpr int TCOL = L / E;
  constexpr int NB = E / R;
  constexpr int LR = L / R;
  C v[E];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * TCOL;
#pragma unroll
    for (int r = 0; r < R; ++r) v[b * R + r] = lds[lay(c, j + r * LR)];
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * TCOL;
    const int k = j % NS;
    if (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        C w = tw[r * k * (L / (NS * R))];
        if (DIR > 0) w.y = -w.y;
        v[b * R + r] = cmul(v[b * R + r], w);
      }
    }
    Bfly<R, DIR, T>::run(&v[b * R]);
    const int base = (j / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) lds[lay(c, base + r * NS)] = v[b * R + r];
  }
  __syncthreads();
}

template <typename T, int L, int E, int DIR, int NS, class Lay>
__device__ __forceinline__ void run_stages(Radices<>, cplx<T>*, const cplx<T>*, int, int, Lay) {}

template <typename T, int L, int E, int DIR, int NS, class Lay, int R, int... Rest>
__device__ __forceinline__ void run_stages(Radices<R, Rest...>, cplx<T>* lds,
                                           const cplx<T>* __restrict__ tw, int c, int t,
                                           Lay lay) {
  stockham_stage<T, L, E, R, NS, DIR>(lds, tw, c, t, lay);
  run_stages<T, L, E, DIR, NS * R>(Radices<Rest...>{}, lds, tw, c, t, lay);
}

template <typename T, int L, int DIR, class Lay>
__device__ __forceinline__ void fft_lds(cplx<T>* lds, const cplx<T>* __restrict__ tw, int c,
                                        int t, Lay lay) {
  run_stages<T, L, FftCfg<L>::E, DIR, 1>(typename FftCfg<L>::radices{}, lds, tw, c, t, lay);
}

// ---- register-chained Stockham passes -------------------------------------
// For power-of-two lengths (first radix = E) thread t keeps E values of its
// column or row in registers: before the first stage slot s holds element
// t + s L/E, after the last stage (radix Rl) slot s holds element
// t + out_m(s) L/E.  The first stage therefore reads global memory directly
// and the last stage's outputs go straight to global memory (or to the next
// fused step): two LDS passes fewer than the LDS-staged form above.
template <class R>
struct RadixInfo;
template <int R>
struct RadixInfo<Radices<R>> {
  static constexpr int first = R, last = R;
};
template <int R, int R2, int... Rest>
struct RadixInfo<Radices<R, R2, Rest...>> {
  static constexpr int first = R, last = RadixInfo<Radices<R2, Rest...>>::last;
};

__host__ __device__ constexpr bool rc_len_ok(int L) {
  return L >= 8 && L <= 512 && (L & (L - 1)) == 0;
}

template <int L, int V = 0>
__host__ __device__ constexpr int out_m(int s) {
  constexpr int E = FftCfgV<L, V>::E;
  constexpr int Rl = RadixInfo<typename FftCfgV<L, V>::radices>::last;
  return s / Rl + (s % Rl) * (E / Rl);
}

// Column tiles: TC columns of length L, element (e, c) at row
// e ^ ((e / E) & 1) (E = values per thread, a power of two): rows Ea..Ea+E-1
// swap neighbours when bit log2(E) of the row is set, so the first-stage
// stores of threads t and t + 1 (rows Et + r) fall in different LDS banks
// when a row is narrower than the 128 bytes of banks.
template <int TC, int E = 8>
struct LaySwz {
  __device__ __forceinline__ int operator()(int c, int e) const {
    return (e ^ ((e / E) & 1)) * TC + c;
  }
};
// Row tiles: row c of rs slots, element e at e + e / 8 (one pad slot per 8),
// which spreads the stride-8 first-stage stores over all banks.
struct LayRowPad {
  int rs;
  __device__ __forceinline__ int operator()(int c, int e) const { return c * rs + e + (e >> 3); }
};
__host__ __device__ constexpr int row_pad_len(int M) { return M + M / 8 + 1; }

// One Stockham stage on NT tiles in lock-step (one twiddle load serves every
// tile; tile q starts ts elements after tile q - 1).  Arithmetic in T; the
// LDS tiles hold CL values (T, or the narrower storage type of a pass whose
// arithmetic is wider: each exchange then rounds once).  FIRST: the inputs are
// already in v (no LDS read).  LAST: the outputs stay in v (no LDS write).  A
// stage that reads and then writes LDS synchronises in between; a stage that
// writes LDS ends with a barrier.  A FIRST stage writes without a barrier
// before it: the caller makes sure no thread still reads the LDS it writes.
//
// Twiddles come from the per-stage table that follows the L plain twiddles
// of the axis (see stage_twiddle_len): the stage with prefix NS owns NS * R
// entries at OFF, entry k * R + r = exp(-2 pi i r k / (NS R)), so a thread's
// R - 1 twiddles of one stage are contiguous.
// Barrier policies of the register-chained stages.  BarWg: __syncthreads().
// BarLds: waits for this wave's LDS accesses only and joins the workgroup,
// for kernels that keep direct-to-LDS prefetches of other LDS in flight
// across the stages (and wait for those explicitly).
struct BarWg {
  __device__ __forceinline__ static void sync() { __syncthreads(); }
};
struct BarLds {
  __device__ __forceinline__ static void sync() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  }
};

template <typename T, int L, int R, int NS, int OFF, int DIR, int NT, bool FIRST, bool LAST,
          class Lay, int Q0, int NQ, class CL, int V, class Bar = BarWg>
__device__ __forceinline__ void rstage(cplx<T> (&v)[NT][(FftCfgV<L, V>::E)], CL* lds, int ts,
                                       const cplx<T>* __restrict__ tws, int c, int t, Lay lay) {
  using TL = typename ScalarOf<CL>::type;
  using C = cplx<T>;
  constexpr int E = FftCfgV<L, V>::E;
  constexpr int TCOL = L / E;
  constexpr int NB = E / R;
  constexpr int LR = L / R;
  // The stage's twiddles are loaded first, so their latency overlaps the
  // LDS reads and the barrier below (FP64: w only; FP32: w^1 .. w^(R-1)).
  constexpr int NTW = NS > 1 ? (sizeof(T) == 8 ? 1 : R - 1) : 1;
  C twv[NB][NTW];
  if (NS > 1) {
    // 32-bit byte offsets from the table's uniform base.
    const char* twb = reinterpret_cast<const char*>(tws);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const uint32_t k0 = (uint32_t)(OFF + ((t + b * TCOL) % NS) * R);
#pragma unroll
      for (int r = 0; r < NTW; ++r)
        twv[b][r] = *reinterpret_cast<const C*>(twb + (k0 + (uint32_t)(r + 1)) * (uint32_t)sizeof(C));
    }
  }
  if (!FIRST) {
#pragma unroll
    for (int q = Q0; q < Q0 + NQ; ++q)
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r)
          v[q][b * R + r] = to_cplx<T>(lds[(q - Q0) * ts + lay(c, t + b * TCOL + r * LR)]);
    if (!LAST) Bar::sync();
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * TCOL;
    const int k = j % NS;
    if (NS > 1) {
      if constexpr (sizeof(T) == 8) {
        // FP64: entry r is w^r, w = entry 1; the powers are formed in
        // registers (error ~ r ulp of FP64), which keeps one twiddle live
        // instead of R - 1.
        C w1 = twv[b][0];
        if (DIR > 0) w1.y = -w1.y;
        C w = w1;
#pragma unroll
        for (int r = 1; r < R; ++r) {
#pragma unroll
          for (int q = Q0; q < Q0 + NQ; ++q) v[q][b * R + r] = cmul(v[q][b * R + r], w);
          if (r + 1 < R) w = cmul(w, w1);
        }
      } else {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          C w = twv[b][r - 1];
          if (DIR > 0) w.y = -w.y;
#pragma unroll
          for (int q = Q0; q < Q0 + NQ; ++q) v[q][b * R + r] = cmul(v[q][b * R + r], w);
        }
      }
    }
#pragma unroll
    for (int q = Q0; q < Q0 + NQ; ++q) Bfly<R, DIR, T>::run(&v[q][b * R]);
    if (!LAST) {
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int q = Q0; q < Q0 + NQ; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r)
          lds[(q - Q0) * ts + lay(c, base + r * NS)] = to_cplx<TL>(v[q][b * R + r]);
    }
  }
  if (!LAST) Bar::sync();
}

template <typename T, int L, int DIR, int NT, int NS, int OFF, bool FIRST, int Q0, int NQ,
          class Lay, class CL, int V, class Bar>
__device__ __forceinline__ void rstages_(Radices<>, cplx<T> (&)[NT][(FftCfgV<L, V>::E)], CL*,
                                         int, const cplx<T>*, int, int, Lay) {}

template <typename T, int L, int DIR, int NT, int NS, int OFF, bool FIRST, int Q0, int NQ,
          class Lay, class CL, int V, class Bar, int R, int... Rest>
__device__ __forceinline__ void rstages_(Radices<R, Rest...>,
                                         cplx<T> (&v)[NT][(FftCfgV<L, V>::E)], CL* lds, int ts,
                                         const cplx<T>* __restrict__ tws, int c, int t,
                                         Lay lay) {
  rstage<T, L, R, NS, OFF, DIR, NT, FIRST, sizeof...(Rest) == 0, Lay, Q0, NQ, CL, V, Bar>(
      v, lds, ts, tws, c, t, lay);
  rstages_<T, L, DIR, NT, NS * R, (NS > 1 ? OFF + NS * R : OFF), false, Q0, NQ, Lay, CL, V, Bar>(
      Radices<Rest...>{}, v, lds, ts, tws, c, t, lay);
}

// Host: length of the per-stage table of an L-point register-chained plan,
// and its entries as indices m into the plain table exp(-2 pi i m / L).
template <int L, int NS = 1, int R, int... Rest>
constexpr int stage_twiddle_len_(Radices<R, Rest...>) {
  if constexpr (sizeof...(Rest) == 0)
    return NS > 1 ? NS * R : 0;
  else
    return (NS > 1 ? NS * R : 0) + stage_twiddle_len_<L, NS * R>(Radices<Rest...>{});
}
template <int L, int V = 0>
constexpr int stage_twiddle_len() {
  if constexpr (FftCfgV<L, V>::fast)
    return stage_twiddle_len_<L>(typename FftCfgV<L, V>::radices{});
  else
    return 0;
}
// An axis table holds the L plain twiddles, then the stage tables of
// variant 0, 1, ... in turn; variant V's starts stage_twiddle_base<L, V>()
// entries after the plain part.
template <int L, int V>
constexpr int stage_twiddle_base() {
  if constexpr (V == 0)
    return 0;
  else
    return stage_twiddle_base<L, V - 1>() + stage_twiddle_len<L, V - 1>();
}

// All stages of plan variant V; tw is the axis's twiddle table.  Tiles
// Q0 .. Q0 + NQ - 1 of v run in lock-step (LDS tile q - Q0).
template <typename T, int L, int DIR, int NT, int NS, bool FIRST, int V, class Lay, int Q0 = 0,
          int NQ = NT, class Bar = BarWg, class CL>
__device__ __forceinline__ void rstages(cplx<T> (&v)[NT][(FftCfgV<L, V>::E)], CL* lds, int ts,
                                        const cplx<T>* __restrict__ tw, int c, int t, Lay lay) {
  rstages_<T, L, DIR, NT, NS, 0, FIRST, Q0, NQ, Lay, CL, V, Bar>(
      typename FftCfgV<L, V>::radices{}, v, lds, ts, tw + L + stage_twiddle_base<L, V>(), c, t,
      lay);
}
template <int L, int NS = 1, int OFF = 0, int R, int... Rest, class Put>
void stage_twiddles_(Radices<R, Rest...>, Put put) {
  if (NS > 1)
    for (int k = 0; k < NS; ++k)
      for (int r = 0; r < R; ++r) put(OFF + k * R + r, r * k * (L / (NS * R)));
  if constexpr (sizeof...(Rest) > 0)
    stage_twiddles_<L, NS * R, (NS > 1 ? OFF + NS * R : OFF)>(Radices<Rest...>{}, put);
}


// Register-chained strided column pass over [A][L][G][B] (see k_fft_cols),
// TC columns per workgroup, one launch of A * G * ceil(B / TC) workgroups.
// Output o < nout holds element (o - off) mod L (offm = off mod L; nout may
// exceed L).  In place when in == out: every workgroup reads its whole tile
// before writing it.  Arithmetic in T; loads Tin; LDS and stores in To.
// Waves per SIMD a register-chained column kernel is held to when its
// arithmetic is wider than its LDS tiles (the occupancy of the narrow form).
// Variant 1 holds twice the values per thread and is sized for half the waves.
template <typename T, typename To, int V = 0>
constexpr int rc_wide_waves() {
  return sizeof(T) > sizeof(To) ? (V == 1 ? 4 : 8) : 1;
}

template <typename T, int L, int DIR, int TC, typename Tin = T, typename To = T, int V = 0>
__global__ void __launch_bounds__(TC*(L / FftCfgV<L, V>::E))
__attribute__((amdgpu_waves_per_eu(rc_wide_waves<T, To, V>())))
    k_fft_cols_rc(const cplx<Tin>* in, int64_t inA, int64_t inL, int64_t inG, cplx<To>* out,
                  int64_t outA, int64_t outL, int64_t outG, int B, int tilesB, int G, int nout,
                  int offm, const cplx<T>* __restrict__ tw) {
  using C = cplx<T>;
  constexpr int E = FftCfgV<L, V>::E;
  constexpr int TCOL = L / E;
  using Rad = typename FftCfgV<L, V>::radices;
  static_assert(RadixInfo<Rad>::first == E, "first radix must equal E");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<To>* lds = reinterpret_cast<cplx<To>*>(smem_raw);
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int ag = tile / tilesB;
  const int b0 = (tile - ag * tilesB) * TC;
  const int a = ag / G, g = ag - a * G;
  const int c = threadIdx.x % TC, t = threadIdx.x / TC;
  const bool colok = b0 + c < B;
  // Uniform bases and 32-bit per-thread byte offsets (the launcher checks
  // that a tile's offsets stay below 4 GB).
  const char* srcb =
      reinterpret_cast<const char*>(in + (int64_t)a * inA + (int64_t)g * inG + b0);
  char* dstb = reinterpret_cast<char*>(out + (int64_t)a * outA + (int64_t)g * outG + b0);
  const uint32_t inLb = (uint32_t)inL * (uint32_t)sizeof(cplx<Tin>);
  const uint32_t outLb = (uint32_t)outL * (uint32_t)sizeof(cplx<To>);
  const uint32_t cin = (uint32_t)c * (uint32_t)sizeof(cplx<Tin>);
  const uint32_t cout = (uint32_t)c * (uint32_t)sizeof(cplx<To>);
  C v[1][E];
  {
    // Columns past B read column b0 (discarded), so the E loads are
    // unconditional and issue back to back before the first use.
    const uint32_t cinc = colok ? cin : 0u;
    cplx<Tin> xs[E];
#pragma unroll
    for (int s = 0; s < E; ++s)
      xs[s] = ld_stream(
          reinterpret_cast<const cplx<Tin>*>(srcb + (cinc + (uint32_t)(t + s * TCOL) * inLb)));
#pragma unroll
    for (int s = 0; s < E; ++s) v[0][s] = colok ? C{T(xs[s].x), T(xs[s].y)} : C{T(0), T(0)};
  }
  rstages<T, L, DIR, 1, 1, true, V>(v, lds, 0, tw, c, t, LaySwz<TC, E>{});
  if (!colok) return;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    int o = t + out_m<L, V>(s) * TCOL + offm;
    if (o >= L) o -= L;
    for (; o < nout; o += L)
      st_stream(reinterpret_cast<cplx<To>*>(dstb + (cout + (uint32_t)o * outLb)),
                to_cplx<To>(v[0][s]));
  }
}

// ---- persistent column pass with direct-to-LDS prefetch ---------------------
// The column pass above holds two workgroups (32 waves) per CU, each loading
// its tile, transforming it through LDS and storing it in turn; with both
// often in the same phase, HBM idles while they compute.  This form runs one
// workgroup per CU over a sequence of tiles: while it
// transforms one tile in one LDS buffer, the next streams into the other with
// global_load_lds_dwordx4 (no registers held), and the stage barriers wait
// only for LDS (BarLds) so the prefetch stays in flight across them.  Tiles
// are TC columns of 128-byte rows (16 FP32 / 8 FP64 complex); the stage
// twiddles sit in LDS, so no other global load is issued behind the prefetch.
// The stages and twiddles are those of k_fft_cols_rc; the compiler may fuse
// their multiplies and adds differently, so an FP32 output can differ from
// k_fft_cols_rc's in its last bit (2 of 268M values over a 512^3 pair).

__device__ __forceinline__ uint32_t lds_u32(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}

// 16 bytes per lane from gbase + lane_off (gbase uniform) into LDS at
// lds + 16 * lane.  M0 (the LDS base of the transfer) is saved and restored;
// the wave waits for the transfer with an explicit vmcnt(0).  No outputs.
__device__ __forceinline__ void lds_dma16(const char* gbase, uint32_t lane_off, uint32_t lds) {
  uint32_t saved;
  asm volatile(
      "s_mov_b32 %0, m0\n\t"
      "s_mov_b32 m0, %1\n\t"
      "s_nop 0\n\t"
      "global_load_lds_dwordx4 %2, %3\n\t"
      "s_nop 0\n\t"
      "s_mov_b32 m0, %0"
      : "=&s"(saved)
      : "s"(lds), "v"(lane_off), "s"(gbase)
      : "memory");
}

__device__ __forceinline__ void wait_vm_all() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// NROWS rows of ROWB bytes (16 <= ROWB <= 1024, a power of two) from src
// (row pitch `pitch` bytes; src uniform) into buf row-major, by the NTHR
// threads of the workgroup: block k of 1 KB holds rows k * 1024 / ROWB
// onward, one wave instruction each.
template <int NROWS, int ROWB, int NTHR>
__device__ __forceinline__ void lds_fetch_rows(const char* src, uint32_t pitch, void* buf) {
  constexpr int kRpb = 1024 / ROWB;  // rows per block
  constexpr int kLpr = ROWB / 16;    // lanes per row
  constexpr int kBlocks = NROWS / kRpb;
  constexpr int kWaves = NTHR / 64;
  static_assert(NROWS % kRpb == 0 && kBlocks % kWaves == 0, "whole blocks per wave");
  constexpr int kPer = kBlocks / kWaves;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x / 64);
  const int l = threadIdx.x % 64;
  const uint32_t lane_off = (uint32_t)(l / kLpr) * pitch + (uint32_t)(l % kLpr) * 16u;
  const uint32_t lb = lds_u32(buf);
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int blk = w * kPer + j;
    lds_dma16(src + (uint64_t)(blk * kRpb) * pitch, lane_off,
              __builtin_amdgcn_readfirstlane(lb + (uint32_t)blk * 1024u));
  }
}

// A column tile (L rows of 128 bytes) into buf.
template <int L, int TC>
__device__ __forceinline__ void pipe_fetch(const char* src, uint32_t rowb, void* buf) {
  lds_fetch_rows<L, 128, TC*(L / 8)>(src, rowb, buf);
}

template <int L, int TC, typename Ts>
constexpr size_t pipe_lds_bytes() {
  return (size_t)L * TC * sizeof(cplx<Ts>);
}

// One tile: its raw rows are in cur; after the transform cur is refilled
// with tile tt + 2 G (G = gridDim.x; the other buffer already holds tile
// tt + G).
template <typename T, int L, int DIR, int TC, typename Ts>
__device__ __forceinline__ void pipe_tile(const cplx<Ts>* in, int64_t inA, uint32_t inLb,
                                          cplx<Ts>* out, int64_t outA, uint32_t outLb, int B,
                                          int tilesB, int ntiles, int tt, const cplx<T>* twl,
                                          cplx<Ts>* cur) {
  using C = cplx<T>;
  constexpr int E = FftCfgV<L, 0>::E;
  constexpr int TCOL = L / E;
  const int c = threadIdx.x % TC, t = threadIdx.x / TC;
  C v[1][E];
#pragma unroll
  for (int s = 0; s < E; ++s) v[0][s] = to_cplx<T>(cur[(t + s * TCOL) * TC + c]);
  BarLds::sync();  // the raw tile is in registers: cur takes the exchanges
  rstages<T, L, DIR, 1, 1, true, 0, LaySwz<TC, E>, 0, 1, BarLds>(v, cur, 0, twl, c, t,
                                                                 LaySwz<TC, E>{});
  wait_vm_all();   // this wave's prefetch of tile tt + G (and its earlier stores) done
  BarLds::sync();  // every wave's; and every wave has read cur for the last time
  const int nt = tt + 2 * (int)gridDim.x;
  if (nt < ntiles) {
    const int a = nt / tilesB;
    const int b0 = (nt - a * tilesB) * TC;
    pipe_fetch<L, TC>(reinterpret_cast<const char*>(in + (int64_t)a * inA + b0), inLb, cur);
  }
  const int a = tt / tilesB;
  const int b0 = (tt - a * tilesB) * TC;
  if (b0 + c < B) {
    char* dstb = reinterpret_cast<char*>(out + (int64_t)a * outA + b0);
    const uint32_t lane_off = (uint32_t)c * (uint32_t)sizeof(cplx<Ts>) + (uint32_t)t * outLb;
#pragma unroll
    for (int s = 0; s < E; ++s) {
      char* rowp = dstb + (uint64_t)(out_m<L, 0>(s) * TCOL) * outLb;  // uniform
      st_stream(reinterpret_cast<cplx<Ts>*>(rowp + lane_off), to_cplx<Ts>(v[0][s]));
    }
  }
}

// Column pass over [A][L][B] (slab stride inA / outA, row pitch inL / outL
// complex elements of Ts), all L outputs, plan variant 0.  Launch
// min(ntiles, resident workgroups) workgroups of TC * L / 8 threads;
// workgroup b takes tiles b, b + G, ... (G = gridDim.x), so the chip works
// through the grid slab by slab (a contiguous range per workgroup measured
// 10 % slower).  tw is the axis table (plain part, then the stage table).
// The launcher checks that each tile's 128-byte rows lie inside the rows
// (tilesB * TC <= inL).
template <typename T, int L, int DIR, int TC, typename Ts>
__global__ void __launch_bounds__(TC*(L / 8))
    k_fft_cols_pipe(const cplx<Ts>* in, int64_t inA, int64_t inL, cplx<Ts>* out, int64_t outA,
                    int64_t outL, int B, int tilesB, int ntiles, const cplx<T>* __restrict__ tw) {
  static_assert(FftCfgV<L, 0>::fast && FftCfgV<L, 0>::E == 8, "register-chained plan of E = 8");
  static_assert(TC * sizeof(cplx<Ts>) == 128, "tiles of 128-byte rows");
  constexpr int TWN = L + stage_twiddle_len<L, 0>();
  __shared__ __align__(16) cplx<Ts> bufA[L * TC];
  __shared__ __align__(16) cplx<Ts> bufB[L * TC];
  __shared__ __align__(16) cplx<T> twl[TWN];
  const uint32_t inLb = (uint32_t)inL * (uint32_t)sizeof(cplx<Ts>);
  const uint32_t outLb = (uint32_t)outL * (uint32_t)sizeof(cplx<Ts>);
  const int G = gridDim.x;
  int tt = blockIdx.x;
  if (tt >= ntiles) return;
  auto fetch = [&](int k, cplx<Ts>* buf) {
    const int a = k / tilesB;
    const int b0 = (k - a * tilesB) * TC;
    pipe_fetch<L, TC>(reinterpret_cast<const char*>(in + (int64_t)a * inA + b0), inLb, buf);
  };
  fetch(tt, bufA);
  for (int i = L + (int)threadIdx.x; i < TWN; i += blockDim.x) twl[i] = tw[i];
  wait_vm_all();
  __syncthreads();
  if (tt + G < ntiles) fetch(tt + G, bufB);
  for (;;) {
    pipe_tile<T, L, DIR, TC, Ts>(in, inA, inLb, out, outA, outLb, B, tilesB, ntiles, tt, twl,
                                 bufA);
    tt += G;
    if (tt >= ntiles) break;
    pipe_tile<T, L, DIR, TC, Ts>(in, inA, inLb, out, outA, outLb, B, tilesB, ntiles, tt, twl,
                                 bufB);
    tt += G;
    if (tt >= ntiles) break;
  }
}

// Whether a column view's per-tile byte offsets fit the 32-bit offsets of
// k_fft_cols_rc.
template <typename To, typename Tin>
bool rc_offsets_fit(int L, int tc, int64_t inL, int64_t outL, int nout) {
  const uint64_t lim = 0xffffffffull;
  const uint64_t in_span = ((uint64_t)tc + (uint64_t)(L - 1) * (uint64_t)inL) * sizeof(cplx<Tin>);
  const uint64_t out_span =
      ((uint64_t)tc + (uint64_t)(nout - 1) * (uint64_t)outL) * sizeof(cplx<To>);
  return in_span < lim && out_span < lim;
}

// Loads the real sample pair (2e, 2e + 1) of a row as one vector access.
template <typename Tin>
struct PairOf;
template <>
struct PairOf<float> {
  using type = float2;
};
template <>
struct PairOf<double> {
  using type = double2;
};

// Register-chained R2C rows: a real row of 2M samples packed as M complex
// values, a length-M FFT, then the Hermitian post-process (see
// k_fft_rows_r2c).  Block = tr rows x M/E threads, LDS tr row_pad_len(M).
// Arithmetic in T; LDS and outputs in To.
// in2 (optional): rows from nrows1 on come from in2 (row - nrows1), so one
// launch transforms two grids into consecutive output rows.
template <typename T, typename Tin, int M, class Loader, typename To = T, int V = 0>
__global__ void __launch_bounds__(1024)
    k_fft_rows_r2c_rc(const Tin* in, int64_t ldi, cplx<To>* out, int64_t ldo, int64_t nrows,
                      int tr, const cplx<T>* __restrict__ twM, const cplx<T>* __restrict__ twN,
                      Loader ld, const Tin* in2 = nullptr, int64_t nrows1 = 0) {
  using C = cplx<T>;
  constexpr int E = FftCfgV<M, V>::E;
  constexpr int TCOL = M / E;
  constexpr int RS = row_pad_len(M);
  using Rad = typename FftCfgV<M, V>::radices;
  static_assert(RadixInfo<Rad>::first == E, "first radix must equal E");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<To>* lds = reinterpret_cast<cplx<To>*>(smem_raw);
  const LayRowPad lay{RS};
  const int c = threadIdx.x / TCOL, t = threadIdx.x % TCOL;
  const int64_t row = (int64_t)blockIdx.x * tr + c;
  const bool rowok = row < nrows;
  C v[1][E];
  {
    // Rows past the end read row 0 (discarded), so the E loads are
    // unconditional and issue back to back before the first use.
    using P = typename PairOf<Tin>::type;
    const int64_t rowc = rowok ? row : 0;
    const P* src = reinterpret_cast<const P*>(
        in2 && rowc >= nrows1 ? in2 + (rowc - nrows1) * ldi : in + rowc * ldi);
    const auto lr = ld.row(rowc);
    P pv[E];
#pragma unroll
    for (int s = 0; s < E; ++s) pv[s] = ld_stream(src + t + s * TCOL);
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const int e = t + s * TCOL;
      const C z{T(lr((double)pv[s].x, 2 * e)), T(lr((double)pv[s].y, 2 * e + 1))};
      v[0][s] = rowok ? z : C{T(0), T(0)};
    }
  }
  rstages<T, M, -1, 1, 1, true, V>(v, lds, 0, twM, c, t, lay);
  __syncthreads();  // the last stage's LDS reads precede the stores below
#pragma unroll
  for (int s = 0; s < E; ++s) lds[lay(c, t + out_m<M, V>(s) * TCOL)] = to_cplx<To>(v[0][s]);
  __syncthreads();
  // Hermitian half: X[k] = (Z[k] + conj Z[M-k])/2 - i W^k (Z[k] - conj Z[M-k])/2
  if (!rowok) return;
  cplx<To>* dst = out + row * ldo;
  // FP64: W^k for k = t, t + TCOL, ... by repeated products with W^TCOL
  // (error ~ M / TCOL ulp of FP64) rather than one table load per k.
  C w = twN[t];
  const C wstep = twN[TCOL];
  auto post = [&](int k, const C& zk, const C& zm) {
    const T ax = zk.x + zm.x, ay = zk.y - zm.y;
    const T bx = zk.x - zm.x, by = zk.y + zm.y;
    if constexpr (sizeof(T) != 8) w = twN[k];
    const T px = w.x * by + w.y * bx;
    const T py = -(w.x * bx - w.y * by);
    st_stream(dst + k, to_cplx<To>(C{T(0.5) * (ax + px), T(0.5) * (ay + py)}));
    if constexpr (sizeof(T) == 8) w = cmul(w, wstep);
  };
  if constexpr (TCOL % 8 == 0) {
    // Thread t's outputs k = t + j TCOL sit at lay(c, t) + j (TCOL + TCOL / 8)
    // and their mirrors M - k at lay(c, M - t) - j (TCOL + TCOL / 8) (the pad
    // slot every 8 elements moves by TCOL / 8 per step): constant offsets.
    // Thread 0 reads Z[0] for both at k = 0 and also writes k = M.
    constexpr int kStep = TCOL + TCOL / 8;
    const int bz = lay(c, t), bm = lay(c, M - t);
#pragma unroll
    for (int j = 0; j < E; ++j) {
      const C zk = to_cplx<T>(lds[bz + j * kStep]);
      const C zm = to_cplx<T>(lds[j == 0 && t == 0 ? bz : bm - j * kStep]);
      post(t + j * TCOL, zk, zm);
    }
    if (t == 0) {
      const C z0 = to_cplx<T>(lds[bz]);
      post(M, z0, z0);
    }
  } else {
    for (int k = t; k <= M; k += TCOL)
      post(k, to_cplx<T>(lds[lay(c, k == M ? 0 : k)]), to_cplx<T>(lds[lay(c, k == 0 ? 0 : M - k)]));
  }
}

// Register-chained C2R rows (see k_fft_rows_c2r): the pre-process runs in LDS,
// the length-M FFT starts from registers and the real samples go straight
// from registers to global memory.  nsel == 0 writes whole rows; nsel > 0
// writes sample (o - off) mod 2M at o < nsel (offm = off mod 2M).  In place
// when the real rows overlay the complex rows: every workgroup reads its rows
// before writing them.
template <typename T, typename Tout, int M, int V = 0>
__global__ void __launch_bounds__(1024)
    k_fft_rows_c2r_rc(const cplx<T>* in, int64_t ldi, Tout* out, int64_t ldo, int64_t nrows,
                      int tr, const cplx<T>* __restrict__ twM, const cplx<T>* __restrict__ twN,
                      int nsel, int offm) {
  using C = cplx<T>;
  constexpr int E = FftCfgV<M, V>::E;
  constexpr int TCOL = M / E;
  constexpr int RS = row_pad_len(M);
  using Rad = typename FftCfgV<M, V>::radices;
  static_assert(RadixInfo<Rad>::first == E, "first radix must equal E");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const LayRowPad lay{RS};
  const int c = threadIdx.x / TCOL, t = threadIdx.x % TCOL;
  const int64_t row = (int64_t)blockIdx.x * tr + c;
  const bool rowok = row < nrows;
  // Thread t handles k = t + j TCOL <= M (j <= E, j = E only at t = 0) and
  // the pairs p = t + j TCOL <= M / 2 (j <= E / 2, j = E / 2 only at t = 0).
  // Every global load (row values, twiddles) is issued before the first
  // use, from valid addresses (rows past the end read row 0, discarded).
  constexpr int NP = E / 2 + 1;
  const C* src = in + (rowok ? row : 0) * ldi;
  C xs[E + 1], wps[NP], wqs[NP];
#pragma unroll
  for (int j = 0; j <= E; ++j) xs[j] = src[min(t + j * TCOL, M)];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int p = min(t + j * TCOL, M / 2);
    wps[j] = twN[p];
    wqs[j] = twN[M - p];
  }
#pragma unroll
  for (int j = 0; j <= E; ++j) {
    const int k = t + j * TCOL;
    if (k <= M) {
      C x = rowok ? xs[j] : C{T(0), T(0)};
      if (k == 0 || k == M) x.y = T(0);
      lds[lay(c, k)] = x;
    }
  }
  __syncthreads();
  // Z[k] = (X[k] + conj X[M-k]) + i W^-k (X[k] - conj X[M-k]) for k < M, in
  // pairs (p, M - p) so the update is in place.
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int p = t + j * TCOL;
    if (p > M / 2) continue;
    const C xp = lds[lay(c, p)];
    const C xq = lds[lay(c, M - p)];
    const C wp = wps[j];
    const C wq = wqs[j];
    {
      const T ax = xp.x + xq.x, ay = xp.y - xq.y;
      const T bx = xp.x - xq.x, by = xp.y + xq.y;
      const T cx = wp.x * bx + wp.y * by, cy = wp.x * by - wp.y * bx;
      lds[lay(c, p)] = C{ax - cy, ay + cx};
    }
    if (p != 0 && p != M - p) {
      const T ax = xq.x + xp.x, ay = xq.y - xp.y;
      const T bx = xq.x - xp.x, by = xq.y + xp.y;
      const T cx = wq.x * bx + wq.y * by, cy = wq.x * by - wq.y * bx;
      lds[lay(c, M - p)] = C{ax - cy, ay + cx};
    }
  }
  __syncthreads();
  C v[1][E];
#pragma unroll
  for (int s = 0; s < E; ++s) v[0][s] = lds[lay(c, t + s * TCOL)];
  __syncthreads();  // every first-stage input is read before the first stage stores
  rstages<T, M, +1, 1, 1, true, V>(v, lds, 0, twM, c, t, lay);
  if (!rowok) return;
  Tout* dst = out + row * ldo;
  if (nsel == 0) {
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const int e = t + out_m<M, V>(s) * TCOL;
      dst[2 * e] = Tout(v[0][s].x);
      dst[2 * e + 1] = Tout(v[0][s].y);
    }
  } else {
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const int e = t + out_m<M, V>(s) * TCOL;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        int o = 2 * e + h + offm;
        if (o >= 2 * M) o -= 2 * M;
        const Tout val = Tout(h ? v[0][s].y : v[0][s].x);
        for (; o < nsel; o += 2 * M) dst[o] = val;
      }
    }
  }
}

// ---- strided column pass ---------------------------------------------------
// Transforms along L over [A][L][G][B] complex data (G groups of B contiguous
// columns):
//   input  element (a, e, g, b) at a*inA  + e*inL  + g*inG  + b,
//   output element (a, o, g, b) at a*outA + o*outL + g*outG + b,
// with o in [0, nout).  off == 0 writes outputs o = e (nout == L: every
// output); off != 0 writes outputs o < nout holding element e = (o - off)
// mod L (the pruned inverse: only the lags a separation subgrid needs, off =
// nout / 2; nout may equal or exceed L).  In place when in == out.
// grid = (G * tilesB, A), tilesB = ceil(B / tc); block = tc * (L / E);
// LDS = L * tc complex.
// The input may be stored in another precision Tin (converted on load).
template <typename T, int L, int DIR, typename Tin = T>
__global__ void __launch_bounds__(1024)
    k_fft_cols(const cplx<Tin>* in, int64_t inA, int64_t inL, int64_t inG, cplx<T>* out,
               int64_t outA, int64_t outL, int64_t outG, int B, int tilesB, int tc, int nout,
               int off, const cplx<T>* __restrict__ tw) {
  using C = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int g = blockIdx.x / tilesB;
  const int b0 = (blockIdx.x - g * tilesB) * tc;
  const int a = blockIdx.y;
  const cplx<Tin>* src = in + (int64_t)a * inA + (int64_t)g * inG + b0;
  C* dst = out + (int64_t)a * outA + (int64_t)g * outG + b0;
  const int nthr = blockDim.x;
  const int tid = threadIdx.x;
  const int ncol = min(tc, B - b0);
  // L * tc elements over tc * (L / E) threads: exactly E per thread, all
  // loads issued before the first LDS store.
  constexpr int E = FftCfg<L>::E;
  {
    C tmp[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int idx = tid + q * nthr;
      const int e = idx / tc, c = idx - e * tc;
      if (c < ncol) {
        const cplx<Tin> x = src[e * inL + c];
        tmp[q] = C{T(x.x), T(x.y)};
      } else {
        tmp[q] = C{T(0), T(0)};
      }
    }
#pragma unroll
    for (int q = 0; q < E; ++q) lds[tid + q * nthr] = tmp[q];
  }
  __syncthreads();
  fft_lds<T, L, DIR>(lds, tw, tid % tc, tid / tc, LayColFast{tc});
  for (int idx = tid; idx < nout * tc; idx += nthr) {
    const int o = idx / tc, c = idx - o * tc;
    const int e = off == 0 ? o : ((o - off) % L + L) % L;
    if (c < ncol) dst[o * outL + c] = lds[e * tc + c];
  }
}

// Direct DFT fallback, one column per block (any L up to LDS capacity); same
// data layout and pruning as k_fft_cols.  grid = (G * B, A).
template <typename T, int DIR, typename Tin = T>
__global__ void __launch_bounds__(256)
    k_dft_cols(const cplx<Tin>* in, int64_t inA, int64_t inL, int64_t inG, cplx<T>* out,
               int64_t outA, int64_t outL, int64_t outG, int B, int L, int nout, int off,
               const cplx<T>* __restrict__ tw) {
  using C = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int g = blockIdx.x / B;
  const int b = blockIdx.x - g * B;
  const cplx<Tin>* src = in + (int64_t)blockIdx.y * inA + (int64_t)g * inG + b;
  C* dst = out + (int64_t)blockIdx.y * outA + (int64_t)g * outG + b;
  for (int e = threadIdx.x; e < L; e += blockDim.x) {
    const cplx<Tin> x = src[e * inL];
    lds[e] = C{T(x.x), T(x.y)};
  }
  __syncthreads();
  for (int o = threadIdx.x; o < nout; o += blockDim.x) {
    const int k = off == 0 ? o : ((o - off) % L + L) % L;
    T sx = 0, sy = 0;
    int idx = 0;
    for (int e = 0; e < L; ++e) {
      C w = tw[idx];
      if (DIR > 0) w.y = -w.y;
      const C x = lds[e];
      sx += x.x * w.x - x.y * w.y;
      sy += x.x * w.y + x.y * w.x;
      idx += k;
      if (idx >= L) idx -= L;
    }
    dst[o * outL] = C{sx, sy};  // each block owns its column
  }
}

// ---- row (z) pass: real <-> Hermitian half --------------------------------
// Loader: row(r) gives the per-row functor f with f(double v, int k) the
// value transformed for the k-th input sample v of row r (the row's own
// quantities are formed once per thread, not per sample).  Rows are
// [nrows][ldi] (input type Tin) and [nrows][ldo] complex outputs.
struct LoadCopy {
  struct Row {
    __device__ __forceinline__ double operator()(double v, int) const { return v; }
  };
  __device__ __forceinline__ Row row(int64_t) const { return Row{}; }
  void shift_x(int) {}  // rows restart at a later x slab: nothing to adjust
};

// block = tr * (M / E) threads, handling rows blockIdx.x*tr ... ; LDS tr*(M+1).
template <typename T, typename Tin, int M, class Loader>
__global__ void __launch_bounds__(1024) k_fft_rows_r2c(const Tin* in, int64_t ldi,
                                                       cplx<T>* out, int64_t ldo,
                                                       int64_t nrows, int tr,
                                                       const cplx<T>* __restrict__ twM,
                                                       const cplx<T>* __restrict__ twN,
                                                       Loader ld) {
  using C = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  constexpr int RS = M + 1;
  const int64_t row0 = (int64_t)blockIdx.x * tr;
  const int nthr = blockDim.x, tid = threadIdx.x;
  // tr * M samples over tr * (M / E) threads: exactly E pairs per thread.
  {
    constexpr int E = FftCfg<M>::E;
    C tmp[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int idx = tid + q * nthr;
      const int c = idx / M, m = idx - c * M;
      const int64_t row = row0 + c;
      C z{T(0), T(0)};
      if (row < nrows) {
        const Tin* p = in + row * ldi + 2 * m;
        const auto lr = ld.row(row);
        z.x = T(lr((double)p[0], 2 * m));
        z.y = T(lr((double)p[1], 2 * m + 1));
      }
      tmp[q] = z;
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int idx = tid + q * nthr;
      const int c = idx / M, m = idx - c * M;
      lds[c * RS + m] = tmp[q];
    }
  }
  __syncthreads();
  constexpr int TCOL = M / FftCfg<M>::E;
  fft_lds<T, M, -1>(lds, twM, tid / TCOL, tid % TCOL, LayRow{RS});
  // Hermitian half: X[k] = (Z[k] + conj Z[M-k])/2 - i W^k (Z[k] - conj Z[M-k])/2
  for (int idx = tid; idx < tr * (M + 1); idx += nthr) {
    const int c = idx / (M + 1), k = idx - c * (M + 1);
    const int64_t row = row0 + c;
    if (row >= nrows) continue;
    const C zk = lds[c * RS + (k == M ? 0 : k)];
    const C zm = lds[c * RS + (k == 0 ? 0 : M - k)];
    const T ax = zk.x + zm.x, ay = zk.y - zm.y;
    const T bx = zk.x - zm.x, by = zk.y + zm.y;
    const C w = twN[k];
    const T px = w.x * by + w.y * bx;
    const T py = -(w.x * bx - w.y * by);
    out[row * ldo + k] = C{T(0.5) * (ax + px), T(0.5) * (ay + py)};
  }
}

// Inverse: Hermitian half rows -> real rows (unnormalised; imaginary parts of
// the DC and Nyquist terms are ignored, as FFTW's c2r does).
// nsel == 0 writes the whole real row; nsel > 0 writes only samples
// n = (o - off) mod 2M for o < nsel (the lags a separation subgrid needs).
template <typename T, typename Tout, int M>
__global__ void __launch_bounds__(1024) k_fft_rows_c2r(const cplx<T>* in, int64_t ldi,
                                                       Tout* out, int64_t ldo,
                                                       int64_t nrows, int tr,
                                                       const cplx<T>* __restrict__ twM,
                                                       const cplx<T>* __restrict__ twN,
                                                       int nsel, int off) {
  using C = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  constexpr int RS = M + 1;
  const int64_t row0 = (int64_t)blockIdx.x * tr;
  const int nthr = blockDim.x, tid = threadIdx.x;
  {
    // Elements 0..M-1 of each row: exactly E per thread; element M below.
    constexpr int E = FftCfg<M>::E;
    C tmp[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int idx = tid + q * nthr;
      const int c = idx / M, k = idx - c * M;
      const int64_t row = row0 + c;
      C x{T(0), T(0)};
      if (row < nrows) x = in[row * ldi + k];
      if (k == 0) x.y = T(0);
      tmp[q] = x;
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int idx = tid + q * nthr;
      const int c = idx / M, k = idx - c * M;
      lds[c * RS + k] = tmp[q];
    }
    if (tid < tr) {
      const int64_t row = row0 + tid;
      C x{T(0), T(0)};
      if (row < nrows) x = in[row * ldi + M];
      x.y = T(0);
      lds[tid * RS + M] = x;
    }
  }
  __syncthreads();
  // Z[k] = (X[k] + conj X[M-k]) + i W^-k (X[k] - conj X[M-k]) for k < M,
  // computed in pairs (p, M-p) so the update can be in place.
  for (int idx = tid; idx < tr * (M / 2 + 1); idx += nthr) {
    const int c = idx / (M / 2 + 1), p = idx - c * (M / 2 + 1);
    const C xp = lds[c * RS + p];
    const C xq = lds[c * RS + (M - p)];
    const C wp = twN[p];       // exp(-2 pi i p / N); use conj -> W^-p
    const C wq = twN[M - p];
    // Z = A + i conj(w) B with A = X[k] + conj X[M-k], B = X[k] - conj X[M-k];
    // conj(w) B = (w.x bx + w.y by, w.x by - w.y bx), and i (cx + i cy) = (-cy, cx).
    {
      const T ax = xp.x + xq.x, ay = xp.y - xq.y;
      const T bx = xp.x - xq.x, by = xp.y + xq.y;
      const T cx = wp.x * bx + wp.y * by, cy = wp.x * by - wp.y * bx;
      lds[c * RS + p] = C{ax - cy, ay + cx};
    }
    if (p != 0 && p != M - p) {
      const T ax = xq.x + xp.x, ay = xq.y - xp.y;
      const T bx = xq.x - xp.x, by = xq.y + xp.y;
      const T cx = wq.x * bx + wq.y * by, cy = wq.x * by - wq.y * bx;
      lds[c * RS + (M - p)] = C{ax - cy, ay + cx};
    }
  }
  __syncthreads();
  constexpr int TCOL = M / FftCfg<M>::E;
  fft_lds<T, M, +1>(lds, twM, tid / TCOL, tid % TCOL, LayRow{RS});
  if (nsel == 0) {
    for (int idx = tid; idx < tr * M; idx += nthr) {
      const int c = idx / M, m = idx - c * M;
      const int64_t row = row0 + c;
      if (row >= nrows) continue;
      const C z = lds[c * RS + m];
      Tout* p = out + row * ldo + 2 * m;
      p[0] = Tout(z.x);
      p[1] = Tout(z.y);
    }
  } else {
    const T* ldr = reinterpret_cast<const T*>(lds);
    for (int idx = tid; idx < tr * nsel; idx += nthr) {
      const int c = idx / nsel, o = idx - c * nsel;
      const int64_t row = row0 + c;
      if (row >= nrows) continue;
      const int n = ((o - off) % (2 * M) + 2 * M) % (2 * M);
      out[row * ldo + o] = Tout(ldr[2 * (c * RS) + n]);
    }
  }
}

// Direct-DFT fallback rows (any n2): one block per row.
template <typename T, typename Tin, class Loader>
__global__ void __launch_bounds__(256) k_dft_rows_r2c(const Tin* in, int64_t ldi,
                                                      cplx<T>* out, int64_t ldo,
                                                      int n2, const cplx<T>* __restrict__ tw,
                                                      Loader ld) {
  using C = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* x = reinterpret_cast<T*>(smem_raw);
  const int64_t row = blockIdx.x;
  const auto lr = ld.row(row);
  for (int n = threadIdx.x; n < n2; n += blockDim.x)
    x[n] = T(lr((double)in[row * ldi + n], n));
  __syncthreads();
  // Every global read of the row happened before the barrier above, so the
  // output row may alias the input row (in-place transform).
  const int nh = n2 / 2 + 1;
  for (int k = threadIdx.x; k < nh; k += blockDim.x) {
    T sx = 0, sy = 0;
    int idx = 0;
    for (int n = 0; n < n2; ++n) {
      const C w = tw[idx];
      sx += x[n] * w.x;
      sy += x[n] * w.y;
      idx += k;
      if (idx >= n2) idx -= n2;
    }
    out[row * ldo + k] = C{sx, sy};
  }
}

template <typename T, typename Tout>
__global__ void __launch_bounds__(256) k_dft_rows_c2r(const cplx<T>* in, int64_t ldi,
                                                      Tout* out, int64_t ldo, int n2,
                                                      const cplx<T>* __restrict__ tw, int nsel,
                                                      int off) {
  using C = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* X = reinterpret_cast<C*>(smem_raw);
  const int64_t row = blockIdx.x;
  const int nh = n2 / 2 + 1;
  for (int k = threadIdx.x; k < nh; k += blockDim.x) {
    C v = in[row * ldi + k];
    if (k == 0 || (2 * k == n2)) v.y = T(0);
    X[k] = v;
  }
  __syncthreads();
  const int nout = nsel == 0 ? n2 : nsel;
  for (int o = threadIdx.x; o < nout; o += blockDim.x) {
    const int n = nsel == 0 ? o : ((o - off) % n2 + n2) % n2;
    T s = 0;
    int idx = 0;
    for (int k = 0; k < n2; ++k) {
      const C xk = k < nh ? X[k] : C{X[n2 - k].x, -X[n2 - k].y};
      const C w = tw[idx];  // exp(-2 pi i n k / n2); the inverse uses conj(w)
      s += xk.x * w.x + xk.y * w.y;
      idx += n;
      if (idx >= n2) idx -= n2;
    }
    out[row * ldo + o] = Tout(s);
  }
}

}  // namespace fc
