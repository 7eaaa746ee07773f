// gfx950 kernels for the CIRSRADG_WAVE path.
//
//   layer_kernel     k-table interpolation (get_kg.f:308-476), CIA continuum
//                    (nciacon.f, ciaread.f) and random-overlap g-mixing
//                    (noverlapg.f, rankg.f) for one (wavenumber, layer) per
//                    wavefront.  Each overlap round sorts the NG*NG mixed
//                    keys with a register bitonic network, takes an FP64
//                    prefix of the weights and lets NG "bin lanes" rebin in
//                    the reference's sequential order.
//   emission_kernel  per (wavenumber, path): layer optical depth, FP64
//                    transmission prefix product, Planck source and an
//                    affine suffix scan that gives dR/dtau for every layer
//                    in O(L) (cirsradg_wave.f:896-1035 computes the same
//                    derivative with an O(L^2) recurrence), then the
//                    g-quadrature (cirsradg_wave.f:1240-1286).
//   quadrature_kernel  angle-weighted sum over paths (forwardavfovX.f:330,356).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <math.h>

#include "radg_dev.h"

namespace {

constexpr int W = RADG_WAVE;

// --------------------------------------------------------------- helpers
__device__ __forceinline__ long long xcd_item(long long nblocks) {
  // Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
  // range of items so that the layers of one wavenumber (which read the
  // same table slice) share an L2.
  const long long b = blockIdx.x;
  const long long chunk = nblocks / 8;
  return (b % 8) * chunk + b / 8;
}

// 1/n for index arithmetic of the form floor((x + 0.5) / n) with x < 1024
// and n <= 64: the hardware reciprocal (1 ulp) keeps the product far inside
// the 0.5/n margin, where an IEEE division would cost about ten instructions
__device__ __forceinline__ float rcp_idx(int n) { return __builtin_amdgcn_rcpf((float)n); }

// 1/x for x > 0 to within a few FP64 ulps: the hardware reciprocal and two
// Newton steps (each doubles the correct bits), about half the instructions
// of a correctly rounded division
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// The request r of an item, start[r] <= item < start[r + 1], with
// start[nreq] = total.  The requests of a launch class often hold equally
// many items, so the proportional guess is tried first (one pair of loads
// rather than a chain of dependent ones), then a binary search.
// UNI: item is the same in every lane (one item per wavefront), and the
// guess is kept in a scalar register so that the loads below, and every
// address derived from the request, stay scalar.
template <bool UNI>
__device__ __forceinline__ int find_req(const long long *start, int nreq, long long item,
                                        long long total) {
  int g = min(max((int)((float)item * ((float)nreq / (float)total)), 0), nreq - 1);
  if constexpr (UNI) g = __builtin_amdgcn_readfirstlane(g);
  if (start[g] <= item && item < start[g + 1]) return g;
  int lo = 0, hi = nreq;  // start[lo] <= item < start[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (start[mid] <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

// n / d for an item's offset n >= 0 within its request: a 32-bit division
// where n fits (request sizes are bounded far below 2^31 items)
template <bool UNI>
__device__ __forceinline__ int div_item(long long n, int d) {
  const int q = (n <= 0x7fffffffLL) ? (int)((unsigned)n / (unsigned)d) : (int)(n / d);
  // UNI: n is the same in every lane, and the quotient stays scalar
  if constexpr (UNI) return __builtin_amdgcn_readfirstlane(q);
  return q;
}

// A uniform 64-bit value (a field read through a vector load) kept in
// scalar registers, so addresses formed from it stay scalar.
__device__ __forceinline__ long long rfl64(long long x) {
  const unsigned long long u = (unsigned long long)x;
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)u);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// The lane number formed where it is used (two mbcnt in a volatile asm, which
// the compiler neither hoists nor merges with another), so that no copy of
// it is held, and spilled, across the overlap-round loop.
__device__ __forceinline__ int lane_fresh() {
  int l;
  asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
  return l;
}

// Zeros formed where they are stored or used (an asm the compiler does not
// merge), rather than one zero register held across the overlap rounds,
// which the register budget of the layer kernel otherwise spills.
__device__ __forceinline__ double zero_d() {
  double z;
  asm volatile("v_mov_b64 %0, 0" : "=v"(z));
  return z;
}
__device__ __forceinline__ int zero_i() {
  int z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

// Lane (lane ^ m)'s value by ds_bpermute, with the address formed from the
// caller's lane number: a shuffle that forms it from its own lane id has
// that id hoisted out of the overlap-round loop and spilled across it.
__device__ __forceinline__ int bperm_xor(int x, int lane, int m) {
  return __builtin_amdgcn_ds_bpermute((lane ^ m) << 2, x);
}
__device__ __forceinline__ double bperm_xor(double x, int lane, int m) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const int lo = bperm_xor((int)(unsigned)u, lane, m);
  const int hi = bperm_xor((int)(unsigned)(u >> 32), lane, m);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                          (unsigned long long)(unsigned)lo));
}

template <typename T> __device__ __forceinline__ T dexp(T x);
template <> __device__ __forceinline__ float dexp<float>(float x) { return expf(x); }
template <> __device__ __forceinline__ double dexp<double>(double x) { return exp(x); }
template <typename T> __device__ __forceinline__ T dlog(T x);
template <> __device__ __forceinline__ float dlog<float>(float x) { return logf(x); }
template <> __device__ __forceinline__ double dlog<double>(double x) { return log(x); }
// dpexp.f: exp with the argument clamped to [-500, 88.7] (the imod 0 and 1
// transmissions, cirsradg_wave.f:861,888)
template <typename T>
__device__ __forceinline__ T dpexp(T x) {
  const T y = x > (T)-500 ? x : (T)-500;
  return dexp(y >= (T)88.7 ? (T)88.7 : y);
}
template <typename T> __device__ __forceinline__ T dexpm1(T x);
template <> __device__ __forceinline__ float dexpm1<float>(float x) { return expm1f(x); }
template <> __device__ __forceinline__ double dexpm1<double>(double x) { return expm1(x); }
// a / b: IEEE division in FP64, one reciprocal (1 ulp) in FP32
template <typename T> __device__ __forceinline__ T ratio(T a, T b);
template <> __device__ __forceinline__ float ratio<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}
template <> __device__ __forceinline__ double ratio<double>(double a, double b) { return a / b; }
template <typename T> __device__ __forceinline__ T dinf();
template <> __device__ __forceinline__ float dinf<float>() { return __builtin_huge_valf(); }
template <> __device__ __forceinline__ double dinf<double>() { return __builtin_huge_val(); }

// Ascending bitonic networks over 64*E slots held E per lane (lane l owns
// slots l*E .. l*E+E-1), in the form where the first stage of each merge
// compares slot e with its mirror e ^ (k-1) and every later stage compares
// e with e ^ j; every exchange keeps the smaller value at the lower slot, so
// the only lane-dependent quantities are the six lane bits.
constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x >> 1); }

template <typename V>
struct SortOps;

// FP32 keys are >= 0, so their IEEE bit patterns order like their values:
// one u64 holds (key bits << 32 | element index); one 64-bit compare orders
// (key, index) pairs, ties broken by index.
template <>
struct SortOps<unsigned long long> {
  static __device__ __forceinline__ bool lt(unsigned long long a, unsigned long long b) { return a < b; }
  static __device__ __forceinline__ unsigned long long shfl(unsigned long long a, int m) {
    return __shfl_xor(a, m);
  }
};

template <typename V, int E>
__device__ __forceinline__ void bitonic_sort_v(V (&v)[E], int lane) {
  using O = SortOps<V>;
  constexpr int LOGE = ilog2c(E);
  constexpr int LOGP = 6 + LOGE;
#pragma unroll
  for (int ks = 1; ks <= LOGP; ks++) {
    // mirror stage: partner e ^ (2^ks - 1)
    if (ks - 1 >= LOGE) {
      const int lm = ((1 << ks) - 1) >> LOGE;
      const bool lower = ((lane >> (ks - 1 - LOGE)) & 1) == 0;
      V o[E];
#pragma unroll
      for (int q = 0; q < E; q++) o[q] = O::shfl(v[E - 1 - q], lm);
#pragma unroll
      for (int q = 0; q < E; q++) {
        const bool olt = O::lt(o[q], v[q]);
        v[q] = (olt == lower) ? o[q] : v[q];
      }
    } else {
      const int km = (1 << ks) - 1;
#pragma unroll
      for (int q = 0; q < E; q++) {
        const int q2 = q ^ km;
        if (q2 > q) {
          const V x = v[q], y = v[q2];
          const bool sw = O::lt(y, x);
          v[q] = sw ? y : x;
          v[q2] = sw ? x : y;
        }
      }
    }
    // half-cleaners
#pragma unroll
    for (int js = ks - 2; js >= 0; js--) {
      if (js >= LOGE) {
        const int lm = 1 << (js - LOGE);
        const bool lower = ((lane >> (js - LOGE)) & 1) == 0;
#pragma unroll
        for (int q = 0; q < E; q++) {
          const V o = O::shfl(v[q], lm);
          const bool olt = O::lt(o, v[q]);
          v[q] = (olt == lower) ? o : v[q];
        }
      } else {
        const int j = 1 << js;
#pragma unroll
        for (int q = 0; q < E; q++) {
          const int q2 = q ^ j;
          if (q2 > q) {
            const V x = v[q], y = v[q2];
            const bool sw = O::lt(y, x);
            v[q] = sw ? y : x;
            v[q2] = sw ? x : y;
          }
        }
      }
    }
  }
}

template <typename T, int E> struct Sorter;
template <int E> struct Sorter<float, E> {
  static __device__ __forceinline__ void sort(float (&key)[E], int (&idx)[E], int lane) {
    unsigned long long v[E];
#pragma unroll
    for (int q = 0; q < E; q++) {
      const float k = (key[q] > 0.0f) ? key[q] : 0.0f;  // -0 and 0 order alike
      v[q] = ((unsigned long long)__float_as_uint(k) << 32) | (unsigned int)idx[q];
    }
    bitonic_sort_v<unsigned long long, E>(v, lane);
#pragma unroll
    for (int q = 0; q < E; q++) {
      key[q] = __uint_as_float((unsigned int)(v[q] >> 32));
      idx[q] = (int)(unsigned int)(v[q] & 0xffffffffull);
    }
  }
};
// FP64 form: the same network on separate key and index arrays.
__device__ __forceinline__ bool kd_lt(double ka, int ia, double kb, int ib) {
  return (ka < kb) | ((ka == kb) & (ia < ib));
}

template <int E>
__device__ __forceinline__ void bitonic_sort_kd(double (&k)[E], int (&x)[E], int lane) {
  constexpr int LOGE = ilog2c(E);
  constexpr int LOGP = 6 + LOGE;
#pragma unroll
  for (int ks = 1; ks <= LOGP; ks++) {
    if (ks - 1 >= LOGE) {
      const int lm = ((1 << ks) - 1) >> LOGE;
      const bool lower = ((lane >> (ks - 1 - LOGE)) & 1) == 0;
      double ok[E];
      int ox[E];
#pragma unroll
      for (int q = 0; q < E; q++) {
        ok[q] = bperm_xor(k[E - 1 - q], lane, lm);
        ox[q] = bperm_xor(x[E - 1 - q], lane, lm);
      }
#pragma unroll
      for (int q = 0; q < E; q++) {
        const bool take = kd_lt(ok[q], ox[q], k[q], x[q]) == lower;
        k[q] = take ? ok[q] : k[q];
        x[q] = take ? ox[q] : x[q];
      }
    } else {
      const int km = (1 << ks) - 1;
#pragma unroll
      for (int q = 0; q < E; q++) {
        const int q2 = q ^ km;
        if (q2 > q) {
          const bool sw = kd_lt(k[q2], x[q2], k[q], x[q]);
          const double a0 = k[q], a1 = k[q2];
          const int b0 = x[q], b1 = x[q2];
          k[q] = sw ? a1 : a0;
          k[q2] = sw ? a0 : a1;
          x[q] = sw ? b1 : b0;
          x[q2] = sw ? b0 : b1;
        }
      }
    }
#pragma unroll
    for (int js = ks - 2; js >= 0; js--) {
      if (js >= LOGE) {
        const int lm = 1 << (js - LOGE);
        const bool lower = ((lane >> (js - LOGE)) & 1) == 0;
#pragma unroll
        for (int q = 0; q < E; q++) {
          const double ok = bperm_xor(k[q], lane, lm);
          const int ox = bperm_xor(x[q], lane, lm);
          const bool take = kd_lt(ok, ox, k[q], x[q]) == lower;
          k[q] = take ? ok : k[q];
          x[q] = take ? ox : x[q];
        }
      } else {
        const int j = 1 << js;
#pragma unroll
        for (int q = 0; q < E; q++) {
          const int q2 = q ^ j;
          if (q2 > q) {
            const bool sw = kd_lt(k[q2], x[q2], k[q], x[q]);
            const double a0 = k[q], a1 = k[q2];
            const int b0 = x[q], b1 = x[q2];
            k[q] = sw ? a1 : a0;
            k[q2] = sw ? a0 : a1;
            x[q] = sw ? b1 : b0;
            x[q2] = sw ? b0 : b1;
          }
        }
      }
    }
  }
}

template <int E> struct Sorter<double, E> {
  static __device__ __forceinline__ void sort(double (&key)[E], int (&idx)[E], int lane) {
    bitonic_sort_kd<E>(key, idx, lane);
  }
};

// ------------------------------------------- 32-bit sort for FP32 keys
// Value of lane (lane ^ m) for a lane mask m known after unrolling: DPP
// quad permutes, row mirrors and the rotation of a 16-lane row by 8 where
// the pattern exists on gfx9, two DPP moves for 4, a swizzle within 32
// lanes for 16 and 31, and ds_bpermute only across the two wave halves.
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int x) {
  // every lane has a source lane in these patterns, so no lane keeps an old value
  return __builtin_amdgcn_mov_dpp(x, CTRL, 0xF, 0xF, true);
}
template <int M>
__device__ __forceinline__ unsigned xlane(unsigned v, int lane) {
  const int x = (int)v;
  if constexpr (M == 1) return (unsigned)dpp_mov<0xB1>(x);                  // quad_perm 1,0,3,2
  else if constexpr (M == 2) return (unsigned)dpp_mov<0x4E>(x);             // quad_perm 2,3,0,1
  else if constexpr (M == 3) return (unsigned)dpp_mov<0x1B>(x);             // quad_perm 3,2,1,0
  else if constexpr (M == 4) return (unsigned)dpp_mov<0x141>(dpp_mov<0x1B>(x));   // ^3, ^7
  else if constexpr (M == 7) return (unsigned)dpp_mov<0x141>(x);            // row_half_mirror
  else if constexpr (M == 8) return (unsigned)dpp_mov<0x128>(x);            // row_ror:8
  else if constexpr (M == 15) return (unsigned)dpp_mov<0x140>(x);           // row_mirror
  else if constexpr (M == 16 || M == 31)
    return (unsigned)__builtin_amdgcn_ds_swizzle(x, 0x1F | (M << 10));
  else return (unsigned)bperm_xor(x, lane, M);
}

// Exchange across lanes: the lower lane of a pair keeps the smaller value,
// the upper lane the larger, i.e. med3(v, partner, c) with c = 0 in lower
// lanes and c = 0xffffffff in upper lanes.
__device__ __forceinline__ unsigned med3_u32(unsigned a, unsigned b, unsigned c) {
  unsigned r;
  asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
// 0 where bit b of the lane number is 0, 0xffffffff where it is 1
__device__ __forceinline__ unsigned lane_bit_mask(int lane, int b) {
  return (unsigned)(((int)((unsigned)lane << (31 - b))) >> 31);
}

// Ascending bitonic network (same mirror form as above) on 32-bit unsigned
// values, unrolled by template recursion so that every lane pattern is a
// constant: an in-lane compare-exchange is one min and one max, a cross-lane
// one a lane move and a med3.
template <int E, int JS>
__device__ __forceinline__ void u32_half_cleaners(unsigned (&v)[E], int lane) {
  constexpr int LOGE = ilog2c(E);
  if constexpr (JS >= LOGE) {
    const unsigned c = lane_bit_mask(lane, JS - LOGE);
#pragma unroll
    for (int q = 0; q < E; q++) v[q] = med3_u32(v[q], xlane<(1 << (JS - LOGE))>(v[q], lane), c);
  } else {
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int q2 = q ^ (1 << JS);
      if (q2 > q) {
        const unsigned x = v[q], y = v[q2];
        v[q] = min(x, y);
        v[q2] = max(x, y);
      }
    }
  }
  if constexpr (JS > 0) u32_half_cleaners<E, JS - 1>(v, lane);
}

template <int E, int KS>
__device__ __forceinline__ void u32_merge_levels(unsigned (&v)[E], int lane) {
  constexpr int LOGE = ilog2c(E);
  constexpr int LOGP = 6 + LOGE;
  // mirror stage: partner e ^ (2^KS - 1)
  if constexpr (KS - 1 >= LOGE) {
    const unsigned c = lane_bit_mask(lane, KS - 1 - LOGE);
    unsigned o[E];
#pragma unroll
    for (int q = 0; q < E; q++) o[q] = xlane<(((1 << KS) - 1) >> LOGE)>(v[E - 1 - q], lane);
#pragma unroll
    for (int q = 0; q < E; q++) v[q] = med3_u32(v[q], o[q], c);
  } else {
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int q2 = q ^ ((1 << KS) - 1);
      if (q2 > q) {
        const unsigned x = v[q], y = v[q2];
        v[q] = min(x, y);
        v[q2] = max(x, y);
      }
    }
  }
  if constexpr (KS >= 2) u32_half_cleaners<E, KS - 2>(v, lane);
  if constexpr (KS < LOGP) u32_merge_levels<E, KS + 1>(v, lane);
}

template <int E>
__device__ __forceinline__ void bitonic_sort_u32(unsigned (&v)[E], int lane) {
  u32_merge_levels<E, 1>(v, lane);
}
// The same network when each lane's E values are already ascending: the
// merge levels within a lane leave sorted runs unchanged and are skipped.
template <int E>
__device__ __forceinline__ void bitonic_sort_u32_lanesorted(unsigned (&v)[E], int lane) {
  u32_merge_levels<E, ilog2c(E) + 1>(v, lane);
}

// The same network on 64-bit words (the FP64 form's truncated keys): the
// lane moves of xlane on both halves, and one 64-bit compare per exchange.
template <int M>
__device__ __forceinline__ unsigned long long xlane64(unsigned long long v, int lane) {
  const unsigned lo = xlane<M>((unsigned)v, lane), hi = xlane<M>((unsigned)(v >> 32), lane);
  return ((unsigned long long)hi << 32) | lo;
}
// the lower lane of a pair keeps the smaller word, the upper lane the larger
__device__ __forceinline__ unsigned long long cx64(unsigned long long v, unsigned long long o,
                                                   bool upper) {
  return ((o < v) != upper) ? o : v;
}
template <int E, int JS>
__device__ __forceinline__ void u64_half_cleaners(unsigned long long (&v)[E], int lane) {
  constexpr int LOGE = ilog2c(E);
  if constexpr (JS >= LOGE) {
    const bool up = ((lane >> (JS - LOGE)) & 1) != 0;
#pragma unroll
    for (int q = 0; q < E; q++) v[q] = cx64(v[q], xlane64<(1 << (JS - LOGE))>(v[q], lane), up);
  } else {
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int q2 = q ^ (1 << JS);
      if (q2 > q) {
        const unsigned long long x = v[q], y = v[q2];
        const bool sw = y < x;
        v[q] = sw ? y : x;
        v[q2] = sw ? x : y;
      }
    }
  }
  if constexpr (JS > 0) u64_half_cleaners<E, JS - 1>(v, lane);
}
template <int E, int KS>
__device__ __forceinline__ void u64_merge_levels(unsigned long long (&v)[E], int lane) {
  constexpr int LOGE = ilog2c(E);
  constexpr int LOGP = 6 + LOGE;
  if constexpr (KS - 1 >= LOGE) {
    const bool up = ((lane >> (KS - 1 - LOGE)) & 1) != 0;
    unsigned long long o[E];
#pragma unroll
    for (int q = 0; q < E; q++) o[q] = xlane64<(((1 << KS) - 1) >> LOGE)>(v[E - 1 - q], lane);
#pragma unroll
    for (int q = 0; q < E; q++) v[q] = cx64(v[q], o[q], up);
  } else {
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int q2 = q ^ ((1 << KS) - 1);
      if (q2 > q) {
        const unsigned long long x = v[q], y = v[q2];
        const bool sw = y < x;
        v[q] = sw ? y : x;
        v[q2] = sw ? x : y;
      }
    }
  }
  if constexpr (KS >= 2) u64_half_cleaners<E, KS - 2>(v, lane);
  if constexpr (KS < LOGP) u64_merge_levels<E, KS + 1>(v, lane);
}
template <int E>
__device__ __forceinline__ void bitonic_sort_u64(unsigned long long (&v)[E], int lane) {
  u64_merge_levels<E, 1>(v, lane);
}
template <int E>
__device__ __forceinline__ void bitonic_sort_u64_lanesorted(unsigned long long (&v)[E], int lane) {
  u64_merge_levels<E, ilog2c(E) + 1>(v, lane);
}

// After the network has ordered the truncated words (key bits with the low
// log2(64E) mantissa bits replaced by the element number), slots can only be
// out of exact (key, element) order inside a run of equal truncated keys.
// Each slot's exact pair is one 64-bit word, key bits high and element low:
// FP32 keys are >= 0, so the words order like the pairs, and a pair compare
// is one 64-bit compare.  Odd-even transposition on the words, until no
// adjacent pair is out of order (a run of m slots needs at most m phases;
// the bound W*E is never reached on a sorted input).
// lane + 1's and lane - 1's value (DPP wave shifts; 0 where there is no such
// lane)
__device__ __forceinline__ int lane_next(int x) { return __builtin_amdgcn_mov_dpp(x, 0x130, 0xF, 0xF, true); }
__device__ __forceinline__ int lane_prev(int x) { return __builtin_amdgcn_mov_dpp(x, 0x138, 0xF, 0xF, true); }
__device__ __forceinline__ unsigned long long lane_next(unsigned long long x) {
  return ((unsigned long long)(unsigned)lane_next((int)(unsigned)(x >> 32)) << 32) |
         (unsigned)lane_next((int)(unsigned)x);
}
__device__ __forceinline__ unsigned long long lane_prev(unsigned long long x) {
  return ((unsigned long long)(unsigned)lane_prev((int)(unsigned)(x >> 32)) << 32) |
         (unsigned)lane_prev((int)(unsigned)x);
}
// (key, element) as one word
__device__ __forceinline__ unsigned long long key_word(float k, unsigned e) {
  return ((unsigned long long)__float_as_uint(k) << 32) | e;
}

template <int E>
__device__ __forceinline__ void exact_fixup(unsigned long long (&w)[E], int lane) {
  auto unsorted = [&]() {
    bool bad = false;
#pragma unroll
    for (int q = 0; q + 1 < E; q++) bad |= w[q + 1] < w[q];
    const unsigned long long nw = lane_next(w[0]);
    bad |= (lane + 1 < W) && nw < w[E - 1];
    return __any(bad);
  };
  auto in_lane = [&](int q) {
    const unsigned long long a = w[q], b = w[q + 1];
    const bool sw = b < a;
    w[q] = sw ? b : a;
    w[q + 1] = sw ? a : b;
  };
  // pair (slot E-1 of lane l, slot 0 of lane l+1) for every l with en(l)
  auto cross = [&](auto en) {
    const unsigned long long nw = lane_next(w[0]), pw = lane_prev(w[E - 1]);
    const bool lo = (lane + 1 < W) && en(lane) && nw < w[E - 1];
    const bool hi = (lane > 0) && en(lane - 1) && w[0] < pw;
    if (lo) w[E - 1] = nw;
    if (hi) w[0] = pw;
  };
  for (int it = 0; it < W * E && unsorted(); it++) {
    if constexpr (E == 1) {
      cross([](int l) { return (l & 1) == 0; });
      cross([](int l) { return (l & 1) == 1; });
    } else {
#pragma unroll
      for (int q = 0; q + 1 < E; q += 2) in_lane(q);
#pragma unroll
      for (int q = 1; q + 1 < E; q += 2) in_lane(q);
      cross([](int) { return true; });
    }
  }
}

// The FP64 form's exact (key, slot) order inside runs of equal truncated
// keys: odd-even transposition on (double key, slot) pairs, as above.
__device__ __forceinline__ double lane_next(double x) {
  return __longlong_as_double((long long)lane_next((unsigned long long)__double_as_longlong(x)));
}
__device__ __forceinline__ double lane_prev(double x) {
  return __longlong_as_double((long long)lane_prev((unsigned long long)__double_as_longlong(x)));
}
__device__ __forceinline__ bool kd_pair_lt(double ka, int ia, double kb, int ib) {
  return (ka < kb) | ((ka == kb) & (ia < ib));
}
template <int E>
__device__ __forceinline__ void exact_fixup_d(double (&kx)[E], int (&ex)[E], int lane) {
  auto unsorted = [&]() {
    bool bad = false;
#pragma unroll
    for (int q = 0; q + 1 < E; q++) bad |= kd_pair_lt(kx[q + 1], ex[q + 1], kx[q], ex[q]);
    const double nk = lane_next(kx[0]);
    const int ne = lane_next(ex[0]);
    bad |= (lane + 1 < W) && kd_pair_lt(nk, ne, kx[E - 1], ex[E - 1]);
    return __any(bad);
  };
  auto in_lane = [&](int q) {
    if (kd_pair_lt(kx[q + 1], ex[q + 1], kx[q], ex[q])) {
      const double tk = kx[q];
      const int te = ex[q];
      kx[q] = kx[q + 1];
      ex[q] = ex[q + 1];
      kx[q + 1] = tk;
      ex[q + 1] = te;
    }
  };
  auto cross = [&]() {
    const double nk = lane_next(kx[0]), pk = lane_prev(kx[E - 1]);
    const int ne = lane_next(ex[0]), pe = lane_prev(ex[E - 1]);
    const bool lo = (lane + 1 < W) && kd_pair_lt(nk, ne, kx[E - 1], ex[E - 1]);
    const bool hi = (lane > 0) && kd_pair_lt(kx[0], ex[0], pk, pe);
    if (lo) {
      kx[E - 1] = nk;
      ex[E - 1] = ne;
    }
    if (hi) {
      kx[0] = pk;
      ex[0] = pe;
    }
  };
  for (int it = 0; it < W * E && unsorted(); it++) {
#pragma unroll
    for (int q = 0; q + 1 < E; q += 2) in_lane(q);
#pragma unroll
    for (int q = 1; q + 1 < E; q += 2) in_lane(q);
    cross();
  }
}

// ------------------------------------------------ FP64 wave scans by DPP
template <int CTRL, int RM, bool BC>
__device__ __forceinline__ double dpp_d(double x) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, RM, 0xF, BC);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, RM, 0xF, BC);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                          (unsigned long long)(unsigned)lo));
}
// inclusive prefix sum over the 64 lanes: row shifts 1, 2, 4, 8 (zero fill),
// then lane 15 of each row into the next row and lane 31 into rows 2 and 3
__device__ __forceinline__ double wave_incl_sum(double s) {
  s += dpp_d<0x111, 0xF, true>(s);
  s += dpp_d<0x112, 0xF, true>(s);
  s += dpp_d<0x114, 0xF, true>(s);
  s += dpp_d<0x118, 0xF, true>(s);
  s += dpp_d<0x142, 0xA, false>(s);
  s += dpp_d<0x143, 0xC, false>(s);
  return s;
}
// value of lane - 1 (0 in lane 0)
__device__ __forceinline__ double wave_shr1(double s) { return dpp_d<0x138, 0xF, true>(s); }
// A DPP move of a double in which lanes without a source in the pattern, or
// in rows the row mask leaves out, keep `old`.
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_dk(double old, double x) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const unsigned long long o = (unsigned long long)__double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)u, CTRL, RM, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(u >> 32), CTRL,
                                             RM, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                          (unsigned long long)(unsigned)lo));
}
// lane l's double, l uniform
__device__ __forceinline__ double read_lane_d(double x, int l) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)u, l);
  const int hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                          (unsigned long long)(unsigned)lo));
}
// inclusive prefix product over the 64 lanes, the pattern of wave_incl_sum
// with 1 for lanes that have no source
__device__ __forceinline__ double wave_incl_prod(double s) {
  s *= dpp_dk<0x111, 0xF>(1.0, s);
  s *= dpp_dk<0x112, 0xF>(1.0, s);
  s *= dpp_dk<0x114, 0xF>(1.0, s);
  s *= dpp_dk<0x118, 0xF>(1.0, s);
  s *= dpp_dk<0x142, 0xA>(1.0, s);
  s *= dpp_dk<0x143, 0xC>(1.0, s);
  return s;
}
// Inclusive suffix composition over the 64 lanes of the affine maps
// x -> a + b x (lane l's map applied after lane l+1's): row shifts left 1, 2,
// 4, 8 within each 16-lane row (identity where there is no source), then
// each row's lanes compose with the totals of the rows after it, read from
// the rows' first lanes.
__device__ __forceinline__ void wave_suffix_affine(double &a, double &b, int lane) {
#define RADG_SUFFIX_STEP(CT)                                 \
  {                                                          \
    const double oa = dpp_dk<CT, 0xF>(0.0, a);               \
    const double ob = dpp_dk<CT, 0xF>(1.0, b);               \
    a = a + b * oa;                                          \
    b = b * ob;                                              \
  }
  RADG_SUFFIX_STEP(0x101)
  RADG_SUFFIX_STEP(0x102)
  RADG_SUFFIX_STEP(0x104)
  RADG_SUFFIX_STEP(0x108)
#undef RADG_SUFFIX_STEP
  const double a1 = read_lane_d(a, 16), b1 = read_lane_d(b, 16);
  const double a2 = read_lane_d(a, 32), b2 = read_lane_d(b, 32);
  const double a3 = read_lane_d(a, 48), b3 = read_lane_d(b, 48);
  // carries: row 2 takes row 3's total, row 1 rows 2 and 3, row 0 rows 1-3
  const double c1a = a2 + b2 * a3, c1b = b2 * b3;
  const double c0a = a1 + b1 * c1a, c0b = b1 * c1b;
  const int row = lane >> 4;
  const double ca = row == 0 ? c0a : row == 1 ? c1a : row == 2 ? a3 : 0.0;
  const double cb = row == 0 ? c0b : row == 1 ? c1b : row == 2 ? b3 : 1.0;
  a = a + b * ca;
  b = b * cb;
}
template <typename T> __device__ __forceinline__ T dpp_t(T x);
template <> __device__ __forceinline__ float dpp_t<float>(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x138, 0xF, 0xF, true));
}
template <> __device__ __forceinline__ double dpp_t<double>(double x) { return wave_shr1(x); }

// ----------------------------------------------------------- CIA, FP64
// nciacon.f:287-296 and nparacon_all.f:312-320: N2's near-infrared N2-N2
// and N2-H2 terms, the coefficients interpolated per wavenumber on the host
__device__ __forceinline__ void cia_n2nir(const double *n2, double qh2, double qn2, double &sum,
                                          double *dab) {
  const double an2n2 = n2[0], an2h2 = n2[1];
  sum += an2n2 * qn2 * qn2;
  dab[2] += 2 * qn2 * an2n2;
  sum += an2h2 * qn2 * qh2;
  dab[2] += qh2 * an2h2;
  dab[0] += qn2 * an2h2;
}

// ciaread.f:104-196 + nciacon.f:209-303 for one layer.  dab[0..6] as
// DABSORB (already multiplied by the path factor tau).  n2: this
// wavenumber's near-infrared N2 coefficients.
__device__ __forceinline__ void cia_layer(const KernelArgs &a, double vv, int inormal, const LayerConst &lc,
                                          const double *n2, double *absorb, double *dab) {
  double alpha[9], dal[9];
#pragma unroll
  for (int i = 0; i < 9; i++) {
    alpha[i] = 0.0;
    dal[i] = 0.0;
  }
  const double vhi = 1500.0 * a.cia_dnu;
  if (!(vv > vhi || vv < 0.0)) {
    const double *tt = a.cia_temps;
    double temp1 = lc.temp;
    if (temp1 > tt[24]) temp1 = tt[24];
    if (temp1 < tt[0]) temp1 = tt[0];
    // pindex over (i-1)*dnu and over the temperature grid
    int cv = (int)floor(vv / a.cia_dnu) + 1;
    if (cv > 1500) cv = 1500;
    if (cv < 1) cv = 1;
    // pindex keeps the last i with x1 >= x(i): correct a floor that landed
    // one cell high through rounding
    while (cv > 1 && vv < (cv - 1) * a.cia_dnu) cv--;
    while (cv < 1500 && vv >= cv * a.cia_dnu) cv++;
    int ct = 1;
    for (int i = 1; i <= 24; i++) {
      if (temp1 >= tt[i - 1]) ct = i;
      else break;
    }
    const double v_lo = (cv - 1) * a.cia_dnu, v_hi = cv * a.cia_dnu;
    const double t = (vv - v_lo) / (v_hi - v_lo);
    const double u = (temp1 - tt[ct - 1]) / (tt[ct] - tt[ct - 1]);
    const double dudt = 1.0 / (tt[ct] - tt[ct - 1]);
    const double *k11 = a.kcia + ((long long)(cv - 1) * 25 + (ct - 1)) * 9;
    const double *k12 = a.kcia + ((long long)(cv)*25 + (ct - 1)) * 9;
    const double *k22 = a.kcia + ((long long)(cv)*25 + (ct)) * 9;
    const double *k21 = a.kcia + ((long long)(cv - 1) * 25 + (ct)) * 9;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const double a11 = k11[i], a12 = k12[i], a22 = k22[i], a21 = k21[i];
      alpha[i] = (1.0 - t) * (1.0 - u) * a11 + t * (1.0 - u) * a12 + t * u * a22 +
                 (1.0 - t) * u * a21;
      dal[i] = (-(1.0 - t) * a11 - t * a12 + t * a22 + (1.0 - t) * a21) * dudt;
    }
  }
  const double qh2 = lc.qh2, qhe = lc.qhe, qn2 = lc.qn2, qch4 = lc.qch4, qco2 = lc.qco2;
  double sum = 0.0;
  for (int i = 0; i < 7; i++) dab[i] = 0.0;
  // equilibrium (inormal = 0) or normal (3:1) ortho:para H2 pairs
  const int p0 = (inormal == 0) ? 0 : 2;
  sum += alpha[p0] * qh2 * qh2 + alpha[p0 + 1] * qh2 * qhe;
  dab[0] += 2 * qh2 * alpha[p0] + qhe * alpha[p0 + 1];
  dab[1] += qh2 * alpha[p0 + 1];
  dab[5] += qh2 * qh2 * dal[p0] + qh2 * qhe * dal[p0 + 1];
  sum += alpha[4] * qh2 * qn2;
  dab[0] += qn2 * alpha[4];
  dab[2] += qh2 * alpha[4];
  dab[5] += qh2 * qn2 * dal[4];
  sum += alpha[5] * qn2 * qch4;
  dab[2] += qch4 * alpha[5];
  dab[3] += qn2 * alpha[5];
  dab[5] += qn2 * qch4 * dal[5];
  sum += alpha[6] * qn2 * qn2;
  dab[2] += 2 * qn2 * alpha[6];
  dab[5] += qn2 * qn2 * dal[6];
  sum += alpha[7] * qch4 * qch4;
  dab[3] += 2 * qch4 * alpha[7];
  dab[5] += qch4 * qch4 * dal[7];
  sum += alpha[8] * qh2 * qch4;
  dab[0] += qch4 * alpha[8];
  dab[3] += qh2 * alpha[8];
  dab[5] += qh2 * qch4 * dal[8];
  // co2cia.f: non-zero only between 1.06 and 2.55 micron
  {
    const double xl = 1e4 / vv;
    double aco2 = 0.0;
    if (xl >= 2.15 && xl <= 2.55) aco2 = 4e-8;
    if (xl >= 1.7 && xl <= 1.76) aco2 = 6e-9;
    if (xl >= 1.25 && xl < 1.35) aco2 = 1.5e-9;
    if (xl >= 1.125 && xl <= 1.225) aco2 = 0.5 * (0.31 + 0.79) * 1e-9;
    if (xl >= 1.06 && xl < 1.125) aco2 = 0.5 * (0.29 + 0.67) * 1e-9;
    sum += aco2 * qco2 * qco2;
    dab[4] += 2 * qco2 * aco2;
  }
  cia_n2nir(n2, qh2, qn2, sum, dab);
  *absorb = sum * lc.tau;
  for (int i = 0; i < 7; i++) dab[i] *= lc.tau;
}

// CIA with a para-H2 fraction per layer (FLAGH2P = 1): nparacon_all.f with
// fpread_all.f:168-257 on the registered /FPTABLE/.  Only the two para
// fractions the layer's fraction falls between are interpolated in (v, T)
// (each fraction's value is independent of the others).  The fraction index
// follows the reference: x1 = 1 + (fpara - 0.25)/deltaf, j = INT(x1) kept in
// [1, npara-2] with f = 0 below and f = 1 above; fpara < 0 takes the last
// fraction and has no derivative.  dab[6] = d(tau)/d(fpara).
__device__ __forceinline__ void cia_para_layer(const KernelArgs &a, double vv, const LayerConst &lc,
                                               const double *n2, double *absorb, double *dab) {
  const int np = a.npara;
  const double deltaf = a.pfp[1] - a.pfp[0];
  int j0, j1;
  double f = 0.0, has_dfp = 0.0;
  if (lc.fpara < 0.0) {
    j0 = j1 = np - 1;
  } else {
    const double x1 = 1.0 + (lc.fpara - 0.25) / deltaf;
    int j = (int)x1;
    f = x1 - j;
    if (j < 1) {
      j = 1;
      f = 0.0;
    } else if (j > np - 2) {
      j = np - 2;
      f = 1.0;
    }
    j0 = j - 1;
    j1 = j;
    has_dfp = 1.0;
  }
  double h2[2] = {0.0, 0.0}, he[2] = {0.0, 0.0}, dh2[2] = {0.0, 0.0}, dhe[2] = {0.0, 0.0};
  double a5[5], da5[5];
#pragma unroll
  for (int i = 0; i < 5; i++) a5[i] = da5[i] = 0.0;
  if (!(vv > 1500.0 * a.pdnu || vv < 0.0)) {
    const double *tt = a.ptemps;
    double temp1 = lc.temp;
    if (temp1 > tt[24]) temp1 = tt[24];
    if (temp1 < tt[0]) temp1 = tt[0];
    int cv = (int)floor(vv / a.pdnu) + 1;
    if (cv > 1500) cv = 1500;
    if (cv < 1) cv = 1;
    while (cv > 1 && vv < (cv - 1) * a.pdnu) cv--;
    while (cv < 1500 && vv >= cv * a.pdnu) cv++;
    int ct = 1;
    for (int i = 1; i <= 24; i++) {
      if (temp1 >= tt[i - 1]) ct = i;
      else break;
    }
    const double t = (vv - (cv - 1) * a.pdnu) / (cv * a.pdnu - (cv - 1) * a.pdnu);
    const double u = (temp1 - tt[ct - 1]) / (tt[ct] - tt[ct - 1]);
    const double dudt = 1.0 / (tt[ct] - tt[ct - 1]);
    const long long r11 = ((long long)(cv - 1) * 25 + (ct - 1)), r12 = ((long long)cv * 25 + (ct - 1));
    const long long r22 = ((long long)cv * 25 + ct), r21 = ((long long)(cv - 1) * 25 + ct);
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const int i = s ? j1 : j0;
      const double a11 = a.pkh2h2[r11 * np + i], a12 = a.pkh2h2[r12 * np + i];
      const double a22 = a.pkh2h2[r22 * np + i], a21 = a.pkh2h2[r21 * np + i];
      const double b11 = a.pkh2he[r11 * np + i], b12 = a.pkh2he[r12 * np + i];
      const double b22 = a.pkh2he[r22 * np + i], b21 = a.pkh2he[r21 * np + i];
      h2[s] = (1.0 - t) * (1.0 - u) * a11 + t * (1.0 - u) * a12 + t * u * a22 + (1.0 - t) * u * a21;
      he[s] = (1.0 - t) * (1.0 - u) * b11 + t * (1.0 - u) * b12 + t * u * b22 + (1.0 - t) * u * b21;
      dh2[s] = (-(1.0 - t) * a11 - t * a12 + t * a22 + (1.0 - t) * a21) * dudt;
      dhe[s] = (-(1.0 - t) * b11 - t * b12 + t * b22 + (1.0 - t) * b21) * dudt;
    }
#pragma unroll
    for (int i = 0; i < 5; i++) {
      const double a11 = a.pkcia5[r11 * 5 + i], a12 = a.pkcia5[r12 * 5 + i];
      const double a22 = a.pkcia5[r22 * 5 + i], a21 = a.pkcia5[r21 * 5 + i];
      a5[i] = (1.0 - t) * (1.0 - u) * a11 + t * (1.0 - u) * a12 + t * u * a22 + (1.0 - t) * u * a21;
      da5[i] = (-(1.0 - t) * a11 - t * a12 + t * a22 + (1.0 - t) * a21) * dudt;
    }
  }
  const double ah2h2 = (1.0 - f) * h2[0] + f * h2[1], ah2he = (1.0 - f) * he[0] + f * he[1];
  const double dh2dt = (1.0 - f) * dh2[0] + f * dh2[1], dhedt = (1.0 - f) * dhe[0] + f * dhe[1];
  const double dh2dfp = has_dfp * (h2[1] - h2[0]) / deltaf, dhedfp = has_dfp * (he[1] - he[0]) / deltaf;
  const double qh2 = lc.qh2, qhe = lc.qhe, qn2 = lc.qn2, qch4 = lc.qch4, qco2 = lc.qco2;
  double sum = ah2h2 * qh2 * qh2 + ah2he * qh2 * qhe;
  dab[0] = 2 * qh2 * ah2h2 + qhe * ah2he;
  dab[1] = qh2 * ah2he;
  dab[2] = dab[3] = dab[4] = 0.0;
  dab[5] = dh2dt * qh2 * qh2 + dhedt * qh2 * qhe;
  dab[6] = dh2dfp * qh2 * qh2 + dhedfp * qh2 * qhe;
  // H2-N2, N2-CH4, N2-N2, CH4-CH4, H2-CH4
  sum += a5[0] * qh2 * qn2;
  dab[0] += qn2 * a5[0];
  dab[2] += qh2 * a5[0];
  dab[5] += qh2 * qn2 * da5[0];
  sum += a5[1] * qn2 * qch4;
  dab[2] += qch4 * a5[1];
  dab[3] += qn2 * a5[1];
  dab[5] += qn2 * qch4 * da5[1];
  sum += a5[2] * qn2 * qn2;
  dab[2] += 2 * qn2 * a5[2];
  dab[5] += qn2 * qn2 * da5[2];
  sum += a5[3] * qch4 * qch4;
  dab[3] += 2 * qch4 * a5[3];
  dab[5] += qch4 * qch4 * da5[3];
  sum += a5[4] * qh2 * qch4;
  dab[0] += qch4 * a5[4];
  dab[3] += qh2 * a5[4];
  dab[5] += qh2 * qch4 * da5[4];
  {  // co2cia.f
    const double xl = 1e4 / vv;
    double aco2 = 0.0;
    if (xl >= 2.15 && xl <= 2.55) aco2 = 4e-8;
    if (xl >= 1.7 && xl <= 1.76) aco2 = 6e-9;
    if (xl >= 1.25 && xl < 1.35) aco2 = 1.5e-9;
    if (xl >= 1.125 && xl <= 1.225) aco2 = 0.5 * (0.31 + 0.79) * 1e-9;
    if (xl >= 1.06 && xl < 1.125) aco2 = 0.5 * (0.29 + 0.67) * 1e-9;
    sum += aco2 * qco2 * qco2;
    dab[4] += 2 * qco2 * aco2;
  }
  cia_n2nir(n2, qh2, qn2, sum, dab);
  *absorb = sum * lc.tau;
  for (int i = 0; i < 7; i++) dab[i] *= lc.tau;
}

// Rayleigh cross-section per molecule (cm2) at wavenumber vv (cm-1), in FP64
// for every form; pi is the reference's literal 3.1415927.
//   1: rayleighj.f (H2/He mixture, fH2 = 0.864, isotropic)
//   2: rayleighv.f (8.8e-28 / lambda_um^4)
//   3: rayleigha.f (air, refractivity corrected to the layer's P and T,
//      depolarisation factor faniso(0.035))
//   4: rayleighls.f (H2, He, CH4, NH3 in the layer's proportions)
__device__ __forceinline__ double rayleigh_xs(int iray, double vv, const LayerConst &lc) {
  const double pi3 = 3.1415927 * 3.1415927 * 3.1415927;
  if (iray == 2) {
    const double lum = 10000 / vv;
    return 8.8e-28 / ((lum * lum) * (lum * lum));
  }
  if (iray == 1) {
    const double lm = 1e-2 / vv;         // m
    const double x2 = (vv * 1e-4) * (vv * 1e-4);  // micron^-2
    const double n = 0.864 * (13.58e-5 * (1.0 + 7.52e-3 * x2)) +
                     (1 - 0.864) * (3.48e-5 * (1.0 + 2.30e-3 * x2));
    const double nl = (1.013e5 / (1.37971e-23 * 273.15)) * lm * lm;
    return 32 * pi3 * n * n * 1e4 / (3 * (nl * nl));
  }
  if (iray == 3) {
    const double lum = 1e4 / vv;
    const double nu2 = 1.0 / (lum * lum);
    const double p1 = lc.press * 1.013e5, t1 = lc.temp - 273.15;
    const double xcorr = p1 * (1 + p1 * (61.3 - t1) * 1e-10) / (96095.4 * (1 + 0.003661 * t1));
    const double m = (8342.12 + 2406030.0 / (130.0 - nu2) + 15997.0 / (38.9 - nu2)) * 1e-8 * xcorr;
    const double lm = lum * 1e-6;
    const double nl = p1 / (1.37971e-23 * lc.temp) * lm * lm;
    const double mm = m * (m + 2.0);
    const double fan = (6.0 + 3.0 * 0.035) / (6.0 - 7.0 * 0.035);
    return 8 * pi3 * (mm * mm) * 1e4 * fan / (3 * (nl * nl));
  }
  const double wl = 1e4 / vv, iwl2 = 1.0 / (wl * wl);
  const double h2 = (1.0 - lc.fnh3) / (1 + lc.fheh2 + lc.fch4h2);
  const double comp[4] = {h2, lc.fheh2 * h2, lc.fch4h2 * h2, lc.fnh3};
  const double A[4] = {13.58e-5, 3.48e-5, 37.0e-5, 37.0e-5};
  const double B[4] = {7.52e-3, 2.3e-3, 12.0e-3, 12.0e-3};
  const double D[4] = {0.0221, 0.025, .0922, .0922};
  double s = 0.0, w = 0.0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const double nr = 1.0 + A[i] * (1.0 + B[i] * iwl2);
    const double e = nr * nr - 1.0;
    s += e * e * comp[i] * (6 + 3 * D[i]) / (6 - 7 * D[i]);
    w += comp[i];
  }
  const double lo = 2.687e19 * 1e-12;
  return 8.0 * pi3 / (3.0 * ((wl * wl) * (wl * wl)) * (lo * lo)) * 1e-8 * s / w;
}

// H- free-free absorption coefficient (cm4 dyne-1) at wavenumber v0 and
// temperature t: John (1988) / Bell & Berrington (1987) tables 3a (above
// 0.3645 micron) and 3b (0.1823-0.3645 micron) as hmin_ff.f:44-67 sums them,
// each call from a zero sum; negative sums give 0.
__device__ __forceinline__ double hmin_ff_d(double v0, double t) {
  const double wv = 1e4 / v0, th = 5040.0 / t;
  const double w2 = wv * wv, w3 = w2 * wv, w4 = w2 * w2;
  double s = 0.0;
  if (wv > 0.3645) {
    const double A[6] = {0.0, 2483.346, -3449.889, 2200.040, -696.271, 88.283};
    const double B[6] = {0.0, 285.827, -1158.382, 2427.719, -1841.400, 444.517};
    const double C[6] = {0.0, -2054.291, 8746.523, -13651.105, 8624.970, -1863.864};
    const double D[6] = {0.0, 2827.776, -11485.632, 16755.524, -10051.530, 2095.288};
    const double E[6] = {0.0, -1341.537, 5303.609, -7510.494, 4400.067, -901.788};
    const double F[6] = {0.0, 208.952, -812.939, 1132.738, -655.020, 132.985};
#pragma unroll
    for (int k = 0; k < 6; k++)
      s += pow(th, 0.5 * (k + 2)) * (w2 * A[k] + B[k] + C[k] / wv + D[k] / w2 + E[k] / w3 + F[k] / w4);
  } else if (wv < 0.3645 && wv > 0.1823) {
    const double A[4] = {518.1021, 473.2636, -482.2089, 115.5291};
    const double B[4] = {-734.8666, 1443.4137, -737.1616, 169.6374};
    const double C[4] = {1021.1775, -1977.3395, 1096.8827, -245.649};
    const double D[4] = {-479.0721, 922.3575, -521.1341, 114.243};
    const double E[4] = {93.1373, -178.9275, 101.7963, -21.9972};
    const double F[4] = {-6.4285, 12.3600, -7.0571, 1.5097};
#pragma unroll
    for (int k = 0; k < 4; k++)
      s += pow(th, 0.5 * (k + 2)) * (w2 * A[k] + B[k] + C[k] / wv + D[k] / w2 + E[k] / w3 + F[k] / w4);
  }
  return s < 0.0 ? 0.0 : 1e-29 * s;
}

// ngascon.f:136-262 for gas k of layer j at wavenumber vv: its absorption
// (an optical depth).  The wavenumber-only factors come from the host (gw);
// the H2O continuum's temperature scaling and the H- free-free term are
// formed here.
__device__ __forceinline__ double gascon_d(const KernelArgs &a, const DevReq &R, int k, int j,
                                           double vv, const LayerConst &lc, const double *gw) {
  const int kind = R.gcon[k];
  const long long lg = R.lg0 + (long long)j * R.ngas;
  const double amo = a.amo[lg + k], amount = amo * 1e20;
  if (kind == RADG_GC_CH4) return amo * gw[RADG_GCW_CH4];
  if (kind == RADG_GC_NH3) return amo * gw[RADG_GCW_NH3];
  if (kind == RADG_GC_HBF) return amount * gw[RADG_GCW_HBF];
  if (kind == RADG_GC_HFF) {
    // electron.f state as cirsradg_wave.f:590-599 leaves it for this layer
    const double pelec = a.ppg[lg + R.ke80], pph = a.ppg[lg + R.ke91];
    const double amh = a.amo[lg + R.ke91] * 1e20;
    const double nh = (a.ppg[lg + k] * 1.013e6 * pelec) / pph;
    return amh * hmin_ff_d(vv, lc.temp) * nh;
  }
  if (kind == RADG_GC_H2O && gw[RADG_GCW_H2OIN] != 0.0) {
    // ngascon.f:162-207
    const double t = lc.temp, ppress = a.ppg[lg + k];
    const double s96 = gw[RADG_GCW_S296];
    const double tfac = (t - 296.0) / (260.0 - 296.0);
    const double csfft = gw[RADG_GCW_SFAC] * (s96 * pow(gw[RADG_GCW_S260] / s96, tfac));
    const double c11 = 0.5 * a.h2o_c2 / t;
    const double a1 = vv * amount * tanh(c11 * vv);
    const double a2 = a.h2o_ts / (t * a.h2o_ps);
    const double a3 = 1.0e-20 * (ppress * csfft + (lc.press - ppress) * gw[RADG_GCW_F296]);
    return a1 * a2 * a3;
  }
  return 0.0;
}

// 16-bit log10(k) codes (BF: 1 BF16, 2 IEEE half): log10 k = code + off
template <int BF>
__device__ __forceinline__ bool code_nonpos(unsigned short c) {
  return c == (BF == 1 ? RADG_BF16_NONPOS : RADG_F16_NONPOS);
}
template <int BF>
__device__ __forceinline__ float code_value(unsigned short c) {
  if constexpr (BF == 1) return __uint_as_float((unsigned int)c << 16);
  else return __half2float(__ushort_as_half(c));
}

// k and dk/dT of one g-ordinate gg of the table slice at element offset
// base (16-bit forms: offset row ob), bilinear in (ln P, T) between the
// corners i1..i4 (get_kg.f:319-376, get_klblg.f:231-301): in ln k when the
// first ordinate of all four corners is > 0 (KLOG), else in k; on one
// temperature grid, or per pressure row on each row's own grid (NT < 0).
// A NaN result becomes 1e-37 (get_kg.f:455-470).
template <typename T, int BF>
__device__ __forceinline__ void interp_point(const KernelArgs &a, long long base, long long ob, int i1,
                                             int i2, int i3, int i4, int gg, T v, T u, T dudt, T u2,
                                             T dudt2, T &kk, T &dk) {
  bool klog;
  T y1, y2, y3, y4;
  if (BF) {
    // 16-bit log10(k) codes: ln k comes straight from the code
    const unsigned short *sl = reinterpret_cast<const unsigned short *>(a.ktab) + base;
    klog = !(code_nonpos<BF>(sl[i1]) || code_nonpos<BF>(sl[i2]) || code_nonpos<BF>(sl[i3]) ||
             code_nonpos<BF>(sl[i4]));
    const float ln10 = 2.302585092994046f;
    const float offg = a.koff[ob + gg];
    auto dec = [&](unsigned short c) -> float {
      if (code_nonpos<BF>(c)) return klog ? -__builtin_huge_valf() : 0.0f;
      const float l = (code_value<BF>(c) + offg) * ln10;
      return klog ? l : expf(l);
    };
    y1 = (T)dec(sl[i1 + gg]);
    y2 = (T)dec(sl[i2 + gg]);
    y3 = (T)dec(sl[i3 + gg]);
    y4 = (T)dec(sl[i4 + gg]);
  } else {
    const T *sl = reinterpret_cast<const T *>(a.ktab) + base;
    klog = !(sl[i1] <= (T)0 || sl[i2] <= (T)0 || sl[i3] <= (T)0 || sl[i4] <= (T)0);
    y1 = sl[i1 + gg];
    y2 = sl[i2 + gg];
    y3 = sl[i3 + gg];
    y4 = sl[i4 + gg];
    if (klog) {
      y1 = dlog(y1);
      y2 = dlog(y2);
      y3 = dlog(y3);
      y4 = dlog(y4);
    }
  }
  T x, dxdt;
  if (!a.t2) {
    x = ((T)1 - v) * ((T)1 - u) * y1 + v * ((T)1 - u) * y2 + v * u * y3 + ((T)1 - v) * u * y4;
    dxdt = (-((T)1 - v) * y1 - v * y2 + v * y3 + ((T)1 - v) * y4) * dudt;
  } else {
    x = ((T)1 - v) * (((T)1 - u) * y1 + u * y4) + v * (((T)1 - u2) * y2 + u2 * y3);
    dxdt = ((T)1 - v) * (-y1 + y4) * dudt + v * (-y2 + y3) * dudt2;
  }
  if (klog) {
    const T ex = dexp(x);
    kk = ex;
    dk = ex * dxdt;
  } else {
    kk = x;
    dk = dxdt;
  }
  if (kk != kk) kk = (T)1e-37;
  if (dk != dk) dk = (T)1e-37;
}

// ============================================================ layer pass
// LDS of a layer item (layer_layout in radg_dev.h):
//   header (per workgroup)  G[NG+1], delg[NG] (FP64), delg[NG] (T)
//   per wavefront           sSt, sGd [NG] (FP64)  the straddling element's
//                           start and end cumulative weight, per bin
//                           kgn, dkgn [ngas][NG]  interpolated k, dk/dT
//                           bufA, bufB [NS][NG]   overlap row state
//                           sE [NG]               straddling element, per bin
//                           uni                   rankk merge buffers | exact
//                                                 FP32 keys | rebin matrix,
//                                                 bin codes, straddle shares
// with NS = ngas + 2 row slots (K, gas 0..ngas-1, T).

// Orders the LDS accesses of one wavefront: LDS instructions of a wavefront
// complete in issue order, so a compiler barrier is all that lanes reading
// each other's writes need.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Fills the workgroup header: G (g boundaries, G[NG] >= 1, rankg.f:59-65,
// then +inf), the g weights in FP64 and in T, the FP64 weight delg_i delg_j
// of the element in each sort slot (0 for padding) and the cumulative
// weights (rankg.f:59-62).  The host builds it once per table set and form;
// a workgroup copies it, 16 bytes per thread and step.
template <typename T, int E>
__device__ __forceinline__ void layer_header(const KernelArgs &a, char *hdr, int tid, int nthr) {
  const int n16 = layer_header_bytes(a.ng, E, (int)sizeof(T)) >> 4;
  const uint4 *src = reinterpret_cast<const uint4 *>(a.hdr_img);
  uint4 *dst = reinterpret_cast<uint4 *>(hdr);
  for (int i = tid; i < n16; i += nthr) dst[i] = src[i];
}

// ---- row-count rebin of one overlap round --------------------------------
// rankg.f:110-152 walks the sorted elements, adds whole elements to bin ig
// while their cumulative weight stays below the edge G[ig+1], splits the
// element that reaches it (frac below, 1 - frac above) and divides each
// closed bin by its accumulated weight.  So bin b-1 is Q(b) - Q(b-1) over
// its weight, with Q(b) the weighted sum over the elements wholly below
// edge b plus frac_b times the straddler's weighted value.
//
// With the new gas's ordinates non-decreasing (k2_j <= k2_{j+1}) the keys
// of row i, K_i + a2 k2_j, leave the sort in column order, so the elements
// of row i wholly below edge b are its first c_bi columns: c_bi counts the
// row's elements whose gamma (the number of edges at or below the
// cumulative weight through the element) is below b.  Every value is a row
// term plus a column term (noverlapg.f:117-128, 175-187), hence
//   Q(b) = sum_i delg_i (R_i Gw[c_bi] + C[c_bi]) + frac_b w_s v_s,
// with Gw and C prefix sums over the columns.  Lanes own the edges (ngrp
// groups of NG lanes, each over a share of the rows); c_bi is four bytes
// of gamma at a time: x + (0x80 - b) per byte sets the byte's top bit
// exactly when gamma >= b.  All sums are FP64.  Returns false, leaving nxt
// untouched, when an element's weight reaches past two edges: rankg.f
// closes only one bin per element there, which the record rebin follows.
template <int E> struct CountShape;   // RPG: rows per group at most
template <> struct CountShape<4> { static constexpr int RPG = 4; };
template <> struct CountShape<8> { static constexpr int RPG = 7; };
template <> struct CountShape<16> { static constexpr int RPG = 16; };

// The form applies with 4, 8 or 16 sort slots per lane when a row's slots
// start 8-byte aligned and its NG gamma bytes take 4, 5, 6 or 8 words (E =
// 4: 13-16 g-ordinates, E = 8: 17-21, E = 16: 22-24 and 29-32).
template <int E>
__device__ __forceinline__ bool counts_fit(int NG) {
  if constexpr (E < 4 || E > 16) {
    return false;
  } else {
    const int lpr = layer_lpr(NG, E), ngrp = count_groups(NG), nd = (NG + 3) >> 2;
    return NG <= 32 && ngrp >= 1 && (lpr * E) % 8 == 0 && 4 * nd <= lpr * E &&
           (nd == 4 || nd == 5 || nd == 6 || nd == 8) &&
           (NG + ngrp - 1) / ngrp <= CountShape<E>::RPG;
  }
}

// Once per item: the row weights delg_i and the zero rows of rrec.
template <typename T, int E>
__device__ __forceinline__ void counts_init(const int NG, const int ngas, const double *dl64,
                                            char *uni, int lane) {
  const CountLayout cl = count_layout(NG, ngas, E, uni_keep(NG, ngas, E, (int)sizeof(T)));
  double *rrec = reinterpret_cast<double *>(uni + cl.rrec);
  constexpr int RPG = CountShape<E>::RPG;
  const int RR = ngas + 3, ngrp = count_groups(NG), rpg = (NG + ngrp - 1) / ngrp;
  for (int u = lane; u < count_rows(NG, E); u += W) {
    const int g2 = u / RPG, t = u - g2 * RPG, i = g2 * rpg + t;
    const bool ri = t < rpg && i < NG;
    double *rr = rrec + u * RR;
    rr[0] = ri ? dl64[i] : 0.0;
    if (!ri)
      for (int x = 1; x < RR; x++) rr[x] = 0.0;
  }
}

template <typename T, int E, int ND>
__device__ __forceinline__ bool rebin_counts(const int NG, const int ngas, const int r, const T a2,
                                             const int (&idx)[E], const T *rows, T *nxt,
                                             const T *k2, const T *d2, const double *G,
                                             const double *dl64, const double *w64,
                                             const double *Gw, char *uni, double *recGs,
                                             double *recGe, int lane_in) {
  constexpr int RPG = CountShape<E>::RPG;
  // an opaque copy of the lane number keeps the lane-dependent indices
  // below from being hoisted out of the round loop (and spilled there)
  int lane = lane_in;
  asm volatile("" : "+v"(lane));
  const int NS = ngas + 2, RR = ngas + 3;
  const int lpr = layer_lpr(NG, E);
  const CountLayout cl = count_layout(NG, ngas, E, uni_keep(NG, ngas, E, (int)sizeof(T)));
  double *rrec = reinterpret_cast<double *>(uni + cl.rrec);
  unsigned char *gam = reinterpret_cast<unsigned char *>(uni + cl.gam);
  double *ctab = reinterpret_cast<double *>(uni + cl.ctab);
  double *part = reinterpret_cast<double *>(uni + cl.part);
  int *recI = reinterpret_cast<int *>(uni + cl.recI);

  // cumulative weight through each slot, gdist of rankg.f:104-108
  double gd[E];
  double run = 0.0;
#pragma unroll
  for (int q = 0; q < E; q++) {
    run += w64[idx[q]];
    gd[q] = run;
  }
  {
    const double ex = wave_shr1(wave_incl_sum(run));
#pragma unroll
    for (int q = 0; q < E; q++) gd[q] += ex;
  }
  const double gprev = wave_shr1(gd[E - 1]);
  // gamma: b0 edges of G[1..NG] at or below the lane's start (a search),
  // then for each slot b0, plus 1 once its end reaches the next edge
  // G[b0 + 1].  The slot where that first happens straddles the edge
  // (rankg.f:120).  A lane whose end reaches G[b0 + 2] as well (two edges
  // within a lane, or within one element) hands the round to the record
  // rebin; with Gauss-Legendre weights a lane's 8 elements span a bin edge
  // at most once.  Padding slots write gamma bytes past a row's NG real
  // ones, which the counts below mask.
  int b0 = zero_i();
#pragma unroll
  for (int st = 32; st > 0; st >>= 1)
    if (b0 + st <= NG && G[b0 + st] <= gprev) b0 += st;
  const double ge0 = G[b0 + 1], ge1 = G[b0 + 2];
  const bool bad = ge1 <= gd[E - 1];
  bool sprev = false;
#pragma unroll
  for (int q = 0; q < E; q++) {
    const bool sq = ge0 <= gd[q];
    if (sq && !sprev) {  // edge b0 + 1, recorded at b0
      recI[b0] = idx[q];
      recGs[b0] = (q == 0) ? gprev : gd[q - 1];
      recGe[b0] = gd[q];
    }
    sprev = sq;
    gam[idx[q]] = (unsigned char)(b0 + (sq ? 1 : 0));
  }
  if (__any(bad)) return false;
  // edges at or below the total weight
  const int gtot = __builtin_amdgcn_readlane(b0 + (sprev ? 1 : 0), W - 1);

  // column prefix tables: the header's Gw and the new gas's two sums
  {
    const bool jl = lane < NG;
    const double xk = jl ? dl64[lane] * (double)k2[lane] : 0.0;
    const double xd = jl ? dl64[lane] * (double)d2[lane] : 0.0;
    const double ik = wave_incl_sum(xk), id = wave_incl_sum(xd);
    if (jl) {
      ctab[3 * lane + 3] = Gw[lane + 1];
      ctab[3 * lane + 4] = ik;
      ctab[3 * lane + 5] = id;
    }
    if (lane == 0) {
      const double z = zero_d();
      ctab[0] = z;
      ctab[1] = z;
      ctab[2] = z;
    }
  }
  // lane = edge bl + 1 of row group grp, rows i0 + t, t < RPG: rows with
  // t >= rpg or past NG are the zero rows of rrec, whose gamma bytes are
  // whatever follows (their counts are clamped to NG)
  const int ngrp = count_groups(NG);
  const int rpg = (NG + ngrp - 1) / ngrp;
  // (1/NG from an opaque copy of NG, so that it is formed here rather than
  // merged with the round's and held across the sort)
  int ng_g = NG;
  asm volatile("" : "+s"(ng_g));
  const int grp = (int)(((float)lane + 0.5f) * rcp_idx(ng_g));
  const int bl = lane - (int)__umul24(grp, NG);
  const bool act = grp < ngrp;
  const int i0 = act ? (int)__umul24(grp, rpg) : 0;
  // this round's row terms: delg_i K_i, delg_i dK_i/dT and delg_i P_si for
  // the gases already mixed (row i at slot (i / rpg) RPG + i % rpg)
  if (lane < NG) {
    const int g2 = (int)(((float)lane + 0.5f) * rcp_idx(rpg));
    double *rr = rrec + __umul24(__umul24(g2, RPG) + lane - __umul24(g2, rpg), RR);
    const double w = dl64[lane];
    rr[1] = w * (double)rows[lane];
    rr[2] = w * (double)rows[(NS - 1) * NG + lane];
    for (int s = 0; s < r; s++) rr[3 + s] = w * (double)rows[(1 + s) * NG + lane];
  }
  wave_sync();

  // counts: 4 bytes of gamma per word (ND words, ND = ceil(NG / 4)),
  // (x + 0x80 - b) per byte sets the byte's top bit exactly when gamma >= b;
  // the last word's bytes past NG are forced to 0x7f (>= every edge)
  const unsigned pm = (NG & 3) ? (0x7f7f7f7fu << (8 * (NG & 3))) : 0u;
  const unsigned cb = 0x80808080u - (unsigned)(bl + 1) * 0x01010101u;
  double Wt = 0.0, V = 0.0, Z = 0.0, QK = 0.0, QT = 0.0;
  double gi[RPG];
  // LDS pointers with 32-bit offsets, strides as 24-bit products (full rate)
  typedef __attribute__((address_space(3))) const unsigned long long lds_u64;
  typedef __attribute__((address_space(3))) const double lds_f64;
  const int rstride = lpr * E;  // bytes per gamma row
  lds_u64 *rowg0 = (lds_u64 *)(gam + __umul24(i0, rstride));
  lds_f64 *rr0 = (lds_f64 *)(rrec + __umul24(act ? grp : 0, RPG * RR));
  lds_f64 *ctl = (lds_f64 *)ctab;
#pragma unroll
  for (int t = 0; t < RPG; t++) {
    lds_u64 *rowg = rowg0 + __umul24(t, rstride >> 3);
    unsigned ge = 0;  // bytes >= edge
#pragma unroll
    for (int d = 0; d < ND; d += 2) {
      const unsigned long long v = rowg[d >> 1];
      unsigned lo = (unsigned)v;
      if (d == ND - 1) lo |= pm;
      ge += (unsigned)__popc((lo + cb) & 0x80808080u);
      if (d + 1 < ND) {
        unsigned hi = (unsigned)(v >> 32);
        if (d + 1 == ND - 1) hi |= pm;
        ge += (unsigned)__popc((hi + cb) & 0x80808080u);
      }
    }
    const int c = min(4 * ND - (int)ge, NG);
    lds_f64 *ct = ctl + __umul24(c, 3);
    const double gw = ct[0];
    lds_f64 *rr = rr0 + __umul24(t, RR);
    Wt += rr[0] * gw;
    V += rr[0] * ct[1];
    Z += rr[0] * ct[2];
    QK += rr[1] * gw;
    QT += rr[2] * gw;
    gi[t] = gw;
  }
  wave_sync();  // gam and ctab are read: the partial sums take their place
  if (act && grp > 0) {
    double *pp = part + __umul24(__umul24(grp - 1, NG) + bl, 5);
    pp[0] = Wt;
    pp[1] = V;
    pp[2] = Z;
    pp[3] = QK;
    pp[4] = QT;
  }
  wave_sync();
  if (grp == 0) {
    for (int g2 = 1; g2 < ngrp; g2++) {
      const double *pp = part + __umul24(__umul24(g2 - 1, NG) + bl, 5);
      Wt += pp[0];
      V += pp[1];
      Z += pp[2];
      QK += pp[3];
      QT += pp[4];
    }
  }
  // the straddler of edge bl + 1 and its share below the edge
  const bool blane = lane < NG;
  const bool has = blane && (bl + 1 <= gtot);
  double f = 0.0;
  int is = 0;
  double QKf = QK + (double)a2 * V, QNf = V, QTf = QT + (double)a2 * Z, Wtf = Wt;
  if (has) {
    const int s = recI[bl];
    const int l = s / E;
    is = (int)(((float)l + 0.5f) * rcp_idx(lpr));
    const int js = (l - (int)__umul24(is, lpr)) * E + (s & (E - 1));
    const double gs = recGs[bl];
    f = (G[bl + 1] - gs) * rcp_nr(recGe[bl] - gs) * w64[s];
    const T kv = rows[is] + k2[js] * a2;
    const T tv = rows[(NS - 1) * NG + is] + d2[js] * a2;
    Wtf += f;
    QKf += f * (double)kv;
    QNf += f * (double)k2[js];
    QTf += f * (double)tv;
  }
  // bin bl: reached when it is the first or its lower edge was straddled,
  // divided by its weight when closed by a straddle or the last bin
  const double pW = wave_shr1(Wtf), pK = wave_shr1(QKf), pN = wave_shr1(QNf), pT = wave_shr1(QTf);
  const bool reached = blane && (bl == 0 || bl <= gtot);
  const bool norm = (bl + 1 <= gtot) || bl == NG - 1;
  double inv = 1.0;
  if (reached && norm) inv = rcp_nr(Wtf - pW);
  if (blane) {
    nxt[bl] = reached ? (T)((QKf - pK) * inv) : (T)0;
    nxt[(1 + r) * NG + bl] = reached ? (T)((QNf - pN) * inv) : (T)0;
    nxt[(NS - 1) * NG + bl] = reached ? (T)((QTf - pT) * inv) : (T)0;
    if (r + 1 < ngas) {
      const T z = (T)zero_d();
      for (int s4 = r + 1; s4 < ngas; s4++) nxt[(1 + s4) * NG + bl] = z;
    }
  }
  // the payloads of the gases already mixed: row terms only
  for (int s = 0; s < r; s++) {
    double qv = 0.0;
#pragma unroll
    for (int t = 0; t < RPG; t++) qv += rr0[__umul24(t, RR) + 3 + s] * gi[t];
    wave_sync();  // the previous partials are read
    if (act && grp > 0) part[__umul24(grp - 1, NG) + bl] = qv;
    wave_sync();
    if (grp == 0)
      for (int g2 = 1; g2 < ngrp; g2++) qv += part[__umul24(g2 - 1, NG) + bl];
    if (has) qv += f * (double)rows[(1 + s) * NG + is];
    const double pq = wave_shr1(qv);
    if (blane) nxt[(1 + s) * NG + bl] = reached ? (T)((qv - pq) * inv) : (T)0;
  }
  return true;
}

// One (request, wavenumber iw, layer j) item on one wavefront: the k-table
// interpolation of every gas (get_kg.f:308-476), then the random overlap of
// noverlapg.f:64-200 round by round, each round a sort of the NG*NG mixed
// keys (rankg.f:77-96) and a rebin into NG bins (rankg.f:104-152).  Writes
// the layer's k-distribution to ok[i * ostride + j] and its derivatives to
// od[(i * (nks + 1) + s) * ostride + j] (s = kslot[gas] < nks: the gases with
// a k-table, per 1e20 molecules; s = nks: temperature).
//
// Sort slots: lane l holds E consecutive elements of one row of the NG x NG
// key matrix (layer_lpr), so with k2 non-decreasing its keys arrive in
// order and the in-lane merge levels of the network are skipped.  The
// rebin takes, for the key and each payload, an FP64 prefix along the
// sorted order recorded at each bin's closing slot; a bin is its record
// minus the previous bin's.  Rounds whose gas has no k at this wavenumber
// leave the distribution as it is (noverlapg.f:100,166).
template <typename T, int E, int BF>
__device__ __forceinline__ void layer_item(const KernelArgs &a, const DevReq &R, int iw, int j,
                                           const char *hdr, char *ws, T *ok, T *od, long long ostride,
                                           int lane) {
  const int ngas = R.ngas, NG = a.ng;
  const LayerConst lc = a.lay[R.lay0 + j];
  const int NS = ngas + 2;
  const HeaderLayout H = header_layout(NG, E, (int)sizeof(T));
  const double *G = reinterpret_cast<const double *>(hdr + H.G);
  const double *dl64 = reinterpret_cast<const double *>(hdr + H.dl64);
  const T *dlg = reinterpret_cast<const T *>(hdr + H.dlg);
  const LayerLayout lo = layer_layout(NG, ngas, E, (int)sizeof(T));
  double *sSt = reinterpret_cast<double *>(ws + lo.sSt);
  double *sGd = reinterpret_cast<double *>(ws + lo.sGd);
  T *kgn = reinterpret_cast<T *>(ws + lo.kgn);
  T *dkgn = reinterpret_cast<T *>(ws + lo.dkgn);
  T *bufA = reinterpret_cast<T *>(ws + lo.bufA);
  T *bufB = reinterpret_cast<T *>(ws + lo.bufB);
  int *sE = reinterpret_cast<int *>(ws + lo.sE);
  char *uni = ws + lo.uni;
  const float rng = rcp_idx(NG);  // e -> i = floor((e + 0.5) / NG)
  // lane -> (group gq, ordinate or bin gg): W / NG groups of NG lanes
  const int gpp = W / NG;
  const int gq = (int)(((float)lane + 0.5f) * rng);
  const int gg = lane - gq * NG;

  // ---- k-table interpolation, get_kg.f:308-476 -----------------------------
  {
    const T *tab = reinterpret_cast<const T *>(a.ktab);
    const int nt = a.nt;
    const int i1 = (nt * NG * (lc.cp - 1)) + (NG * (lc.ct - 1));
    const int i2 = (nt * NG * lc.cp) + (NG * (lc.ct2 - 1));
    const int i3 = (nt * NG * lc.cp) + (NG * lc.ct2);
    const int i4 = (nt * NG * (lc.cp - 1)) + (NG * lc.ct);
    const T v = (T)lc.v, u = (T)lc.u, dudt = (T)lc.dudt;
    const T u2 = (T)lc.u2, dudt2 = (T)lc.dudt2;
    // bilinear in (ln P, T) on one temperature grid (get_kg.f:349-350), or
    // per pressure row on each row's own grid (NT < 0, :352-357)
    auto bilin = [&](T y1, T y2, T y3, T y4, T &x, T &dxdt) {
      if (!a.t2) {
        x = ((T)1 - v) * ((T)1 - u) * y1 + v * ((T)1 - u) * y2 + v * u * y3 + ((T)1 - v) * u * y4;
        dxdt = (-((T)1 - v) * y1 - v * y2 + v * y3 + ((T)1 - v) * y4) * dudt;
      } else {
        x = ((T)1 - v) * (((T)1 - u) * y1 + u * y4) + v * (((T)1 - u2) * y2 + u2 * y3);
        dxdt = ((T)1 - v) * (-y1 + y4) * dudt + v * (-y2 + y3) * dudt2;
      }
    };
    const float ln10 = 2.302585092994046f;

    // Coincident or zero slices (get_kg.f:319-376), W / NG gases at a time:
    // ln k bilinear when the first ordinate of all four corners is > 0
    // (KLOG), else k itself.
    // gases between two table points, one bit each (a wave-uniform mask)
    unsigned long long between = 0;
    for (int k0 = 0; k0 < ngas; k0 += gpp) {
      const int k = k0 + gq;
      bool merge = false;
      if (gq < gpp && k < ngas) {
        const WaveGas wgk = a.wg[R.wg0 + (long long)iw * ngas + k];
        if (wgk.base < 0 || wgk.coinc) {
          T kk = 0, dk = 0;
          if (wgk.base >= 0)
            interp_point<T, BF>(a, wgk.base, wgk.ob, i1, i2, i3, i4, gg, v, u, dudt, u2, dudt2, kk, dk);
          kgn[k * NG + gg] = kk;
          dkgn[k * NG + gg] = dk;
        } else {
          merge = gg == 0;
        }
      }
      const unsigned long long lanes = __ballot(merge);
      for (int q = 0; q < gpp && k0 + q < ngas; q++)
        if ((lanes >> (q * NG)) & 1ull) between |= 1ull << (k0 + q);
    }

    // Wavenumber between two table points (get_kg.f:379-440): interpolate
    // linear k on the slice at or below it (lanes 0..NG-1) and on the slice
    // above (lanes NG..2NG-1), weight the ordinates delg_i (1 - w2) and
    // delg_i w2, sort the 2NG points and rebin them in the order of rankk.f
    // (whole points while their cumulative weight is <= the next edge, the
    // straddling point split by the fraction below the edge, each closed bin
    // divided by its accumulated weight).  FP64 accumulation in both forms.
    T *rkK = reinterpret_cast<T *>(uni);
    T *rkD = rkK + W, *rkF = rkD + W;
    float *keyM = reinterpret_cast<float *>(rkF + W);  // FP32 form: exact keys
    for (unsigned long long mk = between; mk != 0; mk &= mk - 1) {
      const int k = __builtin_ctzll(mk);
      const WaveGas wgk = a.wg[R.wg0 + (long long)iw * ngas + k];
      const int cnt = 2 * NG;
      const bool act = lane < cnt;
      const int gi = act ? (lane < NG ? lane : lane - NG) : 0;
      T x = (T)0, dxdt = (T)0, fw = (T)0;
      if (act) {
        const long long sb = (lane < NG) ? wgk.base : wgk.base2;
        T y1, y2, y3, y4;
        if (BF) {
          const unsigned short *sl = reinterpret_cast<const unsigned short *>(a.ktab) + sb;
          const float offg = a.koff[((lane < NG) ? wgk.ob : wgk.ob2) + gi];
          auto dec = [&](unsigned short c) -> T {
            if (code_nonpos<BF>(c)) return (T)0;
            return (T)expf((code_value<BF>(c) + offg) * ln10);
          };
          y1 = dec(sl[i1 + gi]);
          y2 = dec(sl[i2 + gi]);
          y3 = dec(sl[i3 + gi]);
          y4 = dec(sl[i4 + gi]);
        } else {
          const T *sl = tab + sb;
          y1 = sl[i1 + gi];
          y2 = sl[i2 + gi];
          y3 = sl[i3 + gi];
          y4 = sl[i4 + gi];
        }
        bilin(y1, y2, y3, y4, x, dxdt);
        const T wl = (lane < NG) ? (T)(1.0 - wgk.w2) : (T)wgk.w2;
        fw = dlg[gi] * wl;
      }
      wave_sync();  // the previous merge's reads of the buffers are done
      rkD[lane] = dxdt;
      rkF[lane] = fw;
      T key[1];
      int idx[1];
      if constexpr (sizeof(T) == 4) {
        const float kq = act ? ((x > 0.0f) ? x : 0.0f) : __builtin_huge_valf();
        keyM[lane] = kq;
        unsigned pk[1] = {(__float_as_uint(kq) & ~(unsigned)(W - 1)) | (unsigned)lane};
        bitonic_sort_u32<1>(pk, lane);
        wave_sync();
        unsigned long long w[1];
        {
          const unsigned e = pk[0] & (unsigned)(W - 1);
          w[0] = key_word(keyM[e], e);
        }
        exact_fixup<1>(w, lane);
        idx[0] = (int)(unsigned)w[0];
        key[0] = __uint_as_float((unsigned)(w[0] >> 32));
      } else {
        key[0] = act ? x : dinf<T>();
        idx[0] = lane;
        wave_sync();
        Sorter<T, 1>::sort(key, idx, lane);
      }
      const T sd = rkD[idx[0]], sf = rkF[idx[0]];
      wave_sync();
      rkK[lane] = key[0];
      rkD[lane] = sd;
      rkF[lane] = sf;
      for (int i = lane; i < NG; i += W) {
        kgn[k * NG + i] = (T)0;
        dkgn[k * NG + i] = (T)0;
      }
      wave_sync();
      double gprev = 0.0, sum = 0.0, ak = 0.0, ad = 0.0;
      int ig = 0;
      for (int I = 0; I < cnt; I++) {
        const double c = (double)rkK[I], dc = (double)rkD[I], wgt = (double)rkF[I];
        const double gcur = gprev + wgt;
        if (ig < NG) {  // once the last bin is closed the remaining points are dropped
          if (gcur <= G[ig + 1]) {
            ak += c * wgt;
            ad += dc * wgt;
            sum += wgt;
          } else {
            const double fr = (G[ig + 1] - gprev) / (gcur - gprev);
            ak += fr * c * wgt;
            ad += fr * dc * wgt;
            sum += fr * wgt;
            if (lane == 0) {
              kgn[k * NG + ig] = (T)(ak / sum);
              dkgn[k * NG + ig] = (T)(ad / sum);
            }
            ig++;
            if (ig < NG) {
              sum = (1.0 - fr) * wgt;
              ak = (1.0 - fr) * c * wgt;
              ad = (1.0 - fr) * dc * wgt;
            }
          }
        }
        gprev = gcur;
      }
      if (ig == NG - 1 && lane == 0) {
        kgn[k * NG + ig] = (T)(ak / sum);
        dkgn[k * NG + ig] = (T)(ad / sum);
      }
      wave_sync();
      for (int i = lane; i < NG; i += W) {  // get_kg.f:443-471
        const T kk = kgn[k * NG + i], dk = dkgn[k * NG + i];
        if (kk != kk) kgn[k * NG + i] = (T)1e-37;
        if (dk != dk) dkgn[k * NG + i] = (T)1e-37;
      }
    }
  }
  wave_sync();

  // ---- random overlap, noverlapg.f:64-200 --------------------------------
  for (int x = lane; x < NG; x += W) sE[x] = 0;  // bin record types, cleared per round
  // Row state: slot 0 = K, slot 1+s = gas s, slot 1+ngas = dK/dT.  Round 1
  // starts from gas 0 scaled by its amount.
  const double *amo = a.amo + rfl64(R.lg0) + (long long)j * ngas;
  {
    const T a1 = (T)amo[0];
    for (int x = lane; x < NS * NG; x += W) {
      const int sl = (int)(((float)x + 0.5f) * rng), i = x - sl * NG;
      const T k1 = kgn[i];
      T val = (T)0;
      if (sl == 0) val = k1 * a1;
      else if (sl == 1) val = k1;
      else if (sl == NS - 1) val = dkgn[i] * a1;
      bufA[x] = val;
    }
  }
  wave_sync();

  const double *w64 = reinterpret_cast<const double *>(hdr + H.w64);
  const double *Gw = reinterpret_cast<const double *>(hdr + H.Gw);

  if constexpr (E >= 4) {
    if (a.rebin && ngas > 1 && counts_fit<E>(NG)) counts_init<T, E>(NG, ngas, dl64, uni, lane);
  }
  T *rows = bufA, *nxt = bufB;
  const int ngas_all = ngas;
  for (int r = 1; r < ngas_all; r++) {
    // Opaque per-round copies of NG, the gas count and the lane number: the
    // round's uniform offsets and masks and its lane indices are formed from
    // them inside the round (scalar and cheap) instead of being hoisted out
    // of the loop and held, spilled, across it.
    int ng_r = NG, ngas_r = ngas_all;
    asm volatile("" : "+s"(ng_r), "+s"(ngas_r));
    const int NG = ng_r, ngas = ngas_r, lane = lane_fresh(), NS = ngas + 2;
    // this lane's sort slots: row li of the running distribution, columns
    // lj0 .. lj0 + E - 1 of the new gas (layer_lpr)
    const int lpr = layer_lpr(NG, E);
    const float rlpr = rcp_idx(lpr), rng = rcp_idx(NG);
    const int li = (int)(((float)lane + 0.5f) * rlpr);
    const int lj0 = (lane - li * lpr) * E;
    const bool lrow = li < NG;
    // (i, j) of slot s
    auto slot_ij = [&](int sl, int &i, int &jx) {
      const int l = sl / E;
      i = (int)(((float)l + 0.5f) * rlpr);
      jx = (l - i * lpr) * E + (sl & (E - 1));
    };
    const T a2 = (T)amo[r];
    const T *k2 = kgn + r * NG;
    const T *d2 = dkgn + r * NG;
    // zero-distribution shortcuts test the last ordinate only
    // (noverlapg.f:89,100,153,166); round 1 tests the raw gas-0 table values
    const bool z1 = (r == 1) ? (kgn[NG - 1] == (T)0) : (rows[NG - 1] == (T)0);
    const bool z2 = (k2[NG - 1] == (T)0);
    if (z1) {
      for (int x = lane; x < NS * NG; x += W) {
        const int sl = (int)(((float)x + 0.5f) * rng), i = x - sl * NG;
        T val = (T)0;
        if (sl == 0) val = k2[i] * a2;
        else if (sl == 1 + r) val = k2[i];
        else if (sl == NS - 1) val = d2[i] * a2;
        nxt[x] = val;
      }
    } else if (z2) {
      // the distribution is unchanged, and its slot for gas r is still zero
      continue;
    } else {
      // keys K_i + k2_j*a2 (noverlapg.f:117-128,175-187); padding slots hold
      // +inf.  Ties are ordered by slot, i.e. by (i, j).  With k2
      // non-decreasing every lane's E keys arrive in order.
      const bool k2up = __all(lane + 1 >= NG || !(k2[lane + 1] < k2[lane]));
      int idx[E];
      T key[E];
      {
        const T K = lrow ? rows[li] : (T)0;
        if constexpr (sizeof(T) == 4) {
          // FP32 keys are >= 0, so their bit patterns order like their
          // values.  Sort one 32-bit word per slot, (key bits with the low
          // log2(64E) bits cleared) | slot, with min/max exchanges; then
          // restore the exact (key, slot) order inside runs of equal
          // truncated keys from the exact keys kept in LDS.
          float *keyE = reinterpret_cast<float *>(uni);
          constexpr unsigned IM = (unsigned)(W * E - 1);
          unsigned pk[E];
#pragma unroll
          for (int q = 0; q < E; q++) {
            const int jx = lj0 + q;
            float kq = __builtin_huge_valf();
            if (lrow && jx < NG) {
              kq = K + k2[jx] * a2;
              kq = (kq > 0.0f) ? kq : 0.0f;  // -0 and 0 order alike
            }
            keyE[lane * E + q] = kq;
            pk[q] = (__float_as_uint(kq) & ~IM) | (unsigned)(lane * E + q);
          }
          if (k2up) bitonic_sort_u32_lanesorted<E>(pk, lane);
          else bitonic_sort_u32<E>(pk, lane);
          wave_sync();
          unsigned long long w[E];
#pragma unroll
          for (int q = 0; q < E; q++) {
            const unsigned e = pk[q] & IM;
            w[q] = key_word(keyE[e], e);
          }
          exact_fixup<E>(w, lane);
#pragma unroll
          for (int q = 0; q < E; q++) idx[q] = (int)(unsigned)w[q];
        } else if constexpr (E >= 2) {
          // FP64 keys are >= 0 too: one 64-bit word per slot, the key's bits
          // with the low log2(64E) mantissa bits replaced by the slot, sorted
          // by the same network.  Two distinct keys can tie in the truncated
          // bits only within 2^-43 of each other; a wavefront that has such a
          // pair forms the exact keys again from the slots and settles the
          // (key, slot) order on them.
          constexpr unsigned long long IM = (unsigned long long)(W * E - 1);
          constexpr unsigned long long INFB = 0x7FF0000000000000ull & ~IM;
          unsigned long long wd[E];
#pragma unroll
          for (int q = 0; q < E; q++) {
            const int jx = lj0 + q;
            double kq = dinf<double>();
            if (lrow && jx < NG) {
              kq = fma((double)k2[jx], (double)a2, (double)K);
              kq = (kq > 0.0) ? kq : 0.0;  // -0 and 0 order alike
            }
            wd[q] = ((unsigned long long)__double_as_longlong(kq) & ~IM) | (unsigned long long)(lane * E + q);
          }
          if (k2up) bitonic_sort_u64_lanesorted<E>(wd, lane);
          else bitonic_sort_u64<E>(wd, lane);
          auto tie = [&](unsigned long long x, unsigned long long y) {
            return ((x ^ y) & ~IM) == 0 && (x & ~IM) != INFB;
          };
          bool t = false;
#pragma unroll
          for (int q = 0; q + 1 < E; q++) t |= tie(wd[q], wd[q + 1]);
          t |= (lane + 1 < W) && tie(wd[E - 1], lane_next(wd[0]));
#pragma unroll
          for (int q = 0; q < E; q++) idx[q] = (int)(wd[q] & IM);
          if (__any(t)) {
            double kx[E];
#pragma unroll
            for (int q = 0; q < E; q++) {
              int i, jx;
              slot_ij(idx[q], i, jx);
              double kq = dinf<double>();
              if (i < NG && jx < NG) {
                kq = fma((double)k2[jx], (double)a2, (double)rows[i]);
                kq = (kq > 0.0) ? kq : 0.0;
              }
              kx[q] = kq;
            }
            exact_fixup_d<E>(kx, idx, lane);
          }
        } else {
#pragma unroll
          for (int q = 0; q < E; q++) {
            const int jx = lj0 + q;
            key[q] = (lrow && jx < NG) ? K + k2[jx] * a2 : dinf<T>();
            idx[q] = lane * E + q;
          }
          Sorter<T, E>::sort(key, idx, lane);
        }
      }
      wave_sync();  // the exact keys are read: the rebin's tables take their place
      bool done = false;
      if constexpr (E >= 4) {
        if (a.rebin && k2up && counts_fit<E>(NG)) {
          // words of gamma per row: 4 (E = 4), 5 or 6 (E = 8), 6 or 8 (E = 16)
          const int nd = (NG + 3) >> 2;
          constexpr int NA = (E == 4) ? 4 : (E == 8) ? 5 : 6;
          constexpr int NB = (E == 4) ? 4 : (E == 8) ? 6 : 8;
          if (nd == NA)
            done = rebin_counts<T, E, NA>(NG, ngas, r, a2, idx, rows, nxt, k2, d2, G, dl64, w64, Gw,
                                          uni, sSt, sGd, lane);
          else if (nd == NB)
            done = rebin_counts<T, E, NB>(NG, ngas, r, a2, idx, rows, nxt, k2, d2, G, dl64, w64, Gw,
                                          uni, sSt, sGd, lane);
        }
      }
      if (!done) {
        // per-bin records (rankg.f:129-152 closes a bin at its straddling
        // element, or at the end): recS[b][vq] the FP64 prefix of quantity vq
        // through the closing slot, recC[b][vq] the closing slot's share of
        // it that belongs to the next bin; recG[b] / recW[b] the same for the
        // weight; rtype[b] 1 closed by a straddle, 2 open at the end, 0 never
        // reached
        const int NQ = ngas + 2;
        double *recS = reinterpret_cast<double *>(uni);
        double *recC = recS + NG * NQ;
        double *recG = sGd, *recW = sSt;
        int *rtype = sE;

        // (i, j) of each slot's element, packed (i << 8) | j; padding gets
        // (NG << 8)
        int ij[E];
        T wq[E];
        double gd[E];
        double run = 0.0;
  #pragma unroll
        for (int q = 0; q < E; q++) {
          int i, jx;
          slot_ij(idx[q], i, jx);
          const bool real = i < NG && jx < NG;
          ij[q] = real ? (i << 8) | jx : (NG << 8);
          // weights and their FP64 cumulative sum gdist (rankg.f:104-108).  The
          // sum is taken over FP64 products: where a run of elements exactly
          // fills a bin, FP32 products would spill a ~1e-8 share of the next
          // (possibly far larger) key across the edge.
          const double w = w64[idx[q]];
          wq[q] = (T)w;
          run += w;
          gd[q] = run;
        }
        {
          const double s = wave_incl_sum(run);
          const double ex = wave_shr1(s);
  #pragma unroll
          for (int q = 0; q < E; q++) gd[q] += ex;
        }
        // cumulative weight before slot 0: the previous lane's last gd
        const double gprev = wave_shr1(gd[E - 1]);

        // bin of each element = number of edges G[1..NG] <= its start; it
        // straddles when its end reaches the next edge (rankg.f:120).  The
        // lane's first bin b0 comes from a search; the next four edges are
        // held in registers and the bin of each slot is b0 plus the number of
        // them at or below its start.  A lane whose slots pass all four takes
        // the edge-by-edge walk instead.
        int b0;
        {
          int lo2 = 0, hi2 = NG;
          while (lo2 < hi2) {
            const int mid = (lo2 + hi2 + 1) >> 1;
            if (G[mid] <= gprev) lo2 = mid;
            else hi2 = mid - 1;
          }
          b0 = lo2;
        }
        constexpr int LA = 4;
        double ge[LA];
  #pragma unroll
        for (int t = 0; t < LA; t++) ge[t] = (b0 + 1 + t <= NG) ? G[b0 + 1 + t] : dinf<double>();
        auto edges_le = [&](double x) {
          int c = 0;
  #pragma unroll
          for (int t = 0; t < LA; t++) c += (ge[t] <= x) ? 1 : 0;
          return c;
        };
        if (!__any(edges_le(gd[E - 1]) >= LA)) {
          int cs = 0;
  #pragma unroll
          for (int q = 0; q < E; q++) {
            const int ce = edges_le(gd[q]);
            const int b = b0 + cs;
            ij[q] |= (b << 16) | ((ce > cs && b < NG) ? (1 << 24) : 0);
            cs = ce;
          }
        } else {
          int b = b0;
  #pragma unroll
          for (int q = 0; q < E; q++) {
            const double st = (q == 0) ? gprev : gd[q - 1];
            if (q > 0)
              while (b < NG && G[b + 1] <= st) b++;
            ij[q] |= (b << 16) | (((b < NG) && !(gd[q] < G[b + 1])) ? (1 << 24) : 0);
          }
        }
        // omf: 1 - frac of rankg.f:121 for a straddling element, formed as a
        // ratio of two FP64 differences
        T omf[E];
  #pragma unroll
        for (int q = 0; q < E; q++) {
          const int b = (ij[q] >> 16) & 255;
          const bool stq = (ij[q] >> 24) != 0;
          const double st = (q == 0) ? gprev : gd[q - 1];
          omf[q] = stq ? ratio<T>((T)(gd[q] - G[b + 1]), (T)(gd[q] - st)) : (T)0;
        }
        // bin records: a straddling element closes its bin; the last slot
        // closes the bin it lies in (or, if it straddles, opens the next one)
  #pragma unroll
        for (int q = 0; q < E; q++) {
          const int e = lane * E + q;
          const int bi = (ij[q] >> 16) & 255;
          const bool stq = (ij[q] >> 24) != 0;
          if (stq) {
            recG[bi] = gd[q];
            recW[bi] = (double)omf[q] * (double)wq[q];
            rtype[bi] = 1;
          }
          if (e == W * E - 1) {
            const int bb = stq ? bi + 1 : bi;
            if (bb < NG) {
              recG[bb] = gd[q];
              recW[bb] = 0.0;
              rtype[bb] = 2;
            }
          }
        }
        // weighted key and payloads: FP64 prefix sums along the sorted order,
        // recorded at the straddling slots and the last slot
        const int NGH = NG << 8;  // (ij & 0xffff) < NGH: a real element
        auto scan_quantity = [&](int vq, auto value) {
          T cq[E];
          double run2 = 0.0;
  #pragma unroll
          for (int q = 0; q < E; q++) {
            const T c = ((ij[q] & 0xffff) < NGH) ? value(q) * wq[q] : (T)0;
            cq[q] = c;
            run2 += (double)c;
          }
          double pq = wave_shr1(wave_incl_sum(run2));
  #pragma unroll
          for (int q = 0; q < E; q++) {
            const int e = lane * E + q;
            pq += (double)cq[q];
            const int bi = (ij[q] >> 16) & 255;
            const bool stq = (ij[q] >> 24) != 0;
            if (stq) {
              recS[bi * NQ + vq] = pq;
              recC[bi * NQ + vq] = (double)omf[q] * (double)cq[q];
            }
            if (e == W * E - 1) {
              const int bb = stq ? bi + 1 : bi;
              if (bb < NG) {
                recS[bb * NQ + vq] = pq;
                recC[bb * NQ + vq] = 0.0;
              }
            }
          }
        };
        const int nq = r + 3;
        scan_quantity(0, [&](int q) { return rows[(ij[q] >> 8) & 255] + k2[ij[q] & 255] * a2; });
        for (int vq = 1; vq < nq; vq++) {
          const int s2 = vq - 1;
          if (s2 < r) {
            const T *src = rows + (1 + s2) * NG;
            scan_quantity(vq, [&](int q) { return src[(ij[q] >> 8) & 255]; });
          } else if (s2 == r) {
            scan_quantity(vq, [&](int q) { return k2[ij[q] & 255]; });
          } else {
            const T *src = rows + (1 + ngas) * NG;
            scan_quantity(vq, [&](int q) { return src[(ij[q] >> 8) & 255] + d2[ij[q] & 255] * a2; });
          }
        }
        wave_sync();
        // bin b = its records minus the previous bin's, plus the share carried
        // over from the element straddling its lower edge; normalised by the
        // accumulated weight when a straddle closed it, or when it is the last
        // bin (rankg.f:129-152)
        if (lane < NG) {
          const int b = lane;
          const int t = rtype[b];
          if (t == 0) {
            for (int sl = 0; sl < NS; sl++) nxt[sl * NG + b] = (T)0;
          } else {
            const bool hp = b > 0;
            const double sw =
                recG[b] - (hp ? recG[b - 1] : 0.0) + (hp ? recW[b - 1] : 0.0) - recW[b];
            const bool norm = (t == 1) || (b == NG - 1);
            const T smT = (T)sw;
            for (int vq = 0; vq < nq; vq++) {
              const double Sv = recS[b * NQ + vq] - (hp ? recS[(b - 1) * NQ + vq] : 0.0) +
                                (hp ? recC[(b - 1) * NQ + vq] : 0.0) - recC[b * NQ + vq];
              T val = (T)Sv;
              if (norm) val = val / smT;
              const int sl = (vq == 0) ? 0 : ((vq - 1 == r + 1) ? 1 + ngas : vq);
              nxt[sl * NG + b] = val;
            }
            for (int s4 = r + 1; s4 < ngas; s4++) nxt[(1 + s4) * NG + b] = (T)0;
          }
          rtype[b] = 0;
        }
      }
    }
    wave_sync();
    T *t = rows;
    rows = nxt;
    nxt = t;
  }

  // ---- the layer's k-distribution and derivatives ------------------------
  // (the lane number and 1/NG formed again here, not held across the rounds)
  const int nks = R.nks;
  int ng_o = NG;
  asm volatile("" : "+s"(ng_o));
  const float rng_o = rcp_idx(ng_o);
  for (int x = lane_fresh(); x < NS * NG; x += W) {
    const int sl = (int)(((float)x + 0.5f) * rng_o), i = x - sl * NG;
    const T val = rows[x];
    const int ks = (sl == 0) ? -2 : (sl == NS - 1) ? nks : (int)R.kslot[sl - 1];
    if (ks == -2) ok[(long long)i * ostride + j] = val;
    else if (ks >= 0) od[((long long)i * (nks + 1) + ks) * ostride + j] = val;
  }
  wave_sync();  // the scratch is free for the next item
}

// Continuum of one (request, wavenumber iw, layer j) item, FP64: dust, the
// gas continua of ngascon.f, the CIA of nciacon.f/ciaread.f and Rayleigh.
// tcon[j] gets the optical depth, dtcon[s * cstride + j] its derivative for
// parameter s (gas slots per 1e20 molecules, then T, then the
// continuum-only parameters).
template <typename T>
__device__ __forceinline__ void continuum_layer(const KernelArgs &a, const DevReq &R, int iw, int j,
                                                T *tcon, T *dtcon, long long cstride) {
  const int ngas = R.ngas;
  const LayerConst lc = a.lay[R.lay0 + j];
  const double vwave = a.vwave[R.wave0 + iw];
  double absorb, dab[7];
  const double vv = (R.ispace == 1) ? 1e4 / vwave : vwave;
  // dust (cirsradg_wave.f:550-577), first as in the reference: extinction at
  // vwave (interpolated on the host) times particles/cm2
  double tdust = 0.0;
  for (int k = 0; k < R.ncont; k++)
    tdust += a.dext[R.dx0 + (long long)iw * R.ncont + k] * a.cont[R.ct0 + (long long)j * R.ncont + k];
  // gas continua (cirsradg_wave.f:590-611): each gas's absorption, and
  // absorb / amount as its d(tau)/d(amount) (here per 1e20 molecules)
  double gcd[RADG_MAXGAS];
  if (R.gcw0 >= 0) {
    const double *gw = a.gcw + R.gcw0 + (long long)iw * RADG_GCW;
    for (int k = 0; k < ngas; k++) {
      gcd[k] = 0.0;
      if (R.gcon[k] == RADG_GC_NONE) continue;
      const double ab = gascon_d(a, R, k, j, vv, lc, gw);
      const double amo = a.amo[R.lg0 + (long long)j * ngas + k];
      tdust += ab;
      if (ab > 0.0 && amo > 0.0) gcd[k] = ab / amo;
    }
  }
  const double *n2 = a.n2nir + 2 * (R.wave0 + iw);
  if (R.flagh2p)
    cia_para_layer(a, vv, lc, n2, &absorb, dab);
  else
    cia_layer(a, vv, R.inormal, lc, n2, &absorb, dab);
  absorb = tdust + absorb;
  // Rayleigh (cirsradg_wave.f:663-691): utotl * xray more extinction and
  // xray * 1e-20 in every gas's d(tau)/d(amount); no T derivative.  The gas
  // rows of dtcon, like dk/dq, are per 1e20 molecules (the 1e-20 is applied
  // once when the Jacobian is written), so the row gets xray itself.
  double dray = 0.0;
  if (R.iray > 0) {
    const double xray = rayleigh_xs(R.iray, vv, lc) * 1e20;
    absorb += lc.utotl * xray;
    dray = xray;
  }
  for (int s = 0; s <= ngas; s++) {
    double d = 0.0;
    for (int l = 0; l < 5; l++)
      if (R.iabs[l] == s) d += dab[l] / lc.utotl;
    if (s == ngas) d += dab[5];
    else d += dray + (R.gcw0 >= 0 ? gcd[s] : 0.0);
    dtcon[(long long)s * cstride + j] = (T)d;
  }
  // the continuum-only parameters: each dust type's extinction
  // (cirsradg_wave.f:573), then the para-H2 fraction (:659-661)
  for (int c = 0; c < R.ncont; c++)
    dtcon[(long long)(ngas + 1 + c) * cstride + j] = (T)a.dext[R.dx0 + (long long)iw * R.ncont + c];
  if (R.flagh2p) dtcon[(long long)(R.nparam - 1) * cstride + j] = (T)dab[6];
  tcon[j] = (T)absorb;
}
}  // namespace

// The layer pass as its own launch: one wavefront per (request,
// wavenumber, layer) item, a.lwaves consecutive items per workgroup sharing
// the header, results to the intermediate block in HBM.
// Occupancy: the FP32 forms with 8 slots per lane are held to 80 VGPRs, six
// wavefronts per SIMD, which is also what their LDS (six 4-wavefront
// workgroups per CU at 20 g-ordinates and 3 gases) allows: the few spills
// this costs are cheaper than the latency that 4 or 5 wavefronts leave
// exposed (measured 11.1, 10.5 and 10.4 ms per 16 spectra at 4, 5 and 6).
template <typename T, int E, int BF>
__global__ __launch_bounds__(256, (E >= 16 ? 1 : (E == 8 && sizeof(T) == 4) ? 6 : 4)) void layer_kernel(
    KernelArgs a, long long nblocks) {
  const int tid = threadIdx.x;
  // the wavefront number is uniform: in a scalar register, every scratch
  // address of the item stays scalar
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & (W - 1);
  extern __shared__ double smem_d[];
  char *hdr = reinterpret_cast<char *>(smem_d);
  layer_header<T, E>(a, hdr, tid, (int)blockDim.x);
  __syncthreads();
  const long long item = xcd_item(nblocks) * a.lwaves + wv;
  if (item >= a.nA) return;
  const int rqc = find_req<true>(a.startA, a.nreqA, item, a.nA);
  const DevReq &R = a.reqs[a.reqmapA[rqc]];
  const int L = R.L, NG = a.ng;
  const long long loc = item - a.startA[rqc];
  const int iw = div_item<true>(loc, L);
  const int j = (int)(loc - (long long)iw * L);
  T *inter = reinterpret_cast<T *>(a.inter) + R.inter0 + (long long)iw * inter_block(NG, L, R.nks, R.nparam);
  layer_item<T, E, BF>(a, R, iw, j, hdr,
                       hdr + layer_header_bytes(NG, E, (int)sizeof(T)) + wv * a.lwstride, inter,
                       inter + (long long)NG * L, L, lane);
}

// The continuum as its own launch: one (request, wavenumber, layer) item
// per thread, into the tcon/dtcon part of the intermediate block.
template <typename T>
__global__ __launch_bounds__(256) void cia_kernel(KernelArgs a) {
  const long long item = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= a.nA) return;
  const int rqc = find_req<false>(a.startA, a.nreqA, item, a.nA);
  const DevReq &R = a.reqs[a.reqmapA[rqc]];
  const int L = R.L, ngas = R.ngas, NG = a.ng;
  const long long loc = item - a.startA[rqc];
  const int iw = div_item<false>(loc, L);
  const int j = (int)(loc - (long long)iw * L);
  T *inter = reinterpret_cast<T *>(a.inter) + R.inter0 + (long long)iw * inter_block(NG, L, R.nks, R.nparam);
  T *tcon = inter + (long long)NG * L + (long long)NG * (R.nks + 1) * L;
  continuum_layer<T>(a, R, iw, j, tcon, tcon + L, L);
}

// The layer pass on line-by-line tables (ILBL = 2, one g-ordinate): one
// (request, wavenumber, layer) item per thread.  Each gas's k and dk/dT
// from its table slice (get_klblg.f:231-301), then the gases summed in
// order with noverlapg1.f's branches for zero terms (:69-156): k = sum of
// amount_i k_i; gas columns k_i, except that a zero running sum before gas
// m >= 3 zeroes the columns of the gases before it (:123-128); the T column
// carries the running sum's derivative through the same branches.  FP64
// sums in every form; results to the rows the layer kernel writes.
template <typename T, int BF>
__global__ __launch_bounds__(256) void lbl_kernel(KernelArgs a) {
  const long long item = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= a.nA) return;
  const int rqc = find_req<false>(a.startA, a.nreqA, item, a.nA);
  const DevReq &R = a.reqs[a.reqmapA[rqc]];
  const int L = R.L, ngas = R.ngas, nks = R.nks;
  const long long loc = item - a.startA[rqc];
  const int iw = div_item<false>(loc, L);
  const int j = (int)(loc - (long long)iw * L);
  T *ok = reinterpret_cast<T *>(a.inter) + R.inter0 + (long long)iw * inter_block(1, L, nks, R.nparam);
  T *od = ok + L;
  const LayerConst lc = a.lay[R.lay0 + j];
  const int nt = a.nt;
  const int i1 = nt * (lc.cp - 1) + (lc.ct - 1);
  const int i2 = nt * lc.cp + (lc.ct2 - 1);
  const int i3 = nt * lc.cp + lc.ct2;
  const int i4 = nt * (lc.cp - 1) + lc.ct;
  const T v = (T)lc.v, u = (T)lc.u, dudt = (T)lc.dudt, u2 = (T)lc.u2, dudt2 = (T)lc.dudt2;
  double s = 0.0, d = 0.0, k1 = 0.0, dk1 = 0.0, a1 = 0.0;
  int zmark = 0;
  for (int k = 0; k < ngas; k++) {
    const WaveGas wgk = a.wg[R.wg0 + (long long)iw * ngas + k];
    T kk = 0, dk = 0;
    if (wgk.base >= 0) interp_point<T, BF>(a, wgk.base, wgk.ob, i1, i2, i3, i4, 0, v, u, dudt, u2, dudt2, kk, dk);
    const double kd = (double)kk, dkd = (double)dk;
    const double am = a.amo[R.lg0 + (long long)j * ngas + k];
    if (k == 0) {  // the first pair waits for gas 2 (:76-108)
      k1 = kd;
      dk1 = dkd;
      a1 = am;
      s = kd * am;  // one gas: the loop never runs; k = amount k
      d = dkd * am;
    } else if (k == 1) {
      if (k1 == 0.0) {
        s = kd * am;
        d = dkd * am;
      } else if (kd == 0.0) {
        s = k1 * a1;
        d = dk1 * a1;
      } else {
        s = k1 * a1 + kd * am;
        d = dk1 * a1 + dkd * am;
      }
    } else if (s == 0.0) {  // :123-130
      s = kd * am;
      d = dkd * am;
      zmark = k;
    } else if (kd != 0.0) {  // :141-146 (kd = 0: :134-138, the sum unchanged)
      s = s + kd * am;
      d = d + dkd * am;
    }
    const int ks = R.kslot[k];
    if (ks >= 0) od[(long long)ks * L + j] = (T)(kd == 0.0 ? 0.0 : kd);
  }
  for (int k = 0; k < zmark; k++) {
    const int ks = R.kslot[k];
    if (ks >= 0) od[(long long)ks * L + j] = (T)0;
  }
  od[(long long)nks * L + j] = (T)d;
  ok[j] = (T)s;
}

// ========================================================= emission kernel
namespace {
__device__ __forceinline__ double planck_d(int ispace, double v, double t) {
  const double c1 = 1.1911e-12, c2 = 1.439;
  double y, a;
  if (ispace == 0) {
    y = v;
    a = c1 * (y * y * y);
  } else {
    y = 1e4 / v;
    a = c1 * (y * y * y * y * y) / 1e4;
  }
  if (y == 0.0) return 0.0;
  return a / (exp(c2 * y / t) - 1.0);
}
__device__ __forceinline__ double planckg_d(int ispace, double v, double t) {
  const double c1 = 1.1911e-12, c2 = 1.439;
  double y, a;
  if (ispace == 0) {
    y = v;
    a = c1 * c2 * (y * y * y * y) / (t * t);
  } else {
    y = 1e4 / v;
    a = c1 * c2 * (y * y * y * y * y * y) * 1e-4 / (t * t);
  }
  const double ex = exp(c2 * y / t);
  return ex * a / ((ex - 1.0) * (ex - 1.0));
}

// calc_nlte_t.f:59-85: the temperature whose Planck radiance is f times the
// LTE one (f from collisions, inlte = 1, or from pressure, inlte = 2).  The
// reference hands cirsradg_wave's wavenumber index (>= 1) to planck_wave and
// invplanck_wave as their spectral-form switch, so both use the 1e4/v form
// whatever ispace is; kept.
__device__ __forceinline__ double nlte_t_d(int inlte, double k1, double a, double t, double p,
                                           double v) {
  double f;
  if (inlte == 1) {
    const double n = (p * 101325. / (1.380658e-23 * t)) * 1.0e-6;
    f = (k1 * n) / (a + k1 * n);
  } else if (inlte == 2) {
    f = p * 1.0e6 / (1.0 + p * 1.0e6);
  } else {
    return t;
  }
  const double c1 = 1.1911e-12, c2 = 1.439, y = 1e4 / v;
  const double b = c1 * (y * y * y * y * y) / 1e4 / (exp(c2 * y / t) - 1.0) * f;
  if (b == 0.0) return 0.0;
  return (c2 * 1.0e4 / v) / log(c1 * 1.0e16 / (b * (v * v * v * v * v)) + 1.0);
}

// e3interp.f:57-77 on the device table [trans | E3 | gradient], RADG_E3N
// points each: E3 at transmission x and dE3/dtr, both linear in the
// interval.  A negative or NaN x gives zero; x >= 1 uses the last interval.
__device__ __forceinline__ void e3interp_d(const double *tab, double x, double &e3, double &de3) {
  if (!(x >= 0.0)) {
    e3 = 0.0;
    de3 = 0.0;
    return;
  }
  const double dtrans = 1.0 / (double)(RADG_E3N - 1);
  int j = 1 + (int)(fmin(x, 2.0) / dtrans);
  if (j >= RADG_E3N) j = RADG_E3N - 1;
  double f = (x - tab[j - 1]) / dtrans;
  if (f > 1.0) f = 1.0;
  const double *te = tab + RADG_E3N, *tg = tab + 2 * RADG_E3N;
  e3 = (1.0 - f) * te[j - 1] + f * te[j];
  de3 = (1.0 - f) * tg[j - 1] + f * tg[j];
}

template <typename T>
struct Kahan {
  T s, c;
  __device__ void init() { s = 0; c = 0; }
  __device__ void add(T x) {
    const T y = x - c;
    const T t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int ds = 5; ds >= 0; ds--) x += __shfl_xor(x, 1 << ds);
  return x;
}

// Reads of the layer results ktot [NG][L] and dkq [NG][nks+1][L] by (g,
// row, byte offset of the lane's layer).  In HBM (GLB): buffer loads with
// the row's offset in a scalar and the lane's 32-bit offset, so no 64-bit
// address is formed per load.  In LDS (the fused kernel): plain reads.
template <typename T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <typename T, bool GLB> struct LayerRows;
template <typename T> struct LayerRows<T, true> {
  __amdgpu_buffer_rsrc_t rk, rdk;
  unsigned rowb, nk1;
  __device__ LayerRows(const T *ktot, const T *dkq, int NG, int L, int nks)
      : rowb((unsigned)L * (unsigned)sizeof(T)), nk1((unsigned)nks + 1u) {
    // raw buffers (stride 0) of the two blocks' byte sizes; word 3 is the
    // gfx9 32-bit data format
    rk = __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(ktot), (short)0, (int)((unsigned)NG * rowb),
                                           0x00020000);
    rdk = __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(dkq), (short)0,
                                            (int)((unsigned)NG * nk1 * rowb), 0x00020000);
  }
  __device__ T k(int g, unsigned lb) const { return buf_ld<T>(rk, lb, (unsigned)g * rowb); }
  __device__ T dk(int g, int kd, unsigned lb) const {
    return buf_ld<T>(rdk, lb, ((unsigned)g * nk1 + (unsigned)kd) * rowb);
  }
};
template <typename T> struct LayerRows<T, false> {
  const char *kt, *dq;
  unsigned rowb, nk1;
  __device__ LayerRows(const T *ktot, const T *dkq, int, int L, int nks)
      : kt(reinterpret_cast<const char *>(ktot)), dq(reinterpret_cast<const char *>(dkq)),
        rowb((unsigned)L * (unsigned)sizeof(T)), nk1((unsigned)nks + 1u) {}
  __device__ T k(int g, unsigned lb) const {
    return *reinterpret_cast<const T *>(kt + (unsigned)g * rowb + lb);
  }
  __device__ T dk(int g, int kd, unsigned lb) const {
    return *reinterpret_cast<const T *>(dq + ((unsigned)g * nk1 + (unsigned)kd) * rowb + lb);
  }
};

// One (request, wavenumber iw, path p) item on one wavefront, reading the
// layer results ktot [NG][L], dkq [NG][ngas+1][L], tcon [L] and
// dtcon [nparam][L], with s2 [nks+1][RL][64] (FP64) of scratch.  EXT: the form with the
// imod 8, 13, 14 and 18 branches and the stellar-spectrum factors, launched
// only for classes that hold such a path or request (their registers would
// cost the other forms occupancy).
// An imod 8 item (cell x atmosphere, cirsradg_wave.f:1037-1082) computes its
// atmosphere path again, every g-ordinate's terms times the cell's
// transmission at that g, and writes the product into its own slots.
template <typename T, int RL, bool EXT, bool GLB>
__device__ __forceinline__ void emission_path(const KernelArgs &a, const DevReq &R, int iw, int p,
                                              const T *ktot, const T *dkq, const T *tcon,
                                              const T *dtcon, double *s2, int lane) {
  const int L = R.L, ngas = R.ngas, NG = a.ng, npath = R.npath, nsw = R.nsw,
            nparam = R.nparam, nks = R.nks;
  const PathInfo pi0 = a.paths[R.path0 + p];
  const bool scr = EXT && pi0.imod == 8;
  const int pg = scr ? pi0.atm : p;  // the path whose layers are computed
  const PathInfo pi = scr ? a.paths[R.path0 + pg] : pi0;
  const int nlays = pi.nlays;
  const int nrows = pi0.nlays;  // Jacobian rows written (imod 8: nlayin(1) <= nlays)
  const double v = a.vwave[R.wave0 + iw];
  const T *delg = (sizeof(T) == 8) ? reinterpret_cast<const T *>(a.delg64)
                                   : reinterpret_cast<const T *>(a.delg32);

  int Lm[RL];
  bool act[RL];
  T sc[RL], tc[RL], db[RL];
  double bb[RL];  // Planck source of each layer, FP64
  double s1[RL], s3[RL];  // sum_g delg_g dR/dtau and sum_g delg_g dR/dT (Planck term)
  // xfac (cirsradg_wave.f:912-925 for imod 3, :1134-1149 with its 2 pi for
  // imod 18) scales each g-ordinate's terms in FP64 before they are rounded
  // to T, as the reference applies it per term: with iform 1 it reaches 1e21
  // and an unscaled FP32 term would fall below FP32's range.
  double xfac = pi.imod == 3 ? R.xfac : (EXT && pi.imod == 18) ? R.xfac18 : 1.0;
  if (EXT && R.solar == 1 && (pi.imod == 3 || pi.imod == 18)) xfac = xfac / a.solar[R.wave0 + iw];
  if (EXT && R.solar == 4 && pi.imod == 0) xfac = a.solar[R.wave0 + iw] / R.solar_area;
#pragma unroll
  for (int q = 0; q < RL; q++) {
    const int m = lane * RL + q;
    act[q] = m < nlays;
    Lm[q] = act[q] ? a.layinc[R.pl0 + (long long)pg * L + m] : 0;
    sc[q] = act[q] ? (T)a.scale[R.pl0 + (long long)pg * L + m] : (T)0;
    tc[q] = act[q] ? tcon[Lm[q]] : (T)0;
    bb[q] = 0;
    db[q] = 0;
    if (act[q] && (pi.imod == 3 || (EXT && pi.imod == 18))) {
      double em = a.emtemp[R.pl0 + (long long)pg * L + m];
      if (R.inlte) em = nlte_t_d(R.inlte, R.nlte_k1, R.nlte_a, em, a.press[R.lay0 + Lm[q]], v);
      bb[q] = planck_d(R.ispace, v, em);
      db[q] = (T)planckg_d(R.ispace, v, em);
    }
    s1[q] = 0.0;
    s3[q] = 0.0;
    for (int kd = 0; kd <= nks; kd++) s2[(kd * RL + q) * W + lane] = 0.0;
  }
  // ground term (cirsradg_wave.f:992-1024; imod 18 :1159,1212-1218, where
  // tsurf > 0 always)
  double radground = 0.0, gradground = 0.0;
  if ((pi.imod == 3 || (EXT && pi.imod == 18)) && pi.surf) {
    if (R.tsurf_le0) {
      const int ob = (nlays - 1) / RL, oq = (nlays - 1) - ob * RL;
      double b0 = 0.0, g0 = 0.0;
#pragma unroll
      for (int q = 0; q < RL; q++)
        if (q == oq) {
          b0 = bb[q];
          g0 = (double)db[q];
        }
      radground = __shfl(b0, ob);
      gradground = __shfl(g0, ob);
    } else {
      const double es = a.esurf[R.wave0 + iw];
      radground = es * planck_d(R.ispace, v, R.tsurf);
      gradground = es * planckg_d(R.ispace, v, R.tsurf);
    }
  }

  // imod 8: the cell path's one layer (imod 13 or 14, cirsradg_wave.f:1084-1102)
  int cL = 0, cim = 0;
  T csc = 0;
  if (scr) {
    cim = a.paths[R.path0 + pi0.cell].imod;
    cL = a.layinc[R.pl0 + (long long)pi0.cell * L];
    csc = (T)a.scale[R.pl0 + (long long)pi0.cell * L];
  }

  // byte offset of each slot's layer in a row of L values: a row's base is
  // uniform, so the loads below take a scalar base and a 32-bit lane offset
  unsigned lb[RL];
#pragma unroll
  for (int q = 0; q < RL; q++) lb[q] = (unsigned)Lm[q] * (unsigned)sizeof(T);
  const LayerRows<T, GLB> rows(ktot, dkq, NG, L, nks);

  // Slots past the path's layers (act false) hold tau = 0, B = dB/dT = 0:
  // their transmission is 1 and their emission 0, so the imod 3 sums below
  // need no masks; their Jacobian sums are never written out.
  Kahan<T> rad, gts;
  rad.init();
  gts.init();
  // FP64 form, up to two layers per lane: loads run one step ahead of their
  // use, each g step starting the next step's ktot loads and its own first
  // NKP dk rows before its scans (measured: 1.77 -> 1.22 ms per 16 spectra;
  // the FP32 form runs slower so, at any register bound)
  constexpr bool PF = sizeof(T) == 8 && RL <= 2;
  constexpr int NKP = PF ? 4 : 0;
  T kn[RL];
  if constexpr (PF) {
#pragma unroll
    for (int q = 0; q < RL; q++) kn[q] = rows.k(0, lb[q]);
  }
  for (int g = 0; g < NG; g++) {
    const T dg = delg[g];
    T kc[RL];
#pragma unroll
    for (int q = 0; q < RL; q++) kc[q] = PF ? kn[q] : rows.k(g, lb[q]);
    if (PF && g + 1 < NG) {
#pragma unroll
      for (int q = 0; q < RL; q++) kn[q] = rows.k(g + 1, lb[q]);
    }
    T dkv[NKP > 0 ? NKP : 1][RL];
#pragma unroll
    for (int kd = 0; kd < NKP; kd++)
#pragma unroll
      for (int q = 0; q < RL; q++) dkv[kd][q] = (kd <= nks) ? rows.dk(g, kd, lb[q]) : (T)0;
    T tau[RL];
#pragma unroll
    for (int q = 0; q < RL; q++) tau[q] = act[q] ? (tc[q] + kc[q]) * sc[q] : (T)0;
    T coR, coG = 0;
    T dRdtau[RL], dT[RL];
    if (pi.imod == 3) {
      T tl[RL], om[RL];
      double incl[RL];
      double run = 1.0;
#pragma unroll
      for (int q = 0; q < RL; q++) {
        tl[q] = dexp(-tau[q]);
        om[q] = -dexpm1(-tau[q]);
        run *= (double)tl[q];
        incl[q] = run;
      }
      // exclusive prefix product over lanes (FP64)
      const double s = wave_incl_prod(run);
      const double ex = dpp_dk<0x138, 0xF>(1.0, s);  // lane - 1's, 1 in lane 0
      const double trN = read_lane_d(s, W - 1);
      // radiance, sum_j (tr_{j-1} - tr_j) B_j (cirsradg_wave.f:946-951)
      double part = 0.0;
      double trp[RL], trx[RL];
#pragma unroll
      for (int q = 0; q < RL; q++) {
        trp[q] = (q == 0) ? ex : ex * incl[q - 1];
        trx[q] = ex * incl[q];
        part += trp[q] * (double)om[q] * bb[q];
      }
      double rg = xfac * read_lane_d(wave_incl_sum(part), W - 1);
      if (pi.surf) {
        rg += xfac * trN * radground;
        coG = (T)(xfac * trN * gradground);
      }
      coR = (T)rg;
      // dR/dtau_m = tr_m (B_m - W_m), W_m the radiance entering layer m from
      // below.  D_m = B_m - W_m is carried by its own recurrence
      //   D_{m-1} = (B_{m-1} - B_m) + t_m D_m,   D_{N-1} = B_{N-1} - R_ground
      // (W_{m-1} = B_m (1 - t_m) + t_m W_m), so it keeps its relative accuracy
      // where B and W nearly cancel.  Affine suffix scan over lanes, FP64.
      const double bup0 = wave_shr1(bb[RL - 1]);  // layer above this lane's first
      double dN;
      {
        const int ob = (nlays - 1) / RL, oq = (nlays - 1) - ob * RL;
        double b0 = 0.0;
#pragma unroll
        for (int q = 0; q < RL; q++)
          if (q == oq) b0 = bb[q];
        dN = read_lane_d(b0, ob) - radground;
      }
      double cm[RL];
#pragma unroll
      for (int q = 0; q < RL; q++) {
        const double bu = (q == 0) ? bup0 : bb[q - 1];
        cm[q] = (act[q] && lane * RL + q > 0) ? bu - bb[q] : 0.0;
      }
      double ca = 0.0, cb = 1.0;  // lane composite, applied to D from below
#pragma unroll
      for (int q = RL - 1; q >= 0; q--) {
        const double tm = (double)tl[q];
        ca = cm[q] + tm * ca;
        cb = tm * cb;
      }
      double ia = ca, ib = cb;  // inclusive suffix composite
      wave_suffix_affine(ia, ib, lane);
      // lane + 1's (the identity in the last lane)
      const double xa = dpp_dk<0x130, 0xF>(0.0, ia), xb = dpp_dk<0x130, 0xF>(1.0, ib);
      double dv = xa + xb * dN;  // D of this lane's deepest active layer
      double Dm[RL];
#pragma unroll
      for (int q = RL - 1; q >= 0; q--) {
        Dm[q] = dv;
        dv = cm[q] + (double)tl[q] * dv;
      }
#pragma unroll
      for (int q = 0; q < RL; q++) {
        dRdtau[q] = (T)(xfac * trx[q] * Dm[q]);
        dT[q] = (T)(xfac * trp[q] * (double)om[q]) * db[q];
      }
    } else if (EXT && (pi.imod == 13 || pi.imod == 14)) {
      // SCR sideband and wideband cells (cirsradg_wave.f:1084-1102): the
      // first layer's transmission, no Jacobian, no surface
      const double e = exp(-(double)__shfl(tau[0], 0));
      coR = (T)(pi.imod == 13 ? 1.0 - e : 0.5 * (1.0 + e));
#pragma unroll
      for (int q = 0; q < RL; q++) dRdtau[q] = dT[q] = (T)0;
    } else if (EXT && pi.imod == 18) {
      // disc-integrated emission (cirsradg_wave.f:1116-1218): layer j
      // weighs B_j by E3(tr_{j-1}) - E3(tr_j), the surface by E3(tr_N);
      // xfac carries the 2 pi.  With g_j = dE3/dtr at tr_j
      // and B_N = 0,
      //   dR/dtau_m = sum_{j>=m} (B_j - B_{j+1}) g_j tr_j,
      // the reference's O(L^2) de3dq recurrence (:1171-1186) summed in
      // closed form; a suffix sum over lanes, FP64.
      // the lane's layer transmissions are formed twice (here for the lane
      // product, below sequentially from the lane's entry value, as the
      // reference's tr = trold*tlayer) so no per-layer array stays live
      double run = 1.0;
#pragma unroll
      for (int q = 0; q < RL; q++) run *= act[q] ? (double)dexp(-tau[q]) : 1.0;
      double s = run;
#pragma unroll
      for (int ds = 0; ds < 6; ds++) {
        const int d = 1 << ds;
        const double o = __shfl_up(s, d);
        if (lane >= d) s *= o;
      }
      double ex = __shfl_up(s, 1);
      if (lane == 0) ex = 1.0;
      double e3x[RL], cq[RL];
      double e3up = 0.5, part = 0.0, tot = 0.0;
      {
        const double bnext0 = __shfl_down(bb[0], 1);  // the layer below this lane's last
        double tr = ex;
#pragma unroll
        for (int q = 0; q < RL; q++) {
          const int m = lane * RL + q;
          tr *= act[q] ? (double)dexp(-tau[q]) : 1.0;
          double gq;
          e3interp_d(a.e3tab, tr, e3x[q], gq);
          const double bn = (q == RL - 1) ? bnext0 : bb[q + 1];
          cq[q] = act[q] ? (bb[q] - (m + 1 < nlays ? bn : 0.0)) * gq * tr : 0.0;
          tot += cq[q];
        }
        const double up = __shfl_up(e3x[RL - 1], 1);
        if (lane > 0) e3up = up;
#pragma unroll
        for (int q = 0; q < RL; q++) {
          const double e3p = q == 0 ? e3up : e3x[q - 1];
          if (act[q]) part += (e3p - e3x[q]) * bb[q];
          dT[q] = act[q] ? (T)(xfac * (e3p - e3x[q]) * (double)db[q]) : (T)0;
        }
      }
      // E3 at the bottom of the path: the last active layer's
      double e3N;
      {
        const int ob = (nlays - 1) / RL, oq = (nlays - 1) - ob * RL;
        double v0 = 0.0;
#pragma unroll
        for (int q = 0; q < RL; q++)
          if (q == oq) v0 = e3x[q];
        e3N = __shfl(v0, ob);
      }
      coR = (T)(xfac * (wave_sum(part) + e3N * radground));
      coG = (T)(xfac * e3N * gradground);
      double ia = tot;
#pragma unroll
      for (int ds = 0; ds < 6; ds++) {
        const int d = 1 << ds;
        const double o = __shfl_down(ia, d);
        if (lane + d < W) ia += o;
      }
      double sv = __shfl_down(ia, 1);
      if (lane == W - 1) sv = 0.0;
#pragma unroll
      for (int q = RL - 1; q >= 0; q--) {
        sv += cq[q];
        dRdtau[q] = act[q] ? (T)(xfac * sv) : (T)0;
      }
    } else {
      // the reference's formulas (:863-895) with e^-tau in T and the rest in
      // FP64: imod 1's 1 - (1 - e^-tau) would cancel in FP32 beyond tau ~ 10
      double part = 0.0;
#pragma unroll
      for (int q = 0; q < RL; q++) part += (double)tau[q];
      const double tmp = wave_sum(part);
      const double e = (double)dpexp((T)-tmp);
      if (pi.imod == 0) {
        // x0: the stellar flux with iform 4 (:862-880), else 1
        const double x0 = EXT ? xfac : 1.0;
        coR = (T)(e * x0);
#pragma unroll
        for (int q = 0; q < RL; q++) dRdtau[q] = act[q] ? (T)(-tmp * e * x0) : (T)0;
      } else {
        const double co = 1.0 - e;
        coR = (T)co;
#pragma unroll
        for (int q = 0; q < RL; q++) dRdtau[q] = act[q] ? (T)(tmp * (1.0 - co)) : (T)0;
      }
#pragma unroll
      for (int q = 0; q < RL; q++) dT[q] = 0;
    }
    if (scr) {  // times the cell at this g-ordinate (cirsradg_wave.f:1063-1081)
      const double e = exp(-(double)((tcon[cL] + ktot[(long long)g * L + cL]) * csc));
      const double fc = cim == 13 ? 1.0 - e : 0.5 * (1.0 + e);
      coR = (T)((double)coR * fc);
      coG = (T)((double)coG * fc);
#pragma unroll
      for (int q = 0; q < RL; q++) {
        dRdtau[q] = (T)((double)dRdtau[q] * fc);
        dT[q] = (T)((double)dT[q] * fc);
      }
    }
    rad.add(coR * dg);
    gts.add(coG * dg);
    // Jacobian: dR/dq = sum_g delg_g dR/dtau scale (dtaucon/dq + dk/dq)
    // (+ the Planck term for T).  scale and dtaucon/dq do not depend on g,
    // so each layer keeps sum_g delg_g dR/dtau (s1), the Planck term's
    // g-sum (s3) and, per k row (the gases with a k-table, then T),
    // sum_g delg_g dR/dtau dk/dq (s2, lane-contiguous in LDS), all in FP64;
    // the columns are formed once after the g loop.
    double rd[RL];
#pragma unroll
    for (int q = 0; q < RL; q++) {
      rd[q] = (double)dRdtau[q] * (double)dg;
      s1[q] += rd[q];
      s3[q] += (double)dT[q] * (double)dg;
    }
#pragma unroll
    for (int kd = 0; kd < NKP; kd++) {
      if (kd > nks) break;
#pragma unroll
      for (int q = 0; q < RL; q++) {
        double &acc = s2[(kd * RL + q) * W + lane];
        acc = fma((double)dkv[kd][q], rd[q], acc);
      }
    }
    for (int kd = NKP; kd <= nks; kd++) {
#pragma unroll
      for (int q = 0; q < RL; q++) {
        double &acc = s2[(kd * RL + q) * W + lane];
        acc = fma((double)rows.dk(g, kd, lb[q]), rd[q], acc);
      }
    }
  }
  // ---- outputs ----
  const long long ob = R.out0 + (long long)iw * npath + p;
  if (lane == 0) {
    reinterpret_cast<T *>(a.out)[ob] = rad.s;
    reinterpret_cast<T *>(a.gts)[ob] = gts.s;
  }
  T *jac = reinterpret_cast<T *>(a.jac) + R.jac0 + ((long long)iw * npath + p) * L * nparam;
  // each lane writes whole rows: zeros, then its isw parameters
  // (cirsradg_wave.f:1243-1247 zeroes every row before the g-sum).  The gas
  // columns stay per 1e20 molecules, as accumulated: per molecule they reach
  // 1e-50 and below in the near infrared, under FP32's range.  Every reader
  // (path quadrature, host copies, state-map GEMM) applies the 1e-20 in FP64.
#pragma unroll
  for (int q = 0; q < RL; q++) {
    const int m = lane * RL + q;
    if (m >= L) continue;
    for (int k = 0; k < nparam; k++) jac[(long long)m * nparam + k] = (T)0;
    if (!act[q] || m >= nrows) continue;
    for (int k1 = 0; k1 < nsw; k1++) {
      const int k = R.isw[k1], kd = R.kdrow[k1];
      double c = (double)dtcon[(long long)k * L + Lm[q]] * s1[q];
      if (kd >= 0) c += s2[(kd * RL + q) * W + lane];
      c *= (double)sc[q];
      if (k == ngas) c += s3[q];
      jac[(long long)m * nparam + k] = (T)c;
    }
  }
}
}  // namespace

// Emission as its own launch: one wavefront per (request, wavenumber, path)
// item of a layers-per-lane class, reading the intermediate block in HBM.
// LDS: the FP64 g-sums s2[nks+1][RL][64].
// The FP32 form's common class (up to 128 layers, no imod 8/13/14/18 path)
// is held to 80 VGPRs: six wavefronts per SIMD, without spills (measured
// 1.22 -> 1.09 ms per 16 spectra against five; eight spill and run slower).
template <typename T, int RL, bool EXT>
__global__ __launch_bounds__(64, (sizeof(T) == 4 && RL <= 2 && !EXT) ? 6 : 1) void emission_kernel(
    KernelArgs a, long long nblocks) {
  const long long item = xcd_item(nblocks);
  if (item >= a.nB) return;
  const int lane = threadIdx.x;
  const int rqc = find_req<true>(a.startB, a.nreqB, item, a.nB);
  const DevReq &R = a.reqs[a.reqmapB[rqc]];
  const int L = R.L, ngas = R.ngas, NG = a.ng, npath = R.npath;
  const long long loc = item - a.startB[rqc];
  const int iw = div_item<true>(loc, npath);
  const int p = (int)(loc - (long long)iw * npath);
  const T *inter = reinterpret_cast<const T *>(a.inter) + R.inter0 +
                   (long long)iw * inter_block(NG, L, R.nks, R.nparam);
  const T *dkq = inter + (long long)NG * L;
  const T *tcon = dkq + (long long)NG * (R.nks + 1) * L;
  extern __shared__ double smem_d[];
  emission_path<T, RL, EXT, true>(a, R, iw, p, inter, dkq, tcon, tcon + L, smem_d, lane);
}

// The whole wavenumber in one workgroup of RADG_FUSED_WAVES wavefronts:
// the continuum of every layer (one layer per thread), the layer items
// (round robin over the wavefronts), then the paths (round robin), with the
// layer results kept in LDS: ktot[NG][L] | dk[NG][ngas+1][L] | tcon[L] |
// dtcon[nparam][L] in T after the header, then one scratch area per
// wavefront, shared in time by its layer items and its paths.  Items are
// (request of the class, wavenumber): startB/reqmapB/nreqB/nB describe them.
template <typename T, int E, int BF, int RL, bool EXT>
__global__ __launch_bounds__(64 * RADG_FUSED_WAVES) void fused_kernel(KernelArgs a, long long nblocks) {
  const long long item = xcd_item(nblocks);
  if (item >= a.nB) return;
  const int tid = threadIdx.x;
  const int wv = tid / W, lane = tid - wv * W;
  const int rqc = find_req<true>(a.startB, a.nreqB, item, a.nB);
  const DevReq &R = a.reqs[a.reqmapB[rqc]];
  const int L = R.L, ngas = R.ngas, NG = a.ng;
  const int iw = (int)(item - a.startB[rqc]);
  extern __shared__ double smem_d[];
  char *hdr = reinterpret_cast<char *>(smem_d);
  const int ts = (int)sizeof(T);
  T *ktot = reinterpret_cast<T *>(hdr + layer_header_bytes(NG, E, ts));
  T *dkq = ktot + (long long)NG * L;
  T *tcon = dkq + (long long)NG * (R.nks + 1) * L;
  T *dtcon = tcon + L;
  char *ws = reinterpret_cast<char *>(ktot) +
             lds_round16(ts * (int)inter_block(NG, L, R.nks, R.nparam)) + wv * a.wsbytes;
  layer_header<T, E>(a, hdr, tid, W * RADG_FUSED_WAVES);
  for (int j = tid; j < L; j += W * RADG_FUSED_WAVES) continuum_layer<T>(a, R, iw, j, tcon, dtcon, L);
  __syncthreads();
  for (int j = wv; j < L; j += RADG_FUSED_WAVES)
    layer_item<T, E, BF>(a, R, iw, j, hdr, ws, ktot, dkq, L, lane);
  __syncthreads();
  double *s2 = reinterpret_cast<double *>(ws);
  for (int p = wv; p < R.npath; p += RADG_FUSED_WAVES)
    emission_path<T, RL, EXT, false>(a, R, iw, p, ktot, dkq, tcon, dtcon, s2, lane);
}

// ======================================================= path quadrature
template <typename T>
__global__ __launch_bounds__(64) void quadrature_kernel(KernelArgs a) {
  const long long item = blockIdx.x;
  if (item >= a.nC) return;
  const int lane = threadIdx.x;
  const int rq = find_req<true>(a.startC, a.nreq, item, a.nC);
  const DevReq &R = a.reqs[rq];
  const int iw = (int)(item - R.itemC0);
  const int L = R.L, npath = R.npath, nparam = R.nparam;
  const double *pw = a.pweight + R.path0;
  const T *out = reinterpret_cast<const T *>(a.out) + R.out0 + (long long)iw * npath;
  const T *gt = reinterpret_cast<const T *>(a.gts) + R.out0 + (long long)iw * npath;
  if (lane == 0) {
    double so = 0.0, sg = 0.0;
    for (int p = 0; p < npath; p++) {
      so += pw[p] * (double)out[p];
      sg += pw[p] * (double)gt[p];
    }
    a.wout[R.wout0 + iw] = so;
    a.wgts[R.wout0 + iw] = sg;
  }
  const T *jac = reinterpret_cast<const T *>(a.jac) + R.jac0 + (long long)iw * npath * L * nparam;
  double *wj = a.wjac + R.wjac0 + (long long)iw * L * nparam;
  const long long per = (long long)L * nparam;
  for (long long x = lane; x < per; x += W) {
    double s = 0.0;
    for (int p = 0; p < npath; p++) s += pw[p] * (double)jac[p * per + x];
    // gas columns are stored per 1e20 molecules
    wj[x] = (x % nparam < R.ngas) ? s * 1e-20 : s;
  }
}

// ============================================================ launchers
namespace {
template <typename T, int E, int BF = 0>
hipError_t launch_layer_E(const KernelArgs &a, size_t lds, hipStream_t st) {
  if (a.lwaves < 1 || a.lwaves > 4) return hipErrorInvalidValue;
  long long nb = ((a.nA + a.lwaves - 1) / a.lwaves + 7) / 8 * 8;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL((layer_kernel<T, E, BF>), dim3((unsigned)nb), dim3(64 * a.lwaves), lds, st, a,
                     nb);
  return hipGetLastError();
}
template <typename T, int RL>
hipError_t launch_em_RL(const KernelArgs &a, bool disc, size_t lds, hipStream_t st) {
  long long nb = ((a.nB + 7) / 8) * 8;
  if (nb == 0) return hipSuccess;
  if (disc)
    hipLaunchKernelGGL((emission_kernel<T, RL, true>), dim3((unsigned)nb), dim3(64), lds, st, a, nb);
  else
    hipLaunchKernelGGL((emission_kernel<T, RL, false>), dim3((unsigned)nb), dim3(64), lds, st, a, nb);
  return hipGetLastError();
}
template <typename T, int E, int BF>
hipError_t launch_fused_E(const KernelArgs &a, int RL, size_t lds, hipStream_t st) {
  long long nb = ((a.nB + 7) / 8) * 8;
  if (nb == 0) return hipSuccess;
  const dim3 blk(64 * RADG_FUSED_WAVES);
  if (RL <= 2) hipLaunchKernelGGL((fused_kernel<T, E, BF, 2, true>), dim3((unsigned)nb), blk, lds, st, a, nb);
  else if (RL == 4) hipLaunchKernelGGL((fused_kernel<T, E, BF, 4, true>), dim3((unsigned)nb), blk, lds, st, a, nb);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}
}  // namespace

// Layer items per workgroup: up to four share the header while the
// workgroup's LDS stays within 48 KiB, so that several workgroups fit a CU.
int radg_layer_waves(int ng, int ngas, int E, int fp64) {
  const int ts = fp64 ? 8 : 4;
  const int h = layer_header_bytes(ng, E, ts), w = layer_layout(ng, ngas, E, ts).total;
  int lw = 4;
  while (lw > 1 && h + lw * w > 48 * 1024) lw >>= 1;
  return lw;
}

int radg_layer_stride(int ng, int ngas, int E, int fp64) {
  return layer_layout(ng, ngas, E, fp64 ? 8 : 4).total;
}

size_t radg_lds_layer_bytes(int ng, int ngas, int E, int fp64) {
  const int ts = fp64 ? 8 : 4;
  return (size_t)layer_header_bytes(ng, E, ts) +
         (size_t)radg_layer_waves(ng, ngas, E, fp64) * (size_t)layer_layout(ng, ngas, E, ts).total;
}

size_t radg_lds_emission_bytes(int RL, int nkrow, int fp64) {
  (void)fp64;  // the g-sums are FP64 in both forms
  return (size_t)RL * W * nkrow * sizeof(double) + 16;
}

// The fused kernel's LDS for requests of up to maxL layers, maxgas gases,
// maxnsw Jacobian parameters and maxblock intermediate elements per
// wavenumber (inter_block), with emission's RL layers per lane; *ws gets the
// scratch bytes per wavefront.  0: no fused form for this (mode, E, RL).
size_t radg_lds_fused_bytes(int mode, int ng, int E, int RL, int maxgas, int maxnsw,
                            long long maxblock, int *ws) {
  const bool have = (mode >= 2) ? (E == 8) : (E == 1 || E == 2 || E == 4 || E == 8);
  if (!have || RL > 4) return 0;
  const int ts = mode == 0 ? 8 : 4;
  const int rl = RL <= 2 ? 2 : 4;
  int w = layer_layout(ng, maxgas, E, ts).total;
  (void)maxnsw;
  const int em = lds_round16(rl * W * (maxgas + 1) * (int)sizeof(double));
  if (w < em) w = em;
  *ws = w;
  return (size_t)layer_header_bytes(ng, E, ts) + (size_t)lds_round16(ts * (int)maxblock) +
         (size_t)RADG_FUSED_WAVES * w;
}

// Slots per lane of an overlap sort: 64 E >= ng^2, with the rows of the
// slot layout (layer_lpr) fitting the 64 lanes.
int radg_layer_E(int ng) {
  const int n = ng * ng;
  int E = 1;
  while (W * E < n || ng * layer_lpr(ng, E) > W) E <<= 1;
  return E;
}

int radg_emission_RL(int maxlays) {
  int RL = 1;
  while (W * RL < maxlays) RL <<= 1;
  return RL;
}

// mode: 0 FP64, 1 mixed (FP32 table), 2 / 3 mixed with BF16 / IEEE-half
// log10(k) tables
hipError_t radg_launch_layer_lds(const KernelArgs &a, int mode, int E, size_t lds, hipStream_t st) {
  if (a.nA > 0) {
    const unsigned nb = (unsigned)((a.nA + 255) / 256);
    if (mode == 0) hipLaunchKernelGGL(cia_kernel<double>, dim3(nb), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(cia_kernel<float>, dim3(nb), dim3(256), 0, st, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (a.lbl) {  // line-by-line tables: one thread per item, no LDS
    if (a.ng != 1) return hipErrorInvalidValue;
    if (a.nA == 0) return hipSuccess;
    const unsigned nb = (unsigned)((a.nA + 255) / 256);
    switch (mode) {
      case 0: hipLaunchKernelGGL((lbl_kernel<double, 0>), dim3(nb), dim3(256), 0, st, a); break;
      case 1: hipLaunchKernelGGL((lbl_kernel<float, 0>), dim3(nb), dim3(256), 0, st, a); break;
      case 2: hipLaunchKernelGGL((lbl_kernel<float, 1>), dim3(nb), dim3(256), 0, st, a); break;
      case 3: hipLaunchKernelGGL((lbl_kernel<float, 2>), dim3(nb), dim3(256), 0, st, a); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  if (mode == 0) {
    switch (E) {
      case 1: return launch_layer_E<double, 1>(a, lds, st);
      case 2: return launch_layer_E<double, 2>(a, lds, st);
      case 4: return launch_layer_E<double, 4>(a, lds, st);
      case 8: return launch_layer_E<double, 8>(a, lds, st);
      case 16: return launch_layer_E<double, 16>(a, lds, st);
    }
  } else if (mode == 2) {
    switch (E) {
      case 1: return launch_layer_E<float, 1, 1>(a, lds, st);
      case 2: return launch_layer_E<float, 2, 1>(a, lds, st);
      case 4: return launch_layer_E<float, 4, 1>(a, lds, st);
      case 8: return launch_layer_E<float, 8, 1>(a, lds, st);
      case 16: return launch_layer_E<float, 16, 1>(a, lds, st);
    }
  } else if (mode == 3) {
    switch (E) {
      case 1: return launch_layer_E<float, 1, 2>(a, lds, st);
      case 2: return launch_layer_E<float, 2, 2>(a, lds, st);
      case 4: return launch_layer_E<float, 4, 2>(a, lds, st);
      case 8: return launch_layer_E<float, 8, 2>(a, lds, st);
      case 16: return launch_layer_E<float, 16, 2>(a, lds, st);
    }
  } else {
    switch (E) {
      case 1: return launch_layer_E<float, 1>(a, lds, st);
      case 2: return launch_layer_E<float, 2>(a, lds, st);
      case 4: return launch_layer_E<float, 4>(a, lds, st);
      case 8: return launch_layer_E<float, 8>(a, lds, st);
      case 16: return launch_layer_E<float, 16>(a, lds, st);
    }
  }
  return hipErrorInvalidValue;
}

// The fused kernel over the wavenumber items a.startB/a.nB of one class.
hipError_t radg_launch_fused(const KernelArgs &a, int mode, int E, int RL, size_t lds, hipStream_t st) {
  if (mode == 0) {
    switch (E) {
      case 1: return launch_fused_E<double, 1, 0>(a, RL, lds, st);
      case 2: return launch_fused_E<double, 2, 0>(a, RL, lds, st);
      case 4: return launch_fused_E<double, 4, 0>(a, RL, lds, st);
      case 8: return launch_fused_E<double, 8, 0>(a, RL, lds, st);
    }
  } else if (mode == 2) {
    if (E == 8) return launch_fused_E<float, 8, 1>(a, RL, lds, st);
  } else if (mode == 3) {
    if (E == 8) return launch_fused_E<float, 8, 2>(a, RL, lds, st);
  } else {
    switch (E) {
      case 1: return launch_fused_E<float, 1, 0>(a, RL, lds, st);
      case 2: return launch_fused_E<float, 2, 0>(a, RL, lds, st);
      case 4: return launch_fused_E<float, 4, 0>(a, RL, lds, st);
      case 8: return launch_fused_E<float, 8, 0>(a, RL, lds, st);
    }
  }
  return hipErrorInvalidValue;
}

hipError_t radg_launch_emission(const KernelArgs &a, int fp64, int RL, bool disc, size_t lds, hipStream_t st) {
  if (fp64) {
    switch (RL) {
      case 1: return launch_em_RL<double, 1>(a, disc, lds, st);
      case 2: return launch_em_RL<double, 2>(a, disc, lds, st);
      case 4: return launch_em_RL<double, 4>(a, disc, lds, st);
      case 8: return launch_em_RL<double, 8>(a, disc, lds, st);
    }
  } else {
    switch (RL) {
      case 1: return launch_em_RL<float, 1>(a, disc, lds, st);
      case 2: return launch_em_RL<float, 2>(a, disc, lds, st);
      case 4: return launch_em_RL<float, 4>(a, disc, lds, st);
      case 8: return launch_em_RL<float, 8>(a, disc, lds, st);
    }
  }
  return hipErrorInvalidValue;
}

hipError_t radg_launch_quadrature(const KernelArgs &a, int fp64, hipStream_t st) {
  if (a.nC == 0) return hipSuccess;
  if (fp64)
    hipLaunchKernelGGL((quadrature_kernel<double>), dim3((unsigned)a.nC), dim3(64), 0, st, a);
  else
    hipLaunchKernelGGL((quadrature_kernel<float>), dim3((unsigned)a.nC), dim3(64), 0, st, a);
  return hipGetLastError();
}

// ======================================================= table storage
// The FP32 and BF16 copies of a registered k-table pool are made on the
// device from its FP64 master, once, on first use.
namespace {
__global__ void to_f32_kernel(const double *src, float *dst, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

// Offset of each (slice, ordinate): the FP32 rounding of the mean of log10 k
// over the slice's positive values at that ordinate (0 when there are none),
// summed in pool order.  per: values per ordinate of a slice (np * nt).
__global__ void slice_offsets_kernel(const double *src, float *off, long long nslice, int per, int ng) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslice * ng) return;
  const long long sl = t / ng;
  const int g = (int)(t - sl * ng);
  const double *p = src + sl * per * ng + g;
  double sum = 0.0;
  int cnt = 0;
  for (int i = 0; i < per; i++) {
    const double k = p[(long long)i * ng];
    if (k > 0.0) {
      sum += log10(k);
      cnt++;
    }
  }
  off[t] = cnt > 0 ? (float)(sum / cnt) : 0.0f;
}

// IEEE half of an FP32 value, round to nearest even, by integer operations
// on its bits (so that the FP32 rounding before it is never folded into one
// rounding from FP64); the values here lie far below the half range's top.
__device__ __forceinline__ unsigned short f32_to_f16_rne(float x) {
  const unsigned int u = __float_as_uint(x);
  const unsigned int sign = (u >> 16) & 0x8000u;
  const int e = (int)((u >> 23) & 0xffu) - 127 + 15;
  unsigned int m = (u & 0x7fffffu) | 0x800000u;  // significand with its leading 1
  if ((u & 0x7fffffffu) == 0u) return (unsigned short)sign;
  if (e >= 31) return (unsigned short)(sign | 0x7c00u);
  // bits to drop: 13 for a normal half, more below the normal range
  const int shift = e >= 1 ? 13 : 14 - e;
  if (shift > 25) return (unsigned short)sign;
  const unsigned int kept = m >> shift;
  const unsigned int rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1);
  unsigned int h = (e >= 1) ? (((unsigned int)e << 10) | (kept & 0x3ffu)) : kept;
  if (rem > half || (rem == half && (h & 1u))) h++;  // a carry moves into the exponent
  return (unsigned short)(sign | h);
}

// code = round16(float(log10 k - off[slice, g])), round to nearest even;
// k <= 0 -> -inf.  BF: 1 BF16, 2 IEEE half.
template <int BF>
__global__ void to_code16_kernel(const double *src, const float *off, unsigned short *dst, long long n,
                                 int ntab, int ng) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double k = src[i];
  if (!(k > 0.0)) {
    dst[i] = (unsigned short)(BF == 1 ? RADG_BF16_NONPOS : RADG_F16_NONPOS);
    return;
  }
  const long long sl = i / ntab;
  const int g = (int)(i % ng);
  const float x = (float)(log10(k) - (double)off[sl * ng + g]);
  if constexpr (BF == 1) {
    unsigned int b = __float_as_uint(x);
    b += 0x7FFFu + ((b >> 16) & 1u);
    dst[i] = (unsigned short)(b >> 16);
  } else {
    dst[i] = f32_to_f16_rne(x);
  }
}
}  // namespace

hipError_t radg_convert_f32(const double *src, float *dst, long long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
  return hipGetLastError();
}

// The per-(slice, ordinate) offsets of a pool of nslice slices of ntab
// values (ng ordinates fastest), then its 16-bit codes (bf: 1 BF16, 2 half).
hipError_t radg_slice_offsets(const double *src, float *off, long long nslice, int ntab, int ng,
                              hipStream_t st) {
  const long long n = nslice * ng;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(slice_offsets_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src,
                     off, nslice, ntab / ng, ng);
  return hipGetLastError();
}

hipError_t radg_convert_code16(int bf, const double *src, const float *off, unsigned short *dst,
                               long long n, int ntab, int ng, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const dim3 grid((unsigned)((n + 255) / 256));
  if (bf == 1) hipLaunchKernelGGL(to_code16_kernel<1>, grid, dim3(256), 0, st, src, off, dst, n, ntab, ng);
  else hipLaunchKernelGGL(to_code16_kernel<2>, grid, dim3(256), 0, st, src, off, dst, n, ntab, ng);
  return hipGetLastError();
}
