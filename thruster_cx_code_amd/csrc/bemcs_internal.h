// Internal helpers shared by the translation units of libbemcs_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/bemcs_hip.h"
#include "bemcs_dmath.h"

namespace bemcs {

void set_error(const char* fmt, ...);

#define BEMCS_CHECK_HIP(expr)                                                   \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      ::bemcs::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,              \
                         hipGetErrorString(e_));                                 \
      return BEMCS_EHIP;                                                         \
    }                                                                            \
  } while (0)

#define BEMCS_CHECK_LAUNCH()                                                    \
  do {                                                                           \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess) {                                                      \
      ::bemcs::set_error("%s:%d kernel launch: %s", __FILE__, __LINE__,          \
                         hipGetErrorString(e_));                                 \
      return BEMCS_EHIP;                                                         \
    }                                                                            \
  } while (0)

#define BEMCS_REQUIRE(cond, msg)                                                \
  do {                                                                           \
    if (!(cond)) {                                                               \
      ::bemcs::set_error("%s", msg);                                             \
      return BEMCS_EINVAL;                                                       \
    }                                                                            \
  } while (0)

inline hipStream_t as_stream(bemcs_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : g);
}

// ---------------------------------------------------------------- division
// Correctly rounded float32 quotient a / b by a fixed divisor.  The FAST form
// is Markstein's reciprocal iteration: q0 = a*r with r = RN(1/b), the
// remainder a - b*q0 taken exactly by an FMA, and one correction
// q = q0 + (a - b*q0)*r.  It costs a multiply and two FMAs instead of the
// scaled Newton sequence of IEEE division.  The library enables it for a
// divisor only after comparing it with IEEE division on every one of the 2^32
// float32 dividends on the device (fastdiv_ok), so the quotients, and every
// gather index and weight derived from them, are bit-identical.
struct Divisor {
  float b;  // divisor
  float r;  // RN(1 / b)
};

template <bool FAST>
__device__ __forceinline__ float div_rn(float a, Divisor d) {
  // Zeros (signed-zero quotient), subnormal remainders and a*r overflow are
  // left to IEEE division; positions never reach those magnitudes, so the
  // branch costs a compare.
  if (FAST && fabsf(a) >= 0x1p-100f && fabsf(a) <= 0x1p100f) {
    float q0 = a * d.r;
    float e = fmaf(-q0, d.b, a);
    return fmaf(e, d.r, q0);
  }
  return a / d.b;
}

// True when div_rn<true> equals IEEE division for divisor b on all float32
// dividends (checked once per divisor and cached; the first check of a new
// divisor synchronises `s`).  BEMCS_FASTDIV=0 in the environment disables it.
bool fastdiv_ok(float b, hipStream_t s);
inline Divisor make_divisor(float b) { return Divisor{b, 1.0f / b}; }

// ----------------------------------------------------------------- indices
// NumPy cast semantics on x86: a float->int conversion that is out of range
// or NaN yields the most negative integer, which the reference's np.clip then
// maps to the lower bound.

// np.clip(np.floor(g).astype(np.int32), 0, hi)
__device__ __forceinline__ int floor_i32_np(float g, int hi) {
  float f = floorf(g);
  int i = (f >= -2147483648.0f && f < 2147483648.0f) ? (int)f : INT32_MIN;
  return i < 0 ? 0 : (i > hi ? hi : i);
}

// floor_i32_np(g, hi) in five instructions, for the push's inner loop: the
// cast maps values it cannot hold (NaN, f >= 2^31, f < -2^31) to INT32_MIN,
// which the clip takes to 0; here NaN and f >= 2^31 become -1 and every
// negative value clips to 0, so every g gives the same index.  hi_f =
// (float)hi, exact for hi < 2^24.
__device__ __forceinline__ int floor_i32_np_med3(float g, float hi_f) {
  const float f = floorf(g);
  const float t = f < 2147483648.0f ? f : -1.0f;
  return (int)__builtin_amdgcn_fmed3f(t, 0.0f, hi_f);
}

// np.clip(np.floor(g).astype(int), 0, hi)   (int64)
__device__ __forceinline__ int floor_i64_np(float g, int hi) {
  float f = floorf(g);
  if (!(f >= -9.2e18f && f < 9.2e18f)) return 0;
  return f < 0.0f ? 0 : (f > (float)hi ? hi : (int)f);
}

// np.clip(np.round(g).astype(int), lo, hi)   (int64, half to even)
__device__ __forceinline__ int round_i64_np(float g, int lo, int hi) {
  float f = rintf(g);
  if (!(f >= -9.2e18f && f < 9.2e18f)) return lo;
  return f < (float)lo ? lo : (f > (float)hi ? hi : (int)f);
}

// np.clip(np.round(g).astype(np.int32), lo, hi)
__device__ __forceinline__ int round_i32_np(float g, int lo, int hi) {
  float f = rintf(g);
  int i = (f >= -2147483648.0f && f < 2147483648.0f) ? (int)f : INT32_MIN;
  return i < lo ? lo : (i > hi ? hi : i);
}

// np.clip((g).astype(int), 0, hi)   (truncation, int64)
__device__ __forceinline__ int trunc_i64_np(float g, int hi) {
  if (!(g >= -9.2e18f && g < 9.2e18f)) return 0;
  float t = truncf(g);
  return t < 0.0f ? 0 : (t > (float)hi ? hi : (int)t);
}

// np.maximum(a, b) for a possibly NaN, b not NaN
__host__ __device__ __forceinline__ double nmax(double a, double b) { return (a < b) ? b : a; }

// np.mod(a, b) for float32 (npy_divmodf)
__device__ __forceinline__ float np_modf(float a, float b) {
  float m = fmodf(a, b);
  if (m != 0.0f) {
    if ((b < 0.0f) != (m < 0.0f)) m = m + b;
  } else {
    m = copysignf(0.0f, b);
  }
  return m;
}

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Reserve one slot per lane with pred set; returns the slot or -1.
template <typename T>
__device__ __forceinline__ int64_t wave_append(T* counter, bool pred) {
  uint64_t mask = __ballot(pred);
  if (mask == 0) return -1;
  int lane = lane_id();
  int leader = __ffsll((unsigned long long)mask) - 1;
  int total = __popcll(mask);
  T base = 0;
  if (lane == leader) base = atomicAdd(counter, (T)total);
  base = __shfl(base, leader);
  uint64_t below = (lane == 0) ? 0ull : (mask & ((~0ull) >> (64 - lane)));
  int off = __popcll(below);
  return pred ? (int64_t)base + off : -1;
}

// Append slot of this lane's element in a compaction with one atomic per
// work-group instead of one per wavefront (every wavefront of a dense
// compaction appends, and the adds to one counter serialise at its L2
// channel).  Every thread of the work-group must call it with the same
// counter; at most 16 wavefronts per work-group.  Returns -1 where pred is
// false.
template <typename T>
__device__ __forceinline__ int64_t block_append(T* counter, bool pred) {
  __shared__ int s_cnt[16];
  __shared__ T s_base;
  const uint64_t mask = __ballot(pred);
  const int wave = (int)(threadIdx.x >> 6), lane = lane_id();
  const int nwaves = (int)((blockDim.x + 63) >> 6);
  if (lane == 0) s_cnt[wave] = __popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = 0;
    for (int w = 0; w < nwaves; ++w) total += s_cnt[w];
    s_base = total ? atomicAdd(counter, (T)total) : (T)0;
  }
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += s_cnt[w];
  const int64_t base = (int64_t)s_base + before;
  __syncthreads();  // s_cnt and s_base are reused by the next call
  const uint64_t below = (lane == 0) ? 0ull : (mask & ((~0ull) >> (64 - lane)));
  return pred ? base + __popcll(below) : -1;
}

// Count of lanes with pred set, added once per wave.
__device__ __forceinline__ void wave_count(uint32_t* counter, bool pred) {
  uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  if (lane_id() == (__ffsll((unsigned long long)mask) - 1)) atomicAdd(counter, (uint32_t)__popcll(mask));
}

// Histogram add of key (key < 0: none).  Runs of equal keys in consecutive
// lanes (cell-sorted particles) are merged into one atomic.
__device__ __forceinline__ void wave_hist_add(uint32_t* hist, int key) {
  int lane = lane_id();
  int prev = __shfl_up(key, 1);
  bool head = (lane == 0) || (key != prev);
  uint64_t heads = __ballot(head);
  if (head && key >= 0) {
    uint64_t after = (lane == 63) ? 0ull : (heads >> (lane + 1));
    int run = after ? (__ffsll((unsigned long long)after)) : (64 - lane);
    atomicAdd(&hist[key], (uint32_t)run);
  }
}

// Adds a non-negative float32 value to a 64-bit fixed-point accumulator as
// rint(v * scale).  Integer sums do not depend on the order of the adds, so
// per-node totals are deterministic.  Values that are not finite, negative or
// beyond 2^53 units set bit 32 of *flags instead of being added.
__device__ __forceinline__ void fixed_add(uint64_t* acc, float v, double scale, uint32_t* flags) {
  double s = (double)v * scale;
  if (!(s >= 0.0 && s <= 9007199254740992.0)) {
    atomicOr(flags, 32u);
    return;
  }
  unsigned long long q = (unsigned long long)rint(s);
  if (q) atomicAdd((unsigned long long*)acc, q);
}

// Fixed point of the bilinear deposit's weight sums: 2^32 units per unit
// weight (signed 64-bit, stored as uint64 two's complement).
constexpr double kCicScale = 4294967296.0;
constexpr double kCicUnit = 1.0 / 4294967296.0;

// Cubic B-spline (knots t[n], coefficients c[n-4]) at x, as FITPACK's splev
// evaluates it: the knot interval l with t[l] <= x < t[l+1] (the last one for
// x at the right end) and fpbspl's Cox-de Boor recursion in its operation
// order.  x must lie in [t[3], t[n-4]].
__device__ __forceinline__ double bspline3(const double* __restrict__ t, const double* __restrict__ c,
                                           int n, double x) {
  int lo = 3, hi = n - 5;  // largest l in [3, n-5] with t[l] <= x
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  const int l = lo;
  double h[4] = {1.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 1; j <= 3; ++j) {
    double hh[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) hh[i] = h[i];
    h[0] = 0.0;
#pragma unroll
    for (int i = 1; i <= 3; ++i) {
      if (i > j) break;
      const double tl = t[l + i], tj = t[l + i - j];
      if (tl != tj) {
        const double f = hh[i - 1] / (tl - tj);
        h[i - 1] = h[i - 1] + f * (tl - x);
        h[i] = f * (x - tj);
      } else {
        h[i] = 0.0;
      }
    }
  }
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) s = s + c[l - 3 + j] * h[j];
  return s;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

}  // namespace bemcs
