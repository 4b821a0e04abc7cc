// Particle kernels of the MI355X time step: gather + Boris push, NGP
// deposit, substep count, fused per-species step (push, hits, SEE, CEX MCC,
// removal, deposit and PPC counts), injection and cell sort.
//
// Layout: one species is a structure of float32 arrays (x, y, vx, vy, vz)
// plus a uint32 tag (ion id | isCEX<<31); one thread owns one particle and
// keeps it in registers for all n_sub substeps, so a step reads and writes
// the particle state from HBM exactly once.  The field is interleaved
// (Ex, Ey) float32 so one 8-byte load fetches both components of a corner.
//
// Arithmetic follows the reference NumPy path (Python/physics_engine.py
// :188-254): float32 positions and velocities, float64 interpolation weights
// and kicks (float32 - int32 promotes to float64 there), float32 drift.  With
// -ffp-contract=off every result is bit-identical to it on the same inputs.
#include <limits.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <unordered_map>

#include "bemcs_internal.h"

namespace bemcs {

constexpr int kBlock = 256;

// ---------------------------------------------------- LDS-privatised counts
// Per-tile privatisation of the NGP and PPC count histograms: a work-group
// counts its particles' nodes in an LDS window of W x H nodes anchored at its
// smallest (ix, iy) (cell-sorted particles of one work-group lie in a few
// cells), with runs of equal nodes in a wavefront merged before the LDS
// atomics, then adds each touched window node to the global histogram with one
// atomic.  Nodes outside the window go straight to global atomics.  Integer
// counts, so the totals do not depend on the order of the adds or on the
// window shape, which the caller picks (count_window: 0 = 64 x 32, else
// log2 W | log2 H << 8 | 1 << 16, W x H <= kWinBins; anything else falls back
// to 64 x 32).
constexpr int kWinLog2 = 11, kWinBins = 1 << kWinLog2;
struct LdsHist {
  uint32_t bins[kWinBins];
  int mins[2][kBlock / 64];
};
struct CountWindow {
  int xs, ys;  // log2 of the window's width and height
};
__host__ __device__ __forceinline__ bool count_window_valid(int32_t w) {
  const int xs = w & 0xff, ys = (w >> 8) & 0xff;
  return w == 0 || ((w >> 16) == 1 && xs + ys <= kWinLog2);
}
__device__ __forceinline__ CountWindow count_window(int32_t w) {
  if (w == 0 || !count_window_valid(w)) return CountWindow{6, 5};
  return CountWindow{w & 0xff, (w >> 8) & 0xff};
}

// Every thread of the work-group must call it with K nodes (ix[k] < 0 adds
// nothing).
template <int K>
__device__ __forceinline__ void block_hist_add(LdsHist& h, uint32_t* hist, int nx, const int (&ix)[K],
                                               const int (&iy)[K], CountWindow win) {
  const int W = 1 << win.xs, H = 1 << win.ys, NB = W * H;
  int mx = INT_MAX, my = INT_MAX;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (ix[k] >= 0) {
      mx = min(mx, ix[k]);
      my = min(my, iy[k]);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    mx = min(mx, __shfl_xor(mx, o));
    my = min(my, __shfl_xor(my, o));
  }
  const int wave = threadIdx.x >> 6;
  if (lane_id() == 0) {
    h.mins[0][wave] = mx;
    h.mins[1][wave] = my;
  }
  for (int k = threadIdx.x; k < NB; k += kBlock) h.bins[k] = 0u;
  __syncthreads();
  int x0 = h.mins[0][0], y0 = h.mins[1][0];
  for (int w = 1; w < kBlock / 64; ++w) {
    x0 = min(x0, h.mins[0][w]);
    y0 = min(y0, h.mins[1][w]);
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool valid = ix[k] >= 0;
    const int lx = ix[k] - x0, ly = iy[k] - y0;
    const bool in_win = valid && lx < W && ly < H;
    wave_hist_add(h.bins, in_win ? (ly << win.xs) + lx : -1);
    wave_hist_add(hist, (valid && !in_win) ? iy[k] * nx + ix[k] : -1);
  }
  __syncthreads();
  if (x0 != INT_MAX) {
    for (int k = threadIdx.x; k < NB; k += kBlock) {
      const uint32_t c = h.bins[k];
      if (c) atomicAdd(&hist[(y0 + (k >> win.xs)) * nx + (x0 + (k & (W - 1)))], c);
    }
  }
}

// ------------------------------------------------------------------ push
// The field at the gather corners: float32, either two arrays or
// interleaved (Ex, Ey) pairs.
typedef float f32x4_a8 __attribute__((ext_vector_type(4), aligned(8)));
template <bool INTERLEAVED, typename T = float>
struct EField {
  using value_type = T;
  const T* ex;
  const T* ey;
  __device__ __forceinline__ void corner(int idx, T& a, T& b) const {
    if (INTERLEAVED) {
      const T* q = ex + 2 * (int64_t)idx;
      a = q[0];
      b = q[1];
    } else {
      a = ex[idx];
      b = ey[idx];
    }
  }
  // (Ex, Ey) at nodes (i0, j0), (i0 + 1, j0), (i0, j0 + 1), (i0 + 1, j0 + 1)
  // as e[0..7] = e00x, e00y, e10x, e10y, e01x, e01y, e11x, e11y.  Interleaved
  // float32 fields load each row pair with one 16-byte load at a 32-bit byte
  // offset from the base (24-bit row product: 0 <= i0, j0 and nx below 2^24,
  // meshes below 2^28 nodes, which the entry points check).
  __device__ __forceinline__ void corners(int i0, int j0, int nx, T (&e)[8]) const {
    if constexpr (INTERLEAVED && sizeof(T) == 4) {
      const uint32_t off = (__umul24((uint32_t)j0, (uint32_t)nx) + (uint32_t)i0) << 3;
      const char* b = reinterpret_cast<const char*>(ex);
      const f32x4_a8 r0 = *reinterpret_cast<const f32x4_a8*>(b + off);
      const f32x4_a8 r1 = *reinterpret_cast<const f32x4_a8*>(b + (off + ((uint32_t)nx << 3)));
      e[0] = r0.x; e[1] = r0.y; e[2] = r0.z; e[3] = r0.w;
      e[4] = r1.x; e[5] = r1.y; e[6] = r1.z; e[7] = r1.w;
    } else {
      const int c00 = j0 * nx + i0, c01 = c00 + nx;
      corner(c00, e[0], e[1]);
      corner(c00 + 1, e[2], e[3]);
      corner(c01, e[4], e[5]);
      corner(c01 + 1, e[6], e[7]);
    }
  }
  // The same corners from the clipped cell coordinates as floats: the node
  // index j0 * nx + i0 is fma(cy, nx, cx), exact on meshes of at most 2^24
  // nodes, and both rows load at one 32-bit byte offset, the second from the
  // uniform base row1 = field + 2 nx floats.
  __device__ __forceinline__ void corners_f(float cx, float cy, float nxf, const char* row1,
                                            T (&e)[8]) const {
    static_assert(INTERLEAVED && sizeof(T) == 4, "float32 interleaved field only");
    const uint32_t off = (uint32_t)fmaf(cy, nxf, cx) << 3;
    const f32x4_a8 r0 = *reinterpret_cast<const f32x4_a8*>(reinterpret_cast<const char*>(ex) + off);
    const f32x4_a8 r1 = *reinterpret_cast<const f32x4_a8*>(row1 + off);
    e[0] = r0.x; e[1] = r0.y; e[2] = r0.z; e[3] = r0.w;
    e[4] = r1.x; e[5] = r1.y; e[6] = r1.z; e[7] = r1.w;
  }
};

// The float64 part of a substep without magnetic field: bilinear weights of
// the fractions (fx, fy) = (gx - i0, gy - j0) in the cell (i0, j0), the field
// e[8] at its corners (as
// EField::corners returns them) gathered, the half kicks (with B = 0 the
// rotation adds exact zeros: v+ = v-) and the float32 drift.
template <typename FT>
__device__ __forceinline__ void kick_drift(double fx, double fy, const FT (&e)[8], float& x, float& y,
                                           float& vx, float& vy, double h, float dt32) {
  double a00 = (1.0 - fx) * (1.0 - fy);
  double a10 = fx * (1.0 - fy);
  double a01 = (1.0 - fx) * fy;
  double a11 = fx * fy;
  double ex = (double)e[0] * a00 + (double)e[2] * a10 + (double)e[4] * a01 + (double)e[6] * a11;
  double ey = (double)e[1] * a00 + (double)e[3] * a10 + (double)e[5] * a01 + (double)e[7] * a11;
  double kx = h * ex;
  double ky = h * ey;
  double ux = (double)vx + kx;
  double uy = (double)vy + ky;
  vx = (float)(ux + kx);
  vy = (float)(uy + ky);
  x = x + (vx * dt32) * 1000.0f;
  y = y + (vy * dt32) * 1000.0f;
}

// One gather + Boris + drift substep.  h = 0.5*q_m*dt (f64), dt32 = f32(dt);
// dx, dy are the float32 cell sizes.
template <class FieldT, bool HAS_B, bool FAST>
__device__ __forceinline__ void boris_substep(float& x, float& y, float& vx, float& vy, float& vz,
                                              const FieldT& E, const float* Bx,
                                              const float* By, const float* Bz, int nx, int ny,
                                              Divisor dx, Divisor dy, double h, float dt32) {
  using FT = typename FieldT::value_type;
  float gx = div_rn<FAST>(x, dx);
  float gy = div_rn<FAST>(y, dy);
  int i0 = floor_i32_np_med3(gx, (float)(nx - 2));
  int j0 = floor_i32_np_med3(gy, (float)(ny - 2));
  FT e[8];
  E.corners(i0, j0, nx, e);
  if (!HAS_B) {
    kick_drift((double)gx - (double)i0, (double)gy - (double)j0, e, x, y, vx, vy, h, dt32);
    return;
  } else {
    double fx = (double)gx - (double)i0;
    double fy = (double)gy - (double)j0;
    double a00 = (1.0 - fx) * (1.0 - fy);
    double a10 = fx * (1.0 - fy);
    double a01 = (1.0 - fx) * fy;
    double a11 = fx * fy;
    const FT e00x = e[0], e00y = e[1], e10x = e[2], e10y = e[3], e01x = e[4], e01y = e[5], e11x = e[6],
             e11y = e[7];
    double ex = (double)e00x * a00 + (double)e10x * a10 + (double)e01x * a01 + (double)e11x * a11;
    double ey = (double)e00y * a00 + (double)e10y * a10 + (double)e01y * a01 + (double)e11y * a11;
    double kx = h * ex;
    double ky = h * ey;
    double ux = (double)vx + kx;
    double uy = (double)vy + ky;
    const int c00 = j0 * nx + i0, c10 = c00 + 1, c01 = c00 + nx, c11 = c01 + 1;
    double bx = (double)Bx[c00] * a00 + (double)Bx[c10] * a10 + (double)Bx[c01] * a01 +
                (double)Bx[c11] * a11;
    double by = (double)By[c00] * a00 + (double)By[c10] * a10 + (double)By[c01] * a01 +
                (double)By[c11] * a11;
    double bz = (double)Bz[c00] * a00 + (double)Bz[c10] * a10 + (double)Bz[c01] * a01 +
                (double)Bz[c11] * a11;
    double uz = (double)vz;
    double tx = h * bx, ty = h * by, tz = h * bz;
    double tt = tx * tx + ty * ty + tz * tz;
    double sx = 2.0 * tx / (1.0 + tt);
    double sy = 2.0 * ty / (1.0 + tt);
    double sz = 2.0 * tz / (1.0 + tt);
    double wx = ux + (uy * tz - uz * ty);
    double wy = uy + (uz * tx - ux * tz);
    double wz = uz + (ux * ty - uy * tx);
    double px = ux + (wy * sz - wz * sy);
    double py = uy + (wz * sx - wx * sz);
    double pz = uz + (wx * sy - wy * sx);
    vx = (float)(px + kx);
    vy = (float)(py + ky);
    vz = (float)pz;
  }
  x = x + (vx * dt32) * 1000.0f;
  y = y + (vy * dt32) * 1000.0f;
}

template <bool HAS_B, bool FAST>
__global__ __launch_bounds__(kBlock) void k_push_boris(float* __restrict__ x, float* __restrict__ y,
                                                        float* __restrict__ vx, float* __restrict__ vy,
                                                        float* __restrict__ vz, int64_t n,
                                                        const float* Ex, const float* Ey,
                                                        const float* Bx, const float* By,
                                                        const float* Bz, int nx, int ny, Divisor dx,
                                                        Divisor dy, double h, float dt32, int n_rep) {
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float px = x[i], py = y[i], pvx = vx[i], pvy = vy[i], pvz = vz[i];
  EField<false> E{Ex, Ey};
  for (int r = 0; r < n_rep; ++r)
    boris_substep<EField<false>, HAS_B, FAST>(px, py, pvx, pvy, pvz, E, Bx, By, Bz, nx, ny, dx, dy, h, dt32);
  x[i] = px;
  y[i] = py;
  vx[i] = pvx;
  vy[i] = pvy;
  vz[i] = pvz;
}

// --------------------------------------------------------------- deposit
__global__ __launch_bounds__(kBlock) void k_deposit(const float* __restrict__ x,
                                                     const float* __restrict__ y, int64_t n,
                                                     const int32_t* n_dev, float dx32, float dy32,
                                                     int nx, int ny, uint32_t* counts) {
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int64_t nn = n_dev ? (int64_t)*n_dev : n;
  if (nn > n) nn = n;
  int key = -1;
  if (i < nn) {
    int ix = round_i32_np(x[i] / dx32, 1, nx - 2);
    int iy = round_i32_np(y[i] / dy32, 1, ny - 2);
    key = iy * nx + ix;
  }
  wave_hist_add(counts, key);
}

// ------------------------------------------------------ bilinear deposit
// Bilinear (cloud-in-cell) deposit of the model manual
// (Python/docs/PY-BEMCS_Manual.tex:215-230).  Index and fractions are those
// of the field gather: g = x / f32(dx), i = clip(floor(g), 0, nx-2),
// h = f64(g) - i, so the deposit is the adjoint of the gather.  Weights are
// rounded to 64-bit fixed point (2^-32 units, two's complement for the small
// negative weights of a node one rounding step past the last cell) and
// summed with integer atomics, so per-node totals are exact and independent
// of the order of the adds.
//
// A work-group takes kCicChunk consecutive particles.  When the nodes they
// touch fit a window of kCicWin nodes (cell-sorted particles: a few rows),
// it accumulates in LDS and adds each touched node to HBM once; otherwise
// it adds to HBM directly.  Within a wavefront, runs of lanes on the same
// cell are first summed by a segmented shuffle reduction, so a run costs one
// atomic per corner.
constexpr int kCicPer = 8;
constexpr int kCicChunk = kBlock * kCicPer;
constexpr int kCicWin = 6144;  // uint64 nodes: 48 KiB of LDS

__device__ __forceinline__ int cic_weights(float px, float py, float dx32, float dy32, int nx, int ny,
                                           long long w[4]) {
  const float gx = px / dx32, gy = py / dy32;
  // non-finite or far outside the mesh: no deposit
  if (!(gx >= -1.0f && gx <= (float)nx && gy >= -1.0f && gy <= (float)ny)) return -1;
  const int ix = floor_i32_np(gx, nx - 2), iy = floor_i32_np(gy, ny - 2);
  const double hx = (double)gx - (double)ix, hy = (double)gy - (double)iy;
  const double ax = 1.0 - hx, ay = 1.0 - hy;
  w[0] = llrint((ax * ay) * kCicScale);
  w[1] = llrint((hx * ay) * kCicScale);
  w[2] = llrint((ax * hy) * kCicScale);
  w[3] = llrint((hx * hy) * kCicScale);
  return iy * nx + ix;
}

// Adds the four corner weights of each lane's cell `key` (< 0: none) to
// dst[key], dst[key+1], dst[key+nx], dst[key+nx+1], one atomic per run of
// equal keys in consecutive lanes.
__device__ __forceinline__ void wave_cic_add(unsigned long long* dst, int key, long long w[4], int nx) {
  const int lane = lane_id();
  const int prev = __shfl_up(key, 1);
  const bool head = lane == 0 || key != prev;
  const uint64_t heads = __ballot(head);
  if (heads != ~0ull) {
    // last lane of this lane's run
    const uint64_t later = lane == 63 ? 0ull : (heads >> (lane + 1));
    const int end = later ? lane + __ffsll((unsigned long long)later) - 1 : 63;
    for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const long long o = __shfl_down(w[c], off);
        if (lane + off <= end) w[c] += o;
      }
    }
  }
  if (head && key >= 0) {
    atomicAdd(dst + key, (unsigned long long)w[0]);
    atomicAdd(dst + key + 1, (unsigned long long)w[1]);
    atomicAdd(dst + key + nx, (unsigned long long)w[2]);
    atomicAdd(dst + key + nx + 1, (unsigned long long)w[3]);
  }
}

__global__ __launch_bounds__(kBlock) void k_deposit_cic(const float* __restrict__ x,
                                                         const float* __restrict__ y, int64_t n,
                                                         const int32_t* n_dev, float dx32, float dy32,
                                                         int nx, int ny, unsigned long long* W) {
  __shared__ unsigned long long win[kCicWin];
  __shared__ int red[2][kBlock / 64];
  int64_t nn = n_dev ? (int64_t)*n_dev : n;
  if (nn > n) nn = n;
  const int64_t start = (int64_t)blockIdx.x * kCicChunk;
  if (start >= nn) return;  // whole work-group
  int key[kCicPer];
  long long w[kCicPer][4];
  int lo = INT_MAX, hi = -1;
#pragma unroll
  for (int k = 0; k < kCicPer; ++k) {
    const int64_t i = start + (int64_t)k * kBlock + threadIdx.x;
    w[k][0] = w[k][1] = w[k][2] = w[k][3] = 0;
    key[k] = i < nn ? cic_weights(x[i], y[i], dx32, dy32, nx, ny, w[k]) : -1;
    if (key[k] >= 0) {
      lo = min(lo, key[k]);
      hi = max(hi, key[k]);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, __shfl_xor(lo, off));
    hi = max(hi, __shfl_xor(hi, off));
  }
  const int wid = threadIdx.x >> 6;
  if (lane_id() == 0) {
    red[0][wid] = lo;
    red[1][wid] = hi;
  }
  __syncthreads();
  lo = red[0][0];
  hi = red[1][0];
#pragma unroll
  for (int v = 1; v < kBlock / 64; ++v) {
    lo = min(lo, red[0][v]);
    hi = max(hi, red[1][v]);
  }
  if (hi < 0) return;  // no particle of the work-group deposits
  const int span = hi - lo + nx + 2;
  if (span <= kCicWin) {
    for (int e = threadIdx.x; e < span; e += kBlock) win[e] = 0ull;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kCicPer; ++k) wave_cic_add(win, key[k] >= 0 ? key[k] - lo : -1, w[k], nx);
    __syncthreads();
    for (int e = threadIdx.x; e < span; e += kBlock) {
      const unsigned long long v = win[e];
      if (v) atomicAdd(W + lo + e, v);
    }
  } else {
#pragma unroll
    for (int k = 0; k < kCicPer; ++k) wave_cic_add(W, key[k], w[k], nx);
  }
}

// rho[c] = f32(f64(rho[c]) + cd * W[c] 2^-32) for fixed-point weight sums W
__global__ __launch_bounds__(kBlock) void k_weights_to_rho(const unsigned long long* __restrict__ W,
                                                            int64_t n, double cd, float* rho) {
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double s = (double)(long long)W[i] * kCicUnit;
  rho[i] = (float)((double)rho[i] + cd * s);
}

// rho[c] = f32(f64(rho[c]) + cd), counts[c] times (np.add.at semantics)
__global__ __launch_bounds__(kBlock) void k_counts_to_rho(const uint32_t* __restrict__ cnt,
                                                           int64_t n, double cd, float* rho) {
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t c = cnt[i];
  if (c == 0) return;
  float r = rho[i];
  for (uint32_t k = 0; k < c; ++k) r = (float)((double)r + cd);
  rho[i] = r;
}

// -------------------------------------------------------------- substeps
// ds = |v| dt + 0.5 |qm E| dt^2 with the reference's mixed float32/float64
// evaluation (:940-962); max over particles folded into *ds_bits (ds >= 0,
// so the float64 bit pattern orders like the value).
// work-groups of the substep-count kernel: enough to fill the chip, few
// enough that their maxima meet in a few thousand atomics
inline unsigned substep_grid(int64_t n) {
  const unsigned g = grid_for(n, kBlock);
  return g < 2048u ? g : 2048u;
}

template <bool B>
__global__ __launch_bounds__(kBlock) void k_substep_ds(bemcs_substeps_case_t one,
                                                        const bemcs_substeps_case_t* __restrict__ many,
                                                        int64_t max_bound) {
  __shared__ unsigned long long smax[kBlock / 64];
  const bemcs_substeps_case_t a = B ? many[blockIdx.z] : one;
  const float* __restrict__ x = a.x;
  const float* __restrict__ y = a.y;
  const float* __restrict__ vx = a.vx;
  const float* __restrict__ vy = a.vy;
  const float* __restrict__ vz = a.vz;
  const int32_t* n_dev = a.n_dev;
  const float* __restrict__ Exy = a.Exy;
  const int nx = a.nx, ny = a.ny;
  const float dx32 = (float)a.dx, dy32 = (float)a.dy, dt32 = (float)a.dt;
  const double qm = a.q_m, dt2 = a.dt2;
  unsigned long long* ds_bits = (unsigned long long*)a.ds_max_bits;
  int64_t n = *n_dev;
  if (n > max_bound) n = max_bound;
  // a capped grid strides over the particles, so the block maxima meet in a
  // few thousand atomics on the one result word instead of one per 256
  // particles
  double ds = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    float px = x[i], py = y[i];
    float gx = px / dx32, gy = py / dy32;
    int i0 = floor_i64_np(gx, nx - 2);
    int j0 = floor_i64_np(gy, ny - 2);
    double fx = (double)gx - (double)i0;
    double fy = (double)gy - (double)j0;
    const float2* E2 = reinterpret_cast<const float2*>(Exy);
    float2 e00 = E2[j0 * nx + i0], e10 = E2[j0 * nx + i0 + 1];
    float2 e01 = E2[(j0 + 1) * nx + i0], e11 = E2[(j0 + 1) * nx + i0 + 1];
    double gxm = 1.0 - fx, gym = 1.0 - fy;
    double ex = (double)e00.x * gxm * gym + (double)e10.x * fx * gym + (double)e01.x * gxm * fy +
                (double)e11.x * fx * fy;
    double ey = (double)e00.y * gxm * gym + (double)e10.y * fx * gym + (double)e01.y * gxm * fy +
                (double)e11.y * fx * fy;
    double ax = qm * ex, ay = qm * ey;
    double a_mag = sqrt(ax * ax + ay * ay);
    float a = vx[i], b = vy[i], c = vz[i];
    float v_mag = sqrtf(a * a + b * b + c * c);
    double d = (double)(v_mag * dt32) + 0.5 * a_mag * dt2;
    if (d >= 0.0 && d > ds) ds = d;  // NaN particles do not set the substep count
  }
  unsigned long long bits = (unsigned long long)__double_as_longlong(ds);
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long other = __shfl_xor(bits, o);
    bits = other > bits ? other : bits;
  }
  if (lane_id() == 0) smax[threadIdx.x / 64] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = smax[0];
    for (int w = 1; w < kBlock / 64; ++w) m = smax[w] > m ? smax[w] : m;
    if (m) atomicMax(ds_bits, m);
  }
}

template <bool B>
__global__ void k_substep_finalize(bemcs_substeps_case_t one,
                                   const bemcs_substeps_case_t* __restrict__ many) {
  const bemcs_substeps_case_t a = B ? many[blockIdx.z] : one;
  const unsigned long long* ds_bits = (const unsigned long long*)a.ds_max_bits;
  const int32_t* n_dev = a.n_dev;
  const double ds_lim = a.frac * (a.dx < a.dy ? a.dx : a.dy) * 1e-3;
  int32_t* n_sub = a.n_sub_out;
  double ds_max = __longlong_as_double((long long)*ds_bits);
  int ns = 1;
  if (*n_dev > 0) {
    double lim = ds_lim > 1e-30 ? ds_lim : 1e-30;
    double c = ceil(ds_max / lim);
    ns = c > 1.0 ? (int)c : 1;
  }
  *n_sub = ns;
}

// User cross-section table (lookup_user_cs, Python/physics_engine.py:475-484):
// 10 ** spline(clip(log10(max(E, 1e-30)), lo, hi)), deterministic log/exp.
__device__ __forceinline__ double user_table(const double* t, const double* c, int n, double lo, double hi,
                                             double E) {
  E = nmax(E, 1e-30);
  double le = det_log(E) * kInvLn10;
  le = le < lo ? lo : (le > hi ? hi : le);
  return det_exp(bspline3(t, c, n, le) * kLn10);
}

// Substeps [s0, s1) of K particles of one thread (no magnetic field, the
// interleaved float32 field).  A substep takes a short path when, for every
// particle of the wavefront, both dividends are at least 2^-100 in magnitude
// and the cell coordinates g = x / dx lie in [-2 h, 2 h], h the largest cell
// index (nx - 2 or ny - 2).  The dividends then lie in [2^-100, 2^100], where
// the certified division is exact: fastdiv_ok accepts divisors of at most
// 1e18 < 2^60, so a dividend above 2^100 would give |g| > 2^39 or NaN from the
// division formula, far above 2 h.  The floor index is the clip of floor(g)
// to [0, h] (no int32 overflow or NaN to map), and g - i is exact in float32:
// the fractional part when i = floor(g), g itself when i = 0 (g below 1,
// negative g included), and by Sterbenz's lemma when i = h <= g <= 2 h, so the
// fraction needs one conversion to float64 instead of two and a float64
// subtraction.  Otherwise the wavefront takes boris_substep.  Each substep
// computes every particle's cell and issues its field loads before the
// float64 arithmetic.  The results equal boris_substep's bit for bit.  FIDX
// forms the node offset in float32 (meshes of at most 2^24 nodes).
constexpr int kElecPerThread = 1;
template <bool FAST, int K, bool FIDX>
__device__ __forceinline__ void push_range_impl(float (&x)[K], float (&y)[K], float (&vx)[K],
                                                float (&vy)[K], const EField<true>& E, int nx, int ny,
                                                Divisor dx, Divisor dy, double h, float dt32, int s0,
                                                int s1) {
  const float hx = (float)(nx - 2), hy = (float)(ny - 2);
  const float gx_max = 2.0f * hx, gy_max = 2.0f * hy;
  const float nxf = (float)nx;
  const char* row1 = reinterpret_cast<const char*>(E.ex) + (size_t)nx * 8;
  for (int s = s0; s < s1; ++s) {
    if (FAST) {
      float gx[K], gy[K];
      uint64_t ok = __builtin_amdgcn_read_exec();
#pragma unroll
      for (int k = 0; k < K; ++k) {
        float q = x[k] * dx.r;
        gx[k] = fmaf(fmaf(-q, dx.b, x[k]), dx.r, q);
        q = y[k] * dy.r;
        gy[k] = fmaf(fmaf(-q, dy.b, y[k]), dy.r, q);
        // lane masks of the conditions, combined in scalar registers
        ok &= __builtin_amdgcn_ballot_w64(fabsf(x[k]) >= 0x1p-100f) &
              __builtin_amdgcn_ballot_w64(fabsf(gx[k]) <= gx_max) &
              __builtin_amdgcn_ballot_w64(fabsf(y[k]) >= 0x1p-100f) &
              __builtin_amdgcn_ballot_w64(fabsf(gy[k]) <= gy_max);
      }
      if (ok == __builtin_amdgcn_read_exec()) {
        double fx[K], fy[K];
        float e[K][8];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const float cx = __builtin_amdgcn_fmed3f(floorf(gx[k]), 0.0f, hx);
          const float cy = __builtin_amdgcn_fmed3f(floorf(gy[k]), 0.0f, hy);
          fx[k] = (double)(gx[k] - cx);
          fy[k] = (double)(gy[k] - cy);
          if (FIDX) {
            E.corners_f(cx, cy, nxf, row1, e[k]);
          } else {
            E.corners((int)cx, (int)cy, nx, e[k]);
          }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) kick_drift(fx[k], fy[k], e[k], x[k], y[k], vx[k], vy[k], h, dt32);
        continue;
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      float vz_unused = 0.0f;
      boris_substep<EField<true>, false, FAST>(x[k], y[k], vx[k], vy[k], vz_unused, E, nullptr, nullptr,
                                               nullptr, nx, ny, dx, dy, h, dt32);
    }
  }
}

// The float32 node offset where the mesh allows it (a uniform branch).
template <bool FAST, int K>
__device__ __forceinline__ void push_range_k(float (&x)[K], float (&y)[K], float (&vx)[K], float (&vy)[K],
                                             const EField<true>& E, int nx, int ny, Divisor dx, Divisor dy,
                                             double h, float dt32, int s0, int s1) {
  if (FAST && (int64_t)nx * ny <= (int64_t{1} << 24))
    push_range_impl<FAST, K, true>(x, y, vx, vy, E, nx, ny, dx, dy, h, dt32, s0, s1);
  else
    push_range_impl<FAST, K, false>(x, y, vx, vy, E, nx, ny, dx, dy, h, dt32, s0, s1);
}

// ---------------------------------------------------------- fused ion step
template <bool FAST, bool B>
__global__ __launch_bounds__(kBlock) void k_ion_step(bemcs_ion_case_t one,
                                                      const bemcs_ion_case_t* __restrict__ many) {
  const bemcs_ion_case_t a = B ? many[blockIdx.z] : one;
  const bemcs_species_t& in = a.in;
  const bemcs_species_t& out = a.out;
  const bemcs_species_t& eout = a.e_out;
  const bemcs_ion_step_t& p = a.p;
  const bool PERIODIC = p.periodic_y != 0;
  const Divisor ddx{p.dx32, 1.0f / p.dx32}, ddy{p.dy32, 1.0f / p.dy32};
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t cnt = *in.count;
  const int64_t n = cnt < in.cap ? cnt : in.cap;
  if (i == 0 && cnt > (int64_t)gridDim.x * kBlock) atomicOr(&p.diag->overflow, 16u);
  const bool active = i < n;
  float x = 0.f, y = 0.f, vx = 0.f, vy = 0.f, vz = 0.f;
  uint32_t tag = 0;
  if (active) {
    x = in.x[i];
    y = in.y[i];
    vx = in.vx[i];
    vy = in.vy[i];
    vz = in.vz[i];
    tag = in.tag[i];
  }
  const float x0 = x, y0 = y;
  const int nx = p.nx, ny = p.ny;
  const uint8_t* __restrict__ bound = p.is_bound;

  // ---- C. push (n_sub substeps, dt/n_sub)
  {
    const int nsub = *p.n_sub;
    const double dts = p.dt / (double)nsub;
    const double h = 0.5 * p.q_m * dts;
    const float dts32 = (float)dts;
    EField<true> E{p.Exy, p.Exy + 1};
    if (active) {
      float xs[1] = {x}, ys[1] = {y}, vxs[1] = {vx}, vys[1] = {vy};
      push_range_k<FAST, 1>(xs, ys, vxs, vys, E, nx, ny, ddx, ddy, h, dts32, 0, nsub);
      x = xs[0];
      y = ys[0];
      vx = vxs[0];
      vy = vys[0];
    }
  }
  if (PERIODIC && active) y = np_modf(y, p.Ly32);

  // ---- D. hits, OOB, diagnostics
  // The impact node is the first electrode node along the segment, else the
  // final node (:1329-1335); the segment is walked for a final-node hit only
  // when the impact node is needed (heat or damage).
  const bool SURF = p.heat_acc != nullptr || p.damage_acc != nullptr;
  bool hit = false, oob = false;
  int impact = 0;
  if (active) {
    int ix = round_i64_np(div_rn<FAST>(x, ddx), 0, nx - 1);
    int iy = round_i64_np(div_rn<FAST>(y, ddy), 0, ny - 1);
    const bool hit_final = bound[iy * nx + ix] != 0;
    impact = iy * nx + ix;
    if (!hit_final || SURF) {
      // the eight sample nodes' electrode flags are loaded together (one
      // memory round trip instead of eight dependent ones), then the first
      // electrode sample is taken, as the reference's walk does
      int cell[8];
      uint8_t flag[8];
#pragma unroll
      for (int k = 1; k <= 8; ++k) {
        float f = (float)k / 8.0f;
        float xi = x0 + (x - x0) * f;
        float yi = y0 + (y - y0) * f;
        int cx = round_i64_np(div_rn<FAST>(xi, ddx), 0, nx - 1);
        int cy = round_i64_np(div_rn<FAST>(yi, ddy), 0, ny - 1);
        cell[k - 1] = cy * nx + cx;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) flag[k] = bound[cell[k]];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (!hit && flag[k]) {
          hit = true;
          impact = cell[k];
        }
      }
    }
    hit = hit || hit_final;
    oob = (x < 0.0f) || (x > p.x_plume32) || isnan(x);
    if (!PERIODIC) oob = oob || (y < 0.0f) || (y > p.Ly32);
  }
  const bool entered = active && (x0 < p.x_entry32) && (x >= p.x_entry32) && (vx > 0.0f);
  wave_count(&p.diag->entered, entered);
  wave_count(&p.diag->lost_to_grid, active && hit && !oob);

  const bool surf = active && hit && x > 0.25f;
  wave_count(&p.diag->surface_hits, surf);
  // impact heating (:1359-1388) and sputter damage (:1419-1438), float32 as
  // the reference evaluates them, summed per node in 64-bit fixed point
  if (surf && SURF) {
    float v2 = vx * vx + vy * vy;
    v2 = v2 + vz * vz;
    if (p.heat_acc) {
      float ej = (p.half_m32 * v2) * p.w32;
      float dT = (ej / p.ccell32) * p.accel32;
      fixed_add(p.heat_acc + impact, dT, 1048576.0, &p.diag->overflow);
    }
    if (p.damage_acc) {
      float E = (p.half_m32 * v2) / p.q32;
      if (E > p.eth32) {
        float dmg = (p.ycoef32 * (E - p.eth32)) * p.w32;
        fixed_add(p.damage_acc + impact, dmg, 16777216.0, &p.diag->overflow);
      }
    }
  }

  // SEE: electrons spawned by grid hits (:1390-1417), purge-checked (:1541-1551)
  {
    bool spawn = false;
    float ex = 0.f, ey = 0.f, evx = 0.f, evy = 0.f, evz = 0.f;
    if (surf) {
      double a = vx, b = vy, c = vz;
      double v2 = a * a + b * b + c * c;
      double E_eV = (0.5 * p.m_ion * v2) / p.q;
      double gamma = p.see_t ? user_table(p.see_t, p.see_c, p.see_n, p.see_lo, p.see_hi, E_eV)
                             : 0.05 + 1e-4 * E_eV;
      gamma = gamma < 0.0 ? 0.0 : (gamma > 1.0 ? 1.0 : gamma);
      const uint32_t pid = tag & 0x7FFFFFFFu;
      u01x2 u = uniforms(p.seed, P_SEE, pid, p.step, 0u);
      if (u.a < gamma) {
        spawn = true;
        ex = x - ((vx * p.dt32) * 1000.0f) / 1.5f;
        ey = y - ((vy * p.dt32) * 1000.0f) / 1.5f;
        u01x2 z01 = normals(p.seed, P_SEE, pid, p.step, 1u);
        u01x2 z23 = normals(p.seed, P_SEE, pid, p.step, 2u);
        evx = (float)(z01.a * p.see_vth);
        evy = (float)(z01.b * p.see_vth);
        evz = (float)(z23.a * p.see_vth);
      }
    }
    wave_count(&p.diag->see_spawned, spawn);
    if (p.see_defer) {
      // staged unpurged; bemcs_see_commit purges them after this step's erosion
      const bemcs_species_t& st = p.see_stage;
      int64_t slot = wave_append(st.count, spawn);
      if (spawn) {
        if (slot < st.cap) {
          st.x[slot] = ex;
          st.y[slot] = ey;
          st.vx[slot] = evx;
          st.vy[slot] = evy;
          st.vz[slot] = evz;
        } else {
          atomicOr(&p.diag->overflow, 1u);
        }
      }
      spawn = false;
    }
    bool keep_e = false;
    int ekey = -1;
    if (spawn) {
      int eix = round_i64_np(div_rn<FAST>(ex, ddx), 0, nx - 1);
      int eiy = round_i64_np(div_rn<FAST>(ey, ddy), 0, ny - 1);
      bool ehit = bound[eiy * nx + eix] != 0;
      bool eout = (ex < 0.0f) || (ex > p.Lx32) || (ey < 0.0f) || (ey > p.Ly32) || isnan(ex);
      keep_e = !(ehit || eout);
      if (keep_e) {
        int dix = round_i32_np(div_rn<FAST>(ex, ddx), 1, nx - 2);
        int diy = round_i32_np(div_rn<FAST>(ey, ddy), 1, ny - 2);
        ekey = diy * nx + dix;
      }
    }
    int64_t slot = wave_append(eout.count, keep_e);
    if (keep_e) {
      if (slot < eout.cap) {
        eout.x[slot] = ex;
        eout.y[slot] = ey;
        eout.vx[slot] = evx;
        eout.vy[slot] = evy;
        eout.vz[slot] = evz;
      } else {
        atomicOr(&p.diag->overflow, 1u);
        ekey = -1;
      }
    }
    wave_count(&p.diag->see_kept, keep_e);
    wave_hist_add(p.e_cnt_next, ekey);
  }

  // exit statistics and theta95 samples (:1451-1488)
  const bool crossed = active && (x0 <= p.x_exit32) && (x > p.x_exit32) && (vx > 0.0f);
  const bool exited = active && (x > p.x_exit32) && (vx > 0.0f);
  wave_count(&p.diag->exited, exited);
  {
    int64_t slot = wave_append(&p.diag->crossed, crossed);
    double s_vx = 0, s_vx2 = 0, s_v = 0, s_v2 = 0, s_E = 0, s_vp = 0, s_vc = 0;
    bool prim_c = false;
    if (crossed) {
      // correctly rounded float32 angle (float64 atan2, then rounded); the
      // scaling to degrees is float32 as in the reference
      p.crossed_angle[slot] = ((float)fabs(atan2((double)vy, (double)vx)) * 180.0f) /
                              3.14159265358979323846f;
      float v = sqrtf(vx * vx + vy * vy + vz * vz);
      float E = ((float)(0.5 * p.m_ion) * (v * v)) / (float)p.q;
      s_vx = vx;
      s_vx2 = (double)vx * vx;
      s_v = v;
      s_v2 = (double)v * v;
      s_E = E;
      prim_c = (tag >> 31) == 0;
      if (prim_c) s_vp = v; else s_vc = v;
    }
    wave_count(&p.diag->crossed_primary, prim_c);
    if (__ballot(crossed)) {
      s_vx = wave_sum(s_vx);
      s_vx2 = wave_sum(s_vx2);
      s_v = wave_sum(s_v);
      s_v2 = wave_sum(s_v2);
      s_E = wave_sum(s_E);
      s_vp = wave_sum(s_vp);
      s_vc = wave_sum(s_vc);
      if (lane_id() == 0) {
        atomicAdd(&p.diag->sum_vx, s_vx);
        atomicAdd(&p.diag->sum_vx2, s_vx2);
        atomicAdd(&p.diag->sum_v, s_v);
        atomicAdd(&p.diag->sum_v2, s_v2);
        atomicAdd(&p.diag->sum_E, s_E);
        atomicAdd(&p.diag->sum_v_primary, s_vp);
        atomicAdd(&p.diag->sum_v_cex, s_vc);
      }
    }
  }

  // removal (:1499-1529); the keep test is implied by the OOB test but kept
  bool keep = active && !(hit || oob);
  keep = keep && ((x <= p.x_plume32) || (vx < 0.0f));

  // ---- H. CEX collisions (:1682-1739), float64 deterministic probability
  bool coll = false;
  if (keep && p.mcc_enabled && (tag >> 31) == 0 && x >= 1.0f && x <= p.Lx32) {
    double x64 = x, y64 = y;
    double nloc;
    if (p.neutral_model == 1) {
      nloc = p.mcc_n0;
    } else {
      double z = nmax((x64 - 1.0) * 1e-3, 0.0);
      double r = y64 * 1e-3;
      double zr = z + p.mcc_rT;
      double R = sqrt(r * r + zr * zr);
      double cos_t = zr / R;
      double qq = p.mcc_rT / nmax(R, 1e-12);
      nloc = (p.mcc_n0a * (1.0 - 1.0 / sqrt(1.0 + qq * qq))) * cos_t;
      nloc = nmax(nloc, 0.0);
    }
    double a = vx, b = vy, c = vz;
    double g = nmax(sqrt(a * a + b * b + c * c), 1.0);
    double sigma;
    if (p.cx_t) {
      const double E = (0.5 * p.m_ion * (g * g)) / p.q;
      sigma = user_table(p.cx_t, p.cx_c, p.cx_n, p.cx_lo, p.cx_hi, E);
    } else if (p.cx_model == 1) {
      double E = (0.5 * p.m_ion * (g * g)) / p.q;
      sigma = nmax((87.3 - 13.6 * (det_log(E) * kInvLn10)) * 1e-20, 0.0);
    } else {
      double t = -0.8821 * det_log(g) + 15.1262;
      sigma = (t * t) * 1e-20;
    }
    double P = 1.0 - det_exp(-(((nloc * sigma) * g) * p.dt));
    const uint32_t pid = tag & 0x7FFFFFFFu;
    u01x2 u = uniforms(p.seed, P_MCC, pid, p.step, 0u);
    if (u.a < P) {
      coll = true;
      double R[10];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        u01x2 w = uniforms(p.seed, P_MCC, pid, p.step, (uint32_t)(k + 1));
        R[2 * k] = w.a;
        R[2 * k + 1] = w.b;
      }
      vx = (float)(p.mcc_vth * (2.0 * (R[0] + R[1] + R[2] - 1.5)));
      vy = (float)(p.mcc_vth * (2.0 * (R[3] + R[4] + R[5] - 1.5)));
      vz = (float)(p.mcc_vth * (2.0 * (R[6] + R[7] + R[8] - 1.5)));
      tag |= 0x80000000u;
    }
  }
  wave_count(&p.diag->mcc_events, coll);

  // ---- write survivors, NGP counts for the next deposit, PPC counts
  int64_t slot = block_append(out.count, keep);
  int dix = -1, diy = -1, pix = -1, piy = -1;
  if (keep) {
    if (slot < out.cap) {
      out.x[slot] = x;
      out.y[slot] = y;
      out.vx[slot] = vx;
      out.vy[slot] = vy;
      out.vz[slot] = vz;
      out.tag[slot] = tag;
      const float gx = div_rn<FAST>(x, ddx), gy = div_rn<FAST>(y, ddy);
      dix = round_i32_np(gx, 1, nx - 2);
      diy = round_i32_np(gy, 1, ny - 2);
      pix = trunc_i64_np(gx, nx - 1);
      piy = trunc_i64_np(gy, ny - 1);
    } else {
      atomicOr(&p.diag->overflow, 2u);
    }
  }
  __shared__ LdsHist hcnt, hppc;
  const CountWindow win = count_window(p.count_window);
  const int dixk[1] = {dix}, diyk[1] = {diy}, pixk[1] = {pix}, piyk[1] = {piy};
  block_hist_add<1>(hcnt, p.cnt_next, nx, dixk, diyk, win);
  block_hist_add<1>(hppc, p.ppc, nx, pixk, piyk, win);
}

// ------------------------------------------------------ fused electron step
// First substep of chunk j of k over n substeps: floor(n j / k).
__device__ __forceinline__ int chunk_begin(int n, int j, int k) {
  if (k <= 1 || j <= 0) return 0;
  if (j >= k) return n;
  return (int)(((int64_t)n * j) / k);
}

// Substeps of one chunk, in place (no wrap, purge or removal).
// Substeps [s0, s1) of one particle in registers.
template <bool FAST>
__global__ __launch_bounds__(kBlock) void k_electron_push(bemcs_species_t sp, bemcs_electron_step_t p,
                                                          int chunk, int64_t n_bound) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int64_t n = *sp.count;
  if (n > sp.cap) n = sp.cap;
  if (n > n_bound) n = n_bound;
  if (i >= n) return;
  const int nsub = *p.n_sub;
  const int s0 = chunk_begin(nsub, chunk, p.sub_chunks);
  const int s1 = chunk_begin(nsub, chunk + 1, p.sub_chunks);
  if (s0 >= s1) return;
  const Divisor ddx{p.dx32, 1.0f / p.dx32}, ddy{p.dy32, 1.0f / p.dy32};
  const double dts = p.dt / (double)nsub;
  const double h = 0.5 * p.q_m * dts;
  const float dts32 = (float)dts;
  // no magnetic field: vz is unchanged
  float x[1] = {sp.x[i]}, y[1] = {sp.y[i]}, vx[1] = {sp.vx[i]}, vy[1] = {sp.vy[i]};
  const EField<true> E{p.Exy, p.Exy + 1};
  push_range_k<FAST, 1>(x, y, vx, vy, E, p.nx, p.ny, ddx, ddy, h, dts32, s0, s1);
  sp.x[i] = x[0];
  sp.y[i] = y[0];
  sp.vx[i] = vx[0];
  sp.vy[i] = vy[0];
}

template <bool FAST, bool B>
__global__ __launch_bounds__(kBlock) void k_electron_step(bemcs_electron_case_t one,
                                                           const bemcs_electron_case_t* __restrict__ many) {
  // kElecPerThread electrons per thread: electron k of the thread is
  // blockIdx.x * kElecPerThread * kBlock + k * kBlock + threadIdx.x
  constexpr int K = kElecPerThread;
  const bemcs_electron_case_t a = B ? many[blockIdx.z] : one;
  const bemcs_species_t& in = a.in;
  const bemcs_species_t& out = a.out;
  const bemcs_electron_step_t& p = a.p;
  const bool PERIODIC = p.periodic_y != 0;
  const Divisor ddx{p.dx32, 1.0f / p.dx32}, ddy{p.dy32, 1.0f / p.dy32};
  const int64_t i_first = (int64_t)blockIdx.x * (K * kBlock) + threadIdx.x;
  const int64_t cnt = *in.count;
  const int64_t n = cnt < in.cap ? cnt : in.cap;
  if (i_first == 0 && cnt > (int64_t)gridDim.x * (K * kBlock)) atomicOr(p.overflow, 16u);
  float x[K], y[K], vx[K], vy[K], vz[K];
  bool active[K];
  bool any = false;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t i = i_first + (int64_t)k * kBlock;
    active[k] = i < n;
    any = any || active[k];
    // an inactive slot pushes zeros (every field index stays clamped to the
    // mesh) and is dropped afterwards
    x[k] = active[k] ? in.x[i] : 0.0f;
    y[k] = active[k] ? in.y[i] : 0.0f;
    vx[k] = active[k] ? in.vx[i] : 0.0f;
    vy[k] = active[k] ? in.vy[i] : 0.0f;
    vz[k] = active[k] ? in.vz[i] : 0.0f;
  }
  const int nx = p.nx, ny = p.ny;
  {
    const int nsub = *p.n_sub;
    const double dts = p.dt / (double)nsub;
    const double h = 0.5 * p.q_m * dts;
    const float dts32 = (float)dts;
    const int s0 = chunk_begin(nsub, p.sub_first, p.sub_chunks);
    const EField<true> E{p.Exy, p.Exy + 1};
    if (any) push_range_k<FAST, K>(x, y, vx, vy, E, nx, ny, ddx, ddy, h, dts32, s0, nsub);
  }
  int dix[K], diy[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (PERIODIC && active[k]) y[k] = np_modf(y[k], p.Ly32);
    bool keep = false;
    if (active[k]) {
      int ix = round_i64_np(div_rn<FAST>(x[k], ddx), 0, nx - 1);
      int iy = round_i64_np(div_rn<FAST>(y[k], ddy), 0, ny - 1);
      bool hit = p.is_bound[iy * nx + ix] != 0;
      bool outb = (x[k] < 0.0f) || (x[k] > p.Lx32) || (y[k] < 0.0f) || (y[k] > p.Ly32) || isnan(x[k]);
      keep = !(hit || outb);
    }
    int64_t slot = block_append(out.count, keep);
    dix[k] = -1;
    diy[k] = -1;
    if (keep) {
      if (slot < out.cap) {
        out.x[slot] = x[k];
        out.y[slot] = y[k];
        out.vx[slot] = vx[k];
        out.vy[slot] = vy[k];
        out.vz[slot] = vz[k];
        dix[k] = round_i32_np(div_rn<FAST>(x[k], ddx), 1, nx - 2);
        diy[k] = round_i32_np(div_rn<FAST>(y[k], ddy), 1, ny - 2);
      } else {
        atomicOr(p.overflow, 4u);
      }
    }
  }
  __shared__ LdsHist hcnt;
  block_hist_add<K>(hcnt, p.cnt_next, nx, dix, diy, count_window(p.count_window));
}

// --------------------------------------------------------------- injection
// Population hold: particles to add so that the count reaches target, at most
// the launch size n.
constexpr uint32_t kHoldIdBit = 0x40000000u;
__device__ __forceinline__ int64_t hold_need(int64_t target, int64_t count, int64_t n) {
  int64_t need = target - count;
  if (need < 0) need = 0;
  return need < n ? need : n;
}

template <bool B>
__global__ __launch_bounds__(kBlock) void k_inject(bemcs_inject_case_t one,
                                                    const bemcs_inject_case_t* __restrict__ many) {
  const bemcs_inject_case_t a = B ? many[blockIdx.z] : one;
  const bemcs_species_t& sp = a.sp;
  const bemcs_inject_t& p = a.p;
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t base = *sp.count;
  const bool hold = p.target > 0 && (p.kind == 3 || p.kind == 4);
  const int64_t n = hold ? hold_need(p.target, base, a.n) : a.n;
  const uint32_t id0 = hold ? *p.id_next : 0u;
  int key = -1;
  if (i < n && base + i < sp.cap) {
    const uint32_t id = id0 + (uint32_t)i;
    float x, y;
    double vx, vy, vz;
    if (p.kind == 3 || p.kind == 4) {
      const uint32_t purpose = p.kind == 3 ? (hold ? P_HOLD_ION : P_INIT_ION)
                                           : (hold ? P_HOLD_ELEC : P_INIT_ELEC);
      x = 0.f;
      y = 0.f;
      for (uint32_t t = 0; t < 32; ++t) {
        u01x2 u = uniforms(p.seed, purpose, id, t, 0u);
        float cx = (float)(p.Lx * u.a);
        float cy = (float)(p.Ly * u.b);
        int bi = round_i64_np(cx / p.dx32, 0, p.nx - 1);
        int bj = round_i64_np(cy / p.dy32, 0, p.ny - 1);
        x = cx;
        y = cy;
        if (!p.is_bound[bj * p.nx + bi]) break;
      }
      u01x2 z01 = normals(p.seed, purpose, id, 0u, 1u);
      u01x2 z23 = normals(p.seed, purpose, id, 0u, 2u);
      if (p.kind == 3) {
        vx = p.v_drift + z01.a * p.v_spread;
        vy = z01.b * p.v_spread;
        vz = z23.a * p.v_spread;
      } else {
        vx = z01.a * p.v_spread;
        vy = z01.b * p.v_spread;
        vz = z23.a * p.v_spread;
      }
    } else {
      const uint32_t purpose = p.kind == 0 ? P_INJ_ION : (p.kind == 1 ? P_NEUT : P_RF_ELEC);
      u01x2 u = uniforms(p.seed, purpose, id, p.step, 0u);
      int band = (p.two_band && u.b >= 0.5) ? 1 : 0;
      double lo = p.y_lo[band], hi = p.y_hi[band];
      double span = p.band_f32 ? (double)((float)hi - (float)lo) : hi - lo;
      y = (float)(lo + span * u.a);
      x = (float)p.x0;
      u01x2 z01 = normals(p.seed, purpose, id, p.step, 1u);
      u01x2 z23 = normals(p.seed, purpose, id, p.step, 2u);
      if (p.kind == 0) {
        vx = p.v_drift + z01.a * p.v_spread;
      } else if (p.kind == 2) {
        vx = fabs(z01.a) * p.v_spread + p.v_drift;
      } else {
        vx = z01.a * p.v_spread;
      }
      vy = z01.b * p.v_spread;
      vz = z23.a * p.v_spread;
    }
    const int64_t s = base + i;
    sp.x[s] = x;
    sp.y[s] = y;
    sp.vx[s] = (float)vx;
    sp.vy[s] = (float)vy;
    sp.vz[s] = (float)vz;
    if (sp.tag) sp.tag[s] = hold ? (kHoldIdBit | (id & (kHoldIdBit - 1u))) : p.pid_base + id;
    if (p.cnt) {
      int dix = round_i32_np(x / p.dx32, 1, p.nx - 2);
      int diy = round_i32_np(y / p.dy32, 1, p.ny - 2);
      key = diy * p.nx + dix;
    }
  }
  if (p.cnt) wave_hist_add(p.cnt, key);
}

template <bool B>
__global__ void k_add_count(bemcs_inject_case_t one, const bemcs_inject_case_t* __restrict__ many) {
  const bemcs_inject_case_t a = B ? many[blockIdx.z] : one;
  int32_t* count = a.sp.count;
  const int64_t cap = a.sp.cap;
  const bool hold = a.p.target > 0 && (a.p.kind == 3 || a.p.kind == 4);
  const int64_t n = hold ? hold_need(a.p.target, *count, a.n) : a.n;
  int64_t c = (int64_t)*count + n;
  if (c > cap) {
    if (a.p.overflow) atomicOr(a.p.overflow, 8u);
    c = cap;
  }
  *count = (int32_t)c;
  if (hold) *a.p.id_next += (uint32_t)n;
}

// --------------------------------------------------------------- cell sort
// Counting sort by gather cell.  Scratch layout per case: hist [nx*ny],
// block sums [1024 + nx*ny/1024], keys [n_bound].
struct SortView {
  uint32_t *hist, *sums, *keys;
  int64_t ncell;  // bins: tiles of 2^shift_x x 2^shift_y gather cells
  int tiles_x;
};
__host__ __device__ __forceinline__ int sort_tiles_x(const bemcs_sort_case_t& a) {
  return ((a.nx - 2) >> a.shift_x) + 1;
}
__host__ __device__ __forceinline__ int64_t sort_bins(const bemcs_sort_case_t& a) {
  return (int64_t)sort_tiles_x(a) * (((a.ny - 2) >> a.shift_y) + 1);
}
__device__ __forceinline__ SortView sort_view(const bemcs_sort_case_t& a) {
  SortView v;
  v.ncell = sort_bins(a);
  v.tiles_x = sort_tiles_x(a);
  v.hist = a.scratch;
  v.sums = a.scratch + v.ncell;
  v.keys = v.sums + 1024 + ((v.ncell + 1023) / 1024);
  return v;
}

template <bool B>
__global__ __launch_bounds__(kBlock) void k_sort_prepare(bemcs_sort_case_t one,
                                                          const bemcs_sort_case_t* __restrict__ many) {
  const bemcs_sort_case_t a = B ? many[blockIdx.z] : one;
  const SortView v = sort_view(a);
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < v.ncell) v.hist[i] = 0u;
  if (i == 0) {
    const int64_t n = *a.in.count, nb = a.n_bound > 0 ? a.n_bound : 0;
    *a.out.count = (int32_t)(n < nb ? n : nb);
  }
}

template <bool B>
__global__ __launch_bounds__(kBlock) void k_cell_hist(bemcs_sort_case_t one,
                                                       const bemcs_sort_case_t* __restrict__ many) {
  const bemcs_sort_case_t a = B ? many[blockIdx.z] : one;
  const SortView v = sort_view(a);
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int key = -1;
  int64_t n = *a.in.count;
  if (n > a.n_bound) n = a.n_bound;
  if (i < n) {
    int i0 = floor_i32_np(a.in.x[i] / (float)a.dx, a.nx - 2);
    int j0 = floor_i32_np(a.in.y[i] / (float)a.dy, a.ny - 2);
    key = (j0 >> a.shift_y) * v.tiles_x + (i0 >> a.shift_x);
    v.keys[i] = (uint32_t)key;
  }
  wave_hist_add(v.hist, key);
}

// exclusive scan of the histogram in place: per-block scan of 1024 entries,
// scan of the block sums, add
template <bool B>
__global__ __launch_bounds__(1024) void k_scan_blocks(bemcs_sort_case_t one,
                                                       const bemcs_sort_case_t* __restrict__ many) {
  __shared__ uint32_t s[1024];
  const bemcs_sort_case_t a = B ? many[blockIdx.z] : one;
  const SortView sv = sort_view(a);
  const int64_t n = sv.ncell;
  if ((int64_t)blockIdx.x * 1024 >= n) return;
  uint32_t* data = sv.hist;
  int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
  uint32_t v = i < n ? data[i] : 0u;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    uint32_t t = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) data[i] = s[threadIdx.x] - v;
  if (threadIdx.x == 1023) sv.sums[blockIdx.x] = s[1023];
}

template <bool B>
__global__ __launch_bounds__(1024) void k_scan_sums(bemcs_sort_case_t one,
                                                     const bemcs_sort_case_t* __restrict__ many) {
  __shared__ uint32_t s[1024];
  const bemcs_sort_case_t a = B ? many[blockIdx.z] : one;
  const SortView sv = sort_view(a);
  uint32_t* sums = sv.sums;
  const int nb = (int)((sv.ncell + 1023) / 1024);
  uint32_t carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    int i = base + (int)threadIdx.x;
    uint32_t v = i < nb ? sums[i] : 0u;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      uint32_t t = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0u;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) sums[i] = carry + s[threadIdx.x] - v;
    uint32_t tot = s[1023];
    __syncthreads();
    carry += tot;
  }
}

template <bool B>
__global__ __launch_bounds__(1024) void k_scan_add(bemcs_sort_case_t one,
                                                    const bemcs_sort_case_t* __restrict__ many) {
  const bemcs_sort_case_t a = B ? many[blockIdx.z] : one;
  const SortView sv = sort_view(a);
  int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
  if (i < sv.ncell) sv.hist[i] += sv.sums[blockIdx.x];
}

template <bool B>
__global__ __launch_bounds__(kBlock) void k_cell_scatter(bemcs_sort_case_t one,
                                                          const bemcs_sort_case_t* __restrict__ many) {
  const bemcs_sort_case_t a = B ? many[blockIdx.z] : one;
  const SortView sv = sort_view(a);
  const bemcs_species_t& in = a.in;
  const bemcs_species_t& out = a.out;
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int64_t n = *in.count;
  if (n > a.n_bound) n = a.n_bound;
  // runs of equal keys in consecutive lanes (input already close to cell
  // order) reserve their slots with one atomic
  const int key = i < n ? (int)sv.keys[i] : -1;
  const int lane = lane_id();
  const int prev = __shfl_up(key, 1);
  const bool head = (lane == 0) || (key != prev);
  const uint64_t heads = __ballot(head);
  uint32_t base = 0;
  if (head && key >= 0) {
    const uint64_t after = (lane == 63) ? 0ull : (heads >> (lane + 1));
    const int run = after ? __ffsll((unsigned long long)after) : (64 - lane);
    base = atomicAdd(&sv.hist[key], (uint32_t)run);
  }
  const uint64_t upto = (lane == 63) ? heads : (heads & ((2ull << lane) - 1ull));
  const int leader = 63 - __clzll((long long)upto);
  base = __shfl(base, leader);
  if (i >= n) return;
  const uint32_t d = base + (uint32_t)(lane - leader);
  out.x[d] = in.x[i];
  out.y[d] = in.y[i];
  out.vx[d] = in.vx[i];
  out.vy[d] = in.vy[i];
  out.vz[d] = in.vz[i];
  if (in.tag) out.tag[d] = in.tag[i];
}


// ------------------------------------------------- division certification
// Counts the float32 dividends a (all 2^32 bit patterns) for which
// div_rn<true>(a, {b, r}) differs from IEEE a / b; two NaNs count as equal.
__global__ __launch_bounds__(256) void k_check_div(float b, float r,
                                                   unsigned long long* mismatches) {
  const Divisor d{b, r};
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  unsigned long long bad = 0;
  for (uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x; k < (1ull << 32); k += stride) {
    float a = __uint_as_float((uint32_t)k);
    float f = div_rn<true>(a, d);
    float e = div_rn<false>(a, d);
    bool same = __float_as_uint(f) == __float_as_uint(e) || (isnan(f) && isnan(e));
    bad += same ? 0 : 1;
  }
  if (bad) atomicAdd(mismatches, bad);
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }

static std::mutex g_div_mu;
static std::unordered_map<uint32_t, bool> g_div_ok;
static unsigned long long* g_div_count = nullptr;

// Runs k_check_div on stream s and waits; returns BEMCS_OK and the count.
static int check_divisor(float b, float r, hipStream_t s, unsigned long long* out) {
  if (!g_div_count) BEMCS_CHECK_HIP(hipMalloc(&g_div_count, sizeof(unsigned long long)));
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, g_div_count, 0ull);
  hipLaunchKernelGGL(k_check_div, dim3(8192), dim3(256), 0, s, b, r, g_div_count);
  BEMCS_CHECK_LAUNCH();
  BEMCS_CHECK_HIP(hipMemcpyAsync(out, g_div_count, sizeof(*out), hipMemcpyDeviceToHost, s));
  BEMCS_CHECK_HIP(hipStreamSynchronize(s));
  return BEMCS_OK;
}

bool fastdiv_ok(float b, hipStream_t s) {
  const char* env = getenv("BEMCS_FASTDIV");
  if (env && env[0] == '0') return false;
  // divisors of at most 1e18 < 2^60: the short path of the push relies on it
  if (!(fabsf(b) >= 1e-30f && fabsf(b) <= 1e18f)) return false;
  uint32_t key;
  memcpy(&key, &b, sizeof(key));
  std::lock_guard<std::mutex> lock(g_div_mu);
  auto it = g_div_ok.find(key);
  if (it != g_div_ok.end()) return it->second;
  unsigned long long bad = 1;
  bool ok = check_divisor(b, 1.0f / b, s, &bad) == BEMCS_OK && bad == 0;
  g_div_ok[key] = ok;
  return ok;
}

}  // namespace bemcs

using namespace bemcs;

extern "C" {

int bemcs_fastdiv_check(float b, float r, uint64_t* mismatches, bemcs_stream_t stream) {
  BEMCS_REQUIRE(mismatches, "bemcs_fastdiv_check: null argument");
  unsigned long long bad = 0;
  std::lock_guard<std::mutex> lock(g_div_mu);
  int rc = check_divisor(b, r, as_stream(stream), &bad);
  *mismatches = bad;
  return rc;
}

int bemcs_fastdiv_enabled(float b, bemcs_stream_t stream) {
  return fastdiv_ok(b, as_stream(stream)) ? 1 : 0;
}

int bemcs_push_boris(float* x, float* y, float* vx, float* vy, float* vz, int64_t n,
                     const float* Ex, const float* Ey, const float* Bx, const float* By,
                     const float* Bz, int32_t nx, int32_t ny, double dx, double dy, double dt,
                     double q_m, int32_t n_rep, bemcs_stream_t stream) {
  BEMCS_REQUIRE(n >= 0 && nx >= 2 && ny >= 2 && n_rep >= 0, "bemcs_push_boris: bad sizes");
  if (n == 0 || n_rep == 0) return BEMCS_OK;
  BEMCS_REQUIRE(x && y && vx && vy && vz && Ex && Ey, "bemcs_push_boris: null pointer");
  bool has_b = Bx || By || Bz;
  BEMCS_REQUIRE(!has_b || (Bx && By && Bz), "bemcs_push_boris: B needs all three components");
  double h = 0.5 * q_m * dt;
  float dt32 = (float)dt;
  hipStream_t s = as_stream(stream);
  const Divisor ddx = make_divisor((float)dx), ddy = make_divisor((float)dy);
  const bool fast = fastdiv_ok(ddx.b, s) && fastdiv_ok(ddy.b, s);
  auto kern = has_b ? (fast ? k_push_boris<true, true> : k_push_boris<true, false>)
                    : (fast ? k_push_boris<false, true> : k_push_boris<false, false>);
  hipLaunchKernelGGL(kern, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, x, y, vx, vy, vz, n, Ex,
                     Ey, Bx, By, Bz, nx, ny, ddx, ddy, h, dt32, n_rep);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_deposit_ngp_counts(const float* x, const float* y, int64_t n, double dx, double dy,
                             int32_t nx, int32_t ny, uint32_t* counts, bemcs_stream_t stream) {
  BEMCS_REQUIRE(n >= 0 && nx >= 3 && ny >= 3, "bemcs_deposit_ngp_counts: bad sizes");
  if (n == 0) return BEMCS_OK;
  BEMCS_REQUIRE(x && y && counts, "bemcs_deposit_ngp_counts: null pointer");
  hipLaunchKernelGGL(k_deposit, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, as_stream(stream), x, y,
                     n, (const int32_t*)nullptr, (float)dx, (float)dy, nx, ny, counts);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_deposit_cic(const float* x, const float* y, int64_t n, const int32_t* n_dev, double dx,
                      double dy, int32_t nx, int32_t ny, uint64_t* weights, bemcs_stream_t stream) {
  BEMCS_REQUIRE(n >= 0 && nx >= 2 && ny >= 2, "bemcs_deposit_cic: bad sizes");
  BEMCS_REQUIRE((int64_t)nx * ny < INT_MAX - 2 * (int64_t)nx, "bemcs_deposit_cic: mesh too large");
  if (n == 0) return BEMCS_OK;
  BEMCS_REQUIRE(x && y && weights, "bemcs_deposit_cic: null pointer");
  BEMCS_REQUIRE(dx > 0.0 && dy > 0.0, "bemcs_deposit_cic: spacing must be positive");
  hipLaunchKernelGGL(k_deposit_cic, dim3(grid_for(n, kCicChunk)), dim3(kBlock), 0, as_stream(stream), x,
                     y, n, n_dev, (float)dx, (float)dy, nx, ny, (unsigned long long*)weights);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_weights_to_rho(const uint64_t* weights, int64_t n_nodes, double charge_density, float* rho,
                         bemcs_stream_t stream) {
  BEMCS_REQUIRE(n_nodes >= 0 && weights && rho, "bemcs_weights_to_rho: bad arguments");
  if (n_nodes == 0) return BEMCS_OK;
  hipLaunchKernelGGL(k_weights_to_rho, dim3(grid_for(n_nodes, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), (const unsigned long long*)weights, n_nodes, charge_density, rho);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_counts_to_rho(const uint32_t* counts, int64_t n_nodes, double charge_density, float* rho,
                        bemcs_stream_t stream) {
  BEMCS_REQUIRE(n_nodes >= 0 && counts && rho, "bemcs_counts_to_rho: bad arguments");
  if (n_nodes == 0) return BEMCS_OK;
  hipLaunchKernelGGL(k_counts_to_rho, dim3(grid_for(n_nodes, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), counts, n_nodes, charge_density, rho);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_substeps(const float* x, const float* y, const float* vx, const float* vy,
                   const float* vz, const int32_t* n_dev, int64_t n_bound, const float* Exy,
                   int32_t nx, int32_t ny, double dx, double dy, double q_m, double dt,
                   double dt2, double frac, uint64_t* ds_max_bits, int32_t* n_sub_out,
                   bemcs_stream_t stream) {
  BEMCS_REQUIRE(n_dev && Exy && ds_max_bits && n_sub_out && nx >= 2 && ny >= 2,
                "bemcs_substeps: bad arguments");
  hipStream_t s = as_stream(stream);
  const bemcs_substeps_case_t a{x, y, vx, vy, vz, n_dev, Exy, nx, ny, dx, dy, q_m, dt, dt2, frac,
                                ds_max_bits, n_sub_out};
  if (n_bound > 0)
    hipLaunchKernelGGL(k_substep_ds<false>, dim3(substep_grid(n_bound)), dim3(kBlock), 0, s, a,
                       (const bemcs_substeps_case_t*)nullptr, n_bound);
  hipLaunchKernelGGL(k_substep_finalize<false>, dim3(1), dim3(1), 0, s, a,
                     (const bemcs_substeps_case_t*)nullptr);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int64_t bemcs_poisson_count(uint32_t seed, uint32_t purpose, uint32_t step, double lam) {
  return poisson_count(seed, purpose, step, lam);
}

double bemcs_det_exp(double x) { return det_exp(x); }
double bemcs_det_log(double x) { return det_log(x); }

void bemcs_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  u32x4 r = philox4x32_10(u32x4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
  out[3] = r.w;
}

int bemcs_ion_step(const bemcs_species_t* in, const bemcs_species_t* out,
                   const bemcs_species_t* e_out, const bemcs_ion_step_t* p, int64_t n_bound,
                   bemcs_stream_t stream) {
  BEMCS_REQUIRE(in && out && e_out && p, "bemcs_ion_step: null argument");
  BEMCS_REQUIRE(in->tag && out->tag && in->count && out->count && e_out->count,
                "bemcs_ion_step: species buffers incomplete");
  BEMCS_REQUIRE(p->Exy && p->is_bound && p->n_sub && p->cnt_next && p->ppc && p->crossed_angle &&
                    p->diag && p->e_cnt_next,
                "bemcs_ion_step: null buffer in parameters");
  BEMCS_REQUIRE(p->nx >= 3 && p->ny >= 3, "bemcs_ion_step: mesh too small");
  BEMCS_REQUIRE((int64_t)p->nx * p->ny < (int64_t{1} << 28), "bemcs_ion_step: mesh of 2^28 nodes or more");
  BEMCS_REQUIRE(count_window_valid(p->count_window), "bemcs_ion_step: count_window beyond 2048 nodes");
  BEMCS_REQUIRE(n_bound <= in->cap && out->cap >= n_bound, "bemcs_ion_step: capacity below bound");
  if (n_bound <= 0) return BEMCS_OK;
  hipStream_t s = as_stream(stream);
  const bool fast = fastdiv_ok(p->dx32, s) && fastdiv_ok(p->dy32, s);
  const bemcs_ion_case_t a{*in, *out, *e_out, *p};
  auto kern = fast ? k_ion_step<true, false> : k_ion_step<false, false>;
  hipLaunchKernelGGL(kern, dim3(grid_for(n_bound, kBlock)), dim3(kBlock), 0, s, a,
                     (const bemcs_ion_case_t*)nullptr);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_electron_step(const bemcs_species_t* in, const bemcs_species_t* out,
                        const bemcs_electron_step_t* p, int64_t n_bound, bemcs_stream_t stream) {
  BEMCS_REQUIRE(in && out && p && in->count && out->count, "bemcs_electron_step: null argument");
  BEMCS_REQUIRE(p->Exy && p->is_bound && p->n_sub && p->cnt_next && p->overflow,
                "bemcs_electron_step: null buffer in parameters");
  BEMCS_REQUIRE(p->nx >= 3 && p->ny >= 3, "bemcs_electron_step: mesh too small");
  BEMCS_REQUIRE((int64_t)p->nx * p->ny < (int64_t{1} << 28), "bemcs_electron_step: mesh of 2^28 nodes or more");
  BEMCS_REQUIRE(count_window_valid(p->count_window), "bemcs_electron_step: count_window beyond 2048 nodes");
  BEMCS_REQUIRE(n_bound <= in->cap, "bemcs_electron_step: capacity below bound");
  if (n_bound <= 0) return BEMCS_OK;
  hipStream_t s = as_stream(stream);
  const bool fast = fastdiv_ok(p->dx32, s) && fastdiv_ok(p->dy32, s);
  const bemcs_electron_case_t a{*in, *out, *p};
  auto kern = fast ? k_electron_step<true, false> : k_electron_step<false, false>;
  hipLaunchKernelGGL(kern, dim3(grid_for(n_bound, kElecPerThread * kBlock)), dim3(kBlock), 0, s, a,
                     (const bemcs_electron_case_t*)nullptr);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_electron_push_chunk(const bemcs_species_t* sp, const bemcs_electron_step_t* p, int32_t chunk,
                              int64_t n_bound, bemcs_stream_t stream) {
  BEMCS_REQUIRE(sp && p && sp->count && p->Exy && p->n_sub, "bemcs_electron_push_chunk: null argument");
  BEMCS_REQUIRE(p->nx >= 3 && p->ny >= 3, "bemcs_electron_push_chunk: mesh too small");
  BEMCS_REQUIRE((int64_t)p->nx * p->ny < (int64_t{1} << 28), "bemcs_electron_push_chunk: mesh of 2^28 nodes or more");
  BEMCS_REQUIRE(p->sub_chunks >= 1 && chunk >= 0 && chunk < p->sub_chunks,
                "bemcs_electron_push_chunk: 0 <= chunk < sub_chunks");
  if (n_bound <= 0) return BEMCS_OK;
  hipStream_t s = as_stream(stream);
  const bool fast = fastdiv_ok(p->dx32, s) && fastdiv_ok(p->dy32, s);
  auto kern = fast ? k_electron_push<true> : k_electron_push<false>;
  hipLaunchKernelGGL(kern, dim3(grid_for(n_bound, kBlock)), dim3(kBlock), 0, s, *sp, *p, (int)chunk, n_bound);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_inject(const bemcs_species_t* sp, const bemcs_inject_t* p, int64_t n,
                 int64_t n_bound_before, bemcs_stream_t stream) {
  BEMCS_REQUIRE(sp && p && sp->count, "bemcs_inject: null argument");
  BEMCS_REQUIRE(p->kind >= 0 && p->kind <= 4, "bemcs_inject: bad kind");
  BEMCS_REQUIRE(p->nx >= 3 && p->ny >= 3, "bemcs_inject: mesh too small");
  BEMCS_REQUIRE((p->kind != 3 && p->kind != 4) || p->is_bound, "bemcs_inject: preload needs is_bound");
  BEMCS_REQUIRE(p->target <= 0 || ((p->kind == 3 || p->kind == 4) && p->id_next && n >= p->target),
                "bemcs_inject: population hold needs kind 3 or 4, id_next and n >= target");
  // a top-up ends at max(count, target); other injections add n
  BEMCS_REQUIRE(p->target > 0 ? (n_bound_before <= sp->cap && p->target <= sp->cap)
                              : n_bound_before + n <= sp->cap,
                "bemcs_inject: capacity exceeded");
  if (n <= 0) return BEMCS_OK;
  hipStream_t s = as_stream(stream);
  const bemcs_inject_case_t a{*sp, *p, n, n_bound_before};
  hipLaunchKernelGGL(k_inject<false>, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, a,
                     (const bemcs_inject_case_t*)nullptr);
  hipLaunchKernelGGL(k_add_count<false>, dim3(1), dim3(1), 0, s, a,
                     (const bemcs_inject_case_t*)nullptr);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_sort_by_cell(const bemcs_species_t* in, const bemcs_species_t* out, int64_t n_bound,
                       int32_t nx, int32_t ny, double dx, double dy, uint32_t* scratch,
                       bemcs_stream_t stream) {
  return bemcs_sort_by_tile(in, out, n_bound, nx, ny, dx, dy, 0, 0, scratch, stream);
}

int bemcs_sort_by_tile(const bemcs_species_t* in, const bemcs_species_t* out, int64_t n_bound,
                       int32_t nx, int32_t ny, double dx, double dy, int32_t shift_x, int32_t shift_y,
                       uint32_t* scratch, bemcs_stream_t stream) {
  BEMCS_REQUIRE(in && out && scratch && in->count && out->count, "bemcs_sort_by_cell: null argument");
  BEMCS_REQUIRE(out->cap >= n_bound && in->cap >= n_bound, "bemcs_sort_by_cell: capacity");
  BEMCS_REQUIRE(nx >= 3 && ny >= 3 && shift_x >= 0 && shift_x <= 8 && shift_y >= 0 && shift_y <= 8,
                "bemcs_sort_by_tile: mesh too small or tile shift outside 0..8");
  hipStream_t s = as_stream(stream);
  const bemcs_sort_case_t a{*in, *out, n_bound, nx, ny, dx, dy, scratch, shift_x, shift_y};
  const bemcs_sort_case_t* none = nullptr;
  const int64_t ncell = sort_bins(a);
  hipLaunchKernelGGL(k_sort_prepare<false>, dim3(grid_for(ncell, kBlock)), dim3(kBlock), 0, s, a, none);
  if (n_bound > 0) {
    const unsigned nb = grid_for(ncell, 1024);
    hipLaunchKernelGGL(k_cell_hist<false>, dim3(grid_for(n_bound, kBlock)), dim3(kBlock), 0, s, a, none);
    hipLaunchKernelGGL(k_scan_blocks<false>, dim3(nb), dim3(1024), 0, s, a, none);
    hipLaunchKernelGGL(k_scan_sums<false>, dim3(1), dim3(1024), 0, s, a, none);
    hipLaunchKernelGGL(k_scan_add<false>, dim3(nb), dim3(1024), 0, s, a, none);
    hipLaunchKernelGGL(k_cell_scatter<false>, dim3(grid_for(n_bound, kBlock)), dim3(kBlock), 0, s, a,
                       none);
  }
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

// ------------------------------------------------------- batched versions
static bool all_fast(const float* div, int32_t n, hipStream_t s) {
  if (!div) return false;
  for (int32_t c = 0; c < 2 * n; ++c)
    if (!fastdiv_ok(div[c], s)) return false;
  return true;
}

int bemcs_ion_step_batched(const bemcs_ion_case_t* cases, int32_t n, int64_t max_bound,
                           const float* divisors_host, bemcs_stream_t stream) {
  BEMCS_REQUIRE(cases && n >= 1, "bemcs_ion_step_batched: bad arguments");
  if (max_bound <= 0) return BEMCS_OK;
  hipStream_t s = as_stream(stream);
  auto kern = all_fast(divisors_host, n, s) ? k_ion_step<true, true> : k_ion_step<false, true>;
  hipLaunchKernelGGL(kern, dim3(grid_for(max_bound, kBlock), 1, n), dim3(kBlock), 0, s,
                     bemcs_ion_case_t{}, cases);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_electron_step_batched(const bemcs_electron_case_t* cases, int32_t n, int64_t max_bound,
                                const float* divisors_host, bemcs_stream_t stream) {
  BEMCS_REQUIRE(cases && n >= 1, "bemcs_electron_step_batched: bad arguments");
  if (max_bound <= 0) return BEMCS_OK;
  hipStream_t s = as_stream(stream);
  auto kern = all_fast(divisors_host, n, s) ? k_electron_step<true, true> : k_electron_step<false, true>;
  hipLaunchKernelGGL(kern, dim3(grid_for(max_bound, kElecPerThread * kBlock), 1, n), dim3(kBlock), 0, s,
                     bemcs_electron_case_t{}, cases);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_substeps_batched(const bemcs_substeps_case_t* cases, int32_t n, int64_t max_bound,
                           bemcs_stream_t stream) {
  BEMCS_REQUIRE(cases && n >= 1, "bemcs_substeps_batched: bad arguments");
  hipStream_t s = as_stream(stream);
  if (max_bound > 0)
    hipLaunchKernelGGL(k_substep_ds<true>, dim3(substep_grid(max_bound), 1, n), dim3(kBlock), 0,
                       s, bemcs_substeps_case_t{}, cases, max_bound);
  hipLaunchKernelGGL(k_substep_finalize<true>, dim3(1, 1, n), dim3(1), 0, s, bemcs_substeps_case_t{},
                     cases);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_inject_batched(const bemcs_inject_case_t* cases, int32_t n, int64_t max_n,
                         bemcs_stream_t stream) {
  BEMCS_REQUIRE(cases && n >= 1, "bemcs_inject_batched: bad arguments");
  if (max_n <= 0) return BEMCS_OK;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_inject<true>, dim3(grid_for(max_n, kBlock), 1, n), dim3(kBlock), 0, s,
                     bemcs_inject_case_t{}, cases);
  hipLaunchKernelGGL(k_add_count<true>, dim3(1, 1, n), dim3(1), 0, s, bemcs_inject_case_t{}, cases);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_sort_by_cell_batched(const bemcs_sort_case_t* cases, int32_t n, int64_t max_bound,
                               int64_t max_nodes, bemcs_stream_t stream) {
  BEMCS_REQUIRE(cases && n >= 1 && max_nodes > 0, "bemcs_sort_by_cell_batched: bad arguments");
  hipStream_t s = as_stream(stream);
  const bemcs_sort_case_t z{};
  hipLaunchKernelGGL(k_sort_prepare<true>, dim3(grid_for(max_nodes, kBlock), 1, n), dim3(kBlock), 0, s,
                     z, cases);
  if (max_bound > 0) {
    const unsigned nb = grid_for(max_nodes, 1024);
    hipLaunchKernelGGL(k_cell_hist<true>, dim3(grid_for(max_bound, kBlock), 1, n), dim3(kBlock), 0, s,
                       z, cases);
    hipLaunchKernelGGL(k_scan_blocks<true>, dim3(nb, 1, n), dim3(1024), 0, s, z, cases);
    hipLaunchKernelGGL(k_scan_sums<true>, dim3(1, 1, n), dim3(1024), 0, s, z, cases);
    hipLaunchKernelGGL(k_scan_add<true>, dim3(nb, 1, n), dim3(1024), 0, s, z, cases);
    hipLaunchKernelGGL(k_cell_scatter<true>, dim3(grid_for(max_bound, kBlock), 1, n), dim3(kBlock), 0,
                       s, z, cases);
  }
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

}  // extern "C"
