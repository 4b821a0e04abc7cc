// Poisson solve of the reference operator (Python/physics_engine.py:525-608)
// by a multigrid-preconditioned conjugate gradient, standing in for SuperLU.
//
// The reference matrix has Dirichlet rows (electrodes and column 0), Neumann
// rows (right edge; top/bottom edges in half_hole mode) and 5-point rows
// (-4, 1, 1, 1, 1) for the remaining "unknown" nodes (periodic in y for
// one_hole/two_holes).  Every Neumann node adjacent to an unknown node P
// mirrors P itself, so eliminating Dirichlet and Neumann rows leaves an SPD
// system on the unknown nodes:
//     d_P V_P - sum_{Q unknown nbr} V_Q = -b_P + sum_{Q other nbr} b_Q,
//     d_P = 4 - #(Neumann neighbours).
// Vectors live on the full ny x nx grid with zeros at eliminated nodes, so the
// fine operator is matrix-free (one byte per node).  Precision: the outer
// flexible CG keeps x, r, p, Ap and all dot products in float64; the
// preconditioner is one float32 V-cycle (bilinear prolongation, Galerkin
// 9-point coarse operators built on the device, damped Jacobi smoothing,
// dense inverse on the coarsest level, at most kCoarseMax nodes).
#include <math.h>
#include <stdlib.h>

#include <string.h>

#include <algorithm>
#include <functional>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "bemcs_internal.h"

namespace bemcs {

constexpr int kMaxLevels = 20;
constexpr int kCoarseMax = 2048;  // default: coarsening stops at or below this many nodes
constexpr int kDenseMax = 4096;  // largest coarsest level the dense solve accepts
constexpr float kOmega0 = 0.8f;  // Jacobi damping, 5-point fine level (one-sweep smoothing)
constexpr float kOmegaC = 0.7f;  // Jacobi damping, Galerkin 9-point levels (one-sweep smoothing)
// Two-sweep smoothing: the degree-2 Chebyshev polynomial of D^-1 A for the
// eigenvalue interval [0.5, 2] (Gershgorin bounds D^-1 A by 2 for these
// diagonally dominant operators, fine and Galerkin alike): sweep weights
// 1 / (1.25 +- 0.75 cos(pi / 4)).  Its largest amplification on [0.5, 2] is
// 1 / T2(5/3) = 0.22, against 0.36 for two Jacobi sweeps damped by 0.8; the
// two factors commute, so the V-cycle stays symmetric whatever their order.
constexpr float kChebA = 0.56169f;
constexpr float kChebB = 1.38953f;

enum : uint8_t { K_BOUND = 0, K_UNKNOWN = 1, K_RIGHT = 2, K_TOP = 3, K_BOTTOM = 4 };

struct OpData;  // operator arrays, shared by handles built from the same electrode map

struct LevelDev {
  int nx, ny, n, periodic;
  const float* st;      // [9][n] coarse stencil (k = (dy+1)*3 + (dx+1)), null on level 0
  const float* inv_d;   // [n] 1/diagonal, 0 at inactive nodes
  const uint8_t* code;  // level 0: diagonal d (0 = inactive)
  float* x;
  float* t;
  float* b;
};


}  // namespace bemcs

struct bemcs_poisson_s {
  int nx, ny, periodic, nlev;
  bemcs::LevelDev lev[bemcs::kMaxLevels];
  std::vector<void*> allocs;
  float* cinv;
  int nc;
  uint8_t* kind;
  int32_t* mirror;   // final mirror (unknown or Dirichlet) of Neumann nodes, -1 otherwise
  int32_t* mirror1;  // first hop of Neumann nodes, -1 otherwise
  double *x, *p, *Ap, *bred, *Vnew;
  double *p1, *Ap1, *Az;  // second direction buffers and A z (multi-level path)
  double* xg;             // solution of the solve before last (Picard extrapolation)
  double* R[2];       // residuals r_k, alternating
  int nparts;         // partial pairs per reduction buffer (level-0 tiles)
  double *pa, *pr, *pi;
  double* pz[2];
  bool have_guess;
  int tile0_pref;     // level-0 coarse tile edge asked for (0: by mesh size, kTC or kTCL)
  bemcs::OpData* op;  // the operator arrays this handle's pointers above lead into
  int op_gen;         // incremented when the handle moves to operator arrays of its own
  bemcs_poisson_batch_s* own;  // solve set of this handle alone
  // operator set-up workspace (in-place updates)
  double* su_st[2];   // float64 9-point stencils of two consecutive coarse levels
  uint8_t* su_act[2];  // their active flags
  double* su_A;       // dense coarsest operator, then its inverse
  double *su_C, *su_R, *su_T;
  double* su_d0;      // its original diagonal
  uint8_t* su_swept;  // pivots taken (0: vanished, generalised inverse)
  int* su_err;
};

namespace bemcs {

// per-case scalars; S_RZ0 + (k & 1) and S_PAP0 + (k & 1) hold r_{k-1}.z_{k-1}
// and p_k.Ap_k of iteration k, stored by the first work-group of the launch
// that sums them for the next iteration
enum { S_RR = 0, S_BB, S_TOL2, S_ITERS, S_RTOL, S_RZ0, S_RZ1, S_PAP0, S_PAP1, S_Q0, S_COUNT = 14 };

// ------------------------------------------------------------- 1D transfer
__host__ __device__ __forceinline__ float w1d(int f, int I, int nc, bool periodic) {
  if ((f & 1) == 0) return (f >> 1) == I ? 1.0f : 0.0f;
  int a = (f - 1) >> 1, b = (f + 1) >> 1;
  if (b >= nc) {
    if (periodic) b -= nc;
    else return a == I ? 1.0f : 0.0f;
  }
  return (a == I ? 0.5f : 0.0f) + (b == I ? 0.5f : 0.0f);
}

// ------------------------------------------------------- operator applies
__device__ __forceinline__ int nbr_s(int i, int x, int y, int nx, int ny, bool per) {
  return y > 0 ? i - nx : (per ? i + (ny - 1) * nx : -1);
}
__device__ __forceinline__ int nbr_n(int i, int x, int y, int nx, int ny, bool per) {
  return y < ny - 1 ? i + nx : (per ? i - (ny - 1) * nx : -1);
}

template <typename T>
__device__ __forceinline__ T apply5(const T* __restrict__ v, const uint8_t* __restrict__ code, int i,
                                    int nx, int ny, bool per) {
  int d = code[i];
  if (d == 0) return (T)0;
  int x = i % nx, y = i / nx;
  T s = v[i - 1] + v[i + 1] + v[nbr_s(i, x, y, nx, ny, per)] + v[nbr_n(i, x, y, nx, ny, per)];
  return (T)d * v[i] - s;
}



// ---------------------------------------------------------- V-cycle kernels
// One launch per level on the way down and one on the way up.  A work-group
// owns a tile and keeps what it needs, with halos, in LDS:
//   down (fine F -> coarse C), one coarse kTC x kTC tile: two damped-Jacobi
//     sweeps from zero on F (x1 = w D^-1 b on a halo-3 region, x2 on halo 2),
//     the residual b - A x2 on halo 1 and its restriction (transpose of the
//     bilinear prolongation) into C.b; x2 is stored to F.x for the owned fine
//     nodes (the 2 kTC x 2 kTC fine nodes under the coarse tile).
//   up (C -> F), one fine kTF x kTF tile: x + P e on a halo-`sweeps` region,
//     then `sweeps` Jacobi sweeps; the result goes to F.t.
// The result of a level is F.t (C.x on the coarsest level, from the dense
// solve).  Halo values are recomputed by every tile that needs them from
// inputs that no block of the same launch writes.
constexpr int kTC = 8;              // coarse tile edge of the level-0 down launch
constexpr int kTF = 2 * kTC;        // fine tile edge of the level-0 up launch
constexpr int kTCL = 16;            // the same for large meshes: 4x fewer partial sums,
constexpr int kTFL = 2 * kTCL;      // which every consumer block adds up
constexpr int kLargeMesh = 600000;  // nodes above which a solve set uses the large tiles
constexpr int kTCc = 8;             // the same on the coarser levels (more,
constexpr int kTFc = 2 * kTCc;      // smaller tiles: these launches are latency bound)
constexpr int kTCs = 4;             // small coarse-level tiles (fewer rounds per phase)
constexpr double kPicardExtrap = 0.5;  // initial-guess extrapolation of Picard solves
constexpr int kTFs = 2 * kTCs;

static_assert(kTC * kTC <= 256 && kTCc * kTCc <= 256, "one coarse node per thread in the restriction");

__device__ __forceinline__ int gnode(const LevelDev& L, int gx, int gy) {
  if (gx < 0 || gx >= L.nx) return -1;
  if (gy < 0 || gy >= L.ny) {
    if (!L.periodic) return -1;
    gy = (gy % L.ny + L.ny) % L.ny;
  }
  return gy * L.nx + gx;
}

// ---------------------------------------------------- block reductions
// Partial sums are pairs (a, b) per block.  A consumer block adds n pairs in
// a fixed order, so every block of a launch obtains the same bits and takes
// the same decisions without a separate scalar kernel.
__device__ __forceinline__ void block_partials(double a, double b, double* part, int slot) {
  __shared__ double sa[4], sb[4];
  a = wave_sum(a);
  b = wave_sum(b);
  if (lane_id() == 0) {
    sa[threadIdx.x / 64] = a;
    sb[threadIdx.x / 64] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * slot] = (sa[0] + sa[1]) + (sa[2] + sa[3]);
    part[2 * slot + 1] = (sb[0] + sb[1]) + (sb[2] + sb[3]);
  }
}

// Four partial sums per block (the level-0 up launch: r.z, r_old.z, z.Az,
// z.Ap) and their fixed-order total.
__device__ __forceinline__ void block_partials4(double a, double b, double c, double d, double* part,
                                                int slot) {
  __shared__ double s4[4][4];
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  d = wave_sum(d);
  if (lane_id() == 0) {
    const int w = threadIdx.x / 64;
    s4[0][w] = a;
    s4[1][w] = b;
    s4[2][w] = c;
    s4[3][w] = d;
  }
  __syncthreads();
  if (threadIdx.x < 4) part[4 * slot + threadIdx.x] =
      (s4[threadIdx.x][0] + s4[threadIdx.x][1]) + (s4[threadIdx.x][2] + s4[threadIdx.x][3]);
}

__device__ __forceinline__ void block_sum4(const double* __restrict__ part, int n, double out[4]) {
  __shared__ double sq[4][256];
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < n; k += 256) {
#pragma unroll
    for (int q = 0; q < 4; ++q) a[q] += part[4 * k + q];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) sq[q][threadIdx.x] = a[q];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
#pragma unroll
      for (int q = 0; q < 4; ++q) sq[q][threadIdx.x] += sq[q][threadIdx.x + o];
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) out[q] = sq[q][0];
  __syncthreads();
}

template <int NT = 256>
__device__ __forceinline__ void block_sum2(const double* __restrict__ part, int n, double& a_out,
                                           double& b_out) {
  __shared__ double sa[NT], sb[NT];
  double a = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < n; k += NT) {
    a += part[2 * k];
    b += part[2 * k + 1];
  }
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int o = NT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sa[threadIdx.x] += sa[threadIdx.x + o];
      sb[threadIdx.x] += sb[threadIdx.x + o];
    }
    __syncthreads();
  }
  a_out = sa[0];
  b_out = sb[0];
  __syncthreads();
}

// ---------------------------------------------------------- CG state
// Flexible (Polak-Ribiere) PCG in float64.  Iteration k:
//   dir   : beta_k = (r_{k-1}.z_{k-1} - r_{k-2}.z_{k-1}) / r_{k-2}.z_{k-2},
//           p = z + beta p, Ap = A z + beta Ap, partials p.Ap
//   step  : alpha = r_{k-1}.z_{k-1} / p.Ap, x += alpha p,
//           r_k = r_{k-1} - alpha Ap, partials r_k.r_k (fused into the first
//           down launch, whose right-hand side is f32(r_k))
//   check : the launch after the step stops the solve when r.r <= tol^2 b.b
//   V-cycle z_k = M r_k, the last up launch adds partials (r_k.z_k, r_{k-1}.z_k)
// r alternates between two vectors (r_k in R[k&1]) and the r.z partials
// between two buffers, so no launch overwrites what another block of the same
// launch reads.
struct CGArgs {
  double* x;
  double* R[2];
  double *p, *Ap;  // single-level path; the multi-level path alternates P[k&1], AP[k&1]
  double* P[2];
  double* AP[2];
  double* Az;      // A z of the last V-cycle (level-0 up launch)
  const double* bred;
  float* b0;       // level-0 right-hand side of the V-cycle
  const float* z;  // level-0 V-cycle output
  const uint8_t* code;
  int nx, ny, per, n;
  int nparts;      // partial pairs per buffer (one per level-0 tile)
  int presum;      // the level-0 up launch's partial sums are added by a launch of their own
  double* pa;      // p.Ap
  double* pr;      // r.r
  double* pz[2];   // (r_k.z_k, r_{k-1}.z_k) of V-cycle k in pz[k&1]
  double* pi;      // init: (r0.r0, b.b)
  double* scal;
  int32_t* done;
};

// a[i] for a two-entry array member without indexing memory (a select of the
// two values keeps argument structs in registers)
__device__ __forceinline__ double* pick2(double* const (&a)[2], int i) { return i ? a[1] : a[0]; }

struct CheckArgs {
  const double* pr;
  int nparts;
  int check;  // this launch directly follows the step: it runs the convergence test
  double* scal;
  int32_t* done;
  // non-null in the last test launch of a graph of a one-case set: the test
  // work-group also writes the scalars to pinned host memory and counts
  // itself in the host's flag (no separate status launch)
  double* host;
  int* flag;
};

// All S_COUNT scalars of the case into fine-grained pinned host memory, then
// the flag (the caller's thread 0 may just have written some of them).
__device__ __forceinline__ void write_status(const CheckArgs& ca) {
  __syncthreads();
  if (threadIdx.x < S_COUNT) ca.host[threadIdx.x] = ca.scal[threadIdx.x];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_fetch_add(ca.flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Convergence test of iteration k by one work-group (all NT threads take
// part): it records r.r and sets the stop flag, and the launches that follow
// exit.  The other work-groups of the launch carry on; after the last
// iteration their output is scratch that nothing reads, and they save every
// work-group adding up all partial sums.
template <int NT = 256>
__device__ __forceinline__ bool test_converged(const CheckArgs& ca, int k) {
  double rr, unused;
  block_sum2<NT>(ca.pr, ca.nparts, rr, unused);
  const bool stop = !(rr > ca.scal[S_TOL2]);  // converged, or NaN
  if (threadIdx.x == 0) {
    ca.scal[S_RR] = rr;
    ca.scal[S_ITERS] = (double)k;
    if (stop) *ca.done = 1;
  }
  return stop;
}

// The same by the first work-group of a launch (the coarsest solve of one-
// and two-level meshes); false for the other work-groups.
template <int NT = 256>
__device__ __forceinline__ bool check_converged(const CheckArgs& ca, int k) {
  if (blockIdx.x != 0 || blockIdx.y != 0) return false;
  return test_converged<NT>(ca, k);
}

// Arguments of one case for the V-cycle launches.  A launch serves one case
// (arguments by value) or a batch (blockIdx.z indexes a device array).
struct DownPack {
  LevelDev F, C;
  CGArgs g;
  CheckArgs ca;
};
struct UpPack {
  LevelDev F, C;  // C.x is the coarse result
  CGArgs g;
};
struct CoarsePack {
  LevelDev C;
  const float* cinv;
  CheckArgs ca;
};

// Launches that run a convergence test can see `done` change while they run
// (block 0 sets it); a block reads it once and shares it, so all of its waves
// take the same branch around the barriers that follow.
__device__ __forceinline__ bool block_done(const int32_t* done) {
  __shared__ int s;
  if (threadIdx.x == 0) s = *(volatile const int32_t*)done;
  __syncthreads();
  return s != 0;
}

// Tolerance and initial convergence test of a solve from the partial sums
// (r0.r0, b.b) of the r0 launch; all threads of the work-group take part.
__device__ __forceinline__ void init_scalars(const CGArgs& g, double rtol) {
  double rr, bb;
  block_sum2(g.pi, g.nparts, rr, bb);
  if (threadIdx.x != 0) return;
  g.scal[S_RTOL] = rtol;
  g.scal[S_RR] = rr;
  g.scal[S_BB] = bb;
  g.scal[S_TOL2] = rtol * rtol * bb;
  g.scal[S_ITERS] = 0.0;
  *g.done = (rr <= g.scal[S_TOL2]) ? 1 : 0;
}

// Workgroups are dispatched to the 8 XCDs round-robin by linear workgroup
// id, and each XCD has its own L2.  The tile launches map ids so that XCD k
// processes the k-th contiguous band of tiles (row-major) in every launch and
// on every level: what a band writes in one launch is read back from the same
// L2 in the next.  The map is a bijection of the (x, y) grid onto itself.
constexpr int kXcds = 8;
__device__ __forceinline__ void xcd_tile(int& tx, int& ty) {
  const int gx = (int)gridDim.x, n = gx * (int)gridDim.y;
  const int id = (int)blockIdx.y * gx + (int)blockIdx.x;
  const int q = n / kXcds, r = n % kXcds, xcd = id % kXcds, j = id / kXcds;
  const int t = xcd < r ? xcd * (q + 1) + j : r * (q + 1) + (xcd - r) * q + j;
  tx = t % gx;
  ty = t / gx;
}

// Batched launches cover the largest case's tile grid (columns rounded up to
// a multiple of 8, so a work-group's XCD is its linear id within its case's
// z-slice modulo 8).  Each case maps its own tiles_x x tiles_y tiles as
// xcd_tile maps a single grid: XCD k takes the k-th contiguous band of the
// case's tiles.  (Bands of the largest grid would leave the upper bands of
// small cases empty and load the first XCDs with the first rows of every
// case.)  Returns false for the work-groups beyond the case's tiles.
__device__ __forceinline__ bool xcd_case_tile(int tiles_x, int tiles_y, int& tx, int& ty) {
  const int id = (int)blockIdx.y * (int)gridDim.x + (int)blockIdx.x;
  const int n = tiles_x * tiles_y;
  const int q = n / kXcds, r = n % kXcds, xcd = id % kXcds, j = id / kXcds;
  if (j >= (xcd < r ? q + 1 : q)) return false;
  const int t = xcd < r ? xcd * (q + 1) + j : r * (q + 1) + (xcd - r) * q + j;
  tx = t % tiles_x;
  ty = t / tiles_x;
  return true;
}

// ------------------------------------------------------------ down launch
// Every global load of a tile is issued up front (clamped addresses, values
// selected afterwards), so a launch waits for memory about once; the
// Jacobi sweeps, the residual and the restriction then run from LDS.  The
// 9-point coefficients of coarse levels are staged in LDS with the values.
template <bool IS9>
__device__ __forceinline__ float tile_apply(const float* c, int W, const float* st, int nst, int ti,
                                            uint8_t code) {
  if (!IS9) {
    const float nb = (c[-1] + c[1]) + (c[-W] + c[W]);
    return (float)code * c[0] - nb;
  }
  float a = 0.0f;
#pragma unroll
  for (int k = 0; k < 9; ++k) a += st[k * nst + ti] * c[(k / 3 - 1) * W + (k % 3 - 1)];
  return a;
}

template <int TC, bool IS9, bool STEP, bool B>
__global__ __launch_bounds__(256) void k_down(DownPack one, const DownPack* __restrict__ many,
                                              int k, double rtol) {
  const DownPack P = B ? many[blockIdx.z] : one;
  const LevelDev& F = P.F;
  const LevelDev& C = P.C;
  const CGArgs& g = P.g;
  const int tiles_x = (C.nx + TC - 1) / TC;
  const int tiles_y = (C.ny + TC - 1) / TC;
  int tx, ty;
  if (B) {
    // the launches with a scalar work-group have one more tile row (below)
    const bool scalar_row = (!IS9 && !STEP) || (k > 0 && P.ca.check);
    if (!xcd_case_tile(tiles_x, tiles_y + (scalar_row ? 1 : 0), tx, ty)) return;
  } else {
    xcd_tile(tx, ty);
  }
  // The launches that compute a global scalar (the initial tolerance and stop
  // flag in the level-0 launch of the initial V-cycle, the convergence test
  // in the first coarse down launch of an iteration) have a grid one tile
  // row taller: the work-group at (0, tiles_y), which has no tile, computes
  // it, while the tile work-groups run beside it.  The other work-groups
  // carry on (after convergence their output is scratch nobody reads).
  const bool scalar_block = tx == 0 && ty == tiles_y;
  if (!IS9 && !STEP) {
    if (scalar_block) {
      init_scalars(g, rtol);
      return;
    }
  } else if (k > 0 && P.ca.check) {
    if (scalar_block) {
      // a solve that has converged keeps its count and residual
      if (!block_done(P.ca.done)) test_converged(P.ca, k);
      if (P.ca.host) write_status(P.ca);
      return;
    }
  }
  if (tx >= tiles_x || ty >= tiles_y) return;
  if ((IS9 || STEP) && !(k > 0 && P.ca.check)) {
    if (*P.ca.done) return;
  } else if (IS9 || STEP) {
    if (block_done(P.ca.done)) return;
  }
  constexpr int TFd = 2 * TC;                 // owned fine nodes per edge
  constexpr int W1 = TFd + 5, W2 = W1 - 2, W3 = W1 - 4;
  constexpr int N1 = W1 * W1, N2 = W2 * W2, N3 = W3 * W3;
  constexpr int IT1 = (N1 + 255) / 256, IT2 = (N2 + 255) / 256;
  __shared__ float sb[N1], sid[N1], s1[N1], s2[N2], sr[N3];
  __shared__ float sst[IS9 ? 9 * N2 : 1];
  __shared__ uint8_t scode[IS9 ? 1 : N2];
  const int CX0 = tx * TC, CY0 = ty * TC;
  const int ox = 2 * CX0 - 3, oy = 2 * CY0 - 3;  // fine coordinates of local (0, 0)
  const double* rp = pick2(g.R, (k + 1) & 1);
  double alpha = 0.0;
  double beta = 0.0;
  // Small level-0 tiles (meshes up to kLargeMesh nodes) form the CG direction
  // in this launch; large tiles take it from the direction launch, which is
  // cheaper there than the wider halo loads of the fused form.
  constexpr bool FUSED = STEP && TC == kTC;
  // fused: A z and the previous Ap; otherwise the direction launch's Ap
  const double* azp = FUSED ? g.Az : g.Ap;
  const double* apo = pick2(g.AP, (k + 1) & 1);  // Ap of iteration k-1 (fused)
  const double* po = pick2(g.P, (k + 1) & 1);
  // The owned nodes' CG values are loaded with the staged region (unfused:
  // p and x; r and Ap, just staged for the halo region, are read again from
  // the caches, which keeps the register count at five waves per SIMD), and
  // the scalar reductions run after both are issued, so the launch waits for
  // memory about once.  This launch writes none of the values it loads.
  constexpr int NS = STEP ? TFd * TFd / 256 : 1;
  double o_r[NS], o_x[NS], o_p[NS], o_ap[NS], o_apo[NS];
  float o_z[NS];
  if (STEP) {
#pragma unroll
    for (int it = 0; it < NS; ++it) {
      const int t = threadIdx.x + it * 256;
      const int gx = 2 * CX0 + t % TFd, gy = 2 * CY0 + t / TFd;
      const int gi = (gx < F.nx && gy < F.ny) ? gy * F.nx + gx : 0;
      if (FUSED) o_r[it] = rp[gi];
      o_x[it] = g.x[gi];
      if (FUSED) {
        o_z[it] = g.z[gi];
        o_p[it] = k > 1 ? po[gi] : 0.0;
        o_ap[it] = azp[gi];
        o_apo[it] = k > 1 ? apo[gi] : 0.0;
      } else {
        o_p[it] = g.p[gi];
      }
    }
  }
  {
    int vg[IT1];
    float vid[IT1], vb[IT1];
    double vr[STEP ? IT1 : 1], va[STEP ? IT1 : 1], vo[STEP ? IT1 : 1];
#pragma unroll
    for (int it = 0; it < IT1; ++it) {
      const int t = min((int)threadIdx.x + it * 256, N1 - 1);
      const int gi = gnode(F, ox + t % W1, oy + t / W1);
      const int gc = gi < 0 ? 0 : gi;
      vg[it] = gi;
      vid[it] = F.inv_d[gc];
      if (STEP) {
        vr[it] = rp[gc];
        va[it] = azp[gc];
        vo[it] = FUSED && k > 1 ? apo[gc] : 0.0;
      } else {
        vb[it] = F.b[gc];
      }
    }
    float vst[IS9 ? IT2 : 1][9];
    uint8_t vcode[IS9 ? 1 : IT2];
    int vg2[IT2];
#pragma unroll
    for (int it = 0; it < IT2; ++it) {
      const int t = min((int)threadIdx.x + it * 256, N2 - 1);
      const int gi = gnode(F, ox + 1 + t % W2, oy + 1 + t / W2);
      const int gc = gi < 0 ? 0 : gi;
      vg2[it] = gi;
      if (IS9) {
#pragma unroll
        for (int q = 0; q < 9; ++q) vst[it][q] = F.st[(size_t)q * F.n + gc];
      } else {
        vcode[it] = F.code[gc];
      }
    }
    if (STEP && !FUSED) {
      double pap, u;
      const double rz = g.scal[S_RZ0 + (k & 1)];  // block 0 of the direction launch
      block_sum2(g.pa, g.nparts, pap, u);
      alpha = pap != 0.0 ? rz / pap : 0.0;
    }
    if (FUSED) {
      // direction and step of iteration k in one launch: beta from the r.z
      // partial sums of the last V-cycle, p.Ap from the expansion
      // (z + beta p).(Az + beta Ap) = z.Az + 2 beta z.Ap + beta^2 p.Ap (A
      // symmetric), so the direction needs no launch and no reduction of its own
      double q[4];
      if (g.presum) {
#pragma unroll
        for (int j = 0; j < 4; ++j) q[j] = g.scal[S_Q0 + j];
      } else {
        block_sum4(pick2(g.pz, (k + 1) & 1), g.nparts, q);
      }
      const double rz = q[0], roz = q[1], zaz = q[2], zap = q[3];
      double pap = zaz;
      if (k > 1) {
        const double rz_old = g.scal[S_RZ0 + ((k - 1) & 1)];
        beta = rz_old != 0.0 ? (rz - roz) / rz_old : 0.0;
        pap = zaz + 2.0 * beta * zap + beta * beta * g.scal[S_PAP0 + ((k - 1) & 1)];
      }
      alpha = pap != 0.0 ? rz / pap : 0.0;
      if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        g.scal[S_RZ0 + (k & 1)] = rz;
        g.scal[S_PAP0 + (k & 1)] = pap;
      }
    }
#pragma unroll
    for (int it = 0; it < IT1; ++it) {
      const int t = threadIdx.x + it * 256;
      if (t >= N1) break;
      const bool in = vg[it] >= 0;
      const float id = in ? vid[it] : 0.0f;
      float b = 0.0f;
      if (in) {
        if (STEP) {
          const double ap = FUSED && k > 1 ? va[it] + beta * vo[it] : va[it];  // Ap of iteration k
          b = (float)(vr[it] - alpha * ap);
        } else {
          b = vb[it];
        }
      }
      sb[t] = b;
      sid[t] = id;
      s1[t] = kChebA * id * b;
    }
#pragma unroll
    for (int it = 0; it < IT2; ++it) {
      const int t = threadIdx.x + it * 256;
      if (t >= N2) break;
      const bool in = vg2[it] >= 0;
      if (IS9) {
#pragma unroll
        for (int q = 0; q < 9; ++q) sst[q * N2 + t] = in ? vst[it][q] : 0.0f;
      } else {
        scode[t] = in ? vcode[it] : (uint8_t)0;
      }
    }
  }
  if (STEP) {
    // the owned nodes take the CG step; their b is the value staged above
    double rr = 0.0;
    double* rn_out = pick2(g.R, k & 1);
    double* pn = pick2(g.P, k & 1);
    double* apn = pick2(g.AP, k & 1);
#pragma unroll
    for (int it = 0; it < NS; ++it) {
      const int t = threadIdx.x + it * 256;
      const int gx = 2 * CX0 + t % TFd, gy = 2 * CY0 + t / TFd;
      if (gx >= F.nx || gy >= F.ny) continue;
      const int gi = gy * F.nx + gx;
      double pi, api;
      if (FUSED) {
        const double zi = (double)o_z[it];
        pi = k > 1 ? zi + beta * o_p[it] : zi;
        api = k > 1 ? o_ap[it] + beta * o_apo[it] : o_ap[it];
        pn[gi] = pi;
        apn[gi] = api;
      } else {
        pi = o_p[it];
        api = g.Ap[gi];
      }
      const double rn = (FUSED ? o_r[it] : rp[gi]) - alpha * api;
      g.x[gi] = o_x[it] + alpha * pi;
      rn_out[gi] = rn;
      F.b[gi] = (float)rn;
      rr += rn * rn;
    }
    block_partials(rr, 0.0, g.pr, ty * tiles_x + tx);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < N2; t += 256) {
    const int lx = t % W2 + 1, ly = t / W2 + 1;
    const int li = ly * W1 + lx;
    const float id = sid[li];
    float v = 0.0f;
    if (id != 0.0f) {
      v = s1[li] + kChebB * id * (sb[li] - tile_apply<IS9>(&s1[li], W1, sst, N2, t, IS9 ? 0 : scode[t]));
    }
    s2[t] = v;
    const int gx = ox + lx, gy = oy + ly;
    if (lx >= 3 && lx < 3 + TFd && ly >= 3 && ly < 3 + TFd && gx < F.nx && gy < F.ny)
      F.x[gy * F.nx + gx] = v;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < N3; t += 256) {
    const int lx = t % W3 + 2, ly = t / W3 + 2;
    const int li = ly * W1 + lx;
    const float id = sid[li];
    float r = 0.0f;
    if (id != 0.0f) {
      const int t2 = (ly - 1) * W2 + (lx - 1);
      r = sb[li] - tile_apply<IS9>(&s2[t2], W2, sst, N2, t2, IS9 ? 0 : scode[t2]);
    }
    sr[t] = r;
  }
  __syncthreads();
  if (threadIdx.x >= TC * TC) return;
  const int cx = CX0 + (int)(threadIdx.x % TC), cy = CY0 + (int)(threadIdx.x / TC);
  if (cx >= C.nx || cy >= C.ny) return;
  const int I = cy * C.nx + cx;
  float s = 0.0f;
  if (C.inv_d[I] != 0.0f) {
    for (int dy = -1; dy <= 1; ++dy) {
      const int fyu = 2 * cy + dy;
      int fy = fyu;
      if (fy < 0 || fy >= F.ny) {
        if (!F.periodic) continue;
        fy = (fy + F.ny) % F.ny;
      }
      const float wy = w1d(fy, cy, C.ny, F.periodic != 0);
      if (wy == 0.0f) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int fx = 2 * cx + dx;
        if (fx < 0 || fx >= F.nx) continue;
        const float wx = w1d(fx, cx, C.nx, false);
        if (wx == 0.0f) continue;
        s += wx * wy * sr[(fyu - (2 * CY0 - 1)) * W3 + (fx - (2 * CX0 - 1))];
      }
    }
  }
  C.b[I] = s;
}

// ----------------------------------------------- two-level down launch
// The down launches of two coarse levels in one (F -> M -> C, 9-point
// operators, non-periodic meshes): a work-group owns a kT2 x kT2 tile of C,
// recomputes the right-hand side of M on the (2 kT2 + 5)^2 region its
// M-level sweeps need (the restriction of F's residual on a region of F
// around it), then runs the M level as the single-level launch does.  Every
// value is computed by the same operations in the same order as two
// single-level launches, so the V-cycle gives the same bits; one launch and
// one global round trip fewer.
constexpr int kT2 = 4;
struct Down2Pack {
  LevelDev F, M, C;
  const int32_t* done;
};

template <bool B>
__global__ __launch_bounds__(256) void k_down2(Down2Pack one, const Down2Pack* __restrict__ many) {
  const Down2Pack P = B ? many[blockIdx.z] : one;
  const LevelDev& F = P.F;
  const LevelDev& M = P.M;
  const LevelDev& C = P.C;
  const int tiles_x = (C.nx + kT2 - 1) / kT2;
  int tx, ty;
  if (B) {
    if (!xcd_case_tile(tiles_x, (C.ny + kT2 - 1) / kT2, tx, ty)) return;
  } else {
    xcd_tile(tx, ty);
  }
  if (tx >= tiles_x || ty * kT2 >= C.ny) return;
  if (*P.done) return;
  // M level: the single-level launch's regions for a kT2 tile of C
  constexpr int MW1 = 2 * kT2 + 5, MW2 = MW1 - 2, MW3 = MW1 - 4;
  constexpr int MN1 = MW1 * MW1, MN2 = MW2 * MW2, MN3 = MW3 * MW3;
  // F level: the same for the MW1 x MW1 region of M
  constexpr int FW1 = 2 * MW1 + 5, FW2 = FW1 - 2, FW3 = FW1 - 4;
  constexpr int FN1 = FW1 * FW1, FN2 = FW2 * FW2, FN3 = FW3 * FW3;
  constexpr int FIT1 = (FN1 + 255) / 256, FIT2 = (FN2 + 255) / 256;
  __shared__ float fb[FN1], fid[FN1], f1[FN1], f2[FN2], fr[FN3];
  __shared__ float fst[9 * FN2];
  __shared__ float mb[MN1], mid[MN1], m1[MN1], m2[MN2], mr[MN3];
  __shared__ float mst[9 * MN2];
  const int CX0 = tx * kT2, CY0 = ty * kT2;
  const int mx0 = 2 * CX0 - 3, my0 = 2 * CY0 - 3;  // M coordinates of the M region's (0, 0)
  const int fx0 = 2 * mx0 - 3, fy0 = 2 * my0 - 3;  // F coordinates of the F region's (0, 0)
  {
    int vg[FIT1];
    float vid[FIT1], vb[FIT1];
#pragma unroll
    for (int it = 0; it < FIT1; ++it) {
      const int t = min((int)threadIdx.x + it * 256, FN1 - 1);
      const int gi = gnode(F, fx0 + t % FW1, fy0 + t / FW1);
      const int gc = gi < 0 ? 0 : gi;
      vg[it] = gi;
      vid[it] = F.inv_d[gc];
      vb[it] = F.b[gc];
    }
    float vst[FIT2][9];
    int vg2[FIT2];
#pragma unroll
    for (int it = 0; it < FIT2; ++it) {
      const int t = min((int)threadIdx.x + it * 256, FN2 - 1);
      const int gi = gnode(F, fx0 + 1 + t % FW2, fy0 + 1 + t / FW2);
      const int gc = gi < 0 ? 0 : gi;
      vg2[it] = gi;
#pragma unroll
      for (int q = 0; q < 9; ++q) vst[it][q] = F.st[(size_t)q * F.n + gc];
    }
#pragma unroll
    for (int it = 0; it < FIT1; ++it) {
      const int t = threadIdx.x + it * 256;
      if (t >= FN1) break;
      const bool in = vg[it] >= 0;
      const float id = in ? vid[it] : 0.0f;
      const float b = in ? vb[it] : 0.0f;
      fb[t] = b;
      fid[t] = id;
      f1[t] = kChebA * id * b;
    }
#pragma unroll
    for (int it = 0; it < FIT2; ++it) {
      const int t = threadIdx.x + it * 256;
      if (t >= FN2) break;
      const bool in = vg2[it] >= 0;
#pragma unroll
      for (int q = 0; q < 9; ++q) fst[q * FN2 + t] = in ? vst[it][q] : 0.0f;
    }
    // M level: 1/diagonal and coefficients (the right-hand side follows)
    for (int t = threadIdx.x; t < MN1; t += 256) {
      const int gi = gnode(M, mx0 + t % MW1, my0 + t / MW1);
      mid[t] = gi >= 0 ? M.inv_d[gi] : 0.0f;
    }
    for (int t = threadIdx.x; t < MN2; t += 256) {
      const int gi = gnode(M, mx0 + 1 + t % MW2, my0 + 1 + t / MW2);
#pragma unroll
      for (int q = 0; q < 9; ++q) mst[q * MN2 + t] = gi >= 0 ? M.st[(size_t)q * M.n + gi] : 0.0f;
    }
  }
  __syncthreads();
  // F: second sweep; the F nodes under this work-group's M tile keep it
  for (int t = threadIdx.x; t < FN2; t += 256) {
    const int lx = t % FW2 + 1, ly = t / FW2 + 1;
    const int li = ly * FW1 + lx;
    const float id = fid[li];
    float v = 0.0f;
    if (id != 0.0f) v = f1[li] + kChebB * id * (fb[li] - tile_apply<true>(&f1[li], FW1, fst, FN2, t, 0));
    f2[t] = v;
    const int gx = fx0 + lx, gy = fy0 + ly;
    if (lx >= 9 && lx < 9 + 4 * kT2 && ly >= 9 && ly < 9 + 4 * kT2 && gx < F.nx && gy < F.ny)
      F.x[gy * F.nx + gx] = v;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < FN3; t += 256) {
    const int lx = t % FW3 + 2, ly = t / FW3 + 2;
    const int li = ly * FW1 + lx;
    const float id = fid[li];
    float r = 0.0f;
    if (id != 0.0f) {
      const int t2 = (ly - 1) * FW2 + (lx - 1);
      r = fb[li] - tile_apply<true>(&f2[t2], FW2, fst, FN2, t2, 0);
    }
    fr[t] = r;
  }
  __syncthreads();
  // restriction to the M region (the single-level restriction, node by node);
  // the M nodes under this work-group's C tile store their right-hand side
  for (int t = threadIdx.x; t < MN1; t += 256) {
    const int cx = mx0 + t % MW1, cy = my0 + t / MW1;
    float s = 0.0f;
    if (cx >= 0 && cx < M.nx && cy >= 0 && cy < M.ny) {
      const int I = cy * M.nx + cx;
      if (M.inv_d[I] != 0.0f) {
        for (int dy = -1; dy <= 1; ++dy) {
          const int fy = 2 * cy + dy;
          if (fy < 0 || fy >= F.ny) continue;
          const float wy = w1d(fy, cy, M.ny, false);
          if (wy == 0.0f) continue;
          for (int dx = -1; dx <= 1; ++dx) {
            const int fx = 2 * cx + dx;
            if (fx < 0 || fx >= F.nx) continue;
            const float wx = w1d(fx, cx, M.nx, false);
            if (wx == 0.0f) continue;
            s += wx * wy * fr[(fy - (fy0 + 2)) * FW3 + (fx - (fx0 + 2))];
          }
        }
      }
      const int lx = t % MW1, ly = t / MW1;
      if (lx >= 3 && lx < 3 + 2 * kT2 && ly >= 3 && ly < 3 + 2 * kT2) M.b[I] = s;
    }
    mb[t] = s;
    m1[t] = kChebA * mid[t] * s;
  }
  __syncthreads();
  // M level, as the single-level launch for the C tile
  for (int t = threadIdx.x; t < MN2; t += 256) {
    const int lx = t % MW2 + 1, ly = t / MW2 + 1;
    const int li = ly * MW1 + lx;
    const float id = mid[li];
    float v = 0.0f;
    if (id != 0.0f) v = m1[li] + kChebB * id * (mb[li] - tile_apply<true>(&m1[li], MW1, mst, MN2, t, 0));
    m2[t] = v;
    const int gx = mx0 + lx, gy = my0 + ly;
    if (lx >= 3 && lx < 3 + 2 * kT2 && ly >= 3 && ly < 3 + 2 * kT2 && gx < M.nx && gy < M.ny)
      M.x[gy * M.nx + gx] = v;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < MN3; t += 256) {
    const int lx = t % MW3 + 2, ly = t / MW3 + 2;
    const int li = ly * MW1 + lx;
    const float id = mid[li];
    float r = 0.0f;
    if (id != 0.0f) {
      const int t2 = (ly - 1) * MW2 + (lx - 1);
      r = mb[li] - tile_apply<true>(&m2[t2], MW2, mst, MN2, t2, 0);
    }
    mr[t] = r;
  }
  __syncthreads();
  if (threadIdx.x >= kT2 * kT2) return;
  const int cx = CX0 + (int)(threadIdx.x % kT2), cy = CY0 + (int)(threadIdx.x / kT2);
  if (cx >= C.nx || cy >= C.ny) return;
  const int I = cy * C.nx + cx;
  float s = 0.0f;
  if (C.inv_d[I] != 0.0f) {
    for (int dy = -1; dy <= 1; ++dy) {
      const int fy = 2 * cy + dy;
      if (fy < 0 || fy >= M.ny) continue;
      const float wy = w1d(fy, cy, C.ny, false);
      if (wy == 0.0f) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int fx = 2 * cx + dx;
        if (fx < 0 || fx >= M.nx) continue;
        const float wx = w1d(fx, cx, C.nx, false);
        if (wx == 0.0f) continue;
        s += wx * wy * mr[(fy - (2 * CY0 - 1)) * MW3 + (fx - (2 * CX0 - 1))];
      }
    }
  }
  C.b[I] = s;
}

// -------------------------------------------------------------- up launch
// The up to four (index, weight) terms of P e at fine node i; absent terms
// have weight 0 and index 0.
__device__ __forceinline__ void pe_terms(const LevelDev& F, const LevelDev& C, int i, int* idx,
                                         float* w) {
  // bilinear prolongation, the weights w1d gives: an even fine index takes its
  // coarse node with weight 1, an odd one the two coarse nodes beside it with
  // 1/2 each, or the lower one alone with weight 1 where the upper one is past
  // a non-periodic edge (periodic in y: the upper row wraps)
  const int fx = i % F.nx, fy = i / F.nx;
  const int x0 = fx >> 1, y0 = fy >> 1;
  const bool x1ok = (fx & 1) && x0 + 1 < C.nx;
  int y1 = y0 + 1;
  bool y1ok = (fy & 1) != 0;
  if (y1 >= C.ny) {
    if (F.periodic) y1 -= C.ny;
    else y1ok = false;
  }
  const float wx0 = x1ok ? 0.5f : 1.0f, wx1 = x1ok ? 0.5f : 0.0f;
  const float wy0 = y1ok ? 0.5f : 1.0f, wy1 = y1ok ? 0.5f : 0.0f;
  const int x1 = x1ok ? x0 + 1 : 0;
  idx[0] = y0 * C.nx + x0;
  idx[1] = x1ok ? y0 * C.nx + x1 : 0;
  idx[2] = y1ok ? y1 * C.nx + x0 : 0;
  idx[3] = (y1ok && x1ok) ? y1 * C.nx + x1 : 0;
  w[0] = wy0 * wx0;
  w[1] = wy0 * wx1;
  w[2] = wy1 * wx0;
  w[3] = wy1 * wx1;
}

// C.x must point at the coarse result (C.t, or C.x of the coarsest level).
template <int TF, int SWEEPS, bool IS9, bool DOT, bool B>
__global__ __launch_bounds__(256) void k_up(UpPack one, const UpPack* __restrict__ many, int k) {
  const UpPack P = B ? many[blockIdx.z] : one;
  const LevelDev& F = P.F;
  const LevelDev& C = P.C;
  const CGArgs& g = P.g;
  const int tiles_x = (F.nx + TF - 1) / TF;
  int tx, ty;
  if (B) {
    if (!xcd_case_tile(tiles_x, (F.ny + TF - 1) / TF, tx, ty)) return;
  } else {
    xcd_tile(tx, ty);
  }
  if (tx >= tiles_x || ty * TF >= F.ny) return;
  if (*g.done) return;
  // the level-0 launch of the fused form (small level-0 tiles) also computes
  // z on a ring around its tile, for A z at its own nodes
  constexpr bool FUSED = DOT && TF == kTF;
  constexpr int EXT = FUSED ? 1 : 0, H = SWEEPS + EXT;
  constexpr int W0 = TF + 2 * H, W1 = W0 - 2;  // the coefficient region is W1
  constexpr int N0 = W0 * W0, N1 = W1 * W1;
  constexpr int IT0 = (N0 + 255) / 256, IT1 = (N1 + 255) / 256;
  constexpr int TZ = TF + 2 * EXT;  // region of the final sweep
  __shared__ float sx[N0], sb[N0], sid[N0], s1[SWEEPS == 2 ? N1 : 1];
  __shared__ float sst[IS9 ? 9 * N1 : 1];
  __shared__ uint8_t scode[IS9 ? 1 : N1];
  __shared__ float sz[FUSED ? TZ * TZ : 1];
  const float om = IS9 ? kOmegaC : kOmega0;
  const int ox = tx * TF - H, oy = ty * TF - H;
  // fused: the owned nodes' r_k, r_{k-1}, Ap_{k-1} and operator code, used
  // after the sweeps, are loaded now so their latency overlaps the sweeps
  // (unfused: r_k and r_{k-1} only, loaded after the staging below)
  constexpr int NOWN = DOT ? (TF * TF + 255) / 256 : 1;
  double o_r[NOWN], o_ro[NOWN], o_ap[NOWN];
  int o_code[NOWN];
  if (FUSED) {
#pragma unroll
    for (int it = 0; it < NOWN; ++it) {
      const int t = min((int)threadIdx.x + it * 256, TF * TF - 1);
      const int gx = ox + H + t % TF, gy = oy + H + t / TF;
      const int gc = (gx < F.nx && gy < F.ny) ? gy * F.nx + gx : 0;
      o_r[it] = pick2(g.R, k & 1)[gc];
      o_ro[it] = pick2(g.R, (k + 1) & 1)[gc];
      o_ap[it] = k > 0 ? pick2(g.AP, k & 1)[gc] : 0.0;
      o_code[it] = F.code[gc];
    }
  }
  {
    int vg[IT0];
    float vx[IT0], vb[IT0], vid[IT0], vc[IT0][4], vw[IT0][4];
#pragma unroll
    for (int it = 0; it < IT0; ++it) {
      const int t = min((int)threadIdx.x + it * 256, N0 - 1);
      const int gi = gnode(F, ox + t % W0, oy + t / W0);
      const int gc = gi < 0 ? 0 : gi;
      vg[it] = gi;
      vid[it] = F.inv_d[gc];
      vb[it] = F.b[gc];
      vx[it] = F.x[gc];
      int ci[4];
      pe_terms(F, C, gc, ci, vw[it]);
#pragma unroll
      for (int q = 0; q < 4; ++q) vc[it][q] = C.x[ci[q]];
    }
    float vst[IS9 ? IT1 : 1][9];
    uint8_t vcode[IS9 ? 1 : IT1];
    int vg1[IT1];
#pragma unroll
    for (int it = 0; it < IT1; ++it) {
      const int t = min((int)threadIdx.x + it * 256, N1 - 1);
      const int gi = gnode(F, ox + 1 + t % W1, oy + 1 + t / W1);
      const int gc = gi < 0 ? 0 : gi;
      vg1[it] = gi;
      if (IS9) {
#pragma unroll
        for (int q = 0; q < 9; ++q) vst[it][q] = F.st[(size_t)q * F.n + gc];
      } else {
        vcode[it] = F.code[gc];
      }
    }
#pragma unroll
    for (int it = 0; it < IT0; ++it) {
      const int t = threadIdx.x + it * 256;
      if (t >= N0) break;
      const bool in = vg[it] >= 0;
      const float id = in ? vid[it] : 0.0f;
      float v = 0.0f;
      if (id != 0.0f) {
        const float pe = ((vw[it][0] * vc[it][0] + vw[it][1] * vc[it][1]) +
                          (vw[it][2] * vc[it][2] + vw[it][3] * vc[it][3]));
        v = vx[it] + pe;
      }
      sx[t] = v;
      sb[t] = in ? vb[it] : 0.0f;
      sid[t] = id;
    }
#pragma unroll
    for (int it = 0; it < IT1; ++it) {
      const int t = threadIdx.x + it * 256;
      if (t >= N1) break;
      const bool in = vg1[it] >= 0;
      if (IS9) {
#pragma unroll
        for (int q = 0; q < 9; ++q) sst[q * N1 + t] = in ? vst[it][q] : 0.0f;
      } else {
        scode[t] = in ? vcode[it] : (uint8_t)0;
      }
    }
  }
  if (DOT && !FUSED) {
    // unfused: the owned nodes' r_k and r_{k-1}, loaded once the staged
    // values are in LDS (fewer live registers) so the sweeps hide their latency
#pragma unroll
    for (int it = 0; it < NOWN; ++it) {
      const int t = min((int)threadIdx.x + it * 256, TF * TF - 1);
      const int gx = ox + H + t % TF, gy = oy + H + t / TF;
      const int gc = (gx < F.nx && gy < F.ny) ? gy * F.nx + gx : 0;
      o_r[it] = pick2(g.R, k & 1)[gc];
      o_ro[it] = pick2(g.R, (k + 1) & 1)[gc];
    }
  }
  __syncthreads();
  if (SWEEPS == 2) {
    for (int t = threadIdx.x; t < N1; t += 256) {
      const int lx = t % W1 + 1, ly = t / W1 + 1;
      const int li = ly * W0 + lx;
      const float id = sid[li];
      float v = 0.0f;
      if (id != 0.0f)
        v = sx[li] + kChebB * id * (sb[li] - tile_apply<IS9>(&sx[li], W0, sst, N1, t, IS9 ? 0 : scode[t]));
      s1[t] = v;
    }
    __syncthreads();
  }
  double rz = 0.0, roz = 0.0, zaz = 0.0, zap = 0.0;
  constexpr int NZ = (TZ * TZ + 255) / 256;
#pragma unroll
  for (int it = 0; it < NZ; ++it) {
    const int t = threadIdx.x + it * 256;
    if (t >= TZ * TZ) break;
    const int lx = t % TZ + SWEEPS, ly = t / TZ + SWEEPS;
    const int gx = ox + lx, gy = oy + ly;
    const int li = ly * W0 + lx;
    const int t1 = (ly - 1) * W1 + (lx - 1);  // index in the coefficient region
    const float id = sid[li];
    float v = 0.0f;
    if (id != 0.0f) {
      const float* c = SWEEPS == 2 ? &s1[t1] : &sx[li];
      const int W = SWEEPS == 2 ? W1 : W0;
      v = c[0] + (SWEEPS == 2 ? kChebA : om) * id * (sb[li] - tile_apply<IS9>(c, W, sst, N1, t1, IS9 ? 0 : scode[t1]));
    }
    if (FUSED) {
      sz[t] = v;
      continue;
    }
    if (gx >= F.nx || gy >= F.ny) continue;
    const int gi = gy * F.nx + gx;
    F.t[gi] = v;
    if (DOT && it < NOWN) {  // unfused: TZ = TF, so t indexes the owned nodes as above
      rz += o_r[it] * (double)v;
      roz += o_ro[it] * (double)v;
    }
  }
  if (FUSED) {
    __syncthreads();
    // owned nodes: z, A z with the level-0 operator (neighbours as the
    // single-level direction kernel takes them), partial sums
#pragma unroll
    for (int it = 0; it < NOWN; ++it) {
      const int t = threadIdx.x + it * 256;
      if (t >= TF * TF) break;
      const int ux = t % TF + EXT, uy = t / TF + EXT;  // position in the sz region
      const int gx = ox + SWEEPS + ux, gy = oy + SWEEPS + uy;
      if (gx >= F.nx || gy >= F.ny) continue;
      const int gi = gy * F.nx + gx;
      const float v = sz[uy * TZ + ux];
      F.t[gi] = v;
      double az = 0.0;
      const int d = o_code[it];
      if (d) {
        const float zs = (gy > 0 || F.periodic) ? sz[(uy - 1) * TZ + ux] : v;
        const float zn = (gy < F.ny - 1 || F.periodic) ? sz[(uy + 1) * TZ + ux] : v;
        const double nb =
            (double)sz[uy * TZ + ux - 1] + (double)sz[uy * TZ + ux + 1] + (double)zs + (double)zn;
        az = (double)d * (double)v - nb;
      }
      g.Az[gi] = az;
      rz += o_r[it] * (double)v;
      roz += o_ro[it] * (double)v;
      zaz += (double)v * az;
      if (k > 0) zap += (double)v * o_ap[it];
    }
  }
  if (FUSED) block_partials4(rz, roz, zaz, zap, pick2(g.pz, k & 1), ty * tiles_x + tx);
  else if (DOT) block_partials(rz, roz, pick2(g.pz, k & 1), ty * tiles_x + tx);
}

// coarsest level: x = C^-1 b.  The inverse is stored column-major and is
// symmetric up to rounding, so row i is read as the contiguous column i: one
// coalesced 256-byte load per wavefront and step.  One wavefront per row and
// four per work-group spread the matrix reads over many CUs; a row's
// work-group has the same id, hence the same XCD, in every launch, so its
// rows stay in that XCD's L2 between V-cycles.  The lane partial sums are
// added by a fixed butterfly (deterministic).
constexpr int kCoarseRowsPerBlock = 4;
template <bool B>
__global__ __launch_bounds__(64 * kCoarseRowsPerBlock) void k_coarse_solve(CoarsePack one,
                                                                           const CoarsePack* __restrict__ many,
                                                                           int k) {
  const CoarsePack P = B ? many[blockIdx.z] : one;
  const LevelDev& C = P.C;
  if ((int)blockIdx.x * kCoarseRowsPerBlock >= C.n) return;
  if (k > 0 && P.ca.check) {
    if (block_done(P.ca.done) || check_converged<64 * kCoarseRowsPerBlock>(P.ca, k)) return;
  } else if (*P.ca.done) {
    return;
  }
  const int nc = C.n;
  const int lane = lane_id();
  const int row = (int)blockIdx.x * kCoarseRowsPerBlock + (int)(threadIdx.x >> 6);
  if (row >= nc) return;
  const float* __restrict__ col = P.cinv + (size_t)row * nc;
  const float* __restrict__ b = C.b;
  // eight accumulators per lane over chunks of 8 x 64 columns: the 16 loads
  // of a chunk are in flight together (the row is latency bound, not
  // bandwidth bound: it comes from the XCD's L2)
  constexpr int U = 8;
  float sa[U];
#pragma unroll
  for (int q = 0; q < U; ++q) sa[q] = 0.0f;
  int j0 = 0;
  for (; j0 + U * 64 <= nc; j0 += U * 64) {
    float cv[U], bv[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      cv[q] = col[j0 + q * 64 + lane];
      bv[q] = b[j0 + q * 64 + lane];
    }
#pragma unroll
    for (int q = 0; q < U; ++q) sa[q] += cv[q] * bv[q];
  }
#pragma unroll
  for (int q = 0; q < U; ++q) {
    const int j = j0 + q * 64 + lane;
    if (j < nc) sa[q] += col[j] * b[j];
  }
  float s = ((sa[0] + sa[1]) + (sa[2] + sa[3])) + ((sa[4] + sa[5]) + (sa[6] + sa[7]));
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) C.x[row] = s;
}

// ----------------------------------------------------------- CG kernels
// Node loops over g.nparts blocks, one partial pair per block.  Nodes are
// split into min(8, nparts) contiguous bands (row-major, so bands of rows
// like the tile bands of the V-cycle launches) and band k is strided over
// the blocks whose id is k modulo the band count, which run on XCD k.
__device__ __forceinline__ int band_count(int nparts) { return nparts < kXcds ? nparts : kXcds; }
__device__ __forceinline__ int band_start(int n, int nparts) {
  const int nb = band_count(nparts), band = (int)blockIdx.x % nb, j = (int)blockIdx.x / nb;
  return (int)((long long)n * band / nb) + j * 256 + (int)threadIdx.x;
}
__device__ __forceinline__ int band_end(int n, int nparts) {
  const int nb = band_count(nparts), band = (int)blockIdx.x % nb;
  return (int)((long long)n * (band + 1) / nb);
}
__device__ __forceinline__ int band_step(int nparts) {
  const int nb = band_count(nparts), band = (int)blockIdx.x % nb;
  return ((nparts - band + nb - 1) / nb) * 256;
}
#define GRID_LOOP(i, n)                                                                  \
  for (int i = band_start((n), g.nparts), i##_end = band_end((n), g.nparts),            \
           i##_step = band_step(g.nparts);                                              \
       i < i##_end; i += i##_step)
#define CG_CASE()                               \
  const CGArgs g = B ? many[blockIdx.z] : one;  \
  if ((int)blockIdx.x >= g.nparts) return

// r0 = b - A x -> R[0], R[1] = 0, b0 = f32(r0); partials (r0.r0, b.b)
template <bool B>
__global__ __launch_bounds__(256) void k_cg_init(CGArgs one, const CGArgs* __restrict__ many) {
  CG_CASE();
  double rr = 0, bb = 0;
  GRID_LOOP(i, g.n) {
    double ri = g.code[i] ? g.bred[i] - apply5<double>(g.x, g.code, i, g.nx, g.ny, g.per != 0) : 0.0;
    g.R[0][i] = ri;
    g.R[1][i] = 0.0;
    g.b0[i] = (float)ri;
    rr += ri * ri;
    bb += g.bred[i] * g.bred[i];
  }
  block_partials(rr, bb, g.pi, blockIdx.x);
}

// tolerance and the initial convergence test
template <bool B>
__global__ __launch_bounds__(256) void k_cg_init_scalar(CGArgs one, const CGArgs* __restrict__ many,
                                                        double rtol) {
  CG_CASE();
  init_scalars(g, rtol);
}

// The PCG scalars of every case written straight into fine-grained pinned
// host memory by the last launch of a PCG graph, which then adds 1 to a
// pinned counter (the host clears it before launching and polls it, which
// wakes it sooner than a blocking stream synchronisation).
__global__ __launch_bounds__(256) void k_status(const double* __restrict__ scal, double* host, int count,
                                                int* flag) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) host[i] = scal[i];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_fetch_add(flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// A batch's sums of the level-0 up launch's partial quads (r.z, r_old.z,
// z.Az, z.Ap) of iteration k, one work-group per case, for the step launch.
template <bool B>
__global__ __launch_bounds__(256) void k_sum_quads(CGArgs one, const CGArgs* __restrict__ many, int k) {
  const CGArgs g = B ? many[blockIdx.z] : one;
  if (*g.done) return;
  double q[4];
  block_sum4(pick2(g.pz, (k + 1) & 1), g.nparts, q);
  if (threadIdx.x < 4) g.scal[S_Q0 + threadIdx.x] = q[threadIdx.x];
}

// p = z + beta p ; Ap = A z + beta Ap ; partials p.Ap.  A thread's nodes
// are taken kDirBatch at a time: every load of a batch is issued before any
// of its stores (stores to p and Ap would otherwise keep the next node's
// loads waiting), and the first batch before the reduction that gives beta.
// Each thread adds its nodes in increasing order, as a one-node loop would.
constexpr int kDirBatch = 4;
struct DirNodes {
  int d[kDirBatch];
  float zc[kDirBatch], zw[kDirBatch], ze[kDirBatch], zs[kDirBatch], zn[kDirBatch];
  double p[kDirBatch], ap[kDirBatch];
};

__device__ __forceinline__ void dir_load(const CGArgs& g, int k, int base, int end, int step, DirNodes& v) {
#pragma unroll
  for (int u = 0; u < kDirBatch; ++u) {
    const int i = base + u * step;
    const int c = i < end ? i : base;  // base < end: a node of this thread
    const int y = c / g.nx;
    const int sn = y > 0 ? c - g.nx : (g.per ? c + (g.ny - 1) * g.nx : c);
    const int nn = y < g.ny - 1 ? c + g.nx : (g.per ? c - (g.ny - 1) * g.nx : c);
    v.d[u] = g.code[c];
    v.zc[u] = g.z[c];
    // a node with a nonzero code has both x neighbours (the clamps only keep
    // the loads of the other nodes in bounds)
    v.zw[u] = g.z[c > 0 ? c - 1 : c];
    v.ze[u] = g.z[c < g.n - 1 ? c + 1 : c];
    v.zs[u] = g.z[sn];
    v.zn[u] = g.z[nn];
    v.p[u] = k > 1 ? g.p[c] : 0.0;
    v.ap[u] = k > 1 ? g.Ap[c] : 0.0;
  }
}

__device__ __forceinline__ void dir_store(const CGArgs& g, int k, double beta, int base, int end, int step,
                                          const DirNodes& v, double& s) {
#pragma unroll
  for (int u = 0; u < kDirBatch; ++u) {
    const int i = base + u * step;
    if (i >= end) break;
    double az = 0.0;
    if (v.d[u]) {
      const double nb = (double)v.zw[u] + (double)v.ze[u] + (double)v.zs[u] + (double)v.zn[u];
      az = (double)v.d[u] * (double)v.zc[u] - nb;
    }
    const double pi = k > 1 ? (double)v.zc[u] + beta * v.p[u] : (double)v.zc[u];
    const double api = k > 1 ? az + beta * v.ap[u] : az;
    g.p[i] = pi;
    g.Ap[i] = api;
    s += pi * api;
  }
}

template <bool B>
__global__ __launch_bounds__(256) void k_cg_dir(CGArgs one, const CGArgs* __restrict__ many, int k) {
  CG_CASE();
  if (*g.done) return;
  const int i0 = band_start(g.n, g.nparts), end = band_end(g.n, g.nparts), step = band_step(g.nparts);
  DirNodes v;
  if (i0 < end) dir_load(g, k, i0, end, step, v);
  double beta = 0.0;
  if (k > 1 || blockIdx.x == 0) {
    double rz, roz;
    block_sum2(pick2(g.pz, (k + 1) & 1), g.nparts, rz, roz);
    if (k > 1) {
      const double rz_old = g.scal[S_RZ0 + ((k - 1) & 1)];  // block 0 of iteration k-1
      beta = rz_old != 0.0 ? (rz - roz) / rz_old : 0.0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) g.scal[S_RZ0 + (k & 1)] = rz;
  }
  double s = 0;
  for (int base = i0; base < end; base += kDirBatch * step) {
    if (base != i0) dir_load(g, k, base, end, step, v);
    dir_store(g, k, beta, base, end, step, v, s);
  }
  block_partials(s, 0.0, g.pa, blockIdx.x);
}

// single-level hierarchy: the step without a fused down sweep
template <bool B>
__global__ __launch_bounds__(256) void k_cg_step(CGArgs one, const CGArgs* __restrict__ many, int k) {
  CG_CASE();
  if (*g.done) return;
  double rz, pap, u;
  block_sum2(pick2(g.pz, (k + 1) & 1), g.nparts, rz, u);
  block_sum2(g.pa, g.nparts, pap, u);
  const double a = pap != 0.0 ? rz / pap : 0.0;
  const double* rp = pick2(g.R, (k + 1) & 1);
  double* rn_out = pick2(g.R, k & 1);
  double rr = 0;
  GRID_LOOP(i, g.n) {
    g.x[i] += a * g.p[i];
    const double rn = rp[i] - a * g.Ap[i];
    rn_out[i] = rn;
    g.b0[i] = (float)rn;
    rr += rn * rn;
  }
  block_partials(rr, 0.0, g.pr, blockIdx.x);
}

// single-level hierarchy: partials (r_k.z, r_{k-1}.z)
template <bool B>
__global__ __launch_bounds__(256) void k_cg_dotz(CGArgs one, const CGArgs* __restrict__ many, int k) {
  CG_CASE();
  if (*g.done) return;
  double a = 0, b = 0;
  GRID_LOOP(i, g.n) {
    const double zi = g.z[i];
    a += pick2(g.R, k & 1)[i] * zi;
    b += pick2(g.R, (k + 1) & 1)[i] * zi;
  }
  block_partials(a, b, pick2(g.pz, k & 1), blockIdx.x);
}

// --------------------------------------------------- reduced system set-up
// b_red[P] = -b[P] + sum_{non-unknown nbrs Q} b[Q] at unknown nodes, 0 else;
// x0 = V at unknown nodes (initial guess)
__global__ __launch_bounds__(256) void k_reduce_rhs(const double* __restrict__ b,
                                                    const double* __restrict__ V,
                                                    const uint8_t* __restrict__ kind, double* bred,
                                                    double* x, int nx, int ny, int per, int guess) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nx * ny) return;
  if (kind[i] != K_UNKNOWN) {
    bred[i] = 0.0;
    x[i] = 0.0;
    return;
  }
  int xx = i % nx, y = i / nx;
  int nb[4] = {i - 1, i + 1, nbr_s(i, xx, y, nx, ny, per != 0), nbr_n(i, xx, y, nx, ny, per != 0)};
  double s = -b[i];
  for (int k = 0; k < 4; ++k)
    if (kind[nb[k]] != K_UNKNOWN) s += b[nb[k]];
  bred[i] = s;
  if (guess) x[i] = V[i];
}

// Full solution from the reduced one: unknown -> x, Dirichlet -> b (value),
// Neumann -> value of its mirror plus the row's right-hand sides on the way.
__global__ __launch_bounds__(256) void k_assemble(const double* __restrict__ x,
                                                  const double* __restrict__ bfull,
                                                  const uint8_t* __restrict__ kind,
                                                  const int32_t* __restrict__ mirror,
                                                  const int32_t* __restrict__ mirror1, double* V,
                                                  int n) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uint8_t k = kind[i];
  if (k == K_UNKNOWN) {
    V[i] = x[i];
  } else if (k == K_BOUND) {
    V[i] = bfull[i];
  } else {
    int m = mirror[i];
    double base = kind[m] == K_UNKNOWN ? x[m] : bfull[m];
    int m1 = mirror1[i];
    double extra = (m1 != m) ? bfull[m1] : 0.0;
    V[i] = (base + extra) + bfull[i];
  }
}

// E = -grad V (np.gradient, edge_order 1) with the periodic Ey fix; float32
__device__ __forceinline__ void efield_at(const double* __restrict__ V, int nx, int ny, double dx_m,
                                          double dy_m, int per, float* Exy, int i) {
  int x = i % nx, y = i / nx;
  double ex, ey;
  if (x == 0) ex = (V[i] - V[i + 1]) / dx_m;
  else if (x == nx - 1) ex = (V[i - 1] - V[i]) / dx_m;
  else ex = (V[i - 1] - V[i + 1]) / (2.0 * dx_m);
  if (per && y == 0) ey = (V[(ny - 1) * nx + x] - V[nx + x]) / (2.0 * dy_m);
  else if (per && y == ny - 1) ey = (V[(ny - 2) * nx + x] - V[x]) / (2.0 * dy_m);
  else if (y == 0) ey = (V[i] - V[i + nx]) / dy_m;
  else if (y == ny - 1) ey = (V[i - nx] - V[i]) / dy_m;
  else ey = (V[i - nx] - V[i + nx]) / (2.0 * dy_m);
  reinterpret_cast<float2*>(Exy)[i] = make_float2((float)ex, (float)ey);
}

__global__ __launch_bounds__(256) void k_efield(const double* __restrict__ V, int nx, int ny,
                                                double dx_m, double dy_m, int per, float* Exy) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nx * ny) efield_at(V, nx, ny, dx_m, dy_m, per, Exy, i);
}

// ------------------------------------------------------------ Picard loop
// Per-case arguments of the Picard loop (:886-902) and the field update.
struct PicardPack {
  const uint32_t* cnt_ion;
  const uint32_t* cnt_e;
  const unsigned long long* w_ion;  // bilinear weight sums (null: NGP counts)
  const unsigned long long* w_e;
  double cd_ion, cd_e;
  float* rho;
  const double* V_fixed;
  double* V;
  double* Vnew;
  const uint8_t* kind;
  const int32_t* mirror;
  double* bred;
  const double* x;
  double* xg;       // x of the solve before last (initial-guess extrapolation)
  // the case's PCG stop flag: the launches between two solves run only while
  // it is 1 (the solve before them has converged and they have not run yet)
  const int32_t* done;
  const double* xr; // last solution, relaxed into V at the start of a sweep
  double* xw;       // x, writable (initial-guess extrapolation)
  int nx, ny, per, n;
  double coeff, Vp, Te, qn0, one_m_omega, omega, dx_m, dy_m;
  float* Exy;
};

// Picard phases launched on their own (rho from the deposit and the
// right-hand side are part of k_picard_start)
enum { PIC_RELAX = 2, PIC_EFIELD = 3 };

// rho at node i from NGP counts as the reference's two np.add.at deposits
// produce it: per particle rho = f32(f64(rho) + cd), ions first, then
// electrons; or from the bilinear weight sums, as bemcs_weights_to_rho on a
// zero rho
__device__ __forceinline__ float deposit_rho(const PicardPack& P, int i) {
  float r = 0.0f;
  if (P.w_ion) {
    r = (float)((double)r + P.cd_ion * ((double)(long long)P.w_ion[i] * kCicUnit));
    r = (float)((double)r + P.cd_e * ((double)(long long)P.w_e[i] * kCicUnit));
  } else {
    for (uint32_t k = P.cnt_ion[i]; k > 0; --k) r = (float)((double)r + P.cd_ion);
    for (uint32_t k = P.cnt_e[i]; k > 0; --k) r = (float)((double)r + P.cd_e);
  }
  return r;
}

// The launches between two solves of a Picard call (and after the last one)
// are enqueued before the host has read whether the solve converged within
// the iterations it launched.  They run only for a case whose stop flag is 1;
// a case that still needs iterations skips them, and the host enqueues them
// again after its further iterations.  In a batch, k_mark_started then sets
// the flag of every case that ran them to 2, so a repeat skips those cases
// (the PCG launches skip any nonzero flag; a solve's first V-cycle resets it).
__device__ __forceinline__ bool stage_blocked(const int32_t* done) { return *done != 1; }

__global__ void k_mark_started(int32_t* done, int n) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n && done[c] == 1) done[c] = 2;
}

template <int PHASE, bool B>
__global__ __launch_bounds__(256) void k_picard(PicardPack one, const PicardPack* __restrict__ many) {
  const PicardPack P = B ? many[blockIdx.z] : one;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.n) return;
  if (stage_blocked(P.done)) return;
  if (PHASE == PIC_RELAX) {
    // Dirichlet values from V_fixed, Neumann rows mirror;
    // V = (1-omega) V + omega V_new
    const uint8_t k = P.kind[i];
    double vn;
    if (k == K_UNKNOWN) vn = P.x[i];
    else if (k == K_BOUND) vn = P.V_fixed[i];
    else {
      const int m = P.mirror[i];
      vn = P.kind[m] == K_UNKNOWN ? P.x[m] : P.V_fixed[m];
    }
    P.Vnew[i] = vn;
    P.V[i] = P.one_m_omega * P.V[i] + P.omega * vn;
  } else {
    efield_at(P.V, P.nx, P.ny, P.dx_m, P.dy_m, P.per, P.Exy, i);
  }
}

// Initial guess of Picard solve j >= 2 from the last two solutions:
// x_j^0 = x_{j-1} + theta (x_{j-1} - x_{j-2}); the solutions of successive
// Picard sweeps change by a nearly constant factor, so the guess starts the
// solve closer to its solution.  theta = 0 only saves x_{j-1}.
template <bool B>
__global__ __launch_bounds__(256) void k_picard_extrap(PicardPack one, const PicardPack* __restrict__ many,
                                                       double theta) {
  const PicardPack P = B ? many[blockIdx.z] : one;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.n || stage_blocked(P.done)) return;
  const double x1 = P.xw[i];
  if (theta != 0.0) P.xw[i] = x1 + theta * (x1 - P.xg[i]);
  P.xg[i] = x1;
}

// One launch for the start of Picard iteration j: the relaxation of
// iteration j-1 (RELAX), the right-hand side of iteration j and the initial
// PCG residual r0 = b - A x with its partials (r0.r0, b.b).  Each node needs
// only its own V (relaxed here) and neighbouring x and V_fixed, which no
// thread writes, so the three steps fuse exactly.
// save (second sweep of a call with extrapolation): relax from the solve's x
// and save it as the solution before last, which a separate extrapolation
// launch with theta = 0 would otherwise copy first.
template <bool RELAX, bool B>
__global__ __launch_bounds__(256) void k_picard_start(PicardPack pone, const PicardPack* __restrict__ pmany,
                                                      CGArgs one, const CGArgs* __restrict__ many, int save) {
  const CGArgs g = B ? many[blockIdx.z] : one;
  if ((int)blockIdx.x >= g.nparts) return;
  if (RELAX && stage_blocked(g.done)) return;
  const PicardPack P = B ? pmany[blockIdx.z] : pone;
  double rr = 0.0, bb = 0.0;
  GRID_LOOP(i, P.n) {
    const uint8_t kd = P.kind[i];
    double v = P.V[i];
    if (RELAX) {
      const double* xr = save ? P.x : P.xr;
      if (save) P.xg[i] = P.x[i];
      double vn;
      if (kd == K_UNKNOWN) vn = xr[i];
      else if (kd == K_BOUND) vn = P.V_fixed[i];
      else {
        const int m = P.mirror[i];
        vn = P.kind[m] == K_UNKNOWN ? xr[m] : P.V_fixed[m];
      }
      P.Vnew[i] = vn;
      v = P.one_m_omega * v + P.omega * vn;
      P.V[i] = v;
    }
    // the first sweep of a call computes rho from the deposit (every node)
    float rho_i;
    if (RELAX) {
      rho_i = kd == K_UNKNOWN ? P.rho[i] : 0.0f;
    } else {
      rho_i = deposit_rho(P, i);
      P.rho[i] = rho_i;
    }
    double b = 0.0;
    if (kd == K_UNKNOWN) {
      const double rho_e = P.qn0 * det_exp(((v < P.Vp ? v : P.Vp) - P.Vp) / P.Te);
      const double rt = (double)rho_i + rho_e;
      b = -((-P.coeff) * rt);
      const int xx = i % P.nx, y = i / P.nx;
      const int nb[4] = {i - 1, i + 1, nbr_s(i, xx, y, P.nx, P.ny, P.per != 0),
                         nbr_n(i, xx, y, P.nx, P.ny, P.per != 0)};
      for (int k = 0; k < 4; ++k)
        if (P.kind[nb[k]] == K_BOUND) b += P.V_fixed[nb[k]];
    }
    P.bred[i] = b;
    const double ri = g.code[i] ? b - apply5<double>(g.x, g.code, i, g.nx, g.ny, g.per != 0) : 0.0;
    g.R[0][i] = ri;
    g.R[1][i] = 0.0;
    g.b0[i] = (float)ri;
    rr += ri * ri;
    bb += b * b;
  }
  block_partials(rr, bb, g.pi, blockIdx.x);
}

// ------------------------------------------------------ operator set-up
// The reduced operator of build_sparse_matrix (:525-626) and its multigrid
// hierarchy are built on the device from the electrode map, into the
// handle's existing arrays: level shapes depend only on (nx, ny), so a new
// map (erosion, thermal bow) updates a handle in place and its captured
// launch graphs stay valid.
//   k_setup_kind     node kinds (Dirichlet, right / top / bottom Neumann, unknown)
//   k_setup_mirror   Neumann mirrors (first hop and final node)
//   k_setup_level0   level-0 diagonal 4 - #Neumann neighbours, 1/diagonal
//   k_galerkin       coarse 9-point operator P^T A P (bilinear P), one thread
//                    per coarse node gathering its fine contributions in a
//                    fixed order (deterministic)
//   k_galerkin_finish  drop couplings of inactive nodes, 1/diagonal, float32 copy
//   k_dense_*, k_gj_*  inverse of the coarsest operator by blocked in-place
//                    Gauss-Jordan, 16-wide panels.  The coarse operator is
//                    positive semidefinite: fine nodes enclosed by electrodes
//                    (eroded cells) can make the prolongation columns of their
//                    coarse parents dependent.  A pivot that has fallen below
//                    1e-10 of its original diagonal is skipped and its row and
//                    column are zeroed at the end, which gives a generalised
//                    inverse that solves every consistent coarse system.
constexpr int kGJPanel = 16;
enum : int { SETUP_ERR_CHAIN = 1, SETUP_ERR_MIRROR = 2, SETUP_ERR_EDGE = 4 };

__device__ __forceinline__ uint8_t node_kind(const uint8_t* is_bound, int i, int nx, int ny, bool per) {
  const int x = i % nx, y = i / nx;
  if (is_bound[i]) return K_BOUND;
  if (x == nx - 1) return K_RIGHT;
  if (!per && y == ny - 1) return K_TOP;
  if (!per && y == 0) return K_BOTTOM;
  return K_UNKNOWN;
}

__device__ __forceinline__ int mirror1_of(const uint8_t* kind, int i, int nx, int ny) {
  const int x = i % nx, y = i / nx;
  const uint8_t k = kind[i];
  if (k == K_RIGHT) return y * nx + (nx - 2);
  if (k == K_TOP) return (ny - 2) * nx + x;
  if (k == K_BOTTOM) return nx + x;
  return -1;
}

__global__ __launch_bounds__(256) void k_setup_kind(const uint8_t* __restrict__ is_bound, uint8_t* kind,
                                                    int nx, int ny, int per) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nx * ny) kind[i] = node_kind(is_bound, i, nx, ny, per != 0);
}

__global__ __launch_bounds__(256) void k_setup_mirror(const uint8_t* __restrict__ kind, int32_t* mirror,
                                                      int32_t* mirror1, int nx, int ny, int* err) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nx * ny) return;
  int m = mirror1_of(kind, i, nx, ny);
  mirror1[i] = m;
  if (m < 0) {
    mirror[i] = -1;
    return;
  }
  int hops = 0;
  while (kind[m] != K_UNKNOWN && kind[m] != K_BOUND && hops < 4) {
    m = mirror1_of(kind, m, nx, ny);
    ++hops;
  }
  if (hops >= 2) atomicOr(err, SETUP_ERR_CHAIN);
  mirror[i] = m;
}

__global__ __launch_bounds__(256) void k_setup_level0(const uint8_t* __restrict__ kind,
                                                      const int32_t* __restrict__ mirror1, uint8_t* code,
                                                      float* inv0, int nx, int ny, int per, int* err) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nx * ny) return;
  if (kind[i] != K_UNKNOWN) {
    code[i] = 0;
    inv0[i] = 0.0f;
    return;
  }
  const int x = i % nx, y = i / nx;
  const int s = y > 0 ? i - nx : (per ? i + (ny - 1) * nx : -1);
  const int nn = y < ny - 1 ? i + nx : (per ? i - (ny - 1) * nx : -1);
  if (x == 0 || x == nx - 1 || s < 0 || nn < 0) {
    atomicOr(err, SETUP_ERR_EDGE);
    code[i] = 0;
    inv0[i] = 0.0f;
    return;
  }
  const int nb[4] = {i - 1, i + 1, s, nn};
  int d = 4;
  for (int k = 0; k < 4; ++k) {
    const uint8_t t = kind[nb[k]];
    if (t == K_RIGHT || t == K_TOP || t == K_BOTTOM) {
      if (mirror1[nb[k]] != i) atomicOr(err, SETUP_ERR_MIRROR);
      --d;
    }
  }
  code[i] = (uint8_t)d;
  inv0[i] = (float)(1.0 / (double)d);
}

// 1D bilinear prolongation parents of fine index f (coarse nc), weights 1 or 1/2.
__device__ __forceinline__ int parents1d(int f, int nc, bool per, int* idx, double* w) {
  if ((f & 1) == 0) {
    idx[0] = f >> 1;
    w[0] = 1.0;
    return 1;
  }
  int a = (f - 1) >> 1, b = (f + 1) >> 1;
  if (b >= nc) {
    if (per) b -= nc;
    else {
      idx[0] = a;
      w[0] = 1.0;
      return 1;
    }
  }
  if (a == b) {
    idx[0] = a;
    w[0] = 1.0;
    return 1;
  }
  idx[0] = a;
  w[0] = 0.5;
  idx[1] = b;
  w[1] = 0.5;
  return 2;
}

// Fine level read by the Galerkin product: level 0 from kinds and codes
// (5-point, -1 to unknown neighbours), coarser levels from their float64
// 9-point stencils.
struct FineOp {
  int nx, ny;
  const uint8_t* kind;  // level 0
  const uint8_t* code;  // level 0
  const double* st;     // level >= 1: [9][n]
  const uint8_t* act;   // level >= 1
};

template <bool L0>
__device__ __forceinline__ bool fine_active(const FineOp& F, int f) {
  return L0 ? F.kind[f] == K_UNKNOWN : F.act[f] != 0;
}

template <bool L0>
__device__ __forceinline__ double fine_coef(const FineOp& F, int k, int f, int g) {
  if (!L0) return F.st[(size_t)k * F.nx * F.ny + f];
  if (k == 4) return (double)F.code[f];
  if (k == 1 || k == 3 || k == 5 || k == 7) return F.kind[g] == K_UNKNOWN ? -1.0 : 0.0;
  return 0.0;
}

template <bool L0>
__global__ __launch_bounds__(256) void k_galerkin(FineOp F, double* stC, uint8_t* actC, int ncx,
                                                  int ncy, int per) {
  const int I = blockIdx.x * 256 + threadIdx.x;
  const int nC = ncx * ncy;
  if (I >= nC) return;
  const int cx = I % ncx, cy = I / ncx;
  double acc[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) acc[k] = 0.0;
  for (int oy = -1; oy <= 1; ++oy) {
    int fy = 2 * cy + oy;
    if (fy < 0 || fy >= F.ny) {
      if (!per) continue;
      fy = (fy + F.ny) % F.ny;
    }
    int pyi[2];
    double pyw[2];
    const int npy = parents1d(fy, ncy, per != 0, pyi, pyw);
    double wy = 0.0;
    for (int a = 0; a < npy; ++a)
      if (pyi[a] == cy) wy = pyw[a];
    if (wy == 0.0) continue;
    for (int ox = -1; ox <= 1; ++ox) {
      const int fx = 2 * cx + ox;
      if (fx < 0 || fx >= F.nx) continue;
      int pxi[2];
      double pxw[2];
      const int npx = parents1d(fx, ncx, false, pxi, pxw);
      double wx = 0.0;
      for (int a = 0; a < npx; ++a)
        if (pxi[a] == cx) wx = pxw[a];
      if (wx == 0.0) continue;
      const int f = fy * F.nx + fx;
      if (!fine_active<L0>(F, f)) continue;
      for (int k = 0; k < 9; ++k) {
        int gx = fx + (k % 3 - 1), gy = fy + (k / 3 - 1);
        if (gx < 0 || gx >= F.nx) continue;
        if (gy < 0 || gy >= F.ny) {
          if (!per) continue;
          gy = (gy + F.ny) % F.ny;
        }
        const int g = gy * F.nx + gx;
        const double a = fine_coef<L0>(F, k, f, g);
        if (a == 0.0 || !fine_active<L0>(F, g)) continue;
        const double w1 = wx * wy * a;
        int qxi[2], qyi[2];
        double qxw[2], qyw[2];
        const int nqx = parents1d(gx, ncx, false, qxi, qxw);
        const int nqy = parents1d(gy, ncy, per != 0, qyi, qyw);
        for (int by = 0; by < nqy; ++by)
          for (int bx = 0; bx < nqx; ++bx) {
            const int ddx = qxi[bx] - cx;
            int ddy = qyi[by] - cy;
            if (per) {
              if (ddy > 1) ddy -= ncy;
              if (ddy < -1) ddy += ncy;
            }
            acc[(ddy + 1) * 3 + (ddx + 1)] += w1 * qxw[bx] * qyw[by];
          }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) stC[(size_t)k * nC + I] = acc[k];
  actC[I] = acc[4] > 1e-12 ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_galerkin_finish(double* stC, const uint8_t* __restrict__ actC,
                                                         float* st32, float* inv, int ncx, int ncy,
                                                         int per) {
  const int I = blockIdx.x * 256 + threadIdx.x;
  const int nC = ncx * ncy;
  if (I >= nC) return;
  const int cx = I % ncx, cy = I / ncx;
  const bool aI = actC[I] != 0;
  for (int k = 0; k < 9; ++k) {
    int gx = cx + (k % 3 - 1), gy = cy + (k / 3 - 1);
    bool ok = gx >= 0 && gx < ncx;
    if (gy < 0 || gy >= ncy) {
      if (per) gy = (gy + ncy) % ncy;
      else ok = false;
    }
    double v = stC[(size_t)k * nC + I];
    if (!ok || !aI || !actC[gy * ncx + gx]) v = 0.0;
    stC[(size_t)k * nC + I] = v;
    st32[(size_t)k * nC + I] = (float)v;
  }
  inv[I] = aI ? (float)(1.0 / stC[(size_t)4 * nC + I]) : 0.0f;
}

// Dense matrix of the coarsest operator, row-major; inactive rows are the
// identity (the active block is decoupled from them).
template <bool L0>
__global__ __launch_bounds__(256) void k_dense_assemble(FineOp L, double* A, double* d0, int per) {
  const int n = L.nx * L.ny;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int i = (int)(e / n), j = (int)(e % n);
  double v = 0.0;
  if (!fine_active<L0>(L, i)) {
    v = (i == j) ? 1.0 : 0.0;
  } else {
    const int cx = i % L.nx, cy = i / L.nx;
    for (int k = 0; k < 9; ++k) {
      int gx = cx + (k % 3 - 1), gy = cy + (k / 3 - 1);
      if (gx < 0 || gx >= L.nx) continue;
      if (gy < 0 || gy >= L.ny) {
        if (!per) continue;
        gy = (gy + L.ny) % L.ny;
      }
      const int g = gy * L.nx + gx;
      if (g != j || !fine_active<L0>(L, g)) continue;
      v += fine_coef<L0>(L, k, i, g);
    }
  }
  A[e] = v;
  if (i == j) d0[i] = v;
}

// Panel step of the blocked in-place Gauss-Jordan inversion, part 1: every
// work-group inverts the p x p diagonal block T = A[K,K]^-1 in LDS (scalar
// in-place Gauss-Jordan over the pivots that have not vanished; the others
// keep zero rows and columns in T), then C[i,:] = A[i,K] T for its rows and
// R[:,j] = A[K,j] for its columns; work-group 0 stores T and the panel's
// pivot flags.
__global__ __launch_bounds__(256) void k_gj_panel(const double* __restrict__ A, const double* __restrict__ d0,
                                                  int n, int K0, int p, double* C, double* R, double* Tg,
                                                  uint8_t* swept) {
  __shared__ double L[kGJPanel][kGJPanel];
  __shared__ double rowk[kGJPanel], colk[kGJPanel];
  __shared__ uint8_t ok[kGJPanel];
  const int t = threadIdx.x;
  for (int e = t; e < kGJPanel * kGJPanel; e += 256) {
    const int ti = e / kGJPanel, tj = e % kGJPanel;
    L[ti][tj] = (ti < p && tj < p) ? A[(size_t)(K0 + ti) * n + K0 + tj] : 0.0;
  }
  if (t < kGJPanel) ok[t] = 0;
  for (int k = 0; k < p; ++k) {
    __syncthreads();
    const double piv = L[k][k];
    if (!(fabs(piv) > 1e-10 * fabs(d0[K0 + k]))) continue;  // vanished pivot (same in every thread)
    if (t < p) {
      rowk[t] = ((t == k) ? 1.0 : L[k][t]) / piv;  // new row k
      colk[t] = L[t][k];                            // multipliers
    }
    if (t == 0) ok[k] = 1;
    __syncthreads();
    for (int e = t; e < p * kGJPanel; e += 256) {
      const int ti = e / kGJPanel, tj = e % kGJPanel;
      if (tj >= p) continue;
      if (ti == k) L[ti][tj] = rowk[tj];
      else L[ti][tj] = ((tj == k) ? 0.0 : L[ti][tj]) - colk[ti] * rowk[tj];
    }
  }
  __syncthreads();
  for (int e = t; e < p * kGJPanel; e += 256) {
    const int ti = e / kGJPanel, tj = e % kGJPanel;
    if (tj < p && !(ok[ti] && ok[tj])) L[ti][tj] = 0.0;
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    for (int e = t; e < p * kGJPanel; e += 256) {
      const int ti = e / kGJPanel, tj = e % kGJPanel;
      if (tj < p) Tg[ti * kGJPanel + tj] = L[ti][tj];
    }
    if (t < p) swept[K0 + t] = ok[t];
  }
  for (int64_t e = (int64_t)blockIdx.x * 256 + t; e < (int64_t)n * p; e += (int64_t)gridDim.x * 256) {
    const int i = (int)(e / p), a = (int)(e % p);
    double c = 0.0;
    for (int q = 0; q < p; ++q) c += A[(size_t)i * n + K0 + q] * L[q][a];
    C[(size_t)i * kGJPanel + a] = c;
    R[(size_t)a * n + i] = A[(size_t)(K0 + a) * n + i];  // p rows x n columns, same index space
  }
}

// Part 2: the panel update of every entry from C, R and T only (the panel's
// vanished pivots count as rows and columns outside it).  A thread owns one
// column j of kGJRows consecutive rows and keeps that column's p entries of R
// in registers; each entry is formed by the same operations in the same order
// as a one-entry-per-thread update.
constexpr int kGJRows = 4;
__global__ __launch_bounds__(256) void k_gj_update(double* A, int n, int K0, int p,
                                                   const double* __restrict__ C,
                                                   const double* __restrict__ R,
                                                   const double* __restrict__ Tg,
                                                   const uint8_t* __restrict__ swept) {
  const int j = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (j >= n) return;
  const bool kj = j >= K0 && j < K0 + p && swept[j];
  double r[kGJPanel];
#pragma unroll
  for (int a = 0; a < kGJPanel; ++a) r[a] = a < p ? R[(size_t)a * n + j] : 0.0;
  const int i0 = (int)blockIdx.y * kGJRows;
#pragma unroll
  for (int ii = 0; ii < kGJRows; ++ii) {
    const int i = i0 + ii;
    if (i >= n) break;
    const bool ki = i >= K0 && i < K0 + p && swept[i];
    const size_t e = (size_t)i * n + j;
    double v;
    if (ki && kj) {
      v = Tg[(i - K0) * kGJPanel + (j - K0)];
    } else if (ki) {
      v = 0.0;
#pragma unroll
      for (int a = 0; a < kGJPanel; ++a)
        if (a < p) v += Tg[(i - K0) * kGJPanel + a] * r[a];
    } else if (kj) {
      v = -C[(size_t)i * kGJPanel + (j - K0)];
    } else {
      v = A[e];
#pragma unroll
      for (int a = 0; a < kGJPanel; ++a)
        if (a < p) v -= C[(size_t)i * kGJPanel + a] * r[a];
    }
    A[e] = v;
  }
}

template <bool L0>
__global__ __launch_bounds__(256) void k_dense_store(FineOp L, const double* __restrict__ A,
                                                     const uint8_t* __restrict__ swept, float* cinv) {
  const int n = L.nx * L.ny;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int i = (int)(e / n), j = (int)(e % n);
  const bool on = fine_active<L0>(L, i) && fine_active<L0>(L, j) && swept[i] && swept[j];
  cinv[(size_t)j * n + i] = on ? (float)A[e] : 0.0f;  // column-major
}

static void* dalloc(bemcs_poisson_s* h, size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  h->allocs.push_back(p);
  return p;
}

// ------------------------------------------------------- shared operators
// The operator of a mesh -- node kinds and Neumann mirrors, level-0 codes,
// the coarse stencils, every level's inverse diagonal, the dense inverse of
// the coarsest level -- depends only on the mesh shape and the electrode map,
// not on voltages or charge.  Sweep points of one geometry (the reference
// sweeps beam current or voltages at fixed grids,
// Python/benchmarks/benchmark_perveance.py:11-75) therefore share one set of
// operator arrays: a handle created for an electrode map that an existing
// handle was built from points at that handle's arrays instead of building
// its own (the set-up is deterministic, so the arrays would be bit-identical).
// Batched launches then read one operator per geometry, from the caches.  A
// handle whose electrode map changes (bemcs_poisson_update) first moves to
// arrays of its own when it shares them.
struct OpData {
  int nx, ny, per, nlev, nc;
  std::vector<uint8_t> ib;  // electrode map the arrays were built from (keyed ops only)
  bool keyed;               // in the registry: later handles may share it
  int refs;
  std::vector<void*> allocs;
  float* inv_d[kMaxLevels];
  uint8_t* code0;
  float* st[kMaxLevels];
  float* cinv;
  uint8_t* kind;
  int32_t *mirror, *mirror1;
  double* su_st[2];
  uint8_t* su_act[2];
  double *su_A, *su_C, *su_R, *su_T, *su_d0;
  uint8_t* su_swept;
  int* su_err;
};

static std::mutex g_op_mu;
static std::vector<OpData*> g_ops;

static void free_op(OpData* op) {
  for (void* p : op->allocs) (void)hipFree(p);
  delete op;
}

// Operator arrays (and set-up workspace) for the level shapes of h.
static OpData* new_op(const bemcs_poisson_s* h) {
  OpData* op = new OpData();
  op->nx = h->nx;
  op->ny = h->ny;
  op->per = h->periodic;
  op->nlev = h->nlev;
  op->nc = h->nc;
  op->keyed = false;
  op->refs = 1;
  bool ok = true;
  auto A = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (ok && hipMalloc(&p, bytes) == hipSuccess) {
      op->allocs.push_back(p);
      return p;
    }
    ok = false;
    return nullptr;
  };
  const int n = h->nx * h->ny, nc = h->nc;
  for (int l = 0; l < h->nlev; ++l) {
    const int nl = h->lev[l].n;
    op->inv_d[l] = (float*)A(sizeof(float) * nl);
    op->st[l] = l == 0 ? nullptr : (float*)A(sizeof(float) * 9 * nl);
  }
  op->code0 = (uint8_t*)A(n);
  op->cinv = (float*)A(sizeof(float) * (size_t)nc * nc);
  op->kind = (uint8_t*)A(n);
  op->mirror = (int32_t*)A(sizeof(int32_t) * n);
  op->mirror1 = (int32_t*)A(sizeof(int32_t) * n);
  const int n1 = h->nlev > 1 ? h->lev[1].n : 1;
  op->su_st[0] = (double*)A(sizeof(double) * 9 * n1);
  op->su_st[1] = (double*)A(sizeof(double) * 9 * n1);
  op->su_act[0] = (uint8_t*)A(n1);
  op->su_act[1] = (uint8_t*)A(n1);
  op->su_A = (double*)A(sizeof(double) * (size_t)nc * nc);
  op->su_C = (double*)A(sizeof(double) * (size_t)nc * kGJPanel);
  op->su_R = (double*)A(sizeof(double) * (size_t)nc * kGJPanel);
  op->su_T = (double*)A(sizeof(double) * kGJPanel * kGJPanel);
  op->su_d0 = (double*)A(sizeof(double) * nc);
  op->su_swept = (uint8_t*)A(nc);
  op->su_err = (int*)A(sizeof(int));
  if (!ok) {
    free_op(op);
    return nullptr;
  }
  return op;
}

// Points the handle's operator pointers at op's arrays.
static void attach(bemcs_poisson_s* h, OpData* op) {
  h->op = op;
  for (int l = 0; l < h->nlev; ++l) {
    h->lev[l].inv_d = op->inv_d[l];
    h->lev[l].st = op->st[l];
    h->lev[l].code = l == 0 ? op->code0 : nullptr;
  }
  h->cinv = op->cinv;
  h->kind = op->kind;
  h->mirror = op->mirror;
  h->mirror1 = op->mirror1;
  for (int k = 0; k < 2; ++k) {
    h->su_st[k] = op->su_st[k];
    h->su_act[k] = op->su_act[k];
  }
  h->su_A = op->su_A;
  h->su_C = op->su_C;
  h->su_R = op->su_R;
  h->su_T = op->su_T;
  h->su_d0 = op->su_d0;
  h->su_swept = op->su_swept;
  h->su_err = op->su_err;
}

// A keyed operator built from electrode map ib for h's shapes (one more
// reference), or null.
static OpData* find_op(const bemcs_poisson_s* h, const uint8_t* ib) {
  const size_t n = (size_t)h->nx * h->ny;
  std::lock_guard<std::mutex> lock(g_op_mu);
  for (OpData* op : g_ops) {
    if (op->keyed && op->nx == h->nx && op->ny == h->ny && op->per == h->periodic && op->nlev == h->nlev &&
        op->nc == h->nc && memcmp(op->ib.data(), ib, n) == 0) {
      ++op->refs;
      return op;
    }
  }
  return nullptr;
}

static void register_op(OpData* op, const uint8_t* ib) {
  op->ib.assign(ib, ib + (size_t)op->nx * op->ny);
  std::lock_guard<std::mutex> lock(g_op_mu);
  op->keyed = true;
  g_ops.push_back(op);
}

// True when op has no other handle: it then leaves the registry (its arrays
// are about to change) in the same critical section, so no handle created
// meanwhile can start sharing it.
static bool claim_private(OpData* op) {
  std::lock_guard<std::mutex> lock(g_op_mu);
  if (op->refs > 1) return false;
  if (op->keyed) {
    op->keyed = false;
    op->ib.clear();
    g_ops.erase(std::remove(g_ops.begin(), g_ops.end(), op), g_ops.end());
  }
  return true;
}

static void release_op(OpData* op) {
  if (!op) return;
  {
    std::lock_guard<std::mutex> lock(g_op_mu);
    if (--op->refs > 0) return;
    if (op->keyed) g_ops.erase(std::remove(g_ops.begin(), g_ops.end(), op), g_ops.end());
  }
  free_op(op);
}



static FineOp fine_op(const bemcs_poisson_s* h, int l) {
  FineOp F{};
  F.nx = h->lev[l].nx;
  F.ny = h->lev[l].ny;
  if (l == 0) {
    F.kind = h->kind;
    F.code = h->lev[0].code;
  } else {
    F.st = h->su_st[(l - 1) & 1];
    F.act = h->su_act[(l - 1) & 1];
  }
  return F;
}

// Builds the operator of electrode map is_bound (device) into h; asynchronous
// on s; validation failures are OR-ed into h->su_err.
static int setup_operator(bemcs_poisson_s* h, const uint8_t* is_bound, hipStream_t s) {
  const int nx = h->nx, ny = h->ny, n = nx * ny, per = h->periodic;
  BEMCS_CHECK_HIP(hipMemsetAsync(h->su_err, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_setup_kind, dim3(grid_for(n, 256)), dim3(256), 0, s, is_bound, h->kind, nx, ny, per);
  hipLaunchKernelGGL(k_setup_mirror, dim3(grid_for(n, 256)), dim3(256), 0, s, h->kind, h->mirror,
                     h->mirror1, nx, ny, h->su_err);
  hipLaunchKernelGGL(k_setup_level0, dim3(grid_for(n, 256)), dim3(256), 0, s, h->kind, h->mirror1,
                     const_cast<uint8_t*>(h->lev[0].code), const_cast<float*>(h->lev[0].inv_d), nx, ny,
                     per, h->su_err);
  for (int l = 1; l < h->nlev; ++l) {
    const LevelDev& Cl = h->lev[l];
    FineOp F = fine_op(h, l - 1);
    double* st = h->su_st[(l - 1) & 1];
    uint8_t* act = h->su_act[(l - 1) & 1];
    if (l == 1)
      hipLaunchKernelGGL(k_galerkin<true>, dim3(grid_for(Cl.n, 256)), dim3(256), 0, s, F, st, act, Cl.nx,
                         Cl.ny, per);
    else
      hipLaunchKernelGGL(k_galerkin<false>, dim3(grid_for(Cl.n, 256)), dim3(256), 0, s, F, st, act, Cl.nx,
                         Cl.ny, per);
    hipLaunchKernelGGL(k_galerkin_finish, dim3(grid_for(Cl.n, 256)), dim3(256), 0, s, st, act,
                       const_cast<float*>(Cl.st), const_cast<float*>(Cl.inv_d), Cl.nx, Cl.ny, per);
  }
  const int lc = h->nlev - 1;
  const FineOp Lc = fine_op(h, lc);
  const int nc = h->nc;
  const int64_t nn = (int64_t)nc * nc;
  const unsigned gd = grid_for(nn, 256);
  if (lc == 0) hipLaunchKernelGGL(k_dense_assemble<true>, dim3(gd), dim3(256), 0, s, Lc, h->su_A, h->su_d0, per);
  else hipLaunchKernelGGL(k_dense_assemble<false>, dim3(gd), dim3(256), 0, s, Lc, h->su_A, h->su_d0, per);
  for (int K0 = 0; K0 < nc; K0 += kGJPanel) {
    const int p = std::min(kGJPanel, nc - K0);
    const unsigned gp = std::min<unsigned>(grid_for((int64_t)nc * p, 256), 1024u);
    hipLaunchKernelGGL(k_gj_panel, dim3(gp), dim3(256), 0, s, h->su_A, h->su_d0, nc, K0, p, h->su_C, h->su_R,
                       h->su_T, h->su_swept);
    hipLaunchKernelGGL(k_gj_update, dim3((nc + 255) / 256, (nc + kGJRows - 1) / kGJRows), dim3(256), 0, s,
                       h->su_A, nc, K0, p, h->su_C, h->su_R, h->su_T, h->su_swept);
  }
  if (lc == 0)
    hipLaunchKernelGGL(k_dense_store<true>, dim3(gd), dim3(256), 0, s, Lc, h->su_A, h->su_swept, h->cinv);
  else
    hipLaunchKernelGGL(k_dense_store<false>, dim3(gd), dim3(256), 0, s, Lc, h->su_A, h->su_swept, h->cinv);
  BEMCS_CHECK_LAUNCH();
  h->have_guess = false;  // the next Picard call restarts from V on the new unknowns
  return BEMCS_OK;
}

static int setup_status(bemcs_poisson_s* h, hipStream_t s) {
  int err = 0;
  BEMCS_CHECK_HIP(hipMemcpyAsync(&err, h->su_err, sizeof(int), hipMemcpyDeviceToHost, s));
  BEMCS_CHECK_HIP(hipStreamSynchronize(s));
  if (err & SETUP_ERR_CHAIN) {
    set_error("bemcs_poisson: Neumann chain longer than two nodes");
    return BEMCS_EINVAL;
  }
  if (err & SETUP_ERR_MIRROR) {
    set_error("bemcs_poisson: Neumann neighbour mirrors another node");
    return BEMCS_EINVAL;
  }
  if (err & SETUP_ERR_EDGE) {
    set_error("bemcs_poisson: unknown node on the domain edge (column 0 must be Dirichlet)");
    return BEMCS_EINVAL;
  }
  return BEMCS_OK;
}

static int upload(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}

static void destroy_set(bemcs_poisson_batch_s* S);

static void destroy(bemcs_poisson_s* h) {
  if (!h) return;
  destroy_set(h->own);
  for (void* p : h->allocs) (void)hipFree(p);
  release_op(h->op);
  delete h;
}

// ------------------------------------------------------------- solve sets
// A solve set is one or more cases whose PCG iterations advance together: a
// launch of one V-cycle position serves every case that has that level
// (blockIdx.z indexes the case's argument pack), so a batch of uneven meshes
// costs one launch chain per iteration.  Each handle owns a set of one for its
// own solves; bemcs_poisson_batch_create builds sets over several handles.
// Per case, the set owns the PCG scalars and the stop flag; vectors and
// partial buffers belong to the handle (a handle must not be used by two
// sets at the same time).
enum OpKind { OP_INIT, OP_INIT_SCALAR, OP_DIR, OP_STEP_PLAIN, OP_DOWN, OP_DOWN2, OP_COARSE, OP_UP, OP_DOTZ,
              OP_SUMQ };

struct Op {
  OpKind kind;
  int level;
  int tile;   // coarse tile edge of a down / up launch
  dim3 grid;  // grid.z = number of cases in the op
  void* packs_dev;
  // the single-case path passes the (only) pack by value
  CGArgs cg;
  DownPack down;
  Down2Pack down2;
  UpPack up;
  CoarsePack coarse;
};

}  // namespace bemcs

struct bemcs_poisson_batch_s {
  std::vector<bemcs_poisson_s*> cases;
  std::vector<int> gens;  // each handle's op_gen when the set was made
  int n;
  int post_sweeps;
  bool use_graphs;
  std::vector<bemcs::CGArgs> cg;  // per case
  void* cg_dev;                   // the same on the device
  std::vector<bemcs::Op> init_ops, iter_ops;
  // iter_ops[0, iter_split): the step and every case's convergence test;
  // [iter_split, end): the rest of the V-cycle, which a converged case skips
  size_t iter_split;
  std::vector<void*> allocs;
  double* scal;        // [n][S_COUNT]
  int32_t* done;       // [n]
  double* host_scal;   // fine-grained pinned copy of scal, written by k_status
  double* host_scal_dev;  // its device address
  int* host_flag;      // fine-grained pinned: k_status launches finished since the host cleared it
  int* host_flag_dev;
  bemcs::PicardPack* pic_dev;
  bemcs::PicardPack* pic_host;  // pinned staging
  hipStream_t cap_stream;
  std::vector<hipGraphExec_t> iter_graph;
  // PCG iterations launched with a solve's first graph: the count the same
  // solve needed last time, keyed by the Picard call's length and the solve's
  // index in it (within a call the counts follow a stable pattern, e.g. 5 2 2
  // 2 2: the first solve sees the new charge, the next ones only the
  // relaxation), so a solve usually needs no extra launch and host round trip
  std::map<int, std::vector<int>> pred;
  int tile0;  // coarse tile edge of the level-0 launches (kTC or kTCL)
  int tilec;  // coarse tile edge forced on the coarser levels (kTCc or kTCs), 0: per level
  bool fuse_down2;  // the last two coarse down launches as one (opt-in: slower at 1024x384)
  double extrap;  // theta of the Picard initial-guess extrapolation (0: off)
  int pred0;      // iterations launched with a solve position's first graph
  // one case with three or more levels: a graph's last launch (the level-1
  // down launch with the convergence test) writes the status itself
  bool status_in_test;
  std::vector<hipGraphExec_t> pcg_graph;  // init + iterations 1..N in one graph, per N
  std::vector<double> pcg_rtol;
};

namespace bemcs {

using SolveSet = bemcs_poisson_batch_s;

static void destroy_set(SolveSet* S) {
  if (!S) return;
  for (hipGraphExec_t e : S->iter_graph)
    if (e) (void)hipGraphExecDestroy(e);
  for (hipGraphExec_t e : S->pcg_graph)
    if (e) (void)hipGraphExecDestroy(e);
  if (S->cap_stream) (void)hipStreamDestroy(S->cap_stream);
  for (void* p : S->allocs) (void)hipFree(p);
  if (S->host_scal) (void)hipHostFree(S->host_scal);
  if (S->host_flag) (void)hipHostFree(S->host_flag);
  if (S->pic_host) (void)hipHostFree(S->pic_host);
  delete S;
}

static CGArgs cg_args_of(const bemcs_poisson_s* h, double* scal, int32_t* done, int tile0) {
  CGArgs g;
  g.x = h->x;
  g.R[0] = h->R[0];
  g.R[1] = h->R[1];
  g.p = h->p;
  g.Ap = h->Ap;
  g.P[0] = h->p;
  g.P[1] = h->p1;
  g.AP[0] = h->Ap;
  g.AP[1] = h->Ap1;
  g.Az = h->Az;
  g.bred = h->bred;
  g.b0 = h->lev[0].b;
  g.z = h->nlev == 1 ? h->lev[0].x : h->lev[0].t;
  g.code = h->lev[0].code;
  g.nx = h->nx;
  g.ny = h->ny;
  g.per = h->periodic;
  g.n = h->nx * h->ny;
  const int tf = 2 * tile0;
  g.nparts = ((h->nx + tf - 1) / tf) * ((h->ny + tf - 1) / tf);
  g.pa = h->pa;
  g.pr = h->pr;
  g.pz[0] = h->pz[0];
  g.pz[1] = h->pz[1];
  g.pi = h->pi;
  g.scal = scal;
  g.done = done;
  return g;
}

static unsigned cdiv(int a, int b) { return (unsigned)((a + b - 1) / b); }

// Builds the ops of one V-cycle (k = 0: the initial one; iterations add
// OP_DIR in front) with their device packs.
// Pack bytes to upload, and the op whose packs_dev receives the address.
struct PackUpload {
  std::vector<uint8_t> bytes;
  std::vector<Op>* ops;
  size_t index;
};

// Coarse tile edge of the launches between level l and l+1 (l >= 1): 8x8
// coarse tiles while the launch has at least 1024 of them (four per CU),
// 4x4 below, where each launch is a chain of latency-bound phases and
// smaller tiles take fewer rounds per phase.
static int coarse_tile(const SolveSet* S, int l) {
  if (S->tilec) return S->tilec;
  // batches: the 8 x 8 tiles on every coarse level (measured +2% on a
  // 64-case sweep); the tile edge changes no result, only the decomposition
  if (S->n > 1) return kTCc;
  long tiles = 0;
  for (const bemcs_poisson_s* h : S->cases)
    if (h->nlev - 1 > l) tiles += (long)cdiv(h->lev[l + 1].nx, kTCc) * cdiv(h->lev[l + 1].ny, kTCc);
  return tiles >= 1024 ? kTCc : kTCs;
}

// Batched tile launches: grid columns rounded up to a multiple of the XCD
// count, so every case's z-slice starts on XCD 0 (xcd_case_tile).
static unsigned round_up8(unsigned v) { return (v + kXcds - 1) / kXcds * kXcds; }

static bool build_ops(SolveSet* S, bool iteration, std::vector<Op>& ops,
                      std::vector<PackUpload>& uploads) {
  const int n = S->n;
  int lmax = 0, nparts_max = 0;
  for (int c = 0; c < n; ++c) {
    lmax = std::max(lmax, S->cases[c]->nlev);
    nparts_max = std::max(nparts_max, S->cg[c].nparts);
  }
  auto add_cg = [&](OpKind kind, std::vector<int> idx, dim3 grid) {
    if (idx.empty()) return;
    Op op{};
    op.kind = kind;
    op.grid = dim3(grid.x, grid.y, (unsigned)idx.size());
    std::vector<uint8_t> bytes(sizeof(CGArgs) * idx.size());
    for (size_t j = 0; j < idx.size(); ++j) memcpy(&bytes[j * sizeof(CGArgs)], &S->cg[idx[j]], sizeof(CGArgs));
    op.cg = S->cg[idx[0]];
    ops.push_back(op);
    uploads.push_back({bytes, &ops, ops.size() - 1});
  };
  std::vector<int> all(n), plain;
  for (int c = 0; c < n; ++c) {
    all[c] = c;
    if (S->cases[c]->nlev == 1) plain.push_back(c);
  }
  if (!iteration) {
    // the caller has launched the r0 kernel; multi-level cases take their
    // initial scalars in the level-0 down launch
    add_cg(OP_INIT_SCALAR, plain, dim3(1, 1));
  } else {
    // with small level-0 tiles, multi-level cases form the direction inside
    // the level-0 down launch
    add_cg(OP_DIR, S->tile0 == kTC ? plain : all, dim3(nparts_max, 1));
    if (S->tile0 == kTC && n > 1) {
      std::vector<int> multi;
      for (int c = 0; c < n; ++c)
        if (S->cases[c]->nlev > 1) multi.push_back(c);
      add_cg(OP_SUMQ, multi, dim3(1, 1));
    }
    add_cg(OP_STEP_PLAIN, plain, dim3(nparts_max, 1));
  }
  // the last two down launches as one (every case with the same depth,
  // non-periodic, the first of them below the level that runs the
  // convergence test)
  bool fuse2 = S->fuse_down2 && lmax >= 5;
  for (int c = 0; c < n; ++c)
    fuse2 = fuse2 && S->cases[c]->nlev == lmax && !S->cases[c]->periodic;
  for (int l = 0; l + 1 < lmax; ++l) {
    if (fuse2 && l == lmax - 3) {
      std::vector<Down2Pack> packs;
      unsigned gx = 1, gy = 1;
      for (int c = 0; c < n; ++c) {
        bemcs_poisson_s* h = S->cases[c];
        Down2Pack P;
        P.F = h->lev[l];
        P.M = h->lev[l + 1];
        P.C = h->lev[l + 2];
        P.done = S->cg[c].done;
        gx = std::max(gx, cdiv(P.C.nx, kT2));
        gy = std::max(gy, cdiv(P.C.ny, kT2));
        packs.push_back(P);
      }
      Op op{};
      op.kind = OP_DOWN2;
      op.level = l;
      op.grid = dim3(n > 1 ? round_up8(gx) : gx, gy, (unsigned)packs.size());
      op.down2 = packs[0];
      std::vector<uint8_t> bytes(sizeof(Down2Pack) * packs.size());
      memcpy(bytes.data(), packs.data(), bytes.size());
      ops.push_back(op);
      uploads.push_back({bytes, &ops, ops.size() - 1});
      ++l;  // the next level's down is part of this launch
      continue;
    }
    std::vector<DownPack> packs;
    unsigned gx = 1, gy = 1;
    const int tc = l == 0 ? S->tile0 : coarse_tile(S, l);
    for (int c = 0; c < n; ++c) {
      bemcs_poisson_s* h = S->cases[c];
      if (h->nlev - 1 <= l) continue;
      DownPack P;
      P.F = h->lev[l];
      P.C = h->lev[l + 1];
      P.g = S->cg[c];
      P.ca = CheckArgs{h->pr, S->cg[c].nparts, l == 1 ? 1 : 0, S->cg[c].scal, S->cg[c].done};
      gx = std::max(gx, cdiv(P.C.nx, tc));
      gy = std::max(gy, cdiv(P.C.ny, tc));
      packs.push_back(P);
    }
    if (packs.empty()) continue;
    // one more tile row for the work-group that computes the launch's global
    // scalars (initial tolerance, or the convergence test)
    if ((!iteration && l == 0) || (iteration && l == 1)) gy += 1;
    Op op{};
    op.kind = OP_DOWN;
    op.level = l;
    op.tile = tc;
    op.grid = dim3(n > 1 ? round_up8(gx) : gx, gy, (unsigned)packs.size());
    op.down = packs[0];
    std::vector<uint8_t> bytes(sizeof(DownPack) * packs.size());
    memcpy(bytes.data(), packs.data(), bytes.size());
    ops.push_back(op);
    uploads.push_back({bytes, &ops, ops.size() - 1});
  }
  {
    std::vector<CoarsePack> packs;
    unsigned gx = 1;
    for (int c = 0; c < n; ++c) {
      bemcs_poisson_s* h = S->cases[c];
      CoarsePack P;
      P.C = h->lev[h->nlev - 1];
      P.cinv = h->cinv;
      P.ca = CheckArgs{h->pr, S->cg[c].nparts, h->nlev <= 2 ? 1 : 0, S->cg[c].scal, S->cg[c].done};
      gx = std::max(gx, cdiv(P.C.n, kCoarseRowsPerBlock));
      packs.push_back(P);
    }
    Op op{};
    op.kind = OP_COARSE;
    op.grid = dim3(gx, 1, (unsigned)packs.size());
    op.coarse = packs[0];
    std::vector<uint8_t> bytes(sizeof(CoarsePack) * packs.size());
    memcpy(bytes.data(), packs.data(), bytes.size());
    ops.push_back(op);
    uploads.push_back({bytes, &ops, ops.size() - 1});
  }
  for (int l = lmax - 2; l >= 0; --l) {
    std::vector<UpPack> packs;
    unsigned gx = 1, gy = 1;
    const int tf = l == 0 ? 2 * S->tile0 : 2 * coarse_tile(S, l);
    for (int c = 0; c < n; ++c) {
      bemcs_poisson_s* h = S->cases[c];
      if (h->nlev - 1 <= l) continue;
      UpPack P;
      P.F = h->lev[l];
      P.C = h->lev[l + 1];
      if (l + 1 != h->nlev - 1) P.C.x = P.C.t;  // the coarse level's result
      P.g = S->cg[c];
      gx = std::max(gx, cdiv(P.F.nx, tf));
      gy = std::max(gy, cdiv(P.F.ny, tf));
      packs.push_back(P);
    }
    if (packs.empty()) continue;
    Op op{};
    op.kind = OP_UP;
    op.level = l;
    op.tile = tf / 2;
    op.grid = dim3(n > 1 ? round_up8(gx) : gx, gy, (unsigned)packs.size());
    op.up = packs[0];
    std::vector<uint8_t> bytes(sizeof(UpPack) * packs.size());
    memcpy(bytes.data(), packs.data(), bytes.size());
    ops.push_back(op);
    uploads.push_back({bytes, &ops, ops.size() - 1});
  }
  add_cg(OP_DOTZ, plain, dim3(nparts_max, 1));
  return true;
}

// `status`: this launch ends a graph of a set that writes its status from
// the test work-group (set->status_in_test).
static void launch_op(SolveSet* S, const Op& op, int k, double rtol, hipStream_t s, bool status = false) {
  const bool B = S->n > 1;
  const dim3 blk(256);
  const CGArgs* cgm = (const CGArgs*)op.packs_dev;
  const bool two = S->post_sweeps > 1;
  switch (op.kind) {
    case OP_INIT:
      if (B) hipLaunchKernelGGL(k_cg_init<true>, op.grid, blk, 0, s, op.cg, cgm);
      else hipLaunchKernelGGL(k_cg_init<false>, op.grid, blk, 0, s, op.cg, cgm);
      break;
    case OP_SUMQ:
      if (B) hipLaunchKernelGGL(k_sum_quads<true>, op.grid, blk, 0, s, op.cg, cgm, k);
      else hipLaunchKernelGGL(k_sum_quads<false>, op.grid, blk, 0, s, op.cg, cgm, k);
      break;
    case OP_INIT_SCALAR:
      if (B) hipLaunchKernelGGL(k_cg_init_scalar<true>, op.grid, blk, 0, s, op.cg, cgm, rtol);
      else hipLaunchKernelGGL(k_cg_init_scalar<false>, op.grid, blk, 0, s, op.cg, cgm, rtol);
      break;
    case OP_DIR:
      if (B) hipLaunchKernelGGL(k_cg_dir<true>, op.grid, blk, 0, s, op.cg, cgm, k);
      else hipLaunchKernelGGL(k_cg_dir<false>, op.grid, blk, 0, s, op.cg, cgm, k);
      break;
    case OP_STEP_PLAIN:
      if (B) hipLaunchKernelGGL(k_cg_step<true>, op.grid, blk, 0, s, op.cg, cgm, k);
      else hipLaunchKernelGGL(k_cg_step<false>, op.grid, blk, 0, s, op.cg, cgm, k);
      break;
    case OP_DOTZ:
      if (B) hipLaunchKernelGGL(k_cg_dotz<true>, op.grid, blk, 0, s, op.cg, cgm, k);
      else hipLaunchKernelGGL(k_cg_dotz<false>, op.grid, blk, 0, s, op.cg, cgm, k);
      break;
    case OP_DOWN2: {
      const Down2Pack* pm = (const Down2Pack*)op.packs_dev;
      if (B) hipLaunchKernelGGL(k_down2<true>, op.grid, blk, 0, s, op.down2, pm);
      else hipLaunchKernelGGL(k_down2<false>, op.grid, blk, 0, s, op.down2, pm);
      break;
    }
    case OP_DOWN: {
      const DownPack* pm = (const DownPack*)op.packs_dev;
      DownPack d = op.down;  // the one-case pack goes by value
      if (status) {
        d.ca.host = S->host_scal_dev;
        d.ca.flag = S->host_flag_dev;
      }
      if (op.level == 0 && S->tile0 == kTCL) {
        if (k > 0) {
          if (B) hipLaunchKernelGGL((k_down<kTCL, false, true, true>), op.grid, blk, 0, s, d, pm, k, rtol);
          else hipLaunchKernelGGL((k_down<kTCL, false, true, false>), op.grid, blk, 0, s, d, pm, k, rtol);
        } else {
          if (B) hipLaunchKernelGGL((k_down<kTCL, false, false, true>), op.grid, blk, 0, s, d, pm, k, rtol);
          else hipLaunchKernelGGL((k_down<kTCL, false, false, false>), op.grid, blk, 0, s, d, pm, k, rtol);
        }
      } else if (op.level == 0) {
        if (k > 0) {
          if (B) hipLaunchKernelGGL((k_down<kTC, false, true, true>), op.grid, blk, 0, s, d, pm, k, rtol);
          else hipLaunchKernelGGL((k_down<kTC, false, true, false>), op.grid, blk, 0, s, d, pm, k, rtol);
        } else {
          if (B) hipLaunchKernelGGL((k_down<kTC, false, false, true>), op.grid, blk, 0, s, d, pm, k, rtol);
          else hipLaunchKernelGGL((k_down<kTC, false, false, false>), op.grid, blk, 0, s, d, pm, k, rtol);
        }
      } else if (op.tile == kTCs) {
        if (B) hipLaunchKernelGGL((k_down<kTCs, true, false, true>), op.grid, blk, 0, s, d, pm, k, rtol);
        else hipLaunchKernelGGL((k_down<kTCs, true, false, false>), op.grid, blk, 0, s, d, pm, k, rtol);
      } else {
        if (B) hipLaunchKernelGGL((k_down<kTCc, true, false, true>), op.grid, blk, 0, s, d, pm, k, rtol);
        else hipLaunchKernelGGL((k_down<kTCc, true, false, false>), op.grid, blk, 0, s, d, pm, k, rtol);
      }
      break;
    }
    case OP_COARSE: {
      const CoarsePack* pm = (const CoarsePack*)op.packs_dev;
      if (B) hipLaunchKernelGGL(k_coarse_solve<true>, op.grid, dim3(64 * kCoarseRowsPerBlock), 0, s, op.coarse, pm, k);
      else hipLaunchKernelGGL(k_coarse_solve<false>, op.grid, dim3(64 * kCoarseRowsPerBlock), 0, s, op.coarse, pm, k);
      break;
    }
    case OP_UP: {
      const UpPack* pm = (const UpPack*)op.packs_dev;
      if (op.level == 0 && S->tile0 == kTCL) {
        if (two) {
          if (B) hipLaunchKernelGGL((k_up<kTFL, 2, false, true, true>), op.grid, blk, 0, s, op.up, pm, k);
          else hipLaunchKernelGGL((k_up<kTFL, 2, false, true, false>), op.grid, blk, 0, s, op.up, pm, k);
        } else {
          if (B) hipLaunchKernelGGL((k_up<kTFL, 1, false, true, true>), op.grid, blk, 0, s, op.up, pm, k);
          else hipLaunchKernelGGL((k_up<kTFL, 1, false, true, false>), op.grid, blk, 0, s, op.up, pm, k);
        }
      } else if (op.level == 0) {
        if (two) {
          if (B) hipLaunchKernelGGL((k_up<kTF, 2, false, true, true>), op.grid, blk, 0, s, op.up, pm, k);
          else hipLaunchKernelGGL((k_up<kTF, 2, false, true, false>), op.grid, blk, 0, s, op.up, pm, k);
        } else {
          if (B) hipLaunchKernelGGL((k_up<kTF, 1, false, true, true>), op.grid, blk, 0, s, op.up, pm, k);
          else hipLaunchKernelGGL((k_up<kTF, 1, false, true, false>), op.grid, blk, 0, s, op.up, pm, k);
        }
      } else if (op.tile == kTCs) {
        if (two) {
          if (B) hipLaunchKernelGGL((k_up<kTFs, 2, true, false, true>), op.grid, blk, 0, s, op.up, pm, k);
          else hipLaunchKernelGGL((k_up<kTFs, 2, true, false, false>), op.grid, blk, 0, s, op.up, pm, k);
        } else {
          if (B) hipLaunchKernelGGL((k_up<kTFs, 1, true, false, true>), op.grid, blk, 0, s, op.up, pm, k);
          else hipLaunchKernelGGL((k_up<kTFs, 1, true, false, false>), op.grid, blk, 0, s, op.up, pm, k);
        }
      } else {
        if (two) {
          if (B) hipLaunchKernelGGL((k_up<kTFc, 2, true, false, true>), op.grid, blk, 0, s, op.up, pm, k);
          else hipLaunchKernelGGL((k_up<kTFc, 2, true, false, false>), op.grid, blk, 0, s, op.up, pm, k);
        } else {
          if (B) hipLaunchKernelGGL((k_up<kTFc, 1, true, false, true>), op.grid, blk, 0, s, op.up, pm, k);
          else hipLaunchKernelGGL((k_up<kTFc, 1, true, false, false>), op.grid, blk, 0, s, op.up, pm, k);
        }
      }
      break;
    }
  }
}

static int create_set(bemcs_poisson_s* const* hs, int n, SolveSet** out) {
  *out = nullptr;
  SolveSet* S = new SolveSet();
  S->n = n;
  S->cases.assign(hs, hs + n);
  for (int c = 0; c < n; ++c) S->gens.push_back(hs[c]->op_gen);
  const char* ps = getenv("BEMCS_MG_POST_SWEEPS");
  S->post_sweeps = (ps && ps[0] == '1') ? 1 : 2;
  const char* pg = getenv("BEMCS_POISSON_GRAPHS");
  S->use_graphs = !(pg && pg[0] == '0');
  // a handle asks for the large level-0 tiles explicitly or by its mesh size;
  // the set takes them if any of its handles does
  S->tile0 = kTC;
  for (int c = 0; c < n; ++c) {
    const int pref = hs[c]->tile0_pref;
    if (pref == kTCL || (pref == 0 && hs[c]->nx * hs[c]->ny > kLargeMesh)) S->tile0 = kTCL;
  }
  const char* tl = getenv("BEMCS_MG_TILE0");
  if (tl && (atoi(tl) == kTC || atoi(tl) == kTCL)) S->tile0 = atoi(tl);
  const char* ex = getenv("BEMCS_PICARD_EXTRAP");
  S->extrap = ex ? atof(ex) : kPicardExtrap;
  const char* pp = getenv("BEMCS_PCG_PREDICT");
  S->pred0 = pp && atoi(pp) >= 1 ? atoi(pp) : 8;
  const char* f2 = getenv("BEMCS_MG_FUSE");
  S->fuse_down2 = f2 && f2[0] == '1';
  S->tilec = 0;
  const char* tcs = getenv("BEMCS_MG_TILEC");
  if (tcs && (atoi(tcs) == kTCc || atoi(tcs) == kTCs)) S->tilec = atoi(tcs);
  bool ok = true;
  auto A = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (ok && hipMalloc(&p, bytes) == hipSuccess) {
      S->allocs.push_back(p);
      return p;
    }
    ok = false;
    return nullptr;
  };
  S->scal = (double*)A(sizeof(double) * S_COUNT * n);
  S->done = (int32_t*)A(sizeof(int32_t) * n);
  S->pic_dev = (PicardPack*)A(sizeof(PicardPack) * n);
  if (ok) ok = hipMemset(S->scal, 0, sizeof(double) * S_COUNT * n) == hipSuccess;
  if (ok) ok = hipMemset(S->done, 0, sizeof(int32_t) * n) == hipSuccess;
  if (ok)
    ok = hipHostMalloc((void**)&S->host_scal, sizeof(double) * S_COUNT * n, hipHostMallocCoherent) == hipSuccess;
  if (ok) ok = hipHostGetDevicePointer((void**)&S->host_scal_dev, S->host_scal, 0) == hipSuccess;
  if (ok) ok = hipHostMalloc((void**)&S->host_flag, 64, hipHostMallocCoherent) == hipSuccess;
  if (ok) ok = hipHostGetDevicePointer((void**)&S->host_flag_dev, S->host_flag, 0) == hipSuccess;
  if (ok) ok = hipHostMalloc((void**)&S->pic_host, sizeof(PicardPack) * n, 0) == hipSuccess;
  if (ok) {
    for (int c = 0; c < n; ++c)
    {
      S->cg.push_back(cg_args_of(hs[c], S->scal + S_COUNT * c, S->done + c, S->tile0));
      // a batch adds each case's level-0 partial sums once per iteration
      // (the same fixed-order sum as a case alone adds in every work-group)
      S->cg.back().presum = n > 1 && S->tile0 == kTC ? 1 : 0;
    }
    std::vector<PackUpload> uploads;
    build_ops(S, false, S->init_ops, uploads);
    build_ops(S, true, S->iter_ops, uploads);
    // cases with three or more levels test convergence in the level-1 down
    // launch, cases with one or two in the coarsest solve
    bool coarse_test = false;
    for (int c = 0; c < n; ++c) coarse_test = coarse_test || hs[c]->nlev <= 2;
    S->status_in_test = n == 1 && hs[0]->nlev >= 3;
    S->iter_split = S->iter_ops.size();
    for (size_t j = 0; j < S->iter_ops.size(); ++j) {
      const Op& op = S->iter_ops[j];
      if (coarse_test ? op.kind == OP_COARSE : (op.kind == OP_DOWN && op.level == 1)) S->iter_split = j + 1;
    }
    S->cg_dev = A(sizeof(CGArgs) * n);
    if (ok) ok = upload(S->cg_dev, S->cg.data(), sizeof(CGArgs) * n) == 0;
    for (auto& u : uploads) {
      void* d = A(u.bytes.size());
      (*u.ops)[u.index].packs_dev = d;
      if (ok) ok = upload(d, u.bytes.data(), u.bytes.size()) == 0;
    }
  }
  if (!ok) {
    destroy_set(S);
    set_error("Poisson solve set: device allocation or upload failed");
    return BEMCS_EHIP;
  }
  *out = S;
  return BEMCS_OK;
}

// Captures what `enqueue` launches into an executable graph (on the set's
// private capture stream: nothing runs).  Launch arguments are fixed per set
// and iteration number, so a graph is captured once and replayed.
template <typename F>
static hipError_t capture(SolveSet* S, F enqueue, hipGraphExec_t* out) {
  *out = nullptr;
  hipError_t e = hipSuccess;
  if (!S->cap_stream) e = hipStreamCreateWithFlags(&S->cap_stream, hipStreamNonBlocking);
  if (e != hipSuccess) return e;
  e = hipStreamBeginCapture(S->cap_stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  enqueue(S->cap_stream);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(S->cap_stream, &g);
  if (e != hipSuccess) return e;
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  return e;
}

static void launch_status(SolveSet* S, hipStream_t s) {
  hipLaunchKernelGGL(k_status, dim3(1), dim3(256), 0, s, (const double*)S->scal, S->host_scal_dev,
                     S_COUNT * S->n, S->host_flag_dev);
}

// Waits until `graphs` k_status launches (the ends of the graphs enqueued on
// `s` since the flag was cleared) have written host_scal.  Polls the pinned
// counter, and the stream every 1024 polls, so a stream that finishes (or
// fails) without reaching the count still ends the wait.
static int wait_status(SolveSet* S, int graphs, hipStream_t s) {
  for (unsigned spins = 1;; ++spins) {
    if (__atomic_load_n(S->host_flag, __ATOMIC_ACQUIRE) >= graphs) return BEMCS_OK;
    if ((spins & 1023) == 0) {
      const hipError_t q = hipStreamQuery(s);
      if (q == hipErrorNotReady) continue;
      BEMCS_CHECK_HIP(q);
      if (__atomic_load_n(S->host_flag, __ATOMIC_ACQUIRE) >= graphs) return BEMCS_OK;
      set_error("Poisson MG-PCG: stream finished without a status");
      return BEMCS_EHIP;
    }
  }
}

// Launches of iteration k's op range [b, e).
static void launch_iter_ops(SolveSet* S, int k, size_t b, size_t e, hipStream_t s) {
  for (size_t j = b; j < e; ++j) launch_op(S, S->iter_ops[j], k, 0.0, s);
}

// The end of a graph: iteration k up to its convergence test, then the
// status (from the test launch itself when the set allows it).
static void launch_graph_tail(SolveSet* S, int k, hipStream_t s) {
  if (S->status_in_test) {
    launch_iter_ops(S, k, 0, S->iter_split - 1, s);
    launch_op(S, S->iter_ops[S->iter_split - 1], k, 0.0, s, true);
  } else {
    launch_iter_ops(S, k, 0, S->iter_split, s);
    launch_status(S, s);
  }
}

// A graph ends after the launch that runs the convergence test of its last
// iteration (nothing after the test is needed if the solve has converged);
// the next graph, if the solve needs one, first completes that V-cycle.
// Iteration k >= 2 as one graph: the rest of iteration k-1, then iteration k
// up to its test.
static int enqueue_iter(SolveSet* S, int k, hipStream_t s) {
  auto body = [&](hipStream_t cs) {
    launch_iter_ops(S, k - 1, S->iter_split, S->iter_ops.size(), cs);
    launch_graph_tail(S, k, cs);
  };
  if (!S->use_graphs) {
    body(s);
    return BEMCS_OK;
  }
  if ((int)S->iter_graph.size() <= k) S->iter_graph.resize(k + 1, nullptr);
  if (!S->iter_graph[k]) BEMCS_CHECK_HIP(capture(S, body, &S->iter_graph[k]));
  BEMCS_CHECK_HIP(hipGraphLaunch(S->iter_graph[k], s));
  return BEMCS_OK;
}

// The initial V-cycle and iterations 1..N (the last up to its test) as one
// graph (one host launch).
static int enqueue_pcg_head(SolveSet* S, int N, double rtol, hipStream_t s) {
  auto body = [&](hipStream_t cs) {
    for (const Op& op : S->init_ops) launch_op(S, op, 0, rtol, cs);
    for (int k = 1; k < N; ++k) launch_iter_ops(S, k, 0, S->iter_ops.size(), cs);
    launch_graph_tail(S, N, cs);
  };
  if (!S->use_graphs) {
    body(s);
    return BEMCS_OK;
  }
  if ((int)S->pcg_graph.size() <= N) {
    S->pcg_graph.resize(N + 1, nullptr);
    S->pcg_rtol.resize(N + 1, -1.0);
  }
  if (!S->pcg_graph[N] || S->pcg_rtol[N] != rtol) {
    if (S->pcg_graph[N]) (void)hipGraphExecDestroy(S->pcg_graph[N]);
    S->pcg_graph[N] = nullptr;
    BEMCS_CHECK_HIP(capture(S, body, &S->pcg_graph[N]));
    S->pcg_rtol[N] = rtol;
  }
  BEMCS_CHECK_HIP(hipGraphLaunch(S->pcg_graph[N], s));
  return BEMCS_OK;
}

// PCG on the reduced systems already in each case's b_red with initial guess
// x.  Iterations advance for all cases together; converged cases skip.  The
// initial V-cycle and the predicted number of iterations go out as one graph;
// further iterations, if needed, one graph each.  `next`, if given, enqueues
// the launches that follow the solve (guarded by the stop flags) after every
// graph, before the host waits for the graph's status.
static int pcg(SolveSet* S, int call_len, int index, double rtol, int max_iter, int* iters, double* relres,
               hipStream_t s, const std::function<void()>& next = nullptr) {
  std::vector<int>& pred = S->pred[call_len];
  if ((int)pred.size() <= index) pred.resize(index + 1, S->pred0);
  int launched = std::min(std::max(1, pred[index]), max_iter);
  __atomic_store_n(S->host_flag, 0, __ATOMIC_RELEASE);
  int rc = enqueue_pcg_head(S, launched, rtol, s);
  int graphs = 1;
  if (rc) return rc;
  BEMCS_CHECK_LAUNCH();
  int batch = 0;
  const int n = S->n;
  for (;;) {
    for (int j = 0; j < batch && launched < max_iter; ++j) {
      if (j == 0) {
        __atomic_store_n(S->host_flag, 0, __ATOMIC_RELEASE);
        graphs = 0;
      }
      ++graphs;
      rc = enqueue_iter(S, ++launched, s);
      if (rc) return rc;
    }
    if (next) next();
    BEMCS_CHECK_LAUNCH();
    rc = wait_status(S, graphs, s);
    if (rc) return rc;
    bool all = true;
    for (int c = 0; c < n && all; ++c) {
      const double* sc = S->host_scal + S_COUNT * c;
      const double rr = sc[S_RR];
      if (!(rr <= sc[S_TOL2]) && rr == rr) all = false;
    }
    if (all || launched >= max_iter) break;
    batch = 1;
  }
  int worst_it = 0, bad = -1;
  double bad_rel = 0;
  for (int c = 0; c < n; ++c) {
    const double* sc = S->host_scal + S_COUNT * c;
    const double rr = sc[S_RR], bb = sc[S_BB];
    const int it = (int)sc[S_ITERS];
    const double rel = bb > 0 ? sqrt(rr / bb) : sqrt(rr);
    worst_it = std::max(worst_it, it);
    if (iters) iters[c] = it;
    if (relres) relres[c] = rel;
    if (!(rr <= sc[S_TOL2]) && bad < 0) {
      bad = c;
      bad_rel = rel;
    }
  }
  pred[index] = worst_it;
  if (bad >= 0) {
    set_error("Poisson MG-PCG: case %d relative residual %.3e above tolerance %.3e after %d iterations",
              bad, bad_rel, rtol, launched);
    return BEMCS_ENOCONV;
  }
  return BEMCS_OK;
}

static PicardPack picard_pack(bemcs_poisson_s* h, const bemcs_picard_t* p, bool extrap) {
  PicardPack P;
  P.cnt_ion = p->cnt_ion;
  P.cnt_e = p->cnt_e;
  P.w_ion = (const unsigned long long*)p->w_ion;
  P.w_e = (const unsigned long long*)p->w_e;
  P.cd_ion = p->cd_ion;
  P.cd_e = p->cd_e;
  P.rho = p->rho_out;
  P.V_fixed = p->V_fixed;
  P.V = p->V;
  P.Vnew = h->Vnew;
  P.kind = h->kind;
  P.mirror = h->mirror;
  P.bred = h->bred;
  P.x = h->x;
  P.xw = h->x;
  P.xg = h->xg;
  // with extrapolation the solve's x holds the next initial guess when a
  // sweep starts, and the last solution has been saved to xg
  P.xr = extrap ? h->xg : h->x;
  P.nx = h->nx;
  P.ny = h->ny;
  P.per = h->periodic;
  P.n = h->nx * h->ny;
  P.coeff = p->coeff;
  P.Vp = p->V_plasma;
  P.Te = p->Te_up;
  P.qn0 = -p->q * p->n0;
  P.one_m_omega = 1.0 - p->omega;
  P.omega = p->omega;
  P.dx_m = p->dx_m;
  P.dy_m = p->dy_m;
  P.Exy = p->Exy;
  return P;
}

template <bool RELAX>
static void launch_picard_start(SolveSet* S, hipStream_t s, int save = 0) {
  unsigned gx = 1;
  for (int c = 0; c < S->n; ++c) gx = std::max(gx, (unsigned)S->cg[c].nparts);
  const CGArgs* cgm = (const CGArgs*)S->cg_dev;
  if (S->n > 1)
    hipLaunchKernelGGL((k_picard_start<RELAX, true>), dim3(gx, 1, S->n), dim3(256), 0, s, S->pic_host[0],
                       (const PicardPack*)S->pic_dev, S->cg[0], cgm, save);
  else
    hipLaunchKernelGGL((k_picard_start<RELAX, false>), dim3(gx, 1, 1), dim3(256), 0, s, S->pic_host[0],
                       (const PicardPack*)S->pic_dev, S->cg[0], cgm, save);
}

// r0 = b - A x of a solve without a Picard right-hand side
static void launch_cg_init(SolveSet* S, hipStream_t s) {
  unsigned gx = 1;
  for (int c = 0; c < S->n; ++c) gx = std::max(gx, (unsigned)S->cg[c].nparts);
  const CGArgs* cgm = (const CGArgs*)S->cg_dev;
  if (S->n > 1) hipLaunchKernelGGL(k_cg_init<true>, dim3(gx, 1, S->n), dim3(256), 0, s, S->cg[0], cgm);
  else hipLaunchKernelGGL(k_cg_init<false>, dim3(gx, 1, 1), dim3(256), 0, s, S->cg[0], cgm);
}

template <int PHASE>
static void launch_picard(SolveSet* S, unsigned gx, hipStream_t s) {
  if (S->n > 1)
    hipLaunchKernelGGL((k_picard<PHASE, true>), dim3(gx, 1, S->n), dim3(256), 0, s, S->pic_host[0],
                       (const PicardPack*)S->pic_dev);
  else
    hipLaunchKernelGGL((k_picard<PHASE, false>), dim3(gx, 1, 1), dim3(256), 0, s, S->pic_host[0],
                       (const PicardPack*)S->pic_dev);
}

static void launch_picard_extrap(SolveSet* S, unsigned gx, double theta, hipStream_t s) {
  if (S->n > 1)
    hipLaunchKernelGGL(k_picard_extrap<true>, dim3(gx, 1, S->n), dim3(256), 0, s, S->pic_host[0],
                       (const PicardPack*)S->pic_dev, theta);
  else
    hipLaunchKernelGGL(k_picard_extrap<false>, dim3(gx, 1, 1), dim3(256), 0, s, S->pic_host[0],
                       (const PicardPack*)S->pic_dev, theta);
}

// The Picard loop of every case of the set (:886-902), then E = -grad V.
static int picard(SolveSet* S, const bemcs_picard_t* ps, int32_t* iters_total, double* worst_relres,
                  hipStream_t s) {
  const int n = S->n;
  const int iterations = ps[0].iterations;
  const double rtol = ps[0].rtol;
  const int max_iter = ps[0].max_iter;
  unsigned gx = 1;
  for (int c = 0; c < n; ++c) {
    const bemcs_picard_t* p = &ps[c];
    BEMCS_REQUIRE(((p->cnt_ion && p->cnt_e) || (p->w_ion && p->w_e)) && p->V_fixed && p->V &&
                      p->rho_out && p->Exy,
                  "bemcs_poisson_picard: bad arguments");
    BEMCS_REQUIRE(p->iterations == iterations && p->rtol == rtol && p->max_iter == max_iter,
                  "bemcs_poisson_batch_picard: iterations, rtol and max_iter must agree across cases");
    S->pic_host[c] = picard_pack(S->cases[c], p, S->extrap != 0.0);
    S->pic_host[c].done = S->done + c;
    gx = std::max(gx, grid_for(S->pic_host[c].n, 256));
  }
  if (n > 1)
    BEMCS_CHECK_HIP(hipMemcpyAsync(S->pic_dev, S->pic_host, sizeof(PicardPack) * n,
                                   hipMemcpyHostToDevice, s));
  for (int c = 0; c < n; ++c) {
    bemcs_poisson_s* h = S->cases[c];
    if (h->have_guess) continue;
    hipLaunchKernelGGL(k_reduce_rhs, dim3(grid_for(h->nx * h->ny, 256)), dim3(256), 0, s,
                       ps[c].V_fixed, ps[c].V, h->kind, h->bred, h->x, h->nx, h->ny, h->periodic, 1);
    h->have_guess = true;
  }
  BEMCS_CHECK_LAUNCH();
  std::vector<int> it(n, 0), total(n, 0);
  std::vector<double> rel(n, 0.0), worst(n, 0.0);
  int rc = BEMCS_OK;
  // the right-hand side + r0 of the first sweep
  launch_picard_start<false>(S, s);
  BEMCS_CHECK_LAUNCH();
  for (int k = 0; k < iterations && rc == BEMCS_OK; ++k) {
    // what follows solve k, enqueued before its status is read: the initial
    // guess and start of sweep k+1 (relaxation of sweep k, right-hand side,
    // r0), or after the last solve the relaxation and E = -grad V
    auto next = [&]() {
      if (k + 1 < iterations) {
        // the second sweep has one solution to save and none to extrapolate
        // from: its start launch saves it
        const bool save = S->extrap != 0.0 && k + 1 == 1;
        if (S->extrap != 0.0 && !save) launch_picard_extrap(S, gx, S->extrap, s);
        launch_picard_start<true>(S, s, save ? 1 : 0);
      } else {
        launch_picard<PIC_RELAX>(S, gx, s);
        launch_picard<PIC_EFIELD>(S, gx, s);
      }
      if (n > 1) hipLaunchKernelGGL(k_mark_started, dim3(cdiv(n, 256)), dim3(256), 0, s, S->done, n);
    };
    rc = pcg(S, iterations, k, rtol, max_iter, it.data(), rel.data(), s, next);
    for (int c = 0; c < n; ++c) {
      total[c] += it[c];
      worst[c] = std::max(worst[c], rel[c]);
    }
  }
  for (int c = 0; c < n; ++c) {
    if (iters_total) iters_total[c] = total[c];
    if (worst_relres) worst_relres[c] = worst[c];
  }
  return rc;
}

}  // namespace bemcs

using namespace bemcs;

extern "C" {

int bemcs_poisson_create(const uint8_t* is_bound_host, int32_t nx, int32_t ny, int32_t periodic_y,
                         bemcs_poisson_t* out) {
  return bemcs_poisson_create_opt(is_bound_host, nx, ny, periodic_y, kCoarseMax, out);
}

int bemcs_poisson_create_opt(const uint8_t* is_bound_host, int32_t nx, int32_t ny, int32_t periodic_y,
                             int32_t coarse_max, bemcs_poisson_t* out) {
  BEMCS_REQUIRE(is_bound_host && out && nx >= 3 && ny >= 3 && coarse_max >= 16 && coarse_max <= kDenseMax,
                "bemcs_poisson_create: bad arguments (16 <= coarse_max <= 4096)");
  *out = nullptr;
  const int n = nx * ny;
  const bool per = periodic_y != 0;
  // level shapes (they depend only on the mesh): halve until the coarsest
  // level has at most kCoarseMax nodes
  int lnx[kMaxLevels], lny[kMaxLevels], nlev = 1;
  lnx[0] = nx;
  lny[0] = ny;
  while (nlev < kMaxLevels) {
    const int fx = lnx[nlev - 1], fy = lny[nlev - 1];
    if (fx * fy <= coarse_max) break;
    const int ncx = (fx + 1) / 2, ncy = (fy + 1) / 2;
    if (ncx < 3 || ncy < (per ? 4 : 3)) break;
    lnx[nlev] = ncx;
    lny[nlev] = ncy;
    ++nlev;
  }
  const int nc = lnx[nlev - 1] * lny[nlev - 1];
  BEMCS_REQUIRE(nc <= kDenseMax, "bemcs_poisson_create: coarsest level too large for the dense solve");

  bemcs_poisson_s* h = new bemcs_poisson_s();
  h->nx = nx;
  h->ny = ny;
  h->periodic = per ? 1 : 0;
  h->nlev = nlev;
  h->have_guess = false;
  h->tile0_pref = 0;
  h->op = nullptr;
  h->op_gen = 0;
  h->own = nullptr;
  h->nc = nc;
  bool ok = true;
  auto A = [&](size_t bytes) {
    void* p = dalloc(h, bytes);
    if (!p) ok = false;
    return p;
  };
  // the handle's own state: multigrid vectors of every level, CG vectors,
  // partial sums (the operator arrays come from an OpData)
  for (int l = 0; l < nlev && ok; ++l) {
    LevelDev& D = h->lev[l];
    D.nx = lnx[l];
    D.ny = lny[l];
    D.n = D.nx * D.ny;
    D.periodic = per ? 1 : 0;
    D.x = (float*)A(sizeof(float) * D.n);
    D.t = (float*)A(sizeof(float) * D.n);
    D.b = (float*)A(sizeof(float) * D.n);
    if (ok) ok = hipMemset(D.x, 0, sizeof(float) * D.n) == hipSuccess;
    if (ok) ok = hipMemset(D.t, 0, sizeof(float) * D.n) == hipSuccess;
    if (ok) ok = hipMemset(D.b, 0, sizeof(float) * D.n) == hipSuccess;
  }
  if (ok) {
    h->x = (double*)A(sizeof(double) * n);
    h->R[0] = (double*)A(sizeof(double) * n);
    h->R[1] = (double*)A(sizeof(double) * n);
    h->p = (double*)A(sizeof(double) * n);
    h->Ap = (double*)A(sizeof(double) * n);
    h->p1 = (double*)A(sizeof(double) * n);
    h->Ap1 = (double*)A(sizeof(double) * n);
    h->Az = (double*)A(sizeof(double) * n);
    h->xg = (double*)A(sizeof(double) * n);
    h->bred = (double*)A(sizeof(double) * n);
    h->Vnew = (double*)A(sizeof(double) * n);
    h->nparts = ((nx + kTF - 1) / kTF) * ((ny + kTF - 1) / kTF);
    h->pa = (double*)A(sizeof(double) * 2 * h->nparts);
    h->pr = (double*)A(sizeof(double) * 2 * h->nparts);
    h->pi = (double*)A(sizeof(double) * 2 * h->nparts);
    h->pz[0] = (double*)A(sizeof(double) * 4 * h->nparts);
    h->pz[1] = (double*)A(sizeof(double) * 4 * h->nparts);
    if (ok) ok = hipMemset(h->p, 0, sizeof(double) * n) == hipSuccess;
    if (ok) ok = hipMemset(h->Ap, 0, sizeof(double) * n) == hipSuccess;
    if (ok) ok = hipMemset(h->x, 0, sizeof(double) * n) == hipSuccess;
  }
  int rc = BEMCS_OK;
  if (ok) {
    // the operator: shared with a handle built from the same electrode map,
    // else built here (and offered to later handles)
    OpData* op = find_op(h, is_bound_host);
    if (op) {
      attach(h, op);
    } else {
      op = new_op(h);
      ok = op != nullptr;
      if (ok) {
        attach(h, op);
        uint8_t* ib = nullptr;
        ok = hipMalloc(&ib, n) == hipSuccess;
        if (ok) ok = upload(ib, is_bound_host, n) == 0;
        if (ok) {
          rc = setup_operator(h, ib, nullptr);
          if (rc == BEMCS_OK) rc = setup_status(h, nullptr);
        }
        if (ib) (void)hipFree(ib);
        if (ok && rc == BEMCS_OK) register_op(op, is_bound_host);
      }
    }
  }
  if (ok && rc == BEMCS_OK) ok = create_set(&h, 1, &h->own) == BEMCS_OK;
  if (!ok || rc != BEMCS_OK) {
    destroy(h);
    if (rc == BEMCS_OK) {
      set_error("bemcs_poisson_create: device allocation or upload failed");
      rc = BEMCS_EHIP;
    }
    return rc;
  }
  *out = h;
  return BEMCS_OK;
}

int bemcs_poisson_update(bemcs_poisson_t h, const uint8_t* is_bound, bemcs_stream_t stream) {
  BEMCS_REQUIRE(h && is_bound, "bemcs_poisson_update: bad arguments");
  hipStream_t s = as_stream(stream);
  if (!claim_private(h->op)) {
    // other handles read these arrays: build into arrays of its own; the
    // handle's own solve set is made again for the new pointers (other sets
    // that hold the handle refuse to run until they are made again)
    OpData* mine = new_op(h);
    if (!mine) {
      set_error("bemcs_poisson_update: device allocation failed");
      return BEMCS_EHIP;
    }
    OpData* old = h->op;
    attach(h, mine);
    SolveSet* S = nullptr;
    const int rc = create_set(&h, 1, &S);
    if (rc != BEMCS_OK) {
      attach(h, old);
      free_op(mine);
      return rc;
    }
    release_op(old);
    ++h->op_gen;
    destroy_set(h->own);
    h->own = S;
    S->gens[0] = h->op_gen;
  }
  int rc = setup_operator(h, is_bound, s);
  if (rc != BEMCS_OK) return rc;
  return setup_status(h, s);
}

int bemcs_poisson_free(bemcs_poisson_t h) {
  destroy(h);
  return BEMCS_OK;
}

int bemcs_poisson_levels(bemcs_poisson_t h) { return h ? h->nlev : 0; }

int bemcs_poisson_set_level0_tile(bemcs_poisson_t h, int32_t tile) {
  BEMCS_REQUIRE(h && (tile == 0 || tile == kTC || tile == kTCL),
                "bemcs_poisson_set_level0_tile: tile must be 0, 8 or 16");
  if (h->tile0_pref == tile) return BEMCS_OK;
  const int old = h->tile0_pref;
  h->tile0_pref = tile;
  SolveSet* S = nullptr;
  const int rc = create_set(&h, 1, &S);
  if (rc != BEMCS_OK) {
    h->tile0_pref = old;
    return rc;
  }
  destroy_set(h->own);
  h->own = S;
  return BEMCS_OK;
}

int bemcs_poisson_solve(bemcs_poisson_t h, const double* b, double* V, double rtol,
                        int32_t max_iter, int32_t* iters, double* relres, bemcs_stream_t stream) {
  BEMCS_REQUIRE(h && b && V && rtol > 0 && max_iter > 0, "bemcs_poisson_solve: bad arguments");
  hipStream_t s = as_stream(stream);
  const int n = h->nx * h->ny;
  hipLaunchKernelGGL(k_reduce_rhs, dim3(grid_for(n, 256)), dim3(256), 0, s, b, V, h->kind, h->bred,
                     h->x, h->nx, h->ny, h->periodic, 1);
  BEMCS_CHECK_LAUNCH();
  int it = 0;
  double rel = 0;
  launch_cg_init(h->own, s);
  int rc = pcg(h->own, 0, 0, rtol, max_iter, &it, &rel, s);
  if (iters) *iters = it;
  if (relres) *relres = rel;
  hipLaunchKernelGGL(k_assemble, dim3(grid_for(n, 256)), dim3(256), 0, s, h->x, b, h->kind,
                     h->mirror, h->mirror1, V, n);
  BEMCS_CHECK_LAUNCH();
  return rc;
}

int bemcs_poisson_picard(bemcs_poisson_t h, const bemcs_picard_t* p, int32_t* iters_total,
                         double* worst_relres, bemcs_stream_t stream) {
  BEMCS_REQUIRE(h && p, "bemcs_poisson_picard: bad arguments");
  return picard(h->own, p, iters_total, worst_relres, as_stream(stream));
}

int bemcs_poisson_batch_create(const bemcs_poisson_t* handles, int32_t n, bemcs_poisson_batch_t* out) {
  BEMCS_REQUIRE(handles && out && n >= 1, "bemcs_poisson_batch_create: bad arguments");
  *out = nullptr;
  for (int c = 0; c < n; ++c) {
    BEMCS_REQUIRE(handles[c], "bemcs_poisson_batch_create: null handle");
    for (int d = 0; d < c; ++d)
      BEMCS_REQUIRE(handles[c] != handles[d], "bemcs_poisson_batch_create: a handle appears twice");
  }
  return create_set(handles, n, out);
}

int bemcs_poisson_batch_free(bemcs_poisson_batch_t b) {
  destroy_set(b);
  return BEMCS_OK;
}

int bemcs_poisson_batch_size(bemcs_poisson_batch_t b) { return b ? b->n : 0; }

int bemcs_poisson_batch_picard(bemcs_poisson_batch_t b, const bemcs_picard_t* params,
                               int32_t* iters_total, double* worst_relres, bemcs_stream_t stream) {
  BEMCS_REQUIRE(b && params, "bemcs_poisson_batch_picard: bad arguments");
  for (int c = 0; c < b->n; ++c)
    BEMCS_REQUIRE(b->cases[c]->op_gen == b->gens[c],
                  "bemcs_poisson_batch_picard: a handle of the set moved to operator arrays of its own "
                  "(bemcs_poisson_update of a shared operator); create the set again");
  return picard(b, params, iters_total, worst_relres, as_stream(stream));
}

int bemcs_efield(const double* V, int32_t nx, int32_t ny, double dx_m, double dy_m,
                 int32_t periodic_y, float* Exy, bemcs_stream_t stream) {
  BEMCS_REQUIRE(V && Exy && nx >= 2 && ny >= 2, "bemcs_efield: bad arguments");
  hipLaunchKernelGGL(k_efield, dim3(grid_for((int64_t)nx * ny, 256)), dim3(256), 0,
                     as_stream(stream), V, nx, ny, dx_m, dy_m, periodic_y, Exy);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

}  // extern "C"
