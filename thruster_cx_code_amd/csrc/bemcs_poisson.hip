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
// 9-point coarse operators built on the host, damped Jacobi smoothing, dense
// inverse on the coarsest level, at most kCoarseMax nodes).
#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <vector>

#include "bemcs_internal.h"

namespace bemcs {

constexpr int kMaxLevels = 20;
constexpr int kCoarseMax = 512;  // coarsening stops at or below this many nodes
constexpr int kDenseMax = 4096;  // largest coarsest level the dense solve accepts
constexpr int kFuseBelow = 32768;  // levels this small fuse residual+restrict, prolong+sweep
constexpr int kRedBlock = 256;
constexpr int kMaxRedBlocks = 1024;
constexpr float kOmega0 = 0.8f;  // Jacobi damping, 5-point fine level
constexpr float kOmegaC = 0.7f;  // Jacobi damping, Galerkin 9-point levels
constexpr int kNu = 2;           // pre- and post-smoothing sweeps

enum : uint8_t { K_BOUND = 0, K_UNKNOWN = 1, K_RIGHT = 2, K_TOP = 3, K_BOTTOM = 4 };

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
  double *x, *r, *p, *Ap, *rold, *bred, *Vnew;
  double* part;      // reduction partials
  double* scal;      // device scalars, see S_* below
  int32_t* done;
  double* host_scal; // pinned copy
  bool have_guess;
  float* rscr;        // float32 residual scratch (largest level)
  double* part2;     // step partials
  int pred_iters;    // iterations the last solve needed (launch that many first)
  int post_sweeps;    // Jacobi sweeps after prolongation (1 or 2)
};

namespace bemcs {

enum { S_RZ = 0, S_PAP, S_RR, S_BB, S_ALPHA, S_BETA, S_TOL2, S_ITERS, S_RTOL, S_COUNT = 16 };

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

__device__ __forceinline__ float apply9(const LevelDev& L, const float* __restrict__ v, int i) {
  int x = i % L.nx, y = i / L.nx;
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    int dx = k % 3 - 1, dy = k / 3 - 1;
    int xx = x + dx, yy = y + dy;
    if (xx < 0 || xx >= L.nx) continue;
    if (yy < 0 || yy >= L.ny) {
      if (!L.periodic) continue;
      yy = (yy + L.ny) % L.ny;
    }
    float c = L.st[(size_t)k * L.n + i];
    s += c * v[yy * L.nx + xx];
  }
  return s;
}

__device__ __forceinline__ float applyL(const LevelDev& L, const float* v, int i) {
  return L.st ? apply9(L, v, i) : apply5<float>(v, L.code, i, L.nx, L.ny, L.periodic != 0);
}

// --------------------------------------------------------- V-cycle phases
// Device functions shared by the multi-block kernels and the single
// work-group tail kernel.  A V-cycle does per level: two Jacobi sweeps from
// zero (pre), residual fused with restriction, prolongation fused with the
// first post sweep, second post sweep.  The post sweeps write t and the
// pointers are swapped twice, so x is the output buffer again afterwards.

// first Jacobi iterate from x = 0
__device__ __forceinline__ float x1_at(const LevelDev& L, float om, int j) {
  return om * L.inv_d[j] * L.b[j];
}

__device__ __forceinline__ float apply_x1(const LevelDev& L, float om, int i) {
  int x = i % L.nx, y = i / L.nx;
  if (!L.st) {
    int d = L.code[i];
    int s_ = y > 0 ? i - L.nx : (L.periodic ? i + (L.ny - 1) * L.nx : i);
    int n_ = y < L.ny - 1 ? i + L.nx : (L.periodic ? i - (L.ny - 1) * L.nx : i);
    float sum = x1_at(L, om, i - 1) + x1_at(L, om, i + 1) + x1_at(L, om, s_) + x1_at(L, om, n_);
    return (float)d * x1_at(L, om, i) - sum;
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    int xx = x + (k % 3 - 1), yy = y + (k / 3 - 1);
    if (xx < 0 || xx >= L.nx) continue;
    if (yy < 0 || yy >= L.ny) {
      if (!L.periodic) continue;
      yy = (yy + L.ny) % L.ny;
    }
    s += L.st[(size_t)k * L.n + i] * x1_at(L, om, yy * L.nx + xx);
  }
  return s;
}

__device__ __forceinline__ void ph_pre(const LevelDev& L, float om, int i) {
  float id = L.inv_d[i];
  if (id == 0.0f) {
    L.x[i] = 0.0f;
    return;
  }
  float x1 = om * id * L.b[i];
  L.x[i] = x1 + om * id * (L.b[i] - apply_x1(L, om, i));
}

__device__ __forceinline__ float resid_at(const LevelDev& F, int f) {
  return F.inv_d[f] != 0.0f ? F.b[f] - applyL(F, F.x, f) : 0.0f;
}

__device__ __forceinline__ void ph_resid_restrict(const LevelDev& F, const LevelDev& C, int I) {
  if (C.inv_d[I] == 0.0f) {
    C.b[I] = 0.0f;
    return;
  }
  int cx = I % C.nx, cy = I / C.nx;
  float s = 0.0f;
  for (int oy = -1; oy <= 1; ++oy) {
    int fy = 2 * cy + oy;
    if (fy < 0 || fy >= F.ny) {
      if (!F.periodic) continue;
      fy = (fy + F.ny) % F.ny;
    }
    float wy = w1d(fy, cy, C.ny, F.periodic != 0);
    if (wy == 0.0f) continue;
    for (int ox = -1; ox <= 1; ++ox) {
      int fx = 2 * cx + ox;
      if (fx < 0 || fx >= F.nx) continue;
      float wx = w1d(fx, cx, C.nx, false);
      if (wx == 0.0f) continue;
      s += wx * wy * resid_at(F, fy * F.nx + fx);
    }
  }
  C.b[I] = s;
}

// coarse correction P e at fine node i (0 at inactive fine nodes)
__device__ __forceinline__ float pe_at(const LevelDev& F, const LevelDev& C, int i) {
  if (F.inv_d[i] == 0.0f) return 0.0f;
  int fx = i % F.nx, fy = i / F.nx;
  int ax0 = fx >> 1, ax1 = -1;
  if (fx & 1) {
    ax0 = (fx - 1) >> 1;
    ax1 = (fx + 1) >> 1;
    if (ax1 >= C.nx) ax1 = -1;
  }
  int ay0 = fy >> 1, ay1 = -1;
  if (fy & 1) {
    ay0 = (fy - 1) >> 1;
    ay1 = (fy + 1) >> 1;
    if (ay1 >= C.ny) ay1 = F.periodic ? ay1 - C.ny : -1;
  }
  float s = 0.0f;
  int xs[2] = {ax0, ax1}, ys[2] = {ay0, ay1};
  for (int a = 0; a < 2; ++a) {
    if (ys[a] < 0) continue;
    float wy = w1d(fy, ys[a], C.ny, F.periodic != 0);
    for (int b = 0; b < 2; ++b) {
      if (xs[b] < 0) continue;
      float wx = w1d(fx, xs[b], C.nx, false);
      s += wx * wy * C.x[ys[a] * C.nx + xs[b]];
    }
  }
  return s;
}

__device__ __forceinline__ float xnew_at(const LevelDev& F, const LevelDev& C, int j) {
  return F.x[j] + pe_at(F, C, j);
}

// x <- x + P e, then one Jacobi sweep; result in t
__device__ __forceinline__ void ph_post1(const LevelDev& F, const LevelDev& C, float om, int i) {
  float id = F.inv_d[i];
  if (id == 0.0f) {
    F.t[i] = 0.0f;
    return;
  }
  int x = i % F.nx, y = i / F.nx;
  float xi = xnew_at(F, C, i);
  float ax;
  if (!F.st) {
    int d = F.code[i];
    int s_ = y > 0 ? i - F.nx : (F.periodic ? i + (F.ny - 1) * F.nx : i);
    int n_ = y < F.ny - 1 ? i + F.nx : (F.periodic ? i - (F.ny - 1) * F.nx : i);
    float sum = xnew_at(F, C, i - 1) + xnew_at(F, C, i + 1) + xnew_at(F, C, s_) + xnew_at(F, C, n_);
    ax = (float)d * xi - sum;
  } else {
    ax = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      int xx = x + (k % 3 - 1), yy = y + (k / 3 - 1);
      if (xx < 0 || xx >= F.nx) continue;
      if (yy < 0 || yy >= F.ny) {
        if (!F.periodic) continue;
        yy = (yy + F.ny) % F.ny;
      }
      int j = yy * F.nx + xx;
      ax += F.st[(size_t)k * F.n + i] * (j == i ? xi : xnew_at(F, C, j));
    }
  }
  F.t[i] = xi + om * id * (F.b[i] - ax);
}

__device__ __forceinline__ void ph_jacobi(const LevelDev& L, float om, int i) {
  float id = L.inv_d[i];
  if (id == 0.0f) {
    L.t[i] = 0.0f;
    return;
  }
  L.t[i] = L.x[i] + om * id * (L.b[i] - applyL(L, L.x, i));
}

__device__ __forceinline__ float omega_of(const LevelDev& L) { return L.st ? kOmegaC : kOmega0; }

__global__ __launch_bounds__(256) void k_pre(LevelDev L, const int32_t* __restrict__ done) {
  if (*done) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < L.n) ph_pre(L, omega_of(L), i);
}
__global__ __launch_bounds__(256) void k_resid_restrict(LevelDev F, LevelDev C, const int32_t* __restrict__ done) {
  if (*done) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < C.n) ph_resid_restrict(F, C, i);
}
__global__ __launch_bounds__(256) void k_post1(LevelDev F, LevelDev C, const int32_t* __restrict__ done) {
  if (*done) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < F.n) ph_post1(F, C, omega_of(F), i);
}
__global__ __launch_bounds__(256) void k_jacobi(LevelDev L, const int32_t* __restrict__ done) {
  if (*done) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < L.n) ph_jacobi(L, omega_of(L), i);
}

__global__ __launch_bounds__(256) void k_residual(LevelDev L, float* r, const int32_t* __restrict__ done) {
  if (*done) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < L.n) r[i] = resid_at(L, i);
}

// restriction of a stored residual r (fine) into C.b
__global__ __launch_bounds__(256) void k_restrict(LevelDev F, const float* __restrict__ r, LevelDev C,
                                                  const int32_t* __restrict__ done) {
  if (*done) return;
  int I = blockIdx.x * 256 + threadIdx.x;
  if (I >= C.n) return;
  if (C.inv_d[I] == 0.0f) {
    C.b[I] = 0.0f;
    return;
  }
  int cx = I % C.nx, cy = I / C.nx;
  float s = 0.0f;
  for (int oy = -1; oy <= 1; ++oy) {
    int fy = 2 * cy + oy;
    if (fy < 0 || fy >= F.ny) {
      if (!F.periodic) continue;
      fy = (fy + F.ny) % F.ny;
    }
    float wy = w1d(fy, cy, C.ny, F.periodic != 0);
    if (wy == 0.0f) continue;
    for (int ox = -1; ox <= 1; ++ox) {
      int fx = 2 * cx + ox;
      if (fx < 0 || fx >= F.nx) continue;
      float wx = w1d(fx, cx, C.nx, false);
      if (wx == 0.0f) continue;
      s += wx * wy * r[fy * F.nx + fx];
    }
  }
  C.b[I] = s;
}

// x += P e
__global__ __launch_bounds__(256) void k_prolong(LevelDev F, LevelDev C, const int32_t* __restrict__ done) {
  if (*done) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < F.n) F.x[i] += pe_at(F, C, i);
}

// coarsest level: x = C^-1 b with the inverse stored column-major.  A block
// owns 64 rows (one per lane); its 16 waves split the columns, and the partial
// sums are added in LDS, so each thread runs only nc/16 independent loads.
constexpr int kCoarseWaves = 16;
__global__ __launch_bounds__(64 * kCoarseWaves) void k_coarse_solve(LevelDev C,
                                                                    const float* __restrict__ cinv,
                                                                    const int32_t* __restrict__ done) {
  __shared__ float sb[kDenseMax];
  if (*done) return;
  __shared__ float part[kCoarseWaves][64];
  const int nc = C.n;
  for (int j = threadIdx.x; j < nc; j += blockDim.x) sb[j] = C.b[j];
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row = blockIdx.x * 64 + lane;
  float s = 0.0f;
  if (row < nc) {
    const int chunk = (nc + kCoarseWaves - 1) / kCoarseWaves;
    const int j0 = w * chunk, j1 = min(nc, j0 + chunk);
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    int j = j0;
    for (; j + 4 <= j1; j += 4) {
      s0 += cinv[(size_t)j * nc + row] * sb[j];
      s1 += cinv[(size_t)(j + 1) * nc + row] * sb[j + 1];
      s2 += cinv[(size_t)(j + 2) * nc + row] * sb[j + 2];
      s3 += cinv[(size_t)(j + 3) * nc + row] * sb[j + 3];
    }
    for (; j < j1; ++j) s0 += cinv[(size_t)j * nc + row] * sb[j];
    s = (s0 + s1) + (s2 + s3);
  }
  part[w][lane] = s;
  __syncthreads();
  if (w == 0 && row < nc) {
    float t = 0.0f;
    for (int k = 0; k < kCoarseWaves; ++k) t += part[k][lane];
    C.x[row] = t;
  }
}

// ----------------------------------------------------------- CG kernels
// Vector kernels write per-block partial sums (fixed-order wave and block
// reductions); a one-block kernel adds the partials in a fixed order and
// updates the scalars, so every reduction is deterministic.
struct CGArgs {
  double *x, *r, *rold, *p, *Ap;
  const double* bred;
  float* b0;      // level-0 right-hand side of the V-cycle (float32 r)
  const float* z; // level-0 V-cycle output
  const uint8_t* code;
  int nx, ny, per, n;
  double* part;   // 2 * kMaxRedBlocks: partials of dir and dotz
  double* part2;  // 2 * kMaxRedBlocks: partials of step (r.r)
  double* scal;
  int32_t* done;
};

// grid-stride loops so the number of partials stays <= kMaxRedBlocks
#define GRID_LOOP(i, n) for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < (n); i += gridDim.x * kRedBlock)

__device__ __forceinline__ void block_partials(double a, double b, double* part) {
  __shared__ double sa[kRedBlock / 64], sb[kRedBlock / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  if (lane_id() == 0) {
    sa[threadIdx.x / 64] = a;
    sb[threadIdx.x / 64] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double pa = 0, pb = 0;
    for (int w = 0; w < kRedBlock / 64; ++w) {
      pa += sa[w];
      pb += sb[w];
    }
    part[2 * blockIdx.x] = pa;
    part[2 * blockIdx.x + 1] = pb;
  }
}

// r = b - A x, b0 = f32(r), p = Ap = 0 ; partials (r.r, b.b)
__global__ __launch_bounds__(kRedBlock) void k_cg_init(CGArgs g) {
  double rr = 0, bb = 0;
  GRID_LOOP(i, g.n) {
    double ri = g.code[i] ? g.bred[i] - apply5<double>(g.x, g.code, i, g.nx, g.ny, g.per != 0) : 0.0;
    g.r[i] = ri;
    g.b0[i] = (float)ri;
    g.p[i] = 0.0;
    g.Ap[i] = 0.0;
    rr += ri * ri;
    bb += g.bred[i] * g.bred[i];
  }
  block_partials(rr, bb, g.part);
}

// p = z + beta p ; Ap = A z + beta Ap ; partial p.Ap
__global__ __launch_bounds__(kRedBlock) void k_cg_dir(CGArgs g) {
  if (*g.done) return;
  const double beta = g.scal[S_BETA];
  double s = 0;
  GRID_LOOP(i, g.n) {
    double az = 0.0;
    int d = g.code[i];
    if (d) {
      int x = i % g.nx, y = i / g.nx;
      int sn = y > 0 ? i - g.nx : (g.per ? i + (g.ny - 1) * g.nx : i);
      int nn = y < g.ny - 1 ? i + g.nx : (g.per ? i - (g.ny - 1) * g.nx : i);
      double nb = (double)g.z[i - 1] + (double)g.z[i + 1] + (double)g.z[sn] + (double)g.z[nn];
      az = (double)d * (double)g.z[i] - nb;
      (void)x;
    }
    double pi = (double)g.z[i] + beta * g.p[i];
    double api = az + beta * g.Ap[i];
    g.p[i] = pi;
    g.Ap[i] = api;
    s += pi * api;
  }
  block_partials(s, 0.0, g.part);
}

// x += a p ; r_old = r ; r -= a Ap ; b0 = f32(r) ; partial r.r
__global__ __launch_bounds__(kRedBlock) void k_cg_step(CGArgs g) {
  if (*g.done) return;
  const double a = g.scal[S_ALPHA];
  double rr = 0;
  GRID_LOOP(i, g.n) {
    g.x[i] += a * g.p[i];
    double ro = g.r[i];
    g.rold[i] = ro;
    double rn = ro - a * g.Ap[i];
    g.r[i] = rn;
    g.b0[i] = (float)rn;
    rr += rn * rn;
  }
  block_partials(rr, 0.0, g.part2);
}

// partials (r.z, r_old.z)
__global__ __launch_bounds__(kRedBlock) void k_cg_dotz(CGArgs g) {
  if (*g.done) return;
  double a = 0, b = 0;
  GRID_LOOP(i, g.n) {
    double zi = g.z[i];
    a += g.r[i] * zi;
    b += g.rold[i] * zi;
  }
  block_partials(a, b, g.part);
}

enum { M_INIT = 0, M_ALPHA = 1, M_BETA0 = 3, M_BETA = 4 };

// fixed-order sum of n partial pairs by one block
__device__ __forceinline__ void sum_partials(const double* part, int nparts, double* a_out,
                                             double* b_out) {
  __shared__ double sa[256], sb[256];
  double a = 0, b = 0;
  for (int k = threadIdx.x; k < nparts; k += 256) {
    a += part[2 * k];
    b += part[2 * k + 1];
  }
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sa[threadIdx.x] += sa[threadIdx.x + o];
      sb[threadIdx.x] += sb[threadIdx.x + o];
    }
    __syncthreads();
  }
  *a_out = sa[0];
  *b_out = sb[0];
}

// scalar recurrence: INIT (tolerance, done), ALPHA = rz / p.Ap, BETA
// (Polak-Ribiere beta, rz) together with the convergence test on the r.r
// partials of the preceding step kernel (an iteration that converged still
// runs its V-cycle, whose output is then unused).
__global__ __launch_bounds__(256) void k_cg_scalar(const double* part, const double* part2, int nparts,
                                                   double* scal, int32_t* done, int mode) {
  if (mode != M_INIT && *done) return;
  double a, b, rr = 0.0, unused;
  sum_partials(part, nparts, &a, &b);
  if (mode == M_BETA) {
    __syncthreads();
    sum_partials(part2, nparts, &rr, &unused);
  }
  if (threadIdx.x != 0) return;
  if (mode == M_INIT) {
    const double rt = scal[S_RTOL];
    scal[S_RR] = a;
    scal[S_BB] = b;
    scal[S_TOL2] = rt * rt * b;
    scal[S_ITERS] = 0.0;
    scal[S_BETA] = 0.0;
    *done = (a <= scal[S_TOL2]) ? 1 : 0;
  } else if (mode == M_ALPHA) {
    scal[S_PAP] = a;
    scal[S_ALPHA] = a != 0.0 ? scal[S_RZ] / a : 0.0;
  } else {
    if (mode == M_BETA) {
      scal[S_RR] = rr;
      scal[S_ITERS] += 1.0;
      if (rr <= scal[S_TOL2] || !(rr == rr)) {
        *done = 1;
        return;
      }
    }
    const double rz_old = scal[S_RZ];
    scal[S_BETA] = (mode == M_BETA0 || rz_old == 0.0) ? 0.0 : (a - b) / rz_old;
    scal[S_RZ] = a;
  }
}

__global__ void k_set_rtol(double* scal, double rtol) { scal[S_RTOL] = rtol; }

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

// Picard right-hand side (:891-898) directly in reduced form:
// rho (float32) from NGP counts, rho_total = rho + rho_e(V),
// b_red = coeff*rho_total + sum of Dirichlet neighbour values.
__global__ __launch_bounds__(256) void k_picard_rhs(const float* __restrict__ rho,
                                                    const double* __restrict__ V,
                                                    const double* __restrict__ Vf,
                                                    const uint8_t* __restrict__ kind, double* bred,
                                                    int nx, int ny, int per, double coeff,
                                                    double Vp, double Te, double qn0) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nx * ny) return;
  if (kind[i] != K_UNKNOWN) {
    bred[i] = 0.0;
    return;
  }
  double v = V[i];
  double rho_e = qn0 * det_exp(((v < Vp ? v : Vp) - Vp) / Te);
  double rt = (double)rho[i] + rho_e;
  double s = -((-coeff) * rt);
  int xx = i % nx, y = i / nx;
  int nb[4] = {i - 1, i + 1, nbr_s(i, xx, y, nx, ny, per != 0), nbr_n(i, xx, y, nx, ny, per != 0)};
  for (int k = 0; k < 4; ++k)
    if (kind[nb[k]] == K_BOUND) s += Vf[nb[k]];
  bred[i] = s;
}

// rho from NGP counts as the reference's two np.add.at deposits produce it:
// per particle rho = f32(f64(rho) + cd), ions first, then electrons.
__global__ __launch_bounds__(256) void k_rho_from_counts(const uint32_t* __restrict__ ci,
                                                         const uint32_t* __restrict__ ce, int n,
                                                         double cdi, double cde, float* rho) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float r = 0.0f;
  for (uint32_t k = ci[i]; k > 0; --k) r = (float)((double)r + cdi);
  for (uint32_t k = ce[i]; k > 0; --k) r = (float)((double)r + cde);
  rho[i] = r;
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

// Picard variant: Dirichlet values from V_fixed, Neumann rows have b = 0;
// V = (1-omega) V + omega V_new.
__global__ __launch_bounds__(256) void k_picard_relax(const double* __restrict__ x,
                                                      const double* __restrict__ Vf,
                                                      const uint8_t* __restrict__ kind,
                                                      const int32_t* __restrict__ mirror, double* V,
                                                      double* Vnew, int n, double one_m_omega,
                                                      double omega) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uint8_t k = kind[i];
  double vn;
  if (k == K_UNKNOWN) vn = x[i];
  else if (k == K_BOUND) vn = Vf[i];
  else {
    int m = mirror[i];
    vn = kind[m] == K_UNKNOWN ? x[m] : Vf[m];
  }
  Vnew[i] = vn;
  V[i] = one_m_omega * V[i] + omega * vn;
}

// E = -grad V (np.gradient, edge_order 1) with the periodic Ey fix; float32
__global__ __launch_bounds__(256) void k_efield(const double* __restrict__ V, int nx, int ny,
                                                double dx_m, double dy_m, int per, float* Exy) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nx * ny) return;
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

// ------------------------------------------------------------- host set-up
struct HostLevel {
  int nx, ny;
  bool periodic;
  std::vector<double> st;  // [9][n]
  std::vector<uint8_t> active;
};

static void* dalloc(bemcs_poisson_s* h, size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  h->allocs.push_back(p);
  return p;
}

// Galerkin coarse operator P^T A P for bilinear P (w1d tensor product),
// P rows zero at inactive fine nodes.
static HostLevel coarsen(const HostLevel& F) {
  HostLevel C;
  C.nx = (F.nx + 1) / 2;
  C.ny = (F.ny + 1) / 2;
  C.periodic = F.periodic;
  const int nF = F.nx * F.ny, nC = C.nx * C.ny;
  C.st.assign((size_t)9 * nC, 0.0);
  C.active.assign(nC, 0);
  auto parents = [](int f, int nf, int nc, bool per, int* idx, double* w) -> int {
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
  };
  for (int f = 0; f < nF; ++f) {
    if (!F.active[f]) continue;
    int fx = f % F.nx, fy = f / F.nx;
    int pxi[2], pyi[2];
    double pxw[2], pyw[2];
    int npx = parents(fx, F.nx, C.nx, false, pxi, pxw);
    int npy = parents(fy, F.ny, C.ny, F.periodic, pyi, pyw);
    for (int k = 0; k < 9; ++k) {
      double a = F.st[(size_t)k * nF + f];
      if (a == 0.0) continue;
      int gx = fx + (k % 3 - 1), gy = fy + (k / 3 - 1);
      if (gx < 0 || gx >= F.nx) continue;
      if (gy < 0 || gy >= F.ny) {
        if (!F.periodic) continue;
        gy = (gy + F.ny) % F.ny;
      }
      int g = gy * F.nx + gx;
      if (!F.active[g]) continue;
      int qxi[2], qyi[2];
      double qxw[2], qyw[2];
      int nqx = parents(gx, F.nx, C.nx, false, qxi, qxw);
      int nqy = parents(gy, F.ny, C.ny, F.periodic, qyi, qyw);
      for (int ay = 0; ay < npy; ++ay)
        for (int ax = 0; ax < npx; ++ax) {
          int I = pyi[ay] * C.nx + pxi[ax];
          double w1 = pxw[ax] * pyw[ay] * a;
          for (int by = 0; by < nqy; ++by)
            for (int bx = 0; bx < nqx; ++bx) {
              int ddx = qxi[bx] - pxi[ax];
              int ddy = qyi[by] - pyi[ay];
              if (C.periodic) {
                if (ddy > 1) ddy -= C.ny;
                if (ddy < -1) ddy += C.ny;
              }
              int kk = (ddy + 1) * 3 + (ddx + 1);
              C.st[(size_t)kk * nC + I] += w1 * qxw[bx] * qyw[by];
            }
        }
    }
  }
  for (int I = 0; I < nC; ++I) C.active[I] = C.st[(size_t)4 * nC + I] > 1e-12 ? 1 : 0;
  // drop couplings to inactive coarse nodes and from inactive rows
  for (int I = 0; I < nC; ++I) {
    int cx = I % C.nx, cy = I / C.nx;
    for (int k = 0; k < 9; ++k) {
      int gx = cx + (k % 3 - 1), gy = cy + (k / 3 - 1);
      bool ok = gx >= 0 && gx < C.nx;
      if (gy < 0 || gy >= C.ny) {
        if (C.periodic) gy = (gy + C.ny) % C.ny;
        else ok = false;
      }
      if (!ok || !C.active[I] || !C.active[gy * C.nx + gx]) C.st[(size_t)k * nC + I] = 0.0;
    }
  }
  return C;
}

// Dense inverse of the active block (Gauss-Jordan with partial pivoting),
// returned column-major: element (i, j) at j * n + i.
static std::vector<float> dense_inverse(const HostLevel& L) {
  const int n = L.nx * L.ny;
  std::vector<int> act;
  for (int i = 0; i < n; ++i)
    if (L.active[i]) act.push_back(i);
  const int m = (int)act.size();
  std::vector<int> pos(n, -1);
  for (int k = 0; k < m; ++k) pos[act[k]] = k;
  std::vector<double> A((size_t)m * m, 0.0), Inv((size_t)m * m, 0.0);
  for (int k = 0; k < m; ++k) {
    int I = act[k], cx = I % L.nx, cy = I / L.nx;
    for (int s = 0; s < 9; ++s) {
      double c = L.st[(size_t)s * n + I];
      if (c == 0.0) continue;
      int gx = cx + (s % 3 - 1), gy = cy + (s / 3 - 1);
      if (gy < 0 || gy >= L.ny) gy = (gy + L.ny) % L.ny;
      int j = pos[gy * L.nx + gx];
      if (j >= 0) A[(size_t)k * m + j] += c;
    }
    Inv[(size_t)k * m + k] = 1.0;
  }
  for (int c = 0; c < m; ++c) {
    int piv = c;
    for (int r = c + 1; r < m; ++r)
      if (fabs(A[(size_t)r * m + c]) > fabs(A[(size_t)piv * m + c])) piv = r;
    if (piv != c)
      for (int j = 0; j < m; ++j) {
        std::swap(A[(size_t)c * m + j], A[(size_t)piv * m + j]);
        std::swap(Inv[(size_t)c * m + j], Inv[(size_t)piv * m + j]);
      }
    double d = A[(size_t)c * m + c];
    if (fabs(d) < 1e-300) continue;
    for (int j = 0; j < m; ++j) {
      A[(size_t)c * m + j] /= d;
      Inv[(size_t)c * m + j] /= d;
    }
    for (int r = 0; r < m; ++r) {
      if (r == c) continue;
      double f = A[(size_t)r * m + c];
      if (f == 0.0) continue;
      for (int j = 0; j < m; ++j) {
        A[(size_t)r * m + j] -= f * A[(size_t)c * m + j];
        Inv[(size_t)r * m + j] -= f * Inv[(size_t)c * m + j];
      }
    }
  }
  std::vector<float> out((size_t)n * n, 0.0f);
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b) out[(size_t)act[b] * n + act[a]] = (float)Inv[(size_t)a * m + b];
  return out;
}

static int upload(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}

static void destroy(bemcs_poisson_s* h) {
  if (!h) return;

  for (void* p : h->allocs) (void)hipFree(p);
  if (h->host_scal) (void)hipHostFree(h->host_scal);
  delete h;
}

// --------------------------------------------------------- V-cycle driver
// Per level above the coarsest: fused pre-smoothing (2 Jacobi sweeps), then
// residual + restriction (one fused kernel on small levels, two kernels on
// large ones where recomputing residuals costs more than a launch), then on
// the way up prolongation + two Jacobi sweeps (the first fused with the
// prolongation on small levels).  The coarsest level is a dense solve.
static void vcycle(bemcs_poisson_s* h, hipStream_t s) {
  const int last = h->nlev - 1;
  const int32_t* dn = h->done;
  for (int l = 0; l < last; ++l) {
    LevelDev& L = h->lev[l];
    LevelDev& C = h->lev[l + 1];
    hipLaunchKernelGGL(k_pre, dim3(grid_for(L.n, 256)), dim3(256), 0, s, L, dn);
    if (L.n <= kFuseBelow) {
      hipLaunchKernelGGL(k_resid_restrict, dim3(grid_for(C.n, 256)), dim3(256), 0, s, L, C, dn);
    } else {
      hipLaunchKernelGGL(k_residual, dim3(grid_for(L.n, 256)), dim3(256), 0, s, L, h->rscr, dn);
      hipLaunchKernelGGL(k_restrict, dim3(grid_for(C.n, 256)), dim3(256), 0, s, L,
                         (const float*)h->rscr, C, dn);
    }
  }
  hipLaunchKernelGGL(k_coarse_solve, dim3((h->lev[last].n + 63) / 64), dim3(64 * kCoarseWaves), 0, s,
                     h->lev[last], (const float*)h->cinv, dn);
  for (int l = last - 1; l >= 0; --l) {
    LevelDev& L = h->lev[l];
    LevelDev& C = h->lev[l + 1];
    unsigned g = grid_for(L.n, 256);
    if (L.n <= kFuseBelow) {
      hipLaunchKernelGGL(k_post1, dim3(g), dim3(256), 0, s, L, C, dn);
      std::swap(L.x, L.t);
    } else {
      hipLaunchKernelGGL(k_prolong, dim3(g), dim3(256), 0, s, L, C, dn);
      hipLaunchKernelGGL(k_jacobi, dim3(g), dim3(256), 0, s, L, dn);
      std::swap(L.x, L.t);
    }
    if (h->post_sweeps > 1) {
      hipLaunchKernelGGL(k_jacobi, dim3(g), dim3(256), 0, s, L, dn);
      std::swap(L.x, L.t);
    }
  }
}

static unsigned red_grid(int n) {
  unsigned g = grid_for(n, kRedBlock);
  return g > (unsigned)kMaxRedBlocks ? (unsigned)kMaxRedBlocks : g;
}

static CGArgs cg_args(bemcs_poisson_s* h) {
  CGArgs g;
  g.x = h->x;
  g.r = h->r;
  g.rold = h->rold;
  g.p = h->p;
  g.Ap = h->Ap;
  g.bred = h->bred;
  g.b0 = h->lev[0].b;
  g.z = h->lev[0].x;
  g.code = h->lev[0].code;
  g.nx = h->nx;
  g.ny = h->ny;
  g.per = h->periodic;
  g.n = h->nx * h->ny;
  g.part = h->part;
  g.part2 = h->part2;
  g.scal = h->scal;
  g.done = h->done;
  return g;
}


static void enqueue_init(bemcs_poisson_s* h, const CGArgs& g, hipStream_t s) {
  const unsigned rg = red_grid(g.n);
  hipLaunchKernelGGL(k_cg_init, dim3(rg), dim3(kRedBlock), 0, s, g);
  hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, s, (const double*)h->part,
                     (const double*)h->part2, (int)rg, h->scal, h->done, (int)M_INIT);
  vcycle(h, s);
  hipLaunchKernelGGL(k_cg_dotz, dim3(rg), dim3(kRedBlock), 0, s, g);
  hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, s, (const double*)h->part,
                     (const double*)h->part2, (int)rg, h->scal, h->done, (int)M_BETA0);
}

static void enqueue_iter(bemcs_poisson_s* h, const CGArgs& g, hipStream_t s) {
  const unsigned rg = red_grid(g.n);
  {
    hipLaunchKernelGGL(k_cg_dir, dim3(rg), dim3(kRedBlock), 0, s, g);
    hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, s, (const double*)h->part,
                       (const double*)h->part2, (int)rg, h->scal, h->done, (int)M_ALPHA);
    hipLaunchKernelGGL(k_cg_step, dim3(rg), dim3(kRedBlock), 0, s, g);
    vcycle(h, s);
    hipLaunchKernelGGL(k_cg_dotz, dim3(rg), dim3(kRedBlock), 0, s, g);
    hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, s, (const double*)h->part,
                       (const double*)h->part2, (int)rg, h->scal, h->done, (int)M_BETA);
  }
}

// PCG on the reduced system already in h->bred with initial guess h->x.
static int pcg(bemcs_poisson_s* h, double rtol, int max_iter, int* iters, double* relres,
               hipStream_t s) {
  const CGArgs g = cg_args(h);
  hipLaunchKernelGGL(k_set_rtol, dim3(1), dim3(1), 0, s, h->scal, rtol);
  enqueue_init(h, g, s);
  BEMCS_CHECK_LAUNCH();
  // launch as many iterations as the previous solve needed, then check
  int launched = 0;
  int batch = std::max(1, h->pred_iters);
  for (;;) {
    for (int k = 0; k < batch && launched < max_iter; ++k, ++launched) enqueue_iter(h, g, s);
    BEMCS_CHECK_LAUNCH();
    BEMCS_CHECK_HIP(hipMemcpyAsync(h->host_scal, h->scal, sizeof(double) * S_COUNT,
                                   hipMemcpyDeviceToHost, s));
    BEMCS_CHECK_HIP(hipStreamSynchronize(s));
    double rr = h->host_scal[S_RR], tol2 = h->host_scal[S_TOL2];
    if (rr <= tol2 || !(rr == rr) || launched >= max_iter) break;
    batch = 1;
  }
  double rr = h->host_scal[S_RR], bb = h->host_scal[S_BB];
  int it = (int)h->host_scal[S_ITERS];
  h->pred_iters = it;
  if (iters) *iters = it;
  double rel = bb > 0 ? sqrt(rr / bb) : sqrt(rr);
  if (relres) *relres = rel;
  if (!(rr <= h->host_scal[S_TOL2])) {
    set_error("Poisson MG-PCG: relative residual %.3e above tolerance %.3e after %d iterations",
              rel, rtol, it);
    return BEMCS_ENOCONV;
  }
  return BEMCS_OK;
}

}  // namespace bemcs

using namespace bemcs;

extern "C" {

int bemcs_poisson_create(const uint8_t* is_bound_host, int32_t nx, int32_t ny, int32_t periodic_y,
                         bemcs_poisson_t* out) {
  BEMCS_REQUIRE(is_bound_host && out && nx >= 3 && ny >= 3, "bemcs_poisson_create: bad arguments");
  *out = nullptr;
  const int n = nx * ny;
  const bool per = periodic_y != 0;
  // node kinds, as build_sparse_matrix assigns them
  std::vector<uint8_t> kind(n, K_UNKNOWN);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      int i = y * nx + x;
      bool b = is_bound_host[i] != 0;
      if (b) kind[i] = K_BOUND;
      else if (x == nx - 1) kind[i] = K_RIGHT;
      else if (!per && y == ny - 1) kind[i] = K_TOP;
      else if (!per && y == 0) kind[i] = K_BOTTOM;
    }
  auto mirror1_of = [&](int i) -> int {
    int x = i % nx, y = i / nx;
    if (kind[i] == K_RIGHT) return y * nx + (nx - 2);
    if (kind[i] == K_TOP) return (ny - 2) * nx + x;
    if (kind[i] == K_BOTTOM) return nx + x;
    return -1;
  };
  std::vector<int32_t> mirror(n, -1), mirror1(n, -1);
  for (int i = 0; i < n; ++i) {
    int m = mirror1_of(i);
    if (m < 0) continue;
    mirror1[i] = m;
    int hops = 0;
    while (kind[m] != K_UNKNOWN && kind[m] != K_BOUND && hops < 4) {
      m = mirror1_of(m);
      ++hops;
    }
    BEMCS_REQUIRE(hops < 2, "bemcs_poisson_create: Neumann chain longer than two nodes");
    mirror[i] = m;
  }
  // fine operator: diagonal 4 - #Neumann neighbours at unknown nodes
  HostLevel L0;
  L0.nx = nx;
  L0.ny = ny;
  L0.periodic = per;
  L0.st.assign((size_t)9 * n, 0.0);
  L0.active.assign(n, 0);
  std::vector<uint8_t> code(n, 0);
  for (int i = 0; i < n; ++i) {
    if (kind[i] != K_UNKNOWN) continue;
    int x = i % nx, y = i / nx;
    BEMCS_REQUIRE(x > 0 && x < nx - 1, "bemcs_poisson_create: unknown node on x edge");
    int s = y > 0 ? i - nx : (per ? i + (ny - 1) * nx : -1);
    int nn = y < ny - 1 ? i + nx : (per ? i - (ny - 1) * nx : -1);
    BEMCS_REQUIRE(s >= 0 && nn >= 0, "bemcs_poisson_create: unknown node on y edge");
    int nb[4] = {i - 1, i + 1, s, nn};
    int kk[4] = {3, 5, 1, 7};
    int d = 4;
    for (int k = 0; k < 4; ++k) {
      uint8_t t = kind[nb[k]];
      if (t == K_RIGHT || t == K_TOP || t == K_BOTTOM) {
        BEMCS_REQUIRE(mirror1[nb[k]] == i, "bemcs_poisson_create: Neumann neighbour mirrors another node");
        --d;
      } else if (t == K_UNKNOWN) {
        L0.st[(size_t)kk[k] * n + i] = -1.0;
      }
    }
    code[i] = (uint8_t)d;
    L0.st[(size_t)4 * n + i] = d;
    L0.active[i] = 1;
  }
  std::vector<HostLevel> H;
  H.push_back(std::move(L0));
  while ((int)H.size() < kMaxLevels) {
    const HostLevel& F = H.back();
    int ncx = (F.nx + 1) / 2, ncy = (F.ny + 1) / 2;
    if (F.nx * F.ny <= kCoarseMax) break;
    if (ncx < 3 || ncy < (per ? 4 : 3)) break;
    H.push_back(coarsen(F));
  }
  BEMCS_REQUIRE(H.back().nx * H.back().ny <= kDenseMax,
                "bemcs_poisson_create: coarsest level too large for the dense solve");

  bemcs_poisson_s* h = new bemcs_poisson_s();
  h->nx = nx;
  h->ny = ny;
  h->periodic = per ? 1 : 0;
  h->nlev = (int)H.size();
  h->host_scal = nullptr;
  h->have_guess = false;
  h->pred_iters = 8;
  const char* ps = getenv("BEMCS_MG_POST_SWEEPS");
  h->post_sweeps = (ps && ps[0] == '1') ? 1 : 2;
  bool ok = true;
  auto A = [&](size_t bytes) {
    void* p = dalloc(h, bytes);
    if (!p) ok = false;
    return p;
  };
  for (int l = 0; l < h->nlev && ok; ++l) {
    const HostLevel& HL = H[l];
    LevelDev& D = h->lev[l];
    D.nx = HL.nx;
    D.ny = HL.ny;
    D.n = HL.nx * HL.ny;
    D.periodic = per ? 1 : 0;
    std::vector<float> inv(D.n, 0.0f);
    for (int i = 0; i < D.n; ++i)
      if (HL.active[i]) inv[i] = (float)(1.0 / HL.st[(size_t)4 * D.n + i]);
    float* dinv = (float*)A(sizeof(float) * D.n);
    if (!ok) break;
    ok = upload(dinv, inv.data(), sizeof(float) * D.n) == 0;
    D.inv_d = dinv;
    if (l == 0) {
      uint8_t* dc = (uint8_t*)A(D.n);
      if (!ok) break;
      ok = upload(dc, code.data(), D.n) == 0;
      D.code = dc;
      D.st = nullptr;
    } else {
      std::vector<float> st((size_t)9 * D.n);
      for (size_t k = 0; k < st.size(); ++k) st[k] = (float)HL.st[k];
      float* ds = (float*)A(sizeof(float) * st.size());
      if (!ok) break;
      ok = upload(ds, st.data(), sizeof(float) * st.size()) == 0;
      D.st = ds;
      D.code = nullptr;
    }
    D.x = (float*)A(sizeof(float) * D.n);
    D.t = (float*)A(sizeof(float) * D.n);
    D.b = (float*)A(sizeof(float) * D.n);
    if (ok) ok = hipMemset(D.x, 0, sizeof(float) * D.n) == hipSuccess;
    if (ok) ok = hipMemset(D.t, 0, sizeof(float) * D.n) == hipSuccess;
    if (ok) ok = hipMemset(D.b, 0, sizeof(float) * D.n) == hipSuccess;
  }
  if (ok) {
    std::vector<float> cinv = dense_inverse(H.back());
    h->nc = H.back().nx * H.back().ny;
    h->cinv = (float*)A(sizeof(float) * cinv.size());
    if (ok) ok = upload(h->cinv, cinv.data(), sizeof(float) * cinv.size()) == 0;
    h->kind = (uint8_t*)A(n);
    if (ok) ok = upload(h->kind, kind.data(), n) == 0;
    h->mirror = (int32_t*)A(sizeof(int32_t) * n);
    if (ok) ok = upload(h->mirror, mirror.data(), sizeof(int32_t) * n) == 0;
    h->mirror1 = (int32_t*)A(sizeof(int32_t) * n);
    if (ok) ok = upload(h->mirror1, mirror1.data(), sizeof(int32_t) * n) == 0;
    h->x = (double*)A(sizeof(double) * n);
    h->r = (double*)A(sizeof(double) * n);
    h->p = (double*)A(sizeof(double) * n);
    h->Ap = (double*)A(sizeof(double) * n);
    h->rold = (double*)A(sizeof(double) * n);
    h->bred = (double*)A(sizeof(double) * n);
    h->Vnew = (double*)A(sizeof(double) * n);
    h->part = (double*)A(sizeof(double) * 2 * kMaxRedBlocks);
    h->scal = (double*)A(sizeof(double) * S_COUNT);
    h->done = (int32_t*)A(sizeof(int32_t) * 4);
    h->rscr = (float*)A(sizeof(float) * n);
    h->part2 = (double*)A(sizeof(double) * 2 * kMaxRedBlocks);
    if (ok) ok = hipMemset(h->done, 0, sizeof(int32_t) * 4) == hipSuccess;
    if (ok) ok = hipMemset(h->p, 0, sizeof(double) * n) == hipSuccess;
    if (ok) ok = hipMemset(h->Ap, 0, sizeof(double) * n) == hipSuccess;
    if (ok) ok = hipMemset(h->x, 0, sizeof(double) * n) == hipSuccess;
    if (ok) ok = hipMemset(h->scal, 0, sizeof(double) * S_COUNT) == hipSuccess;
    if (ok) ok = hipHostMalloc((void**)&h->host_scal, sizeof(double) * S_COUNT, 0) == hipSuccess;
  }
  if (!ok) {
    destroy(h);
    set_error("bemcs_poisson_create: device allocation or upload failed");
    return BEMCS_EHIP;
  }
  *out = h;
  return BEMCS_OK;
}

int bemcs_poisson_free(bemcs_poisson_t h) {
  destroy(h);
  return BEMCS_OK;
}

int bemcs_poisson_levels(bemcs_poisson_t h) { return h ? h->nlev : 0; }

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
  int rc = pcg(h, rtol, max_iter, &it, &rel, s);
  if (iters) *iters = it;
  if (relres) *relres = rel;
  hipLaunchKernelGGL(k_assemble, dim3(grid_for(n, 256)), dim3(256), 0, s, h->x, b, h->kind,
                     h->mirror, h->mirror1, V, n);
  BEMCS_CHECK_LAUNCH();
  return rc;
}

int bemcs_poisson_picard(bemcs_poisson_t h, const bemcs_picard_t* p, int32_t* iters_total,
                         double* worst_relres, bemcs_stream_t stream) {
  BEMCS_REQUIRE(h && p && p->cnt_ion && p->cnt_e && p->V_fixed && p->V && p->rho_out && p->Exy,
                "bemcs_poisson_picard: bad arguments");
  hipStream_t s = as_stream(stream);
  const int n = h->nx * h->ny;
  const unsigned g = grid_for(n, 256);
  hipLaunchKernelGGL(k_rho_from_counts, dim3(g), dim3(256), 0, s, p->cnt_ion, p->cnt_e, n,
                     p->cd_ion, p->cd_e, p->rho_out);
  if (!h->have_guess) {
    hipLaunchKernelGGL(k_reduce_rhs, dim3(g), dim3(256), 0, s, p->V_fixed, p->V, h->kind, h->bred,
                       h->x, h->nx, h->ny, h->periodic, 1);
    h->have_guess = true;
  }
  BEMCS_CHECK_LAUNCH();
  int total = 0;
  double worst = 0;
  const double qn0 = -p->q * p->n0;
  for (int k = 0; k < p->iterations; ++k) {
    hipLaunchKernelGGL(k_picard_rhs, dim3(g), dim3(256), 0, s, p->rho_out, p->V, p->V_fixed,
                       h->kind, h->bred, h->nx, h->ny, h->periodic, p->coeff, p->V_plasma,
                       p->Te_up, qn0);
    BEMCS_CHECK_LAUNCH();
    int it = 0;
    double rel = 0;
    int rc = pcg(h, p->rtol, p->max_iter, &it, &rel, s);
    total += it;
    worst = rel > worst ? rel : worst;
    if (rc) {
      if (iters_total) *iters_total = total;
      if (worst_relres) *worst_relres = worst;
      return rc;
    }
    hipLaunchKernelGGL(k_picard_relax, dim3(g), dim3(256), 0, s, h->x, p->V_fixed, h->kind,
                       h->mirror, p->V, h->Vnew, n, 1.0 - p->omega, p->omega);
    BEMCS_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_efield, dim3(g), dim3(256), 0, s, p->V, h->nx, h->ny, p->dx_m, p->dy_m,
                     h->periodic, p->Exy);
  BEMCS_CHECK_LAUNCH();
  if (iters_total) *iters_total = total;
  if (worst_relres) *worst_relres = worst;
  return BEMCS_OK;
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
