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
// inverse on the coarsest level).  Levels at or below kTailNodes nodes run in a
// single work-group kernel so a V-cycle costs few launches.
#include <math.h>

#include <algorithm>
#include <vector>

#include "bemcs_internal.h"

namespace bemcs {

constexpr int kMaxLevels = 20;
constexpr int kTailNodes = 8192;
constexpr int kCoarseMax = 600;
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
  float* r;
};

struct TailArgs {
  LevelDev lev[kMaxLevels];
  int first, last;  // tail covers levels first..last (last = coarsest)
  const float* cinv;
  int nc;
};

}  // namespace bemcs

struct bemcs_poisson_s {
  int nx, ny, periodic, nlev, tail_first;
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
  bemcs::TailArgs* tail_dev;
};

namespace bemcs {

enum { S_RZ = 0, S_PAP, S_RR, S_BB, S_ALPHA, S_BETA, S_TOL2, S_ITERS, S_COUNT = 16 };

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
// Each phase is a device function over [begin, n) with stride, used both by
// the multi-block kernels and by the single-block tail kernel.
__device__ __forceinline__ void ph_jacobi_first(const LevelDev& L, float omega, int i) {
  L.x[i] = omega * L.inv_d[i] * L.b[i];
}
__device__ __forceinline__ void ph_jacobi(const LevelDev& L, float omega, int i) {
  float ax = applyL(L, L.x, i);
  L.t[i] = L.x[i] + omega * L.inv_d[i] * (L.b[i] - ax);
}
__device__ __forceinline__ void ph_residual(const LevelDev& L, int i) {
  float ax = applyL(L, L.x, i);
  L.r[i] = L.inv_d[i] != 0.0f ? L.b[i] - ax : 0.0f;
}
__device__ __forceinline__ void ph_restrict(const LevelDev& F, const LevelDev& C, int I) {
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
      s += wx * wy * F.r[fy * F.nx + fx];
    }
  }
  C.b[I] = C.inv_d[I] != 0.0f ? s : 0.0f;
}
__device__ __forceinline__ void ph_prolong(const LevelDev& F, const LevelDev& C, int i) {
  if (F.inv_d[i] == 0.0f) return;
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
  F.x[i] += s;
}

__global__ __launch_bounds__(256) void k_jacobi_first(LevelDev L, float omega, const int32_t* done) {
  if (*done) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < L.n) ph_jacobi_first(L, omega, i);
}
__global__ __launch_bounds__(256) void k_jacobi(LevelDev L, float omega, const int32_t* done) {
  if (*done) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < L.n) ph_jacobi(L, omega, i);
}
__global__ __launch_bounds__(256) void k_residual(LevelDev L, const int32_t* done) {
  if (*done) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < L.n) ph_residual(L, i);
}
__global__ __launch_bounds__(256) void k_restrict(LevelDev F, LevelDev C, const int32_t* done) {
  if (*done) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < C.n) ph_restrict(F, C, i);
}
__global__ __launch_bounds__(256) void k_prolong(LevelDev F, LevelDev C, const int32_t* done) {
  if (*done) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < F.n) ph_prolong(F, C, i);
}

// All levels first..last in one work-group: pre-smooth, residual, restrict
// down to the coarsest, dense solve, then prolong and post-smooth back up.
__global__ __launch_bounds__(1024) void k_vcycle_tail(const TailArgs* __restrict__ A,
                                                       const int32_t* done) {
  if (*done) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int first = A->first, last = A->last;
  for (int l = first; l < last; ++l) {
    LevelDev L = A->lev[l];
    float om = L.st ? kOmegaC : kOmega0;
    for (int i = tid; i < L.n; i += nt) ph_jacobi_first(L, om, i);
    __syncthreads();
    for (int s = 1; s < kNu; ++s) {
      for (int i = tid; i < L.n; i += nt) ph_jacobi(L, om, i);
      __syncthreads();
      for (int i = tid; i < L.n; i += nt) L.x[i] = L.t[i];
      __syncthreads();
    }
    for (int i = tid; i < L.n; i += nt) ph_residual(L, i);
    __syncthreads();
    LevelDev C = A->lev[l + 1];
    for (int i = tid; i < C.n; i += nt) ph_restrict(L, C, i);
    __syncthreads();
  }
  {
    LevelDev C = A->lev[last];
    const int nc = A->nc;
    for (int i = tid; i < nc; i += nt) {
      const float* row = A->cinv + (size_t)i * nc;
      float s = 0.0f;
      for (int j = 0; j < nc; ++j) s += row[j] * C.b[j];
      C.x[i] = s;
    }
    __syncthreads();
  }
  for (int l = last - 1; l >= first; --l) {
    LevelDev L = A->lev[l];
    LevelDev C = A->lev[l + 1];
    float om = L.st ? kOmegaC : kOmega0;
    for (int i = tid; i < L.n; i += nt) ph_prolong(L, C, i);
    __syncthreads();
    for (int s = 0; s < kNu; ++s) {
      for (int i = tid; i < L.n; i += nt) ph_jacobi(L, om, i);
      __syncthreads();
      for (int i = tid; i < L.n; i += nt) L.x[i] = L.t[i];
      __syncthreads();
    }
  }
}

// ----------------------------------------------------------- CG kernels
__device__ __forceinline__ void block_reduce2(double a, double b, double* part) {
  __shared__ double sa[kRedBlock / 64], sb[kRedBlock / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  if (lane_id() == 0) {
    sa[threadIdx.x / 64] = a;
    sb[threadIdx.x / 64] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0, tb = 0;
    for (int w = 0; w < kRedBlock / 64; ++w) {
      ta += sa[w];
      tb += sb[w];
    }
    part[2 * blockIdx.x] = ta;
    part[2 * blockIdx.x + 1] = tb;
  }
}

// grid-stride loops so the number of partials stays <= kMaxRedBlocks
#define GRID_LOOP(i, n) for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < (n); i += gridDim.x * kRedBlock)

// r = b - A x ; partials (r.r, b.b)
__global__ __launch_bounds__(kRedBlock) void k_cg_init_residual(const double* __restrict__ b,
                                                                  const double* __restrict__ x,
                                                                  double* r, const uint8_t* code,
                                                                  int nx, int ny, int per, int n,
                                                                  double* part) {
  double rr = 0, bb = 0;
  GRID_LOOP(i, n) {
    double ri = code[i] ? b[i] - apply5<double>(x, code, i, nx, ny, per != 0) : 0.0;
    r[i] = ri;
    rr += ri * ri;
    bb += b[i] * b[i];
  }
  block_reduce2(rr, bb, part);
}

// b0 (float) = r ; used before the first V-cycle
__global__ __launch_bounds__(256) void k_to_f32(const double* __restrict__ r, float* b0, int n,
                                                const int32_t* done) {
  if (*done) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) b0[i] = (float)r[i];
}

// p = z ; partial (r.z, 0)
__global__ __launch_bounds__(kRedBlock) void k_cg_p_init(const double* __restrict__ r,
                                                           const float* __restrict__ z, double* p,
                                                           int n, double* part) {
  double rz = 0;
  GRID_LOOP(i, n) {
    double zi = z[i];
    p[i] = zi;
    rz += r[i] * zi;
  }
  block_reduce2(rz, 0.0, part);
}

// Ap = A p ; partial (p.Ap)
__global__ __launch_bounds__(kRedBlock) void k_cg_apply(const double* __restrict__ p, double* Ap,
                                                          const uint8_t* code, int nx, int ny,
                                                          int per, int n, double* part,
                                                          const int32_t* done) {
  if (*done) return;
  double s = 0;
  GRID_LOOP(i, n) {
    double a = apply5<double>(p, code, i, nx, ny, per != 0);
    Ap[i] = a;
    s += p[i] * a;
  }
  block_reduce2(s, 0.0, part);
}

// x += a p ; r_old = r ; r -= a Ap ; b0 = f32(r) ; partial (r.r)
__global__ __launch_bounds__(kRedBlock) void k_cg_update(double* x, double* r, double* rold,
                                                           const double* __restrict__ p,
                                                           const double* __restrict__ Ap, float* b0,
                                                           int n, const double* scal, double* part,
                                                           const int32_t* done) {
  if (*done) return;
  const double a = scal[S_ALPHA];
  double rr = 0;
  GRID_LOOP(i, n) {
    x[i] += a * p[i];
    double ro = r[i];
    rold[i] = ro;
    double rn = ro - a * Ap[i];
    r[i] = rn;
    b0[i] = (float)rn;
    rr += rn * rn;
  }
  block_reduce2(rr, 0.0, part);
}

// partials (r.z, r_old.z)
__global__ __launch_bounds__(kRedBlock) void k_cg_dot_z(const double* __restrict__ r,
                                                          const double* __restrict__ rold,
                                                          const float* __restrict__ z, int n,
                                                          double* part, const int32_t* done) {
  if (*done) return;
  double a = 0, b = 0;
  GRID_LOOP(i, n) {
    double zi = z[i];
    a += r[i] * zi;
    b += rold[i] * zi;
  }
  block_reduce2(a, b, part);
}

// p = z + beta p
__global__ __launch_bounds__(256) void k_cg_p_update(double* p, const float* __restrict__ z, int n,
                                                     const double* scal, const int32_t* done) {
  if (*done) return;
  const double beta = scal[S_BETA];
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = (double)z[i] + beta * p[i];
}

// Fixed-order sum of the partials and the scalar recurrence.
// mode 0: after init residual   -> S_RR, S_BB, tolerance, done if converged
// mode 1: after p init          -> S_RZ
// mode 2: after apply           -> S_ALPHA = rz / pAp
// mode 3: after update          -> S_RR, done if converged, iteration count
// mode 4: after dot_z           -> S_BETA (Polak-Ribiere), S_RZ
__global__ __launch_bounds__(256) void k_cg_scalar(const double* part, int nparts, double* scal,
                                                   int32_t* done, int mode, double rtol) {
  __shared__ double sa[256], sb[256];
  if (mode != 0 && mode != 1 && *done) return;
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
  if (threadIdx.x != 0) return;
  a = sa[0];
  b = sb[0];
  if (mode == 0) {
    scal[S_RR] = a;
    scal[S_BB] = b;
    scal[S_TOL2] = rtol * rtol * b;
    scal[S_ITERS] = 0;
    *done = (a <= scal[S_TOL2]) ? 1 : 0;
  } else if (mode == 1) {
    scal[S_RZ] = a;
  } else if (mode == 2) {
    scal[S_PAP] = a;
    scal[S_ALPHA] = a != 0.0 ? scal[S_RZ] / a : 0.0;
  } else if (mode == 3) {
    scal[S_RR] = a;
    scal[S_ITERS] += 1.0;
    if (a <= scal[S_TOL2] || !(a == a)) *done = 1;
  } else {
    double rz_old = scal[S_RZ];
    double beta = rz_old != 0.0 ? (a - b) / rz_old : 0.0;
    scal[S_BETA] = beta;
    scal[S_RZ] = a;
  }
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

// Dense inverse of the active block (Gauss-Jordan with partial pivoting).
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
    for (int b = 0; b < m; ++b) out[(size_t)act[a] * n + act[b]] = (float)Inv[(size_t)a * m + b];
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
static void vcycle(bemcs_poisson_s* h, hipStream_t s) {
  const int tf = h->tail_first;
  for (int l = 0; l < tf; ++l) {
    LevelDev& L = h->lev[l];
    float om = L.st ? kOmegaC : kOmega0;
    unsigned g = grid_for(L.n, 256);
    hipLaunchKernelGGL(k_jacobi_first, dim3(g), dim3(256), 0, s, L, om, h->done);
    for (int k = 1; k < kNu; ++k) {
      hipLaunchKernelGGL(k_jacobi, dim3(g), dim3(256), 0, s, L, om, h->done);
      std::swap(L.x, L.t);
    }
    hipLaunchKernelGGL(k_residual, dim3(g), dim3(256), 0, s, L, h->done);
    LevelDev& C = h->lev[l + 1];
    hipLaunchKernelGGL(k_restrict, dim3(grid_for(C.n, 256)), dim3(256), 0, s, L, C, h->done);
  }
  hipLaunchKernelGGL(k_vcycle_tail, dim3(1), dim3(1024), 0, s, (const TailArgs*)h->tail_dev, h->done);
  for (int l = tf - 1; l >= 0; --l) {
    LevelDev& L = h->lev[l];
    LevelDev& C = h->lev[l + 1];
    float om = L.st ? kOmegaC : kOmega0;
    unsigned g = grid_for(L.n, 256);
    hipLaunchKernelGGL(k_prolong, dim3(g), dim3(256), 0, s, L, C, h->done);
    for (int k = 0; k < kNu; ++k) {
      hipLaunchKernelGGL(k_jacobi, dim3(g), dim3(256), 0, s, L, om, h->done);
      std::swap(L.x, L.t);
    }
  }
}

// The level-0 float32 V-cycle output (z) after vcycle(); the ping-pong swaps
// are even per level, so it is always lev[0].x.
static inline float* vc_out(bemcs_poisson_s* h) { return h->lev[0].x; }

static unsigned red_grid(int n) {
  unsigned g = grid_for(n, kRedBlock);
  return g > (unsigned)kMaxRedBlocks ? (unsigned)kMaxRedBlocks : g;
}

// PCG on the reduced system already in h->bred with initial guess h->x.
static int pcg(bemcs_poisson_s* h, double rtol, int max_iter, int* iters, double* relres,
               hipStream_t s) {
  const int n = h->nx * h->ny;
  const unsigned rg = red_grid(n);
  const unsigned g = grid_for(n, 256);
  LevelDev& L0 = h->lev[0];
  BEMCS_CHECK_HIP(hipMemsetAsync(h->done, 0, sizeof(int32_t), s));
  hipLaunchKernelGGL(k_cg_init_residual, dim3(rg), dim3(kRedBlock), 0, s, h->bred, h->x, h->r,
                     L0.code, h->nx, h->ny, h->periodic, n, h->part);
  hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, s, h->part, (int)rg, h->scal, h->done, 0,
                     rtol);
  hipLaunchKernelGGL(k_to_f32, dim3(g), dim3(256), 0, s, h->r, L0.b, n, h->done);
  vcycle(h, s);
  hipLaunchKernelGGL(k_cg_p_init, dim3(rg), dim3(kRedBlock), 0, s, h->r, vc_out(h), h->p, n, h->part);
  hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, s, h->part, (int)rg, h->scal, h->done, 1,
                     rtol);
  BEMCS_CHECK_LAUNCH();
  int it = 0;
  const int group = 4;
  while (it < max_iter) {
    int ng = std::min(group, max_iter - it);
    for (int k = 0; k < ng; ++k) {
      hipLaunchKernelGGL(k_cg_apply, dim3(rg), dim3(kRedBlock), 0, s, h->p, h->Ap, L0.code, h->nx,
                         h->ny, h->periodic, n, h->part, h->done);
      hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, s, h->part, (int)rg, h->scal, h->done,
                         2, rtol);
      hipLaunchKernelGGL(k_cg_update, dim3(rg), dim3(kRedBlock), 0, s, h->x, h->r, h->rold, h->p,
                         h->Ap, L0.b, n, h->scal, h->part, h->done);
      hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, s, h->part, (int)rg, h->scal, h->done,
                         3, rtol);
      vcycle(h, s);
      hipLaunchKernelGGL(k_cg_dot_z, dim3(rg), dim3(kRedBlock), 0, s, h->r, h->rold, vc_out(h), n,
                         h->part, h->done);
      hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, s, h->part, (int)rg, h->scal, h->done,
                         4, rtol);
      hipLaunchKernelGGL(k_cg_p_update, dim3(g), dim3(256), 0, s, h->p, vc_out(h), n, h->scal,
                         h->done);
    }
    BEMCS_CHECK_LAUNCH();
    it += ng;
    BEMCS_CHECK_HIP(hipMemcpyAsync(h->host_scal, h->scal, sizeof(double) * S_COUNT,
                                   hipMemcpyDeviceToHost, s));
    BEMCS_CHECK_HIP(hipStreamSynchronize(s));
    double rr = h->host_scal[S_RR], tol2 = h->host_scal[S_TOL2];
    if (rr <= tol2 || !(rr == rr)) break;
  }
  BEMCS_CHECK_HIP(hipMemcpyAsync(h->host_scal, h->scal, sizeof(double) * S_COUNT,
                                 hipMemcpyDeviceToHost, s));
  BEMCS_CHECK_HIP(hipStreamSynchronize(s));
  double rr = h->host_scal[S_RR], bb = h->host_scal[S_BB];
  if (iters) *iters = (int)h->host_scal[S_ITERS];
  double rel = bb > 0 ? sqrt(rr / bb) : sqrt(rr);
  if (relres) *relres = rel;
  if (!(rr <= h->host_scal[S_TOL2])) {
    set_error("Poisson MG-PCG: relative residual %.3e above tolerance %.3e after %d iterations",
              rel, rtol, (int)h->host_scal[S_ITERS]);
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
  BEMCS_REQUIRE(H.back().nx * H.back().ny <= 4096,
                "bemcs_poisson_create: coarsest level too large for the dense solve");

  bemcs_poisson_s* h = new bemcs_poisson_s();
  h->nx = nx;
  h->ny = ny;
  h->periodic = per ? 1 : 0;
  h->nlev = (int)H.size();
  h->host_scal = nullptr;
  h->have_guess = false;
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
    D.r = (float*)A(sizeof(float) * D.n);
    if (ok) ok = hipMemset(D.x, 0, sizeof(float) * D.n) == hipSuccess;
  }
  if (ok) {
    int tf = h->nlev - 1;
    while (tf > 0 && h->lev[tf - 1].n <= kTailNodes) --tf;
    h->tail_first = tf;
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
    h->tail_dev = (TailArgs*)A(sizeof(TailArgs));
    if (ok) ok = hipMemset(h->x, 0, sizeof(double) * n) == hipSuccess;
    if (ok) ok = hipMemset(h->scal, 0, sizeof(double) * S_COUNT) == hipSuccess;
    if (ok) ok = hipHostMalloc((void**)&h->host_scal, sizeof(double) * S_COUNT, 0) == hipSuccess;
    if (ok) {
      TailArgs ta;
      for (int l = 0; l < h->nlev; ++l) ta.lev[l] = h->lev[l];
      ta.first = h->tail_first;
      ta.last = h->nlev - 1;
      ta.cinv = h->cinv;
      ta.nc = h->nc;
      ok = upload(h->tail_dev, &ta, sizeof(TailArgs)) == 0;
    }
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
