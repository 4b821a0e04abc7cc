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
  double *x, *p, *Ap, *bred, *Vnew;
  double* R[2];       // residuals r_k, alternating
  int nparts;         // partial pairs per reduction buffer (level-0 tiles)
  double *pa, *pr, *pi;
  double* pz[2];
  double* scal;       // device scalars, see S_* below
  int32_t* done;
  double* host_scal;  // pinned copy
  bool have_guess;
  int pred_iters;     // iterations the last solve needed (launch that many first)
  int post_sweeps;    // Jacobi sweeps after prolongation (1 or 2)
  bool use_graphs;    // PCG iterations replayed from captured HIP graphs
  hipStream_t cap_stream;                  // private stream used only for capture
  std::vector<hipGraphExec_t> iter_graph;  // [k] = iteration k (dir + V-cycle)
  hipGraphExec_t init_graph;               // initial V-cycle
};

namespace bemcs {

enum { S_RR = 0, S_BB, S_TOL2, S_ITERS, S_RTOL, S_COUNT = 8 };

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
constexpr int kTCc = 8;             // the same on the coarser levels (more,
constexpr int kTFc = 2 * kTCc;      // smaller tiles: these launches are latency bound)

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
  double *p, *Ap;
  const double* bred;
  float* b0;       // level-0 right-hand side of the V-cycle
  const float* z;  // level-0 V-cycle output
  const uint8_t* code;
  int nx, ny, per, n;
  int nparts;      // partial pairs per buffer (one per level-0 tile)
  double* pa;      // p.Ap
  double* pr;      // r.r
  double* pz[2];   // (r_k.z_k, r_{k-1}.z_k) of V-cycle k in pz[k&1]
  double* pi;      // init: (r0.r0, b.b)
  double* scal;
  int32_t* done;
};

struct CheckArgs {
  const double* pr;
  int nparts;
  double* scal;
  int32_t* done;
  int k;  // iteration whose step precedes this launch; 0 = no check
};

// Convergence test of iteration ca.k (every block computes it identically;
// all NT threads of the block take part).
template <int NT = 256>
__device__ __forceinline__ bool check_converged(const CheckArgs& ca) {
  double rr, unused;
  block_sum2<NT>(ca.pr, ca.nparts, rr, unused);
  const bool stop = !(rr > ca.scal[S_TOL2]);  // converged, or NaN
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    ca.scal[S_RR] = rr;
    ca.scal[S_ITERS] = (double)ca.k;
    if (stop) *ca.done = 1;
  }
  return stop;
}

// Launches that run a convergence test can see `done` change while they run
// (block 0 sets it); a block reads it once and shares it, so all of its waves
// take the same branch around the barriers that follow.
__device__ __forceinline__ bool block_done(const int32_t* done) {
  __shared__ int s;
  if (threadIdx.x == 0) s = *(volatile const int32_t*)done;
  __syncthreads();
  return s != 0;
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

template <int TC, bool IS9, bool STEP>
__global__ __launch_bounds__(256) void k_down(LevelDev F, LevelDev C, CGArgs g, int k,
                                              CheckArgs ca) {
  if (ca.k > 0) {
    if (block_done(ca.done) || check_converged(ca)) return;
  } else if (*ca.done) {
    return;
  }
  constexpr int TFd = 2 * TC;                 // owned fine nodes per edge
  constexpr int W1 = TFd + 5, W2 = W1 - 2, W3 = W1 - 4;
  constexpr int N1 = W1 * W1, N2 = W2 * W2, N3 = W3 * W3;
  constexpr int IT1 = (N1 + 255) / 256, IT2 = (N2 + 255) / 256;
  __shared__ float sb[N1], sid[N1], s1[N1], s2[N2], sr[N3];
  __shared__ float sst[IS9 ? 9 * N2 : 1];
  __shared__ uint8_t scode[IS9 ? 1 : N2];
  const float om = IS9 ? kOmegaC : kOmega0;
  const int CX0 = blockIdx.x * TC, CY0 = blockIdx.y * TC;
  const int ox = 2 * CX0 - 3, oy = 2 * CY0 - 3;  // fine coordinates of local (0, 0)
  const double* rp = g.R[(k + 1) & 1];
  double alpha = 0.0;
  if (STEP) {
    double rz, pap, u;
    block_sum2(g.pz[(k + 1) & 1], g.nparts, rz, u);
    block_sum2(g.pa, g.nparts, pap, u);
    alpha = pap != 0.0 ? rz / pap : 0.0;
  }
  {
    int vg[IT1];
    float vid[IT1], vb[IT1];
    double vr[STEP ? IT1 : 1], va[STEP ? IT1 : 1];
#pragma unroll
    for (int it = 0; it < IT1; ++it) {
      const int t = min((int)threadIdx.x + it * 256, N1 - 1);
      const int gi = gnode(F, ox + t % W1, oy + t / W1);
      const int gc = gi < 0 ? 0 : gi;
      vg[it] = gi;
      vid[it] = F.inv_d[gc];
      if (STEP) {
        vr[it] = rp[gc];
        va[it] = g.Ap[gc];
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
#pragma unroll
    for (int it = 0; it < IT1; ++it) {
      const int t = threadIdx.x + it * 256;
      if (t >= N1) break;
      const bool in = vg[it] >= 0;
      const float id = in ? vid[it] : 0.0f;
      float b = 0.0f;
      if (in) b = STEP ? (float)(vr[it] - alpha * va[it]) : vb[it];
      sb[t] = b;
      sid[t] = id;
      s1[t] = om * id * b;
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
    double* rn_out = g.R[k & 1];
#pragma unroll
    for (int it = 0; it < TFd * TFd / 256; ++it) {
      const int t = threadIdx.x + it * 256;
      const int gx = 2 * CX0 + t % TFd, gy = 2 * CY0 + t / TFd;
      if (gx >= F.nx || gy >= F.ny) continue;
      const int gi = gy * F.nx + gx;
      const double rn = rp[gi] - alpha * g.Ap[gi];
      g.x[gi] += alpha * g.p[gi];
      rn_out[gi] = rn;
      F.b[gi] = (float)rn;
      rr += rn * rn;
    }
    block_partials(rr, 0.0, g.pr, blockIdx.y * gridDim.x + blockIdx.x);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < N2; t += 256) {
    const int lx = t % W2 + 1, ly = t / W2 + 1;
    const int li = ly * W1 + lx;
    const float id = sid[li];
    float v = 0.0f;
    if (id != 0.0f) {
      v = s1[li] + om * id * (sb[li] - tile_apply<IS9>(&s1[li], W1, sst, N2, t, IS9 ? 0 : scode[t]));
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

// -------------------------------------------------------------- up launch
// The up to four (index, weight) terms of P e at fine node i; absent terms
// have weight 0 and index 0.
__device__ __forceinline__ void pe_terms(const LevelDev& F, const LevelDev& C, int i, int* idx,
                                         float* w) {
  const int fx = i % F.nx, fy = i / F.nx;
  int xs[2] = {fx >> 1, -1}, ys[2] = {fy >> 1, -1};
  if (fx & 1) {
    xs[0] = (fx - 1) >> 1;
    xs[1] = (fx + 1) >> 1;
    if (xs[1] >= C.nx) xs[1] = -1;
  }
  if (fy & 1) {
    ys[0] = (fy - 1) >> 1;
    ys[1] = (fy + 1) >> 1;
    if (ys[1] >= C.ny) ys[1] = F.periodic ? ys[1] - C.ny : -1;
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const bool ok = ys[a] >= 0 && xs[b] >= 0;
      idx[2 * a + b] = ok ? ys[a] * C.nx + xs[b] : 0;
      w[2 * a + b] = ok ? w1d(fy, ys[a], C.ny, F.periodic != 0) * w1d(fx, xs[b], C.nx, false) : 0.0f;
    }
}

// C.x must point at the coarse result (C.t, or C.x of the coarsest level).
template <int TF, int SWEEPS, bool IS9, bool DOT>
__global__ __launch_bounds__(256) void k_up(LevelDev F, LevelDev C, CGArgs g, int k,
                                            const int32_t* __restrict__ done) {
  if (*done) return;
  constexpr int W0 = TF + 2 * SWEEPS, W1 = W0 - 2;  // the coefficient region is W1
  constexpr int N0 = W0 * W0, N1 = W1 * W1;
  constexpr int IT0 = (N0 + 255) / 256, IT1 = (N1 + 255) / 256;
  __shared__ float sx[N0], sb[N0], sid[N0], s1[SWEEPS == 2 ? N1 : 1];
  __shared__ float sst[IS9 ? 9 * N1 : 1];
  __shared__ uint8_t scode[IS9 ? 1 : N1];
  const float om = IS9 ? kOmegaC : kOmega0;
  const int ox = blockIdx.x * TF - SWEEPS, oy = blockIdx.y * TF - SWEEPS;
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
  __syncthreads();
  if (SWEEPS == 2) {
    for (int t = threadIdx.x; t < N1; t += 256) {
      const int lx = t % W1 + 1, ly = t / W1 + 1;
      const int li = ly * W0 + lx;
      const float id = sid[li];
      float v = 0.0f;
      if (id != 0.0f)
        v = sx[li] + om * id * (sb[li] - tile_apply<IS9>(&sx[li], W0, sst, N1, t, IS9 ? 0 : scode[t]));
      s1[t] = v;
    }
    __syncthreads();
  }
  double rz = 0.0, roz = 0.0;
  for (int t = threadIdx.x; t < TF * TF; t += 256) {
    const int lx = t % TF + SWEEPS, ly = t / TF + SWEEPS;
    const int gx = ox + lx, gy = oy + ly;
    if (gx >= F.nx || gy >= F.ny) continue;
    const int gi = gy * F.nx + gx;
    const int li = ly * W0 + lx;
    const int t1 = (ly - 1) * W1 + (lx - 1);  // index in the coefficient region
    const float id = sid[li];
    float v = 0.0f;
    if (id != 0.0f) {
      const float* c = SWEEPS == 2 ? &s1[t1] : &sx[li];
      const int W = SWEEPS == 2 ? W1 : W0;
      v = c[0] + om * id * (sb[li] - tile_apply<IS9>(c, W, sst, N1, t1, IS9 ? 0 : scode[t1]));
    }
    F.t[gi] = v;
    if (DOT) {
      rz += g.R[k & 1][gi] * (double)v;
      roz += g.R[(k + 1) & 1][gi] * (double)v;
    }
  }
  if (DOT) block_partials(rz, roz, g.pz[k & 1], blockIdx.y * gridDim.x + blockIdx.x);
}

// coarsest level: x = C^-1 b with the inverse stored column-major.  A block
// owns 64 rows (one per lane); its 16 waves split the columns, and the partial
// sums are added in LDS, so each thread runs only nc/16 independent loads.
constexpr int kCoarseWaves = 16;
__global__ __launch_bounds__(64 * kCoarseWaves) void k_coarse_solve(LevelDev C,
                                                                    const float* __restrict__ cinv,
                                                                    CheckArgs ca) {
  __shared__ float sb[kDenseMax];
  __shared__ float part[kCoarseWaves][64];
  if (ca.k > 0) {
    if (block_done(ca.done) || check_converged<64 * kCoarseWaves>(ca)) return;
  } else if (*ca.done) {
    return;
  }
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
    for (int q = 0; q < kCoarseWaves; ++q) t += part[q][lane];
    C.x[row] = t;
  }
}

// ----------------------------------------------------------- CG kernels
// grid-stride loops over g.nparts blocks, one partial pair per block
#define GRID_LOOP(i, n) for (int i = blockIdx.x * 256 + threadIdx.x; i < (n); i += gridDim.x * 256)

// r0 = b - A x -> R[0], R[1] = 0, b0 = f32(r0); partials (r0.r0, b.b)
__global__ __launch_bounds__(256) void k_cg_init(CGArgs g) {
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
__global__ __launch_bounds__(256) void k_cg_init_scalar(CGArgs g, double rtol) {
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

// p = z + beta p ; Ap = A z + beta Ap ; partials p.Ap
__global__ __launch_bounds__(256) void k_cg_dir(CGArgs g, int k) {
  if (*g.done) return;
  double beta = 0.0;
  if (k > 1) {
    double rz, roz, rz_old, u;
    block_sum2(g.pz[(k + 1) & 1], g.nparts, rz, roz);
    block_sum2(g.pz[k & 1], g.nparts, rz_old, u);
    beta = rz_old != 0.0 ? (rz - roz) / rz_old : 0.0;
  }
  double s = 0;
  GRID_LOOP(i, g.n) {
    double az = 0.0;
    const int d = g.code[i];
    if (d) {
      const int x = i % g.nx, y = i / g.nx;
      const int sn = y > 0 ? i - g.nx : (g.per ? i + (g.ny - 1) * g.nx : i);
      const int nn = y < g.ny - 1 ? i + g.nx : (g.per ? i - (g.ny - 1) * g.nx : i);
      const double nb = (double)g.z[i - 1] + (double)g.z[i + 1] + (double)g.z[sn] + (double)g.z[nn];
      az = (double)d * (double)g.z[i] - nb;
      (void)x;
    }
    const double pi = k > 1 ? (double)g.z[i] + beta * g.p[i] : (double)g.z[i];
    const double api = k > 1 ? az + beta * g.Ap[i] : az;
    g.p[i] = pi;
    g.Ap[i] = api;
    s += pi * api;
  }
  block_partials(s, 0.0, g.pa, blockIdx.x);
}

// single-level hierarchy: the step without a fused down sweep
__global__ __launch_bounds__(256) void k_cg_step(CGArgs g, int k) {
  if (*g.done) return;
  double rz, pap, u;
  block_sum2(g.pz[(k + 1) & 1], g.nparts, rz, u);
  block_sum2(g.pa, g.nparts, pap, u);
  const double a = pap != 0.0 ? rz / pap : 0.0;
  const double* rp = g.R[(k + 1) & 1];
  double* rn_out = g.R[k & 1];
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
__global__ __launch_bounds__(256) void k_cg_dotz(CGArgs g, int k) {
  if (*g.done) return;
  double a = 0, b = 0;
  GRID_LOOP(i, g.n) {
    const double zi = g.z[i];
    a += g.R[k & 1][i] * zi;
    b += g.R[(k + 1) & 1][i] * zi;
  }
  block_partials(a, b, g.pz[k & 1], blockIdx.x);
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
  for (hipGraphExec_t e : h->iter_graph)
    if (e) (void)hipGraphExecDestroy(e);
  if (h->init_graph) (void)hipGraphExecDestroy(h->init_graph);
  if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->host_scal) (void)hipHostFree(h->host_scal);
  delete h;
}

// --------------------------------------------------------- V-cycle driver
static CGArgs cg_args(bemcs_poisson_s* h) {
  CGArgs g;
  g.x = h->x;
  g.R[0] = h->R[0];
  g.R[1] = h->R[1];
  g.p = h->p;
  g.Ap = h->Ap;
  g.bred = h->bred;
  g.b0 = h->lev[0].b;
  g.z = h->nlev == 1 ? h->lev[0].x : h->lev[0].t;
  g.code = h->lev[0].code;
  g.nx = h->nx;
  g.ny = h->ny;
  g.per = h->periodic;
  g.n = h->nx * h->ny;
  g.nparts = h->nparts;
  g.pa = h->pa;
  g.pr = h->pr;
  g.pz[0] = h->pz[0];
  g.pz[1] = h->pz[1];
  g.pi = h->pi;
  g.scal = h->scal;
  g.done = h->done;
  return g;
}

// V-cycle k: k = 0 is the initial one (its right-hand side b0 was written by
// k_cg_init); for k >= 1 the CG step of iteration k is fused into the first
// down launch and the launch after it runs the convergence test.
static void enqueue_vcycle(bemcs_poisson_s* h, const CGArgs& g, int k, hipStream_t s) {
  const int last = h->nlev - 1;
  CheckArgs ck{h->pr, h->nparts, h->scal, h->done, 0};
  if (last == 0) {
    if (k > 0) hipLaunchKernelGGL(k_cg_step, dim3(h->nparts), dim3(256), 0, s, g, k);
    ck.k = k;
    hipLaunchKernelGGL(k_coarse_solve, dim3((h->lev[0].n + 63) / 64), dim3(64 * kCoarseWaves), 0, s,
                       h->lev[0], (const float*)h->cinv, ck);
    hipLaunchKernelGGL(k_cg_dotz, dim3(h->nparts), dim3(256), 0, s, g, k);
    return;
  }
  for (int l = 0; l < last; ++l) {
    const LevelDev& F = h->lev[l];
    const LevelDev& C = h->lev[l + 1];
    CheckArgs c = ck;
    c.k = (l == 1 && k > 0) ? k : 0;
    if (l == 0) {
      dim3 grid((C.nx + kTC - 1) / kTC, (C.ny + kTC - 1) / kTC);
      if (k > 0) hipLaunchKernelGGL((k_down<kTC, false, true>), grid, dim3(256), 0, s, F, C, g, k, c);
      else hipLaunchKernelGGL((k_down<kTC, false, false>), grid, dim3(256), 0, s, F, C, g, k, c);
    } else {
      dim3 grid((C.nx + kTCc - 1) / kTCc, (C.ny + kTCc - 1) / kTCc);
      hipLaunchKernelGGL((k_down<kTCc, true, false>), grid, dim3(256), 0, s, F, C, g, k, c);
    }
  }
  ck.k = (last == 1 && k > 0) ? k : 0;
  hipLaunchKernelGGL(k_coarse_solve, dim3((h->lev[last].n + 63) / 64), dim3(64 * kCoarseWaves), 0,
                     s, h->lev[last], (const float*)h->cinv, ck);
  for (int l = last - 1; l >= 0; --l) {
    const LevelDev& F = h->lev[l];
    LevelDev C = h->lev[l + 1];
    if (l + 1 != last) C.x = C.t;  // the coarse level's result
    const bool two = h->post_sweeps > 1;
    if (l == 0) {
      dim3 grid((F.nx + kTF - 1) / kTF, (F.ny + kTF - 1) / kTF);
      if (two) hipLaunchKernelGGL((k_up<kTF, 2, false, true>), grid, dim3(256), 0, s, F, C, g, k, h->done);
      else hipLaunchKernelGGL((k_up<kTF, 1, false, true>), grid, dim3(256), 0, s, F, C, g, k, h->done);
    } else {
      dim3 grid((F.nx + kTFc - 1) / kTFc, (F.ny + kTFc - 1) / kTFc);
      if (two) hipLaunchKernelGGL((k_up<kTFc, 2, true, false>), grid, dim3(256), 0, s, F, C, g, k, h->done);
      else hipLaunchKernelGGL((k_up<kTFc, 1, true, false>), grid, dim3(256), 0, s, F, C, g, k, h->done);
    }
  }
}

// Captures what `enqueue` launches into an executable graph (on the private
// capture stream: nothing runs).  The launches' arguments are fixed per
// handle and iteration, so a graph is captured once and replayed.
template <typename F>
static hipError_t capture(bemcs_poisson_s* h, F enqueue, hipGraphExec_t* out) {
  *out = nullptr;
  hipError_t e = hipSuccess;
  if (!h->cap_stream) e = hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking);
  if (e != hipSuccess) return e;
  e = hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  enqueue(h->cap_stream);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(h->cap_stream, &g);
  if (e != hipSuccess) return e;
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  return e;
}

static int enqueue_init(bemcs_poisson_s* h, const CGArgs& g, double rtol, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_init, dim3(h->nparts), dim3(256), 0, s, g);
  hipLaunchKernelGGL(k_cg_init_scalar, dim3(1), dim3(256), 0, s, g, rtol);
  if (!h->use_graphs) {
    enqueue_vcycle(h, g, 0, s);
    return BEMCS_OK;
  }
  if (!h->init_graph)
    BEMCS_CHECK_HIP(capture(h, [&](hipStream_t cs) { enqueue_vcycle(h, g, 0, cs); }, &h->init_graph));
  BEMCS_CHECK_HIP(hipGraphLaunch(h->init_graph, s));
  return BEMCS_OK;
}

static int enqueue_iter(bemcs_poisson_s* h, const CGArgs& g, int k, hipStream_t s) {
  if (!h->use_graphs) {
    hipLaunchKernelGGL(k_cg_dir, dim3(h->nparts), dim3(256), 0, s, g, k);
    enqueue_vcycle(h, g, k, s);
    return BEMCS_OK;
  }
  if ((int)h->iter_graph.size() <= k) h->iter_graph.resize(k + 1, nullptr);
  if (!h->iter_graph[k]) {
    auto body = [&](hipStream_t cs) {
      hipLaunchKernelGGL(k_cg_dir, dim3(h->nparts), dim3(256), 0, cs, g, k);
      enqueue_vcycle(h, g, k, cs);
    };
    BEMCS_CHECK_HIP(capture(h, body, &h->iter_graph[k]));
  }
  BEMCS_CHECK_HIP(hipGraphLaunch(h->iter_graph[k], s));
  return BEMCS_OK;
}

// PCG on the reduced system already in h->bred with initial guess h->x.
static int pcg(bemcs_poisson_s* h, double rtol, int max_iter, int* iters, double* relres,
               hipStream_t s) {
  const CGArgs g = cg_args(h);
  int rc = enqueue_init(h, g, rtol, s);
  if (rc) return rc;
  BEMCS_CHECK_LAUNCH();
  // launch as many iterations as the previous solve needed, then check
  int launched = 0;
  int batch = std::max(1, h->pred_iters);
  for (;;) {
    for (int j = 0; j < batch && launched < max_iter; ++j) {
      rc = enqueue_iter(h, g, ++launched, s);
      if (rc) return rc;
    }
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
  const char* pg = getenv("BEMCS_POISSON_GRAPHS");
  h->use_graphs = !(pg && pg[0] == '0');
  h->cap_stream = nullptr;
  h->init_graph = nullptr;
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
    h->R[0] = (double*)A(sizeof(double) * n);
    h->R[1] = (double*)A(sizeof(double) * n);
    h->p = (double*)A(sizeof(double) * n);
    h->Ap = (double*)A(sizeof(double) * n);
    h->bred = (double*)A(sizeof(double) * n);
    h->Vnew = (double*)A(sizeof(double) * n);
    h->nparts = ((nx + kTF - 1) / kTF) * ((ny + kTF - 1) / kTF);
    h->pa = (double*)A(sizeof(double) * 2 * h->nparts);
    h->pr = (double*)A(sizeof(double) * 2 * h->nparts);
    h->pi = (double*)A(sizeof(double) * 2 * h->nparts);
    h->pz[0] = (double*)A(sizeof(double) * 2 * h->nparts);
    h->pz[1] = (double*)A(sizeof(double) * 2 * h->nparts);
    h->scal = (double*)A(sizeof(double) * S_COUNT);
    h->done = (int32_t*)A(sizeof(int32_t) * 4);
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
