// Electrode-surface physics of the time step (sim_mode Thermal / Erosion /
// Both) and the RF modulation of a grid potential:
//   k_erosion     sputter damage into the damage map and cell failure
//                 (Python/physics_engine.py:1419-1449)
//   k_see_commit  purge and append of SEE electrons staged by the ion step,
//                 after erosion (:1390-1417, :1534-1561)
//   k_thermal     impact heat + radiative cooling + 10 conduction sweeps +
//                 300 K floor + per-grid partial sums, one launch
//                 (:1359-1388, :1563-1617)
//   k_grid_means  per-grid mean temperatures from the partial sums
//   k_conduction_sweep  one conduction sweep with the drop-in signature of
//                 thermal_conduction_taichi (:147-171)
//   k_rf_set      V_fixed of the RF grid = Vdc + A sin(2 pi f t) (:984-991)
//   k_electrodes  electrode map, Dirichlet values and grid masks of the grids,
//                 bow included (:791-843)
//   k_grid_sums   per-grid temperature sums (grid means at a domain build)
#include <mutex>

#include "bemcs_internal.h"

// Scratch of bemcs_thermal_noalloc: the per-tile partial sums of the grid
// means, in a file-scope device array instead of a caller workspace, for
// callers that cannot allocate.  Each GPU holds its own copy until the process
// ends.  In host code `scratch` names a host-side stand-in, not the device
// array: host code uses only the address hipGetSymbolAddress returns for it on
// each device (thermal_scratch below).
constexpr int64_t kThermalScratchBytes = 1 << 20;
static __device__ __align__(256) unsigned char scratch[kThermalScratchBytes];

namespace bemcs {
namespace {

constexpr int kBlock = 256;
constexpr int kThTX = 64;       // thermal tile: 64 x 32 nodes per work-group
constexpr int kThTY = 32;
constexpr int kMaxSweeps = 16;  // halo of the LDS tile = number of sweeps
constexpr int kMaxGrids = 8;
constexpr double kHeatUnit = 1.0 / 1048576.0;    // 2^-20 K
constexpr double kDamageUnit = 1.0 / 16777216.0;  // 2^-24

__global__ __launch_bounds__(kBlock) void k_erosion(bemcs_erosion_t a) {
  if (a.diag->surface_hits == 0) return;
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n_nodes) return;
  const unsigned long long q = a.damage_acc[i];
  double d = a.damage_map[i];
  bool write = false;
  if (q) {
    d = d + (double)q * kDamageUnit;
    a.damage_acc[i] = 0;
    write = true;
  }
  if (d > a.thresh && a.is_bound[i]) {
    a.eroded_depth[i] += a.dy;
    a.is_bound[i] = 0;
    d = 0.0;
    write = true;
    atomicAdd(a.n_failed, 1u);
  }
  if (write) a.damage_map[i] = d;
}

__global__ __launch_bounds__(kBlock) void k_see_commit(bemcs_species_t st, bemcs_species_t eo,
                                                       const uint8_t* __restrict__ bound,
                                                       uint32_t* e_cnt_next, bemcs_ion_diag_t* diag,
                                                       int nx, int ny, float dx32, float dy32,
                                                       float Lx32, float Ly32) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t cnt = *st.count;
  const int64_t n = cnt < st.cap ? cnt : st.cap;
  if (i == 0 && n > (int64_t)gridDim.x * kBlock) atomicOr(&diag->overflow, 16u);
  const bool active = i < n;
  float ex = 0.f, ey = 0.f;
  bool keep = false;
  int key = -1;
  if (active) {
    ex = st.x[i];
    ey = st.y[i];
    int eix = round_i64_np(ex / dx32, 0, nx - 1);
    int eiy = round_i64_np(ey / dy32, 0, ny - 1);
    bool ehit = bound[eiy * nx + eix] != 0;
    bool eout = (ex < 0.0f) || (ex > Lx32) || (ey < 0.0f) || (ey > Ly32) || isnan(ex);
    keep = !(ehit || eout);
    if (keep) {
      int dix = round_i32_np(ex / dx32, 1, nx - 2);
      int diy = round_i32_np(ey / dy32, 1, ny - 2);
      key = diy * nx + dix;
    }
  }
  int64_t slot = wave_append(eo.count, keep);
  if (keep) {
    if (slot < eo.cap) {
      eo.x[slot] = ex;
      eo.y[slot] = ey;
      eo.vx[slot] = st.vx[i];
      eo.vy[slot] = st.vy[i];
      eo.vz[slot] = st.vz[i];
    } else {
      atomicOr(&diag->overflow, 1u);
      key = -1;
    }
  }
  wave_count(&diag->see_kept, keep);
  wave_hist_add(e_cnt_next, key);
}

// Temperature after this step's impact heat and, on electrode nodes, the
// radiative cooling T -= c (T^4 - 300^4) in float32 (T^4 from float64:
// T*T is exact, one rounding to float64 and one to float32).
__device__ __forceinline__ float thermal_source(float v, unsigned long long heat, bool is_bound,
                                                float cool32, float t300_4) {
  if (heat) v = (float)((double)v + (double)heat * kHeatUnit);
  if (is_bound) {
    double t = (double)v;
    double t2 = t * t;
    float t4 = (float)(t2 * t2);
    float dcool = cool32 * (t4 - t300_4);
    v = v - dcool;
  }
  return v;
}

// Per-grid sum and node count of the work-group, reduced in a fixed order.
__device__ void grid_partials(const double (&sum)[kMaxGrids], const double (&cnt)[kMaxGrids],
                              int n_grids, double* out, double* red) {
  const int wave = threadIdx.x / 64;
  for (int g = 0; g < n_grids; ++g) {
    double s = wave_sum(sum[g]);
    double c = wave_sum(cnt[g]);
    if (lane_id() == 0) {
      red[2 * wave] = s;
      red[2 * wave + 1] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double ts = 0.0, tc = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) {
        ts += red[2 * w];
        tc += red[2 * w + 1];
      }
      out[2 * g] = ts;
      out[2 * g + 1] = tc;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_thermal(bemcs_thermal_t a) {
  extern __shared__ unsigned char lds_raw[];
  __shared__ double red[2 * (kBlock / 64)];
  const int H = a.sweeps;
  const int EX = kThTX + 2 * H, EY = kThTY + 2 * H, NE = EX * EY;
  float* cur = reinterpret_cast<float*>(lds_raw);
  float* nxt = cur + NE;
  uint8_t* fl = reinterpret_cast<uint8_t*>(nxt + NE);
  const int nx = a.nx, ny = a.ny;
  const int tx0 = blockIdx.x * kThTX, ty0 = blockIdx.y * kThTY;
  const uint8_t* __restrict__ bound = a.is_bound;
  const unsigned long long* heat = reinterpret_cast<const unsigned long long*>(a.heat_acc);

  int any = 0;
  for (int p = threadIdx.x; p < kThTX * kThTY; p += kBlock) {
    int gx = tx0 + p % kThTX, gy = ty0 + p / kThTX;
    if (gx < nx && gy < ny && bound[gy * nx + gx]) any = 1;
  }
  any = __syncthreads_or(any);

  double gsum[kMaxGrids], gcnt[kMaxGrids];
#pragma unroll
  for (int g = 0; g < kMaxGrids; ++g) gsum[g] = gcnt[g] = 0.0;
  const int ng = a.n_grids;

  if (!any) {
    // no electrode node in the tile: only impact heat can change a node
    for (int p = threadIdx.x; p < kThTX * kThTY; p += kBlock) {
      int gx = tx0 + p % kThTX, gy = ty0 + p / kThTX;
      if (gx >= nx || gy >= ny) continue;
      int i = gy * nx + gx;
      float v = a.T_in[i];
      if (heat) v = thermal_source(v, heat[i], false, a.cool32, a.t300_4_32);
      a.T_out[i] = v;
      if (a.heat_clear) a.heat_clear[i] = 0;
      uint8_t bits = a.grid_bits ? a.grid_bits[i] : 0;
      for (int g = 0; g < ng; ++g)
        if (bits & (1u << g)) {
          gsum[g] += (double)v;
          gcnt[g] += 1.0;
        }
    }
  } else {
    // extended tile (halo H) into LDS with the pointwise source applied;
    // flag bit 0 = electrode, bit 1 = updated by conduction (interior electrode)
    for (int p = threadIdx.x; p < NE; p += kBlock) {
      int gx = tx0 - H + p % EX, gy = ty0 - H + p / EX;
      float v = 0.0f;
      uint8_t f = 0;
      if (gx >= 0 && gx < nx && gy >= 0 && gy < ny) {
        int i = gy * nx + gx;
        bool b = bound[i] != 0;
        v = thermal_source(a.T_in[i], heat ? heat[i] : 0ull, b, a.cool32, a.t300_4_32);
        if (b) {
          f = 1;
          if (gx >= 1 && gx <= nx - 2 && gy >= 1 && gy <= ny - 2) f = 3;
        }
      }
      cur[p] = v;
      fl[p] = f;
    }
    __syncthreads();
    // Jacobi sweeps; the outermost ring of the extended tile is never
    // updated, so after s sweeps the nodes at distance >= s from it are exact
    // and the own tile (distance H) is exact after H sweeps.
    const float fx = a.fo_x32, fy = a.fo_y32;
    for (int s = 0; s < H; ++s) {
      for (int p = threadIdx.x; p < NE; p += kBlock) {
        int ex = p % EX, ey = p / EX;
        float c = cur[p];
        float o = c;
        if ((fl[p] & 2) && ex > 0 && ex < EX - 1 && ey > 0 && ey < EY - 1) {
          float l = (fl[p - 1] & 1) ? cur[p - 1] : c;
          float r = (fl[p + 1] & 1) ? cur[p + 1] : c;
          float d = (fl[p - EX] & 1) ? cur[p - EX] : c;
          float u = (fl[p + EX] & 1) ? cur[p + EX] : c;
          float dT = fx * ((l - 2.0f * c) + r) + fy * ((d - 2.0f * c) + u);
          float t = c + dT;
          o = (300.0f > t) ? 300.0f : t;
        }
        nxt[p] = o;
      }
      __syncthreads();
      float* tmp = cur;
      cur = nxt;
      nxt = tmp;
    }
    for (int p = threadIdx.x; p < kThTX * kThTY; p += kBlock) {
      int ox = p % kThTX, oy = p / kThTX;
      int gx = tx0 + ox, gy = ty0 + oy;
      if (gx >= nx || gy >= ny) continue;
      int e = (oy + H) * EX + (ox + H);
      float v = cur[e];
      if (fl[e] & 1) v = (v < 300.0f) ? 300.0f : v;  // np.maximum: NaN stays
      int i = gy * nx + gx;
      a.T_out[i] = v;
      if (a.heat_clear) a.heat_clear[i] = 0;
      uint8_t bits = a.grid_bits ? a.grid_bits[i] : 0;
      for (int g = 0; g < ng; ++g)
        if (bits & (1u << g)) {
          gsum[g] += (double)v;
          gcnt[g] += 1.0;
        }
    }
  }
  if (ng > 0) {
    double* out = reinterpret_cast<double*>(a.workspace) +
                  (size_t)(blockIdx.y * gridDim.x + blockIdx.x) * (2 * ng);
    grid_partials(gsum, gcnt, ng, out, red);
  }
}

__global__ __launch_bounds__(kBlock) void k_grid_means(const double* __restrict__ part, int nblocks,
                                                       int n_grids, float* T_grids) {
  __shared__ double red[2 * (kBlock / 64)];
  for (int g = 0; g < n_grids; ++g) {
    double s = 0.0, c = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) {
      s += part[(size_t)b * 2 * n_grids + 2 * g];
      c += part[(size_t)b * 2 * n_grids + 2 * g + 1];
    }
    s = wave_sum(s);
    c = wave_sum(c);
    const int wave = threadIdx.x / 64;
    if (lane_id() == 0) {
      red[2 * wave] = s;
      red[2 * wave + 1] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double ts = 0.0, tc = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) {
        ts += red[2 * w];
        tc += red[2 * w + 1];
      }
      if (tc > 0.0) T_grids[g] = (float)(ts / tc);
    }
    __syncthreads();
  }
}

// One explicit conduction sweep with thermal_conduction_taichi's signature
// (Python/physics_engine.py:147-171) and the NumPy loop's semantics (:1598-1613):
// T_new = T everywhere except interior electrode nodes (mask == 1, not on the
// outer ring), which take max(T + dT, 300) with insulated (non-electrode)
// neighbours replaced by the node itself.  float32 in the loop's operation
// order: dT = Fo_x ((l - 2c) + r) + Fo_y ((d - 2c) + u).
__global__ __launch_bounds__(kBlock) void k_conduction_sweep(const float* __restrict__ T,
                                                             float* __restrict__ T_new,
                                                             const int32_t* __restrict__ mask,
                                                             int nx, int ny, float fx, float fy) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int64_t)nx * ny) return;
  const int ix = (int)(i % nx), iy = (int)(i / nx);
  const float c = T[i];
  float o = c;
  if (mask[i] == 1 && ix >= 1 && ix <= nx - 2 && iy >= 1 && iy <= ny - 2) {
    const float l = mask[i - 1] == 1 ? T[i - 1] : c;
    const float r = mask[i + 1] == 1 ? T[i + 1] : c;
    const float d = mask[i - nx] == 1 ? T[i - nx] : c;
    const float u = mask[i + nx] == 1 ? T[i + nx] : c;
    const float dT = fx * ((l - 2.0f * c) + r) + fy * ((d - 2.0f * c) + u);
    const float t = c + dT;
    o = (300.0f > t) ? 300.0f : t;  // Python max(t, 300.0): NaN stays
  }
  T_new[i] = o;
}

// RF grid nodes: their DC potential is that of the last grid covering them
// (grids are rasterised in order, later ones overwrite, :826-827).
struct GridV {
  double v[kMaxGrids];
};

__global__ __launch_bounds__(kBlock) void k_rf_set(double* V_fixed, const uint8_t* __restrict__ bits,
                                                   int64_t n, int n_grids, GridV gv, int rf, double v_rf) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const unsigned b = bits[i] & ((1u << n_grids) - 1u);
  if (!((b >> rf) & 1u)) return;
  const int last = 31 - __clz(b);
  V_fixed[i] = gv.v[last] + v_rf;
}

// Electrode map of the grids (:791-843): node coordinates as np.linspace,
// bowed grid faces x + delta eta^2, apertures with chamfer, in grid order;
// column 0 is the plasma boundary (Dirichlet).
__global__ __launch_bounds__(kBlock) void k_electrodes(bemcs_electrodes_t a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int nx = a.nx, ny = a.ny;
  if (i >= (int64_t)nx * ny) return;
  const int ix = (int)(i % nx), iy = (int)(i / nx);
  const double X = ix == nx - 1 ? a.Lx : (double)ix * (a.Lx / (double)(nx - 1));
  const double Y = iy == ny - 1 ? a.Ly : (double)iy * (a.Ly / (double)(ny - 1));
  uint8_t bound = 0, bits = 0;
  double vf = 0.0;
  for (int g = 0; g < a.n_grids; ++g) {
    const bemcs_grid_geom_t& G = a.grids[g];
    double dxb = 0.0;
    if (fabs(G.delta) > 1e-6) {
      double eta;
      if (a.geometry == 2) {
        const double yw = 1.5 * a.screen_r + 0.5 * a.pitch;
        const double m = a.pitch * 0.5 > 1e-6 ? a.pitch * 0.5 : 1e-6;
        eta = fmin(fmax(1.0 - fabs(Y - yw) / m, 0.0), 1.0);
        dxb = G.delta * (eta * eta);
      } else if (a.geometry == 1) {
        const double ym = 1.5 * a.screen_r;
        const double m = a.Ly * 0.5 > 1e-6 ? a.Ly * 0.5 : 1e-6;
        eta = fmin(fmax(1.0 - fabs(Y - ym) / m, 0.0), 1.0);
        dxb = G.delta * (eta * eta);
      } else {
        const double Lc = a.Ly - G.r;
        if (Lc > 0) {
          eta = fmin(fmax((a.Ly - Y) / Lc, 0.0), 1.0);
          dxb = G.delta * (eta * eta);
        }
      }
    }
    const bool inx = (X >= G.gstart + dxb) && (X <= G.gend + dxb);
    bool mask = inx;
    for (int h = 0; h < a.n_holes; ++h) {
      double lr = G.r - fmax(0.0, X - G.gstart - dxb) * G.tan_cham;
      lr = fmax(lr, 0.0);
      if (inx && fabs(Y - a.hole_y[h]) <= lr) mask = false;
    }
    if (mask) {
      bound = 1;
      vf = G.V;
      bits |= (uint8_t)(1u << g);
    }
  }
  if (ix == 0) {
    bound = 1;
    vf = a.V_plasma;
  }
  a.is_bound[i] = bound;
  a.V_fixed[i] = vf;
  a.grid_bits[i] = bits;
}

// Per-grid sums of T over a node range (grid stride), partials per work-group.
__global__ __launch_bounds__(kBlock) void k_grid_sums(const float* __restrict__ T,
                                                      const uint8_t* __restrict__ bits, int64_t n,
                                                      int n_grids, double* part) {
  __shared__ double red[2 * (kBlock / 64)];
  double gsum[kMaxGrids], gcnt[kMaxGrids];
#pragma unroll
  for (int g = 0; g < kMaxGrids; ++g) gsum[g] = gcnt[g] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const uint8_t b = bits[i];
    if (!b) continue;
    const double v = (double)T[i];
    for (int g = 0; g < n_grids; ++g)
      if (b & (1u << g)) {
        gsum[g] += v;
        gcnt[g] += 1.0;
      }
  }
  grid_partials(gsum, gcnt, n_grids, part + (size_t)blockIdx.x * 2 * n_grids, red);
}

inline int thermal_blocks(int nx, int ny, dim3* grid) {
  *grid = dim3((unsigned)((nx + kThTX - 1) / kThTX), (unsigned)((ny + kThTY - 1) / kThTY), 1);
  return (int)(grid->x * grid->y);
}

}  // namespace
}  // namespace bemcs

using namespace bemcs;

extern "C" {

int bemcs_erosion(const bemcs_erosion_t* p, bemcs_stream_t stream) {
  BEMCS_REQUIRE(p && p->damage_acc && p->damage_map && p->eroded_depth && p->is_bound && p->diag &&
                    p->n_failed && p->n_nodes > 0,
                "bemcs_erosion: bad arguments");
  hipLaunchKernelGGL(k_erosion, dim3(grid_for(p->n_nodes, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), *p);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_see_commit(const bemcs_species_t* stage, const bemcs_species_t* e_out,
                     const uint8_t* is_bound, uint32_t* e_cnt_next, bemcs_ion_diag_t* diag,
                     int32_t nx, int32_t ny, float dx32, float dy32, float Lx32, float Ly32,
                     int64_t n_bound, bemcs_stream_t stream) {
  BEMCS_REQUIRE(stage && e_out && is_bound && e_cnt_next && diag && nx >= 3 && ny >= 3 &&
                    n_bound >= 0 && stage->count && e_out->count,
                "bemcs_see_commit: bad arguments");
  int64_t n = n_bound < stage->cap ? n_bound : stage->cap;
  if (n == 0) return BEMCS_OK;
  hipLaunchKernelGGL(k_see_commit, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, as_stream(stream),
                     *stage, *e_out, is_bound, e_cnt_next, diag, nx, ny, dx32, dy32, Lx32, Ly32);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int64_t bemcs_thermal_workspace_bytes(int32_t nx, int32_t ny, int32_t n_grids) {
  if (nx < 1 || ny < 1 || n_grids < 0) return 0;
  dim3 g;
  int nb = thermal_blocks(nx, ny, &g);
  return (int64_t)nb * 2 * (n_grids > 0 ? n_grids : 1) * (int64_t)sizeof(double);
}

int bemcs_thermal(const bemcs_thermal_t* p, bemcs_stream_t stream) {
  BEMCS_REQUIRE(p && p->T_in && p->T_out && p->T_in != p->T_out && p->is_bound && p->nx >= 3 &&
                    p->ny >= 3 && p->sweeps >= 0 && p->sweeps <= kMaxSweeps && p->n_grids >= 0 &&
                    p->n_grids <= kMaxGrids,
                "bemcs_thermal: bad arguments (T_in != T_out, sweeps <= 16, n_grids <= 8)");
  BEMCS_REQUIRE(p->n_grids == 0 || (p->grid_bits && p->workspace && p->T_grids),
                "bemcs_thermal: grid_bits, workspace and T_grids are required when n_grids > 0");
  dim3 grid;
  int nb = thermal_blocks(p->nx, p->ny, &grid);
  const int EX = kThTX + 2 * p->sweeps, EY = kThTY + 2 * p->sweeps;
  size_t lds = (size_t)EX * EY * (2 * sizeof(float) + 1);
  hipLaunchKernelGGL(k_thermal, grid, dim3(kBlock), lds, as_stream(stream), *p);
  BEMCS_CHECK_LAUNCH();
  if (p->n_grids > 0) {
    hipLaunchKernelGGL(k_grid_means, dim3(1), dim3(kBlock), 0, as_stream(stream),
                       reinterpret_cast<const double*>(p->workspace), nb, p->n_grids, p->T_grids);
    BEMCS_CHECK_LAUNCH();
  }
  return BEMCS_OK;
}

int bemcs_thermal_conduction(const float* T, float* T_new, const int32_t* mask, int32_t nx, int32_t ny,
                             float fo_x, float fo_y, bemcs_stream_t stream) {
  BEMCS_REQUIRE(T && T_new && mask && nx >= 1 && ny >= 1 && T != T_new,
                "bemcs_thermal_conduction: bad arguments (T and T_new distinct, nx, ny >= 1)");
  const int64_t n = (int64_t)nx * ny;
  hipLaunchKernelGGL(k_conduction_sweep, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, as_stream(stream),
                     T, T_new, mask, (int)nx, (int)ny, fo_x, fo_y);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_rf_set(double* V_fixed, const uint8_t* grid_bits, int64_t n_nodes, int32_t n_grids,
                 const double* grid_V, int32_t rf_grid, double v_rf, bemcs_stream_t stream) {
  BEMCS_REQUIRE(V_fixed && grid_bits && grid_V && n_nodes > 0 && n_grids >= 1 && n_grids <= kMaxGrids &&
                    rf_grid >= 0 && rf_grid < n_grids,
                "bemcs_rf_set: bad arguments (1 <= n_grids <= 8, 0 <= rf_grid < n_grids)");
  GridV gv{};
  for (int g = 0; g < n_grids; ++g) gv.v[g] = grid_V[g];
  hipLaunchKernelGGL(k_rf_set, dim3(grid_for(n_nodes, kBlock)), dim3(kBlock), 0, as_stream(stream),
                     V_fixed, grid_bits, n_nodes, (int)n_grids, gv, (int)rf_grid, v_rf);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_electrodes(const bemcs_electrodes_t* p, bemcs_stream_t stream) {
  BEMCS_REQUIRE(p && p->is_bound && p->V_fixed && p->grid_bits && p->nx >= 3 && p->ny >= 3 &&
                    p->n_grids >= 0 && p->n_grids <= kMaxGrids && p->n_holes >= 0 && p->n_holes <= 2,
                "bemcs_electrodes: bad arguments (n_grids <= 8, n_holes <= 2)");
  const int64_t n = (int64_t)p->nx * p->ny;
  hipLaunchKernelGGL(k_electrodes, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *p);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};
}  // namespace

static int thermal_scratch(void** out) {
  int dev = 0;
  BEMCS_CHECK_HIP(hipGetDevice(&dev));
  BEMCS_REQUIRE(dev >= 0 && dev < kMaxDevices, "bemcs_thermal_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    BEMCS_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = ptr;
  }
  *out = g_scratch_dev[dev];
  return BEMCS_OK;
}

int bemcs_thermal_noalloc(const bemcs_thermal_t* p, bemcs_stream_t stream) {
  BEMCS_REQUIRE(p, "bemcs_thermal_noalloc: null argument");
  const int64_t need = bemcs_thermal_workspace_bytes(p->nx, p->ny, p->n_grids);
  BEMCS_REQUIRE(need <= kThermalScratchBytes,
                "bemcs_thermal_noalloc: mesh too large for the 1 MiB scratch "
                "(bemcs_thermal_workspace_bytes(nx, ny, n_grids) must be <= 1048576)");
  void* ws = nullptr;
  const int rc = thermal_scratch(&ws);
  if (rc != BEMCS_OK) return rc;
  bemcs_thermal_t q = *p;
  q.workspace = ws;
  return bemcs_thermal(&q, stream);
}

int bemcs_grid_means(const float* T, const uint8_t* grid_bits, int32_t nx, int32_t ny, int32_t n_grids,
                     void* workspace, float* T_grids, bemcs_stream_t stream) {
  BEMCS_REQUIRE(T && grid_bits && workspace && T_grids && nx >= 1 && ny >= 1 && n_grids >= 1 &&
                    n_grids <= kMaxGrids,
                "bemcs_grid_means: bad arguments");
  dim3 g;
  const int nb = thermal_blocks(nx, ny, &g);
  hipLaunchKernelGGL(k_grid_sums, dim3(nb), dim3(kBlock), 0, as_stream(stream), T, grid_bits,
                     (int64_t)nx * ny, (int)n_grids, reinterpret_cast<double*>(workspace));
  BEMCS_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_grid_means, dim3(1), dim3(kBlock), 0, as_stream(stream),
                     reinterpret_cast<const double*>(workspace), nb, (int)n_grids, T_grids);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

}  // extern "C"
