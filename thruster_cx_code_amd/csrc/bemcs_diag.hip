// Step diagnostics on the device: theta95 (np.percentile, linear), PPC
// statistics, saddle potential gather, error reporting.
#include <stdarg.h>

#include "bemcs_internal.h"

namespace bemcs {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// k-th smallest (0-based) of n non-negative float32 values, by radix select
// on the bit patterns (11 + 11 + 10 bits), one work-group.
// Exclusive prefix sum over a 1024-thread block (wave shuffles, then the 16
// wave totals); every thread also gets the block total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum /* LDS [16] */) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  if (w == 0) {
    uint32_t t = lane < 16 ? wsum[lane] : 0u;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      uint32_t y = __shfl_up(t, o);
      if (lane >= o) t += y;
    }
    if (lane < 16) wsum[lane] = t;
  }
  __syncthreads();
  const uint32_t base = w > 0 ? wsum[w - 1] : 0u;
  __syncthreads();
  return base + x - v;
}

// k-th smallest (0-based) of n non-negative float32 bit patterns by a 3-pass
// MSD radix select (11/11/10 bits) in one 1024-thread block.  Each pass
// histograms the keys that match the digits chosen so far, and a block scan
// of the histogram finds the bucket holding the k-th key.
__device__ uint32_t select_kth(const uint32_t* __restrict__ keys, uint32_t n, uint32_t k,
                               uint32_t* hist /* LDS, 2048 */) {
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t s_digit, s_k;
  uint32_t prefix = 0, mask = 0;
  const int shifts[3] = {21, 10, 0};
  const int widths[3] = {11, 11, 10};
  for (int pass = 0; pass < 3; ++pass) {
    const int sh = shifts[pass];
    const uint32_t dm = (1u << widths[pass]) - 1u;
    for (int b = threadIdx.x; b < 2048; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      uint32_t v = keys[i];
      if ((v & mask) == prefix) atomicAdd(&hist[(v >> sh) & dm], 1u);
    }
    __syncthreads();
    const uint32_t h0 = hist[2 * threadIdx.x], h1 = hist[2 * threadIdx.x + 1];
    const uint32_t before = block_excl_scan(h0 + h1, wsum);
    if (k >= before && k < before + h0) {
      s_digit = 2 * threadIdx.x;
      s_k = k - before;
    } else if (k >= before + h0 && k < before + h0 + h1) {
      s_digit = 2 * threadIdx.x + 1;
      s_k = k - before - h0;
    }
    __syncthreads();
    prefix |= s_digit << sh;
    mask |= dm << sh;
    k = s_k;
    __syncthreads();
  }
  return prefix;
}

// np.percentile of float32 data evaluates its index and the interpolation in
// float32: q = f32(percent)/f32(100), vi = f32(n-1)*q, gamma = vi - floor(vi),
// lerp a + (b-a)*gamma, or b - (b-a)*(1-gamma) when gamma >= 0.5.
template <bool B>
__global__ __launch_bounds__(1024) void k_percentile(bemcs_percentile_case_t one,
                                                     const bemcs_percentile_case_t* __restrict__ many) {
  __shared__ uint32_t hist[2048];
  const bemcs_percentile_case_t c = B ? many[blockIdx.z] : one;
  const float* __restrict__ a = c.angles;
  const uint32_t* n_dev = c.n_dev;
  const float q_percent = (float)c.q_percent;
  const uint32_t min_count = c.min_count;
  double* out = c.out;
  const uint32_t n = *n_dev;
  if (n <= min_count || n == 0) {
    if (threadIdx.x == 0) *out = __builtin_nan("");
    return;
  }
  const uint32_t* keys = reinterpret_cast<const uint32_t*>(a);
  const float q = q_percent / 100.0f;
  const float vi = (float)(n - 1) * q;
  const float fl = floorf(vi);
  uint32_t lo = (uint32_t)fl;
  if (lo > n - 1) lo = n - 1;
  uint32_t hi = lo + 1 < n ? lo + 1 : n - 1;
  const float g = vi - fl;
  const uint32_t klo = select_kth(keys, n, lo, hist);
  // rank hi = lo + 1 (or lo at the end): klo again when more than hi keys
  // are <= klo, else the smallest key above klo; one pass instead of a
  // second selection
  __shared__ uint32_t s_le, s_next;
  if (threadIdx.x == 0) {
    s_le = 0;
    s_next = 0xFFFFFFFFu;
  }
  __syncthreads();
  uint32_t le = 0, next = 0xFFFFFFFFu;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t v = keys[i];
    le += v <= klo;
    if (v > klo && v < next) next = v;
  }
  if (le) atomicAdd(&s_le, le);
  if (next != 0xFFFFFFFFu) atomicMin(&s_next, next);
  __syncthreads();
  const uint32_t khi = hi == lo || s_le > hi ? klo : s_next;
  if (threadIdx.x == 0) {
    float va = __uint_as_float(klo), vb = __uint_as_float(khi);
    float d = vb - va;
    *out = (double)(g >= 0.5f ? vb - d * (1.0f - g) : va + d * g);
  }
}

// Occupied and low (1-2 particle) cells of the per-step PPC map, and the
// running float64 PPC sum.  Grid-stride over at most kPpcBlocks blocks with
// one pair of counter atomics per block.
constexpr int kPpcBlocks = 256;
template <bool B>
__global__ __launch_bounds__(1024) void k_ppc_stats(bemcs_ppc_case_t one,
                                                    const bemcs_ppc_case_t* __restrict__ many) {
  __shared__ uint32_t s_occ[16], s_low[16];
  const bemcs_ppc_case_t c = B ? many[blockIdx.z] : one;
  const uint32_t* __restrict__ ppc = c.ppc;
  const int64_t n = c.n_nodes;
  double* accum = c.accum;
  uint32_t* stats = c.stats;
  uint32_t occ = 0, low = 0;
  // two nodes per round, both loaded before either is used (the counts are
  // integers and each accumulator element is its own, so order is free)
  const int64_t stride = (int64_t)gridDim.x * 1024;
  for (int64_t i0 = (int64_t)blockIdx.x * 1024 + threadIdx.x; i0 < n; i0 += 2 * stride) {
    const int64_t i1 = i0 + stride;
    const bool has1 = i1 < n;
    const uint32_t c0 = ppc[i0];
    const uint32_t c1 = has1 ? ppc[i1] : 0u;
    const double a0 = accum ? accum[i0] : 0.0;
    const double a1 = accum && has1 ? accum[i1] : 0.0;
    if (accum) {
      accum[i0] = a0 + (double)c0;
      if (has1) accum[i1] = a1 + (double)c1;
    }
    occ += (c0 > 0) + (c1 > 0);
    low += (c0 > 0 && c0 < 3) + (c1 > 0 && c1 < 3);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    occ += __shfl_xor(occ, o);
    low += __shfl_xor(low, o);
  }
  if ((threadIdx.x & 63) == 0) {
    s_occ[threadIdx.x >> 6] = occ;
    s_low[threadIdx.x >> 6] = low;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t a = 0, b = 0;
    for (int w = 0; w < 16; ++w) {
      a += s_occ[w];
      b += s_low[w];
    }
    if (a) atomicAdd(&stats[0], a);
    if (b) atomicAdd(&stats[1], b);
  }
}

// --------------------------------------------------- per-step bookkeeping
template <bool B>
__global__ __launch_bounds__(256) void k_step_begin(bemcs_step_begin_t one,
                                                    const bemcs_step_begin_t* __restrict__ many) {
  const bemcs_step_begin_t p = B ? many[blockIdx.z] : one;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < p.n_nodes) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (p.node_zero[a]) p.node_zero[a][i] = 0u;
  }
  if (i < p.diag_bytes) p.diag[i] = 0;
  if (i == 0) {
    p.ds_bits[0] = 0ull;
    p.ds_bits[1] = 0ull;
    *p.count_out_i = 0;
    *p.count_out_e = 0;
    p.ppc_stats[0] = 0u;
    p.ppc_stats[1] = 0u;
    p.nin_out[0] = *p.count_in_i;
    p.nin_out[1] = *p.count_in_e;
    if (p.count_out_x) *p.count_out_x = 0;
  }
}

template <bool B>
__global__ __launch_bounds__(64) void k_step_end(bemcs_step_end_t one,
                                                 const bemcs_step_end_t* __restrict__ many) {
  const bemcs_step_end_t p = B ? many[blockIdx.z] : one;
  double m = __builtin_inf();
  for (int k = threadIdx.x; k < p.n_saddle; k += 64) m = fmin(m, p.V[p.saddle_idx[k]]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmin(m, __shfl_xor(m, o));
  if (threadIdx.x == 0) {
    p.counts_out[0] = *p.count_i;
    p.counts_out[1] = *p.count_e;
    p.nsub_out[0] = p.n_sub[0];
    p.nsub_out[1] = p.n_sub[1];
    *p.minpot_out = p.n_saddle > 0 ? m : __builtin_nan("");
  }
}

}  // namespace bemcs

using namespace bemcs;

extern "C" {

int bemcs_abi_version(void) { return BEMCS_ABI_VERSION; }

const char* bemcs_last_error(void) { return g_err; }

// ------------------------------------------------------------ HIP graphs
// Thread-local capture, so launches of other host threads (and their
// allocations) are unaffected; the stream may have work pending, which the
// graph does not depend on and which its launches follow in stream order.
int bemcs_graph_capture_begin(bemcs_stream_t stream) {
  BEMCS_CHECK_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
  return BEMCS_OK;
}

int bemcs_graph_capture_end(bemcs_stream_t stream, int32_t abort, void** exec_out) {
  BEMCS_REQUIRE(exec_out, "bemcs_graph_capture_end: null argument");
  *exec_out = nullptr;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(as_stream(stream), &g);
  if (abort || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    if (abort) return BEMCS_OK;
    BEMCS_CHECK_HIP(e);
  }
  hipGraphExec_t x = nullptr;
  const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  BEMCS_CHECK_HIP(ie);
  *exec_out = x;
  return BEMCS_OK;
}

int bemcs_graph_launch(void* exec, bemcs_stream_t stream) {
  BEMCS_REQUIRE(exec, "bemcs_graph_launch: null graph");
  BEMCS_CHECK_HIP(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(exec), as_stream(stream)));
  return BEMCS_OK;
}

int bemcs_graph_destroy(void* exec) {
  if (exec) BEMCS_CHECK_HIP(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(exec)));
  return BEMCS_OK;
}

int bemcs_percentile(const float* angles, const uint32_t* n_dev, int64_t n_bound, double q_percent,
                     uint32_t min_count, uint32_t* scratch, double* out, bemcs_stream_t stream) {
  BEMCS_REQUIRE(angles && n_dev && out, "bemcs_percentile: null argument");
  BEMCS_REQUIRE(q_percent >= 0.0 && q_percent <= 100.0, "bemcs_percentile: q must be in [0, 100]");
  (void)scratch;
  (void)n_bound;
  const bemcs_percentile_case_t c{angles, n_dev, q_percent, min_count, 0u, out};
  hipLaunchKernelGGL(k_percentile<false>, dim3(1), dim3(1024), 0, as_stream(stream), c,
                     (const bemcs_percentile_case_t*)nullptr);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_ppc_stats(const uint32_t* ppc, int64_t n_nodes, double* accum, uint32_t* stats,
                    bemcs_stream_t stream) {
  BEMCS_REQUIRE(ppc && stats && n_nodes > 0, "bemcs_ppc_stats: bad arguments");
  unsigned g = grid_for(n_nodes, 1024);
  if (g > (unsigned)kPpcBlocks) g = kPpcBlocks;
  const bemcs_ppc_case_t c{ppc, n_nodes, accum, stats};
  hipLaunchKernelGGL(k_ppc_stats<false>, dim3(g), dim3(1024), 0, as_stream(stream), c,
                     (const bemcs_ppc_case_t*)nullptr);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_step_begin(const bemcs_step_begin_t* p, bemcs_stream_t stream) {
  BEMCS_REQUIRE(p && p->ds_bits && p->count_out_i && p->count_out_e && p->ppc_stats &&
                    p->count_in_i && p->count_in_e && p->nin_out && (p->diag || p->diag_bytes == 0) &&
                    p->n_nodes >= 0 && p->diag_bytes >= 0,
                "bemcs_step_begin: bad arguments");
  int64_t n = p->n_nodes > p->diag_bytes ? p->n_nodes : p->diag_bytes;
  hipLaunchKernelGGL(k_step_begin<false>, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), *p,
                     (const bemcs_step_begin_t*)nullptr);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_step_end(const bemcs_step_end_t* p, bemcs_stream_t stream) {
  BEMCS_REQUIRE(p && p->count_i && p->count_e && p->n_sub && p->counts_out && p->nsub_out &&
                    p->minpot_out && p->n_saddle >= 0 && (p->n_saddle == 0 || (p->V && p->saddle_idx)),
                "bemcs_step_end: bad arguments");
  hipLaunchKernelGGL(k_step_end<false>, dim3(1), dim3(64), 0, as_stream(stream), *p,
                     (const bemcs_step_end_t*)nullptr);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_step_begin_batched(const bemcs_step_begin_t* cases, int32_t n, int64_t max_nodes,
                             bemcs_stream_t stream) {
  BEMCS_REQUIRE(cases && n >= 1 && max_nodes >= 0, "bemcs_step_begin_batched: bad arguments");
  // the clears cover the larger of the node arrays and the diagnostics record
  const int64_t m = max_nodes > (int64_t)sizeof(bemcs_ion_diag_t) ? max_nodes
                                                                 : (int64_t)sizeof(bemcs_ion_diag_t);
  hipLaunchKernelGGL(k_step_begin<true>, dim3(grid_for(m, 256), 1, n), dim3(256), 0, as_stream(stream),
                     bemcs_step_begin_t{}, cases);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_step_end_batched(const bemcs_step_end_t* cases, int32_t n, bemcs_stream_t stream) {
  BEMCS_REQUIRE(cases && n >= 1, "bemcs_step_end_batched: bad arguments");
  hipLaunchKernelGGL(k_step_end<true>, dim3(1, 1, n), dim3(64), 0, as_stream(stream),
                     bemcs_step_end_t{}, cases);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_ppc_stats_batched(const bemcs_ppc_case_t* cases, int32_t n, bemcs_stream_t stream) {
  BEMCS_REQUIRE(cases && n >= 1, "bemcs_ppc_stats_batched: bad arguments");
  hipLaunchKernelGGL(k_ppc_stats<true>, dim3(kPpcBlocks / 4, 1, n), dim3(1024), 0, as_stream(stream),
                     bemcs_ppc_case_t{}, cases);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_percentile_batched(const bemcs_percentile_case_t* cases, int32_t n,
                             bemcs_stream_t stream) {
  BEMCS_REQUIRE(cases && n >= 1, "bemcs_percentile_batched: bad arguments");
  hipLaunchKernelGGL(k_percentile<true>, dim3(1, 1, n), dim3(1024), 0, as_stream(stream),
                     bemcs_percentile_case_t{}, cases);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

}  // extern "C"
