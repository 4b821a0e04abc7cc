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
__device__ uint32_t select_kth(const uint32_t* __restrict__ keys, uint32_t n, uint32_t k,
                               uint32_t* hist /* LDS, 2048 */) {
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
    __shared__ uint32_t s_digit, s_k;
    if (threadIdx.x == 0) {
      uint32_t acc = 0, d = 0;
      for (d = 0; d <= dm; ++d) {
        if (acc + hist[d] > k) break;
        acc += hist[d];
      }
      s_digit = d;
      s_k = k - acc;
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
__global__ __launch_bounds__(1024) void k_percentile(const float* __restrict__ a,
                                                     const uint32_t* n_dev, float q_percent,
                                                     uint32_t min_count, double* out) {
  __shared__ uint32_t hist[2048];
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
  uint32_t klo = select_kth(keys, n, lo, hist);
  uint32_t khi = select_kth(keys, n, hi, hist);
  if (threadIdx.x == 0) {
    float va = __uint_as_float(klo), vb = __uint_as_float(khi);
    float d = vb - va;
    *out = (double)(g >= 0.5f ? vb - d * (1.0f - g) : va + d * g);
  }
}

__global__ __launch_bounds__(256) void k_ppc_stats(const uint32_t* __restrict__ ppc, int64_t n,
                                                   double* accum, uint32_t* stats) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  uint32_t c = 0;
  if (i < n) {
    c = ppc[i];
    if (accum) accum[i] += (double)c;
  }
  wave_count(&stats[0], c > 0);
  wave_count(&stats[1], c > 0 && c < 3);
}

}  // namespace bemcs

using namespace bemcs;

extern "C" {

int bemcs_abi_version(void) { return BEMCS_ABI_VERSION; }

const char* bemcs_last_error(void) { return g_err; }

int bemcs_percentile(const float* angles, const uint32_t* n_dev, int64_t n_bound, double q_percent,
                     uint32_t min_count, uint32_t* scratch, double* out, bemcs_stream_t stream) {
  BEMCS_REQUIRE(angles && n_dev && out, "bemcs_percentile: null argument");
  BEMCS_REQUIRE(q_percent >= 0.0 && q_percent <= 100.0, "bemcs_percentile: q must be in [0, 100]");
  (void)scratch;
  (void)n_bound;
  hipLaunchKernelGGL(k_percentile, dim3(1), dim3(1024), 0, as_stream(stream), angles, n_dev,
                     (float)q_percent, min_count, out);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

int bemcs_ppc_stats(const uint32_t* ppc, int64_t n_nodes, double* accum, uint32_t* stats,
                    bemcs_stream_t stream) {
  BEMCS_REQUIRE(ppc && stats && n_nodes > 0, "bemcs_ppc_stats: bad arguments");
  hipLaunchKernelGGL(k_ppc_stats, dim3(grid_for(n_nodes, 256)), dim3(256), 0, as_stream(stream),
                     ppc, n_nodes, accum, stats);
  BEMCS_CHECK_LAUNCH();
  return BEMCS_OK;
}

}  // extern "C"
