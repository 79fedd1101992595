// nsx_stats.hip -- device-side simulation statistics: SimulationEngine::step phase 3
// (engine.rs:617-670) without its O(N + M) host passes per step.
//
// With stats on, every update kernel adds its block's spike count of step t into a striped
// per-step counter (kStatStripes atomics per step, not one per block on one address) and the HH
// kernels add their block's sum of |synaptic currents| into a per-block running total.  After
// each nsx_step chunk the host folds the per-step counts into the reference's running values:
// mean_firing_rate is the mean of the per-neuron low-pass rates, and because that filter is
// linear it is the same low-pass applied to the population's per-step rate; active_neurons,
// weight_updates and synchrony_index come from the per-step counts.  Weight mean / std are an
// on-demand two-pass reduction over all synapses with a fixed summation order.
#include <math.h>

#include <mutex>
#include <vector>

#include "nsx_internal.h"

namespace nsx {

constexpr uint32_t kWBlocks = 1024;

// Per-block FP64 sums of w (pass 0) or of (w - mean)^2 (pass 1); a thread sums its strided
// elements in index order, the block combines its threads in a fixed tree.
__global__ __launch_bounds__(256) void k_weight_partials(const float* w, uint64_t m,
                                                         double mean, int pass, double* part) {
  __shared__ double red[256];
  double a = 0.0;
  for (uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x; k < m;
       k += (uint64_t)gridDim.x * 256) {
    const double x = (double)w[k];
    if (pass == 0) {
      a += x;
    } else {
      const double d = x - mean;
      a += d * d;
    }
  }
  red[threadIdx.x] = a;
  __syncthreads();
  for (uint32_t h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Mean and population standard deviation of the weights by two passes of kWBlocks per-block
// partial sums into `part` (kWBlocks doubles of device memory), summed on the host in block order.
static nsx_status weight_stats_into(nsx_ctx* ctx, double* part, double* mean, double* stddev) {
  nsx_status s = flush_pending(ctx, true);
  if (s) return s;
  double h[kWBlocks];
  double mu = 0.0;
  for (int pass = 0; pass < 2; ++pass) {
    hipLaunchKernelGGL(k_weight_partials, dim3(kWBlocks), dim3(256), 0, ctx->stream, ctx->w,
                       ctx->m, mu, pass, part);
    NSX_HIP(ctx, hipGetLastError());
    NSX_HIP(ctx, hipMemcpyAsync(h, part, kWBlocks * 8, hipMemcpyDeviceToHost, ctx->stream));
    NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double a = 0.0;
    for (uint32_t b = 0; b < kWBlocks; ++b) a += h[b];
    if (pass == 0) mu = a / (double)ctx->m;
    else *stddev = sqrt(a / (double)ctx->m);
  }
  *mean = mu;
  return NSX_OK;
}

// Device memory of nsx_weight_stats_noalloc: the kWBlocks per-block partial sums of one pass.
// Every block writes its own entry, so nothing needs clearing between calls.  The grid is fixed
// at kWBlocks blocks whatever the synapse count (its loop strides), so no input needs more.  Each
// GPU the library runs on holds one copy until the process ends; in host code the array's name
// is not a device address -- scratch_address() gives each device's.
constexpr uint32_t kScratchDoubles = kWBlocks;
static __device__ __align__(256) double scratch[kScratchDoubles];
static_assert(kWBlocks <= kScratchDoubles, "the partial sums fit the scratch array");
constexpr int kScratchDevices = 64;
static double* g_scratch[kScratchDevices] = {};
static std::mutex g_scratch_mu;

// The scratch array's address on the context's device, looked up once per device with that
// device current (hipGetSymbolAddress) and kept.
static nsx_status scratch_address(nsx_ctx* ctx, double** out) {
  const int dev = ctx->device;
  if (dev < 0 || dev >= kScratchDevices) return set_error(ctx, NSX_E_INVALID, "device index");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    int cur = -1;
    NSX_HIP(ctx, hipGetDevice(&cur));
    if (cur != dev) NSX_HIP(ctx, hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess) return hip_error(ctx, e, "hipGetSymbolAddress(scratch)");
    g_scratch[dev] = (double*)p;
  }
  *out = g_scratch[dev];
  return NSX_OK;
}

// Fold the per-step counts of steps [first, first + n) (n <= stat_cap) into the host stats.
nsx_status stats_fold(nsx_ctx* ctx, uint64_t first, uint64_t n) {
  if (!n) return NSX_OK;
  const uint32_t cap = ctx->p.stat_cap;
  std::vector<uint32_t> h((size_t)n * kStatStripes);
  uint64_t k = 0;
  while (k < n) {
    const uint64_t t = first + k;
    const uint64_t piece = std::min<uint64_t>(n - k, cap - t % cap);
    NSX_HIP(ctx, hipMemcpy(h.data() + k * kStatStripes,
                           ctx->d.stat_counts + (t % cap) * kStatStripes,
                           piece * kStatStripes * 4, hipMemcpyDeviceToHost));
    k += piece;
  }
  const double dt = ctx->cfg.dt_ms, nn = (double)ctx->n;
  const double alpha = dt / 100.0, inst = 1000.0 / dt;  // engine.rs:636-640
  auto& hs = ctx->hs;
  for (uint64_t q = 0; q < n; ++q) {
    uint64_t c = 0;
    for (uint32_t r = 0; r < kStatStripes; ++r) c += h[q * kStatStripes + r];
    hs.active = std::max(hs.active, c);
    hs.last = c;
    const double mean_inst = inst * (double)c / nn;
    hs.rate += (mean_inst - hs.rate) * alpha;
    hs.steps += 1;
  }
  return NSX_OK;
}

}  // namespace nsx

using namespace nsx;

extern "C" {

nsx_status nsx_set_stats(nsx_ctx* ctx, int enable) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (enable && !ctx->d.stat_counts) {
    nsx_status s;
    ctx->p.stat_cap = 1u << 16;
    if ((s = dev_alloc(ctx, (void**)&ctx->d.stat_counts,
                       (size_t)ctx->p.stat_cap * kStatStripes * 4, true)))
      return s;
    if ((s = dev_alloc(ctx, (void**)&ctx->d.lfp_part, (size_t)ctx->p.nblk * 8, true))) return s;
  }
  if (ctx->d.lfp_part)
    NSX_HIP(ctx, hipMemsetAsync(ctx->d.lfp_part, 0, (size_t)ctx->p.nblk * 8, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->p.stats_on = enable ? 1u : 0u;
  ctx->hs = {};
  destroy_graph(ctx);  // the kernels' parameters changed
  return NSX_OK;
}

nsx_status nsx_weight_stats(nsx_ctx* ctx, double* mean, double* stddev) {
  if (!ctx || !mean || !stddev) return set_error(ctx, NSX_E_INVALID, "ctx, mean, std required");
  if (!ctx->has_network) return set_error(ctx, NSX_E_STATE, "no network loaded");
  *mean = 0.0;
  *stddev = 0.0;
  if (ctx->m == 0) return NSX_OK;  // the reference leaves both at 0 without synapses
  double* d = nullptr;
  nsx_status s = dev_alloc(ctx, (void**)&d, kWBlocks * 8, false);
  if (s) return s;
  s = weight_stats_into(ctx, d, mean, stddev);
  dev_free(ctx, d);
  return s;
}

nsx_status nsx_weight_stats_noalloc(nsx_ctx* ctx, double* mean, double* stddev) {
  if (!ctx || !mean || !stddev) return set_error(ctx, NSX_E_INVALID, "ctx, mean, std required");
  if (!ctx->has_network) return set_error(ctx, NSX_E_STATE, "no network loaded");
  *mean = 0.0;
  *stddev = 0.0;
  if (ctx->m == 0) return NSX_OK;
  double* d = nullptr;
  nsx_status s = scratch_address(ctx, &d);
  if (s) return s;
  return weight_stats_into(ctx, d, mean, stddev);
}

nsx_status nsx_read_sim_stats(nsx_ctx* ctx, nsx_sim_stats* out) {
  if (!ctx || !out) return set_error(ctx, NSX_E_INVALID, "ctx and out are required");
  if (!ctx->p.stats_on) return set_error(ctx, NSX_E_STATE, "stats are off (nsx_set_stats)");
  nsx_sim_stats r{};
  nsx_step_stats st{};
  nsx_status s = nsx_step(ctx, 0, &st);  // totals only
  if (s) return s;
  r.total_spikes = st.total_spikes;
  r.output_spikes = 0;
  r.active_neurons = ctx->hs.active;
  r.weight_updates = ctx->hs.last;
  r.sim_time_ms = st.sim_time_ms;
  r.mean_firing_rate = ctx->n ? ctx->hs.rate : 0.0;
  r.synchrony_index = ctx->hs.steps && ctx->n ? (double)ctx->hs.last / (double)ctx->n : 0.0;
  if ((s = nsx_weight_stats(ctx, &r.weight_mean, &r.weight_std))) return s;
  std::vector<double> lp(ctx->p.nblk);
  NSX_HIP(ctx, hipMemcpy(lp.data(), ctx->d.lfp_part, lp.size() * 8, hipMemcpyDeviceToHost));
  double a = 0.0;
  for (double x : lp) a += x;
  r.lfp = a;
  r.seed = ctx->cfg.seed;
  r.steps = ctx->hs.steps;
  *out = r;
  return NSX_OK;
}

}  // extern "C"
