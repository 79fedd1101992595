// nsx_batch.hip -- batched evolution population: many independent networks of uneven size
// stepped as one concatenated network.
//
// Replaces the per-genome loop of Population::evolve_generation (population.rs:116-124), where
// rayon builds one CPU engine per genome (Genome::build_engine, genome.rs:151-169) and runs it.
// Layout: network i occupies global neurons [offset_i, offset_i + n_i), offset_i a multiple of
// kBlock, so each neuron block belongs to one network; the step kernels read the block's network
// (seed, local index, E/I split) from a small table.  The concatenated CSR is the concatenation
// of the per-network CSRs with columns shifted by offset_i: the batch generator emits every
// network's synapses in that network's own target-major order and one stable sort by source
// builds all rows, so each network's rows are exactly those a solo build would produce.  All
// random draws are keyed by (network seed, local index, step), so a network's spikes inside the
// batch are bit-identical to its sequential run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "nsx_device.h"
#include "nsx_internal.h"
#include "nsx_kernels.h"

namespace nsx {

// Synapse idx of the concatenated list: find its network (binary search on syn_off), then the
// network's own generator draw.
__global__ void k_gen_batch(const BatchNet* nets, uint32_t n_nets, uint64_t m, uint32_t* src,
                            uint32_t* dst, float* w, uint8_t* delay) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m) return;
  uint32_t lo = 0, hi = n_nets - 1;
  while (lo < hi) {  // largest net with syn_off <= idx
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (nets[mid].syn_off <= idx) lo = mid;
    else hi = mid - 1;
  }
  const BatchNet b = nets[lo];
  const uint64_t li = idx - b.syn_off;
  const uint32_t j = (uint32_t)(li / b.k_in), k = (uint32_t)(li % b.k_in);
  uint32_t s;
  gen_synapse(b.seed, b.n, b.n_exc, b.w_exc_lo, b.w_exc_hi, b.w_inh_lo, b.w_inh_hi, b.d_lo,
              b.d_span, j, k, s, w[idx], delay[idx]);
  src[idx] = b.offset + s;
  dst[idx] = b.offset + j;
}

// Per-network spike totals and sums of squared per-neuron counts (integer atomics: exact and
// order-independent).
__global__ __launch_bounds__(kBlock) void k_batch_stats(StepParams p, DevPtrs s,
                                                        unsigned long long* spikes,
                                                        unsigned long long* sq) {
  __shared__ unsigned long long red[2][kBlock / 64];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const NeuronId id = neuron_id(p, s, i, blockIdx.x);
  unsigned long long c = 0;
  if (i < p.n && id.local < id.n) c = s.count[i];
  unsigned long long c2 = c * c;
  for (int off = 32; off > 0; off >>= 1) {
    c += __shfl_xor(c, off);
    c2 += __shfl_xor(c2, off);
  }
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wid] = c;
    red[1][wid] = c2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0, b = 0;
    for (uint32_t k = 0; k < kBlock / 64; ++k) {
      a += red[0][k];
      b += red[1][k];
    }
    const uint32_t net = s.blk_net[blockIdx.x];
    if (a) atomicAdd(&spikes[net], a);
    if (b) atomicAdd(&sq[net], b);
  }
}

}  // namespace nsx

using namespace nsx;

extern "C" {

nsx_status nsx_batch_create(const nsx_config* base, const nsx_batch_net* nets, uint32_t n_nets,
                            nsx_ctx** out) {
  if (!out) return set_error(nullptr, NSX_E_INVALID, "out is NULL");
  *out = nullptr;
  if (!base || !nets || n_nets < 1) return set_error(nullptr, NSX_E_INVALID, "no networks");
  if (base->struct_size != sizeof(nsx_config))
    return set_error(nullptr, NSX_E_INVALID, "nsx_config.struct_size mismatch");
  if (base->model == NSX_MODEL_HH_SURROGATE)
    return set_error(nullptr, NSX_E_INVALID, "a batched population runs LIF, Izhikevich or HH");
  std::vector<BatchNet> tab(n_nets);
  uint64_t off = 0, syn = 0;
  for (uint32_t i = 0; i < n_nets; ++i) {
    const nsx_batch_net& q = nets[i];
    if (q.n_neurons < 2 || q.n_exc > q.n_neurons)
      return set_error(nullptr, NSX_E_INVALID, "network needs n_neurons >= 2, n_exc <= n_neurons");
    if (q.k_in > 0 && (q.d_lo < 1 || q.d_hi < q.d_lo || q.d_hi >= base->ring_slots))
      return set_error(nullptr, NSX_E_INVALID,
                       "delays must satisfy 1 <= d_lo <= d_hi <= ring_slots-1");
    BatchNet& b = tab[i];
    b.seed = q.seed;
    b.syn_off = syn;
    b.offset = (uint32_t)off;
    b.n = q.n_neurons;
    b.n_exc = q.n_exc;
    b.k_in = q.k_in;
    b.d_lo = q.d_lo;
    b.d_span = q.k_in ? q.d_hi - q.d_lo + 1 : 1;
    b.w_exc_lo = q.w_exc_lo;
    b.w_exc_hi = q.w_exc_hi;
    b.w_inh_lo = q.w_inh_lo;
    b.w_inh_hi = q.w_inh_hi;
    off += ((uint64_t)q.n_neurons + kBlock - 1) / kBlock * kBlock;
    syn += (uint64_t)q.n_neurons * q.k_in;
    if (off >= 0xffffffffull) return set_error(nullptr, NSX_E_INVALID, "population too large");
  }
  if (syn >= (1ull << 31))
    return set_error(nullptr, NSX_E_INVALID, "more than 2^31-1 synapses in one population");
  nsx_config cfg = *base;
  cfg.n_neurons = off;
  cfg.n_exc = off;
  nsx_ctx* ctx = nullptr;
  nsx_status s = nsx_create(&cfg, &ctx);
  if (s) return s;
  auto fail = [&](nsx_status st) {
    set_error(nullptr, st, ctx->err);
    nsx_destroy(ctx);
    return st;
  };
  const uint32_t nblk = ctx->p.nblk;
  std::vector<uint32_t> blk(nblk);
  for (uint32_t i = 0; i < n_nets; ++i) {
    const uint32_t b0 = tab[i].offset / kBlock;
    const uint32_t b1 = (uint32_t)((tab[i].offset + (uint64_t)tab[i].n + kBlock - 1) / kBlock);
    for (uint32_t b = b0; b < b1; ++b) blk[b] = i;
  }
  BatchNet* d_tab = nullptr;
  uint32_t* d_blk = nullptr;
  if ((s = dev_alloc(ctx, (void**)&d_tab, n_nets * sizeof(BatchNet), false))) return fail(s);
  if ((s = dev_alloc(ctx, (void**)&d_blk, (size_t)nblk * 4, false))) return fail(s);
  hipError_t e = hipMemcpy(d_tab, tab.data(), n_nets * sizeof(BatchNet), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_blk, blk.data(), (size_t)nblk * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(hip_error(ctx, e, "batch tables"));
  ctx->d.bnet = d_tab;
  ctx->d.blk_net = d_blk;
  ctx->p.batch = 1;
  ctx->p.n_nets = n_nets;
  ctx->nets_h = tab;
  if ((s = nsx_reset_state(ctx))) return fail(s);  // per-network parameters and identities
  // the concatenated synapse list, then one stable sort by source
  uint32_t *ds = nullptr, *dd = nullptr;
  float* dw = nullptr;
  uint8_t* dl = nullptr;
  if ((s = dev_alloc(ctx, (void**)&ds, syn * 4, false)) ||
      (s = dev_alloc(ctx, (void**)&dd, syn * 4, false)) ||
      (s = dev_alloc(ctx, (void**)&dw, syn * 4, false)) ||
      (s = dev_alloc(ctx, (void**)&dl, syn, false)))
    return fail(s);
  if (syn) {
    hipLaunchKernelGGL(k_gen_batch, dim3((uint32_t)((syn + 255) / 256)), dim3(256), 0,
                       ctx->stream, d_tab, n_nets, syn, ds, dd, dw, dl);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(hip_error(ctx, e, "k_gen_batch"));
  }
  s = build_csr_device(ctx, ds, dd, dw, dl, syn);
  dev_free(ctx, ds);
  dev_free(ctx, dd);
  dev_free(ctx, dw);
  dev_free(ctx, dl);
  if (s) return fail(s);
  *out = ctx;
  return NSX_OK;
}

nsx_status nsx_batch_layout(nsx_ctx* ctx, uint64_t* offsets) {
  if (!ctx || !offsets) return set_error(ctx, NSX_E_INVALID, "ctx and offsets are required");
  if (!ctx->p.batch) return set_error(ctx, NSX_E_STATE, "not a batched population");
  for (size_t i = 0; i < ctx->nets_h.size(); ++i) offsets[i] = ctx->nets_h[i].offset;
  offsets[ctx->nets_h.size()] = ctx->n;
  return NSX_OK;
}

nsx_status nsx_batch_stats(nsx_ctx* ctx, uint64_t* spikes, uint64_t* spikes_sq) {
  if (!ctx || !spikes) return set_error(ctx, NSX_E_INVALID, "ctx and spikes are required");
  if (!ctx->p.batch) return set_error(ctx, NSX_E_STATE, "not a batched population");
  const uint32_t k = ctx->p.n_nets;
  unsigned long long* d = nullptr;
  nsx_status s = dev_alloc(ctx, (void**)&d, (size_t)k * 16, true);
  if (s) return s;
  hipLaunchKernelGGL(k_batch_stats, dim3(ctx->p.nblk), dim3(kBlock), 0, ctx->stream, ctx->p,
                     ctx->d, d, d + k);
  hipError_t e = hipGetLastError();
  std::vector<unsigned long long> h((size_t)k * 2);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), d, (size_t)k * 16, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dev_free(ctx, d);
  if (e != hipSuccess) return hip_error(ctx, e, "nsx_batch_stats");
  for (uint32_t i = 0; i < k; ++i) {
    spikes[i] = h[i];
    if (spikes_sq) spikes_sq[i] = h[k + i];
  }
  return NSX_OK;
}

}  // extern "C"
