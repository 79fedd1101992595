// nsx_csr.hip -- network construction on the device.
//
// Network::finalize (src/network/mod.rs:79-106) is a stable counting sort of the synapse list
// by source: rows in source order, insertion order kept inside a row.  On the device this is a
// stable LSD radix sort of (source, insertion index) pairs; col / weight / delay are then
// gathered through the permutation, so all three end up in CSR slot order (the reference
// permutes only adjacency_indices, see engine.rs:244-247).  Row pointers are 64-bit.
#include <hipcub/hipcub.hpp>

#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "nsx_device.h"
#include "nsx_internal.h"
#include "nsx_kernels.h"

namespace nsx {

// Element-wise kernels stride over their domain: a dispatch counts its work-items in 32 bits,
// so a launch is capped (grid_for) and every thread takes every (grid size)-th element.
#define GRID_STRIDE(k, count) \
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < (count); \
       k += (uint64_t)gridDim.x * blockDim.x)

__global__ void k_check_list(const uint32_t* src, const uint32_t* dst, const uint8_t* delay,
                             uint64_t m, uint32_t n, uint32_t D, unsigned int* bad) {
  GRID_STRIDE(k, m) {
    if (src[k] >= n || dst[k] >= n || delay[k] < 1u || delay[k] >= D) atomicOr(bad, 1u);
  }
}

__global__ void k_iota(uint32_t* v, uint64_t m) {
  GRID_STRIDE(k, m) v[k] = (uint32_t)k;
}

// row_ptr[r] = first sorted position whose key >= r (keys sorted ascending)
__global__ void k_row_ptr(const uint32_t* keys, uint64_t m, uint32_t n, uint64_t* row_ptr) {
  GRID_STRIDE(k, m + 1) {
    uint32_t lo = (k == 0) ? 0u : keys[k - 1] + 1u;
    uint32_t hi = (k == m) ? n : keys[k];
    for (uint32_t r = lo; r <= hi && r <= n; ++r) row_ptr[r] = k;
    if (k == m && m == 0)
      for (uint32_t r = 0; r <= n; ++r) row_ptr[r] = 0;
  }
}

__global__ void k_gather(const uint32_t* perm, const uint32_t* dst_in, const float* w_in,
                         const uint8_t* d_in, uint64_t m, uint32_t* col, float* w,
                         uint8_t* delay) {
  GRID_STRIDE(k, m) {
    uint32_t q = perm[k];
    col[k] = dst_in[q];
    w[k] = w_in[q];
    delay[k] = d_in[q];
  }
}

// Fixed in-degree generator: synapse (j, k) from Philox(seed; j, k, CONN).
__global__ void k_gen_fixed_indegree(uint64_t seed, uint32_t n, uint32_t n_exc, uint32_t k_in,
                                     float w_exc_lo, float w_exc_hi, float w_inh_lo,
                                     float w_inh_hi, uint32_t d_lo, uint32_t d_span,
                                     uint32_t* src, uint32_t* dst, float* w, uint8_t* delay) {
  GRID_STRIDE(idx, (uint64_t)n * k_in) {
    uint32_t j = (uint32_t)(idx / k_in), k = (uint32_t)(idx % k_in);
    gen_synapse(seed, n, n_exc, w_exc_lo, w_exc_hi, w_inh_lo, w_inh_hi, d_lo, d_span, j, k,
                src[idx], w[idx], delay[idx]);
    dst[idx] = j;
  }
}

// ---- direct fixed in-degree build (no global sort; any number of synapses) ------------------
struct ToU64 {
  __host__ __device__ unsigned long long operator()(uint32_t x) const { return x; }
};

// Synapse (j, k) of the generator is row entry number (j * k_in + k) in insertion order; CSR by
// source in that order is what a stable sort by source gives (Network::finalize,
// network/mod.rs:79-106).  Pass 1 counts out-degrees, pass 2 drops every synapse into its row
// at an atomic cursor together with k, and a per-row sort by (target, k) restores insertion
// order.

__global__ void k_gen_count(uint64_t seed, uint32_t n, uint32_t n_exc, uint32_t k_in,
                            uint32_t* cnt) {
  GRID_STRIDE(idx, (uint64_t)n * k_in) {
    const uint32_t j = (uint32_t)(idx / k_in), k = (uint32_t)(idx % k_in);
    U4 r = draw(seed, j, k, kStreamConn);
    uint32_t src = (uint32_t)(((uint64_t)r.x * (uint64_t)(n - 1)) >> 32);
    if (src >= j) src += 1;
    atomicAdd(&cnt[src], 1u);
  }
}

__global__ void k_gen_fill(uint64_t seed, uint32_t n, uint32_t n_exc, uint32_t k_in,
                           float w_exc_lo, float w_exc_hi, float w_inh_lo, float w_inh_hi,
                           uint32_t d_lo, uint32_t d_span, const uint64_t* row_ptr,
                           uint32_t* cursor, uint32_t* col, float* w, uint8_t* delay,
                           uint8_t* kk) {
  GRID_STRIDE(idx, (uint64_t)n * k_in) {
    const uint32_t j = (uint32_t)(idx / k_in), k = (uint32_t)(idx % k_in);
    uint32_t src;
    float wv;
    uint8_t dv;
    gen_synapse(seed, n, n_exc, w_exc_lo, w_exc_hi, w_inh_lo, w_inh_hi, d_lo, d_span, j, k,
                src, wv, dv);
    const uint64_t pos = row_ptr[src] + atomicAdd(&cursor[src], 1u);
    col[pos] = j;
    w[pos] = wv;
    delay[pos] = dv;
    kk[pos] = (uint8_t)k;
  }
}

// Every row filled exactly: the cursors equal the row lengths.
__global__ void k_check_fill(const uint64_t* row_ptr, const uint32_t* cursor, uint32_t n,
                             unsigned int* bad) {
  GRID_STRIDE(i, n) {
    if ((uint64_t)cursor[i] != row_ptr[i + 1] - row_ptr[i]) atomicOr(bad, 1u);
  }
}

// One wavefront per row (rows of at most 256 entries): bitonic sort of (col, k) in LDS, then the
// row's col / w / delay rewritten in that order.
__global__ __launch_bounds__(256) void k_row_sort(const uint64_t* row_ptr, uint32_t n,
                                                  uint32_t* col, float* w, uint8_t* delay,
                                                  const uint8_t* kk) {
  __shared__ unsigned long long key[4][256];
  __shared__ uint32_t scol[4][256];
  __shared__ float sw[4][256];
  __shared__ uint8_t sd[4][256];
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  for (uint64_t row = (uint64_t)blockIdx.x * 4 + wid; row < n;
       row += (uint64_t)gridDim.x * 4) {
    const uint64_t beg = row_ptr[row];
    const uint32_t len = (uint32_t)(row_ptr[row + 1] - beg);
    if (len < 2) continue;  // wave-uniform
    uint32_t P = 2;
    while (P < len) P <<= 1;
    for (uint32_t e = lane; e < P; e += 64) {
      if (e < len) {
        const uint32_t c = col[beg + e];
        scol[wid][e] = c;
        sw[wid][e] = w[beg + e];
        sd[wid][e] = delay[beg + e];
        // key: (target, k) in the high 40 bits, the entry's position in the low 8 bits
        key[wid][e] = ((unsigned long long)c << 24) | ((unsigned long long)kk[beg + e] << 8) | e;
      } else {
        key[wid][e] = ~0ull;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t size = 2; size <= P; size <<= 1) {
      for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (uint32_t e = lane; e < P; e += 64) {
          const uint32_t o = e ^ stride;
          if (o > e) {
            const bool up = (e & size) == 0;
            const unsigned long long a = key[wid][e], b = key[wid][o];
            if ((a > b) == up) {
              key[wid][e] = b;
              key[wid][o] = a;
            }
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    for (uint32_t e = lane; e < len; e += 64) {
      const uint32_t from = (uint32_t)(key[wid][e] & 0xffu);
      col[beg + e] = scol[wid][from];
      w[beg + e] = sw[wid][from];
      delay[beg + e] = sd[wid][from];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // LDS is reused by the next row
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

__global__ void k_check_csr(const uint64_t* row_ptr, const uint32_t* col, const uint8_t* delay,
                            uint32_t n, uint64_t m, uint32_t D, unsigned int* bad) {
  GRID_STRIDE(k, (m > n ? m : (uint64_t)n) + 1) {
    if (k < n) {
      if (row_ptr[k] > row_ptr[k + 1]) atomicOr(bad, 1u);
    }
    if (k == 0 && (row_ptr[0] != 0 || row_ptr[n] != m)) atomicOr(bad, 2u);
    if (k < m) {
      if (col[k] >= n) atomicOr(bad, 4u);
      if (delay && (delay[k] < 1u || delay[k] >= D)) atomicOr(bad, 8u);
    }
  }
}

// Excitatory test of the transpose build: i < n_exc, and in a batched population (blk_net set,
// n_exc = all neurons) local index < its network's n_exc.
__device__ inline bool csc_exc(uint32_t i, uint32_t n_exc, const uint32_t* blk_net,
                               const BatchNet* bnet) {
  if (i >= n_exc) return false;
  if (!blk_net) return true;
  const BatchNet& b = bnet[blk_net[i / kBlock]];
  return i - b.offset < b.n_exc;
}

// E->E transpose: count per excitatory target, then scatter (slot, source).
__global__ void k_csc_count(const uint64_t* row_ptr, const uint32_t* col, uint32_t n_exc,
                            const uint32_t* blk_net, const BatchNet* bnet,
                            unsigned long long* cnt) {
  for (uint32_t i = blockIdx.x; i < n_exc; i += gridDim.x) {  // a block per source row
    if (!csc_exc(i, n_exc, blk_net, bnet)) continue;
    for (uint64_t k = row_ptr[i] + threadIdx.x; k < row_ptr[i + 1]; k += blockDim.x) {
      uint32_t j = col[k];
      if (csc_exc(j, n_exc, blk_net, bnet)) atomicAdd(&cnt[j], 1ull);
    }
  }
}

__global__ void k_csc_fill(const uint64_t* row_ptr, const uint32_t* col, uint32_t n_exc,
                           const uint32_t* blk_net, const BatchNet* bnet,
                           unsigned long long* cursor, uint64_t* slot, uint32_t* srcs) {
  for (uint32_t i = blockIdx.x; i < n_exc; i += gridDim.x) {
    if (!csc_exc(i, n_exc, blk_net, bnet)) continue;
    for (uint64_t k = row_ptr[i] + threadIdx.x; k < row_ptr[i + 1]; k += blockDim.x) {
      uint32_t j = col[k];
      if (csc_exc(j, n_exc, blk_net, bnet)) {
        unsigned long long p = atomicAdd(&cursor[j], 1ull);
        slot[p] = k;
        srcs[p] = i;
      }
    }
  }
}

// Every target's in-slots filled: cursor[j] reached csc_ptr[j + 1].
__global__ void k_check_csc(const uint64_t* csc_ptr, const unsigned long long* cursor,
                            uint32_t n_exc, unsigned int* bad) {
  GRID_STRIDE(j, n_exc) {
    if (cursor[j] != csc_ptr[j + 1]) atomicOr(bad, 1u);
  }
}

// Blocks for `m` items at `b` per block, capped: the kernels stride over the rest.
static inline uint32_t blocks_for(uint64_t m, uint32_t b) {
  const uint64_t g = (m + b - 1) / b;
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(g, 1), 1u << 16);
}

static nsx_status free_network(nsx_ctx* ctx) {
  // pull delivery: the input the old network's spikes still owe joins the ring while its rows
  // are here; the tables belong to that network
  if (ctx->p.pull_on) (void)pull_materialize(ctx);
  pull_release(ctx);
  dev_free(ctx, ctx->row_ptr);
  dev_free(ctx, ctx->col);
  dev_free(ctx, ctx->w);
  dev_free(ctx, ctx->delay);
  dev_free(ctx, ctx->csc_ptr);
  dev_free(ctx, ctx->csc_slot);
  dev_free(ctx, ctx->csc_src);
  ctx->row_ptr = nullptr;
  ctx->col = nullptr;
  ctx->w = nullptr;
  ctx->delay = nullptr;
  ctx->csc_ptr = nullptr;
  ctx->csc_slot = nullptr;
  ctx->csc_src = nullptr;
  ctx->has_network = false;
  ctx->m = 0;
  ctx->csc_m = 0;
  return NSX_OK;
}

__global__ void k_min_delay(const uint8_t* delay, uint64_t m, unsigned int* out) {
  unsigned int mn = 255u;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m;
       k += (uint64_t)gridDim.x * blockDim.x)
    mn = min(mn, (unsigned int)delay[k]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mn = min(mn, (unsigned int)__shfl_xor((int)mn, o));
  if ((threadIdx.x & 63u) == 0) atomicMin(out, mn);
}

nsx_status network_min_delay(nsx_ctx* ctx, uint32_t* d_min) {
  *d_min = (uint32_t)ctx->D;
  if (ctx->m == 0) return NSX_OK;
  unsigned int* d = nullptr;
  nsx_status s = dev_alloc(ctx, (void**)&d, 4, false);
  if (s) return s;
  const unsigned int init = 255u;
  unsigned int h = 255u;
  hipError_t e = hipMemcpyAsync(d, &init, 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_min_delay, dim3(blocks_for(ctx->m, 256 * 16)), dim3(256), 0,
                       ctx->stream, ctx->delay, ctx->m, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&h, d, 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dev_free(ctx, d);
  if (e != hipSuccess) return hip_error(ctx, e, "minimum delay");
  *d_min = std::min<uint32_t>(h, (uint32_t)ctx->D);
  return NSX_OK;
}

static void publish_network(nsx_ctx* ctx) {
  ctx->d_min = 0;  // recomputed before the next step
  ctx->d.row_ptr = ctx->row_ptr;
  ctx->d.col = ctx->col;
  ctx->d.w = ctx->w;
  ctx->d.delay = ctx->delay;
  ctx->d.csc_ptr = ctx->csc_ptr;
  ctx->d.csc_slot = ctx->csc_slot;
  ctx->d.csc_src = ctx->csc_src;
  ctx->has_network = true;
  (void)flush_pending(ctx, false);  // depressions pending on the old weights are dropped
  // binned delivery: events in flight join the ring; buffers (or the pull tables) are re-sized
  // before the next step
  if (ctx->delivery != NSX_DELIVERY_ATOMIC) {
    (void)bin_materialize(ctx);
    ctx->bin_dirty = true;
  }
  // a captured graph holds the old pointers
  if (ctx->graph_exec) {
    (void)hipGraphExecDestroy(ctx->graph_exec);
    ctx->graph_exec = nullptr;
  }
  if (ctx->graph) {
    (void)hipGraphDestroy(ctx->graph);
    ctx->graph = nullptr;
  }
}

nsx_status build_csc_device(nsx_ctx* ctx) {
  const uint32_t ne = (uint32_t)ctx->cfg.n_exc;
  nsx_status s;
  if ((s = dev_alloc(ctx, (void**)&ctx->csc_ptr, ((size_t)ne + 1) * 8, true))) return s;
  if (ne == 0) return NSX_OK;
  unsigned long long* cnt = nullptr;
  if ((s = dev_alloc(ctx, (void**)&cnt, (size_t)ne * 8, true))) return s;
  const uint32_t* bt = ctx->p.batch ? ctx->d.blk_net : nullptr;  // batch: per-network E/I
  hipLaunchKernelGGL(k_csc_count, dim3(blocks_for(ne, 1)), dim3(128), 0, ctx->stream,
                     ctx->row_ptr, ctx->col, ne, bt, ctx->d.bnet, cnt);
  NSX_HIP(ctx, hipGetLastError());
  // exclusive scan into csc_ptr[1..]
  size_t tmp_bytes = 0;
  NSX_HIP(ctx, hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, cnt,
                                                 (unsigned long long*)ctx->csc_ptr + 1, (int)ne,
                                                 ctx->stream));
  void* tmp = nullptr;
  if ((s = dev_alloc(ctx, &tmp, tmp_bytes, false))) return s;
  NSX_HIP(ctx, hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, cnt,
                                                 (unsigned long long*)ctx->csc_ptr + 1, (int)ne,
                                                 ctx->stream));
  uint64_t tot = 0;
  NSX_HIP(ctx, hipMemcpyAsync(&tot, ctx->csc_ptr + ne, 8, hipMemcpyDeviceToHost, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dev_free(ctx, tmp);
  if ((s = dev_alloc(ctx, (void**)&ctx->csc_slot, tot * 8, false))) return s;
  if ((s = dev_alloc(ctx, (void**)&ctx->csc_src, tot * 4, false))) return s;
  NSX_HIP(ctx, hipMemcpyAsync(cnt, ctx->csc_ptr, (size_t)ne * 8, hipMemcpyDeviceToDevice,
                              ctx->stream));
  hipLaunchKernelGGL(k_csc_fill, dim3(blocks_for(ne, 1)), dim3(128), 0, ctx->stream,
                     ctx->row_ptr, ctx->col, ne, bt, ctx->d.bnet, cnt, ctx->csc_slot,
                     ctx->csc_src);
  NSX_HIP(ctx, hipGetLastError());
  unsigned int* bad = nullptr;
  if ((s = dev_alloc(ctx, (void**)&bad, 4, true))) return s;
  hipLaunchKernelGGL(k_check_csc, dim3(blocks_for(ne, 256)), dim3(256), 0, ctx->stream,
                     ctx->csc_ptr, cnt, ne, bad);
  NSX_HIP(ctx, hipGetLastError());
  unsigned int hb = 0;
  NSX_HIP(ctx, hipMemcpyAsync(&hb, bad, 4, hipMemcpyDeviceToHost, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dev_free(ctx, bad);
  dev_free(ctx, cnt);
  if (hb) return set_error(ctx, NSX_E_HIP, "transpose index incomplete");
  ctx->csc_m = tot;
  return NSX_OK;
}

// Sort (src, insertion index) stably, build row_ptr, gather col/w/delay. Consumes the inputs.
nsx_status build_csr_device(nsx_ctx* ctx, uint32_t* d_src, uint32_t* d_dst, float* d_w,
                            uint8_t* d_delay, uint64_t m) {
  const uint32_t n = (uint32_t)ctx->n;
  nsx_status s;
  if (m >= (1ull << 31))
    return set_error(ctx, NSX_E_INVALID, "more than 2^31-1 synapses per build call");
  free_network(ctx);
  uint32_t *keys_out = nullptr, *perm_in = nullptr, *perm_out = nullptr;
  if ((s = dev_alloc(ctx, (void**)&keys_out, m * 4, false))) return s;
  if ((s = dev_alloc(ctx, (void**)&perm_in, m * 4, false))) return s;
  if ((s = dev_alloc(ctx, (void**)&perm_out, m * 4, false))) return s;
  if (m) {
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(m, 256)), dim3(256), 0, ctx->stream, perm_in, m);
    NSX_HIP(ctx, hipGetLastError());
  }
  int end_bit = 1;
  while (end_bit < 32 && (1ull << end_bit) < (uint64_t)n) ++end_bit;
  size_t tmp_bytes = 0;
  NSX_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, d_src, keys_out, perm_in,
                                                  perm_out, (int)m, 0, end_bit, ctx->stream));
  void* tmp = nullptr;
  if ((s = dev_alloc(ctx, &tmp, tmp_bytes, false))) return s;
  if (m)
    NSX_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, d_src, keys_out, perm_in,
                                                    perm_out, (int)m, 0, end_bit, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dev_free(ctx, tmp);
  dev_free(ctx, perm_in);
  if ((s = dev_alloc(ctx, (void**)&ctx->row_ptr, ((size_t)n + 1) * 8, false))) return s;
  if ((s = dev_alloc(ctx, (void**)&ctx->col, m * 4, false))) return s;
  if ((s = dev_alloc(ctx, (void**)&ctx->w, m * 4, false))) return s;
  if ((s = dev_alloc(ctx, (void**)&ctx->delay, m, false))) return s;
  hipLaunchKernelGGL(k_row_ptr, dim3(blocks_for(m + 1, 256)), dim3(256), 0, ctx->stream,
                     keys_out, m, n, ctx->row_ptr);
  NSX_HIP(ctx, hipGetLastError());
  if (m) {
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(m, 256)), dim3(256), 0, ctx->stream, perm_out,
                       d_dst, d_w, d_delay, m, ctx->col, ctx->w, ctx->delay);
    NSX_HIP(ctx, hipGetLastError());
  }
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dev_free(ctx, keys_out);
  dev_free(ctx, perm_out);
  ctx->m = m;
  if (ctx->cfg.plasticity != NSX_PLAST_NONE && (s = build_csc_device(ctx))) return s;
  publish_network(ctx);
  return NSX_OK;
}

nsx_status validate_csr_device(nsx_ctx* ctx) {
  unsigned int* bad = nullptr;
  nsx_status s;
  if ((s = dev_alloc(ctx, (void**)&bad, 4, true))) return s;
  uint64_t cover = std::max<uint64_t>(ctx->m, ctx->n);
  hipLaunchKernelGGL(k_check_csr, dim3(blocks_for(cover + 1, 256)), dim3(256), 0, ctx->stream,
                     ctx->row_ptr, ctx->col, ctx->delay, (uint32_t)ctx->n, ctx->m,
                     (uint32_t)ctx->D, bad);
  NSX_HIP(ctx, hipGetLastError());
  unsigned int hb = 0;
  NSX_HIP(ctx, hipMemcpyAsync(&hb, bad, 4, hipMemcpyDeviceToHost, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dev_free(ctx, bad);
  if (hb & 1u) return set_error(ctx, NSX_E_BADCSR, "row_ptr is not monotone");
  if (hb & 2u) return set_error(ctx, NSX_E_BADCSR, "row_ptr[0] != 0 or row_ptr[n] != m");
  if (hb & 4u) return set_error(ctx, NSX_E_BADCSR, "column index out of range");
  if (hb & 8u) return set_error(ctx, NSX_E_BADCSR, "delay outside [1, ring_slots-1]");
  return NSX_OK;
}

}  // namespace nsx

using namespace nsx;

extern "C" {

nsx_status nsx_build_csr(nsx_ctx* ctx, const uint32_t* src, const uint32_t* dst,
                         const float* weight, const uint8_t* delay, uint64_t m) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (m && (!src || !dst || !weight || !delay))
    return set_error(ctx, NSX_E_INVALID, "NULL synapse array");
  uint32_t *ds = nullptr, *dd = nullptr;
  float* dw = nullptr;
  uint8_t* dl = nullptr;
  unsigned int* bad = nullptr;
  nsx_status s;
  if ((s = dev_alloc(ctx, (void**)&ds, m * 4, false)) ||
      (s = dev_alloc(ctx, (void**)&dd, m * 4, false)) ||
      (s = dev_alloc(ctx, (void**)&dw, m * 4, false)) ||
      (s = dev_alloc(ctx, (void**)&dl, m, false)) || (s = dev_alloc(ctx, (void**)&bad, 4, true)))
    return s;
  if (m) {
    NSX_HIP(ctx, hipMemcpyAsync(ds, src, m * 4, hipMemcpyHostToDevice, ctx->stream));
    NSX_HIP(ctx, hipMemcpyAsync(dd, dst, m * 4, hipMemcpyHostToDevice, ctx->stream));
    NSX_HIP(ctx, hipMemcpyAsync(dw, weight, m * 4, hipMemcpyHostToDevice, ctx->stream));
    NSX_HIP(ctx, hipMemcpyAsync(dl, delay, m, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_check_list, dim3(blocks_for(m, 256)), dim3(256), 0, ctx->stream, ds, dd,
                       dl, m, (uint32_t)ctx->n, (uint32_t)ctx->D, bad);
    NSX_HIP(ctx, hipGetLastError());
  }
  unsigned int hb = 0;
  NSX_HIP(ctx, hipMemcpyAsync(&hb, bad, 4, hipMemcpyDeviceToHost, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dev_free(ctx, bad);
  if (hb) {
    dev_free(ctx, ds); dev_free(ctx, dd); dev_free(ctx, dw); dev_free(ctx, dl);
    return set_error(ctx, NSX_E_BADCSR,
                     "synapse with source/target >= n_neurons or delay outside [1, ring_slots-1]");
  }
  s = build_csr_device(ctx, ds, dd, dw, dl, m);
  dev_free(ctx, ds); dev_free(ctx, dd); dev_free(ctx, dw); dev_free(ctx, dl);
  return s;
}

nsx_status nsx_generate_fixed_indegree(nsx_ctx* ctx, uint32_t k_in, float w_exc_lo,
                                       float w_exc_hi, float w_inh_lo, float w_inh_hi,
                                       uint32_t d_lo, uint32_t d_hi) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (ctx->n < 2) return set_error(ctx, NSX_E_INVALID, "generator needs n_neurons >= 2");
  if (d_lo < 1 || d_hi < d_lo || d_hi >= (uint32_t)ctx->D)
    return set_error(ctx, NSX_E_INVALID, "delays must satisfy 1 <= d_lo <= d_hi <= ring_slots-1");
  if (k_in > 256) return set_error(ctx, NSX_E_INVALID, "k_in must be <= 256");
  const uint32_t n = (uint32_t)ctx->n;
  const uint64_t m = (uint64_t)n * k_in;
  const uint32_t d_span = d_hi - d_lo + 1;
  nsx_status s;
  free_network(ctx);
  uint32_t *cnt = nullptr, *cur = nullptr;
  uint8_t* kk = nullptr;
  if ((s = dev_alloc(ctx, (void**)&cnt, (size_t)n * 4, true))) return s;
  if ((s = dev_alloc(ctx, (void**)&ctx->row_ptr, ((size_t)n + 1) * 8, true))) return s;
  if (m) {
    hipLaunchKernelGGL(k_gen_count, dim3(blocks_for(m, 256)), dim3(256), 0, ctx->stream,
                       ctx->cfg.seed, n, (uint32_t)ctx->cfg.n_exc, k_in, cnt);
    NSX_HIP(ctx, hipGetLastError());
  }
  // row_ptr[1..n] = inclusive sum of the out-degrees (64-bit accumulation); the longest row
  // must fit the row sort
  size_t tmp_bytes = 0, tmp2 = 0;
  uint32_t* dmax = nullptr;
  if ((s = dev_alloc(ctx, (void**)&dmax, 4, true))) return s;
  hipcub::TransformInputIterator<unsigned long long, ToU64, const uint32_t*> wide(cnt, ToU64());
  NSX_HIP(ctx, hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, wide,
                                                 (unsigned long long*)ctx->row_ptr + 1, (int)n,
                                                 ctx->stream));
  NSX_HIP(ctx, hipcub::DeviceReduce::Max(nullptr, tmp2, cnt, dmax, (int)n, ctx->stream));
  void* tmp = nullptr;
  if ((s = dev_alloc(ctx, &tmp, std::max(tmp_bytes, tmp2), false))) return s;
  NSX_HIP(ctx, hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, wide,
                                                 (unsigned long long*)ctx->row_ptr + 1, (int)n,
                                                 ctx->stream));
  NSX_HIP(ctx, hipcub::DeviceReduce::Max(tmp, tmp2, cnt, dmax, (int)n, ctx->stream));
  uint32_t maxdeg = 0;
  NSX_HIP(ctx, hipMemcpyAsync(&maxdeg, dmax, 4, hipMemcpyDeviceToHost, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dev_free(ctx, tmp);
  dev_free(ctx, dmax);
  if (maxdeg > 256) {
    dev_free(ctx, cnt);
    free_network(ctx);
    return set_error(ctx, NSX_E_INVALID, "an out-degree above 256 (rows are sorted in LDS)");
  }
  cur = cnt;  // the counts become cursors
  NSX_HIP(ctx, hipMemsetAsync(cur, 0, (size_t)n * 4, ctx->stream));
  if ((s = dev_alloc(ctx, (void**)&ctx->col, m * 4, false))) return s;
  if ((s = dev_alloc(ctx, (void**)&ctx->w, m * 4, false))) return s;
  if ((s = dev_alloc(ctx, (void**)&ctx->delay, m, false))) return s;
  if ((s = dev_alloc(ctx, (void**)&kk, m, false))) return s;
  if (m) {
    hipLaunchKernelGGL(k_gen_fill, dim3(blocks_for(m, 256)), dim3(256), 0, ctx->stream,
                       ctx->cfg.seed, n, (uint32_t)ctx->cfg.n_exc, k_in, w_exc_lo, w_exc_hi,
                       w_inh_lo, w_inh_hi, d_lo, d_span, ctx->row_ptr, cur, ctx->col, ctx->w,
                       ctx->delay, kk);
    NSX_HIP(ctx, hipGetLastError());
    unsigned int* bad = nullptr;
    if ((s = dev_alloc(ctx, (void**)&bad, 4, true))) return s;
    hipLaunchKernelGGL(k_check_fill, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream,
                       ctx->row_ptr, cur, n, bad);
    NSX_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_row_sort, dim3(blocks_for(n, 4)), dim3(256), 0, ctx->stream,
                       ctx->row_ptr, n, ctx->col, ctx->w, ctx->delay, kk);
    NSX_HIP(ctx, hipGetLastError());
    unsigned int hb = 0;
    NSX_HIP(ctx, hipMemcpyAsync(&hb, bad, 4, hipMemcpyDeviceToHost, ctx->stream));
    NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    dev_free(ctx, bad);
    if (hb) {
      dev_free(ctx, kk);
      dev_free(ctx, cnt);
      free_network(ctx);
      return set_error(ctx, NSX_E_HIP, "generator rows not filled exactly");
    }
  }
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dev_free(ctx, kk);
  dev_free(ctx, cnt);
  ctx->m = m;
  if ((s = validate_csr_device(ctx))) {
    free_network(ctx);
    return s;
  }
  if (ctx->cfg.plasticity != NSX_PLAST_NONE && (s = build_csc_device(ctx))) return s;
  publish_network(ctx);
  return NSX_OK;
}

nsx_status nsx_upload_csr(nsx_ctx* ctx, const uint64_t* row_ptr, const uint32_t* col,
                          const float* weight, const uint8_t* delay, uint64_t n, uint64_t m) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (n != ctx->n) return set_error(ctx, NSX_E_INVALID, "n does not match the context");
  if (!row_ptr || (m && (!col || !weight || !delay)))
    return set_error(ctx, NSX_E_INVALID, "NULL CSR array");
  free_network(ctx);
  nsx_status s;
  if ((s = dev_alloc(ctx, (void**)&ctx->row_ptr, (n + 1) * 8, false)) ||
      (s = dev_alloc(ctx, (void**)&ctx->col, m * 4, false)) ||
      (s = dev_alloc(ctx, (void**)&ctx->w, m * 4, false)) ||
      (s = dev_alloc(ctx, (void**)&ctx->delay, m, false)))
    return s;
  NSX_HIP(ctx, hipMemcpyAsync(ctx->row_ptr, row_ptr, (n + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
  if (m) {
    NSX_HIP(ctx, hipMemcpyAsync(ctx->col, col, m * 4, hipMemcpyHostToDevice, ctx->stream));
    NSX_HIP(ctx, hipMemcpyAsync(ctx->w, weight, m * 4, hipMemcpyHostToDevice, ctx->stream));
    NSX_HIP(ctx, hipMemcpyAsync(ctx->delay, delay, m, hipMemcpyHostToDevice, ctx->stream));
  }
  ctx->m = m;
  if ((s = validate_csr_device(ctx))) {
    free_network(ctx);
    return s;
  }
  if (ctx->cfg.plasticity != NSX_PLAST_NONE && (s = build_csc_device(ctx))) return s;
  publish_network(ctx);
  return NSX_OK;
}

nsx_status nsx_read_csr(nsx_ctx* ctx, uint64_t* row_ptr, uint32_t* col, float* weight,
                        uint8_t* delay) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (!ctx->has_network) return set_error(ctx, NSX_E_STATE, "no network loaded");
  if (weight) {
    nsx_status s = flush_pending(ctx, true);
    if (s) return s;
  }
  if (row_ptr)
    NSX_HIP(ctx, hipMemcpyAsync(row_ptr, ctx->row_ptr, (ctx->n + 1) * 8, hipMemcpyDeviceToHost,
                                ctx->stream));
  if (ctx->m) {
    if (col) NSX_HIP(ctx, hipMemcpyAsync(col, ctx->col, ctx->m * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (weight) NSX_HIP(ctx, hipMemcpyAsync(weight, ctx->w, ctx->m * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (delay) NSX_HIP(ctx, hipMemcpyAsync(delay, ctx->delay, ctx->m, hipMemcpyDeviceToHost, ctx->stream));
  }
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NSX_OK;
}

}  // extern "C"
