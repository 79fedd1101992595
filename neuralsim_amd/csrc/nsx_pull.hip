// nsx_pull.hip -- pull delivery (gfx950, wave64) for networks without plasticity.
//
// The reference pushes each spike's events into pending_input[t + d] (engine.rs:414-494); the
// push kernels here do the same with ring atomics or bucketed records.  Pull delivery computes
// the same sums from the other side: after the update of step t, one workgroup per target tile
// gathers the input of step t + 1 -- for d = 1 .. D - 1, the delay-d synapses into the tile of
// every neuron that spiked at step t + 1 - d -- into LDS int64 accumulators and adds them into
// ring slot t + 1.  The integer sums equal the pushed ones bit for bit.  No event is written to
// or read back from HBM and no atomic leaves the CU; the synapse data a tile reads is exactly its
// slices, because every row keeps a copy of its synapses sorted by (target range, delay) with a
// u16 offset table (k_pull_sort).  A tile covers a range of at most kPullTE / channels targets of
// one network, so in a batched population a tile reads only its own network's rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "nsx_device.h"
#include "nsx_internal.h"
#include "nsx_kernels.h"

namespace nsx {

// One thread per row: the row's synapses re-sorted by key = part(target) * (D - 1) + d - 1,
// stable (targets keep their CSR order within a key), into pdsyn at the row's CSR position; the
// key starts and the degree into the row's offset table.  bad: 1 a degree of 65536 or more, 2 a
// target outside its network (no row of a batched population may leave its network).
__global__ void k_pull_sort(StepParams p, DevPtrs s, const PullNet* nets, uint2* dsyn,
                            uint16_t* doff, unsigned int* bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  const PullNet nt = nets[p.batch ? s.blk_net[i / kBlock] : 0u];
  const uint32_t local = i - nt.offset;
  if (local >= nt.n) return;  // padding of a batched population: an empty row
  const uint64_t beg = s.row_ptr[i], end = s.row_ptr[i + 1];
  if (end - beg >= 65536u) {
    atomicOr(bad, 1u);
    return;
  }
  const uint32_t keys = nt.stride - 1u, dm = p.D - 1u;
  uint32_t cnt[kPullMaxKeys];
  for (uint32_t k = 0; k < keys; ++k) cnt[k] = 0;
  for (uint64_t k = beg; k < end; ++k) {
    const uint32_t jl = s.col[k] - nt.offset;
    const uint32_t key = (jl / nt.L) * dm + (uint32_t)s.delay[k] - 1u;
    if (jl >= nt.n || key >= keys) {
      atomicOr(bad, 2u);
      return;
    }
    cnt[key] += 1u;
  }
  uint16_t* off = doff + nt.doff_base + (uint64_t)local * nt.stride;
  uint32_t run = 0;
  for (uint32_t k = 0; k < keys; ++k) {
    off[k] = (uint16_t)run;
    const uint32_t c = cnt[k];
    cnt[k] = run;
    run += c;
  }
  off[keys] = (uint16_t)run;
  for (uint64_t k = beg; k < end; ++k) {
    const uint32_t key = ((s.col[k] - nt.offset) / nt.L) * dm + (uint32_t)s.delay[k] - 1u;
    dsyn[beg + cnt[key]] = make_uint2(s.col[k], __float_as_uint(s.w[k]));
    cnt[key] += 1u;
  }
}

// Order LDS traffic of one wavefront (all lanes write, then any lane reads).
__device__ inline void pull_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr uint32_t kPullThreads = 1024, kPullWaves = kPullThreads / 64;
constexpr uint32_t kPullSpk = 1024;  // (spike, delay) items staged in LDS per pass
constexpr uint32_t kPullUnroll = 8;  // synapses per lane in flight while walking the slices

// Walk the slices of the ns staged items: each wavefront takes 64 items at a time, looks up
// their slices (the key (tile's part, delay) of the source's row, whose start was staged with
// the spike), and walks them flattened, kPullUnroll synapses per lane per round with every load
// issued before the LDS atomics.
__device__ inline void pull_walk(const StepParams& p, const DevPtrs& s, const PullTile& tl,
                                 long long* acc, const uint32_t* spk, const uint32_t* srow,
                                 const uint8_t* spd, uint32_t ns, uint64_t (*ibeg)[64],
                                 uint32_t (*iend)[64], int32_t (*ient)[64], uint32_t lane,
                                 uint32_t wid) {
  for (uint32_t i0 = wid * 64; i0 < ns; i0 += kPullThreads) {
    const uint32_t it = i0 + lane;
    uint64_t beg = 0;
    uint32_t len = 0;
    int32_t ent = 0;
    if (it < ns) {
      const uint32_t src = spk[it], d = spd[it];
      const uint16_t* o =
          s.pdoff + tl.doff_base + (uint64_t)(src - tl.slo) * tl.stride + tl.key0 + d - 1u;
      const uint32_t x0 = o[0], x1 = o[1];
      beg = (uint64_t)srow[it] + x0;
      len = x1 - x0;
      const uint32_t ch = (p.ch > 1 && !exc_neuron(p, s, src)) ? 1u : 0u;
      ent = (int32_t)(ch * tl.L) - (int32_t)tl.lo;
    }
    uint32_t inc = len;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o);
      if (lane >= o) inc += y;
    }
    const uint32_t tot = __builtin_amdgcn_readfirstlane(__shfl(inc, 63));
    ibeg[wid][lane] = beg;
    iend[wid][lane] = inc;
    ient[wid][lane] = ent;
    pull_wave_sync();
    for (uint32_t f0 = 0; f0 < tot; f0 += 64 * kPullUnroll) {
      uint2 x[kPullUnroll];
      int32_t en[kPullUnroll];
#pragma unroll
      for (uint32_t u = 0; u < kPullUnroll; ++u) {
        // a flat index past the end repeats the last synapse's load (in bounds), adds nothing
        const uint32_t f = min(f0 + u * 64 + lane, tot - 1u);
        uint32_t l = 0;  // the item of flat index f: lanes whose inclusive end is <= f
#pragma unroll
        for (uint32_t st = 32; st > 0; st >>= 1)
          if (iend[wid][l + st - 1] <= f) l += st;
        const uint32_t start = l ? iend[wid][l - 1] : 0u;
        x[u] = s.pdsyn[ibeg[wid][l] + (f - start)];
        en[u] = ient[wid][l];
      }
#pragma unroll
      for (uint32_t u = 0; u < kPullUnroll; ++u)
        if (f0 + u * 64 + lane < tot)
          atomicAdd((unsigned long long*)&acc[(uint32_t)((int32_t)x[u].x + en[u])],
                    (unsigned long long)quantize(__uint_as_float(x[u].y), p.scale_in));
    }
    pull_wave_sync();
  }
}

// Arrival at step a of the spikes of steps [max(*pull_t0, a - D + 1), e_hi] for tile blockIdx.x.
// Step mode (mat = 0): t = *tbase + off is the step just updated, a = t + 1, e_hi = t; step t's
// spikes are read from the current segments, and the tile of part 0 of their network copies them
// into the history and counts the events they emit.  Materialise mode (mat = 1): a = *tbase +
// off, e_hi = e_cap, every spike from the history.  The (spike, delay) items of all those steps
// are staged in LDS in passes of kPullSpk and their slices walked (pull_walk) into the LDS
// accumulators of (channel of the source, target); the accumulators are then added into ring
// slot a (entries that received nothing are not rewritten; no other kernel writes that slot
// meanwhile).
__global__ __launch_bounds__(kPullThreads) void k_pull(StepParams p, DevPtrs s, int32_t off,
                                                        int32_t e_cap, uint32_t mat) {
  __shared__ long long acc[kPullTE];
  __shared__ uint32_t spk[kPullSpk];
  __shared__ uint32_t srow[kPullSpk];
  __shared__ uint8_t spd[kPullSpk];
  __shared__ uint64_t ibeg[kPullWaves][64];
  __shared__ uint32_t iend[kPullWaves][64];
  __shared__ int32_t ient[kPullWaves][64];
  __shared__ uint32_t wsum[kPullWaves];
  __shared__ unsigned long long wev[kPullWaves];
  __shared__ uint32_t bend[kPullThreads];  // inclusive end of each block's items in the chunk
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
  const PullTile tl = s.ptile[blockIdx.x];
  const uint32_t D = p.D, nblk = p.nblk;
  const int32_t t = *s.tbase + off;
  const int32_t a = mat ? t : t + 1;
  const int32_t e_hi = mat ? e_cap : t;
  const int32_t e_lo = max(*s.pull_t0, a - (int32_t)(D - 1u));
  const bool copy = !mat && tl.key0 == 0;  // this tile keeps its network's history
  for (uint32_t k = tid; k < kPullTE; k += kPullThreads) acc[k] = 0;
  const uint32_t b0 = tl.slo / kBlock, b1 = (tl.shi + kBlock - 1u) / kBlock;
  unsigned long long ev = 0;
  uint32_t fill = 0;  // items staged (uniform)
  __syncthreads();
  for (int32_t e = e_lo; e <= e_hi; ++e) {
    const uint32_t d = (uint32_t)(a - e);
    const bool cur = !mat && e == t;
    const size_t hs = (size_t)((uint32_t)e % D) * nblk;
    const uint32_t* cnt = cur ? s.bcnt : s.phist_cnt + hs;
    const uint32_t* seg = cur ? s.spikes : s.phist_spk + hs * kBlock;
    const uint32_t* segr = s.phist_row + hs * kBlock;
    for (uint32_t bc = b0; bc < b1; bc += kPullThreads) {
      const uint32_t b = bc + tid;
      const uint32_t c = b < b1 ? cnt[b] : 0u;
      if (cur && copy && b < b1) s.phist_cnt[hs + b] = c;
      uint32_t incl = c;
#pragma unroll
      for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
      }
      if (lane == 63) wsum[wid] = incl;
      __syncthreads();
      uint32_t pre = 0, S = 0;
#pragma unroll
      for (uint32_t q = 0; q < kPullWaves; ++q) {
        const uint32_t x = wsum[q];
        pre += q < wid ? x : 0u;
        S += x;
      }
      bend[tid] = pre + incl;
      __syncthreads();  // wsum is rewritten by the next chunk; bend complete
      const uint32_t nb = min(kPullThreads, b1 - bc);
      for (uint32_t done = 0; done < S;) {
        const uint32_t take = min(S - done, kPullSpk - fill);
        // item q of the chunk: block i with bend[i - 1] <= q < bend[i], its spike q - bend[i - 1]
        for (uint32_t q = done + tid; q < done + take; q += kPullThreads) {
          uint32_t lo = 0, hi = nb;  // first block whose end exceeds q
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (bend[mid] <= q) lo = mid + 1;
            else hi = mid;
          }
          const uint32_t k = q - (lo ? bend[lo - 1] : 0u);
          const uint32_t bb = bc + lo;
          const uint32_t src = seg[bb * kBlock + k];
          uint32_t r0;
          if (cur) {  // step t's spike: its row start, once (the history keeps it)
            const uint64_t rb = s.row_ptr[src];
            r0 = (uint32_t)rb;
            if (copy) {
              s.phist_spk[(hs + bb) * kBlock + k] = src;
              s.phist_row[(hs + bb) * kBlock + k] = r0;
              ev += s.row_ptr[src + 1] - rb;
            }
          } else {
            r0 = segr[bb * kBlock + k];
          }
          spk[fill + q - done] = src;
          srow[fill + q - done] = r0;
          spd[fill + q - done] = (uint8_t)d;
        }
        fill += take;
        done += take;
        __syncthreads();
        if (fill == kPullSpk) {
          pull_walk(p, s, tl, acc, spk, srow, spd, fill, ibeg, iend, ient, lane, wid);
          fill = 0;
          __syncthreads();  // the items are restaged
        }
      }
      __syncthreads();  // bend is rewritten by the next chunk
    }
  }
  if (fill) pull_walk(p, s, tl, acc, spk, srow, spd, fill, ibeg, iend, ient, lane, wid);
  if (copy) {  // step t's events, added to the network's first block
#pragma unroll
    for (uint32_t o = 32; o > 0; o >>= 1) ev += __shfl_xor(ev, o);
    if (lane == 0) wev[wid] = ev;
  }
  __syncthreads();
  if (copy && tid == 0) {
    unsigned long long tot = 0;
    for (uint32_t q = 0; q < kPullWaves; ++q) tot += wev[q];
    s.part[kPartEvents * nblk + b0] += tot;
  }
  // ring slot a += accumulators, eight entries per thread per round with the loads first
  long long* ring = s.ring + (size_t)((uint32_t)a % D) * p.ch * p.n;
  const uint32_t tn = tl.hi - tl.lo, ne = p.ch * tl.L;
  for (uint32_t k0 = 0; k0 < ne; k0 += 8 * kPullThreads) {
    long long add[8], r[8];
    size_t idx[8];
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) {
      const uint32_t k = k0 + u * kPullThreads + tid;
      const uint32_t c = k / tl.L, jl = k - c * tl.L;
      add[u] = k < ne && jl < tn ? acc[k] : 0;
      idx[u] = (size_t)c * p.n + tl.lo + jl;
      r[u] = add[u] != 0 ? ring[idx[u]] : 0;
    }
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u)
      if (add[u] != 0) ring[idx[u]] = r[u] + add[u];
  }
}

// ---- host side ------------------------------------------------------------------------------

// Part length and parts of a network of n neurons and m synapses: at least ceil(n / Lmax) parts
// (Lmax x channels LDS accumulators), and more for a network of many synapses, so that no tile
// gathers much more than m_tile synapses' events (the kernel lasts as long as its busiest tile),
// within kPullMaxKeys keys per row; L = ceil(n / parts).
static void pull_parts(const nsx_ctx* ctx, uint32_t n, uint64_t m, uint64_t m_tile, uint32_t& L,
                       uint32_t& parts) {
  const uint32_t lmax = kPullTE / (uint32_t)ctx->ch;
  parts = std::max<uint32_t>((n + lmax - 1) / lmax, 1u);
  const uint32_t kmax = std::max<uint32_t>(kPullMaxKeys / ((uint32_t)ctx->D - 1u), 1u);
  if (m_tile) {
    const uint64_t by_m = (m + m_tile - 1) / m_tile;
    parts = (uint32_t)std::max<uint64_t>(parts, std::min<uint64_t>({by_m, kmax, n}));
  }
  L = std::max<uint32_t>((n + parts - 1) / parts, 1u);  // equal parts
}

// The networks of the context: a batch's, or the whole network as one.
static std::vector<PullNet> pull_nets(const nsx_ctx* ctx, uint64_t& doff_total) {
  std::vector<PullNet> nets;
  doff_total = 0;
  // a batch: tiles of about m / CUs synapses at most
  uint64_t m_tile = 0;
  if (ctx->p.batch) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
            hipSuccess ||
        cus < 1)
      cus = 1;
    m_tile = std::max<uint64_t>(ctx->m / (uint64_t)cus, 1u << 16);
  }
  auto add = [&](uint32_t offset, uint32_t n, uint64_t m) {
    PullNet q{};
    pull_parts(ctx, n, m, m_tile, q.L, q.parts);
    q.offset = offset;
    q.n = n;
    q.stride = q.parts * ((uint32_t)ctx->D - 1u) + 1u;
    q.doff_base = doff_total;
    doff_total += (uint64_t)n * q.stride;
    nets.push_back(q);
  };
  if (ctx->p.batch) {
    for (const BatchNet& b : ctx->nets_h) add(b.offset, b.n, (uint64_t)b.n * b.k_in);
  } else {
    add(0, (uint32_t)ctx->n, ctx->m);
  }
  return nets;
}

bool pull_supported(const nsx_ctx* ctx) {
  if (!ctx->has_network || ctx->cfg.plasticity != NSX_PLAST_NONE || ctx->backend_mode ||
      ctx->D < 2 || ctx->D > 256 || ctx->n >= (1ull << 31) ||  // delays fit a byte
      ctx->m >= (1ull << 32))                                   // row starts fit 32 bits
    return false;
  uint64_t total = 0;
  for (const PullNet& q : pull_nets(ctx, total))
    if (q.stride - 1u > kPullMaxKeys) return false;
  return true;
}

void pull_release(nsx_ctx* ctx) {
  DevPtrs& d = ctx->d;
  dev_free(ctx, (void*)d.pdsyn);
  dev_free(ctx, (void*)d.pdoff);
  dev_free(ctx, (void*)d.ptile);
  dev_free(ctx, d.phist_spk);
  dev_free(ctx, d.phist_row);
  dev_free(ctx, d.phist_cnt);
  dev_free(ctx, d.pull_t0);
  d.pdsyn = nullptr;
  d.pdoff = nullptr;
  d.ptile = nullptr;
  d.phist_spk = d.phist_row = d.phist_cnt = nullptr;
  d.pull_t0 = nullptr;
  ctx->p.pull_on = 0;
  ctx->p.pull_ntiles = 0;
  ctx->pull_built = false;
}

// The spikes of steps before ctx->t have their input in the ring: the gathers start at ctx->t.
nsx_status pull_clear(nsx_ctx* ctx) {
  if (!ctx->p.pull_on) return NSX_OK;
  const int32_t t0 = (int32_t)ctx->t;
  NSX_HIP(ctx, hipMemcpy(ctx->d.pull_t0, &t0, 4, hipMemcpyHostToDevice));
  return NSX_OK;
}

nsx_status pull_configure(nsx_ctx* ctx) {
  if (!pull_supported(ctx))
    return set_error(ctx, NSX_E_INVALID,
                     "pull delivery needs a loaded network without plasticity, not the GpuBackend "
                     "mode, and (tiles of a network) x (D - 1) <= 256");
  if (!ctx->pull_built) {
    uint64_t doff_total = 0;
    const std::vector<PullNet> nets = pull_nets(ctx, doff_total);
    std::vector<PullTile> tiles;
    for (const PullNet& q : nets)
      for (uint32_t part = 0; part < q.parts; ++part) {
        PullTile tl{};
        tl.doff_base = q.doff_base;
        tl.lo = q.offset + part * q.L;
        tl.hi = std::min(q.offset + q.n, tl.lo + q.L);
        tl.slo = q.offset;
        tl.shi = q.offset + q.n;
        tl.key0 = part * ((uint32_t)ctx->D - 1u);
        tl.stride = q.stride;
        tl.L = q.L;
        if (tl.lo < tl.hi) tiles.push_back(tl);
      }
    DevPtrs& d = ctx->d;
    PullNet* dn = nullptr;
    unsigned int* bad = nullptr;
    nsx_status s;
    if ((s = dev_alloc(ctx, (void**)&d.pdsyn, std::max<uint64_t>(ctx->m, 1) * 8, false)) ||
        (s = dev_alloc(ctx, (void**)&d.pdoff, std::max<uint64_t>(doff_total, 1) * 2, false)) ||
        (s = dev_alloc(ctx, (void**)&d.ptile, tiles.size() * sizeof(PullTile), false)) ||
        (s = dev_alloc(ctx, (void**)&d.phist_spk, (size_t)ctx->D * ctx->p.nblk * kBlock * 4,
                       false)) ||
        (s = dev_alloc(ctx, (void**)&d.phist_row, (size_t)ctx->D * ctx->p.nblk * kBlock * 4,
                       false)) ||
        (s = dev_alloc(ctx, (void**)&d.phist_cnt, (size_t)ctx->D * ctx->p.nblk * 4, true)) ||
        (s = dev_alloc(ctx, (void**)&d.pull_t0, 4, true)) ||
        (s = dev_alloc(ctx, (void**)&dn, nets.size() * sizeof(PullNet), false)) ||
        (s = dev_alloc(ctx, (void**)&bad, 4, true))) {
      dev_free(ctx, dn);
      dev_free(ctx, bad);
      pull_release(ctx);
      return s;
    }
    hipError_t e = hipMemcpy((void*)d.ptile, tiles.data(), tiles.size() * sizeof(PullTile),
                             hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(dn, nets.data(), nets.size() * sizeof(PullNet), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_pull_sort, dim3((uint32_t)((ctx->n + 255) / 256)), dim3(256), 0,
                         ctx->stream, ctx->p, ctx->d, dn, (uint2*)d.pdsyn, (uint16_t*)d.pdoff,
                         bad);
      e = hipGetLastError();
    }
    unsigned int hb = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&hb, bad, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, dn);
    dev_free(ctx, bad);
    if (e != hipSuccess) {
      pull_release(ctx);
      return hip_error(ctx, e, "pull delivery tables");
    }
    if (hb) {
      pull_release(ctx);
      return set_error(ctx, NSX_E_INVALID,
                       hb & 1u ? "pull delivery needs degrees below 65536"
                               : "pull delivery needs every target inside its source's network");
    }
    ctx->p.pull_ntiles = (uint32_t)tiles.size();
    ctx->pull_built = true;
  }
  ctx->p.pull_on = 1;
  return pull_clear(ctx);
}

hipError_t launch_pull(nsx_ctx* ctx, int32_t off) {
  hipLaunchKernelGGL(k_pull, dim3(std::max(ctx->p.pull_ntiles, 1u)), dim3(kPullThreads), 0,
                     ctx->stream, ctx->p, ctx->d, off, 0, 0u);
  return hipGetLastError();
}

// The gathers of steps up to ctx->t have run; the spikes of steps pull_t0 .. t - 1 still owe
// input to steps t + 1 .. t + D - 2, which is added into the ring now.
nsx_status pull_materialize(nsx_ctx* ctx) {
  if (!ctx->p.pull_on) return NSX_OK;
  const int64_t t = (int64_t)ctx->t;
  for (int64_t ta = t + 1; t > 0 && ta <= t + ctx->D - 2; ++ta) {
    hipLaunchKernelGGL(k_pull, dim3(std::max(ctx->p.pull_ntiles, 1u)), dim3(kPullThreads), 0,
                       ctx->stream, ctx->p, ctx->d, (int32_t)(ta - ctx->tbase_host),
                       (int32_t)(t - 1), 1u);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(ctx, e, "pull delivery arrival");
  }
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return pull_clear(ctx);
}

}  // namespace nsx
