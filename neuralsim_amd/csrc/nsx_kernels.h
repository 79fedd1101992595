// nsx_kernels.h -- device helpers shared by the step kernels (plain models and the surrogate).
#pragma once
#include "nsx_device.h"
#include "nsx_internal.h"

namespace nsx {

__device__ inline float trace_at(const float* tab, int32_t tab_n, int32_t k) {
  return (k >= 0 && k < tab_n) ? tab[k] : 0.0f;
}

// Trace value at step t of a record last bumped at step s (lazy decay through the table).
__device__ inline float trace_now(float v, int32_t s, const float* tab, int32_t tab_n,
                                  int32_t t) {
  const float q = trace_at(tab, tab_n, t - s);
  return v * q;
}

// StdpTrace::on_pre_spike / on_post_spike (plasticity.rs:144-154) on a lazily decayed record:
// the per-step decay and the elapsed-time factor are both P[t - s].
__device__ inline TraceRec trace_bump(const TraceRec rec, const float* tab, int32_t tab_n,
                                      int32_t t) {
  const float q = trace_at(tab, tab_n, t - rec.s);
  float x = rec.v * q;
  x = x * q;
  return TraceRec{x + 1.0f, t, rec.v, rec.s};
}

// Synapse k of target j of a fixed in-degree network (n neurons, [0, n_exc) excitatory): source
// uniform over the other n-1 neurons, weight uniform in the source type's range, delay uniform in
// [d_lo, d_lo + d_span).  Counter-based: the same (seed, j, k) gives the same synapse wherever it
// is generated (alone or inside a batched population).
__device__ inline void gen_synapse(uint64_t seed, uint32_t n, uint32_t n_exc, float w_exc_lo,
                                   float w_exc_hi, float w_inh_lo, float w_inh_hi, uint32_t d_lo,
                                   uint32_t d_span, uint32_t j, uint32_t k, uint32_t& src,
                                   float& w, uint8_t& delay) {
  U4 r = draw(seed, j, k, kStreamConn);
  uint32_t s = (uint32_t)(((uint64_t)r.x * (uint64_t)(n - 1)) >> 32);
  if (s >= j) s += 1;
  float u = u01(r.y);
  const bool exc = s < n_exc;
  float lo = exc ? w_exc_lo : w_inh_lo, hi = exc ? w_exc_hi : w_inh_hi;
  float span = hi - lo;
  float x = span * u;
  src = s;
  w = lo + x;
  delay = (uint8_t)(d_lo + (uint32_t)(((uint64_t)r.z * d_span) >> 32));
}

// RNG key, local index and network size of neuron i of block blk (see NeuronId).  The table
// entry is read with relaxed atomic loads: plain loads here and the parameter block's fields
// in the other branch get merged into one load through a selected pointer, a vector load the
// kernel then has to wait for.
__device__ inline NeuronId neuron_id(const StepParams& p, const DevPtrs& s, uint32_t i,
                                     uint32_t blk) {
  if (!p.batch) return NeuronId{p.seed, i, p.n, p.n_exc};
  const BatchNet* b = s.bnet + s.blk_net[blk];
  constexpr int kWg = __HIP_MEMORY_SCOPE_WORKGROUP;
  const uint64_t seed = __hip_atomic_load(&b->seed, __ATOMIC_RELAXED, kWg);
  const uint32_t off = __hip_atomic_load(&b->offset, __ATOMIC_RELAXED, kWg);
  const uint32_t n = __hip_atomic_load(&b->n, __ATOMIC_RELAXED, kWg);
  const uint32_t ne = __hip_atomic_load(&b->n_exc, __ATOMIC_RELAXED, kWg);
  return NeuronId{seed, i - off, n, ne};
}

// Whether neuron i is excitatory: in [0, n_exc) of the network, or of its own network in a
// batched population (the block table names it).  The single network's test costs no load.
__device__ inline bool exc_neuron(const StepParams& p, const DevPtrs& s, uint32_t i) {
  if (!p.batch) return i < p.n_exc;
  const BatchNet& b = s.bnet[s.blk_net[i / kBlock]];
  return i - b.offset < b.n_exc;
}

// Drive words of one step for a block's neurons.  One Philox draw covers four consecutive
// neurons of a network (local indices 4g .. 4g+3; blocks start at a multiple of 256 of their
// network): wavefront 0 draws the block's 64 groups into w[0 .. kBlock).  The caller synchronises
// the block before drive_draw reads them.  `id` is the calling thread's NeuronId, computed
// before the neuron's loads were issued: looking the network's key up here again would be a
// load after them, and waiting for it (memory counters retire in order) would wait for all of
// them before the draws could start.
template <int DRIVE>
__device__ inline void drive_stage(const StepParams& p, const DevPtrs& s, int32_t t, uint32_t* w,
                                   const NeuronId& id) {
  if (DRIVE == NSX_DRIVE_NONE) return;
  if (threadIdx.x < kBlock / 4) {
    const uint32_t local0 = id.local - threadIdx.x;  // the block's first neuron
    const U4 r = draw(id.seed, (uint64_t)(local0 >> 2) + threadIdx.x, (uint32_t)t,
                      kStreamDrive);
    w[4 * threadIdx.x] = r.x;
    w[4 * threadIdx.x + 1] = r.y;
    w[4 * threadIdx.x + 2] = r.z;
    w[4 * threadIdx.x + 3] = r.w;
  }
}

// Per-step drive value of the calling thread's neuron from the staged words (value added to
// channel 0 / the input current).  Poisson: word (local & 3) by inversion.  Gaussian: the words
// 2k, 2k+1 of pair k = (local & 3) / 2 by Box-Muller (1958), the cosine for the even neuron and
// the sine for the odd one -- the angle word advanced by three quarter turns, exact on the 24-bit
// grid of the angle.
template <int DRIVE>
__device__ inline void drive_draw(const StepParams& p, const uint32_t* w, float& a0, float& z) {
  if (DRIVE == NSX_DRIVE_NONE) return;
  const uint32_t k = threadIdx.x;
  if (DRIVE == NSX_DRIVE_POISSON) {
    // the first four thresholds at constant offsets (fetched by scalar loads up front); the
    // rest only in the rare wavefront where some neuron's word passes the fourth (k >= 4)
    const uint32_t x = w[k];
    uint32_t c = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) c += (q < p.npthr && x >= p.pthr[q]) ? 1u : 0u;
    if (p.npthr > 4 && x >= p.pthr[3])
      for (int q = 4; q < p.npthr; ++q) c += (x >= p.pthr[q]) ? 1u : 0u;
    float f = (float)c * p.drive_weight;
    a0 = a0 + f;
  } else {
    const uint32_t b = (k & ~3u) | (k & 2u);
    const uint32_t angle = w[b + 1] + ((k & 1u) ? 0xC0000000u : 0u);
    z = normal_(w[b], angle);
  }
}

// Pair STDP: the post-synaptic trace bump of neuron i's spike at step t-1, deferred to the update
// of step t so that the synapse pass of step t-1 read the record without it.  `last` is the
// neuron's last spike step as loaded at the start of the update (not loaded, and kNeverStep, in a
// block whose spike segment of step t-1 is empty); each thread bumps its own neuron, before any
// spike_record of step t.
// A list of pending depressions that filled at step t-1 was applied by that step's synapse pass
// and is emptied here.
template <bool PAIR>
__device__ inline void post_bump_prev(const StepParams& p, const DevPtrs& s, uint32_t i,
                                      int32_t last, int32_t t) {
  if (PAIR && last == t - 1 && exc_neuron(p, s, i)) {
    PostRec& r = s.post[i];
    const float v = r.v;
    const int32_t st = r.s;
    const uint32_t np = r.npend;
    const float q = trace_at(s.mtab, p.mtab_n, (t - 1) - st);
    float x = v * q;
    x = x * q;
    r.v = x + 1.0f;
    r.s = t - 1;
    if (np == kPostPend) r.npend = 0;
  }
}

// Neuron i's ring slot of step t: ring_load reads channel 0 (and channel 1 when the network has
// two, TWO), ring_take later converts and clears -- a slot that received nothing is not
// rewritten.  Split so that the update kernels issue all of a neuron's loads at once.
struct RingIn {
  long long* r0;
  long long q0, q1;
};

// Whether ring slot t can hold input: some neuron spiked in [t - D + 1, t - 1] (a spike's events
// arrive 1 .. D - 1 steps later) or the host wrote the ring.  Uniform across the grid, so a step
// of a silent network skips every ring load.
__device__ inline bool ring_active(const StepParams& p, const DevPtrs& s, int32_t t) {
  int32_t m = kNeverStep;
#pragma unroll
  for (uint32_t k = 0; k < kEvStripes; ++k) m = max(m, s.ev_last[k * kEvStride]);
  return m >= t - (int32_t)p.D + 1;
}

template <bool TWO>
__device__ inline RingIn ring_load(const StepParams& p, const DevPtrs& s, uint32_t i, int32_t t,
                                   bool active) {
  if (!active) return RingIn{nullptr, 0, 0};
  const uint32_t slot = (uint32_t)t % p.D;
  RingIn r;
  r.r0 = s.ring + (size_t)slot * p.ch * p.n + i;
  r.q0 = r.r0[0];
  r.q1 = TWO ? r.r0[p.n] : 0;
  return r;
}

template <bool TWO>
__device__ inline void ring_take(const StepParams& p, const RingIn& r, float& a0, float& a1) {
  a0 = dequantize(r.q0, p.scale_out);
  if (r.q0 != 0) r.r0[0] = 0;
  a1 = 0.0f;
  if (TWO) {
    a1 = dequantize(r.q1, p.scale_out);
    if (r.q1 != 0) r.r0[p.n] = 0;
  }
}

// Bookkeeping of a spike of neuron i at step t: count, last spike step; for excitatory neurons
// under pair STDP the pre-synaptic trace bump and the post record's same-step flag, under
// triplet STDP the post record's (previous, last) spike steps.  PL is an nsx_plasticity_kind.
template <int PL>
__device__ inline void spike_record(const StepParams& p, const DevPtrs& s, uint32_t i,
                                    int32_t t) {
  s.count[i] += 1u;
  s.last[i] = t;
  if (PL == NSX_PLAST_PAIR && exc_neuron(p, s, i)) {
    s.pre[i] = trace_bump(s.pre[i], s.ptab, p.ptab_n, t);
    PostRec& r = s.post[i];
    r.last = t;
    const uint32_t np = r.npend;  // < kPostPend: a full list was emptied by post_bump_prev
    if (np < kPostPend) {
      r.pend[np] = t;  // this spike's depression is pending
      r.npend = np + 1;
    }
  }
  if (PL == NSX_PLAST_TRIPLET && exc_neuron(p, s, i)) {
    PostRec r = s.post[i];
    r.s = r.last;
    r.last = t;
    s.post[i] = r;
  }
}

// Block-local compaction of the spiking neurons of this 256-neuron block into its spike segment
// (and the raster when recording).  Called by every thread of the block.
// Adds the block's sum of x (one value per thread) to lfp_part[blk] in a fixed order (double).
// Every thread of the block calls it.
__device__ inline void lfp_accumulate(const DevPtrs& s, float x, uint32_t blk) {
  __shared__ double red[kBlock / 64];
  double v = (double)x;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0;
#pragma unroll
    for (uint32_t k = 0; k < kBlock / 64; ++k) a += red[k];
    s.lfp_part[blk] += a;
  }
}

__device__ inline void block_compact(const StepParams& p, const DevPtrs& s, bool spiked,
                                     uint32_t i, int32_t t, uint32_t blk) {
  // Block-local compaction into this block's spike segment: wave ballots, wave counts in LDS,
  // no global atomics (one atomic per wave on a shared counter serialises at the memory-side
  // atomic unit once tens of thousands of waves fire).
  __shared__ uint32_t s_wcnt[kBlock / 64];
  const unsigned long long mask = __ballot(spiked);
  const uint32_t lane = __lane_id(), wid = threadIdx.x >> 6;
  const uint32_t nset = (uint32_t)__popcll(mask);
  const uint32_t below = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  if (lane == 0) s_wcnt[wid] = nset;
  __syncthreads();
  uint32_t off_w = 0, tot = 0;
#pragma unroll
  for (uint32_t k = 0; k < kBlock / 64; ++k) {
    const uint32_t c = s_wcnt[k];
    off_w += k < wid ? c : 0u;
    tot += c;
  }
  if (spiked) s.spikes[blk * kBlock + off_w + below] = i;
  if (threadIdx.x == 0) {
    s.bcnt[blk] = tot;
    if (tot) s.part[kPartSpikes * p.nblk + blk] += tot;
    if (tot) s.ev_last[(blk % kEvStripes) * kEvStride] = t;
    if (p.stats_on && tot)
      atomicAdd(&s.stat_counts[((uint32_t)t % p.stat_cap) * kStatStripes + blk % kStatStripes],
                tot);
  }
  if (p.raster_on && mask) {  // test/analysis path: append (step, id) pairs
    const int leader = __ffsll((long long)mask) - 1;
    unsigned long long rb = 0;
    if ((int)lane == leader) {
      rb = atomicAdd(&s.counters[kCntRaster], (unsigned long long)nset);
      atomicAdd(&s.step_counts[(uint32_t)t % p.step_counts_cap], nset);
    }
    rb = __shfl(rb, leader);
    if (spiked) {
      unsigned long long pos = rb + below;
      if (pos < p.raster_cap) {
        s.raster_step[pos] = (uint32_t)t;
        s.raster_id[pos] = i;
      }
    }
  }
}

}  // namespace nsx
