// nsx_step.hip -- per-step kernels of the simulation core (gfx950, wave64).
//
// One step t is two launches on one stream:
//   k_update     one thread per neuron: read+clear ring slot t, external drive, model update
//                (LIF / Izhikevich / HH), block-local spike compaction into the block's spike
//                segment, pre-synaptic trace bump; first applies the post-synaptic trace bumps
//                of the block's spikes of step t-1.  Replaces SimulationEngine::step phase 1
//                (engine.rs:267-391) and the lif/izhikevich/hh_update.wgsl shaders.
//   k_synapses   two kinds of wavefront work items in one grid, interleaved:
//                * propagation: a neuron block's spiking rows flattened across the 64 lanes
//                  (coalesced col / weight / delay streams); each slot adds rint(w 2^S) into
//                  ring[(t+d) mod D][channel][target] with a 64-bit integer atomic; E->E slots
//                  get STDP at the pre spike (pair: potentiation by the post trace, then
//                  depression when the target fired at t too; triplet: depression)
//                  (engine.rs:414-473 with delays, engine.rs:512-534; spmv.wgsl);
//                * post-spike: E->E in-slots (transpose index) of a block's spiking excitatory
//                  neurons whose source did not fire at t (pair: depression, engine.rs:524-534;
//                  triplet: potentiation).
//                Every weight has one writer per step, so the kinds run concurrently: the
//                atomic-throughput-bound propagation overlaps the latency-bound depression.
// Kernels read the step index as *tbase + off so that a captured hipGraph of G steps can be
// replayed (k_advance moves tbase by G at the end of the graph).
#include <algorithm>

#include "nsx_device.h"
#include "nsx_hh.h"
#include "nsx_internal.h"
#include "nsx_kernels.h"

namespace nsx {

template <int MODEL, int DRIVE, int PL>
__global__ __launch_bounds__(kBlock) void k_update(StepParams p, DevPtrs s, int32_t off) {
  constexpr bool TWO = MODEL != NSX_MODEL_IZHIKEVICH;
  __shared__ uint32_t s_drive[kBlock];
  const int32_t t = *s.tbase + off;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const NeuronId id = neuron_id(p, s, i, blockIdx.x);
  const bool live = i < p.n && id.local < id.n;  // batch padding neurons stay idle
  // Every load of the neuron's update is issued here, ahead of the block's drive staging and
  // barrier, so that they are in flight together (one memory round trip, not one per field).
  constexpr bool PAIR = PL == NSX_PLAST_PAIR;
  RingIn rg{nullptr, 0, 0};
  float v = 0.0f, w1 = 0.0f, w2 = 0.0f, w3 = 0.0f;  // LIF: v ie ii; Izh: v u; HH: v m h n
  int32_t refr = 0, last = kNeverStep;
  HHSyn y{0.0f, 0.0f, 0.0f, 0.0f};
  // blocks with no spike at t-1 (their segment count is 0) skip the last-spike loads
  const uint32_t prev = PAIR ? s.bcnt[blockIdx.x] : 0u;
  if (live) {
    rg = ring_load<TWO>(p, s, i, t);
    v = s.v[i];
    if (PAIR && prev) last = s.last[i];
    if (MODEL == NSX_MODEL_LIF) {
      w1 = s.s0[i];
      w2 = s.s1[i];
      refr = s.refr[i];
    } else if (MODEL == NSX_MODEL_IZHIKEVICH) {
      w1 = s.u[i];
    } else {
      w1 = s.gm[i];
      w2 = s.gh[i];
      w3 = s.gn[i];
      y = hh_syn_load<DRIVE>(s, i);
    }
  }
  drive_stage<DRIVE>(p, s, t, s_drive);
  if (DRIVE != NSX_DRIVE_NONE) __syncthreads();  // the staged words
  bool spiked = false;
  float lfp = 0.0f;  // HH: this neuron's |synaptic currents|
  if (live) {
    post_bump_prev<PAIR>(p, s, i, last, t);
    float a0, a1;
    ring_take<TWO>(p, rg, a0, a1);
    float z = 0.0f;
    drive_draw<DRIVE>(p, s_drive, a0, z);
    if (MODEL == NSX_MODEL_LIF) {
      // values whose bits do not change are not rewritten (silent neurons: decayed currents)
      float ie = w1 * p.dec_e;
      ie = ie + a0;
      float ii = w2 * p.dec_i;
      ii = ii + a1;
      if (__float_as_uint(ie) != __float_as_uint(w1)) s.s0[i] = ie;
      if (__float_as_uint(ii) != __float_as_uint(w2)) s.s1[i] = ii;
      if (refr > 0) {
        s.refr[i] = refr - 1;
      } else {
        float I = ie + ii;
        float x = v - p.v_rest;
        x = -x;
        float yv = p.lif_r * I;
        x = x + yv;
        x = x * p.lif_k;
        float vn = v + x;
        if (vn >= p.v_th) {
          vn = p.v_reset;
          s.refr[i] = p.refr_steps;
          spiked = true;
        }
        if (__float_as_uint(vn) != __float_as_uint(v)) s.v[i] = vn;
      }
    } else if (MODEL == NSX_MODEL_IZHIKEVICH) {
      float I = a0;
      if (DRIVE == NSX_DRIVE_GAUSSIAN) {
        float q = p.drive_sigma * z;
        q = p.drive_mean + q;
        I = I + q;
      }
      float u = w1;
      if (v >= 30.0f) {
        v = s.c[i];
        u = u + s.d[i];
        spiked = true;
      } else {
        const int ty = id.local < id.n_exc ? 0 : 1;
        const float a = p.izh_a[ty], b = p.izh_b[ty];
        const float v0 = v;
        for (uint32_t k = 0; k < p.substeps; ++k) {
          float t1 = 0.04f * v;
          t1 = t1 * v;
          float t2 = 5.0f * v;
          float acc = t1 + t2;
          acc = acc + 140.0f;
          acc = acc - u;
          acc = acc + I;
          acc = acc * p.izh_h;
          v = v + acc;
        }
        float t3 = b * v0;
        t3 = t3 - u;
        t3 = a * t3;
        t3 = t3 * p.dtf;
        u = u + t3;
      }
      s.v[i] = v;
      s.u[i] = u;
    } else {  // Hodgkin-Huxley with AMPA / GABA-A / NMDA conductances
      const float I = hh_input<DRIVE>(p, y, v, a0, a1, z, lfp);  // held over the sub-steps
      hh_syn_store<DRIVE>(s, i, y);
      float m = w1, h = w2, nn = w3;
      float vn = v;
      for (uint32_t k = 0; k < p.hh_substeps; ++k) hh_substep(vn, m, h, nn, I, p.hh);
      if (p.hh_reset_mode) {
        spiked = vn >= -55.0f;
        if (spiked) vn = -70.0f;
      } else {
        spiked = (v < p.hh_spike_v) && (vn >= p.hh_spike_v);
      }
      s.v[i] = vn;
      s.gm[i] = m;
      s.gh[i] = h;
      s.gn[i] = nn;
    }
    if (spiked) spike_record<PL>(p, s, i, t);
  }
  if (MODEL == NSX_MODEL_HH && p.stats_on) lfp_accumulate(s, lfp, blockIdx.x);
  block_compact(p, s, spiked, i, t, blockIdx.x);
}

// Order LDS traffic of one wavefront (all lanes write, then any lane reads).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Flattened view of up to 64 rows: lane r holds row r (start, source); incl[r] is the
// inclusive prefix of row lengths.  Lanes then walk the concatenated slots 64 at a time.
struct RowChunk {
  uint32_t incl[64];
  uint32_t src[64];
  uint64_t beg[64];
};

// Inclusive wave scan of deg; stores prefix / start / source of each lane's row in LDS and
// returns the total.
__device__ inline uint32_t chunk_setup(RowChunk& rc, uint32_t lane, uint32_t deg, uint64_t beg,
                                       uint32_t src) {
  uint32_t incl = deg;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  rc.incl[lane] = incl;
  rc.src[lane] = src;
  rc.beg[lane] = beg;
  wave_sync();
  return __builtin_amdgcn_readfirstlane(__shfl(incl, 63));
}

// Slot position of flat index f (rows visited in increasing order by each lane).
__device__ inline uint64_t chunk_slot(const RowChunk& rc, uint32_t f, uint32_t& r) {
  while (rc.incl[r] <= f) ++r;
  const uint32_t excl = r ? rc.incl[r - 1] : 0u;
  return rc.beg[r] + (f - excl);
}

// Pair STDP on an E->E synapse whose source fired at step t (a: the source's record, bumped at
// t; pr: the target's).  pair_pending: the target's pending depressions from post spikes after
// the source's previous spike (their pre trace is the one between the source's last two
// spikes), in time order -- the weight the spike delivers.  pair_at_pre: then potentiation by
// the post trace, then depression if the target fired at t as well.
__device__ inline float pair_pending(const StepParams& p, const DevPtrs& s, float w,
                                     const PostRec& pr, const TraceRec& a, int32_t t) {
#pragma unroll
  for (uint32_t q = 0; q < kPostPend; ++q) {
    const int32_t sp = pr.pend[q];
    if (q < pr.npend && sp > a.s_prev && sp < t) {
      const float x = trace_now(a.v_prev, a.s_prev, s.ptab, p.ptab_n, sp);
      const float dlt = p.a_minus * x;
      w = w - dlt;
      w = clampf_(w, p.w_min, p.w_max);
    }
  }
  return w;
}

__device__ inline float pair_at_pre(const StepParams& p, const DevPtrs& s, float w,
                                    const PostRec& pr, const TraceRec& a, int32_t t) {
  // the source fired at t, so its pre trace is the record bumped at t (factor P[0] = 1)
  const float x = trace_now(pr.v, pr.s, s.mtab, p.mtab_n, t);
  const float dlt = p.a_plus * x;
  w = w + dlt;
  w = clampf_(w, p.w_min, p.w_max);
  if (pr.last == t) {  // target fired at t as well: depression follows
    const float y = p.a_minus * a.v;
    w = w - y;
    w = clampf_(w, p.w_min, p.w_max);
  }
  return w;
}

// Deferred pair depressions of post neuron j's pending spikes pend[0, np) on one in-synapse whose
// source has record a: a spike at step sp applies when the source has not fired since (a.s <
// sp; otherwise the source's propagation applied it, or owned it as a same-step pair).
__device__ inline float pair_flush(const StepParams& p, const DevPtrs& s, float w,
                                   const int32_t* pend, uint32_t np, const TraceRec& a,
                                   bool& changed) {
#pragma unroll
  for (uint32_t q = 0; q < kPostPend; ++q) {
    const int32_t sp = pend[q];
    if (q < np && a.s < sp) {
      const float x = trace_now(a.v, a.s, s.ptab, p.ptab_n, sp);
      const float dlt = p.a_minus * x;
      w = w - dlt;
      w = clampf_(w, p.w_min, p.w_max);
      changed = true;
    }
  }
  return w;
}

// Propagation work item: part q of neuron block b.  PL: STDP applied on the spiking rows' E->E
// slots -- pair: the target's pending depressions, potentiation by the post trace (then
// depression if the target fired at t); triplet: depression by the target's last spike.
template <int PL>
__device__ void propagate_item(const StepParams& p, const DevPtrs& s, RowChunk& rc, int32_t t,
                               uint32_t b, uint32_t q, uint32_t cnt, uint32_t lane) {
  const uint32_t P = p.prop_split;
  if (cnt <= q) return;
  const uint32_t mine = (cnt - q + P - 1) / P;  // rows q, q+P, ... of the segment
  const uint32_t tm = (uint32_t)t % p.D;
  const size_t plane = (size_t)p.ch * p.n;
  unsigned long long ev = 0, ltp = 0;
  for (uint32_t c0 = 0; c0 < mine; c0 += 64) {
    const uint32_t nr = min(64u, mine - c0);
    uint32_t i = 0, deg = 0;
    uint64_t beg = 0;
    if (lane < nr) {
      i = s.spikes[b * kBlock + q + (c0 + lane) * P];
      beg = s.row_ptr[i];
      deg = (uint32_t)(s.row_ptr[i + 1] - beg);
    }
    const uint32_t total = chunk_setup(rc, lane, deg, beg, i);
    ev += total;
    uint32_t r = 0;
    for (uint32_t f = lane; f < total; f += 128) {
      // two slots per lane: both sets of loads are issued before the atomics
      const bool two = f + 64 < total;
      const uint64_t k0 = chunk_slot(rc, f, r);
      const uint32_t src0 = rc.src[r];
      uint32_t r1 = r;
      const uint64_t k1 = two ? chunk_slot(rc, f + 64, r1) : k0;
      const uint32_t src1 = rc.src[r1];
      const uint32_t j0 = s.col[k0], j1 = s.col[k1];
      float w0 = s.w[k0], w1 = s.w[k1];
      const uint32_t d0 = s.delay[k0], d1 = s.delay[k1];
      const bool pot0 = PL != NSX_PLAST_NONE && src0 < p.n_exc && j0 < p.n_exc;
      const bool pot1 = PL != NSX_PLAST_NONE && two && src1 < p.n_exc && j1 < p.n_exc;
      PostRec pr0{0.0f, kNeverStep, kNeverStep, 0u, {}}, pr1{0.0f, kNeverStep, kNeverStep, 0u, {}};
      if (pot0) pr0 = s.post[j0];  // 32 bytes: trace, last spike, pending spikes
      if (pot1) pr1 = s.post[j1];
      TraceRec a0{}, a1{};
      if (PL == NSX_PLAST_PAIR) {  // the target's pending depressions precede the delivery
        if (pot0) {
          a0 = s.pre[src0];
          w0 = pair_pending(p, s, w0, pr0, a0, t);
        }
        if (pot1) {
          a1 = s.pre[src1];
          w1 = pair_pending(p, s, w1, pr1, a1, t);
        }
      }
      uint32_t sl0 = tm + d0, sl1 = tm + d1;
      sl0 = sl0 >= p.D ? sl0 - p.D : sl0;
      sl1 = sl1 >= p.D ? sl1 - p.D : sl1;
      const uint32_t ch0 = (p.ch > 1 && src0 >= p.n_exc) ? 1u : 0u;
      const uint32_t ch1 = (p.ch > 1 && src1 >= p.n_exc) ? 1u : 0u;
      atomicAdd((unsigned long long*)(s.ring + sl0 * plane + (size_t)ch0 * p.n + j0),
                (unsigned long long)quantize(w0, p.scale_in));
      if (two)
        atomicAdd((unsigned long long*)(s.ring + sl1 * plane + (size_t)ch1 * p.n + j1),
                  (unsigned long long)quantize(w1, p.scale_in));
      if (PL == NSX_PLAST_TRIPLET && (pot0 || pot1)) {
        // pre spike at t: depression by the target's last spike (same step included)
        if (pot0) {
          const float x = trace_at(s.mtab, p.mtab_n, t - pr0.last);
          const float dlt = p.a_minus * x;
          const float wn = w0 - dlt;
          s.w[k0] = clampf_(wn, p.w_min, p.w_max);
          ltp += 1;
        }
        if (pot1) {
          const float x = trace_at(s.mtab, p.mtab_n, t - pr1.last);
          const float dlt = p.a_minus * x;
          const float wn = w1 - dlt;
          s.w[k1] = clampf_(wn, p.w_min, p.w_max);
          ltp += 1;
        }
      }
      if (PL == NSX_PLAST_PAIR && (pot0 || pot1)) {
        if (pot0) {
          s.w[k0] = pair_at_pre(p, s, w0, pr0, a0, t);
          ltp += 1;
        }
        if (pot1) {
          s.w[k1] = pair_at_pre(p, s, w1, pr1, a1, t);
          ltp += 1;
        }
      }
      r = r1;
    }
    wave_sync();
  }
  if (PL != NSX_PLAST_NONE) {
#pragma unroll
    for (uint32_t o = 32; o > 0; o >>= 1) ltp += __shfl_xor(ltp, o);
  }
  if (lane == 0) {
    // parts of one block accumulate into that block's counter
    if (P == 1) {
      s.part[kPartEvents * p.nblk + b] += ev;
      if (PL != NSX_PLAST_NONE) s.part[kPartLtp * p.nblk + b] += ltp;
    } else {
      atomicAdd(&s.part[kPartEvents * p.nblk + b], ev);
      if (PL != NSX_PLAST_NONE) atomicAdd(&s.part[kPartLtp * p.nblk + b], ltp);
    }
  }
}

// Post-spike work item: part q of neuron block b (excitatory neurons only), over the E->E
// in-slots of its spiking neurons -- pair: only neurons whose list of pending depressions
// filled at t, all of them applied (pair_flush); triplet: potentiation by the source's last
// spike and the target's previous spike, for sources that did not fire at t.
template <int PL>
__device__ void depress_item(const StepParams& p, const DevPtrs& s, RowChunk& rc, int32_t t,
                             uint32_t b, uint32_t q, uint32_t cnt, uint32_t lane) {
  const uint32_t P = p.prop_split;
  if (cnt <= q) return;
  const uint32_t mine = (cnt - q + P - 1) / P;
  unsigned long long slots = 0;
  for (uint32_t c0 = 0; c0 < mine; c0 += 64) {
    const uint32_t nr = min(64u, mine - c0);
    uint32_t j = 0, deg = 0;
    uint64_t beg = 0;
    if (lane < nr) {
      j = s.spikes[b * kBlock + q + (c0 + lane) * P];
      if (j < p.n_exc && (PL != NSX_PLAST_PAIR || s.post[j].npend == kPostPend)) {
        beg = s.csc_ptr[j];
        deg = (uint32_t)(s.csc_ptr[j + 1] - beg);
      }
    }
    const uint32_t total = chunk_setup(rc, lane, deg, beg, j);
    slots += total;
    uint32_t r = 0;
    for (uint32_t f = lane; f < total; f += 128) {
      const bool two = f + 64 < total;
      const uint64_t e0 = chunk_slot(rc, f, r);
      uint32_t r1 = r;
      const uint64_t e1 = two ? chunk_slot(rc, f + 64, r1) : e0;
      const uint32_t i0 = s.csc_src[e0], i1 = s.csc_src[e1];
      const uint64_t k0 = s.csc_slot[e0], k1 = s.csc_slot[e1];
      if (PL == NSX_PLAST_TRIPLET) {
        const int32_t l0 = s.last[i0], l1 = s.last[i1];
        // the source's propagation item owns same-step pairs; a source that never fired
        // pairs with nothing
        const bool own0 = l0 != t && l0 != kNeverStep;
        const bool own1 = two && l1 != t && l1 != kNeverStep;
        float w0 = 0.0f, w1 = 0.0f;
        if (own0) w0 = s.w[k0];
        if (own1) w1 = s.w[k1];
        if (own0) {
          const int32_t pv = s.post[rc.src[r]].s;
          const float x1 = trace_at(s.ptab, p.ptab_n, t - l0);
          const float x2 = trace_at(s.ytab, p.ytab_n, t - pv);
          float dw = p.a_plus * x1;
          const float y = p.a3 * x2;
          dw = dw + y;
          const float wn = w0 + dw;
          s.w[k0] = clampf_(wn, p.w_min, p.w_max);
        }
        if (own1) {
          const int32_t pv = s.post[rc.src[r1]].s;
          const float x1 = trace_at(s.ptab, p.ptab_n, t - l1);
          const float x2 = trace_at(s.ytab, p.ytab_n, t - pv);
          float dw = p.a_plus * x1;
          const float y = p.a3 * x2;
          dw = dw + y;
          const float wn = w1 + dw;
          s.w[k1] = clampf_(wn, p.w_min, p.w_max);
        }
        r = r1;
        continue;
      }
      // pair: the row's post neuron has a full list of pending depressions
      const PostRec& pj0 = s.post[rc.src[r]];
      const PostRec& pj1 = s.post[rc.src[r1]];
      const TraceRec a0 = s.pre[i0], a1 = s.pre[i1];
      // a source that fired at or after every pending spike has none left on this slot
      const bool own0 = a0.s < pj0.pend[kPostPend - 1];
      const bool own1 = two && a1.s < pj1.pend[kPostPend - 1];
      float w0 = 0.0f, w1 = 0.0f;
      if (own0) w0 = s.w[k0];
      if (own1) w1 = s.w[k1];
      if (own0) {
        bool ch = false;
        const float wn = pair_flush(p, s, w0, pj0.pend, kPostPend, a0, ch);
        if (ch) s.w[k0] = wn;
      }
      if (own1) {
        bool ch = false;
        const float wn = pair_flush(p, s, w1, pj1.pend, kPostPend, a1, ch);
        if (ch) s.w[k1] = wn;
      }
      r = r1;
    }
    wave_sync();
  }
  if (lane == 0) {
    if (P == 1)
      s.part[kPartLtd * p.nblk + b] += slots;
    else
      atomicAdd(&s.part[kPartLtd * p.nblk + b], slots);
  }
}

// Work items: propagation (nblk * P) and post-spike STDP (nblk_e * P), interleaved so that both
// kinds are resident on every CU.  The grid is capped (and fixed, for graph capture): each
// wavefront owns items wave, wave + W, wave + 2W, ... (W wavefronts in the grid), looks up the
// block spike counts of 64 of them at once (a lane each), and works only on those with spikes,
// so the empty blocks of a sparse step cost one parallel load per 64 items.
__device__ inline void decode_item(uint32_t wk, uint32_t both, uint32_t nprop, uint32_t nltd,
                                   uint32_t P, bool& prop, uint32_t& b, uint32_t& q) {
  uint32_t item;
  if (wk < both) {
    prop = (wk & 1u) == 0u;
    item = wk >> 1;
  } else {
    prop = nprop > nltd;
    item = wk - both / 2;
  }
  b = item / P;
  q = item - b * P;
}

template <int PL>
__global__ __launch_bounds__(kBlock) void k_synapses(StepParams p, DevPtrs s, int32_t off) {
  __shared__ RowChunk sh[kBlock / 64];
  const int32_t t = *s.tbase + off;
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * kBlock + threadIdx.x) >> 6);
  const uint32_t P = p.prop_split;
  const uint32_t nprop = p.nblk * P;
  const uint32_t nltd = PL != NSX_PLAST_NONE ? ((p.n_exc + kBlock - 1) / kBlock) * P : 0u;
  const uint32_t both = 2 * min(nprop, nltd);
  const uint32_t nwork = nprop + nltd;
  const uint32_t nwaves = (gridDim.x * kBlock) >> 6;
  for (uint64_t base = wave; base < nwork; base += 64ull * nwaves) {
    const uint64_t mine = base + (uint64_t)lane * nwaves;
    uint32_t c = 0, qq = 0;
    if (mine < nwork) {
      bool pr;
      uint32_t bb;
      decode_item((uint32_t)mine, both, nprop, nltd, P, pr, bb, qq);
      c = s.bcnt[bb];
    }
    unsigned long long act = __ballot(mine < nwork && c > qq);
    while (act) {
      const uint32_t l = (uint32_t)(__ffsll((long long)act) - 1);
      act &= act - 1ull;
      bool prop;
      uint32_t b, q;
      decode_item((uint32_t)(base + (uint64_t)l * nwaves), both, nprop, nltd, P, prop, b, q);
      const uint32_t cnt = __builtin_amdgcn_readfirstlane(__shfl(c, (int)l));
      if (prop)
        propagate_item<PL>(p, s, sh[wid], t, b, q, cnt, lane);
      else if (PL != NSX_PLAST_NONE)
        depress_item<PL>(p, s, sh[wid], t, b, q, cnt, lane);
    }
  }
}

__global__ void k_advance(int32_t* tbase, int32_t g) { *tbase += g; }

// ---- launch helpers ------------------------------------------------------------------------
typedef void (*UpdateFn)(StepParams, DevPtrs, int32_t);

template <int MODEL, int DRIVE>
static UpdateFn pick_plast(int pl) {
  switch (pl) {
    case NSX_PLAST_PAIR: return k_update<MODEL, DRIVE, NSX_PLAST_PAIR>;
    case NSX_PLAST_TRIPLET: return k_update<MODEL, DRIVE, NSX_PLAST_TRIPLET>;
    default: return k_update<MODEL, DRIVE, NSX_PLAST_NONE>;
  }
}
template <int MODEL>
static UpdateFn pick_drive(int drive, int pl) {
  switch (drive) {
    case NSX_DRIVE_POISSON: return pick_plast<MODEL, NSX_DRIVE_POISSON>(pl);
    case NSX_DRIVE_GAUSSIAN: return pick_plast<MODEL, NSX_DRIVE_GAUSSIAN>(pl);
    case NSX_DRIVE_OU: return pick_plast<MODEL, NSX_DRIVE_OU>(pl);
    default: return pick_plast<MODEL, NSX_DRIVE_NONE>(pl);
  }
}
UpdateFn pick_update(int model, int drive, int pl) {
  switch (model) {
    case NSX_MODEL_LIF: return pick_drive<NSX_MODEL_LIF>(drive, pl);
    case NSX_MODEL_IZHIKEVICH: return pick_drive<NSX_MODEL_IZHIKEVICH>(drive, pl);
    default: return pick_drive<NSX_MODEL_HH>(drive, pl);
  }
}

// One wavefront per work item when every block spikes, capped: sparse steps leave most waves
// idle and the cap bounds their cost.
uint32_t syn_grid(const nsx_ctx* ctx) {
  const bool plast = ctx->cfg.plasticity != NSX_PLAST_NONE;
  const uint64_t waves = (uint64_t)ctx->p.nblk * (plast ? 2 : 1) * ctx->p.prop_split;
  const uint64_t blocks = (waves + kBlock / 64 - 1) / (kBlock / 64);
  return (uint32_t)std::min<uint64_t>(blocks, 16384);
}

// Enqueue the launches of one step with offset `off` from *tbase.
hipError_t enqueue_step(nsx_ctx* ctx, int32_t off, hipEvent_t* evs) {
  const StepParams& p = ctx->p;
  const int pl = (int)ctx->cfg.plasticity;
  UpdateFn upd = pick_update((int)ctx->cfg.model, (int)ctx->cfg.drive, pl);
  const uint32_t gu = (uint32_t)((ctx->n + kBlock - 1) / kBlock);
  const uint32_t gs = syn_grid(ctx);
  if (evs) (void)hipEventRecord(evs[0], ctx->stream);
  if (ctx->cfg.model == NSX_MODEL_HH_SURROGATE) {
    const hipError_t e = launch_update_surrogate(ctx, off);
    if (e != hipSuccess) return e;
  } else {
    hipLaunchKernelGGL(upd, dim3(gu), dim3(kBlock), 0, ctx->stream, p, ctx->d, off);
  }
  if (evs) (void)hipEventRecord(evs[1], ctx->stream);
  if (pl == NSX_PLAST_PAIR)
    hipLaunchKernelGGL(k_synapses<NSX_PLAST_PAIR>, dim3(gs), dim3(kBlock), 0, ctx->stream, p,
                       ctx->d, off);
  else if (pl == NSX_PLAST_TRIPLET)
    hipLaunchKernelGGL(k_synapses<NSX_PLAST_TRIPLET>, dim3(gs), dim3(kBlock), 0, ctx->stream, p,
                       ctx->d, off);
  else
    hipLaunchKernelGGL(k_synapses<NSX_PLAST_NONE>, dim3(gs), dim3(kBlock), 0, ctx->stream, p,
                       ctx->d, off);
  if (evs) (void)hipEventRecord(evs[2], ctx->stream);
  return hipGetLastError();
}

hipError_t enqueue_advance(nsx_ctx* ctx, int32_t g) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, ctx->stream, ctx->d.tbase, g);
  return hipGetLastError();
}

// Every excitatory neuron's pending depressions applied to its E->E in-synapses (apply) or
// dropped, and its list emptied.  A full list was applied by the synapse pass of the step it
// filled in and is only emptied.
__global__ void k_pend_flush(StepParams p, DevPtrs s, int apply) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < p.n_exc;
       j += gridDim.x * blockDim.x) {
    PostRec& r = s.post[j];
    const uint32_t np = r.npend;
    if (np == 0) continue;
    if (apply && np < kPostPend) {
      int32_t pend[kPostPend];
#pragma unroll
      for (uint32_t q = 0; q < kPostPend; ++q) pend[q] = r.pend[q];
      for (uint64_t e = s.csc_ptr[j]; e < s.csc_ptr[j + 1]; ++e) {
        const TraceRec a = s.pre[s.csc_src[e]];
        const uint64_t k = s.csc_slot[e];
        bool ch = false;
        const float w = pair_flush(p, s, s.w[k], pend, np, a, ch);
        if (ch) s.w[k] = w;
      }
    }
    r.npend = 0;
  }
}

nsx_status flush_pending(nsx_ctx* ctx, bool apply) {
  if (ctx->cfg.plasticity != NSX_PLAST_PAIR || !ctx->csc_ptr || ctx->p.n_exc == 0)
    return NSX_OK;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((ctx->p.n_exc + 255) / 256, 1u << 14);
  hipLaunchKernelGGL(k_pend_flush, dim3(blocks), dim3(256), 0, ctx->stream, ctx->p, ctx->d,
                     apply ? 1 : 0);
  NSX_HIP(ctx, hipGetLastError());
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NSX_OK;
}

}  // namespace nsx
