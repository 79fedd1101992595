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

// One step of neuron block blk (256 neurons, one per thread) at step t: the whole update of
// k_update, shared with the persistent kernel (k_persist), which runs it for the blocks in turn.
template <int MODEL, int DRIVE, int PL>
__device__ inline void update_block(const StepParams& p, const DevPtrs& s, int32_t t,
                                    uint32_t blk) {
  constexpr bool TWO = MODEL != NSX_MODEL_IZHIKEVICH;
  __shared__ uint32_t s_drive[kBlock];
  const uint32_t i = blk * kBlock + threadIdx.x;
  const NeuronId id = neuron_id(p, s, i, blk);
  const bool live = i < p.n && id.local < id.n;  // batch padding neurons stay idle
  // Every load of the neuron's update is issued here, ahead of the block's drive staging and
  // barrier, so that they are in flight together (one memory round trip, not one per field).
  constexpr bool PAIR = PL == NSX_PLAST_PAIR;
  RingIn rg{nullptr, 0, 0};
  float v = 0.0f, w1 = 0.0f, w2 = 0.0f, w3 = 0.0f;  // LIF: v ie ii; Izh: v u; HH: v m h n
  int32_t refr = 0, last = kNeverStep;
  HHSyn y{0.0f, 0.0f, 0.0f, 0.0f};
  // blocks with no spike at t-1 (their segment count is 0) skip the last-spike loads
  const uint32_t prev = PAIR ? s.bcnt[blk] : 0u;
  const bool ring_in = ring_active(p, s, t);
  if (live) {
    rg = ring_load<TWO>(p, s, i, t, ring_in);
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
  drive_stage<DRIVE>(p, s, t, s_drive, id);
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
  if (MODEL == NSX_MODEL_HH && p.stats_on) lfp_accumulate(s, lfp, blk);
  block_compact(p, s, spiked, i, t, blk);
}

template <int MODEL, int DRIVE, int PL>
__global__ __launch_bounds__(kBlock) void k_update(StepParams p, DevPtrs s, int32_t off) {
  update_block<MODEL, DRIVE, PL>(p, s, *s.tbase + off, blockIdx.x);
}

// ---- LIF / Izhikevich update, four neurons per thread -------------------------------------
// The same step as k_update for the two point models, arranged so that no work waits on another
// wavefront: thread l of a wavefront updates neurons 4l .. 4l+3 of one 256-neuron segment, so
// its one Philox draw is exactly its neurons' drive words (no LDS staging, no barrier), its
// state arrives in 16-byte loads, and the segment's spike list is compacted by the wavefront
// alone (a prefix sum of per-lane counts).  Spike segments, drive words, arithmetic and
// records are those of k_update, neuron for neuron.

// Four consecutive floats / ints at index i0 (a multiple of 4): one 16-byte load when all four
// lie inside the array (n elements), else element loads of those that do.
template <typename T>
__device__ inline void load4(const T* a, uint32_t i0, uint32_t n, T o[4]) {
  static_assert(sizeof(T) == 4, "4-byte elements");
  if (i0 + 4 <= n) {
    const uint4 q = *reinterpret_cast<const uint4*>(a + i0);
    o[0] = __builtin_bit_cast(T, q.x);
    o[1] = __builtin_bit_cast(T, q.y);
    o[2] = __builtin_bit_cast(T, q.z);
    o[3] = __builtin_bit_cast(T, q.w);
  } else {
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) o[k] = i0 + k < n ? a[i0 + k] : T(0);
  }
}

// Four consecutive ring entries (int64) of one channel plane: two 16-byte loads when aligned.
__device__ inline void ring4(const long long* r, uint32_t i0, uint32_t n, bool aligned,
                             long long o[4]) {
  if (aligned && i0 + 4 <= n) {
    const longlong2 a = *reinterpret_cast<const longlong2*>(r + i0);
    const longlong2 b = *reinterpret_cast<const longlong2*>(r + i0 + 2);
    o[0] = a.x;
    o[1] = a.y;
    o[2] = b.x;
    o[3] = b.y;
  } else {
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) o[k] = i0 + k < n ? r[i0 + k] : 0ll;
  }
}

// Write back four consecutive values when any of them changed (bit mask `chg`): one 16-byte
// store when all four lie inside the array, else element stores of the changed ones.  Silent
// neurons cost no write; active ones a coalesced one.
template <typename T>
__device__ inline void store4(T* a, uint32_t i0, uint32_t n, const T v[4], uint32_t chg) {
  static_assert(sizeof(T) == 4, "4-byte elements");
  if (!chg) return;
  if (i0 + 4 <= n) {
    *reinterpret_cast<uint4*>(a + i0) =
        make_uint4(__builtin_bit_cast(uint32_t, v[0]), __builtin_bit_cast(uint32_t, v[1]),
                   __builtin_bit_cast(uint32_t, v[2]), __builtin_bit_cast(uint32_t, v[3]));
  } else {
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
      if (chg & (1u << k)) a[i0 + k] = v[k];
  }
}

// Clear four ring entries of one plane after they were read (nz: bit k set when entry k held
// input).
__device__ inline void ring_clear4(long long* r, uint32_t i0, uint32_t n, bool aligned,
                                   uint32_t nz) {
  if (!nz) return;
  if (aligned && i0 + 4 <= n) {
    *reinterpret_cast<longlong2*>(r + i0) = make_longlong2(0, 0);
    *reinterpret_cast<longlong2*>(r + i0 + 2) = make_longlong2(0, 0);
  } else {
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
      if (nz & (1u << k)) r[i0 + k] = 0;
  }
}

template <int MODEL, int DRIVE, int PL>
__global__ __launch_bounds__(kBlock) void k_update4(StepParams p, DevPtrs s, int32_t off) {
  static_assert(MODEL == NSX_MODEL_LIF || MODEL == NSX_MODEL_IZHIKEVICH, "point models");
  constexpr bool TWO = MODEL == NSX_MODEL_LIF;
  constexpr bool PAIR = PL == NSX_PLAST_PAIR;
  const int32_t t = *s.tbase + off;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t b = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);  // this wave's segment
  if (b >= p.nblk) return;  // a whole wavefront; the kernel has no block-wide barrier
  const uint32_t i0 = b * kBlock + 4u * lane;
  const NeuronId id = neuron_id(p, s, i0, b);
  bool live[4];
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) live[k] = i0 + k < p.n && id.local + k < id.n;
  const uint32_t prev = PAIR ? s.bcnt[b] : 0u;  // the segment's spikes at t-1
  const bool ring_in = ring_active(p, s, t);
  // every load first
  long long q0[4] = {0, 0, 0, 0}, q1[4] = {0, 0, 0, 0};
  long long* r0 = s.ring + (size_t)((uint32_t)t % p.D) * p.ch * p.n;
  const bool ring_al = (p.n & 1u) == 0u;  // 16-byte aligned pairs of int64
  if (ring_in) {
    ring4(r0, i0, p.n, ring_al, q0);
    if (TWO) ring4(r0 + p.n, i0, p.n, ring_al, q1);
  }
  float v[4], w1[4], w2[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  int32_t refr[4] = {0, 0, 0, 0}, last[4] = {kNeverStep, kNeverStep, kNeverStep, kNeverStep};
  load4(s.v, i0, p.n, v);
  if (MODEL == NSX_MODEL_LIF) {
    load4(s.s0, i0, p.n, w1);
    load4(s.s1, i0, p.n, w2);
    load4(s.refr, i0, p.n, refr);
  } else {
    load4(s.u, i0, p.n, w1);
  }
  if (PAIR && prev) load4(s.last, i0, p.n, last);
  U4 dr{0u, 0u, 0u, 0u};
  if (DRIVE != NSX_DRIVE_NONE)
    dr = draw(id.seed, (uint64_t)(id.local >> 2), (uint32_t)t, kStreamDrive);
  const uint32_t words[4] = {dr.x, dr.y, dr.z, dr.w};
  uint32_t spk = 0;  // bit k: neuron i0 + k spiked
  uint32_t nz0 = 0, nz1 = 0;  // ring entries that held input (cleared below)
  // LIF: the new currents, counter and V, and which of them changed
  float ie4[4] = {w1[0], w1[1], w1[2], w1[3]}, ii4[4] = {w2[0], w2[1], w2[2], w2[3]};
  float vn4[4] = {v[0], v[1], v[2], v[3]};
  int32_t rf4[4] = {refr[0], refr[1], refr[2], refr[3]};
  uint32_t c_ie = 0, c_ii = 0, c_v = 0, c_rf = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    if (!live[k]) continue;
    const uint32_t i = i0 + k;
    post_bump_prev<PAIR>(p, s, i, last[k], t);
    float a0 = dequantize(q0[k], p.scale_out), a1 = 0.0f;
    if (q0[k] != 0) nz0 |= 1u << k;
    if (TWO) {
      a1 = dequantize(q1[k], p.scale_out);
      if (q1[k] != 0) nz1 |= 1u << k;
    }
    float z = 0.0f;
    if (DRIVE == NSX_DRIVE_POISSON) {
      const uint32_t x = words[k];
      uint32_t c = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) c += (q < p.npthr && x >= p.pthr[q]) ? 1u : 0u;
      if (p.npthr > 4 && x >= p.pthr[3])
        for (int q = 4; q < p.npthr; ++q) c += (x >= p.pthr[q]) ? 1u : 0u;
      float f = (float)c * p.drive_weight;
      a0 = a0 + f;
    } else if (DRIVE == NSX_DRIVE_GAUSSIAN) {
      const uint32_t angle = words[(k & 2u) + 1] + ((k & 1u) ? 0xC0000000u : 0u);
      z = normal_(words[k & 2u], angle);
    }
    bool spiked = false;
    if (MODEL == NSX_MODEL_LIF) {
      float ie = w1[k] * p.dec_e;
      ie = ie + a0;
      float ii = w2[k] * p.dec_i;
      ii = ii + a1;
      ie4[k] = ie;
      ii4[k] = ii;
      if (__float_as_uint(ie) != __float_as_uint(w1[k])) c_ie |= 1u << k;
      if (__float_as_uint(ii) != __float_as_uint(w2[k])) c_ii |= 1u << k;
      if (refr[k] > 0) {
        rf4[k] = refr[k] - 1;
        c_rf |= 1u << k;
      } else {
        float I = ie + ii;
        float x = v[k] - p.v_rest;
        x = -x;
        float yv = p.lif_r * I;
        x = x + yv;
        x = x * p.lif_k;
        float vn = v[k] + x;
        if (vn >= p.v_th) {
          vn = p.v_reset;
          rf4[k] = p.refr_steps;
          c_rf |= 1u << k;
          spiked = true;
        }
        vn4[k] = vn;
        if (__float_as_uint(vn) != __float_as_uint(v[k])) c_v |= 1u << k;
      }
    } else {
      float I = a0;
      if (DRIVE == NSX_DRIVE_GAUSSIAN) {
        float q = p.drive_sigma * z;
        q = p.drive_mean + q;
        I = I + q;
      }
      float vv = v[k], u = w1[k];
      if (vv >= 30.0f) {
        vv = s.c[i];
        u = u + s.d[i];
        spiked = true;
      } else {
        const int ty = id.local + k < id.n_exc ? 0 : 1;
        const float a = p.izh_a[ty], bb = p.izh_b[ty];
        const float v0 = vv;
        for (uint32_t m = 0; m < p.substeps; ++m) {
          float t1 = 0.04f * vv;
          t1 = t1 * vv;
          float t2 = 5.0f * vv;
          float acc = t1 + t2;
          acc = acc + 140.0f;
          acc = acc - u;
          acc = acc + I;
          acc = acc * p.izh_h;
          vv = vv + acc;
        }
        float t3 = bb * v0;
        t3 = t3 - u;
        t3 = a * t3;
        t3 = t3 * p.dtf;
        u = u + t3;
      }
      v[k] = vv;
      w1[k] = u;
    }
    if (spiked) {
      spike_record<PL>(p, s, i, t);
      spk |= 1u << k;
    }
  }
  ring_clear4(r0, i0, p.n, ring_al, nz0);
  if (TWO) ring_clear4(r0 + p.n, i0, p.n, ring_al, nz1);
  if (MODEL == NSX_MODEL_LIF) {
    store4(s.s0, i0, p.n, ie4, c_ie);
    store4(s.s1, i0, p.n, ii4, c_ii);
    store4(s.refr, i0, p.n, rf4, c_rf);
    store4(s.v, i0, p.n, vn4, c_v);
  }
  if (MODEL == NSX_MODEL_IZHIKEVICH) {  // v and u are always written
    if (i0 + 4 <= p.n && live[3]) {
      *reinterpret_cast<float4*>(s.v + i0) = make_float4(v[0], v[1], v[2], v[3]);
      *reinterpret_cast<float4*>(s.u + i0) = make_float4(w1[0], w1[1], w1[2], w1[3]);
    } else {
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k)
        if (live[k]) {
          s.v[i0 + k] = v[k];
          s.u[i0 + k] = w1[k];
        }
    }
  }
  // the segment's spike list: per-lane counts, wavefront prefix sum, neuron order
  const uint32_t c = (uint32_t)__popc(spk);
  uint32_t incl = c;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  const uint32_t tot = __builtin_amdgcn_readfirstlane(__shfl(incl, 63));
  uint32_t pos = incl - c;
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k)
    if (spk & (1u << k)) s.spikes[b * kBlock + pos++] = i0 + k;
  if (lane == 0) {
    s.bcnt[b] = tot;
    if (tot) {
      s.part[kPartSpikes * p.nblk + b] += tot;
      s.ev_last[(b % kEvStripes) * kEvStride] = t;
      if (p.stats_on)
        atomicAdd(&s.stat_counts[((uint32_t)t % p.stat_cap) * kStatStripes + b % kStatStripes],
                  tot);
    }
  }
  if (p.raster_on && tot) {  // test/analysis path: append (step, id) pairs
    unsigned long long rb = 0;
    if (lane == 0) {
      rb = atomicAdd(&s.counters[kCntRaster], (unsigned long long)tot);
      atomicAdd(&s.step_counts[(uint32_t)t % p.step_counts_cap], tot);
    }
    rb = __shfl(rb, 0);
    uint32_t q = incl - c;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
      if (spk & (1u << k)) {
        const unsigned long long at = rb + q++;
        if (at < p.raster_cap) {
          s.raster_step[at] = (uint32_t)t;
          s.raster_id[at] = i0 + k;
        }
      }
  }
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
  float pre_v[64];       // pair STDP: the row source's pre record (v, v_prev, s_prev), loaded
  float pre_vp[64];      // once per row rather than once per slot
  int32_t pre_sp[64];
  uint32_t it_end[64];   // batched rows: inclusive prefix of the items' row counts
  uint32_t it_base[64];  // and each item's first spike-segment index
  uint32_t chan[64];     // ring channel of the row's input (two-channel models)
};

// Ring channel of source i's synapses: 1 for an inhibitory source (outside [0, n_exc) of its
// own network) in a two-channel model, else 0.
__device__ inline uint32_t source_channel(const StepParams& p, const DevPtrs& s, uint32_t i) {
  if (p.ch < 2) return 0u;
  if (!p.batch) return i >= p.n_exc ? 1u : 0u;
  const BatchNet& b = s.bnet[s.blk_net[i / kBlock]];
  return i - b.offset >= b.n_exc ? 1u : 0u;
}

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

// Binned delivery: a record is the event's bucket (d - 1, target tile) and the target's LDS
// entry in its tile's accumulator (channel-major), packed as (bucket << kBinEntryBits) | entry,
// with the delivered weight's bits.
constexpr uint32_t kBinEntryBits = 14;  // tile entries <= 16384
static_assert(kMaxBins << kBinEntryBits <= (1u << 27), "bucket and entry share one word");

// The slots of one row chunk (chunk_setup done): propagation and the STDP at the pre spike.
// Each lane takes NS slots per round, 64 apart, and issues all their loads before the first
// atomic or record store.  A slot past the chunk's end repeats the previous slot's loads (in
// bounds) and stores nothing.  Returns the potentiated / depressed-at-pre slot count of this
// lane.  NS: kPropSlots in the ring-atomic kernels (their occupancy bound); the binned pass's
// 1024-thread workgroups run at 4 waves per SIMD anyway, so it takes more slots in flight --
// measured: 4 under STDP (1M network 133.1 -> 130.8 us per step; 6 spills registers and ran
// 132.9), 6 without plasticity (population 0.1796 -> 0.1741 ms per step; 4: 0.1755, 8: 0.1740).
constexpr uint32_t kPropSlots = 2;
constexpr uint32_t kBinSlotsPlastic = 4, kBinSlotsPlain = 6;

template <int PL, bool BIN>
__device__ unsigned long long propagate_chunk(const StepParams& p, const DevPtrs& s,
                                              const RowChunk& rc, int32_t t, uint32_t total,
                                              uint32_t lane, uint32_t* pos, uint2* bev) {
  constexpr uint32_t NS =
      !BIN ? kPropSlots : PL == NSX_PLAST_NONE ? kBinSlotsPlain : kBinSlotsPlastic;
  const uint32_t tm = (uint32_t)t % p.D;
  const size_t plane = (size_t)p.ch * p.n;
  unsigned long long ltp = 0;
  uint32_t r = 0;
  for (uint32_t f = lane; f < total; f += 64 * NS) {
    bool has[NS], pot[NS];
    uint64_t k[NS];
    uint32_t rr[NS], src[NS], j[NS], d[NS];
    float w[NS];
#pragma unroll
    for (uint32_t q = 0; q < NS; ++q) {
      has[q] = q == 0 || f + 64 * q < total;
      uint32_t rq = q ? rr[q - 1] : r;
      k[q] = has[q] ? chunk_slot(rc, f + 64 * q, rq) : k[q - 1];
      rr[q] = rq;
      src[q] = rc.src[rq];
    }
#pragma unroll
    for (uint32_t q = 0; q < NS; ++q) j[q] = s.col[k[q]];
#pragma unroll
    for (uint32_t q = 0; q < NS; ++q) w[q] = s.w[k[q]];
#pragma unroll
    for (uint32_t q = 0; q < NS; ++q) d[q] = s.delay[k[q]];
    PostRec pr[NS];
    TraceRec a[NS];
#pragma unroll
    for (uint32_t q = 0; q < NS; ++q) {
      pot[q] = PL != NSX_PLAST_NONE && has[q] && exc_neuron(p, s, src[q]) &&
               exc_neuron(p, s, j[q]);
      pr[q] = PostRec{0.0f, kNeverStep, kNeverStep, 0u, {}};
      if (pot[q]) pr[q] = s.post[j[q]];  // 32 bytes: trace, last spike, pending spikes
      a[q] = TraceRec{};
    }
    if (PL == NSX_PLAST_PAIR) {  // the target's pending depressions precede the delivery
#pragma unroll
      for (uint32_t q = 0; q < NS; ++q) {
        if (pot[q]) {
          a[q] = TraceRec{rc.pre_v[rr[q]], 0, rc.pre_vp[rr[q]], rc.pre_sp[rr[q]]};
          w[q] = pair_pending(p, s, w[q], pr[q], a[q], t);
        }
      }
    }
#pragma unroll
    for (uint32_t q = 0; q < NS; ++q) {
      uint32_t sl = tm + d[q];
      sl = sl >= p.D ? sl - p.D : sl;
      const uint32_t ch = p.ch > 1 ? rc.chan[rr[q]] : 0u;
      bool atom = has[q];
      if (BIN && pos) {
        // appended to the workgroup's region of the step's scratch (pos[0] its cursor): one LDS
        // add per wavefront and slot round, the round's records stored contiguously, and
        // counted per bucket in pos[1 + bucket] (k_bin_place sorts them into their buckets).
        // Lane 0 runs every round that any lane runs.  Past the region's end: ring atomics.
        const unsigned long long bal = __ballot(has[q]);
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(pos, (uint32_t)__popcll(bal));
        base = __shfl(base, 0);
        const uint32_t k = base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
        if (has[q] && k < p.bin_gcap) {
          const uint32_t bin = (d[q] - 1u) * p.bin_ntiles + (j[q] >> p.bin_tshift);
          const uint32_t entry = (ch << p.bin_tshift) | (j[q] & ((1u << p.bin_tshift) - 1u));
          bev[k] = make_uint2((bin << kBinEntryBits) | entry, __float_as_uint(w[q]));
          atomicAdd(&pos[1 + bin], 1u);
          atom = false;
        }
      }
      if (atom) {
        atomicAdd((unsigned long long*)(s.ring + sl * plane + (size_t)ch * p.n + j[q]),
                  (unsigned long long)quantize(w[q], p.scale_in));
      }
    }
    if (PL == NSX_PLAST_TRIPLET) {
      // pre spike at t: depression by the target's last spike (same step included)
#pragma unroll
      for (uint32_t q = 0; q < NS; ++q) {
        if (pot[q]) {
          const float x = trace_at(s.mtab, p.mtab_n, t - pr[q].last);
          const float dlt = p.a_minus * x;
          const float wn = w[q] - dlt;
          s.w[k[q]] = clampf_(wn, p.w_min, p.w_max);
          ltp += 1;
        }
      }
    }
    if (PL == NSX_PLAST_PAIR) {
#pragma unroll
      for (uint32_t q = 0; q < NS; ++q) {
        if (pot[q]) {
          s.w[k[q]] = pair_at_pre(p, s, w[q], pr[q], a[q], t);
          ltp += 1;
        }
      }
    }
    r = rr[NS - 1];
  }
  return ltp;
}

// Per-block counters of a propagation item: events and potentiated slots of this wavefront.
template <int PL>
__device__ inline void item_counters(const StepParams& p, const DevPtrs& s, uint32_t b,
                                     unsigned long long ev, unsigned long long ltp,
                                     uint32_t lane) {
  if (PL != NSX_PLAST_NONE) {
#pragma unroll
    for (uint32_t o = 32; o > 0; o >>= 1) ltp += __shfl_xor(ltp, o);
  }
  if (lane == 0) {
    // parts of one block (or, in an epoch, several steps' items of one block) accumulate into
    // that block's counter
    if (p.prop_split == 1 && !p.part_atomic) {
      s.part[kPartEvents * p.nblk + b] += ev;
      if (PL != NSX_PLAST_NONE) s.part[kPartLtp * p.nblk + b] += ltp;
    } else {
      atomicAdd(&s.part[kPartEvents * p.nblk + b], ev);
      if (PL != NSX_PLAST_NONE) atomicAdd(&s.part[kPartLtp * p.nblk + b], ltp);
    }
  }
}

// Propagation with a row-split threshold T = p.row_split: the spiking rows of block b are cut
// into pieces of at most T slots (a row of d slots into ceil(d / T)), the pieces numbered in row
// order, and part q of P takes pieces q, q + P, ...  A high-degree row is then walked by several
// wavefronts instead of one.  Each slot is still delivered and, under STDP, updated exactly once.
template <int PL>
__device__ void propagate_pieces(const StepParams& p, const DevPtrs& s, RowChunk& rc, int32_t t,
                                 uint32_t b, uint32_t q, uint32_t cnt, uint32_t lane) {
  const uint32_t P = p.prop_split, T = p.row_split;
  unsigned long long ev = 0, ltp = 0;
  uint32_t before = 0;  // pieces of the rows of earlier row chunks
  for (uint32_t r0 = 0; r0 < cnt; r0 += 64) {
    const uint32_t nr = min(64u, cnt - r0);
    uint32_t i = 0, deg = 0;
    if (lane < nr) {
      i = s.spikes[b * kBlock + r0 + lane];
      deg = (uint32_t)(s.row_ptr[i + 1] - s.row_ptr[i]);
    }
    uint32_t incl = (deg + T - 1) / T;  // pieces of this lane's row, then their prefix
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    const uint32_t tot = __builtin_amdgcn_readfirstlane(__shfl(incl, 63));
    rc.it_end[lane] = incl;
    rc.it_base[lane] = i;
    wave_sync();
    // this part's pieces of the chunk: g with (before + g) mod P == q
    for (uint32_t g0 = (q + P - before % P) % P; g0 < tot; g0 += 64 * P) {
      const uint32_t g = g0 + lane * P;
      uint32_t pi = 0, pdeg = 0;
      uint64_t pbeg = 0;
      if (g < tot) {
        uint32_t r = 0;
#pragma unroll
        for (uint32_t st = 32; st > 0; st >>= 1)
          if (rc.it_end[r + st - 1] <= g) r += st;
        const uint32_t k = g - (r ? rc.it_end[r - 1] : 0u);  // piece k of row r
        pi = rc.it_base[r];
        const uint64_t rb = s.row_ptr[pi];
        const uint32_t rd = (uint32_t)(s.row_ptr[pi + 1] - rb);
        pbeg = rb + (uint64_t)k * T;
        pdeg = min(T, rd - k * T);
        if (p.ch > 1) rc.chan[lane] = source_channel(p, s, pi);
        if (PL == NSX_PLAST_PAIR && exc_neuron(p, s, pi)) {
          const TraceRec a = s.pre[pi];
          rc.pre_v[lane] = a.v;
          rc.pre_vp[lane] = a.v_prev;
          rc.pre_sp[lane] = a.s_prev;
        }
      }
      const uint32_t total = chunk_setup(rc, lane, pdeg, pbeg, pi);
      ev += total;
      ltp += propagate_chunk<PL, false>(p, s, rc, t, total, lane, nullptr, nullptr);
      wave_sync();
    }
    before += tot;
    wave_sync();  // it_end / it_base are rewritten by the next row chunk
  }
  item_counters<PL>(p, s, b, ev, ltp, lane);
}

// Propagation work item: part q of neuron block b.  PL: STDP applied on the spiking rows' E->E
// slots -- pair: the target's pending depressions, potentiation by the post trace (then
// depression if the target fired at t); triplet: depression by the target's last spike.
// PIECES: the row-split instance (p.row_split > 0, atomic delivery); a separate kernel instance
// keeps the whole-row path's registers unchanged.
template <int PL, bool BIN, bool PIECES = false>
__device__ void propagate_item(const StepParams& p, const DevPtrs& s, RowChunk& rc, int32_t t,
                               uint32_t b, uint32_t q, uint32_t cnt, uint32_t lane,
                               uint32_t* pos, uint2* bev) {
  const uint32_t P = p.prop_split;
  if constexpr (PIECES && !BIN) {
    if (cnt) propagate_pieces<PL>(p, s, rc, t, b, q, cnt, lane);
    return;
  }
  if (cnt <= q) return;
  const uint32_t mine = (cnt - q + P - 1) / P;  // rows q, q+P, ... of the segment
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
    if (p.ch > 1 && lane < nr) rc.chan[lane] = source_channel(p, s, i);
    if (PL == NSX_PLAST_PAIR && lane < nr && exc_neuron(p, s, i)) {
      const TraceRec a = s.pre[i];
      rc.pre_v[lane] = a.v;
      rc.pre_vp[lane] = a.v_prev;
      rc.pre_sp[lane] = a.s_prev;
    }
    const uint32_t total = chunk_setup(rc, lane, deg, beg, i);
    ev += total;
    ltp += propagate_chunk<PL, BIN>(p, s, rc, t, total, lane, pos, bev);
    wave_sync();
  }
  item_counters<PL>(p, s, b, ev, ltp, lane);
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
      if (exc_neuron(p, s, j) && (PL != NSX_PLAST_PAIR || s.post[j].npend == kPostPend)) {
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

// The work-item loop of the synapse pass (and of binned delivery's counting pass, which must
// visit exactly the same propagation items in the same workgroups): wavefront w of the grid
// owns items w, w + W, ...; fn(prop, b, q, cnt) runs for the items whose block spiked.
template <int PL, uint32_t NT, typename Fn>
__device__ inline void for_items(const StepParams& p, const DevPtrs& s, Fn&& fn) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * NT + threadIdx.x) >> 6);
  const uint32_t P = p.prop_split;
  const uint32_t nprop = p.nblk * P;
  const uint32_t nltd = PL != NSX_PLAST_NONE ? ((p.n_exc + kBlock - 1) / kBlock) * P : 0u;
  const uint32_t both = 2 * min(nprop, nltd);
  const uint32_t nwork = nprop + nltd;
  const uint32_t nwaves = (gridDim.x * NT) >> 6;
  for (uint64_t base = wave; base < nwork; base += 64ull * nwaves) {
    const uint64_t mine = base + (uint64_t)lane * nwaves;
    uint32_t c = 0, qq = 0;
    if (mine < nwork) {
      bool pr;
      uint32_t bb;
      decode_item((uint32_t)mine, both, nprop, nltd, P, pr, bb, qq);
      c = s.bcnt[bb];
    }
    // with a row-split threshold a part may hold pieces even when the block has fewer rows
    unsigned long long act = __ballot(mine < nwork && c > (p.row_split ? 0u : qq));
    while (act) {
      const uint32_t l = (uint32_t)(__ffsll((long long)act) - 1);
      act &= act - 1ull;
      bool prop;
      uint32_t b, q;
      decode_item((uint32_t)(base + (uint64_t)l * nwaves), both, nprop, nltd, P, prop, b, q);
      const uint32_t cnt = __builtin_amdgcn_readfirstlane(__shfl(c, (int)l));
      fn(prop, b, q, cnt);
    }
  }
}

// The same items and wavefronts as for_items, but a batch's propagation items are handed over
// together: prop(mine, b, q, cnt) runs once per batch on every lane (lane l: item l of the batch,
// mine = it is a propagation item with rows), then post(b, q, cnt) once per post-spike item.
template <int PL, uint32_t NT, typename PropFn, typename PostFn>
__device__ inline void for_batches(const StepParams& p, const DevPtrs& s, PropFn&& prop,
                                   PostFn&& post) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * NT + threadIdx.x) >> 6);
  const uint32_t P = p.prop_split;
  const uint32_t nprop = p.nblk * P;
  const uint32_t nltd = PL != NSX_PLAST_NONE ? ((p.n_exc + kBlock - 1) / kBlock) * P : 0u;
  const uint32_t both = 2 * min(nprop, nltd);
  const uint32_t nwork = nprop + nltd;
  const uint32_t nwaves = (gridDim.x * NT) >> 6;
  for (uint64_t base = wave; base < nwork; base += 64ull * nwaves) {
    const uint64_t item = base + (uint64_t)lane * nwaves;
    uint32_t c = 0, qq = 0, bb = 0;
    bool pr = false;
    if (item < nwork) {
      decode_item((uint32_t)item, both, nprop, nltd, P, pr, bb, qq);
      c = s.bcnt[bb];
    }
    const bool live = item < nwork && c > qq;
    prop(live && pr, bb, qq, c);
    unsigned long long act = __ballot(live && !pr);
    while (act) {
      const uint32_t l = (uint32_t)(__ffsll((long long)act) - 1);
      act &= act - 1ull;
      const uint32_t b = __builtin_amdgcn_readfirstlane(__shfl(bb, (int)l));
      const uint32_t q = __builtin_amdgcn_readfirstlane(__shfl(qq, (int)l));
      const uint32_t cnt = __builtin_amdgcn_readfirstlane(__shfl(c, (int)l));
      post(b, q, cnt);
    }
  }
}

// Seven wavefronts per SIMD (72 VGPRs) for the pair-STDP instance: fits without spills.  NT threads
// per workgroup (64 .. 512; chosen by the autotuner): work items are per wavefront, the size only
// changes how the wavefronts are grouped onto CUs.
template <int PL, uint32_t NT, bool PIECES>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(7, 8)))
void k_synapses(StepParams p, DevPtrs s, int32_t off) {
  __shared__ RowChunk sh[NT / 64];
  const int32_t t = *s.tbase + off;
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  for_items<PL, NT>(p, s, [&](bool prop, uint32_t b, uint32_t q, uint32_t cnt) {
    if (prop)
      propagate_item<PL, false, PIECES>(p, s, sh[wid], t, b, q, cnt, lane, nullptr, nullptr);
    else if (PL != NSX_PLAST_NONE)
      depress_item<PL>(p, s, sh[wid], t, b, q, cnt, lane);
  });
}

// ---- persistent stepping --------------------------------------------------------------------
// nsx_set_persistent: a window of steps in one cooperative launch of a grid that is resident at
// once.  Per step: the update of every neuron block (update_block, the workgroups taking the
// blocks in turn), a grid barrier, the synapse pass with ring atomics (for_items over the same
// work items as k_synapses), a grid barrier.  Every step's arithmetic and memory effects are
// those of k_update and k_synapses, so the run is bit-identical to per-step launches.
//
// Grid barrier: bar[0] counts arrivals and only grows within a launch (zeroed before it), so
// barrier k of a workgroup waits until bar[0] reaches (k + 1) x gridDim.x.  Every wavefront's
// writes reach the device scope through its own release fence before the workgroup's arrival;
// the acquire fences after the wait (every wavefront's, past the workgroup barrier) make the
// other workgroups' writes visible.  The wait is bounded: past kBarrierSpins polls, or when
// another workgroup has given up (bar[1] != 0), a workgroup sets bar[1] and leaves the kernel,
// so every wavefront exits; the host then reports the failure.
constexpr uint32_t kBarrierSpins = 1u << 22;

__device__ inline bool grid_sync(uint32_t* bar, uint32_t& gen) {
  __shared__ uint32_t s_ok;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t target = (gen + 1u) * gridDim.x;
    atomicAdd(&bar[0], 1u);
    uint32_t ok = 1u, spins = 0u;
    while (__hip_atomic_load(&bar[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
          ++spins > kBarrierSpins) {
        atomicOr(&bar[1], 1u);
        ok = 0u;
        break;
      }
      __builtin_amdgcn_s_sleep(4);
    }
    s_ok = ok;
  }
  __syncthreads();
  __threadfence();
  gen += 1u;
  return s_ok != 0u;
}

// The grid holds kPersistPerCu workgroups per CU: the compiler guarantees that many wavefronts
// per SIMD (registers, scalar ones included, fit), so they are resident together.  The runtime's
// occupancy query counts vector registers and LDS, not scalar registers, and can report more
// workgroups per CU than fit: a grid sized by it alone waits at the first barrier for
// workgroups that never start.
constexpr uint32_t kPersistPerCu = 5;

template <int MODEL, int DRIVE, int PL>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kPersistPerCu)))
void k_persist(StepParams p, DevPtrs s, int32_t off, uint32_t nsteps, uint32_t* bar) {
  __shared__ RowChunk sh[kBlock / 64];
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  uint32_t gen = 0u;
  for (uint32_t k = 0; k < nsteps; ++k) {
    const int32_t t = *s.tbase + off + (int32_t)k;
    for (uint32_t b = blockIdx.x; b < p.nblk; b += gridDim.x) {
      update_block<MODEL, DRIVE, PL>(p, s, t, b);
      __syncthreads();  // the next block reuses the drive words' and wave counts' LDS
    }
    if (!grid_sync(bar, gen)) return;
    for_items<PL, kBlock>(p, s, [&](bool prop, uint32_t b, uint32_t q, uint32_t cnt) {
      if (prop) {
        if (p.row_split)  // pieces of high-degree rows, as k_synapses<PL, NT, true>
          propagate_item<PL, false, true>(p, s, sh[wid], t, b, q, cnt, lane, nullptr, nullptr);
        else
          propagate_item<PL, false, false>(p, s, sh[wid], t, b, q, cnt, lane, nullptr, nullptr);
      } else if (PL != NSX_PLAST_NONE) {
        depress_item<PL>(p, s, sh[wid], t, b, q, cnt, lane);
      }
    });
    if (!grid_sync(bar, gen)) return;
  }
}

__global__ void k_zero_bar(uint32_t* bar) {
  if (threadIdx.x < 2) bar[threadIdx.x] = 0u;
}

// ---- epoch stepping -------------------------------------------------------------------------
// Networks without plasticity whose shortest delay is d_min >= 2 steps: the spikes of steps
// t .. t + d_min - 1 arrive at t + d_min at the earliest, so those steps' updates depend on no
// propagation among them.  k_epoch_update advances `len` <= d_min steps in one launch with each
// neuron's state in registers (one load and one store per epoch instead of per step), compacting
// each step's spikes into that step's segment; k_synapses_epoch then propagates the spikes of all
// `len` steps, each to ring[(t_k + d) mod D].  The arithmetic of a step is k_update's, operation
// for operation, so an epoch run equals the per-step run bit for bit.
template <int MODEL, int DRIVE>
__global__ __launch_bounds__(kBlock) void k_epoch_update(StepParams p, DevPtrs s, EpochSeg ep,
                                                         int32_t off, uint32_t len) {
  constexpr bool TWO = MODEL != NSX_MODEL_IZHIKEVICH;
  __shared__ uint32_t s_drive[kBlock];
  const int32_t t0 = *s.tbase + off;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const NeuronId id = neuron_id(p, s, i, blockIdx.x);
  const bool live = i < p.n && id.local < id.n;
  // latest spike step before the epoch: slot t can hold input only if it is >= t - D + 1 (the
  // epoch's own spikes arrive after it)
  int32_t m0 = kNeverStep;
#pragma unroll
  for (uint32_t k = 0; k < kEvStripes; ++k) m0 = max(m0, s.ev_last[k * kEvStride]);
  float v = 0.0f, w1 = 0.0f, w2 = 0.0f, w3 = 0.0f;  // LIF: v ie ii; Izh: v u; HH: v m h n
  int32_t refr = 0, last = kNeverStep;
  uint32_t nspk = 0;
  HHSyn y{0.0f, 0.0f, 0.0f, 0.0f};
  if (live) {
    v = s.v[i];
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
  const float v_in = v, w1_in = w1, w2_in = w2;
  const int32_t refr_in = refr;
  const size_t plane = (size_t)p.ch * p.n;
  // step k's ring slot is loaded during step k - 1
  long long qa = 0, qb = 0;
  if (live && m0 >= t0 - (int32_t)p.D + 1) {
    const long long* r = s.ring + (size_t)((uint32_t)t0 % p.D) * plane + i;
    qa = r[0];
    if (TWO) qb = r[p.n];
  }
  for (uint32_t k = 0; k < len; ++k) {
    const int32_t t = t0 + (int32_t)k;
    long long* rk = s.ring + (size_t)((uint32_t)t % p.D) * plane + i;
    const long long a = qa, b = qb;
    qa = 0;
    qb = 0;
    if (live && k + 1 < len && m0 >= t + 1 - (int32_t)p.D + 1) {
      const long long* rn = s.ring + (size_t)((uint32_t)(t + 1) % p.D) * plane + i;
      qa = rn[0];
      if (TWO) qb = rn[p.n];
    }
    __syncthreads();  // the previous step's readers of the block's LDS are done
    drive_stage<DRIVE>(p, s, t, s_drive, id);
    if (DRIVE != NSX_DRIVE_NONE) __syncthreads();
    bool spiked = false;
    float lfp = 0.0f;
    if (live) {
      float a0 = dequantize(a, p.scale_out);
      if (a != 0) rk[0] = 0;
      float a1 = 0.0f;
      if (TWO) {
        a1 = dequantize(b, p.scale_out);
        if (b != 0) rk[p.n] = 0;
      }
      float z = 0.0f;
      drive_draw<DRIVE>(p, s_drive, a0, z);
      if (MODEL == NSX_MODEL_LIF) {
        float ie = w1 * p.dec_e;
        ie = ie + a0;
        float ii = w2 * p.dec_i;
        ii = ii + a1;
        w1 = ie;
        w2 = ii;
        if (refr > 0) {
          refr = refr - 1;
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
            refr = p.refr_steps;
            spiked = true;
          }
          v = vn;
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
          const float aa = p.izh_a[ty], bb = p.izh_b[ty];
          const float v0 = v;
          for (uint32_t q = 0; q < p.substeps; ++q) {
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
          float t3 = bb * v0;
          t3 = t3 - u;
          t3 = aa * t3;
          t3 = t3 * p.dtf;
          u = u + t3;
        }
        w1 = u;
      } else {
        const float I = hh_input<DRIVE>(p, y, v, a0, a1, z, lfp);
        float vn = v;
        for (uint32_t q = 0; q < p.hh_substeps; ++q) hh_substep(vn, w1, w2, w3, I, p.hh);
        if (p.hh_reset_mode) {
          spiked = vn >= -55.0f;
          if (spiked) vn = -70.0f;
        } else {
          spiked = (v < p.hh_spike_v) && (vn >= p.hh_spike_v);
        }
        v = vn;
      }
      if (spiked) {
        nspk += 1u;
        last = t;
      }
    }
    if (MODEL == NSX_MODEL_HH && p.stats_on) lfp_accumulate(s, lfp, blockIdx.x);
    DevPtrs sk = s;
    sk.spikes = ep.spikes + ((size_t)k * p.nblk) * kBlock;
    sk.bcnt = ep.bcnt + (size_t)k * p.nblk;
    block_compact(p, sk, spiked, i, t, blockIdx.x);
  }
  if (live) {
    if (MODEL == NSX_MODEL_LIF) {  // unchanged values are not rewritten, as in k_update
      if (__float_as_uint(w1) != __float_as_uint(w1_in)) s.s0[i] = w1;
      if (__float_as_uint(w2) != __float_as_uint(w2_in)) s.s1[i] = w2;
      if (refr != refr_in) s.refr[i] = refr;
      if (__float_as_uint(v) != __float_as_uint(v_in)) s.v[i] = v;
    } else if (MODEL == NSX_MODEL_IZHIKEVICH) {
      s.v[i] = v;
      s.u[i] = w1;
    } else {
      s.v[i] = v;
      s.gm[i] = w1;
      s.gh[i] = w2;
      s.gn[i] = w3;
      hh_syn_store<DRIVE>(s, i, y);
    }
    if (nspk) {
      s.count[i] += nspk;
      s.last[i] = last;
    }
  }
}

// The synapse pass of an epoch: propagation items (step k, block b, part q) of every step,
// handed to the wavefronts as for_items does for one step.
template <int PL, bool PIECES>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7, 8)))
void k_synapses_epoch(StepParams p, DevPtrs s, EpochSeg ep, int32_t off, uint32_t len) {
  __shared__ RowChunk sh[kBlock / 64];
  const int32_t t0 = *s.tbase + off;
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * kBlock + threadIdx.x) >> 6);
  const uint32_t P = p.prop_split;
  const uint64_t per = (uint64_t)p.nblk * P;
  const uint64_t nwork = per * len;
  const uint32_t nwaves = (gridDim.x * kBlock) >> 6;
  for (uint64_t base = wave; base < nwork; base += 64ull * nwaves) {
    const uint64_t mine = base + (uint64_t)lane * nwaves;
    uint32_t c = 0, qq = 0;
    if (mine < nwork) {
      const uint32_t kk = (uint32_t)(mine / per);
      const uint32_t rem = (uint32_t)(mine - (uint64_t)kk * per);
      const uint32_t bb = rem / P;
      qq = rem - bb * P;
      c = ep.bcnt[(size_t)kk * p.nblk + bb];
    }
    unsigned long long act = __ballot(mine < nwork && c > (p.row_split ? 0u : qq));
    while (act) {
      const uint32_t l = (uint32_t)(__ffsll((long long)act) - 1);
      act &= act - 1ull;
      const uint64_t item = base + (uint64_t)l * nwaves;
      const uint32_t kk = __builtin_amdgcn_readfirstlane((uint32_t)(item / per));
      const uint32_t rem = (uint32_t)(item - (uint64_t)kk * per);
      const uint32_t b = __builtin_amdgcn_readfirstlane(rem / P);
      const uint32_t q = rem - b * P;
      const uint32_t cnt = __builtin_amdgcn_readfirstlane(__shfl(c, (int)l));
      DevPtrs sk = s;
      sk.spikes = ep.spikes + ((size_t)kk * p.nblk) * kBlock;
      propagate_item<PL, false, PIECES>(p, sk, sh[wid], t0 + (int32_t)kk, b, q, cnt, lane,
                                        nullptr, nullptr);
    }
  }
}

// ---- binned delivery ------------------------------------------------------------------------
// Per step t, after the update: (1) k_synapses_bin runs the synapse pass; each workgroup
// appends its events' records to its own fixed-capacity region of the step's scratch (one LDS
// cursor add per wavefront and round, stores contiguous) and counts them per bucket (delay,
// target tile) in LDS, then writes its row of counts; events past its region's end take the
// ring atomics; (2) k_bin_scan turns the [groups][buckets] counts into per-group offsets within
// each bucket and bucket totals; (3) k_bin_place, one workgroup per region, sorts the region
// into the buckets' positions of the step's event buffer (start_b + the group's offset) and
// publishes the bucket starts; (4) k_bin_arrive, one workgroup per target tile, sums the
// tile's buckets of every emission step whose events arrive at t + 1 into LDS int64
// accumulators and adds them into ring slot (t + 1) mod D.  Integer sums do not depend on the
// order of the events, so the ring equals the one atomic delivery builds.

// Spiking rows of all the propagation items of a wavefront's batch (lane l: item (b, q, cnt),
// `mine` when it is a propagation item whose segment part has rows), walked together in chunks
// of 64 rows: far fewer dependent memory rounds than one item at a time when items are small.
// fn(rc, total) runs once per chunk after chunk_setup; returns the rows' slot count.
template <typename Fn>
__device__ inline unsigned long long batch_rows(const StepParams& p, const DevPtrs& s,
                                                RowChunk& rc, uint32_t lane, bool mine,
                                                uint32_t b, uint32_t q, uint32_t cnt,
                                                bool pre_rec, Fn&& fn) {
  const uint32_t P = p.prop_split;
  const uint32_t m = mine ? (cnt - q + P - 1) / P : 0u;
  uint32_t incl = m;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  const uint32_t R = __builtin_amdgcn_readfirstlane(__shfl(incl, 63));
  if (R == 0) return 0;
  rc.it_end[lane] = incl;
  rc.it_base[lane] = b * kBlock + q;
  wave_sync();
  unsigned long long ev = 0;
  for (uint32_t r0 = 0; r0 < R; r0 += 64) {
    const uint32_t r = r0 + lane;
    uint32_t i = 0, deg = 0;
    uint64_t beg = 0;
    if (r < R) {
      uint32_t l = 0;  // item of row r: lanes whose inclusive prefix is <= r come before it
#pragma unroll
      for (uint32_t st = 32; st > 0; st >>= 1)
        if (rc.it_end[l + st - 1] <= r) l += st;
      const uint32_t k = r - (l ? rc.it_end[l - 1] : 0u);
      i = s.spikes[rc.it_base[l] + k * P];
      beg = s.row_ptr[i];
      deg = (uint32_t)(s.row_ptr[i + 1] - beg);
      if (p.ch > 1) rc.chan[lane] = source_channel(p, s, i);
      if (pre_rec && exc_neuron(p, s, i)) {
        const TraceRec a = s.pre[i];
        rc.pre_v[lane] = a.v;
        rc.pre_vp[lane] = a.v_prev;
        rc.pre_sp[lane] = a.s_prev;
      }
    }
    const uint32_t total = chunk_setup(rc, lane, deg, beg, i);
    ev += total;
    fn(total);
    wave_sync();
  }
  return ev;
}

// kScanCols buckets per workgroup of kScanThreads: thread (column c, row group rg) holds its
// column's counts of groups [rg * per, (rg + 1) * per) in registers (per <= kScanPer, groups <=
// 1024), the row-group sums of a column are scanned in LDS, and the exclusive offsets are written
// back in place; the column totals go to bin_tot.  Each load and store instruction covers row
// pieces of kScanCols counts: 32 make them whole 128-byte lines (measured on the 1M network:
// bucket passes 34.3 -> 32.8 us per step against 16 columns in 256 threads).
constexpr uint32_t kScanThreads = 1024, kScanCols = 32;
constexpr uint32_t kScanRows = kScanThreads / kScanCols, kScanPer = 1024 / kScanRows;
__global__ __launch_bounds__(kScanThreads) void k_bin_scan(StepParams p, DevPtrs s) {
  __shared__ uint32_t part[kScanRows][kScanCols];
  const uint32_t nb = p.bin_nbins, G = p.bin_groups;
  const uint32_t c = threadIdx.x % kScanCols, rg = threadIdx.x / kScanCols;
  const uint32_t col = blockIdx.x * kScanCols + c;
  const uint32_t per = (G + kScanRows - 1) / kScanRows;
  const uint32_t r0 = rg * per;
  uint32_t v[kScanPer], sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanPer; ++k) {
    const uint32_t r = r0 + k;
    v[k] = (k < per && r < G && col < nb) ? s.bin_hist[(size_t)r * nb + col] : 0u;
    sum += v[k];
  }
  part[rg][c] = sum;
  __syncthreads();
  uint32_t run = 0, tot = 0;
#pragma unroll
  for (uint32_t g = 0; g < kScanRows; ++g) {
    const uint32_t x = part[g][c];
    run += g < rg ? x : 0u;
    tot += x;
  }
#pragma unroll
  for (uint32_t k = 0; k < kScanPer; ++k) {
    const uint32_t r = r0 + k;
    if (k < per && r < G && col < nb) s.bin_hist[(size_t)r * nb + col] = run;
    run += v[k];
  }
  if (rg == 0 && col < nb) s.bin_tot[col] = tot;
}

template <int PL>
__global__ __launch_bounds__(kBinThreads)
void k_synapses_bin(StepParams p, DevPtrs s, int32_t off) {
  __shared__ RowChunk sh[kBinThreads / 64];
  __shared__ uint32_t cnt[1 + kMaxBins];  // [0] the region's append cursor, [1 + b] bucket b
  const int32_t t = *s.tbase + off;
  const uint32_t nb = p.bin_nbins, lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  for (uint32_t k = threadIdx.x; k <= nb; k += kBinThreads) cnt[k] = 0u;
  __syncthreads();
  uint2* scr = s.bin_scr + (size_t)blockIdx.x * p.bin_gcap;
  RowChunk& rc = sh[wid];
  for_batches<PL, kBinThreads>(
      p, s,
      [&](bool mine, uint32_t b, uint32_t q, uint32_t c) {
        unsigned long long ltp = 0;
        const unsigned long long ev = batch_rows(
            p, s, rc, lane, mine, b, q, c, PL == NSX_PLAST_PAIR, [&](uint32_t total) {
              ltp += propagate_chunk<PL, true>(p, s, rc, t, total, lane, cnt, scr);
            });
        if (PL != NSX_PLAST_NONE) {
#pragma unroll
          for (uint32_t o = 32; o > 0; o >>= 1) ltp += __shfl_xor(ltp, o);
        }
        // totals only: a batch's counts go to the block of its first propagation item
        const unsigned long long any = __ballot(mine);
        if (any) {
          const uint32_t b0 = __builtin_amdgcn_readfirstlane(
              __shfl(b, (int)(__ffsll((long long)any) - 1)));
          if (lane == 0) {
            atomicAdd(&s.part[kPartEvents * p.nblk + b0], ev);
            if (PL != NSX_PLAST_NONE) atomicAdd(&s.part[kPartLtp * p.nblk + b0], ltp);
          }
        }
      },
      [&](uint32_t b, uint32_t q, uint32_t c) {
        if (PL != NSX_PLAST_NONE) depress_item<PL>(p, s, rc, t, b, q, c, lane);
      });
  __syncthreads();
  uint32_t* row = s.bin_hist + (size_t)blockIdx.x * nb;
  for (uint32_t k = threadIdx.x; k < nb; k += kBinThreads) row[k] = cnt[1 + k];
  if (threadIdx.x == 0)
    s.bin_reg[blockIdx.x] = make_uint2(blockIdx.x * p.bin_gcap, min(cnt[0], p.bin_gcap));
}

// Placement of step t's records: workgroup g moves its region of the scratch into the buckets'
// positions of the step's event buffer, start_b + off_(g,b) + (its records of bucket b so far),
// in chunks of bin_place_ch records: counted per bucket in LDS (each record's rank within its
// bucket kept), the counts scanned into chunk starts, the records staged in bucket order and
// copied out in that order, so each bucket's run is written contiguously.  Group 0 publishes
// the step's bucket starts (every step).  Dynamic LDS: gpos[nb], cst[nb], the staged chunk.
// PER records per thread and chunk: kPlacePer with one workgroup per CU (as many groups as CUs),
// kPlacePer2 when the groups outnumber the CUs and two workgroups share each CU (64 VGPRs,
// <= 80 KB of LDS each).
constexpr uint32_t kPlacePer = 16, kPlacePer2 = 4;
template <uint32_t PER>
__global__ __launch_bounds__(kBinThreads)
__attribute__((amdgpu_waves_per_eu(PER > kPlacePer2 ? 4 : 8)))
void k_bin_place(StepParams p, DevPtrs s, int32_t off) {
  extern __shared__ uint32_t place_lds[];
  __shared__ uint32_t wsum[kBinThreads / 64];
  const int32_t t = *s.tbase + off;
  const uint32_t nb = p.bin_nbins, g = blockIdx.x, tid = threadIdx.x;
  const uint32_t lane = tid & 63u, wid = tid >> 6;
  const uint32_t slot = (uint32_t)t % p.D;
  uint32_t* gpos = place_lds;
  uint32_t* cst = place_lds + nb;
  uint2* stage = reinterpret_cast<uint2*>(place_lds + 2 * nb);  // 8 nb bytes: 8-byte aligned
  const uint32_t* offr = s.bin_hist + (size_t)g * nb;
  // exclusive scan of 8 values per thread over the workgroup (wsum: the waves' totals); returns
  // the thread's exclusive prefix and the total
  auto scan8 = [&](const uint32_t v[8], uint32_t& total) {
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) sum += v[k];
    uint32_t incl = sum;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    uint32_t run = incl - sum;
    total = 0;
#pragma unroll
    for (uint32_t q = 0; q < kBinThreads / 64; ++q) {
      run += q < wid ? wsum[q] : 0u;
      total += wsum[q];
    }
    __syncthreads();  // wsum is read before the next scan writes it
    return run;
  };
  {  // bucket starts; this group's positions; group 0 publishes the table, every step
    uint32_t v[8], total = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) v[k] = tid * 8 + k < nb ? s.bin_tot[tid * 8 + k] : 0u;
    uint32_t run = scan8(v, total);
    uint32_t* tab = s.bin_tab + (size_t)slot * (nb + 1);
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
      const uint32_t b = tid * 8 + k;
      if (b < nb) {
        gpos[b] = run + offr[b];
        if (g == 0) tab[b] = run;
      }
      run += v[k];
    }
    if (g == 0 && tid == 0) tab[nb] = total;
  }
  const uint2 reg = s.bin_reg[g];
  const uint2* src = s.bin_scr + reg.x;
  uint2* ev = s.bin_ev + (size_t)slot * p.bin_cap;
  const uint32_t CH = p.bin_place_ch, mask = (1u << kBinEntryBits) - 1u;
  for (uint32_t c0 = 0; c0 < reg.y; c0 += CH) {
    const uint32_t n = min(CH, reg.y - c0);
    for (uint32_t b = tid; b < nb; b += kBinThreads) cst[b] = 0u;
    __syncthreads();
    uint2 x[PER];
    uint32_t rk[PER];
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
      const uint32_t i = u * kBinThreads + tid;
      x[u] = i < n ? src[c0 + i] : make_uint2(0u, 0u);
    }
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
      const uint32_t i = u * kBinThreads + tid;
      rk[u] = i < n ? atomicAdd(&cst[x[u].x >> kBinEntryBits], 1u) : 0u;
    }
    __syncthreads();
    {  // the chunk's counts -> chunk starts, in place
      uint32_t c[8], total = 0;
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k) c[k] = tid * 8 + k < nb ? cst[tid * 8 + k] : 0u;
      uint32_t run = scan8(c, total);
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k) {
        if (tid * 8 + k < nb) cst[tid * 8 + k] = run;
        run += c[k];
      }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
      const uint32_t i = u * kBinThreads + tid;
      if (i < n) stage[cst[x[u].x >> kBinEntryBits] + rk[u]] = x[u];
    }
    __syncthreads();
    for (uint32_t i = tid; i < n; i += kBinThreads) {
      const uint2 y = stage[i];
      const uint32_t b = y.x >> kBinEntryBits;
      ev[gpos[b] + (i - cst[b])] = make_uint2(y.x & mask, y.y);
    }
    __syncthreads();
    for (uint32_t b = tid; b < nb; b += kBinThreads)
      gpos[b] += (b + 1u < nb ? cst[b + 1] : n) - cst[b];
    __syncthreads();  // the next chunk clears cst
  }
}

// Arrival at step ta = *tbase + off of the events emitted at steps e in [ta - D + 1, ta - 1]
// with e <= emax (emission steps after emax have not run), for target tile blockIdx.x and
// tile size TE (entries = channels x neurons).  The tile's bucket ranges are staged in LDS and
// walked as one flattened list, eight events per thread in flight.
template <uint32_t TE>
__global__ __launch_bounds__(kBinThreads) void k_bin_arrive(StepParams p, DevPtrs s, int32_t off,
                                                            int32_t emax) {
  __shared__ long long acc[TE];
  __shared__ uint32_t seg_lo[kMaxArriveSegs];
  __shared__ uint32_t seg_end[kMaxArriveSegs];  // inclusive prefix of segment lengths
  __shared__ uint32_t seg_buf[kMaxArriveSegs];
  const int32_t ta = *s.tbase + off;
  const uint32_t nb = p.bin_nbins, T = blockIdx.x, nseg = p.D - 1u;
  for (uint32_t k = threadIdx.x; k < TE; k += kBinThreads) acc[k] = 0;
  // segment d - 1: bucket (d - 1, T) of emission step ta - d
  uint32_t len = 0, lo = 0, bufi = 0;
  if (threadIdx.x < nseg) {
    const int32_t e = ta - (int32_t)(threadIdx.x + 1u);
    if (e >= 0 && e <= emax) {
      bufi = (uint32_t)e % p.D;
      const uint32_t* tab = s.bin_tab + (size_t)bufi * (nb + 1);
      const uint32_t bin = threadIdx.x * p.bin_ntiles + T;
      lo = tab[bin];
      len = tab[bin + 1] - lo;
    }
  }
  // inclusive scan of len over the (at most kMaxArriveSegs) segments
  __shared__ uint32_t wtot[kBinThreads / 64];
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  uint32_t incl = len;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  if (lane == 63) wtot[wid] = incl;
  __syncthreads();
  uint32_t pre = 0, total = 0;
#pragma unroll
  for (uint32_t q = 0; q < kBinThreads / 64; ++q) {
    const uint32_t x = wtot[q];
    pre += q < wid ? x : 0u;
    total += x;
  }
  if (threadIdx.x < nseg) {
    seg_lo[threadIdx.x] = lo;
    seg_end[threadIdx.x] = pre + incl;
    seg_buf[threadIdx.x] = bufi;
  }
  __syncthreads();
  const uint2* ev = s.bin_ev;
  const size_t cap = p.bin_cap;
  // Loads are unconditional (an index past the end is clamped to the last event) so that the
  // eight of a round are in flight together; the segment of a flat index is found by walking
  // forward from the previous one (indices grow by 1024, segments are about that long).
  for (uint32_t f0 = 0; f0 < total; f0 += 8 * kBinThreads) {
    uint2 x[8];
    uint32_t a = 0;
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) {
      const uint32_t f = min(f0 + u * kBinThreads + threadIdx.x, total - 1u);
      while (seg_end[a] <= f) ++a;  // stops at the segment holding f (f < seg_end[nseg-1])
      const uint32_t start = a ? seg_end[a - 1] : 0u;
      x[u] = ev[(size_t)seg_buf[a] * cap + seg_lo[a] + (f - start)];
    }
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) {
      const uint32_t f = f0 + u * kBinThreads + threadIdx.x;
      if (f < total)
        atomicAdd((unsigned long long*)&acc[x[u].x],
                  (unsigned long long)quantize(__uint_as_float(x[u].y), p.scale_in));
    }
  }
  __syncthreads();
  // ring slot ta += tile accumulators, four entries per thread per round, loads first
  const uint32_t tn = 1u << p.bin_tshift;
  const uint32_t slot = (uint32_t)ta % p.D;
  long long* ring = s.ring + (size_t)slot * p.ch * p.n;
  for (uint32_t k0 = 0; k0 < TE; k0 += 4 * kBinThreads) {
    size_t idx[4];
    long long r[4], add[4];
    bool wr[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint32_t k = k0 + u * kBinThreads + threadIdx.x;
      const uint32_t c = k >> p.bin_tshift;
      const uint64_t j = (uint64_t)T * tn + (k & (tn - 1u));
      const bool in = c < p.ch && j < p.n;
      add[u] = acc[k];
      wr[u] = in && add[u] != 0;
      idx[u] = in ? (size_t)c * p.n + j : 0u;
      r[u] = ring[idx[u]];
    }
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u)
      if (wr[u]) ring[idx[u]] = r[u] + add[u];
  }
}

__global__ void k_advance(int32_t* tbase, int32_t g) { *tbase += g; }

// ---- launch helpers ------------------------------------------------------------------------
typedef void (*UpdateFn)(StepParams, DevPtrs, int32_t);

// The point models (LIF, Izhikevich) run four neurons per thread (k_update4) on large networks,
// where there are wavefronts enough for every SIMD either way; below kUpdate4Segments segments
// one neuron per thread keeps four times the wavefronts in flight (measured on the 1M network:
// 18 us one per thread, 20 us four per thread; on 10^8 neurons: 588 and 482 us).
constexpr uint32_t kUpdate4Segments = 8192;

constexpr bool point_model(int model) {
  return model == NSX_MODEL_LIF || model == NSX_MODEL_IZHIKEVICH;
}

static bool use_update4(const nsx_ctx* ctx) {
  if (!point_model((int)ctx->cfg.model)) return false;
  if (ctx->update_width) return ctx->update_width == 4;
  return ctx->p.nblk >= kUpdate4Segments;
}

template <int MODEL, int DRIVE, int PL>
static UpdateFn update_kernel(bool four) {
  if constexpr (point_model(MODEL)) {
    if (four) return k_update4<MODEL, DRIVE, PL>;
  }
  return k_update<MODEL, DRIVE, PL>;
}

template <int MODEL, int DRIVE>
static UpdateFn pick_plast(int pl, bool four) {
  switch (pl) {
    case NSX_PLAST_PAIR: return update_kernel<MODEL, DRIVE, NSX_PLAST_PAIR>(four);
    case NSX_PLAST_TRIPLET: return update_kernel<MODEL, DRIVE, NSX_PLAST_TRIPLET>(four);
    default: return update_kernel<MODEL, DRIVE, NSX_PLAST_NONE>(four);
  }
}
template <int MODEL>
static UpdateFn pick_drive(int drive, int pl, bool four) {
  switch (drive) {
    case NSX_DRIVE_POISSON: return pick_plast<MODEL, NSX_DRIVE_POISSON>(pl, four);
    case NSX_DRIVE_GAUSSIAN: return pick_plast<MODEL, NSX_DRIVE_GAUSSIAN>(pl, four);
    case NSX_DRIVE_OU: return pick_plast<MODEL, NSX_DRIVE_OU>(pl, four);
    default: return pick_plast<MODEL, NSX_DRIVE_NONE>(pl, four);
  }
}
UpdateFn pick_update(int model, int drive, int pl, bool four) {
  switch (model) {
    case NSX_MODEL_LIF: return pick_drive<NSX_MODEL_LIF>(drive, pl, four);
    case NSX_MODEL_IZHIKEVICH: return pick_drive<NSX_MODEL_IZHIKEVICH>(drive, pl, four);
    default: return pick_drive<NSX_MODEL_HH>(drive, pl, four);
  }
}

// One wavefront per work item when every block spikes, capped: sparse steps leave most waves
// idle and the cap bounds their cost.
uint32_t syn_grid(const nsx_ctx* ctx) {
  const bool plast = ctx->cfg.plasticity != NSX_PLAST_NONE;
  const uint64_t waves = std::min<uint64_t>(
      (uint64_t)ctx->p.nblk * (plast ? 2 : 1) * ctx->p.prop_split, 65536);
  const uint32_t wpb = ctx->syn_block / 64;
  return (uint32_t)std::max<uint64_t>((waves + wpb - 1) / wpb, 1);
}

typedef void (*SynFn)(StepParams, DevPtrs, int32_t);

template <int PL, bool PIECES>
static SynFn syn_kernel_nt(uint32_t nt) {
  switch (nt) {
    case 64: return k_synapses<PL, 64, PIECES>;
    case 128: return k_synapses<PL, 128, PIECES>;
    case 512: return k_synapses<PL, 512, PIECES>;
    default: return k_synapses<PL, 256, PIECES>;
  }
}

template <bool PIECES>
static SynFn syn_kernel_pl(int pl, uint32_t nt) {
  switch (pl) {
    case NSX_PLAST_PAIR: return syn_kernel_nt<NSX_PLAST_PAIR, PIECES>(nt);
    case NSX_PLAST_TRIPLET: return syn_kernel_nt<NSX_PLAST_TRIPLET, PIECES>(nt);
    default: return syn_kernel_nt<NSX_PLAST_NONE, PIECES>(nt);
  }
}

static SynFn syn_kernel(int pl, uint32_t nt, bool pieces) {
  return pieces ? syn_kernel_pl<true>(pl, nt) : syn_kernel_pl<false>(pl, nt);
}

typedef void (*PersistFn)(StepParams, DevPtrs, int32_t, uint32_t, uint32_t*);

template <int MODEL, int DRIVE>
static PersistFn persist_plast(int pl) {
  switch (pl) {
    case NSX_PLAST_PAIR: return k_persist<MODEL, DRIVE, NSX_PLAST_PAIR>;
    case NSX_PLAST_TRIPLET: return k_persist<MODEL, DRIVE, NSX_PLAST_TRIPLET>;
    default: return k_persist<MODEL, DRIVE, NSX_PLAST_NONE>;
  }
}
template <int MODEL>
static PersistFn persist_drive(int drive, int pl) {
  switch (drive) {
    case NSX_DRIVE_POISSON: return persist_plast<MODEL, NSX_DRIVE_POISSON>(pl);
    case NSX_DRIVE_GAUSSIAN: return persist_plast<MODEL, NSX_DRIVE_GAUSSIAN>(pl);
    case NSX_DRIVE_OU: return persist_plast<MODEL, NSX_DRIVE_OU>(pl);
    default: return persist_plast<MODEL, NSX_DRIVE_NONE>(pl);
  }
}
static PersistFn pick_persist(int model, int drive, int pl) {
  switch (model) {
    case NSX_MODEL_LIF: return persist_drive<NSX_MODEL_LIF>(drive, pl);
    case NSX_MODEL_IZHIKEVICH: return persist_drive<NSX_MODEL_IZHIKEVICH>(drive, pl);
    default: return persist_drive<NSX_MODEL_HH>(drive, pl);
  }
}

// Whether the loaded network can step persistently: a plain update model (not the surrogate,
// not the GpuBackend mode), ring-atomic delivery, one update launch per step (no epochs).
bool persist_applies(const nsx_ctx* ctx) {
  return ctx->persistent && ctx->cfg.model != NSX_MODEL_HH_SURROGATE && !ctx->backend_mode &&
         ctx->delivery == NSX_DELIVERY_ATOMIC && !ctx->p.bin_on && !ctx->p.pull_on &&
         epoch_len(ctx) == 1;
}

// One cooperative launch of `len` steps from offset `off`; the grid is kPersistPerCu
// workgroups per CU, fewer if the runtime's occupancy query reports fewer (the launch fails
// when the grid cannot be resident at once, rather than deadlock).
hipError_t launch_persist(nsx_ctx* ctx, int32_t off, uint32_t len) {
  const PersistFn f = pick_persist((int)ctx->cfg.model, (int)ctx->cfg.drive,
                                   (int)ctx->cfg.plasticity);
  int per_cu = 0, cus = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)f, kBlock, 0);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
  if (e != hipSuccess) return e;
  const uint32_t grid = (uint32_t)std::min(std::max(per_cu, 1), (int)kPersistPerCu) *
                       (uint32_t)std::max(cus, 1);
  hipLaunchKernelGGL(k_zero_bar, dim3(1), dim3(64), 0, ctx->stream, ctx->d_bar);
  StepParams p = ctx->p;
  DevPtrs d = ctx->d;
  uint32_t* bar = ctx->d_bar;
  void* args[] = {&p, &d, &off, &len, &bar};
  return hipLaunchCooperativeKernel((const void*)f, dim3(grid), dim3(kBlock), args, 0,
                                    ctx->stream);
}

// Arrival of the bucketed events at step *tbase + off (emission steps up to emax).
hipError_t launch_arrive(nsx_ctx* ctx, int32_t off, int32_t emax) {
  const StepParams& p = ctx->p;
  const dim3 g(p.bin_ntiles), b(kBinThreads);
  switch (ctx->bin_te) {
    case 4096: hipLaunchKernelGGL(k_bin_arrive<4096>, g, b, 0, ctx->stream, p, ctx->d, off, emax); break;
    case 8192: hipLaunchKernelGGL(k_bin_arrive<8192>, g, b, 0, ctx->stream, p, ctx->d, off, emax); break;
    default: hipLaunchKernelGGL(k_bin_arrive<16384>, g, b, 0, ctx->stream, p, ctx->d, off, emax); break;
  }
  return hipGetLastError();
}

// Enqueue the launches of one step with offset `off` from *tbase.  Profiling events: binned
// delivery [0] update [1] synapse pass [2] bucket passes [3] arrival [4]; otherwise [0] update
// [1] [2] synapse pass (or pull) [3] [4].
hipError_t enqueue_step(nsx_ctx* ctx, int32_t off, hipEvent_t* evs) {
  const StepParams& p = ctx->p;
  const int pl = (int)ctx->cfg.plasticity;
  const bool four = use_update4(ctx);
  UpdateFn upd = pick_update((int)ctx->cfg.model, (int)ctx->cfg.drive, pl, four);
  // one thread per neuron, or per four neurons (a wavefront per segment)
  const uint32_t gu = four ? (p.nblk + kBlock / 64 - 1) / (kBlock / 64)
                           : (uint32_t)((ctx->n + kBlock - 1) / kBlock);
  const uint32_t gs = syn_grid(ctx);
  if (evs) (void)hipEventRecord(evs[0], ctx->stream);
  if (ctx->cfg.model == NSX_MODEL_HH_SURROGATE) {
    const hipError_t e = launch_update_surrogate(ctx, off);
    if (e != hipSuccess) return e;
  } else {
    hipLaunchKernelGGL(upd, dim3(gu), dim3(kBlock), 0, ctx->stream, p, ctx->d, off);
  }
  if (evs) (void)hipEventRecord(evs[1], ctx->stream);
  if (p.pull_on) {
    if (evs) (void)hipEventRecord(evs[2], ctx->stream);
    const hipError_t e = launch_pull(ctx, off);
    if (e != hipSuccess) return e;
    if (evs) (void)hipEventRecord(evs[3], ctx->stream);
  } else if (p.bin_on) {
    // events: update | synapse pass | count scan and placement | arrival (kinds 0, 1, 2, 3)
    const dim3 gb(p.bin_groups), bb(kBinThreads);
    if (pl == NSX_PLAST_PAIR)
      hipLaunchKernelGGL(k_synapses_bin<NSX_PLAST_PAIR>, gb, bb, 0, ctx->stream, p, ctx->d, off);
    else if (pl == NSX_PLAST_TRIPLET)
      hipLaunchKernelGGL(k_synapses_bin<NSX_PLAST_TRIPLET>, gb, bb, 0, ctx->stream, p, ctx->d,
                         off);
    else
      hipLaunchKernelGGL(k_synapses_bin<NSX_PLAST_NONE>, gb, bb, 0, ctx->stream, p, ctx->d, off);
    if (evs) (void)hipEventRecord(evs[2], ctx->stream);
    hipLaunchKernelGGL(k_bin_scan, dim3((p.bin_nbins + kScanCols - 1) / kScanCols),
                       dim3(kScanThreads), 0, ctx->stream, p, ctx->d);
    if (p.bin_place_ch > kPlacePer2 * kBinThreads)
      hipLaunchKernelGGL(k_bin_place<kPlacePer>, gb, bb, ctx->bin_place_lds, ctx->stream, p,
                         ctx->d, off);
    else
      hipLaunchKernelGGL(k_bin_place<kPlacePer2>, gb, bb, ctx->bin_place_lds, ctx->stream, p,
                         ctx->d, off);
    if (evs) (void)hipEventRecord(evs[3], ctx->stream);
    const hipError_t e = launch_arrive(ctx, off + 1, INT32_MAX);
    if (e != hipSuccess) return e;
  } else {
    if (evs) (void)hipEventRecord(evs[2], ctx->stream);
    hipLaunchKernelGGL(syn_kernel(pl, ctx->syn_block, p.row_split != 0), dim3(gs),
                       dim3(ctx->syn_block), 0, ctx->stream, p, ctx->d, off);
    if (evs) (void)hipEventRecord(evs[3], ctx->stream);
  }
  if (evs) (void)hipEventRecord(evs[4], ctx->stream);
  return hipGetLastError();
}

typedef void (*EpochFn)(StepParams, DevPtrs, EpochSeg, int32_t, uint32_t);
typedef void (*EpochSynFn)(StepParams, DevPtrs, EpochSeg, int32_t, uint32_t);

template <int MODEL>
static EpochFn epoch_pick_drive(int drive) {
  switch (drive) {
    case NSX_DRIVE_POISSON: return k_epoch_update<MODEL, NSX_DRIVE_POISSON>;
    case NSX_DRIVE_GAUSSIAN: return k_epoch_update<MODEL, NSX_DRIVE_GAUSSIAN>;
    case NSX_DRIVE_OU: return k_epoch_update<MODEL, NSX_DRIVE_OU>;
    default: return k_epoch_update<MODEL, NSX_DRIVE_NONE>;
  }
}

static EpochFn pick_epoch(int model, int drive) {
  switch (model) {
    case NSX_MODEL_LIF: return epoch_pick_drive<NSX_MODEL_LIF>(drive);
    case NSX_MODEL_IZHIKEVICH: return epoch_pick_drive<NSX_MODEL_IZHIKEVICH>(drive);
    default: return epoch_pick_drive<NSX_MODEL_HH>(drive);
  }
}

uint32_t epoch_len(const nsx_ctx* ctx) {
  if (ctx->epoch_cap < 2 || ctx->cfg.plasticity != NSX_PLAST_NONE || ctx->backend_mode ||
      ctx->p.bin_on || ctx->delivery != NSX_DELIVERY_ATOMIC || ctx->d_min < 2)
    return 1;
  return std::min(ctx->d_min, ctx->epoch_cap);
}

nsx_status epoch_prepare(nsx_ctx* ctx) {
  if (ctx->has_network && ctx->d_min == 0) {
    nsx_status s = network_min_delay(ctx, &ctx->d_min);
    if (s) return s;
  }
  const uint32_t E = epoch_len(ctx);
  const uint32_t atomic = E > 1 ? 1u : 0u;
  if (atomic != ctx->p.part_atomic) {
    ctx->p.part_atomic = atomic;
    destroy_graph(ctx);  // captured launches hold the old parameters
  }
  if (E > 1 && ctx->epoch_alloc < E) {
    dev_free(ctx, ctx->ep.spikes);
    dev_free(ctx, ctx->ep.bcnt);
    ctx->ep = EpochSeg{nullptr, nullptr};
    ctx->epoch_alloc = 0;
    destroy_graph(ctx);
    nsx_status s;
    if ((s = dev_alloc(ctx, (void**)&ctx->ep.spikes, (size_t)E * ctx->p.nblk * kBlock * 4, false)) ||
        (s = dev_alloc(ctx, (void**)&ctx->ep.bcnt, (size_t)E * ctx->p.nblk * 4, true)))
      return s;
    ctx->epoch_alloc = E;
  }
  return NSX_OK;
}

hipError_t enqueue_epoch(nsx_ctx* ctx, int32_t off, uint32_t len, hipEvent_t* evs) {
  const StepParams& p = ctx->p;
  if (evs) (void)hipEventRecord(evs[0], ctx->stream);
  if (ctx->cfg.model == NSX_MODEL_HH_SURROGATE) {
    const hipError_t e = launch_epoch_surrogate(ctx, off, len);
    if (e != hipSuccess) return e;
  } else {
    hipLaunchKernelGGL(pick_epoch((int)ctx->cfg.model, (int)ctx->cfg.drive), dim3(p.nblk),
                       dim3(kBlock), 0, ctx->stream, p, ctx->d, ctx->ep, off, len);
  }
  if (evs) {
    (void)hipEventRecord(evs[1], ctx->stream);
    (void)hipEventRecord(evs[2], ctx->stream);
  }
  const uint64_t waves = (uint64_t)len * p.nblk * p.prop_split;
  const uint32_t gs = (uint32_t)std::min<uint64_t>((waves + kBlock / 64 - 1) / (kBlock / 64), 16384);
  EpochSynFn syn = p.row_split ? k_synapses_epoch<NSX_PLAST_NONE, true>
                               : k_synapses_epoch<NSX_PLAST_NONE, false>;
  hipLaunchKernelGGL(syn, dim3(std::max(gs, 1u)), dim3(kBlock), 0, ctx->stream, p, ctx->d, ctx->ep,
                     off, len);
  if (evs) {
    (void)hipEventRecord(evs[3], ctx->stream);
    (void)hipEventRecord(evs[4], ctx->stream);
  }
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

// ---- binned delivery: host side ------------------------------------------------------------

nsx_status bin_clear(nsx_ctx* ctx) {
  if (ctx->p.pull_on) return pull_clear(ctx);
  if (!ctx->p.bin_on) return NSX_OK;
  NSX_HIP(ctx, hipMemsetAsync(ctx->d.bin_tab, 0,
                              (size_t)ctx->D * (ctx->p.bin_nbins + 1) * 4, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NSX_OK;
}

// The arrival of step t (the next step) has run; the events of steps up to t - 1 that arrive
// at t + 1 .. t + D - 2 are still in the buffers.
nsx_status bin_materialize(nsx_ctx* ctx) {
  if (ctx->p.pull_on) return pull_materialize(ctx);
  if (!ctx->p.bin_on) return NSX_OK;
  const int64_t t = (int64_t)ctx->t;
  for (int64_t ta = t + 1; t > 0 && ta <= t + ctx->D - 2; ++ta) {
    const hipError_t e = launch_arrive(ctx, (int32_t)(ta - ctx->tbase_host), (int32_t)(t - 1));
    if (e != hipSuccess) return hip_error(ctx, e, "binned delivery arrival");
  }
  return bin_clear(ctx);
}

// Tile shift, tile count and bucket count of binned delivery with te-entry tiles; false when
// the network does not fit (buckets, ring slots, 32-bit event counts).
static bool bin_shape(const nsx_ctx* ctx, uint32_t te, uint32_t& tshift, uint64_t& ntiles,
                      uint64_t& nbins) {
  tshift = (uint32_t)__builtin_ctz(te) - (ctx->ch == 2 ? 1u : 0u);
  ntiles = (ctx->n + (1ull << tshift) - 1) >> tshift;
  nbins = (uint64_t)(ctx->D - 1) * ntiles;
  return nbins <= kMaxBins && (uint32_t)ctx->D - 1 <= kMaxArriveSegs && ctx->m < (1ull << 32);
}

bool bin_supported(const nsx_ctx* ctx, uint32_t te) {
  uint32_t tshift;
  uint64_t ntiles, nbins;
  return ctx->has_network && bin_shape(ctx, te, tshift, ntiles, nbins);
}

static void bin_release(nsx_ctx* ctx) {
  dev_free(ctx, ctx->d.bin_ev);
  dev_free(ctx, ctx->d.bin_tab);
  dev_free(ctx, ctx->d.bin_hist);
  dev_free(ctx, ctx->d.bin_tot);
  dev_free(ctx, ctx->d.bin_scr);
  dev_free(ctx, ctx->d.bin_reg);
  ctx->d.bin_ev = ctx->d.bin_scr = ctx->d.bin_reg = nullptr;
  ctx->d.bin_tab = ctx->d.bin_hist = ctx->d.bin_tot = nullptr;
  ctx->p.bin_on = 0;
  ctx->bin_te = 0;
  ctx->p.pull_on = 0;  // the pull tables stay built for the network (pull_release frees them)
}

nsx_status bin_configure(nsx_ctx* ctx, uint32_t mode, uint64_t cap, uint32_t te) {
  if (mode != NSX_DELIVERY_ATOMIC && mode != NSX_DELIVERY_BINNED && mode != NSX_DELIVERY_PULL)
    return set_error(ctx, NSX_E_INVALID, "delivery mode must be ATOMIC, BINNED or PULL");
  if (te != 0 && te != 4096 && te != 8192 && te != 16384)
    return set_error(ctx, NSX_E_INVALID, "tile_entries must be 0, 4096, 8192 or 16384");
  nsx_status s = bin_materialize(ctx);  // pending events move into the ring
  if (s) return s;
  bin_release(ctx);
  destroy_graph(ctx);
  ctx->delivery = mode;
  ctx->bin_cap_req = cap;
  ctx->bin_te_req = te;
  ctx->bin_dirty = false;
  if (mode == NSX_DELIVERY_ATOMIC) return NSX_OK;
  if (!ctx->has_network) {  // sized when a network is loaded
    ctx->bin_dirty = true;
    return NSX_OK;
  }
  if (mode == NSX_DELIVERY_PULL) return pull_configure(ctx);
  const uint32_t D = (uint32_t)ctx->D;
  const uint64_t m = ctx->m;
  if (m >= (1ull << 32)) return set_error(ctx, NSX_E_INVALID, "binned delivery needs M < 2^32");
  // automatic tile: the largest whose tiles still number at least 96 (arrival workgroups; on
  // the 1M-neuron network 8192 entries, 122 tiles, measured 133.8 us/step against 135.9 with
  // 16384 (61 tiles) and 141.7 with 4096 (245 tiles, shorter bucket runs))
  if (te == 0) {
    te = 4096;
    for (uint32_t c : {16384u, 8192u})
      if ((ctx->n * ctx->ch + c - 1) / c >= 96) {
        te = c;
        break;
      }
  }
  uint32_t tshift = 0;
  uint64_t ntiles = 0, nbins = 0;
  if (!bin_shape(ctx, te, tshift, ntiles, nbins))
    return set_error(ctx, NSX_E_INVALID, "binned delivery needs (D - 1) * tiles <= 8192");
  int cus = 0, per_cu = 0;
  NSX_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const void* kfn = ctx->cfg.plasticity == NSX_PLAST_PAIR
                        ? (const void*)k_synapses_bin<NSX_PLAST_PAIR>
                        : ctx->cfg.plasticity == NSX_PLAST_TRIPLET
                              ? (const void*)k_synapses_bin<NSX_PLAST_TRIPLET>
                              : (const void*)k_synapses_bin<NSX_PLAST_NONE>;
  NSX_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kBinThreads, 0));
  const uint32_t groups = (uint32_t)std::min(std::max(cus * std::max(per_cu, 1), 1), 1024);
  if (cap == 0) {
    // room for bursts (each synapse workgroup's region holds cap / groups records of a step;
    // its events past that take the ring atomics): an eighth of all synapses in one step, at
    // least 2^24 events, the D step buffers within 4 GB (the 256-network population bursts to
    // 27 M events in a step, 4.7 times its mean)
    const uint64_t mem_cap = (1ull << 32) / (8ull * D);
    cap = std::min<uint64_t>({m, mem_cap, std::max<uint64_t>(1ull << 24, m / 8)});
    cap = std::max<uint64_t>(cap, std::min<uint64_t>(m, 1ull << 16));
  }
  cap = std::max<uint64_t>(std::min<uint64_t>(cap, 0xFFFFFFFFull), 1);
  DevPtrs& d = ctx->d;
  if ((s = dev_alloc(ctx, (void**)&d.bin_ev, (size_t)D * cap * 8, false)) ||
      (s = dev_alloc(ctx, (void**)&d.bin_tab, (size_t)D * (nbins + 1) * 4, true)) ||
      (s = dev_alloc(ctx, (void**)&d.bin_hist, (size_t)groups * nbins * 4, true)) ||
      (s = dev_alloc(ctx, (void**)&d.bin_tot, (size_t)nbins * 4, true)) ||
      (s = dev_alloc(ctx, (void**)&d.bin_scr, (size_t)cap * 8, false)) ||
      (s = dev_alloc(ctx, (void**)&d.bin_reg, (size_t)groups * 8, true))) {
    bin_release(ctx);
    return s;
  }
  // placement chunks: all groups resident at once, so one workgroup per CU takes the rest of
  // the CU's 160 KB of LDS beside gpos and cst (up to kPlacePer records per thread; a region's
  // records of a step usually fit one chunk), and with more groups than CUs two share a CU
  // (80 KB each, up to kPlacePer2 records per thread)
  const bool two = groups > (uint32_t)cus;
  const uint64_t room = (two ? (80u << 10) : (160u << 10)) - 8ull * nbins - 128u;
  const uint32_t ch = (uint32_t)std::min<uint64_t>(
      (two ? kPlacePer2 : kPlacePer) * kBinThreads,
      std::max<uint64_t>(room / 8 / kBinThreads, 1) * kBinThreads);
  const uint32_t place_lds = (uint32_t)(8ull * nbins + 8ull * ch);
  {
    const void* kp = ch > kPlacePer2 * kBinThreads ? (const void*)k_bin_place<kPlacePer>
                                                   : (const void*)k_bin_place<kPlacePer2>;
    const hipError_t e =
        hipFuncSetAttribute(kp, hipFuncAttributeMaxDynamicSharedMemorySize, (int)place_lds);
    if (e != hipSuccess) {
      bin_release(ctx);
      return hip_error(ctx, e, "binned delivery placement LDS");
    }
  }
  ctx->bin_place_lds = place_lds;
  StepParams& p = ctx->p;
  p.bin_tshift = tshift;
  p.bin_ntiles = (uint32_t)ntiles;
  p.bin_nbins = (uint32_t)nbins;
  p.bin_groups = groups;
  p.bin_cap = (uint32_t)cap;
  // a workgroup's events of a step past its region (cap / groups records) take the ring atomics
  p.bin_gcap = (uint32_t)(cap / groups);
  p.bin_place_ch = ch;
  p.bin_on = 1;
  ctx->bin_te = te;
  return NSX_OK;
}

}  // namespace nsx

using namespace nsx;

extern "C" {

nsx_status nsx_set_delivery(nsx_ctx* ctx, uint32_t mode, uint64_t cap_events,
                            uint32_t tile_entries) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  nsx_status s = bin_configure(ctx, mode, cap_events, tile_entries);
  if (s) {
    ctx->delivery = NSX_DELIVERY_ATOMIC;
    ctx->bin_dirty = false;
  }
  return s;
}

nsx_status nsx_get_delivery(nsx_ctx* ctx, uint32_t* mode) {
  if (!ctx || !mode) return set_error(ctx, NSX_E_INVALID, "NULL argument");
  *mode = ctx->delivery;
  return NSX_OK;
}

}  // extern "C"
