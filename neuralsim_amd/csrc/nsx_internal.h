// nsx_internal.h -- host-side context of the simulation core and kernel parameter blocks.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/nsx.h"

namespace nsx {

constexpr int kMaxPoisson = 32;
constexpr uint32_t kBlock = 256;

// Binned delivery (nsx_set_delivery): the synapse pass writes each event into a bucket of
// (delay, target tile) instead of a ring atomic; a tile's buckets are summed in LDS at arrival.
constexpr uint32_t kBinThreads = 1024;    // workgroup of the binning / synapse kernels
constexpr uint32_t kTileEntries = 16384;  // int64 LDS accumulators of one tile (128 KB)
constexpr uint32_t kMaxBins = 8192;       // (D - 1) * tiles
constexpr uint32_t kMaxArriveSegs = 256;  // D - 1 (ring_slots is at most 256)

// Lazily decayed STDP trace of one neuron: value right after its last bump and that step, and
// the same pair before that bump (pair STDP: the trace between the last two spikes).
struct __attribute__((aligned(16))) TraceRec {
  float v;
  int32_t s;
  float v_prev;
  int32_t s_prev;
};

// Pair STDP defers the depression of a post spike: up to kPostPend post spikes of a neuron are
// kept in its record, and each is applied to an in-synapse either when the synapse's source
// next fires (the propagation pass reads the row's weights anyway) or when the list fills, by
// one walk over the neuron's in-synapses for all of them.  Every depression is applied once, in
// time order with the synapse's potentiations, so the weights equal the per-spike updates.
// Four entries keep the record at 32 bytes: measured on the 1M-neuron network, 12 entries in a
// 64-byte record walked the in-synapses a third as often but made every propagation gather
// twice as wide and ran 6 % slower; 2 entries in a 16-byte record (the newest pending spike is
// always `last`, so one step and a count beside it) halved the gathers' working set but walked
// twice as often and ran 5 % slower (141.8 against 134.5 us per step).
constexpr uint32_t kPostPend = 4;

// Post-synaptic record, 32 bytes: the lazily decayed trace {v, s}, the step of the neuron's last
// spike (same-step post spikes seen by the potentiation pass), and the pending post spikes.  The
// trace bump of a spike at step t is applied by the update kernel of step t+1
// (after every pass of step t has read the pre-bump value), which also empties a list that the
// synapse pass of step t flushed.  Triplet STDP keeps no trace: s holds the step of the
// neuron's spike before `last`.
struct __attribute__((aligned(32))) PostRec {
  float v;
  int32_t s;
  int32_t last;
  uint32_t npend;
  int32_t pend[kPostPend];
};
static_assert(sizeof(PostRec) == 32 && sizeof(TraceRec) == 16, "record layouts");

// Hodgkin-Huxley constants of one Euler sub-step.
struct HHConst {
  float g_na, g_k, g_l, e_na, e_k, e_l;
  float dts;         // sub-step length (ms)
  float kv;          // dts / C_m
  uint32_t literal;  // 1: the reference code's current signs (hodgkin_huxley.rs:78-84)
};

// One network of a batched population (device table).  Networks are concatenated, each starting
// at a multiple of kBlock, so every neuron block belongs to exactly one network.
struct BatchNet {
  uint64_t seed;       // the network's own RNG key (connectivity, parameters, drive)
  uint64_t syn_off;    // first synapse of this network in the generated list
  uint32_t offset;     // first global neuron
  uint32_t n, n_exc;   // neurons, excitatory neurons (local indices [0, n_exc))
  uint32_t k_in, d_lo, d_span;
  float w_exc_lo, w_exc_hi, w_inh_lo, w_inh_hi;
};

// Everything a step kernel reads that is constant over a run (passed by value).
struct StepParams {
  uint32_t n, n_exc, D, ch;
  uint32_t nblk;        // neuron blocks of kBlock neurons (spike segments)
  uint32_t prop_split;  // wavefronts sharing one block's spiking rows (propagation / LTD)
  uint64_t seed;
  double scale_in, scale_out;  // 2^S, 2^-S
  // LIF
  float v_rest, v_reset, v_th, lif_r, lif_k;
  int32_t refr_steps;
  // synaptic decay factors exp(-dt/tau)
  float dec_e, dec_i, dec_n;
  // Izhikevich
  float izh_a[2], izh_b[2];
  float izh_h, dtf;
  uint32_t substeps;
  // HH
  HHConst hh;
  uint32_t hh_substeps;  // Euler sub-steps per step, input current held
  float hh_spike_v;
  float e_exc, e_inh, gmax_ampa, gmax_nmda, gmax_gaba, mgc;
  uint32_t hh_reset_mode, use_nmda;
  // drive
  float drive_mean, drive_sigma, drive_weight, ou_k1, ou_k2, ou_k3;
  uint32_t pthr[kMaxPoisson];
  int32_t npthr;
  // STDP (pair: traces; triplet: nearest-spike intervals through the same tables)
  float a_plus, a_minus, w_min, w_max, a3;
  int32_t ptab_n, mtab_n, ytab_n;
  uint32_t step_counts_cap;
  uint32_t raster_on;
  uint64_t raster_cap;
  uint32_t batch;  // 1: neurons belong to the networks of a batched population
  uint32_t n_nets;
  uint32_t stats_on;  // 1: per-step spike counts and the LFP partials are kept (nsx_set_stats)
  uint32_t stat_cap;  // steps of the stat_counts ring
  // binned delivery (bin_on = 1): tiles of 2^bin_tshift neurons, bucket (d - 1) * ntiles + tile
  uint32_t bin_on, bin_tshift, bin_ntiles, bin_nbins, bin_groups;
  uint32_t bin_cap;  // events one step's bucket buffer holds; more fall back to ring atomics
  uint32_t bin_place_ch;  // records k_bin_place sorts per LDS chunk (a multiple of kBinThreads)
  uint32_t bin_gcap;      // records of each synapse workgroup's scratch region per step
  // epoch stepping: several steps per update launch, one synapse pass for all of them; the
  // per-block counters are then added atomically (items of several steps share a block)
  uint32_t part_atomic;
  // propagation with atomic delivery: rows longer than row_split slots are cut into pieces of at
  // most row_split slots, dealt round-robin to the block's prop_split wavefronts (0: whole rows)
  uint32_t row_split;
  // pull delivery (networks without plasticity): the input of step t + 1 is gathered per target
  // tile from the spikes of steps t + 2 - D .. t instead of pushed by a synapse pass
  uint32_t pull_on, pull_ntiles;
};

// Pull delivery.  Each network's targets are cut into `parts` ranges of at most L neurons (L x
// channels <= kPullTE LDS accumulators); every row keeps, besides the CSR, a copy of its synapses
// sorted by (part of the target, delay) -- the key part * (D - 1) + d - 1 -- with the start of
// each key's slice (entries [0, keys)) and the degree (entry keys) in a u16 offset table of
// `keys + 1` entries per row.  A tile is one part of one network.
constexpr uint32_t kPullTE = 16384;    // LDS int64 accumulators of a tile (channels x neurons)
constexpr uint32_t kPullMaxKeys = 256;  // (part, delay) keys per row at most
struct PullNet {
  uint64_t doff_base;  // offset-table entry of the network's first row
  uint32_t offset, n;  // first global neuron, neurons
  uint32_t L, parts, stride;  // part length, parts, offset-table entries per row (keys + 1)
  uint32_t _pad;
};
struct PullTile {
  uint64_t doff_base;  // the network's doff_base
  uint32_t lo, hi;     // target neurons [lo, hi) (global)
  uint32_t slo, shi;   // source neurons: the network, [slo, shi)
  uint32_t key0, stride;  // part * (D - 1); offset-table entries per row
  uint32_t L, _pad;
};

// Epoch stepping (networks without plasticity whose shortest delay is d_min >= 2 steps): no spike
// of steps t .. t + d_min - 1 arrives before t + d_min, so those steps' updates need no
// propagation between them.  One update launch advances all of them (state in registers, each
// step's spikes into its own segment), then one synapse pass propagates every step's spikes.
constexpr uint32_t kMaxEpoch = 128;  // steps per epoch at most (default cap)

// Per-step spike segments of an epoch: step k's block b at spikes[(k * nblk + b) * kBlock],
// its count at bcnt[k * nblk + b].
struct EpochSeg {
  uint32_t* spikes;
  uint32_t* bcnt;
};

// per-step spike counts are added into this many stripes (one atomic per block and step)
constexpr uint32_t kStatStripes = 16;
// latest step with a spike, kept in kEvStripes stripes kEvStride ints apart (one 128-byte line
// each): a ring slot can hold input only if some neuron spiked in the D - 1 steps before it
constexpr uint32_t kEvStripes = 16, kEvStride = 32;

// Device pointers used by the step kernels.
struct DevPtrs {
  float *v, *u, *gm, *gh, *gn, *s0, *s1, *s2, *drv, *c, *d;
  int32_t *refr, *last;
  uint32_t* count;
  TraceRec* pre;
  PostRec* post;
  long long* ring;
  const float *ptab, *mtab, *ytab;  // exp(-k dt / tau) for tau+, tau- and (triplet) tau_y
  uint32_t* spikes;  // per-block spike segments: block b's spikes at [b*kBlock, +bcnt[b])
  uint32_t* bcnt;    // spikes of each neuron block in the current step
  unsigned long long* part;  // per-block running totals, kPart* planes of nblk entries
  uint32_t* nspk;    // GpuBackend drop-in spike list length
  uint32_t *raster_step, *raster_id;
  unsigned long long* counters;  // see kCnt*
  uint32_t* step_counts;
  uint32_t* stat_counts;  // [stat_cap][kStatStripes] spikes per step (stats on)
  double* lfp_part;       // per-block running sum of |synaptic currents| (stats on)
  int32_t* ev_last;       // [kEvStripes * kEvStride] latest spike step per block stripe
  int32_t* tbase;
  // network
  const uint64_t* row_ptr;
  const uint32_t* col;
  float* w;
  const uint8_t* delay;
  const uint64_t* csc_ptr;
  const uint64_t* csc_slot;
  const uint32_t* csc_src;
  // batched population
  const uint32_t* blk_net;  // network of each neuron block
  const BatchNet* bnet;
  // binned delivery
  uint2* bin_ev;       // [D][bin_cap] events {LDS entry, weight bits} of the step t mod D
  uint32_t* bin_tab;   // [D][nbins + 1] bucket starts in that step's buffer (all 0: none)
  uint32_t* bin_hist;  // [groups][nbins] events per group and bucket, then group offsets
  uint32_t* bin_tot;   // [nbins] events per bucket
  uint2* bin_scr;      // [bin_cap] the step's records in synapse-pass order, one region per group
  uint2* bin_reg;      // [groups] each group's region of bin_scr: {start, records}
  // pull delivery
  const uint2* pdsyn;     // synapses re-sorted per row by (target part, delay): {col, w bits}
  const uint16_t* pdoff;  // per row: start of each (part, delay) key's slice, then the degree
  const PullTile* ptile;  // the pull kernel's tiles
  uint32_t* phist_spk;    // [D][nblk * kBlock] spike segments of step e at slot e mod D
  uint32_t* phist_row;    // [D][nblk * kBlock] the row start of each of those spikes
  uint32_t* phist_cnt;    // [D][nblk] their counts
  int32_t* pull_t0;       // first step whose spikes' input the ring does not hold
};

// Identity of a neuron for its RNG streams and E/I type: the network's seed and local index in
// a batched population, the context's seed and the global index otherwise.
struct NeuronId {
  uint64_t seed;
  uint32_t local, n, n_exc;
};

// planes of DevPtrs::part (each nblk entries; entry b written only by the wave owning block b)
enum : int { kPartSpikes = 0, kPartEvents = 1, kPartLtp = 2, kPartLtd = 3, kNumPart = 4 };

// counters[]
enum : int {
  kCntSpikes = 0,
  kCntEvents = 1,
  kCntLtdSlots = 2,
  kCntLtpSlots = 3,
  kCntRaster = 4,
  kCntOverflow = 5,
  kCntExtraInput = 6,
  kNumCounters = 8
};

// HH surrogate of a context (device buffers owned through nsx_ctx::allocs).
struct Surrogate {
  float* master = nullptr;  // FP32 parameters
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  void* image = nullptr;    // bf16 fragments + FP32 layer 1 / biases as the kernels read them
  float* norm = nullptr;    // normalisation and envelope
  float* loss = nullptr;
  float* data = nullptr;    // training transitions, 9 floats each
  uint64_t n_data = 0;
  float* gpart = nullptr;   // per-wavefront partial gradients
  float* lpart = nullptr;
  uint32_t gpart_waves = 0;
  unsigned long long* fb = nullptr;  // per-block fallback counts (surrogate model)
  uint64_t adam_t = 0;
  bool has_params = false, has_norm = false;
};

}  // namespace nsx

struct nsx_ctx {
  nsx_config cfg;
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t n = 0, m = 0;
  int ch = 1, D = 2;
  bool has_network = false;
  uint64_t t = 0;          // next step index
  int32_t tbase_host = 0;  // value held in d.tbase
  nsx::StepParams p{};
  nsx::DevPtrs d{};
  // owned device allocations
  std::vector<void*> allocs;
  uint64_t* row_ptr = nullptr;
  uint32_t* col = nullptr;
  float* w = nullptr;
  uint8_t* delay = nullptr;
  uint64_t* csc_ptr = nullptr;
  uint64_t* csc_slot = nullptr;
  uint32_t* csc_src = nullptr;
  uint64_t csc_m = 0;
  std::vector<float> ptab_h, mtab_h, ytab_h;  // host copies of the trace decay tables
  // graph
  uint32_t graph_steps = 0;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  // profiling
  bool profiling = false;
  double kt_ms[4] = {0, 0, 0, 0};
  uint64_t kt_launches[4] = {0, 0, 0, 0};
  std::vector<hipEvent_t> ev_pool;  // direct launches: 5 events per step
  hipEvent_t ev_call[2] = {nullptr, nullptr};  // bracket every nsx_step call (device time)
  uint32_t last_call_steps = 0;
  std::vector<unsigned long long> part_host;
  uint64_t part_delta[4] = {0, 0, 0, 0};
  uint64_t last_call_first_step = 0;
  // GpuBackend drop-in state
  bool backend_mode = false;
  // persistent stepping: requested (nsx_set_persistent); the grid barrier's two words
  bool persistent = false;
  uint32_t* d_bar = nullptr;
  float* b_input = nullptr;     // uploaded input_current (f32)
  float* b_params = nullptr;    // 6 floats per neuron
  uint32_t* b_spiked = nullptr;
  uint32_t* b_row_ptr32 = nullptr;
  float b_dt = 0.5f, b_time = 0.0f;
  uint32_t b_model = 0;
  nsx::Surrogate sur;
  // batched population
  std::vector<nsx::BatchNet> nets_h;
  // binned delivery: requested mode, capacity override, buffers sized for (n, m, D)
  uint32_t delivery = NSX_DELIVERY_ATOMIC;
  uint64_t bin_cap_req = 0;
  uint32_t bin_te_req = 0;  // tile entries requested (0: automatic)
  uint32_t bin_te = 0;      // tile entries in use
  uint32_t bin_place_lds = 0;  // dynamic LDS bytes of k_bin_place
  uint32_t update_width = 0;  // neurons per thread of the point-model update (0: by size)
  uint32_t syn_block = 256;   // threads per workgroup of the (atomic-delivery) synapse pass
  // epoch stepping: steps per epoch at most (nsx_set_epoch; <= 1 off), the loaded network's
  // shortest delay in steps (0: not known yet), the epoch segments allocated for epoch_alloc steps
  uint32_t epoch_cap = nsx::kMaxEpoch;
  uint32_t d_min = 0;
  uint32_t epoch_alloc = 0;
  nsx::EpochSeg ep{nullptr, nullptr};
  bool bin_dirty = false;  // the network changed: buffers are re-sized before the next step
  bool pull_built = false;  // the pull copy and tables match the loaded network and weights
  // nsx_snapshot_save / restore: one device copy of the mutable buffers, with the step index
  void* snap = nullptr;
  size_t snap_bytes = 0;
  uint64_t snap_t = 0;
  int32_t snap_tb = 0;
  // SimulationStats kept on the host from the device's per-step counts (nsx_set_stats)
  struct {
    double rate = 0.0;         // low-pass mean firing rate (Hz)
    uint64_t active = 0;       // largest per-step spike count
    uint64_t last = 0;         // spikes of the last step
    uint64_t steps = 0;        // steps folded
  } hs;
  decltype(hs) snap_hs;
};

// helpers shared by translation units
namespace nsx {
nsx_status set_error(nsx_ctx* ctx, nsx_status s, const std::string& msg);
nsx_status hip_error(nsx_ctx* ctx, hipError_t e, const char* what);
nsx_status stats_fold(nsx_ctx* ctx, uint64_t first_step, uint64_t n_steps);
void destroy_graph(nsx_ctx* ctx);
// Pair STDP: apply (apply = true) or drop every pending deferred depression and empty the lists.
nsx_status flush_pending(nsx_ctx* ctx, bool apply);
nsx_status dev_alloc(nsx_ctx* ctx, void** p, size_t bytes, bool zero);
void dev_free(nsx_ctx* ctx, void* p);
nsx_status build_csr_device(nsx_ctx* ctx, uint32_t* d_src, uint32_t* d_dst, float* d_w,
                            uint8_t* d_delay, uint64_t m);
nsx_status build_csc_device(nsx_ctx* ctx);
nsx_status validate_csr_device(nsx_ctx* ctx);
hipError_t launch_update_surrogate(nsx_ctx* ctx, int32_t off);
// Binned delivery: every event still in a bucket buffer added into its ring slot, the buffers
// marked empty (the ring then holds all pending input, as with atomic delivery).
nsx_status bin_materialize(nsx_ctx* ctx);
// Mark every bucket buffer empty (after the ring was overwritten).
nsx_status bin_clear(nsx_ctx* ctx);
// Mark the ring as possibly holding input for the next D - 1 steps (after it was written from
// the host), or as empty (never = true, after it was cleared).
nsx_status ring_mark(nsx_ctx* ctx, bool never);
// (Re)allocate the bucket buffers for the current network, or release them.
nsx_status bin_configure(nsx_ctx* ctx, uint32_t mode, uint64_t cap, uint32_t tile_entries);
// Pull delivery: build (or reuse) the re-sorted copy, tables and history for the loaded network;
// free them; add every pending input into the ring and mark the history consumed; mark it
// consumed only; launch the step's gather; refresh the copy's weights.
nsx_status pull_configure(nsx_ctx* ctx);
void pull_release(nsx_ctx* ctx);
nsx_status pull_materialize(nsx_ctx* ctx);
nsx_status pull_clear(nsx_ctx* ctx);
hipError_t launch_pull(nsx_ctx* ctx, int32_t off);
bool pull_supported(const nsx_ctx* ctx);
// Shortest delay (steps) of the loaded network; the ring depth when it has no synapses.
nsx_status network_min_delay(nsx_ctx* ctx, uint32_t* d_min);
// Steps per epoch of the context as configured now (1: per-step launches).
uint32_t epoch_len(const nsx_ctx* ctx);
// The loaded network's shortest delay (computed on the device after a network is installed)
// and the epoch segments sized for epoch_len steps.
nsx_status epoch_prepare(nsx_ctx* ctx);
// Enqueue `len` steps from *tbase + off as one epoch (update launch + synapse pass).  Profiling
// events as enqueue_step's.
hipError_t enqueue_epoch(nsx_ctx* ctx, int32_t off, uint32_t len, hipEvent_t* evs);
hipError_t launch_epoch_surrogate(nsx_ctx* ctx, int32_t off, uint32_t len);
// The loaded network fits binned delivery with tiles of tile_entries entries.
bool bin_supported(const nsx_ctx* ctx, uint32_t tile_entries);
// Persistent stepping (nsx_set_persistent): whether it applies to the context as configured
// now, and one cooperative launch of `len` steps from *tbase + off.
bool persist_applies(const nsx_ctx* ctx);
hipError_t launch_persist(nsx_ctx* ctx, int32_t off, uint32_t len);
}  // namespace nsx

#define NSX_HIP(ctx, expr)                                          \
  do {                                                              \
    hipError_t _e = (expr);                                         \
    if (_e != hipSuccess) return nsx::hip_error((ctx), _e, #expr);  \
  } while (0)
