/*
 * nsx.h -- C ABI of the MI355X (gfx950) NeuralSim simulation core.
 *
 * The library replaces NeuralSim's wgpu `GpuBackend` (src/simulation/gpu_backend.rs:47-1037),
 * the slot `SimulationEngine::step()` dispatches into when a GPU backend is present
 * (src/simulation/engine.rs:254-258).  Two groups of entry points are exported:
 *
 *   1. nsx_backend_*  -- one-to-one replacements of GpuBackend's methods, same argument meaning,
 *      same per-call host round trip.  A Rust `extern "C"` block binds these where
 *      gpu_backend.rs is compiled today.
 *   2. nsx_*          -- the device-resident engine: the network, neuron state, delay rings,
 *      plasticity traces and drives live in HBM; nsx_step() advances many steps per call and
 *      only statistics cross PCIe.  This is what SimulationEngine::step()/simulate_ms()
 *      (engine.rs:254-797) become when the whole step runs on the GPU.
 *
 * Conventions (all entry points):
 *   - Every call returns an nsx_status; nothing aborts or throws across the ABI.  The text of the
 *     last error of a context is available from nsx_last_error(ctx); errors raised before a
 *     context exists are available from nsx_last_error(NULL) (thread-local).
 *   - The caller owns every host buffer.  The library copies, never retains, host pointers.
 *   - One context is driven by one host thread at a time (GpuBackend methods take &mut self).
 *     Separate contexts are independent.
 *   - Calls are synchronous with respect to host-visible outputs.  Streams/graphs are internal.
 *   - Neuron indices [0, n_exc) are excitatory, [n_exc, n_neurons) inhibitory.
 */
#ifndef NSX_H
#define NSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NSX_ABI_VERSION 3u

/* GpuError {NoAdapter, DeviceRequest} (gpu_backend.rs:24-39) widened to every failure the
 * device path can report; maps onto NeuralSimError::Gpu (src/error.rs:44-46). */
typedef enum nsx_status {
  NSX_OK = 0,
  NSX_E_NODEVICE = -1, /* no usable gfx950 device (GpuError::NoAdapter)            */
  NSX_E_OOM = -2,      /* device allocation failed                                 */
  NSX_E_INVALID = -3,  /* bad argument / configuration                             */
  NSX_E_BADCSR = -4,   /* row_ptr not monotone, column or delay out of range        */
  NSX_E_HIP = -5,      /* any other HIP runtime error (GpuError::DeviceRequest)     */
  NSX_E_STATE = -6,    /* call out of order (e.g. step before a network is loaded)  */
  NSX_E_OVERFLOW = -7  /* a bounded device buffer (spike list, raster) overflowed   */
} nsx_status;

/* NeuronModelParams variants (src/neuron/mod.rs:72-94). */
typedef enum nsx_model {
  NSX_MODEL_LIF = 0,
  NSX_MODEL_IZHIKEVICH = 1,
  NSX_MODEL_HH = 2,
  NSX_MODEL_HH_SURROGATE = 3 /* HH whose gating step is a learned MLP (no reference twin) */
} nsx_model;

typedef enum nsx_drive_kind {
  NSX_DRIVE_NONE = 0,
  NSX_DRIVE_POISSON = 1,  /* k ~ Poisson(inputs*rate*dt) events of drive_weight, exc channel */
  NSX_DRIVE_GAUSSIAN = 2, /* I += drive_mean + drive_sigma * N(0,1), fresh every step       */
  NSX_DRIVE_OU = 3        /* Ornstein-Uhlenbeck current (mean, sigma, tau), exact update     */
} nsx_drive_kind;

typedef enum nsx_plasticity_kind {
  NSX_PLAST_NONE = 0,
  NSX_PLAST_PAIR = 1,   /* StdpRule + StdpTrace as engine.rs:496-537 applies them, E->E only */
  NSX_PLAST_TRIPLET = 2 /* TripletStdpRule (plasticity.rs:48-81) applied online, E->E only:
                         * at a pre spike w -= a_minus e^(-(t - t_post)/tau-) with t_post the
                         * post neuron's last spike (same step included); at a post spike
                         * w += a_plus e^(-(t - t_pre)/tau+) + a3_plus e^(-(t - t_post')/tau_y)
                         * with t_pre the pre neuron's last spike before t and t_post' the post
                         * neuron's previous spike: the closed form on nearest-spike pairs */
} nsx_plasticity_kind;

/* CROSSING: standard HH, the membrane equation of hodgkin_huxley.rs:6-9, spike = upward
 * crossing of hh_spike_v.  REFERENCE_RESET: HodgkinHuxleyNeuron::step as written
 * (hodgkin_huxley.rs:46-96): dv = (I - i_na - i_k - i_l)/C_m dt with i_x = g_x (E_x - V), and
 * V >= -55 -> V := -70. */
typedef enum nsx_hh_spike_mode {
  NSX_HH_SPIKE_CROSSING = 0,
  NSX_HH_SPIKE_REFERENCE_RESET = 1
} nsx_hh_spike_mode;

/* Everything SimulationEngine + NeuronModelParams + SynapseDynamics + PlasticityConfig fix for
 * one network.  Fill with nsx_config_init() and override fields. */
typedef struct nsx_config {
  uint32_t struct_size; /* = sizeof(nsx_config); checked                               */
  uint32_t model;       /* nsx_model                                                     */
  uint64_t n_neurons;
  uint64_t n_exc;
  double dt_ms;         /* SimulationEngine.dt (engine.rs:45)                            */
  uint32_t v_substeps;  /* Izhikevich V sub-steps per dt (1 = reference)                 */
  uint32_t ring_slots;  /* D; every delay d (in steps) must satisfy 1 <= d <= D-1         */
  uint32_t frac_bits;   /* S: arrivals are summed as int64 rint(w * 2^S)                 */
  uint32_t record_raster; /* 1: keep every (step, neuron) spike on the device            */
  uint64_t seed;        /* Philox4x32-10 key for drives, jitter and generators           */
  uint64_t raster_capacity; /* spikes kept when record_raster = 1                        */
  /* LIF (NeuronModelParams::Lif) */
  float lif_v_rest, lif_v_reset, lif_v_th, lif_tau_m, lif_t_ref_ms, lif_r;
  /* synaptic time constants: LIF current / HH conductance decay (SynapseDynamics) */
  float tau_exc_ms, tau_inh_ms, tau_nmda_ms;
  /* Izhikevich per type: [0] excitatory, [1] inhibitory (NeuronModelParams::Izhikevich) */
  float izh_a[2], izh_b[2], izh_c[2], izh_d[2];
  float izh_c_jitter;      /* c = c0 + izh_c_jitter * r^2, r ~ U[0,1) per neuron          */
  float izh_d_jitter_frac; /* d = d0 * (1 - izh_d_jitter_frac * r^2)                       */
  /* Hodgkin-Huxley (NeuronModelParams::HodgkinHuxley) and conductance synapses */
  float hh_g_na, hh_g_k, hh_g_l, hh_e_na, hh_e_k, hh_e_l, hh_c_m;
  float hh_spike_v;        /* crossing threshold for NSX_HH_SPIKE_CROSSING                */
  uint32_t hh_spike_mode;  /* nsx_hh_spike_mode                                           */
  float e_exc, e_inh;      /* reversal potentials (AMPA/NMDA 0 mV, GABA-A -70 mV)         */
  float gmax_ampa, gmax_nmda, gmax_gaba; /* SynapseDynamics::conductance_max             */
  float mg_mM;             /* [Mg2+] for the NMDA block (types.rs:88-94); 0 = no NMDA      */
  uint32_t hh_substeps;    /* plain HH Euler sub-steps per dt                              */
  /* drive */
  uint32_t drive;          /* nsx_drive_kind                                             */
  uint32_t drive_inputs;   /* Poisson: independent inputs per neuron                       */
  float drive_rate_hz;     /* Poisson: rate of each input                                  */
  float drive_weight;      /* Poisson: weight of one input event (exc channel units)       */
  float drive_mean, drive_sigma, drive_tau_ms; /* Gaussian / OU                           */
  /* plasticity (PlasticityConfig.stdp = StdpRule, plasticity.rs:8-28) */
  uint32_t plasticity;     /* nsx_plasticity_kind                                        */
  float a_plus, a_minus, tau_plus_ms, tau_minus_ms, w_min, w_max;
  float tau_y_ms, a3_plus; /* triplet                                                      */
  uint32_t _reserved[8];
} nsx_config;

/* Per-call statistics (SimulationStats, engine.rs:21-33, device-side subset). */
typedef struct nsx_step_stats {
  uint64_t steps_done;    /* steps advanced by this call                                  */
  uint64_t step_index;    /* absolute index of the next step                              */
  uint64_t spikes;        /* spikes emitted by this call                                  */
  uint64_t events;        /* synaptic events = sum of out-degree of spiking rows          */
  uint64_t total_spikes;  /* since creation                                               */
  uint64_t total_events;
  double sim_time_ms;     /* (step_index) * dt                                            */
  double wall_ms;         /* device time of this call (hipEvents)                         */
} nsx_step_stats;

/* SoA neuron state view (NeuronArray, src/neuron/mod.rs:127-143, FP32 on the device).
 * Any pointer may be NULL (that field is skipped).  Arrays hold n_neurons entries except
 * the ring (ring_slots * channels * n_neurons int64, slot-major) and the trace arrays. */
typedef struct nsx_state_view {
  float* v;          /* membrane_potential                                               */
  float* u;          /* Izhikevich recovery_variable                                      */
  float* m;          /* hh_m                                                              */
  float* h;          /* hh_h                                                              */
  float* n;          /* hh_n                                                              */
  float* syn0;       /* LIF: I_exc; HH: g_ampa                                            */
  float* syn1;       /* LIF: I_inh; HH: g_gaba                                            */
  float* syn2;       /* HH: g_nmda                                                        */
  float* drive;      /* OU current                                                        */
  int32_t* refractory; /* refractory_counter                                              */
  int32_t* last_spike_step; /* step of last spike (-2^30 = never)                         */
  uint32_t* spike_count;    /* spike_count                                                */
  float* pre_trace;  /* lazy pair-STDP pre trace value at pre_step                       */
  int32_t* pre_step;
  float* post_trace;
  int32_t* post_step;
  float* post_trace2; /* triplet slow post trace                                         */
  int64_t* ring;     /* delay ring, ring_slots*channels*n_neurons                         */
  float* izh_c;      /* Izhikevich per-neuron reset c and increment d (the jittered         */
  float* izh_d;      /* NeuronModelParams::Izhikevich of each neuron, mod.rs:72-94)         */
} nsx_state_view;

/* ---- library ---------------------------------------------------------------------------- */
uint32_t nsx_abi_version(void);
const char* nsx_status_string(nsx_status s);
/* Last error text of ctx, or of the calling thread when ctx is NULL. */
const char* nsx_last_error(const void* ctx);
/* Number of gfx950 devices visible (GpuBackend::new's adapter request, gpu_backend.rs:102-111). */
nsx_status nsx_device_count(int* count);

typedef struct nsx_ctx nsx_ctx;

/* ---- device-resident engine (SimulationEngine::new/step/simulate_ms, engine.rs:83-797) ----- */
void nsx_config_init(nsx_config* cfg, uint32_t model);
nsx_status nsx_create(const nsx_config* cfg, nsx_ctx** out);
void nsx_destroy(nsx_ctx* ctx);

/* Network::finalize (src/network/mod.rs:79-106) on the device: synapses given in insertion
 * order as parallel arrays; builds the CSR by a stable counting sort on src, permuting col,
 * weight and delay with it.  delay is in steps. */
nsx_status nsx_build_csr(nsx_ctx* ctx, const uint32_t* src, const uint32_t* dst,
                         const float* weight, const uint8_t* delay, uint64_t m);
/* Upload an already-built CSR (GpuBackend::upload_csr, gpu_backend.rs:672-687, with 64-bit
 * row_ptr and per-synapse delays).  Validated on the device. */
nsx_status nsx_upload_csr(nsx_ctx* ctx, const uint64_t* row_ptr, const uint32_t* col,
                          const float* weight, const uint8_t* delay, uint64_t n, uint64_t m);
/* Fixed in-degree generator: every target j draws k_in presynaptic sources uniformly from
 * [0,n) \ {j}; weights U[w_exc_lo,w_exc_hi) from excitatory and U[w_inh_lo,w_inh_hi) from
 * inhibitory sources; delays uniform on {d_lo..d_hi} steps.  Insertion order is (j, k)
 * target-major; the CSR is then built exactly as nsx_build_csr would. */
nsx_status nsx_generate_fixed_indegree(nsx_ctx* ctx, uint32_t k_in, float w_exc_lo,
                                       float w_exc_hi, float w_inh_lo, float w_inh_hi,
                                       uint32_t d_lo, uint32_t d_hi);
nsx_status nsx_read_csr(nsx_ctx* ctx, uint64_t* row_ptr, uint32_t* col, float* weight,
                        uint8_t* delay);
nsx_status nsx_network_size(nsx_ctx* ctx, uint64_t* n, uint64_t* m);

/* Initial state (NeuronState::new, src/neuron/mod.rs:45-60): v = resting potential (LIF v_rest,
 * Izhikevich c, HH -65 as HodgkinHuxleyNeuron::reset_state), u = 0, gates (0.05, 0.6, 0.32),
 * no spike yet (last_spike_step = -2^30).  Called by nsx_create. */
nsx_status nsx_reset_state(nsx_ctx* ctx);
nsx_status nsx_upload_state(nsx_ctx* ctx, const nsx_state_view* s);
nsx_status nsx_read_state(nsx_ctx* ctx, nsx_state_view* s);
nsx_status nsx_read_weights(nsx_ctx* ctx, float* weight, uint64_t m);
/* Replace every synapse weight (CSR slot order, m = the network's synapse count); the topology,
 * delays and state are kept and pending deferred depressions are dropped.  Genome::build_network
 * overrides the builder's weights with the genome's this way (genome.rs:144-152). */
nsx_status nsx_upload_weights(nsx_ctx* ctx, const float* weight, uint64_t m);
/* Step index of the next step (net.time / dt).  Drives are keyed by (seed, neuron, step), so a
 * checkpoint restores it together with the state for an exact resume. */
nsx_status nsx_get_step(nsx_ctx* ctx, uint64_t* step);
nsx_status nsx_set_step(nsx_ctx* ctx, uint64_t step);
/* Add a one-off current to neuron i for the next step (SimulationEngine::stimulate). */
nsx_status nsx_stimulate(nsx_ctx* ctx, uint64_t neuron, float current);

/* Advance n_steps steps (engine.rs:254-680, device-resident).  stats may be NULL. */
nsx_status nsx_step(nsx_ctx* ctx, uint64_t n_steps, nsx_step_stats* stats);
nsx_status nsx_simulate_ms(nsx_ctx* ctx, double duration_ms, nsx_step_stats* stats);
/* Launch shape: wavefronts that share one 256-neuron block's spiking rows in propagation and
 * depression (1..64; default 4).  Results do not depend on it. */
nsx_status nsx_set_prop_split(nsx_ctx* ctx, uint32_t waves_per_block);
/* Use a captured hipGraph of `steps_per_graph` steps for nsx_step (0 = direct launches). */
nsx_status nsx_set_graph_steps(nsx_ctx* ctx, uint32_t steps_per_graph);
/* Neurons per thread of the LIF / Izhikevich update: 1, 4, or 0 = by network size (4 from 2^21
 * neurons).  Results do not depend on it; HH always runs one per thread. */
nsx_status nsx_set_update_width(nsx_ctx* ctx, uint32_t neurons_per_thread);
/* High-degree rows: with atomic delivery, a spiking row longer than `slots` synapses is cut into
 * pieces of at most `slots` that the block's wavefronts share (nsx_set_prop_split), so one hub
 * neuron's row is walked by several wavefronts.  0 (default) keeps whole rows; else 16 .. 2^20.
 * Results do not depend on it. */
nsx_status nsx_set_row_split(nsx_ctx* ctx, uint32_t slots);
/* Threads per workgroup of the atomic-delivery synapse pass: 64, 128, 256 (default) or 512.
 * Results do not depend on it. */
nsx_status nsx_set_syn_block(nsx_ctx* ctx, uint32_t threads);
/* Epoch stepping.  In a network without plasticity whose shortest delay is d_min >= 2 steps, no
 * spike of steps t .. t + d_min - 1 arrives before step t + d_min (the reference delivers every
 * spike at the next step, engine.rs:414-473 -- delays honoured here, with their ring).  Such
 * steps are then advanced as one epoch: one update launch for all of them (neuron state held
 * in registers), one synapse pass for all their spikes.  Epochs are min(d_min, max_steps) long;
 * max_steps <= 1 turns this off, the default is 128 (at most 1024).  Binned delivery and the
 * GpuBackend drop-in step one step at a time.  Results are bit-identical to per-step launches.
 * nsx_get_epoch reports the epoch length the next nsx_step will use (1: per step). */
nsx_status nsx_set_epoch(nsx_ctx* ctx, uint32_t max_steps);
nsx_status nsx_get_epoch(nsx_ctx* ctx, uint32_t* steps);

/* Persistent stepping (off by default; no reference counterpart: it replaces the per-step
 * launches of simulate_ms's loop, engine.rs:792-797, and of the population's evaluation,
 * population.rs:116-124).  A window of steps -- graph_steps of them, 1024 when graphs are off
 * -- runs as one cooperative launch of a grid resident at once: per step the update of every
 * neuron block, a grid barrier, the synapse pass with ring atomics, a grid barrier.  Applies
 * with ring-atomic delivery, the point models and plain HH, and one update launch per step (not
 * when epochs take the network over); otherwise nsx_step launches as before.  Results are
 * bit-identical to per-step launches.  The barriers' waits are bounded: should one give up,
 * nsx_step returns NSX_E_HIP and the state is not valid.  nsx_get_persistent reports the
 * request and whether it applies to the loaded network now. */
nsx_status nsx_set_persistent(nsx_ctx* ctx, int enable);
nsx_status nsx_get_persistent(nsx_ctx* ctx, uint32_t* requested, uint32_t* applies);

/* How propagated events reach their ring slot (the reference adds each event into
 * pending_input, engine.rs:414-494).  Results are bit-identical in both modes: the arrivals are
 * int64 sums, whatever their order.
 * ATOMIC: one 64-bit device atomic per event into ring[(t+d) mod D][channel][target].
 * BINNED: the synapse pass writes each event into a bucket of (delay, target tile) of its
 *   step's buffer (two counting passes give each workgroup its offsets); at arrival one
 *   workgroup per tile sums its buckets in LDS and adds them into the ring slot.  Needs
 *   D - 1 <= 256, (D - 1) * tiles <= 8192 (tiles = ceil(N * channels / tile_entries)) and
 *   M < 2^32; a step with more events than cap_events (0: min(M, 2^32 / (8 D), max(2^24, M / 8)),
 *   at least min(M, 2^16)) falls back to atomics.  Costs D * cap_events * 8 bytes.  Chosen before a
 *   network is loaded, BINNED is sized at the next nsx_step; if that network does not meet the
 *   limits, the call fails with NSX_E_INVALID and the engine continues with ATOMIC.
 * PULL (networks without plasticity): no synapse pass.  After the update of step t, one
 *   workgroup per target tile (a range of at most 16384 / channels neurons of one network)
 *   gathers the input of step t + 1 from the spikes of steps t + 2 - D .. t: for each such spike
 *   it reads the slice of the source's row that targets the tile with the delay that arrives
 *   at t + 1, sums in LDS and adds into the ring slot.  The library keeps a copy of the synapses
 *   sorted per row by (target range, delay) with a u16 offset table, and the last D steps'
 *   spike segments.  Needs plasticity NSX_PLAST_NONE, degrees below 65536 and (tiles of the
 *   widest network) * (D - 1) <= 256; costs 8 bytes per synapse plus 2 bytes per (row, key) and
 *   D * N * 4 bytes.  A batched population reads only its own networks' rows per tile. */
#define NSX_DELIVERY_ATOMIC 0u
#define NSX_DELIVERY_BINNED 1u
#define NSX_DELIVERY_PULL 2u
/* tile_entries: 4096, 8192 or 16384 LDS accumulators per tile (0: automatic; BINNED only). */
nsx_status nsx_set_delivery(nsx_ctx* ctx, uint32_t mode, uint64_t cap_events,
                            uint32_t tile_entries);
nsx_status nsx_get_delivery(nsx_ctx* ctx, uint32_t* mode);

/* Raster (record_raster=1): (step, neuron) pairs in step order; within a step, neuron order.
 * *n_out = spikes available; at most cap are copied.  clear=1 empties the raster. */
nsx_status nsx_read_raster(nsx_ctx* ctx, uint32_t* steps, uint32_t* neurons, uint64_t cap,
                           uint64_t* n_out, int clear);
/* Per-step spike counts of the last nsx_step call (cap entries at most). */
nsx_status nsx_read_step_counts(nsx_ctx* ctx, uint32_t* counts, uint64_t cap, uint64_t* n_out);

/* ---- Simulation statistics (SimulationStats, src/simulation/engine.rs:21-33, filled by
 * SimulationEngine::step phase 3, engine.rs:617-670) ---------------------------------------- */
typedef struct nsx_sim_stats {
  uint64_t total_spikes;     /* since creation / the last reset                              */
  uint64_t output_spikes;    /* 0: the device engine has no output population                */
  uint64_t active_neurons;   /* largest spike count of one step (since stats were enabled)   */
  uint64_t weight_updates;   /* spikes of the last step (the reference assigns exactly that) */
  double sim_time_ms;
  double mean_firing_rate;   /* Hz: mean over neurons of the low-pass rate r += (1000/dt *
                              * spiked - r) * dt/100, updated every step                     */
  double synchrony_index;    /* fraction of the neurons that fired in the last step          */
  double weight_mean;        /* over all synapses, two-pass FP64, fixed summation order      */
  double weight_std;         /* population standard deviation                                */
  double lfp;                /* SimulationEngine::lfp (engine.rs:492, 804-806): running sum  *
                              * of |synaptic currents| of the conductance (HH) models; 0     *
                              * for the current-based LIF and Izhikevich synapses            */
  uint64_t seed;
  uint64_t steps;            /* steps folded since stats were enabled                        */
} nsx_sim_stats;

/* Enable (1) or disable (0) the statistics; enabling clears them.  While on, every step adds
 * its spike count to a per-step counter and the HH kernels add |synaptic currents| per block;
 * nsx_step folds the counts after each launch chunk. */
nsx_status nsx_set_stats(nsx_ctx* ctx, int enable);
/* The current SimulationStats (stats on); computes the weight statistics on the device. */
nsx_status nsx_read_sim_stats(nsx_ctx* ctx, nsx_sim_stats* out);
/* Mean and population standard deviation of all synaptic weights (engine.rs:655-666). */
nsx_status nsx_weight_stats(nsx_ctx* ctx, double* mean, double* stddev);
/* nsx_weight_stats without any device allocation: the same result, bit for bit, with the
 * kernels' per-block partial sums in a library-static device array (1,024 doubles = 8 KiB per
 * GPU, held until the process ends) instead of a buffer allocated per call.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory -- a stepping loop under
 * HIP graph capture, or a serving loop that recomputes the statistics every step as
 * SimulationEngine::step's phase 3 does (engine.rs:655-666).  The array's need is fixed by the
 * reduction's grid, not by the synapse count, so every network fits.  All calls on one GPU share
 * that array: calls that overlap on one GPU race (serialise them). */
nsx_status nsx_weight_stats_noalloc(nsx_ctx* ctx, double* mean, double* stddev);

/* Kernel timing: when enabled, every launch of the named kernel class in nsx_step is bracketed
 * by hipEvents on the stream it runs on (steps run as direct launches, not graphs).  Enabling
 * starts a fresh measurement; disabling keeps it readable.  kind: 0 neuron update, 1 synapse
 * pass (propagation and STDP), 2 binned delivery's bucket passes (the scan of the synapse
 * pass's per-bucket counts and the placement of its records), 3 its arrival pass; bytes
 * report the STDP slots of the last call (1: at pre spikes, 2: at post spikes). */
nsx_status nsx_set_profiling(nsx_ctx* ctx, int enable);
nsx_status nsx_kernel_time(nsx_ctx* ctx, int kind, double* total_ms, uint64_t* launches,
                           uint64_t* bytes);

/* Device-side snapshot of everything a step changes: neuron state, delay ring, trace and post
 * records, weights under plasticity, spike segments, counters and the step index (binned
 * delivery's pending buckets are first added into the ring).  One snapshot per context; a save
 * replaces it.  nsx_snapshot_restore returns the engine to the saved step, after which the run
 * repeats bit for bit (drives are keyed by step).  No reference counterpart: the reference
 * re-runs from a checkpoint file (CheckpointManager::load, src/io/checkpoint.rs:87-100); this
 * keeps the copy in HBM, so a measurement or a tuning trial can replay the same steps. */
nsx_status nsx_snapshot_save(nsx_ctx* ctx);
nsx_status nsx_snapshot_restore(nsx_ctx* ctx);

/* ---- GpuBackend drop-in (gpu_backend.rs), FP32, synchronous ------------------------------ */
/* GpuBackend::new(num_neurons, num_edges) (gpu_backend.rs:98) */
nsx_status nsx_backend_new(uint64_t num_neurons, uint64_t num_edges, nsx_ctx** out);
/* GpuBackend::upload_initial_state (640-670); params is [resting, threshold, reset, tau_m,
 * refractory_period, input_resistance] per neuron for LIF, [a, b, c, d, 0, 0] for Izhikevich,
 * [g_na, g_k, g_l, e_na, e_k, e_l] for HH (C_m = 1).  Gates start at 0.05/0.6/0.32
 * (NeuronArray::new, src/neuron/mod.rs:158-160); nsx_backend_upload_hh_gates overrides them. */
nsx_status nsx_backend_upload_initial_state(nsx_ctx* ctx, const float* membrane_potential,
                                            const float* recovery_variable,
                                            const int32_t* refractory_counter,
                                            const float* last_spike_time,
                                            const uint32_t* spike_count, const float* params6,
                                            const uint32_t* is_output);
nsx_status nsx_backend_upload_hh_gates(nsx_ctx* ctx, const float* m, const float* h,
                                       const float* n);
/* GpuBackend::upload_csr (672-687): u32 row_ptr (n+1), u32 col, f32 weight (CSR slot order). */
nsx_status nsx_backend_upload_csr(nsx_ctx* ctx, const uint32_t* row_ptr, const uint32_t* col,
                                  const float* weight);
/* GpuBackend::upload_input_current (689-691) */
nsx_status nsx_backend_upload_input_current(nsx_ctx* ctx, const float* input_current);
/* GpuBackend::write_uniforms (693-696) */
nsx_status nsx_backend_write_uniforms(nsx_ctx* ctx, float dt, float sim_time);
/* GpuBackend::set_model (705-707): 0 Lif, 1 Izhikevich, 2 HodgkinHuxley */
nsx_status nsx_backend_set_model(nsx_ctx* ctx, uint32_t model);
/* GpuBackend::step_neurons / step_lif (711-717, 852-912): run the neuron update of the active
 * model on the uploaded input current and read back spiked, V, u, refractory, spike_count
 * (any output pointer may be NULL). */
nsx_status nsx_backend_step_neurons(nsx_ctx* ctx, uint32_t* spiked, float* membrane_potential,
                                    float* recovery_variable, int32_t* refractory_counter,
                                    uint32_t* spike_count);
/* GpuBackend::step_spmv (921-981): input_current += sum over spiking sources of the weights of
 * their CSR row (int64 fixed point, order-independent), read back input_current. */
nsx_status nsx_backend_step_spmv(nsx_ctx* ctx, float* input_current);

/* ---- launch-shape autotuner --------------------------------------------------------------------
 * No reference counterpart (the reference has one fixed wgpu dispatch shape, gpu_backend.rs:
 * 921-981).  Times trial_steps steps (rounded up to a multiple of 200, the same steps in every
 * trial) for each candidate, in a coordinate search: the propagation split (wavefronts sharing
 * a neuron block's spike segment: 1, 2, 4, 8, 16) at a 50-step graph window; under atomic
 * delivery the row-split threshold for high-degree rows (off, 256, 128, 64 slots) and the
 * synapse pass's workgroup size (128, 256, 512 threads); for LIF / Izhikevich the update's
 * neurons per thread (1, 4); the graph window (10, 25, 100, 200 steps); the other deliveries
 * (atomic, binned with 8192- and 16384-entry tiles where the network allows it) and, when the
 * delivery changed, the split again.  Every trial starts from a device-side snapshot of the
 * mutable state; the state and the step index are restored and the fastest combination set. */
typedef struct nsx_tune_report {
  uint32_t prop_split, graph_steps; /* chosen                                                 */
  uint32_t n_trials, _pad;
  double best_us_per_step;
  double us_per_step[48];           /* per trial, in trial order                              */
  uint32_t trial_split[48], trial_graph[48];
  uint32_t delivery, tile_entries;  /* chosen (tile entries: 0 atomic, 16384 pull)             */
  uint32_t trial_tile[48];          /* per trial: binned tile entries, else 0                  */
  uint32_t row_split, syn_block;    /* chosen row-split threshold (0: whole rows), workgroup  */
  uint32_t update_width, persistent; /* chosen neurons per update thread (0: not a point
                                        model); chosen persistent stepping (0 / 1)           */
  uint32_t trial_row_split[48], trial_syn_block[48], trial_width[48];
  uint32_t trial_delivery[48];      /* per trial: NSX_DELIVERY_ATOMIC, _BINNED or _PULL         */
  uint32_t trial_persistent[48];    /* per trial: 1 for persistent stepping                     */
} nsx_tune_report;

nsx_status nsx_autotune(nsx_ctx* ctx, uint32_t trial_steps, nsx_tune_report* report);

/* ---- batched evolution population ----------------------------------------------------------
 * Replaces the evaluation loop of Population::evolve_generation (src/evolution/population.rs:
 * 116-124), which builds one CPU engine per genome (Genome::build_engine, genome.rs:151-169) and
 * runs RateHomeostasis::evaluate (fitness.rs:59-75) on each in a rayon thread.  Here the
 * networks are concatenated into one context (each padded to a multiple of 256 neurons) and
 * stepped together by the ordinary nsx_step / hipGraph path.  Every random draw of network i is
 * keyed by (seed_i, local neuron index, step), so its spikes, state and counts are bit-identical
 * to a separate context created with the same config, n_neurons = n_i, seed = seed_i and
 * nsx_generate_fixed_indegree(k_in, w ranges, d_lo, d_hi).  Raster ids, state and CSR of the
 * batch use global indices: network i starts at the offset nsx_batch_layout reports. */
typedef struct nsx_batch_net {
  uint32_t n_neurons, n_exc, k_in, d_lo, d_hi, _pad;
  uint64_t seed;
  float w_exc_lo, w_exc_hi, w_inh_lo, w_inh_hi;
} nsx_batch_net;

/* base supplies the model (LIF, Izhikevich or HH), dt, drive, neuron parameters and plasticity
 * rule shared by all networks (its n_neurons, n_exc and seed are ignored).  Each network keeps
 * its own E/I split: its neuron types, in two-channel models the ring channel its inhibitory
 * sources feed, and under pair / triplet STDP its E->E synapses (the learning of each network is
 * that of its separate context: the Lamarckian evaluation of population.rs:126-138 as one
 * batch). */
nsx_status nsx_batch_create(const nsx_config* base, const nsx_batch_net* nets, uint32_t n_nets,
                            nsx_ctx** out);
/* offsets[n_nets + 1]: first global neuron of each network; offsets[n_nets] = total neurons. */
nsx_status nsx_batch_layout(nsx_ctx* ctx, uint64_t* offsets);
/* Per network: total spikes since the last reset and the sum over its neurons of the squared
 * per-neuron spike count (spikes_sq may be NULL).  Exact integers. */
nsx_status nsx_batch_stats(nsx_ctx* ctx, uint64_t* spikes, uint64_t* spikes_sq);

/* ---- learned HH surrogate -------------------------------------------------------------------
 * No counterpart in the reference (which integrates HH only by forward Euler,
 * src/neuron/hodgkin_huxley.rs:47-92).  A network step of the plain integrator is hh_substeps
 * Euler sub-steps of dt_ms/hh_substeps with the input current held, each evaluating the six gate
 * rates at its own V.  The surrogate step evaluates the rates once, at the step-start V, runs the
 * same sub-steps with them frozen, and adds a learned correction: an MLP 5 -> 64 -> 64 -> 4 (ReLU)
 * on the normalised (V, m, h, n, I) predicting (plain step - frozen-rate step) / (|dV| + eps) for
 * (V, m, h, n), dV the frozen-rate step's change of V; gates are clamped to [0, 1].  Weights are bf16 (FP32 masters); arithmetic
 * accumulates in FP32 (layers 2-3 on MFMA).  A context of model NSX_MODEL_HH_SURROGATE steps with
 * it; neurons whose (V, m, h, n, I) leave the training envelope take the plain step instead.
 * Generation and training need a context of model NSX_MODEL_HH or NSX_MODEL_HH_SURROGATE in
 * crossing spike mode and use its HH constants, dt_ms, hh_substeps and OU drive.
 *
 * Parameter vector (NSX_SURROGATE_PARAMS floats): w1[64][5] b1[64] w2[64][64] b2[64] w3[4][64]
 * b3[4].  Normalisation (NSX_SURROGATE_NORM floats): in_mu[5] in_isd[5] out_mu[4] out_sd[4]
 * lo[5] hi[5] eps pad[3]; x_n = (x - in_mu) * in_isd, output y = y_n * out_sd + out_mu,
 * correction = y * (|dV| + eps), envelope lo <= x <= hi per input. */
#define NSX_SURROGATE_PARAMS 4804u
#define NSX_SURROGATE_NORM 32u

/* Training-set generator.  Trajectories of the plain integrator driven by OU currents: a
 * fraction random_fraction with mean ~ U[mu_lo, mu_hi], sigma ~ U[0, sigma_hi],
 * tau ~ U[tau_lo, tau_hi] (ms), the rest with the context's OU drive (6, 3, 5 ms without one);
 * after warm_ms, every stride-th step's state is recorded with its target. */
typedef struct nsx_gen_config {
  uint32_t struct_size; /* sizeof(nsx_gen_config) */
  uint32_t stride;
  uint64_t n_samples;
  uint64_t seed;
  float random_fraction, scale_eps, warm_ms;
  float mu_lo, mu_hi, sigma_hi, tau_lo, tau_hi;
  uint32_t _reserved[8];
} nsx_gen_config;

typedef struct nsx_train_config {
  uint32_t struct_size;     /* sizeof(nsx_train_config)                                     */
  uint32_t steps;           /* Adam steps                                                   */
  uint32_t waves;           /* wavefronts per step (one 64-sample chunk at a time)          */
  uint32_t chunks_per_wave; /* minibatch = waves * chunks_per_wave * 64 samples             */
  uint32_t init;            /* 1: re-initialise parameters (He-uniform) and Adam state      */
  uint32_t sample_offset;   /* added to the step counter that keys minibatch sampling       */
  float lr, lr_min;         /* cosine schedule from lr to lr_min over `steps`               */
  float beta1, beta2, eps;  /* Adam                                                         */
  uint32_t _pad;
  uint64_t seed;            /* initialisation and minibatch sampling                        */
  uint32_t _reserved[8];
} nsx_train_config;

/* Fill the context's training set with ~n_samples states visited by the plain integrator
 * (OU-driven trajectories, half with the context's drive and half with random mean/sigma/tau)
 * and their targets, and derive the normalisation and envelope from it.  *n_out
 * receives the number stored. */
nsx_status nsx_surrogate_generate(nsx_ctx* ctx, const nsx_gen_config* cfg, uint64_t* n_out);
/* Train on the context's training set on the GPU (mean squared error of normalised increments).
 * loss_history (optional) receives one float per step.  Deterministic for a given set and seed. */
nsx_status nsx_surrogate_train(nsx_ctx* ctx, const nsx_train_config* cfg, float* loss_history);
/* Read / load the FP32 parameters and the normalisation (either pointer may be NULL for get). */
nsx_status nsx_surrogate_get(nsx_ctx* ctx, float* params, float* norm);
nsx_status nsx_surrogate_set(nsx_ctx* ctx, const float* params, const float* norm);
/* Evaluate the surrogate (the network step's code path) on count raw states x[count][5] =
 * (V, m, h, n, I).  what = NSX_SURR_CORRECTION: y = the MLP's output in target units (the
 * correction before the |dV| + eps factor);
 * NSX_SURR_NEXT: y = next (V, m, h, n) as the network step computes it (plain step outside the
 * envelope); NSX_SURR_FROZEN: y = next state of the frozen-rate sub-steps alone.
 * inside[count] (optional) receives the envelope flags. */
#define NSX_SURR_CORRECTION 0u
#define NSX_SURR_NEXT 1u
#define NSX_SURR_FROZEN 2u
nsx_status nsx_surrogate_forward(nsx_ctx* ctx, const float* x, uint64_t count, uint32_t what,
                                 float* y, uint8_t* inside);
/* Gradient of the mean squared normalised error of the correction over (x[count][5],
 * target[count][4]) computed by the training kernels (count a multiple of 64); grad in
 * parameter-vector order. */
nsx_status nsx_surrogate_gradient(nsx_ctx* ctx, const float* x, const float* target,
                                  uint64_t count, float* grad, float* loss);
/* Copy samples [first, first+count) of the training set: 9 floats each, the state (V m h n I)
 * and the target (plain step - frozen-rate step of V m h n, divided by |dV| + eps). */
nsx_status nsx_surrogate_read_data(nsx_ctx* ctx, uint64_t first, uint64_t count, float* out);
/* Neuron-steps that took the plain fallback since the last nsx_reset_state. */
nsx_status nsx_surrogate_fallbacks(nsx_ctx* ctx, uint64_t* count);

#ifdef __cplusplus
}
#endif
#endif /* NSX_H */
