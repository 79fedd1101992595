// cpu_port.cpp -- CPU BASELINE ONLY (timed beside the device, never a checker of it).
//
// The reference's CPU step, SimulationEngine::step (src/simulation/engine.rs:254-697), restated
// in C++ with the reference's data layout and parallel structure, so that its time per step on
// the host's cores is the baseline the device replaces:
//   * NeuronArray (src/neuron/mod.rs:127-143) as f64 structure-of-arrays, the per-neuron
//     parameter enum carried per neuron (NeuronModelParams, mod.rs:72-94);
//   * one SynapseState-shaped record per synapse (src/synapse/mod.rs:59-79: source, target,
//     weight, delay, type, StdpTrace, flags, contribution average), CSR adjacency of usize
//     (network/mod.rs:36-48), and the engine's per-synapse conductance vector;
//   * the phases in the reference's order: drive (the reference's single-threaded noise loop,
//     engine.rs:273-287, replaced by the workload's drive), phase 1 neuron update in parallel
//     (OpenMP where the reference uses rayon, engine.rs:313-367), the spike buffer and
//     HashSet (engine.rs:369-397), conductance decay over all synapses (engine.rs:406-412),
//     propagation in spike order (engine.rs:414-473), the conductance current pass over all
//     synapses (engine.rs:475-494), pair STDP over all plastic synapses with per-synapse trace
//     decay and HashSet probes (engine.rs:496-537), the direct-mode current decay
//     (engine.rs:609-614), and phase 3 statistics: low-pass firing rates, synchrony, weight
//     mean/std over all synapses every step (engine.rs:617-670);
//   * population evaluation: one engine per genome, genomes in parallel (population.rs:116-124).
// Rules the workloads do not run (homeostasis, consolidation, R-STDP, STP) are off, as the
// workloads require (PlasticityConfig fields set accordingly).  Delays are stored but, as in the
// reference (engine.rs:429-430), not applied.  The reference's StdRng noise is replaced by the
// workload drive from a SplitMix64 stream: this is a cost baseline, not a parity target.
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <unordered_set>
#include <vector>

#include "../include/nsx.h"

namespace {

enum SynType : uint8_t { AMPA = 0, GABA = 1, GABA_B = 2, NMDA = 3 };

// SynapseDynamics constants (src/synapse/types.rs:14-54, dynamics_for 105-112)
struct Dyn {
  double decay, e_rev, gmax;
};
inline Dyn dynamics_for(uint8_t ty) {
  switch (ty) {
    case AMPA: return {2.0, 0.0, 1.0};
    case GABA: return {10.0, -70.0, 1.0};
    case GABA_B: return {150.0, -90.0, 0.5};
    default: return {100.0, 0.0, 0.5};
  }
}
inline double mg_block(uint8_t ty, double v) {  // nmda_mg_block, types.rs:88-94 ([Mg] = 1 mM)
  return ty == NMDA ? 1.0 / (1.0 + 1.0 / 3.57 * exp(-0.062 * v)) : 1.0;
}

struct StdpTrace {
  double pre_trace = 0.0, post_trace = 0.0, last_pre = -1e9, last_post = -1e9;
};

// SynapseState (src/synapse/mod.rs:59-79); Option<StdpTrace> carried inline with a flag
struct Synapse {
  uint64_t source, target;
  double weight, delay;
  uint8_t type;
  bool is_active, plasticity_enabled, has_trace;
  double conductance;
  StdpTrace trace;
  double contribution_avg;
};

enum Model : uint8_t { LIF = 0, IZH = 1, HH = 2 };
struct Params {  // NeuronModelParams (mod.rs:72-94): 7 f64 + tag
  double p[7];
  uint8_t tag;
};

inline uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline double u01(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

struct port_net {
  nsx_config cfg;
  uint64_t n = 0, m = 0;
  // NeuronArray
  std::vector<double> v, u, last_spike, input_current, hh_m, hh_h, hh_n;
  std::vector<int32_t> refr;
  std::vector<uint64_t> spike_count;
  std::vector<uint8_t> neuron_type, is_output, just_spiked;
  std::vector<Params> params;
  // Network
  std::vector<uint64_t> adjacency_ptr, adjacency_indices;
  std::vector<Synapse> synapses;
  // SimulationEngine
  std::vector<double> conductances, firing_rates, ou;
  std::deque<std::vector<uint8_t>> spike_history;
  bool use_conductance = false, stdp = false;
  double time = 0.0, lfp = 0.0, weight_mean = 0.0, weight_std = 0.0, mean_rate = 0.0;
  uint64_t total_spikes = 0, active = 0, rng = 0;
  std::vector<uint32_t> pthr;  // Poisson inversion thresholds of the drive
};

// IzhikevichNeuron / LifNeuron / HodgkinHuxleyNeuron::step (the reference code as written)
static bool step_neuron(port_net* r, uint64_t i, double dt, double I) {
  const Params& P = r->params[i];
  double& v = r->v[i];
  if (P.tag == LIF) {  // lif.rs:15-61
    if (r->refr[i] > 0) {
      r->refr[i] -= 1;
      return false;
    }
    v += (-(v - P.p[0]) + P.p[5] * I) * dt / P.p[3];
    if (v >= P.p[1]) {
      v = P.p[2];
      r->refr[i] = (int32_t)round(P.p[4] / dt);
      r->spike_count[i] += 1;
      r->just_spiked[i] = 1;
      return true;
    }
    return false;
  }
  if (P.tag == IZH) {  // izhikevich.rs:20-52
    double& u = r->u[i];
    if (v >= 30.0) {
      v = P.p[2];
      u += P.p[3];
      r->spike_count[i] += 1;
      r->just_spiked[i] = 1;
      return true;
    }
    const double dv = (0.04 * v * v + 5.0 * v + 140.0 - u + I) * dt;
    const double du = P.p[0] * (P.p[1] * v - u) * dt;
    v += dv;
    u += du;
    r->just_spiked[i] = 0;
    return false;
  }
  // hodgkin_huxley.rs:46-96
  const double vv = v;
  double am = fabs(vv + 40.0) < 1e-7 ? 1.0 : 0.1 * (vv + 40.0) / (1.0 - exp(-(vv + 40.0) / 10.0));
  double bm = 4.0 * exp(-(vv + 65.0) / 18.0);
  double ah = 0.07 * exp(-(vv + 65.0) / 20.0);
  double bh = 1.0 / (1.0 + exp(-(vv + 35.0) / 10.0));
  double an = fabs(vv + 55.0) < 1e-7 ? 0.1 : 0.01 * (vv + 55.0) / (1.0 - exp(-(vv + 55.0) / 10.0));
  double bn = 0.125 * exp(-(vv + 65.0) / 80.0);
  double& m = r->hh_m[i];
  double& h = r->hh_h[i];
  double& nn = r->hh_n[i];
  m += (am * (1.0 - m) - bm * m) * dt;
  h += (ah * (1.0 - h) - bh * h) * dt;
  nn += (an * (1.0 - nn) - bn * nn) * dt;
  const double i_na = P.p[0] * m * m * m * h * (P.p[3] - vv);
  const double i_k = P.p[1] * nn * nn * nn * nn * (P.p[4] - vv);
  const double i_l = P.p[2] * (P.p[5] - vv);
  v += (I - i_na - i_k - i_l) / P.p[6] * dt;
  if (v >= -55.0) {
    v = -70.0;
    r->spike_count[i] += 1;
    r->just_spiked[i] = 1;
    return true;
  }
  return false;
}

extern "C" {

int port_threads(void) { return omp_get_max_threads(); }

// Build one engine on a CSR network (row_ptr[n + 1], col, weight in CSR slot order).  E->E
// synapses get a StdpTrace when the config has pair STDP; conductance = 1 runs the engine in
// conductance mode (set_conductance_mode, engine.rs:139-142) with AMPA (+ NMDA when mg_mM > 0)
// for excitatory and GABA for inhibitory sources, else in direct-current mode.
port_net* port_create(const nsx_config* cfg, const uint64_t* row_ptr, const uint32_t* col,
                      const float* weight, const uint8_t* delay, int conductance) {
  port_net* r = new port_net();
  r->cfg = *cfg;
  const uint64_t n = cfg->n_neurons, m = row_ptr[n];
  r->n = n;
  r->m = m;
  r->rng = cfg->seed ^ 0x5DEECE66Dull;
  r->use_conductance = conductance != 0;
  r->stdp = cfg->plasticity == NSX_PLAST_PAIR;
  r->v.assign(n, 0.0);
  r->u.assign(n, 0.0);
  r->last_spike.assign(n, -1e9);
  r->input_current.assign(n, 0.0);
  r->hh_m.assign(n, 0.05);
  r->hh_h.assign(n, 0.6);
  r->hh_n.assign(n, 0.32);
  r->refr.assign(n, 0);
  r->spike_count.assign(n, 0);
  r->neuron_type.assign(n, 0);
  r->is_output.assign(n, 0);
  r->just_spiked.assign(n, 0);
  r->params.resize(n);
  r->firing_rates.assign(n, 0.0);
  r->ou.assign(n, cfg->drive_mean);
  for (uint64_t i = 0; i < n; ++i) {
    Params& P = r->params[i];
    const int ty = i < cfg->n_exc ? 0 : 1;
    r->neuron_type[i] = (uint8_t)ty;
    if (cfg->model == NSX_MODEL_LIF) {
      P = {{cfg->lif_v_rest, cfg->lif_v_th, cfg->lif_v_reset, cfg->lif_tau_m, cfg->lif_t_ref_ms,
            cfg->lif_r, 0.0},
           LIF};
      r->v[i] = cfg->lif_v_rest;
    } else if (cfg->model == NSX_MODEL_IZHIKEVICH) {
      P = {{cfg->izh_a[ty], cfg->izh_b[ty], cfg->izh_c[ty], cfg->izh_d[ty], 0, 0, 0}, IZH};
      r->v[i] = cfg->izh_c[ty];
    } else {
      P = {{cfg->hh_g_na, cfg->hh_g_k, cfg->hh_g_l, cfg->hh_e_na, cfg->hh_e_k, cfg->hh_e_l,
            cfg->hh_c_m},
           HH};
      r->v[i] = -65.0;
    }
  }
  r->adjacency_ptr.assign(row_ptr, row_ptr + n + 1);
  r->adjacency_indices.resize(m);
  r->synapses.resize(m);
  r->conductances.assign(m, 0.0);
  const bool nmda = cfg->mg_mM > 0.0f;
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint64_t i = 0; i < n; ++i) {
    const bool exc = i < cfg->n_exc;
    for (uint64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
      Synapse& s = r->synapses[k];
      s.source = i;
      s.target = col[k];
      s.weight = weight[k];
      s.delay = delay ? (double)delay[k] * cfg->dt_ms : 1.0;
      s.type = exc ? (nmda && (k & 1) ? NMDA : AMPA) : GABA;
      s.is_active = true;
      s.conductance = 0.0;
      s.plasticity_enabled = r->stdp && exc && col[k] < cfg->n_exc;
      s.has_trace = s.plasticity_enabled;
      s.trace = StdpTrace();
      s.contribution_avg = 0.0;
      r->adjacency_indices[k] = col[k];
    }
  }
  if (cfg->drive == NSX_DRIVE_POISSON) {
    // inversion thresholds of Poisson(inputs * rate * dt)
    const double lam = (double)cfg->drive_inputs * cfg->drive_rate_hz * cfg->dt_ms / 1000.0;
    double p = exp(-lam), c = p;
    for (int k = 0; k < 64 && c < 1.0 - 1e-16; ++k) {
      r->pthr.push_back((uint32_t)std::min(c * 4294967296.0, 4294967295.0));
      p *= lam / (k + 1);
      c += p;
    }
  }
  return r;
}

void port_destroy(port_net* r) { delete r; }

// Start from a given membrane state (the device's, so both runs share the activity regime).
void port_set_state(port_net* r, const float* v, const float* u) {
  for (uint64_t i = 0; i < r->n; ++i) {
    if (v) r->v[i] = v[i];
    if (u) r->u[i] = u[i];
  }
}

// Run steps; returns wall seconds; spikes and events (sum of out-degrees of spiking rows).
double port_run(port_net* r, uint32_t steps, uint64_t* spikes_out, uint64_t* events_out) {
  const nsx_config& cfg = r->cfg;
  const uint64_t n = r->n, m = r->m;
  const double dt = cfg.dt_ms;
  uint64_t spikes = 0, events = 0;
  std::vector<uint8_t> spiked(n);
  std::vector<uint64_t> buffer;
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t step = 0; step < steps; ++step) {
    // the workload's drive, single-threaded as the reference's noise injection
    for (uint64_t i = 0; i < n; ++i) {
      if (cfg.drive == NSX_DRIVE_POISSON) {
        const uint32_t x = (uint32_t)splitmix(r->rng);
        uint32_t c = 0;
        while (c < r->pthr.size() && x >= r->pthr[c]) ++c;
        r->input_current[i] += c * cfg.drive_weight;
      } else if (cfg.drive == NSX_DRIVE_GAUSSIAN || cfg.drive == NSX_DRIVE_OU) {
        const double u1 = 1.0 - u01(r->rng), u2 = u01(r->rng);
        const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        if (cfg.drive == NSX_DRIVE_GAUSSIAN) {
          r->input_current[i] += cfg.drive_mean + cfg.drive_sigma * z;
        } else {
          const double e = exp(-dt / cfg.drive_tau_ms);
          r->ou[i] = r->ou[i] * e + cfg.drive_mean * (1.0 - e) +
                     cfg.drive_sigma * sqrt(1.0 - e * e) * z;
          r->input_current[i] += r->ou[i];
        }
      }
    }
    // Phase 1: neuron update, parallel over neurons (rayon par_iter_mut, engine.rs:313-367)
    const double now = r->time;
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
      const bool s = step_neuron(r, i, dt, r->input_current[i]);
      if (s) r->last_spike[i] = now;
      spiked[i] = s;
    }
    buffer.clear();
    for (uint64_t i = 0; i < n; ++i)
      if (spiked[i]) buffer.push_back(i);
    r->spike_history.push_back(spiked);  // synchrony window of 100 steps
    while (r->spike_history.size() > 100) r->spike_history.pop_front();
    r->time += dt;
    std::unordered_set<uint64_t> spike_set(buffer.begin(), buffer.end());
    // Phase 2: conductance decay over every synapse
    if (r->use_conductance)
      for (uint64_t k = 0; k < m; ++k) {
        const Dyn dy = dynamics_for(r->synapses[k].type);
        r->conductances[k] *= exp(-dt / dy.decay);
      }
    // propagation in spike order
    const double cdec = exp(-dt / 1000.0);  // contribution_tau 1000 ms
    for (uint64_t i : buffer) {
      for (uint64_t k = r->adjacency_ptr[i]; k < r->adjacency_ptr[i + 1]; ++k) {
        const uint64_t tgt = r->adjacency_indices[k];
        Synapse& s = r->synapses[k];
        if (r->use_conductance) {
          const Dyn dy = dynamics_for(s.type);
          const double gp = s.weight * dy.gmax * mg_block(s.type, r->v[tgt]);
          r->conductances[k] += (s.type == GABA || s.type == GABA_B) ? fabs(gp) : gp;
          s.contribution_avg = s.contribution_avg * cdec + fabs(gp) * (1.0 - cdec);
        } else {
          r->input_current[tgt] += s.weight;
          s.contribution_avg = s.contribution_avg * cdec + fabs(s.weight) * (1.0 - cdec);
        }
      }
      events += r->adjacency_ptr[i + 1] - r->adjacency_ptr[i];
    }
    // conductance currents over every synapse
    if (r->use_conductance)
      for (uint64_t k = 0; k < m; ++k) {
        const double g = r->conductances[k];
        if (fabs(g) < 1e-12) continue;
        const Synapse& s = r->synapses[k];
        const double isyn = g * (dynamics_for(s.type).e_rev - r->v[s.target]);
        r->input_current[s.target] += isyn;
        r->lfp += fabs(isyn);
      }
    // pair STDP over every plastic synapse (StdpRule defaults of the workload)
    if (r->stdp) {
      const double t_now = r->time;
      for (uint64_t k = 0; k < m; ++k) {
        Synapse& s = r->synapses[k];
        if (!s.plasticity_enabled || !s.has_trace) continue;
        StdpTrace& tr = s.trace;
        tr.pre_trace *= exp(-dt / cfg.tau_plus_ms);
        tr.post_trace *= exp(-dt / cfg.tau_minus_ms);
        const bool pre = spike_set.count(s.source) != 0;
        const bool post = spike_set.count(s.target) != 0;
        if (pre) {
          const double w = s.weight + cfg.a_plus * tr.post_trace;
          s.weight = w < cfg.w_min ? cfg.w_min : (w > cfg.w_max ? cfg.w_max : w);
          tr.pre_trace = tr.pre_trace * exp(-(t_now - tr.last_pre) / cfg.tau_plus_ms) + 1.0;
          tr.last_pre = t_now;
        }
        if (post) {
          const double w = s.weight - cfg.a_minus * tr.pre_trace;
          s.weight = w < cfg.w_min ? cfg.w_min : (w > cfg.w_max ? cfg.w_max : w);
          tr.post_trace = tr.post_trace * exp(-(t_now - tr.last_post) / cfg.tau_minus_ms) + 1.0;
          tr.last_post = t_now;
        }
      }
    }
    if (!r->use_conductance)
      for (uint64_t i = 0; i < n; ++i) r->input_current[i] *= 0.92;
    // Phase 3: statistics
    r->total_spikes += buffer.size();
    if (buffer.size() > r->active) r->active = buffer.size();
    double rsum = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
      const double inst = spike_set.count(i) ? 1000.0 / dt : 0.0;
      r->firing_rates[i] += (inst - r->firing_rates[i]) * dt / 100.0;
      rsum += r->firing_rates[i];
    }
    r->mean_rate = n ? rsum / (double)n : 0.0;
    if (m) {
      double ws = 0.0;
      for (uint64_t k = 0; k < m; ++k) ws += r->synapses[k].weight;
      const double mean = ws / (double)m;
      double var = 0.0;
      for (uint64_t k = 0; k < m; ++k) {
        const double d = r->synapses[k].weight - mean;
        var += d * d;
      }
      r->weight_mean = mean;
      r->weight_std = sqrt(var / (double)m);
    }
    spikes += buffer.size();
  }
  const double wall =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (spikes_out) *spikes_out = spikes;
  if (events_out) *events_out = events;
  return wall;
}

// Population evaluation: engines evaluated in parallel, one per thread at a time
// (population.rs:116-124).  secs[k] receives engine k's own run time; returns the wall time.
double port_run_many(port_net** nets, int count, uint32_t steps, double* secs) {
  const auto t0 = std::chrono::steady_clock::now();
  const int saved = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(saved)
  for (int k = 0; k < count; ++k) {
    // inside a genome's engine the per-neuron loop runs on the calling thread (rayon nests)
    omp_set_num_threads(1);
    secs[k] = port_run(nets[k], steps, nullptr, nullptr);
  }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
