"""Host API over the C ABI, shaped like the reference's engine and GPU backend.

* `SimulationEngine`  -- src/simulation/engine.rs:43-823, device-resident (nsx_* entry points).
* `GpuBackend`        -- src/simulation/gpu_backend.rs:47-1037, method for method
                         (nsx_backend_* entry points).

Everything runs through libnsx.so; there is no Python or CPU compute path.
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import check, load


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class SimulationEngine:
    """Device-resident simulation of one network (SimulationEngine, engine.rs:43-80)."""

    _stats_on = False

    def __init__(self, cfg):
        self._lib = load()
        self.cfg = cfg.copy()
        self.cfg.struct_size = C.sizeof(abi.NsxConfig)
        h = C.c_void_p()
        check(self._lib.nsx_create(C.byref(self.cfg), C.byref(h)), None, "nsx_create")
        self._h = h
        self.dt = float(cfg.dt_ms)

    @classmethod
    def from_spec(cls, spec):
        eng = cls(spec.cfg)
        eng.generate_fixed_indegree(*spec.generator_args())
        return eng

    # ---- lifetime -------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.nsx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st, call):
        return check(st, self._h, call)

    # ---- network (Network::finalize, network/mod.rs:79-106) -------------------------------
    def build_csr(self, src, dst, weight, delay):
        src, dst = _c(src, np.uint32), _c(dst, np.uint32)
        weight, delay = _c(weight, np.float32), _c(delay, np.uint8)
        m = len(src)
        assert len(dst) == m and len(weight) == m and len(delay) == m
        self._chk(self._lib.nsx_build_csr(self._h, _ptr(src), _ptr(dst), _ptr(weight),
                                          _ptr(delay), m), "nsx_build_csr")

    finalize = build_csr

    def upload_csr(self, row_ptr, col, weight, delay):
        row_ptr, col = _c(row_ptr, np.uint64), _c(col, np.uint32)
        weight, delay = _c(weight, np.float32), _c(delay, np.uint8)
        n = len(row_ptr) - 1
        self._chk(self._lib.nsx_upload_csr(self._h, _ptr(row_ptr), _ptr(col), _ptr(weight),
                                           _ptr(delay), n, len(col)), "nsx_upload_csr")

    def generate_fixed_indegree(self, k_in, w_exc_lo, w_exc_hi, w_inh_lo, w_inh_hi, d_lo, d_hi):
        self._chk(self._lib.nsx_generate_fixed_indegree(
            self._h, k_in, w_exc_lo, w_exc_hi, w_inh_lo, w_inh_hi, d_lo, d_hi),
            "nsx_generate_fixed_indegree")

    def size(self):
        n, m = C.c_uint64(), C.c_uint64()
        self._chk(self._lib.nsx_network_size(self._h, C.byref(n), C.byref(m)), "nsx_network_size")
        return n.value, m.value

    def read_csr(self):
        n, m = self.size()
        rp = np.empty(n + 1, np.uint64)
        col = np.empty(m, np.uint32)
        w = np.empty(m, np.float32)
        d = np.empty(m, np.uint8)
        self._chk(self._lib.nsx_read_csr(self._h, _ptr(rp), _ptr(col), _ptr(w), _ptr(d)),
                  "nsx_read_csr")
        return rp, col, w, d

    def row_ptr(self):
        """The CSR row pointers alone (u64, n + 1)."""
        n, _ = self.size()
        rp = np.empty(n + 1, np.uint64)
        self._chk(self._lib.nsx_read_csr(self._h, _ptr(rp), None, None, None), "nsx_read_csr")
        return rp

    def weights(self):
        _, m = self.size()
        w = np.empty(m, np.float32)
        self._chk(self._lib.nsx_read_weights(self._h, _ptr(w), m), "nsx_read_weights")
        return w

    def upload_weights(self, weight):
        """Replace every weight (CSR slot order); topology, delays and state are kept."""
        w = _c(weight, np.float32)
        self._chk(self._lib.nsx_upload_weights(self._h, _ptr(w), len(w)), "nsx_upload_weights")

    # ---- state ----------------------------------------------------------------------------
    def read_state(self, fields=None, ring=False):
        n = int(self.cfg.n_neurons)
        fields = fields or list(abi.STATE_FIELDS)
        out, view = {}, abi.NsxStateView()
        for f in fields:
            a = np.empty(n, abi.STATE_FIELDS[f])
            out[f] = a
            setattr(view, f, C.cast(a.ctypes.data, type(getattr(view, f))))
        if ring:
            r = np.empty(int(self.cfg.ring_slots) * abi.channels(self.cfg) * n, np.int64)
            out["ring"] = r
            view.ring = C.cast(r.ctypes.data, C.POINTER(C.c_int64))
        self._chk(self._lib.nsx_read_state(self._h, C.byref(view)), "nsx_read_state")
        return out

    def upload_state(self, state):
        view = abi.NsxStateView()
        keep = []
        for f, a in state.items():
            if f == "ring":
                a = _c(a, np.int64)
                view.ring = C.cast(a.ctypes.data, C.POINTER(C.c_int64))
            else:
                a = _c(a, abi.STATE_FIELDS[f])
                setattr(view, f, C.cast(a.ctypes.data, type(getattr(view, f))))
            keep.append(a)
        self._chk(self._lib.nsx_upload_state(self._h, C.byref(view)), "nsx_upload_state")

    def reset_state(self):
        self._chk(self._lib.nsx_reset_state(self._h), "nsx_reset_state")

    def get_step(self):
        t = C.c_uint64()
        self._chk(self._lib.nsx_get_step(self._h, C.byref(t)), "nsx_get_step")
        return t.value

    def set_step(self, step):
        self._chk(self._lib.nsx_set_step(self._h, int(step)), "nsx_set_step")

    def stimulate(self, neuron_id, current):
        """SimulationEngine::stimulate (engine.rs:156-161): input for the next step."""
        self._chk(self._lib.nsx_stimulate(self._h, int(neuron_id), float(current)),
                  "nsx_stimulate")

    # ---- stepping (engine.rs:254-797) -------------------------------------------------------
    def step(self, n_steps=1):
        st = abi.NsxStepStats()
        self._chk(self._lib.nsx_step(self._h, int(n_steps), C.byref(st)), "nsx_step")
        return st

    def simulate_ms(self, duration_ms):
        st = abi.NsxStepStats()
        self._chk(self._lib.nsx_simulate_ms(self._h, float(duration_ms), C.byref(st)),
                  "nsx_simulate_ms")
        return st

    def set_graph_steps(self, g):
        self._chk(self._lib.nsx_set_graph_steps(self._h, int(g)), "nsx_set_graph_steps")

    def set_prop_split(self, waves_per_block):
        self._chk(self._lib.nsx_set_prop_split(self._h, int(waves_per_block)),
                  "nsx_set_prop_split")

    def set_update_width(self, neurons_per_thread):
        """Neurons per thread of the LIF / Izhikevich update (1, 4; 0 = by network size)."""
        self._chk(self._lib.nsx_set_update_width(self._h, int(neurons_per_thread)),
                  "nsx_set_update_width")

    def set_persistent(self, enable=True):
        """Step windows of steps as one cooperative launch each (ring-atomic delivery, one update
        launch per step); results are unchanged."""
        self._chk(self._lib.nsx_set_persistent(self._h, 1 if enable else 0), "nsx_set_persistent")

    @property
    def persistent(self):
        """(requested, applies to the loaded network now)."""
        r, a = C.c_uint32(), C.c_uint32()
        self._chk(self._lib.nsx_get_persistent(self._h, C.byref(r), C.byref(a)),
                  "nsx_get_persistent")
        return bool(r.value), bool(a.value)

    def set_epoch(self, max_steps):
        """Steps per epoch at most (<= 1: one step per launch)."""
        self._chk(self._lib.nsx_set_epoch(self._h, max_steps), "nsx_set_epoch")

    @property
    def epoch(self):
        """Steps per epoch the next step() will use (1: per-step launches)."""
        e = C.c_uint32()
        self._chk(self._lib.nsx_get_epoch(self._h, C.byref(e)), "nsx_get_epoch")
        return e.value

    def set_delivery(self, mode, cap_events=0, tile_entries=0):
        """How events reach the ring (abi.DELIVERY_ATOMIC / DELIVERY_BINNED / DELIVERY_PULL);
        results are bit-identical.  cap_events bounds one step's bucket buffer (more events:
        atomics that step); tile_entries is 4096, 8192 or 16384 (0: automatic), binned only."""
        self._chk(self._lib.nsx_set_delivery(self._h, int(mode), int(cap_events),
                                             int(tile_entries)), "nsx_set_delivery")

    @property
    def delivery(self):
        m = C.c_uint32()
        self._chk(self._lib.nsx_get_delivery(self._h, C.byref(m)), "nsx_get_delivery")
        return m.value

    def autotune(self, trial_steps=100):
        """Measure and set the fastest launch shape (state unchanged); returns (chosen split,
        chosen window, [(split, window, us per step, tile), ...]) with tile 0 for atomic
        delivery, else the binned tile size.  The other choices (delivery, tile, row-split
        threshold, synapse workgroup, update width) and their trials are in `last_tune` and
        `tune_trials()`."""
        r = abi.NsxTuneReport()
        self._chk(self._lib.nsx_autotune(self._h, int(trial_steps), C.byref(r)), "nsx_autotune")
        self.last_tune = r
        trials = [(r.trial_split[k], r.trial_graph[k], r.us_per_step[k], r.trial_tile[k])
                  for k in range(r.n_trials)]
        return r.prop_split, r.graph_steps, trials

    def tune_trials(self):
        """Every trial of the last autotune as a dict."""
        r = self.last_tune
        return [{"split": r.trial_split[k], "graph": r.trial_graph[k],
                 "delivery": r.trial_delivery[k], "tile": r.trial_tile[k],
                 "row_split": r.trial_row_split[k], "syn_block": r.trial_syn_block[k],
                 "width": r.trial_width[k], "persistent": r.trial_persistent[k],
                 "us_per_step": r.us_per_step[k]}
                for k in range(r.n_trials)]

    def set_row_split(self, slots):
        """Cut spiking rows longer than `slots` into pieces the block's wavefronts share
        (atomic delivery; 0: whole rows)."""
        self._chk(self._lib.nsx_set_row_split(self._h, slots), "nsx_set_row_split")

    def set_syn_block(self, threads):
        """Threads per workgroup of the atomic-delivery synapse pass (64 .. 512)."""
        self._chk(self._lib.nsx_set_syn_block(self._h, threads), "nsx_set_syn_block")

    def snapshot_save(self):
        """Keep a device copy of everything a step changes (one per engine)."""
        self._chk(self._lib.nsx_snapshot_save(self._h), "nsx_snapshot_save")

    def snapshot_restore(self):
        """Return to the saved step; the steps that follow repeat bit for bit."""
        self._chk(self._lib.nsx_snapshot_restore(self._h), "nsx_snapshot_restore")

    def set_profiling(self, on):
        self._chk(self._lib.nsx_set_profiling(self._h, 1 if on else 0), "nsx_set_profiling")

    def kernel_time(self, kind):
        ms, n, b = C.c_double(), C.c_uint64(), C.c_uint64()
        self._chk(self._lib.nsx_kernel_time(self._h, kind, C.byref(ms), C.byref(n), C.byref(b)),
                  "nsx_kernel_time")
        # (total ms, launches, STDP slots touched in the last nsx_step call: kind 1 potentiated,
        #  kind 2 depressed)
        return ms.value, n.value, b.value

    def raster(self, clear=False):
        navail = C.c_uint64()
        self._chk(self._lib.nsx_read_raster(self._h, None, None, 0, C.byref(navail), 0),
                  "nsx_read_raster")
        k = navail.value
        steps = np.empty(k, np.uint32)
        ids = np.empty(k, np.uint32)
        self._chk(self._lib.nsx_read_raster(self._h, _ptr(steps), _ptr(ids), k,
                                            C.byref(navail), 1 if clear else 0),
                  "nsx_read_raster")
        return steps, ids

    def step_counts(self, cap=1 << 20):
        out = np.empty(cap, np.uint32)
        n = C.c_uint64()
        self._chk(self._lib.nsx_read_step_counts(self._h, _ptr(out), cap, C.byref(n)),
                  "nsx_read_step_counts")
        return out[: min(n.value, cap)]

    def stats(self):
        """SimulationStats (engine.rs:21-33): the full set while statistics are enabled
        (set_stats), otherwise the running totals."""
        if self._stats_on:
            r = abi.NsxSimStats()
            self._chk(self._lib.nsx_read_sim_stats(self._h, C.byref(r)), "nsx_read_sim_stats")
            from .stats import SimulationStats
            return SimulationStats.from_abi(r)
        st = self.step(0)
        return {"total_spikes": st.total_spikes, "total_events": st.total_events,
                "sim_time_ms": st.sim_time_ms}

    def set_stats(self, on=True):
        """Keep SimulationStats on the device (per-step spike counts, LFP); clears them."""
        self._chk(self._lib.nsx_set_stats(self._h, 1 if on else 0), "nsx_set_stats")
        self._stats_on = bool(on)

    def lfp(self):
        """SimulationEngine::lfp (engine.rs:804-806); needs set_stats."""
        return self.stats().lfp

    def weight_stats(self):
        """(mean, population std) of all synaptic weights (engine.rs:655-666)."""
        m, s = C.c_double(), C.c_double()
        self._chk(self._lib.nsx_weight_stats(self._h, C.byref(m), C.byref(s)), "nsx_weight_stats")
        return m.value, s.value

    # ---- learned HH surrogate (no reference counterpart) ----------------------------------
    def surrogate_generate(self, n_samples, seed=1, stride=7, random_fraction=0.5,
                           scale_eps=0.5, warm_ms=20.0, mu=(-5.0, 30.0), sigma_hi=8.0,
                           tau=(1.0, 10.0)):
        """Training set from this context's plain HH step (nsx_gen_config); returns its size."""
        g = abi.NsxGenConfig()
        g.struct_size = C.sizeof(abi.NsxGenConfig)
        g.n_samples, g.seed, g.stride = int(n_samples), int(seed), int(stride)
        g.random_fraction, g.scale_eps, g.warm_ms = random_fraction, scale_eps, warm_ms
        g.mu_lo, g.mu_hi = mu
        g.sigma_hi = sigma_hi
        g.tau_lo, g.tau_hi = tau
        n = C.c_uint64()
        self._chk(self._lib.nsx_surrogate_generate(self._h, C.byref(g), C.byref(n)),
                  "nsx_surrogate_generate")
        return n.value

    def surrogate_train(self, steps, waves=512, chunks_per_wave=2, lr=2e-3, lr_min=1e-5,
                        beta1=0.9, beta2=0.999, eps=1e-8, seed=1, init=True, sample_offset=0):
        """Adam on the GPU; returns the per-step loss (mean squared normalised error)."""
        tc = abi.NsxTrainConfig()
        tc.struct_size = C.sizeof(abi.NsxTrainConfig)
        tc.steps, tc.waves, tc.chunks_per_wave = int(steps), int(waves), int(chunks_per_wave)
        tc.init, tc.sample_offset = int(bool(init)), int(sample_offset)
        tc.lr, tc.lr_min, tc.beta1, tc.beta2, tc.eps = lr, lr_min, beta1, beta2, eps
        tc.seed = int(seed)
        hist = np.zeros(int(steps), np.float32)
        self._chk(self._lib.nsx_surrogate_train(self._h, C.byref(tc), _ptr(hist)),
                  "nsx_surrogate_train")
        return hist

    def surrogate_params(self):
        """(FP32 parameter vector, normalisation vector)."""
        p = np.empty(abi.SURROGATE_PARAMS, np.float32)
        nm = np.empty(abi.SURROGATE_NORM, np.float32)
        self._chk(self._lib.nsx_surrogate_get(self._h, _ptr(p), _ptr(nm)), "nsx_surrogate_get")
        return p, nm

    def surrogate_norm(self):
        """Normalisation / envelope vector (available once training data exist)."""
        nm = np.empty(abi.SURROGATE_NORM, np.float32)
        self._chk(self._lib.nsx_surrogate_get(self._h, None, _ptr(nm)), "nsx_surrogate_get")
        return nm

    def set_surrogate_params(self, params, norm):
        p = _c(params, np.float32)
        nm = _c(norm, np.float32)
        if p.size != abi.SURROGATE_PARAMS or nm.size != abi.SURROGATE_NORM:
            raise ValueError("surrogate parameter / normalisation vector of the wrong size")
        self._chk(self._lib.nsx_surrogate_set(self._h, _ptr(p), _ptr(nm)), "nsx_surrogate_set")

    def surrogate_forward(self, x, what=0):
        """Surrogate on raw states x (count x 5 = V m h n I).  what 0: the MLP's correction;
        1: next (V, m, h, n) as the network step computes it; 2: next state of the frozen-rate
        sub-steps alone.  Returns (count x 4, envelope flags)."""
        x = _c(x, np.float32).reshape(-1, 5)
        y = np.empty((x.shape[0], 4), np.float32)
        inside = np.empty(x.shape[0], np.uint8)
        self._chk(self._lib.nsx_surrogate_forward(self._h, _ptr(x), x.shape[0], int(what),
                                                  _ptr(y), _ptr(inside)), "nsx_surrogate_forward")
        return y, inside.astype(bool)

    def surrogate_gradient(self, x, target):
        """Gradient (parameter order) and loss of the mean squared normalised error."""
        x = _c(x, np.float32).reshape(-1, 5)
        t = _c(target, np.float32).reshape(-1, 4)
        g = np.empty(abi.SURROGATE_PARAMS, np.float32)
        loss = C.c_float()
        self._chk(self._lib.nsx_surrogate_gradient(self._h, _ptr(x), _ptr(t), x.shape[0],
                                                   _ptr(g), C.byref(loss)),
                  "nsx_surrogate_gradient")
        return g, loss.value

    def surrogate_data(self, first, count):
        out = np.empty((int(count), 9), np.float32)
        self._chk(self._lib.nsx_surrogate_read_data(self._h, int(first), int(count), _ptr(out)),
                  "nsx_surrogate_read_data")
        return out

    def surrogate_fallbacks(self):
        n = C.c_uint64()
        self._chk(self._lib.nsx_surrogate_fallbacks(self._h, C.byref(n)),
                  "nsx_surrogate_fallbacks")
        return n.value


class BatchedPopulation(SimulationEngine):
    """Many independent networks stepped as one (nsx_batch_*): the evaluation side of
    Population::evolve_generation (population.rs:116-124).  `nets` is a sequence of
    NsxBatchNet; `base` supplies model, dt, drive and neuron parameters."""

    def __init__(self, base, nets):
        self._lib = load()
        self.cfg = base.copy()
        self.cfg.struct_size = C.sizeof(abi.NsxConfig)
        arr = (abi.NsxBatchNet * len(nets))(*nets)
        h = C.c_void_p()
        check(self._lib.nsx_batch_create(C.byref(self.cfg), arr, len(nets), C.byref(h)), None,
              "nsx_batch_create")
        self._h = h
        self.dt = float(base.dt_ms)
        self.nets = list(nets)
        self.offsets = np.empty(len(nets) + 1, np.uint64)
        self._chk(self._lib.nsx_batch_layout(self._h, _ptr(self.offsets)), "nsx_batch_layout")
        self.cfg.n_neurons = int(self.offsets[-1])
        self.cfg.n_exc = int(self.offsets[-1])

    def net_stats(self):
        """Per network: (total spikes, sum of squared per-neuron spike counts)."""
        k = len(self.nets)
        sp, sq = np.empty(k, np.uint64), np.empty(k, np.uint64)
        self._chk(self._lib.nsx_batch_stats(self._h, _ptr(sp), _ptr(sq)), "nsx_batch_stats")
        return sp, sq

    def net_slots(self):
        """Per network, its [first, end) range of synapse slots in the concatenated CSR (a
        network's rows are contiguous)."""
        rp = self.read_csr()[0]
        off = self.offsets.astype(np.int64)
        return [(int(rp[off[i]]), int(rp[off[i] + self.nets[i].n_neurons]))
                for i in range(len(self.nets))]

    def split(self, global_ids):
        """(network index, local index) of global neuron ids."""
        g = np.asarray(global_ids, np.uint64)
        net = np.searchsorted(self.offsets, g, side="right") - 1
        return net, (g - self.offsets[net]).astype(np.uint32)


class GpuBackend:
    """Drop-in for the wgpu GpuBackend (gpu_backend.rs:47-95), same methods and semantics."""

    LIF, IZHIKEVICH, HODGKIN_HUXLEY = 0, 1, 2

    def __init__(self, num_neurons, num_edges):
        """GpuBackend::new (gpu_backend.rs:98)."""
        self._lib = load()
        h = C.c_void_p()
        check(self._lib.nsx_backend_new(int(num_neurons), int(num_edges), C.byref(h)), None,
              "nsx_backend_new")
        self._h = h
        self.n = int(num_neurons)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.nsx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st, call):
        return check(st, self._h, call)

    def upload_initial_state(self, membrane_potential, recovery_variable, refractory_counter,
                             last_spike_time, spike_count, params, is_output):
        a = [_c(membrane_potential, np.float32), _c(recovery_variable, np.float32),
             _c(refractory_counter, np.int32), _c(last_spike_time, np.float32),
             _c(spike_count, np.uint32), _c(params, np.float32).reshape(-1),
             _c(is_output, np.uint32)]
        self._chk(self._lib.nsx_backend_upload_initial_state(self._h, *[_ptr(x) for x in a]),
                  "upload_initial_state")

    def upload_hh_gates(self, m, h, n):
        a = [_c(m, np.float32), _c(h, np.float32), _c(n, np.float32)]
        self._chk(self._lib.nsx_backend_upload_hh_gates(self._h, *[_ptr(x) for x in a]),
                  "upload_hh_gates")

    def upload_csr(self, adjacency_ptr, adjacency_indices, synapse_weights):
        a = [_c(adjacency_ptr, np.uint32), _c(adjacency_indices, np.uint32),
             _c(synapse_weights, np.float32)]
        self._chk(self._lib.nsx_backend_upload_csr(self._h, *[_ptr(x) for x in a]), "upload_csr")

    def upload_input_current(self, input_current):
        a = _c(input_current, np.float32)
        self._chk(self._lib.nsx_backend_upload_input_current(self._h, _ptr(a)),
                  "upload_input_current")

    def write_uniforms(self, dt, sim_time):
        self._chk(self._lib.nsx_backend_write_uniforms(self._h, float(dt), float(sim_time)),
                  "write_uniforms")

    def set_model(self, model):
        self._chk(self._lib.nsx_backend_set_model(self._h, int(model)), "set_model")

    def step_neurons(self):
        n = self.n
        sp, v, u = np.empty(n, np.uint32), np.empty(n, np.float32), np.empty(n, np.float32)
        refr, cnt = np.empty(n, np.int32), np.empty(n, np.uint32)
        self._chk(self._lib.nsx_backend_step_neurons(self._h, _ptr(sp), _ptr(v), _ptr(u),
                                                     _ptr(refr), _ptr(cnt)), "step_neurons")
        return sp, v, u, refr, cnt

    step_lif = step_neurons

    def step_spmv(self):
        out = np.empty(self.n, np.float32)
        self._chk(self._lib.nsx_backend_step_spmv(self._h, _ptr(out)), "step_spmv")
        return out
