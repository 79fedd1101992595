"""Device statistics and file checkpoints, through the C ABI.

* SimulationStats against the reference's phase 3 (engine.rs:617-670) restated in numpy FP64
  from the spike raster and the downloaded weights: per-neuron low-pass rates averaged,
  largest per-step count, last-step count and fraction, two-pass weight mean / std;
* the LFP of an HH network against the sum of |AMPA|, |NMDA|, |GABA| currents recomputed from
  the states before and after each step;
* stats unchanged by how a run is split into nsx_step calls (one call longer than the per-step
  count ring vs several short ones);
* CheckpointManager: save to files, load into a new engine, continue bit for bit; the surrogate
  context keeps its parameters; at the 10k-neuron LIF size the reference's Network JSON is
  written too and loads into an engine with the same network and neuron state.
"""
import numpy as np
import pytest

from neuralsim_amd import abi, checkpoint, configs
from neuralsim_amd.stats import StatsRecorder

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def test_sim_stats_match_reference_phase3(nsx):
    spec = configs.izh_stdp_network(n=3000, seed=31, record_raster=True, raster_capacity=1 << 22)
    steps = 300
    print(f"case stats {spec.name} steps={steps}", flush=True)
    with nsx.SimulationEngine.from_spec(spec) as eng:
        eng.set_stats(True)
        eng.set_graph_steps(25)
        eng.step(steps)
        s = eng.stats()
        st, ids = eng.raster()
        w = eng.weights()
    n, dt = spec.n, spec.cfg.dt_ms
    counts = np.bincount(st.astype(np.int64), minlength=steps)[:steps]
    assert s.total_spikes == len(ids) and s.steps == steps
    rate = np.zeros(n)  # engine.rs:636-646: per-neuron low-pass, then the mean
    for t in range(steps):
        inst = np.zeros(n)
        inst[ids[st == t]] = 1000.0 / dt
        rate += (inst - rate) * dt / 100.0
    assert s.mean_firing_rate == pytest.approx(rate.sum() / n, rel=1e-9)
    assert s.active_neurons == counts.max() and s.weight_updates == counts[-1]
    assert s.synchrony_index == counts[-1] / n
    wd = w.astype(np.float64)
    mean = wd.sum() / len(wd)
    assert s.weight_mean == pytest.approx(mean, rel=1e-12)
    assert s.weight_std == pytest.approx(np.sqrt(((wd - mean) ** 2).sum() / len(wd)), rel=1e-10)
    assert s.lfp == 0.0  # current-based synapses
    assert s.sim_time_ms == steps * dt and s.seed == spec.cfg.seed
    rec = StatsRecorder(100, 1)
    rec.record(steps, s, s.lfp)
    assert rec.to_csv().count("\n") == 2


def test_hh_lfp_matches_currents(nsx):
    spec = configs.hh_network(n=1500, seed=8)
    c = spec.cfg
    print(f"case lfp {spec.name}", flush=True)
    f = np.float32
    want = 0.0
    with nsx.SimulationEngine.from_spec(spec) as eng:
        eng.step(4000)  # 40 ms: spikes have arrived through the 1-2 ms delays
        eng.set_stats(True)
        for _ in range(40):
            v0 = eng.read_state(["v"])["v"]
            eng.step(1)
            g = eng.read_state(["syn0", "syn1", "syn2"])
            de = (f(c.e_exc) - v0).astype(f)
            di = (f(c.e_inh) - v0).astype(f)
            term = np.abs((g["syn0"] * de).astype(f))
            term = (term + np.abs((g["syn2"] * de).astype(f))).astype(f)
            term = (term + np.abs((g["syn1"] * di).astype(f))).astype(f)
            want += term.astype(np.float64).sum()
        got = eng.lfp()
    assert want > 0 and got == pytest.approx(want, rel=1e-9)


def test_stats_independent_of_call_split(nsx):
    spec = configs.lif_network(n=600, seed=5, drive_scale=60.0, k_in=20)
    print(f"case stats split {spec.name}", flush=True)
    out = []
    for calls in ([70_000], [10_000] * 7):
        with nsx.SimulationEngine.from_spec(spec) as eng:
            eng.set_stats(True)
            eng.set_graph_steps(100)
            for k in calls:
                eng.step(k)
            s = eng.stats()
            out.append((s.total_spikes, s.active_neurons, s.weight_updates, s.mean_firing_rate,
                        s.steps))
    assert out[0] == out[1] and out[0][0] > 0 and out[0][4] == 70_000


def test_file_checkpoint_resumes_bit_exactly(nsx, tmp_path):
    spec = configs.izh_stdp_network(n=2500, seed=17, record_raster=True, raster_capacity=1 << 22)
    print(f"case checkpoint {spec.name}", flush=True)
    mgr = checkpoint.CheckpointManager(tmp_path / "cp")
    with nsx.SimulationEngine.from_spec(spec) as a:
        a.step(60)
        path = mgr.save(a, "run60")
        a.raster(clear=True)
        a.step(80)
        ra, wa, sa = a.raster(), a.weights(), a.read_state(ring=True)
    assert path.endswith("run60.nsx.json")
    (m,) = mgr.list_checkpoints()
    assert (m.name, m.neuron_count, m.synapse_count, m.sim_time_ms) == \
        ("run60", spec.n, spec.m, 60 * spec.cfg.dt_ms)
    b, meta = mgr.load("run60")
    with b:
        assert meta == m and b.get_step() == 60
        b.step(80)
        rb, wb, sb = b.raster(), b.weights(), b.read_state(ring=True)
    assert all(np.array_equal(x, y) for x, y in zip(ra, rb))
    assert np.array_equal(_bits(wa), _bits(wb))
    for k in sa:
        if k != "spike_count":
            assert np.array_equal(_bits(sa[k]), _bits(sb[k])), k


def test_file_checkpoint_keeps_surrogate(nsx, tmp_path):
    from oracle import mlp
    spec = configs.hh_surrogate_network(n=512, k_in=1, w_exc=(0.0, 0.0), w_inh=(0.0, 0.0))
    print(f"case checkpoint surrogate {spec.name}", flush=True)
    mgr = checkpoint.CheckpointManager(tmp_path)
    with nsx.SimulationEngine.from_spec(spec) as a:
        a.surrogate_generate(200_000, 3)
        a.set_surrogate_params(mlp.init_params(5), a.surrogate_norm())
        a.step(20)
        mgr.save(a, "sur")
        a.step(30)
        va = a.read_state(["v", "m", "h", "n"])
    b, _ = mgr.load("sur")
    with b:
        assert b.cfg.model == abi.MODEL_HH_SURROGATE
        b.step(30)
        vb = b.read_state(["v", "m", "h", "n"])
    for k in va:
        assert np.array_equal(_bits(va[k]), _bits(vb[k])), k


def test_network_json_checkpoint_lif_10k(nsx, tmp_path):
    """The 10k-neuron / 1M-synapse LIF network: CheckpointManager also writes the reference's
    Network JSON (checkpoint.rs:47-77); load_network builds an engine from it with the same CSR,
    weights, delays and neuron state, and the run continues from there."""
    import json
    spec = configs.lif_network(n=10_000, seed=3, drive_scale=60.0)
    print(f"case network json {spec.name} m={spec.m}", flush=True)
    mgr = checkpoint.CheckpointManager(tmp_path)
    with nsx.SimulationEngine.from_spec(spec) as a:
        a.step(200)
        mgr.save(a, "c0")
        csr_a = a.read_csr()
        sa = a.read_state(["v", "refractory", "last_spike_step", "spike_count"])
        st_a = a.step(500)
    with open(tmp_path / "c0.json") as f:
        net = json.load(f)
    assert len(net["synapses"]) == spec.m and len(net["adjacency_ptr"]) == spec.n + 1
    assert net["time"] == 200 * spec.cfg.dt_ms
    del net
    b, meta = mgr.load_network("c0", spec.cfg)
    with b:
        assert meta.synapse_count == spec.m and b.get_step() == 200
        csr_b = b.read_csr()
        for x, y in zip(csr_a, csr_b):
            assert np.array_equal(_bits(x), _bits(y))
        sb = b.read_state(["v", "refractory", "last_spike_step", "spike_count"])
        for k in sa:
            assert np.array_equal(_bits(sa[k]), _bits(sb[k])), k
        st_b = b.step(500)
    # events in flight and the synaptic currents are not part of a Network: the continued run
    # is not bit-identical, but the activity is the same network's
    assert st_a.spikes > 0 and abs(st_b.spikes - st_a.spikes) < 0.2 * st_a.spikes
