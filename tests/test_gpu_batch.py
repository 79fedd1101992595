"""Batched population: every network of the batch bit-identical to its own sequential run
(raster, state, spike counts), with direct launches and with captured graphs, with atomic and
binned delivery; per-network statistics exact; population-shaped networks bit-identical to the
oracle's FP32 mirror (raster, state, delay ring); the full 256-network population's counters
consistent (spikes per network, events = sum of count x out-degree, padding silent, fitness in
(0, 1])."""
import numpy as np
import pytest

from neuralsim_amd import configs
from neuralsim_amd.population import RateHomeostasis, evaluate

pytestmark = pytest.mark.gpu


def _nets():
    sizes = [1000, 1300, 257, 3000, 511, 2048, 700]
    ks = [20, 50, 200, 30, 100, 80, 1]
    nets = configs.population_nets(len(sizes), seed=17)
    for b, n, k in zip(nets, sizes, ks):
        b.n_neurons, b.n_exc, b.k_in = n, n * 4 // 5, k
    return nets


def _solo(nsx, base, net, steps, fields):
    cfg = configs.population_single(base, net)
    with nsx.SimulationEngine(cfg) as eng:
        eng.generate_fixed_indegree(net.k_in, net.w_exc_lo, net.w_exc_hi, net.w_inh_lo,
                                    net.w_inh_hi, net.d_lo, net.d_hi)
        st = eng.step(steps)
        r = eng.raster()
        return st, r, eng.read_state(fields), eng.read_csr()


@pytest.mark.parametrize("graph,binned", [(0, False), (25, False), (25, True)])
def test_batch_equals_sequential(nsx, graph, binned):
    base = configs.population_base(record_raster=True, raster_capacity=1 << 23)
    nets = _nets()
    steps = 300
    fields = ["v", "u", "spike_count", "last_spike_step"]
    print(f"batch of {len(nets)} networks, {steps} steps, graph {graph}, binned {binned}",
          flush=True)
    with nsx.BatchedPopulation(base, nets) as pop:
        if binned:  # 4096-entry tiles: the networks straddle tile boundaries
            pop.set_delivery(nsx.abi.DELIVERY_BINNED, 0, 4096)
        rp, col, w, d = pop.read_csr()
        if graph:
            pop.set_graph_steps(graph)
        pst = pop.step(steps)
        b_steps, b_ids = pop.raster()
        b_state = pop.read_state(fields)
        spikes, sq = pop.net_stats()
        off = pop.offsets.astype(np.int64)
        net_of, local = pop.split(b_ids)
    total = 0
    for i, net in enumerate(nets):
        st, (s_steps, s_ids), s_state, (srp, scol, sw, sd) = _solo(nsx, base, net, steps, fields)
        a, n = off[i], net.n_neurons
        # the network's rows of the concatenated CSR are its own CSR, columns shifted
        assert np.array_equal(rp[a:a + n + 1] - rp[a], srp)
        lo, hi = int(rp[a]), int(rp[a + n])
        assert np.array_equal(col[lo:hi] - a, scol)
        assert np.array_equal(w[lo:hi].view(np.uint32), sw.view(np.uint32))
        assert np.array_equal(d[lo:hi], sd)
        sel = net_of == i
        assert np.array_equal(b_steps[sel], s_steps) and np.array_equal(local[sel], s_ids), i
        for f in fields:
            assert np.array_equal(b_state[f][a:a + n].view(np.uint32),
                                  s_state[f].view(np.uint32)), (i, f)
        assert spikes[i] == st.total_spikes == len(s_steps)
        c = s_state["spike_count"].astype(np.uint64)
        assert sq[i] == int((c * c).sum())
        total += st.total_spikes
    assert pst.total_spikes == total > 1000
    # padding neurons between networks never fire
    assert np.all(np.isin(net_of, np.arange(len(nets))))
    assert np.all(local < np.array([b.n_neurons for b in nets])[net_of])


def test_population_evaluate(nsx):
    base = configs.population_base()
    nets = _nets()
    res = evaluate(base, nets, RateHomeostasis(5.0, lifetime_ms=200.0), graph_steps=50)
    assert res["steps"] == 200 and len(res["fitness"]) == len(nets)
    assert np.all(res["fitness"] > 0) and np.all(res["fitness"] <= 1)
    assert np.all(res["mean_rate"] >= 0)


def _model_case(model):
    """Base config and networks of a two-channel batch: the LIF and HH workloads' neuron,
    drive and synapse settings, seven networks of ragged sizes."""
    if model == "lif":
        base = configs.lif_network(n=2, seed=1, drive_scale=60.0, record_raster=True,
                                   raster_capacity=1 << 22).cfg
        nets = configs.population_nets(5, seed=23, n_range=(300, 3000), k_range=(10, 120),
                                       w_exc=(0.0, 0.5), w_inh=(-2.0, 0.0), delays=(10, 50))
        return base, nets, 1500, ["v", "syn0", "syn1", "refractory", "spike_count"]
    base = configs.hh_network(n=2, seed=8, record_raster=True, raster_capacity=1 << 20).cfg
    nets = configs.population_nets(4, seed=29, n_range=(300, 1500), k_range=(20, 100),
                                   w_exc=(0.0, 0.004), w_inh=(-0.02, 0.0), delays=(100, 200))
    return base, nets, 3000, ["v", "m", "h", "n", "syn0", "syn1", "syn2", "spike_count"]


@pytest.mark.parametrize("model,binned", [("lif", False), ("lif", True), ("hh", False),
                                         ("hh", True)])
def test_two_channel_batch_equals_sequential(nsx, model, binned):
    """Inhibitory sources feed the second ring channel by their own network's E/I split."""
    base, nets, steps, fields = _model_case(model)
    print(f"two-channel batch {model}: {len(nets)} networks, {steps} steps, binned {binned}",
          flush=True)
    with nsx.BatchedPopulation(base, nets) as pop:
        if binned:
            pop.set_delivery(nsx.abi.DELIVERY_BINNED, 0, 4096)
        pop.set_graph_steps(50)
        pst = pop.step(steps)
        b_steps, b_ids = pop.raster()
        b_state = pop.read_state(fields)
        off = pop.offsets.astype(np.int64)
        net_of, local = pop.split(b_ids)
    total = 0
    for i, net in enumerate(nets):
        st, (s_steps, s_ids), s_state, _ = _solo(nsx, base, net, steps, fields)
        a, n = off[i], net.n_neurons
        sel = net_of == i
        assert np.array_equal(b_steps[sel], s_steps) and np.array_equal(local[sel], s_ids), i
        for f in fields:
            assert np.array_equal(b_state[f][a:a + n].view(np.uint32),
                                  s_state[f].view(np.uint32)), (i, f)
        total += st.total_spikes
    assert pst.total_spikes == total > 100
    # the inhibitory channel carried input
    assert np.any(b_state["syn1"] != 0)


@pytest.mark.parametrize("binned", [False, True])
def test_population_networks_match_mirror(nsx, orc, binned):
    """Networks of the population workload (Izhikevich RS/FS with c/d jitter, 5 Hz Poisson drive,
    delays 1-10 ms, each its own seed and in-degree) run as one batch equal the oracle's FP32
    mirror of each network run alone: raster, state and delay ring, bit for bit."""
    base = configs.population_base(record_raster=True, raster_capacity=1 << 22)
    nets = configs.population_nets(5, seed=41, n_range=(257, 3000), k_range=(20, 200))
    steps = 250
    fields = ["v", "u", "spike_count", "last_spike_step"]
    print(f"population vs mirror: {[b.n_neurons for b in nets]} neurons, "
          f"{[b.k_in for b in nets]} in-degree, binned {binned}", flush=True)
    with nsx.BatchedPopulation(base, nets) as pop:
        if binned:
            pop.set_delivery(nsx.abi.DELIVERY_BINNED, 0, 4096)
        pop.set_graph_steps(50)
        pst = pop.step(steps)
        b_steps, b_ids = pop.raster()
        b_state = pop.read_state(fields, ring=True)
        off = pop.offsets.astype(np.int64)
        N = int(off[-1])
        net_of, local = pop.split(b_ids)
    D = int(base.ring_slots)
    ring = b_state["ring"].reshape(D, 1, N)
    total_ev = 0
    for i, net in enumerate(nets):
        cfg = configs.population_single(base, net)
        spec = configs.NetSpec("pop", cfg, net.k_in, (net.w_exc_lo, net.w_exc_hi),
                               (net.w_inh_lo, net.w_inh_hi), (net.d_lo, net.d_hi), steps)
        mir = orc.MirrorNet.from_spec(spec)
        m_steps, m_ids, m_ev = mir.run(steps)
        m_state = mir.read_state(fields, ring=True)
        mir.close()
        a, n = off[i], net.n_neurons
        sel = net_of == i
        assert np.array_equal(b_steps[sel], m_steps) and np.array_equal(local[sel], m_ids), i
        for f in fields:
            assert np.array_equal(b_state[f][a:a + n].view(np.uint32),
                                  m_state[f].view(np.uint32)), (i, f)
        assert np.array_equal(ring[:, :, a:a + n], m_state["ring"].reshape(D, 1, n)), i
        total_ev += m_ev
    assert pst.total_events == total_ev and pst.total_spikes == len(b_steps) > 500


def test_population_256_properties(nsx):
    """The full population workload (256 networks, 3.7 M neurons, 4.1e8 synapses), 300 steps:
    per-network spike totals equal the sums of their neurons' counts, events equal the sum of
    spike count x out-degree over all neurons, padding neurons never fire, every network's
    fitness lies in (0, 1]."""
    base = configs.population_base()
    nets = configs.population_nets(256)
    steps = 300
    print(f"population 256: {sum(b.n_neurons for b in nets)} neurons, {steps} steps", flush=True)
    with nsx.BatchedPopulation(base, nets) as pop:
        pop.set_graph_steps(50)
        st = pop.step(steps)
        spikes, sq = pop.net_stats()
        cnt = pop.read_state(["spike_count"])["spike_count"].astype(np.uint64)
        rp = pop.row_ptr()
        off = pop.offsets.astype(np.int64)
    outdeg = np.diff(rp.astype(np.int64)).astype(np.uint64)
    assert int((cnt * outdeg).sum()) == st.total_events
    assert int(cnt.sum()) == st.total_spikes == int(spikes.sum())
    live = np.zeros(len(cnt), bool)
    for i, b in enumerate(nets):
        a, n = off[i], b.n_neurons
        assert int(cnt[a:a + n].sum()) == int(spikes[i]), i
        assert int((cnt[a:a + n] * cnt[a:a + n]).sum()) == int(sq[i]), i
        assert int(rp[a + n] - rp[a]) == n * b.k_in, i
        live[a:a + n] = True
    assert not np.any(cnt[~live]) and not np.any(outdeg[~live])
    n = np.array([b.n_neurons for b in nets], np.float64)
    rate = spikes / n / (steps * base.dt_ms / 1e3)
    fit = RateHomeostasis(5.0).fitness(rate)
    assert np.all(fit > 0) and np.all(fit <= 1) and np.count_nonzero(spikes) == len(nets)
