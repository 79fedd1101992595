"""Pull delivery (nsx_set_delivery PULL) through the C ABI.

After each update, one workgroup per target tile gathers the next step's input from the last
D - 1 steps' spikes (rows re-sorted by (target range, delay)) and adds it into the ring slot.
The ring is an int64 sum, so every result must be bit-identical to atomic delivery and to the
oracle's FP32 mirror:
* against the mirror (raster, state, ring) for Izhikevich, LIF (two channels, D = 51) and HH
  (D = 201), with direct launches and graph windows;
* against atomic delivery on networks cut into several tiles, and on batched populations of
  ragged networks (one and two channels);
* across read_state (input owed by past spikes joins the ring), set_step, switching modes
  mid-run, snapshots, weight uploads and replacing the network;
* with the autotuner, which tries it and keeps results unchanged;
* on edge networks: no synapses, a two-slot ring, networks of two to five neurons;
* networks it cannot take (plasticity, too many tiles x delays) are refused and stay atomic.
"""
import numpy as np
import pytest

from neuralsim_amd import abi, configs

pytestmark = pytest.mark.gpu

PULL, BINNED, ATOMIC = abi.DELIVERY_PULL, abi.DELIVERY_BINNED, abi.DELIVERY_ATOMIC


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def _run(nsx, spec, steps, mode=PULL, graph=0, fields=None):
    print(f"case pull {spec.name} n={spec.n} mode={mode} graph={graph} steps={steps}", flush=True)
    with nsx.SimulationEngine.from_spec(spec) as eng:
        eng.set_delivery(mode)
        assert eng.delivery == mode
        if graph:
            eng.set_graph_steps(graph)
        st = eng.step(steps)
        return st, eng.raster(), eng.weights(), eng.read_state(fields, ring=True)


def _same(a, b):
    st_a, (sa, ia), wa, xa = a
    st_b, (sb, ib), wb, xb = b
    assert st_a.spikes == st_b.spikes and st_a.events == st_b.events
    assert np.array_equal(sa, sb) and np.array_equal(ia, ib), "raster differs"
    assert np.array_equal(_bits(wa), _bits(wb)), "weights differ"
    for k in xa:
        assert np.array_equal(_bits(xa[k]), _bits(xb[k])), f"state {k} differs"


def _vs_mirror(orc, spec, steps, got, fields):
    st, (g_steps, g_ids), _, g_state = got
    mir = orc.MirrorNet.from_spec(spec)
    m_steps, m_ids, m_ev = mir.run(steps)
    m_state = mir.read_state(fields, ring=True)
    mir.close()
    assert st.spikes == len(m_steps) and st.events == m_ev
    assert np.array_equal(g_steps, m_steps) and np.array_equal(g_ids, m_ids), "raster differs"
    for k in m_state:
        assert np.array_equal(_bits(g_state[k]), _bits(m_state[k])), f"state {k} differs"


@pytest.mark.parametrize("graph", [0, 25])
def test_izhikevich_pull_matches_mirror(nsx, orc, graph):
    spec = configs.izh_stdp_network(n=3000, seed=4, plasticity=False, record_raster=True,
                                    raster_capacity=1 << 22)
    fields = ["v", "u", "spike_count"]
    got = _run(nsx, spec, 300, graph=graph, fields=fields)
    assert got[0].spikes > 1000
    _vs_mirror(orc, spec, 300, got, fields)


def test_lif_pull_matches_mirror(nsx, orc):
    spec = configs.lif_network(n=5000, seed=2, drive_scale=60.0, record_raster=True,
                               raster_capacity=1 << 22)
    fields = ["v", "syn0", "syn1", "refractory", "spike_count"]
    got = _run(nsx, spec, 1500, graph=50, fields=fields)
    assert got[0].spikes > 500
    _vs_mirror(orc, spec, 1500, got, fields)


def test_hh_pull_matches_mirror(nsx, orc):
    spec = configs.hh_network(n=1500, seed=8, record_raster=True, raster_capacity=1 << 20)
    fields = ["v", "m", "h", "n", "syn0", "syn1", "syn2"]
    got = _run(nsx, spec, 3000, graph=100, fields=fields)
    assert got[0].spikes > 50
    _vs_mirror(orc, spec, 3000, got, fields)


def test_multi_tile_pull_equals_atomic(nsx):
    # 40 000 Izhikevich neurons: 3 tiles x 20 delays; LIF (two channels): 5 tiles x 50 delays
    izh = configs.izh_stdp_network(n=40_000, seed=9, plasticity=False, record_raster=True,
                                   raster_capacity=1 << 23)
    a = _run(nsx, izh, 150, mode=ATOMIC, graph=40)
    assert a[0].events > 500_000
    _same(a, _run(nsx, izh, 150, graph=40))
    lif = configs.lif_network(n=40_000, seed=3, drive_scale=60.0, record_raster=True,
                              raster_capacity=1 << 23)
    _same(_run(nsx, lif, 400, mode=ATOMIC, graph=50), _run(nsx, lif, 400, graph=50))


def _batch_run(nsx, base, nets, steps, mode, graph, fields):
    with nsx.BatchedPopulation(base, nets) as pop:
        pop.set_delivery(mode)
        if graph:
            pop.set_graph_steps(graph)
        st = pop.step(steps)
        return st, pop.raster(), pop.net_stats(), pop.read_state(fields, ring=True)


def _ragged_nets():
    sizes = [1000, 1300, 257, 30000, 511, 2048, 700]  # 30 000: two tiles of one network
    ks = [20, 50, 200, 30, 100, 80, 1]
    nets = configs.population_nets(len(sizes), seed=17)
    for b, n, k in zip(nets, sizes, ks):
        b.n_neurons, b.n_exc, b.k_in = n, n * 4 // 5, k
    return nets


@pytest.mark.parametrize("model", ["izhikevich", "lif"])
def test_batch_pull_equals_atomic(nsx, model):
    if model == "izhikevich":
        base = configs.population_base(record_raster=True, raster_capacity=1 << 23)
        nets, steps, fields = _ragged_nets(), 300, ["v", "u", "spike_count"]
    else:
        base = configs.lif_network(n=2, seed=1, drive_scale=60.0, record_raster=True,
                                   raster_capacity=1 << 22).cfg
        nets = configs.population_nets(5, seed=23, n_range=(300, 12000), k_range=(10, 120),
                                       w_exc=(0.0, 0.5), w_inh=(-2.0, 0.0), delays=(10, 50))
        steps, fields = 1500, ["v", "syn0", "syn1", "refractory", "spike_count"]
    print(f"case pull batch {model}: {len(nets)} networks, {steps} steps", flush=True)
    a = _batch_run(nsx, base, nets, steps, ATOMIC, 25, fields)
    b = _batch_run(nsx, base, nets, steps, PULL, 25, fields)
    assert a[0].spikes == b[0].spikes > 1000 and a[0].events == b[0].events
    assert all(np.array_equal(x, y) for x, y in zip(a[1], b[1]))
    assert all(np.array_equal(x, y) for x, y in zip(a[2], b[2]))
    for k in a[3]:
        assert np.array_equal(_bits(a[3][k]), _bits(b[3][k])), k


def test_pull_state_round_trips(nsx):
    spec = configs.izh_stdp_network(n=20_000, seed=21, plasticity=False, record_raster=True,
                                    raster_capacity=1 << 23)
    print(f"case pull round trips {spec.name}", flush=True)
    ref = _run(nsx, spec, 250, mode=ATOMIC)
    with nsx.SimulationEngine.from_spec(spec) as eng:
        eng.set_delivery(PULL)
        eng.set_graph_steps(20)
        eng.step(60)
        csr = eng.read_csr()
        mid = eng.read_state(ring=True)  # input owed by past spikes joins the ring
        eng.step(40)                     # continue pulling after the move
        eng.set_delivery(ATOMIC)         # switch mid-run
        eng.step(30)
        eng.set_delivery(PULL)
        eng.step(40)
        eng.set_delivery(BINNED, 0, 4096)
        eng.step(30)
        eng.set_delivery(PULL)
        eng.step(50)
        got = (eng.step(0), eng.raster(), eng.weights(), eng.read_state(ring=True))
    _same(ref, (ref[0], got[1], got[2], got[3]))
    full = ref[3]
    with nsx.SimulationEngine(spec.cfg) as eng:  # resume from the mid-run state, pulling
        eng.set_delivery(PULL)
        eng.upload_csr(*csr)
        eng.upload_state(mid)
        eng.set_step(60)
        eng.step(190)
        x = eng.read_state(ring=True)
    for k in full:
        if k != "spike_count":
            assert np.array_equal(_bits(full[k]), _bits(x[k])), k


def test_pull_snapshot_weights_and_replacement(nsx):
    spec = configs.izh_stdp_network(n=10_000, seed=5, plasticity=False, record_raster=True,
                                    raster_capacity=1 << 23)
    print(f"case pull snapshot / weights / replacement {spec.name}", flush=True)
    out = []
    for mode in (ATOMIC, PULL):
        with nsx.SimulationEngine.from_spec(spec) as eng:
            eng.set_delivery(mode)
            eng.set_graph_steps(10)
            eng.step(40)
            eng.snapshot_save()
            eng.step(30)
            first = eng.read_state(["v", "u"], ring=True)
            eng.snapshot_restore()
            eng.step(30)  # the same 30 steps again
            again = eng.read_state(["v", "u"], ring=True)
            for k in first:
                assert np.array_equal(_bits(first[k]), _bits(again[k])), (mode, k)
            w = eng.weights()
            eng.upload_weights((w * 0.75).astype(np.float32))  # owed input keeps its weights
            eng.step(40)
            rp, col, w, d = eng.read_csr()
            eng.upload_csr(rp, col, (w * 1.5).astype(np.float32), d)  # a new network
            eng.step(40)
            out.append((eng.step(0), eng.raster(), eng.weights(), eng.read_state(ring=True)))
    _same(out[0], out[1])


def test_autotune_tries_pull(nsx):
    base = configs.population_base(record_raster=True, raster_capacity=1 << 23)
    nets = _ragged_nets()
    print("case pull autotune", flush=True)
    fields = ["v", "u", "spike_count"]
    ref = _batch_run(nsx, base, nets, 400, ATOMIC, 0, fields)
    with nsx.BatchedPopulation(base, nets) as pop:
        pop.step(200)
        pop.autotune(100)
        trials = pop.tune_trials()
        assert any(t["delivery"] == PULL for t in trials)
        pop.step(200)
        got = (pop.step(0), pop.raster(), pop.net_stats(), pop.read_state(fields, ring=True))
    assert all(np.array_equal(x, y) for x, y in zip(ref[1], got[1]))
    for k in ref[3]:
        assert np.array_equal(_bits(ref[3][k]), _bits(got[3][k])), k


def test_pull_refuses_what_it_cannot_take(nsx):
    print("case pull limits", flush=True)
    plastic = configs.izh_stdp_network(n=3000, seed=4)  # pair STDP
    with nsx.SimulationEngine.from_spec(plastic) as eng:
        with pytest.raises(Exception):
            eng.set_delivery(PULL)
        assert eng.delivery == ATOMIC
        eng.step(10)
    # 200 000 LIF neurons: 25 tiles x 50 delays = 1 250 keys per row, over the 256 limit
    wide = configs.lif_network(n=200_000, seed=3, k_in=5)
    with nsx.SimulationEngine.from_spec(wide) as eng:
        with pytest.raises(Exception):
            eng.set_delivery(PULL)
        assert eng.delivery == ATOMIC
    with nsx.SimulationEngine(wide.cfg) as eng:
        eng.set_delivery(PULL)  # no network yet: accepted, sized later
        eng.generate_fixed_indegree(*wide.generator_args())
        with pytest.raises(Exception):
            eng.step(1)
        assert eng.delivery == ATOMIC
        assert eng.step(20).steps_done == 20


def test_pull_edge_networks(nsx):
    """No synapses; a ring of two slots (every delay one step); networks of a few neurons."""
    print("case pull edge networks", flush=True)
    empty = configs.lif_network(n=3000, seed=6, drive_scale=60.0, k_in=0, record_raster=True,
                                raster_capacity=1 << 20)
    _same(_run(nsx, empty, 300, mode=ATOMIC), _run(nsx, empty, 300, graph=50))
    fields = ["v", "u", "spike_count"]
    base = configs.population_base(record_raster=True, raster_capacity=1 << 22)
    base.ring_slots = 2
    nets = configs.population_nets(6, seed=31, n_range=(200, 5000), delays=(1, 1))
    sizes = [2, 3, 5, 300, 1200, 4000]
    for b, n in zip(nets, sizes):
        b.n_neurons, b.n_exc = n, max(1, n * 4 // 5)
        b.k_in = min(b.k_in, n)
    a = _batch_run(nsx, base, nets, 300, ATOMIC, 0, fields)
    b = _batch_run(nsx, base, nets, 300, PULL, 25, fields)
    assert a[0].spikes == b[0].spikes > 100 and a[0].events == b[0].events
    assert all(np.array_equal(x, y) for x, y in zip(a[1], b[1]))
    for k in a[3]:
        assert np.array_equal(_bits(a[3][k]), _bits(b[3][k])), k
