"""Binned delivery (nsx_set_delivery BINNED) through the C ABI.

Events are bucketed by (delay, target tile) and summed in LDS at arrival instead of one ring
atomic each; the ring is an int64 sum, so every result must be bit-identical to atomic
delivery and to the oracle's FP32 mirror:
* against the mirror (raster, state, ring, weights) for Izhikevich + pair / triplet STDP, LIF
  (two channels) and HH (D = 201 slots), with direct launches and graph windows;
* against atomic delivery on networks of several tiles, for every tile size;
* with a bucket buffer too small for most steps (those steps fall back to atomics);
* across read_state (pending buckets join the ring), upload into an atomic engine, set_step,
  switching modes mid-run and replacing the network;
* networks beyond its limits are refused, and leave the engine atomic.
"""
import numpy as np
import pytest

from neuralsim_amd import abi, configs

pytestmark = pytest.mark.gpu

BINNED, ATOMIC = abi.DELIVERY_BINNED, abi.DELIVERY_ATOMIC


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def _run(nsx, spec, steps, mode=BINNED, cap=0, te=0, graph=0, fields=None):
    print(f"case delivery {spec.name} n={spec.n} mode={mode} cap={cap} te={te} graph={graph} "
          f"steps={steps}", flush=True)
    with nsx.SimulationEngine.from_spec(spec) as eng:
        eng.set_delivery(mode, cap, te)
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
    st, (g_steps, g_ids), g_w, g_state = got
    mir = orc.MirrorNet.from_spec(spec)
    m_steps, m_ids, m_ev = mir.run(steps)
    m_state = mir.read_state(fields, ring=True)
    assert st.spikes == len(m_steps) and st.events == m_ev
    assert np.array_equal(g_steps, m_steps) and np.array_equal(g_ids, m_ids), "raster differs"
    for k in m_state:
        assert np.array_equal(_bits(g_state[k]), _bits(m_state[k])), f"state {k} differs"
    assert np.array_equal(_bits(g_w), _bits(mir.weights())), "weights differ"


@pytest.mark.parametrize("rule,graph", [("pair", 0), ("pair", 32), ("triplet", 25)])
def test_izhikevich_binned_matches_mirror(nsx, orc, rule, graph):
    spec = configs.izh_stdp_network(n=3000, seed=4, record_raster=True, raster_capacity=1 << 22,
                                    rule=rule)
    fields = ["v", "u", "pre_trace", "post_trace", "post_step"]
    got = _run(nsx, spec, 130, graph=graph, fields=fields)
    assert got[0].spikes > 1000
    _vs_mirror(orc, spec, 130, got, fields)


def test_lif_binned_matches_mirror(nsx, orc):
    spec = configs.lif_network(n=5000, seed=2, drive_scale=60.0, record_raster=True,
                               raster_capacity=1 << 22)
    fields = ["v", "syn0", "syn1", "refractory", "spike_count"]
    got = _run(nsx, spec, 1500, graph=50, fields=fields)
    assert got[0].spikes > 500
    _vs_mirror(orc, spec, 1500, got, fields)


def test_hh_binned_matches_mirror(nsx, orc):
    spec = configs.hh_network(n=1500, seed=8, record_raster=True, raster_capacity=1 << 20)
    fields = ["v", "m", "h", "n", "syn0", "syn1", "syn2"]
    got = _run(nsx, spec, 3000, graph=100, fields=fields)
    assert got[0].spikes > 50
    _vs_mirror(orc, spec, 3000, got, fields)


@pytest.mark.parametrize("te", [4096, 8192, 16384])
def test_many_tiles_binned_equals_atomic(nsx, te):
    # 40 000 Izhikevich neurons: 3-10 tiles, 60-200 buckets; LIF (two channels): 5-20 tiles
    izh = configs.izh_stdp_network(n=40_000, seed=9, record_raster=True, raster_capacity=1 << 23)
    a = _run(nsx, izh, 120, mode=ATOMIC, graph=40)
    b = _run(nsx, izh, 120, te=te, graph=40)
    assert a[0].events > 500_000
    _same(a, b)
    lif = configs.lif_network(n=40_000, seed=3, drive_scale=60.0, record_raster=True,
                              raster_capacity=1 << 23)
    _same(_run(nsx, lif, 400, mode=ATOMIC, graph=50), _run(nsx, lif, 400, te=te, graph=50))


def test_bucket_overflow_falls_back_to_atomics(nsx):
    spec = configs.izh_stdp_network(n=20_000, seed=12, record_raster=True, raster_capacity=1 << 23)
    a = _run(nsx, spec, 150, mode=ATOMIC)
    # a 5 000-event buffer: busy steps overflow (atomics), quiet ones are binned
    b = _run(nsx, spec, 150, cap=5000, te=4096, graph=30)
    _same(a, b)


def test_pending_buckets_survive_state_round_trips(nsx):
    spec = configs.izh_stdp_network(n=20_000, seed=21, record_raster=True, raster_capacity=1 << 23)
    print(f"case delivery round trips {spec.name}", flush=True)
    ref = _run(nsx, spec, 200, mode=ATOMIC)
    with nsx.SimulationEngine.from_spec(spec) as eng:
        eng.set_delivery(BINNED, 0, 4096)
        eng.set_graph_steps(20)
        eng.step(60)
        csr = eng.read_csr()
        mid = eng.read_state(ring=True)  # buckets in flight move into the ring
        eng.step(40)                     # continue binned after the move
        eng.set_delivery(ATOMIC)         # switch mid-run
        eng.step(30)
        eng.set_delivery(BINNED, 0, 8192)
        eng.step(70)
        got = (eng.step(0), eng.raster(), eng.weights(), eng.read_state(ring=True))
    got = (ref[0], got[1], got[2], got[3])
    _same(ref, got)
    # resume from the mid-run state in an atomic engine and a binned one
    full = ref[3]
    for mode in (ATOMIC, BINNED):
        with nsx.SimulationEngine(spec.cfg) as eng:
            eng.set_delivery(mode, 0, 4096)
            eng.upload_csr(*csr)
            eng.upload_state(mid)
            eng.set_step(60)
            eng.step(140)
            x = eng.read_state(ring=True)
        for k in full:
            if k != "spike_count":
                assert np.array_equal(_bits(full[k]), _bits(x[k])), (mode, k)


def test_network_replacement_keeps_pending_input(nsx):
    spec = configs.izh_stdp_network(n=20_000, seed=5, plasticity=False, record_raster=True,
                                    raster_capacity=1 << 23)
    print(f"case delivery network replacement {spec.name}", flush=True)
    out = []
    for mode in (ATOMIC, BINNED):
        with nsx.SimulationEngine.from_spec(spec) as eng:
            eng.set_delivery(mode, 0, 4096)
            eng.step(50)
            rp, col, w, d = eng.read_csr()
            eng.upload_csr(rp, col, (w * 0.5).astype(np.float32), d)  # in-flight input stays
            eng.step(50)
            out.append((eng.step(0), eng.raster(), eng.weights(), eng.read_state(ring=True)))
    _same(out[0], out[1])


def test_binned_delivery_edge_networks(nsx):
    """No synapses at all, and a network whose every step overflows a one-event buffer."""
    empty = configs.lif_network(n=3000, seed=6, drive_scale=60.0, k_in=0, record_raster=True,
                                raster_capacity=1 << 20)
    print("case delivery edge: empty network, one-event buffer", flush=True)
    _same(_run(nsx, empty, 300, mode=ATOMIC), _run(nsx, empty, 300, graph=50))
    spec = configs.izh_stdp_network(n=4000, seed=2, record_raster=True, raster_capacity=1 << 22)
    _same(_run(nsx, spec, 100, mode=ATOMIC), _run(nsx, spec, 100, cap=1, te=4096, graph=20))


def test_binned_delivery_rejects_unsupported_shapes(nsx):
    spec = configs.izh_stdp_network(n=3000, seed=4)
    with nsx.SimulationEngine.from_spec(spec) as eng:
        with pytest.raises(Exception):
            eng.set_delivery(BINNED, 0, 1000)  # not a supported tile size
        assert eng.delivery == ATOMIC
        eng.step(10)


def test_binned_delivery_limits(nsx):
    """A network beyond binned delivery's limits: refused when loaded, and when chosen before
    the network (sized at the first step) that step fails and the engine continues atomic."""
    spec = configs.hh_network(n=200_000, seed=3, k_in=5)  # D - 1 = 200, 98 tiles of 2048
    print("case delivery limits hh_200000", flush=True)
    with nsx.SimulationEngine.from_spec(spec) as eng:
        with pytest.raises(Exception):
            eng.set_delivery(BINNED, 0, 4096)
        assert eng.delivery == ATOMIC
    with nsx.SimulationEngine(spec.cfg) as eng:
        eng.set_delivery(BINNED, 0, 4096)  # no network yet: accepted, sized later
        eng.generate_fixed_indegree(*spec.generator_args())
        with pytest.raises(Exception):
            eng.step(1)
        assert eng.delivery == ATOMIC
        st = eng.step(20)
        assert st.steps_done == 20
