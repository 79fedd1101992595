"""The four-neurons-per-thread update (nsx_set_update_width 4, the default from 2^21 neurons)
against the oracle's FP32 mirror, bit for bit, at sizes that exercise its edges: neuron counts
that are not multiples of 4 (a partial last thread, element loads) or odd (unaligned int64 ring
pairs), pair and triplet STDP, LIF with pair STDP at both widths, graph windows; and a batched
population of networks whose sizes leave padding inside segments, against the one-per-thread
update."""
import numpy as np
import pytest

from neuralsim_amd import configs

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def _vs_mirror(nsx, orc, spec, steps, fields, graph, width=4):
    print(f"case width{width} {spec.name} n={spec.n} steps={steps} graph={graph}", flush=True)
    with nsx.SimulationEngine.from_spec(spec) as eng:
        eng.set_update_width(width)
        eng.set_graph_steps(graph)
        st = eng.step(steps)
        g_steps, g_ids = eng.raster()
        g = eng.read_state(fields, ring=True)
        w = eng.weights()
    mir = orc.MirrorNet.from_spec(spec)
    m_steps, m_ids, m_ev = mir.run(steps)
    m = mir.read_state(fields, ring=True)
    assert st.spikes == len(m_steps) and st.events == m_ev
    assert np.array_equal(g_steps, m_steps) and np.array_equal(g_ids, m_ids), "raster differs"
    for k in m:
        assert np.array_equal(_bits(g[k]), _bits(m[k])), k
    assert np.array_equal(_bits(w), _bits(mir.weights())), "weights differ"
    return st


@pytest.mark.parametrize("n,rule,graph", [(3001, "pair", 32), (3002, "triplet", 0)])
def test_izhikevich_width4_matches_mirror(nsx, orc, n, rule, graph):
    spec = configs.izh_stdp_network(n=n, seed=4, record_raster=True, raster_capacity=1 << 22,
                                    rule=rule)
    st = _vs_mirror(nsx, orc, spec, 130, ["v", "u", "pre_trace", "post_trace", "post_step"],
                    graph)
    assert st.spikes > 1000


def test_lif_width4_matches_mirror(nsx, orc):
    spec = configs.lif_network(n=5003, seed=2, drive_scale=60.0, record_raster=True,
                               raster_capacity=1 << 22)
    st = _vs_mirror(nsx, orc, spec, 1500, ["v", "syn0", "syn1", "refractory", "spike_count"],
                    50)
    assert st.spikes > 500


@pytest.mark.parametrize("width", [1, 4])
def test_lif_pair_stdp_matches_mirror(nsx, orc, width):
    """LIF with pair STDP (the 10^8-neuron workload's model and rule) at both update widths."""
    spec = configs.lif_network(n=5001, seed=3, drive_scale=60.0, plasticity=True,
                               record_raster=True, raster_capacity=1 << 22)
    fields = ["v", "syn0", "syn1", "refractory", "spike_count", "pre_trace", "post_trace",
              "post_step"]
    st = _vs_mirror(nsx, orc, spec, 1500, fields, 50, width)
    assert st.spikes > 500
    with nsx.SimulationEngine.from_spec(spec) as eng:
        w0 = eng.weights()
    mir = orc.MirrorNet.from_spec(spec)
    mir.run(1500)
    assert np.any(mir.weights() != w0)  # the rule changed weights


def test_batch_width4_equals_width1(nsx):
    base = configs.population_base(record_raster=True, raster_capacity=1 << 23)
    nets = configs.population_nets(7, seed=17)
    for b, n, k in zip(nets, [1000, 1301, 257, 3000, 511, 2046, 703], [20, 50, 200, 30, 100, 80, 1]):
        b.n_neurons, b.n_exc, b.k_in = n, n * 4 // 5, k
    fields = ["v", "u", "spike_count", "last_spike_step"]
    out = []
    for width in (1, 4):
        print(f"case width{width} batch of {len(nets)}", flush=True)
        with nsx.BatchedPopulation(base, nets) as pop:
            pop.set_update_width(width)
            pop.set_graph_steps(25)
            pop.step(300)
            out.append((pop.raster(), pop.read_state(fields, ring=True), pop.net_stats()))
    (ra, sa, na), (rb, sb, nb) = out
    assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1]) and len(ra[0]) > 1000
    for k in sa:
        assert np.array_equal(_bits(sa[k]), _bits(sb[k])), k
    assert all(np.array_equal(x, y) for x, y in zip(na, nb))
