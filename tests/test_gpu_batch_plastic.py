"""Batched population under pair and triplet STDP: each network learns only on its own E->E
synapses, by its own E/I split, and every network of the batch is bit-identical to its own
engine -- raster, state, pre / post trace records and the learned weights -- with direct
launches, graph windows and binned delivery (the per-network form of population.rs:126-138's
Lamarckian evaluation)."""
import numpy as np
import pytest

from neuralsim_amd import configs

pytestmark = pytest.mark.gpu

FIELDS = ["v", "u", "spike_count", "last_spike_step", "pre_trace", "pre_step", "post_trace",
          "post_step"]


def _base(rule):
    base = configs.population_base(record_raster=True, raster_capacity=1 << 23)
    stdp = configs.izh_stdp_network(n=2, rule=rule).cfg
    base.plasticity = stdp.plasticity
    for f in ("a_plus", "a_minus", "tau_plus_ms", "tau_minus_ms", "tau_y_ms", "a3_plus",
              "w_min", "w_max"):
        setattr(base, f, getattr(stdp, f))
    return base


def _nets():
    # ragged sizes and E/I splits (4/5, 1/2, 4/5, 9/10, 1/3)
    spec = [(700, 560, 50), (1500, 750, 100), (300, 240, 20), (2100, 1890, 80), (900, 300, 60)]
    nets = configs.population_nets(len(spec), seed=41)
    for b, (n, ne, k) in zip(nets, spec):
        b.n_neurons, b.n_exc, b.k_in = n, ne, k
    return nets


def _solo(nsx, base, net, steps):
    cfg = configs.population_single(base, net)
    with nsx.SimulationEngine(cfg) as eng:
        eng.generate_fixed_indegree(net.k_in, net.w_exc_lo, net.w_exc_hi, net.w_inh_lo,
                                    net.w_inh_hi, net.d_lo, net.d_hi)
        w0 = eng.weights()
        st = eng.step(steps)
        return st, eng.raster(), eng.read_state(FIELDS), w0, eng.weights(), eng.read_csr()


@pytest.mark.parametrize("rule", ["pair", "triplet"])
@pytest.mark.parametrize("graph,binned", [(0, False), (25, False), (25, True)])
def test_plastic_batch_equals_sequential(nsx, rule, graph, binned):
    base = _base(rule)
    nets = _nets()
    steps = 400
    print(f"plastic batch {rule}: {len(nets)} networks, {steps} steps, graph {graph}, "
          f"binned {binned}", flush=True)
    with nsx.BatchedPopulation(base, nets) as pop:
        if binned:  # 4096-entry tiles: the networks straddle tile boundaries
            pop.set_delivery(nsx.abi.DELIVERY_BINNED, 0, 4096)
        if graph:
            pop.set_graph_steps(graph)
        w0 = pop.weights()
        pst = pop.step(steps)
        b_steps, b_ids = pop.raster()
        b_state = pop.read_state(FIELDS)
        w1 = pop.weights()
        slots = pop.net_slots()
        off = pop.offsets.astype(np.int64)
        net_of, local = pop.split(b_ids)
    total, learned = 0, 0
    for i, net in enumerate(nets):
        st, (s_steps, s_ids), s_state, sw0, sw1, (rp, col, _, _) = _solo(nsx, base, net, steps)
        a, n = off[i], net.n_neurons
        lo, hi = slots[i]
        assert np.array_equal(w0[lo:hi].view(np.uint32), sw0.view(np.uint32)), i
        sel = net_of == i
        assert np.array_equal(b_steps[sel], s_steps) and np.array_equal(local[sel], s_ids), i
        for f in FIELDS:
            assert np.array_equal(np.ascontiguousarray(b_state[f][a:a + n]).view(np.uint32),
                                  np.ascontiguousarray(s_state[f]).view(np.uint32)), (i, f)
        assert np.array_equal(w1[lo:hi].view(np.uint32), sw1.view(np.uint32)), i
        # only the network's own E->E synapses learned
        src = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
        ee = (src < net.n_exc) & (col < net.n_exc)
        assert np.array_equal(sw1[~ee].view(np.uint32), sw0[~ee].view(np.uint32)), i
        learned += int(np.count_nonzero(sw1[ee] != sw0[ee]))
        total += st.total_spikes
    assert pst.total_spikes == total > 1000
    assert learned > 1000
