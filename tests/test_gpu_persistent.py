"""Persistent stepping on the device, through the C ABI: a window of steps runs as one
cooperative launch whose workgroups update every neuron block, meet at a grid barrier, run the
synapse pass with ring atomics and meet again, step after step.  Every run here equals the same
run stepped one launch per step, bit for bit -- raster, every state field, the delay ring,
weights and counters:

* Izhikevich + pair STDP and + triplet STDP, in windows of 1024 steps and of 50, with the
  row-split threshold cutting the 100-synapse rows into pieces, with statistics on;
* LIF and plain HH with epochs off (their delays span several steps, so epochs would otherwise
  take the network over), and with epochs on, where persistent stepping does not apply;
* the batched population (networks of uneven size in one concatenated CSR), with its
  per-network spike counts;
* binned delivery, where it does not apply.

Should a grid barrier give up waiting, nsx_step returns an error, which fails the test.
"""
import numpy as np
import pytest

from neuralsim_amd import abi, configs

pytestmark = pytest.mark.gpu


def _run(nsx, spec, steps, persistent, graph=0, epoch=None, row_split=None, stats=False):
    with nsx.SimulationEngine.from_spec(spec) as eng:
        if epoch is not None:
            eng.set_epoch(epoch)
        if row_split is not None:
            eng.set_row_split(row_split)
        if stats:
            eng.set_stats(True)
        eng.set_graph_steps(graph)
        eng.set_persistent(persistent)
        applies = eng.persistent
        st = eng.step(steps)
        out = {"applies": applies, "stats": st, "raster": eng.raster(),
               "state": eng.read_state(None, ring=True), "w": eng.weights(),
               "t": eng.get_step()}
        if stats:
            out["sim"] = eng.stats()
        return out


def _same(a, b):
    assert a["t"] == b["t"]
    assert a["stats"].spikes == b["stats"].spikes and a["stats"].events == b["stats"].events
    assert all(np.array_equal(x, y) for x, y in zip(a["raster"], b["raster"]))
    for k in a["state"]:
        assert np.array_equal(np.ascontiguousarray(a["state"][k]).view(np.uint8),
                              np.ascontiguousarray(b["state"][k]).view(np.uint8)), k
    assert np.array_equal(a["w"].view(np.uint32), b["w"].view(np.uint32)), "weights"


@pytest.mark.parametrize("rule,graph,row_split", [("pair", 0, None), ("pair", 50, 64),
                                                  ("triplet", 50, None)])
def test_izh_stdp_persistent_equals_per_step(nsx, rule, graph, row_split):
    spec = configs.izh_stdp_network(n=6000, seed=21, rule=rule, record_raster=True,
                                    raster_capacity=1 << 22)
    steps = 1100  # past one 1024-step window
    print(f"case persistent izh {rule} graph {graph} row_split {row_split}", flush=True)
    a = _run(nsx, spec, steps, True, graph, row_split=row_split)
    b = _run(nsx, spec, steps, False, graph, row_split=row_split)
    assert a["applies"] == (True, True) and b["applies"] == (False, False)
    assert a["stats"].spikes > 2000 and a["stats"].events > 100_000
    _same(a, b)


def test_izh_persistent_statistics(nsx):
    spec = configs.izh_stdp_network(n=4000, seed=5, record_raster=True, raster_capacity=1 << 22)
    steps = 300
    print("case persistent izh statistics", flush=True)
    a = _run(nsx, spec, steps, True, 0, stats=True)
    b = _run(nsx, spec, steps, False, 0, stats=True)
    assert a["applies"][1]
    _same(a, b)
    assert a["sim"] == b["sim"] and a["sim"].total_spikes == a["stats"].spikes


@pytest.mark.parametrize("model", ["lif", "hh"])
def test_epoch_models_persistent(nsx, model):
    if model == "lif":
        spec = configs.lif_network(n=3000, seed=9, drive_scale=60.0, record_raster=True,
                                   raster_capacity=1 << 22)
        steps = 737
    else:
        spec = configs.hh_network(n=1500, seed=13, record_raster=True, raster_capacity=1 << 22)
        steps = 600
    print(f"case persistent {model}", flush=True)
    a = _run(nsx, spec, steps, True, 0, epoch=1)
    b = _run(nsx, spec, steps, False, 0, epoch=1)
    c = _run(nsx, spec, steps, True, 25)  # epochs take the network over: persistent off
    assert a["applies"] == (True, True) and c["applies"] == (True, False)
    assert a["stats"].spikes > 50
    _same(a, b)
    _same(c, b)


def test_population_persistent_equals_per_step(nsx):
    base = configs.population_base(record_raster=True, raster_capacity=1 << 23)
    sizes = [1000, 1300, 257, 3000, 511, 2048, 700]
    ks = [20, 50, 200, 30, 100, 80, 1]
    nets = configs.population_nets(len(sizes), seed=17)
    for b, n, k in zip(nets, sizes, ks):
        b.n_neurons, b.n_exc, b.k_in = n, n * 4 // 5, k
    steps = 400
    fields = ["v", "u", "spike_count", "last_spike_step"]
    out = []
    for persistent, graph in ((True, 25), (True, 0), (False, 25)):
        print(f"case persistent population {persistent} graph {graph}", flush=True)
        with nsx.BatchedPopulation(base, nets) as pop:
            pop.set_graph_steps(graph)
            pop.set_persistent(persistent)
            applies = pop.persistent
            st = pop.step(steps)
            out.append((applies, st, pop.raster(), pop.read_state(fields), pop.net_stats()))
    assert out[0][0] == (True, True) and out[2][0] == (False, False)
    ref = out[2]
    assert ref[1].spikes > 1000
    for got in out[:2]:
        assert got[1].spikes == ref[1].spikes and got[1].events == ref[1].events
        assert all(np.array_equal(x, y) for x, y in zip(got[2], ref[2]))
        for k in fields:
            assert np.array_equal(got[3][k].view(np.uint32), ref[3][k].view(np.uint32)), k
        assert all(np.array_equal(x, y) for x, y in zip(got[4], ref[4]))


def test_binned_delivery_not_persistent(nsx):
    spec = configs.izh_stdp_network(n=4000, seed=3, record_raster=True, raster_capacity=1 << 22)
    print("case persistent with binned delivery", flush=True)
    with nsx.SimulationEngine.from_spec(spec) as eng:
        eng.set_delivery(abi.DELIVERY_BINNED, 0, 4096)
        eng.set_persistent(True)
        eng.step(1)
        assert eng.persistent == (True, False)
