"""Epoch stepping on the device, through the C ABI: in networks without plasticity whose shortest
delay spans d_min >= 2 steps, up to d_min steps run as one update launch (state in registers)
followed by one synapse pass for all their spikes.  Every run here equals the same run stepped one
launch per step, bit for bit -- raster, every state field, the delay ring, counters and the LFP --
and the point-model run equals the oracle's FP32 mirror:

* LIF (delays 1-5 ms at dt 0.1 ms: epochs of 10 steps), direct launches and graph windows whose
  length is not a multiple of the epoch;
* plain HH at dt 0.01 ms (delays 1-2 ms: epochs of 100 steps) and at dt 0.05 ms with five fused
  sub-steps (epochs of 20), with statistics on;
* the HH surrogate (epochs of 20), fallback counts included;
* a batch of LIF networks;
* when epochs apply: plasticity, binned delivery and nsx_set_epoch(1) step one step at a time.
"""
import numpy as np
import pytest

from neuralsim_amd import abi, configs
from oracle import mlp

pytestmark = pytest.mark.gpu

FIELDS = ["v", "u", "m", "h", "n", "syn0", "syn1", "syn2", "drive", "refractory",
          "last_spike_step", "spike_count"]


def _run(nsx, spec, steps, epoch, graph=0, stats=False, setup=None):
    with nsx.SimulationEngine.from_spec(spec) as eng:
        if setup:
            setup(eng)
        eng.set_epoch(epoch)
        if stats:
            eng.set_stats(True)
        eng.set_graph_steps(graph)
        e = eng.epoch
        st = eng.step(steps)
        out = {"epoch": e, "stats": st, "raster": eng.raster(),
               "state": eng.read_state(FIELDS, ring=True)}
        if stats:
            out["lfp"] = eng.lfp()
            out["sim"] = eng.stats()
        if spec.cfg.model == abi.MODEL_HH_SURROGATE:
            out["fb"] = eng.surrogate_fallbacks()
        return out


def _same(a, b):
    assert a["stats"].spikes == b["stats"].spikes and a["stats"].events == b["stats"].events
    assert all(np.array_equal(x, y) for x, y in zip(a["raster"], b["raster"]))
    for k in a["state"]:
        assert np.array_equal(np.ascontiguousarray(a["state"][k]).view(np.uint8),
                              np.ascontiguousarray(b["state"][k]).view(np.uint8)), k


@pytest.mark.parametrize("graph", [0, 25])
def test_lif_epochs_equal_per_step_and_mirror(nsx, orc, graph):
    spec = configs.lif_network(n=3000, seed=9, drive_scale=60.0, record_raster=True,
                               raster_capacity=1 << 22)
    steps = 737
    print(f"case epoch lif {spec.name} graph {graph}", flush=True)
    a = _run(nsx, spec, steps, 128, graph)
    b = _run(nsx, spec, steps, 1, graph)
    assert a["epoch"] == 10 and b["epoch"] == 1
    assert a["stats"].spikes > 1000
    _same(a, b)
    mir = orc.MirrorNet.from_spec(spec)
    ms, mi, mev = mir.run(steps)
    mst = mir.read_state(["v", "syn0", "syn1", "refractory", "spike_count"], ring=True)
    mir.close()
    assert np.array_equal(a["raster"][0], ms) and np.array_equal(a["raster"][1], mi)
    assert a["stats"].events == mev
    for k in mst:
        assert np.array_equal(np.ascontiguousarray(a["state"][k]).view(np.uint8),
                              np.ascontiguousarray(mst[k]).view(np.uint8)), k


@pytest.mark.parametrize("dt,sub,expect", [(0.01, 1, 100), (0.05, 5, 20)])
def test_hh_epochs_equal_per_step(nsx, dt, sub, expect):
    spec = configs.hh_network(n=2000, seed=13, dt_ms=dt, hh_substeps=sub, record_raster=True,
                              raster_capacity=1 << 22)
    steps = 1003 if dt == 0.01 else 401
    print(f"case epoch hh dt {dt} x{sub}", flush=True)
    a = _run(nsx, spec, steps, 128, 50, stats=True)
    b = _run(nsx, spec, steps, 1, 50, stats=True)
    assert a["epoch"] == expect and b["epoch"] == 1
    assert a["stats"].spikes > 100
    _same(a, b)
    assert a["lfp"] == b["lfp"] and a["sim"].mean_firing_rate == b["sim"].mean_firing_rate


def test_surrogate_epochs_equal_per_step(nsx):
    spec = configs.hh_surrogate_network(n=1500, seed=4, record_raster=True,
                                        raster_capacity=1 << 22)
    print("case epoch surrogate", flush=True)
    with nsx.SimulationEngine.from_spec(configs.hh_surrogate_network(n=512, k_in=1)) as g:
        g.surrogate_generate(300_000, 3)
        norm = g.surrogate_norm()
    params = mlp.init_params(11)

    def setup(eng):
        eng.set_surrogate_params(params, norm)

    a = _run(nsx, spec, 330, 128, 0, setup=setup)
    b = _run(nsx, spec, 330, 1, 0, setup=setup)
    assert a["epoch"] == 20 and b["epoch"] == 1
    _same(a, b)
    assert a["fb"] == b["fb"]


def test_lif_batch_epochs_equal_per_step(nsx):
    base = configs.lif_network(n=2, seed=1, drive_scale=60.0, record_raster=True,
                               raster_capacity=1 << 22).cfg
    nets = configs.population_nets(5, seed=23, n_range=(300, 3000), k_range=(10, 120),
                                   w_exc=(0.0, 0.5), w_inh=(-2.0, 0.0), delays=(10, 50))
    out = []
    for epoch in (128, 1):
        with nsx.BatchedPopulation(base, nets) as pop:
            pop.set_epoch(epoch)
            pop.set_graph_steps(50)
            e = pop.epoch
            st = pop.step(500)
            out.append((e, st, pop.raster(), pop.read_state(["v", "syn0", "syn1"], ring=True),
                        pop.net_stats()))
    (ea, sa, ra, xa, na), (eb, sb, rb, xb, nb) = out
    assert ea == 10 and eb == 1 and sa.spikes == sb.spikes > 100 and sa.events == sb.events
    assert all(np.array_equal(x, y) for x, y in zip(ra, rb))
    for k in xa:
        assert np.array_equal(xa[k].view(np.uint8), xb[k].view(np.uint8)), k
    assert all(np.array_equal(x, y) for x, y in zip(na, nb))


def test_epoch_applicability(nsx):
    lif = configs.lif_network(n=1000, seed=2)
    with nsx.SimulationEngine.from_spec(lif) as eng:
        assert eng.epoch == 10
        eng.set_epoch(4)
        assert eng.epoch == 4
        eng.set_epoch(1)
        assert eng.epoch == 1
        eng.set_epoch(0)
        assert eng.epoch == 1
        eng.set_epoch(128)
        eng.set_delivery(abi.DELIVERY_BINNED)
        eng.step(1)
        assert eng.epoch == 1
    with nsx.SimulationEngine.from_spec(configs.lif_network(n=1000, seed=2, plasticity=True)) as e:
        assert e.epoch == 1  # plasticity: one step per launch
    with nsx.SimulationEngine.from_spec(configs.izh_stdp_network(n=1000, plasticity=False)) as e:
        assert e.epoch == 1  # shortest delay one step
