"""The autotuner measures launch shapes from a snapshot and restores it: a run that is tuned
midway is bit-identical to an untuned run (raster, state, STDP weights)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_leaves_simulation_unchanged(nsx):
    spec = nsx.configs.izh_stdp_network(n=6000, seed=12, record_raster=True,
                                        raster_capacity=1 << 22)
    fields = ["v", "u", "pre_trace", "post_trace", "post_step", "spike_count"]
    print("autotune case izh_stdp 6000", flush=True)
    with nsx.SimulationEngine.from_spec(spec) as a:
        a.step(120)
        split, graph, trials = a.autotune(40)
        print("chosen", split, graph, trials)
        # 5 splits, 4 more windows, then atomic vs binned with two tile sizes
        assert len(trials) == 11 and split in (1, 2, 4, 8, 16) and graph in (10, 25, 50, 100, 200)
        assert all(t[2] > 0 for t in trials)
        assert sorted(t[3] for t in trials[-3:]) == [0, 8192, 16384]
        assert a.delivery == (1 if a.last_tune.tile_entries else 0)
        assert a.get_step() == 120
        a.step(130)
        ra, sa, wa = a.raster(), a.read_state(fields), a.weights()
    with nsx.SimulationEngine.from_spec(spec) as b:
        b.step(250)
        rb, sb, wb = b.raster(), b.read_state(fields), b.weights()
    assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1])
    for f in fields:
        assert np.array_equal(sa[f].view(np.uint32), sb[f].view(np.uint32)), f
    assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))
