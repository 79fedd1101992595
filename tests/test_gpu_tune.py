"""The autotuner measures launch shapes from a snapshot and restores it: a run that is tuned
midway is bit-identical to an untuned run (raster, state, STDP weights).  The shapes it chooses
among -- row-split threshold for high-degree rows, synapse workgroup size, propagation split --
leave every result bit-identical, under pair and triplet STDP and in epoch stepping.  Persistent
stepping is among the trials when ring atomics win."""
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
        # 5 splits; row-split thresholds, workgroup sizes and update widths under atomic
        # delivery; 4 windows; the two binned tile sizes; the split again if binned won;
        # persistent stepping (one or two windows) if ring atomics won
        tt = a.tune_trials()
        n_ps = sum(t["persistent"] for t in tt)
        assert len(trials) - n_ps in (18, 22) and split in (1, 2, 4, 8, 16)
        assert n_ps in (1, 2) if a.delivery == 0 else n_ps == 0
        assert a.persistent[0] == bool(a.last_tune.persistent)
        assert graph in (10, 25, 50, 100, 200)
        assert all(t[2] > 0 for t in trials)
        assert {t["row_split"] for t in tt} >= {0, 256, 128, 64}
        assert {t["syn_block"] for t in tt} >= {128, 256, 512}
        assert {t["width"] for t in tt} >= {1, 4}
        assert {t["tile"] for t in tt} == {0, 8192, 16384}
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


@pytest.mark.parametrize("rule", ["pair", "triplet"])
def test_row_split_and_workgroup_are_result_neutral(nsx, rule):
    spec = nsx.configs.izh_stdp_network(n=5000, seed=21, record_raster=True,
                                        raster_capacity=1 << 22, rule=rule)
    fields = ["v", "u", "pre_trace", "post_trace", "post_step", "spike_count"]
    out = []
    for rs, split, nt in [(0, 4, 256), (16, 1, 256), (64, 4, 128), (32, 8, 512), (16, 16, 64)]:
        print(f"case row split {rule} rs={rs} split={split} nt={nt}", flush=True)
        with nsx.SimulationEngine.from_spec(spec) as e:
            e.set_row_split(rs)
            e.set_prop_split(split)
            e.set_syn_block(nt)
            e.set_graph_steps(25)
            st = e.step(250)
            out.append((st.events, e.raster(), e.read_state(fields), e.weights()))
    ev0, r0, s0, w0 = out[0]
    assert ev0 > 10000
    for ev, r, st, w in out[1:]:
        assert ev == ev0
        assert np.array_equal(r[0], r0[0]) and np.array_equal(r[1], r0[1])
        for f in fields:
            assert np.array_equal(st[f].view(np.uint32), s0[f].view(np.uint32)), f
        assert np.array_equal(w.view(np.uint32), w0.view(np.uint32))


def test_row_split_in_epochs(nsx):
    spec = nsx.configs.lif_network(n=3000, seed=9, drive_scale=60.0, record_raster=True,
                                   raster_capacity=1 << 22)
    out = []
    for rs in (0, 16):
        with nsx.SimulationEngine.from_spec(spec) as e:
            e.set_row_split(rs)
            assert e.epoch == 10
            st = e.step(400)
            out.append((st.events, e.raster(), e.read_state(["v", "syn0", "syn1"], ring=True)))
    assert out[0][0] == out[1][0] > 1000
    assert all(np.array_equal(x, y) for x, y in zip(out[0][1], out[1][1]))
    for k in out[0][2]:
        assert np.array_equal(out[0][2][k].view(np.uint8), out[1][2][k].view(np.uint8)), k
