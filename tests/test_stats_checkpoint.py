"""CPU checks of the statistics recorder and the checkpoint manager's bookkeeping: the
reference's own tests (src/io/checkpoint.rs:257-310) restated, Rust's number formatting in the
CSV against hand-worked cases, the recorder's interval and eviction, metadata round trips,
listing order and pruning by simulated time."""
import json
import os

import pytest

from neuralsim_amd import checkpoint, configs
from neuralsim_amd.stats import CSV_HEADER, SimulationStats, StatsRecorder, rust_display


def test_stats_recorder_reference_test():
    """checkpoint.rs:281-291 test_stats_recorder."""
    rec = StatsRecorder(100, 1)
    stats = SimulationStats(sim_time_ms=10.0, total_spikes=5, mean_firing_rate=2.5)
    rec.record(1, stats, 0.1)
    assert len(rec.records()) == 1
    assert rec.to_csv() == CSV_HEADER + "1,10,5,0,2.5,0,0,0,0.1\n"


def test_rust_display_formatting():
    """Rust's `{}` for f64: shortest round-trip digits, plain notation, no trailing '.0'."""
    cases = {1.0: "1", 0.1: "0.1", 2.5: "2.5", -0.0: "-0", 1e-7: "0.0000001",
             1e21: "1000000000000000000000", 123456789.125: "123456789.125",
             0.30000000000000004: "0.30000000000000004", float("nan"): "NaN",
             float("inf"): "inf", float("-inf"): "-inf", 7: "7", 2**64 - 1: "18446744073709551615"}
    for x, want in cases.items():
        assert rust_display(x) == want, (x, rust_display(x))


def test_stats_recorder_interval_and_eviction():
    rec = StatsRecorder(3, 10)
    for step in range(0, 61):
        rec.record(step, SimulationStats(sim_time_ms=step * 0.5, total_spikes=step), step / 4)
    steps = [r.step for r in rec.records()]
    assert steps == [40, 50, 60]  # every 10th step, the oldest evicted beyond 3
    rows = rec.to_csv().splitlines()
    assert rows[0] + "\n" == CSV_HEADER and rows[1] == "40,20,40,0,0,0,0,0,10"


def test_metadata_round_trip():
    """checkpoint.rs:262-279 test_checkpoint_metadata."""
    m = checkpoint.CheckpointMetadata(name="test", sim_time_ms=100.0, neuron_count=1000,
                                      synapse_count=50000, total_spikes=42, timestamp="test",
                                      seed=42)
    d = json.loads(m.to_json())
    assert list(d) == ["name", "sim_time_ms", "neuron_count", "synapse_count", "total_spikes",
                       "timestamp", "seed"]  # the reference's field names, in its order
    r = checkpoint.CheckpointMetadata.from_json(m.to_json())
    assert r == m and r.neuron_count == 1000 and r.seed == 42


def _fake(mgr, name, t):
    meta = checkpoint.CheckpointMetadata(name, t, 10, 20, 0, "x", 1)
    with open(os.path.join(mgr.checkpoint_dir, f"{name}_meta.json"), "w") as f:
        f.write(meta.to_json())
    for suffix in (".json", ".nsx.npz"):
        with open(os.path.join(mgr.checkpoint_dir, name + suffix), "w") as f:
            f.write("{}")


def test_list_and_prune_by_simulated_time(tmp_path):
    mgr = checkpoint.CheckpointManager(tmp_path / "cp")
    mgr.max_checkpoints = 3
    for k, t in enumerate([500.0, 100.0, 400.0, 200.0, 300.0]):
        _fake(mgr, f"c{k}", t)
    assert [m.sim_time_ms for m in mgr.list_checkpoints()] == [500, 400, 300, 200, 100]
    mgr.prune()
    assert sorted(m.name for m in mgr.list_checkpoints()) == ["c0", "c2", "c4"]
    left = sorted(os.listdir(mgr.checkpoint_dir))
    assert "c1.json" not in left and "c1_meta.json" not in left and "c1.nsx.npz" not in left


def test_auto_save_interval(tmp_path):
    mgr = checkpoint.CheckpointManager(tmp_path)
    assert mgr.auto_save_interval_ms == 1000.0 and mgr.max_checkpoints == 10
    assert not mgr.should_auto_save(999.0) and mgr.should_auto_save(1000.0)
    mgr.mark_saved(1000.0)
    assert not mgr.should_auto_save(1999.0) and mgr.should_auto_save(2000.0)


def test_config_round_trip():
    cfg = configs.izh_stdp_network(n=1000, rule="triplet").cfg
    back = checkpoint.config_from_dict(json.loads(json.dumps(checkpoint.config_to_dict(cfg))))
    assert bytes(back) == bytes(cfg)


def test_load_rejects_foreign_documents(tmp_path):
    mgr = checkpoint.CheckpointManager(tmp_path)
    _fake(mgr, "x", 1.0)
    with pytest.raises(ValueError):
        mgr.load("x")
