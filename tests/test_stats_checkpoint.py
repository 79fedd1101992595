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
    for suffix in (".json", ".nsx.json", ".nsx.npz"):
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
    assert "c1.nsx.json" not in left


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


def test_timestamp_matches_rust_duration_debug():
    """chrono_now formats the Duration since the epoch with {:?} (checkpoint.rs:164-171): the
    fraction without trailing zeros, no fraction for whole seconds."""
    f = checkpoint.rust_duration_debug
    assert f(1_733_280_000_500_000_000) == "1733280000.5s"
    assert f(1_733_280_000_000_000_000) == "1733280000s"
    assert f(1_733_280_000_123_456_789) == "1733280000.123456789s"
    assert f(1_733_280_000_000_000_010) == "1733280000.00000001s"
    assert f(1_500_000) == "0.0015s"


def _small_network():
    import numpy as np
    spec = configs.izh_stdp_network(n=6, k_in=2)
    cfg = spec.cfg
    rp = np.array([0, 2, 3, 5, 5, 7, 8], np.uint64)
    col = np.array([1, 4, 0, 1, 5, 0, 2, 3], np.uint32)
    w = np.array([0.5, 1.25, 2.0, 0.0, -1.5, 3.0, -4.0, 6.0], np.float32)
    d = np.array([1, 2, 3, 1, 5, 20, 1, 2], np.uint8)
    state = {"v": np.array([-65.0, -60.5, 30.5, -70.0, -64.0, -65.0], np.float32),
             "u": np.array([-13.0, -12.0, -1.0, 0.0, 2.0, -3.0], np.float32),
             "refractory": np.zeros(6, np.int32),
             "last_spike_step": np.array([-(1 << 30), 9, 4, -(1 << 30), 9, 2], np.int32),
             "spike_count": np.array([0, 3, 1, 0, 2, 5], np.uint32),
             "pre_trace": np.array([0, 1.5, 1.0, 0, 1.2, 1.0], np.float32),
             "pre_step": np.array([-(1 << 30), 9, 4, -(1 << 30), 9, 2], np.int32),
             "post_trace": np.array([0, 1.0, 1.0, 0, 1.0, 1.0], np.float32),
             "post_step": np.array([-(1 << 30), 9, 4, -(1 << 30), 9, 2], np.int32),
             "ring": np.zeros(21 * 6, np.int64)}
    state["ring"][(10 % 21) * 6 + 3] = 3 << 31  # 1.5 arriving at step 10 for neuron 3
    cd = (np.array([-65, -60, -55, -50, -65, -65], np.float32),
          np.array([8, 7, 6, 5, 2, 2], np.float32))
    return cfg, (rp, col, w, d), state, cd


def test_network_json_layout_and_round_trip(tmp_path):
    """The reference's Network JSON (network/mod.rs:36-48): field names and order of Network,
    NeuronArray and SynapseState, serde's externally tagged enums, synapses in CSR slot order so
    synapses[k].target == adjacency_indices[k]; parsed back to the same arrays."""
    import numpy as np
    cfg, csr, state, cd = _small_network()
    path = tmp_path / "net.json"
    checkpoint.write_network_json(path, cfg, csr, state, 10, cd)
    net = json.loads(path.read_text())
    assert list(net) == ["name", "neurons", "adjacency_ptr", "synapses", "adjacency_indices",
                         "time", "neuron_region", "region_names"]
    assert list(net["neurons"]) == ["membrane_potential", "recovery_variable",
                                    "refractory_counter", "last_spike_time", "spike_count",
                                    "neuron_type", "model_params", "input_current", "is_output",
                                    "hh_m", "hh_h", "hh_n", "just_spiked"]
    s0 = net["synapses"][0]
    assert list(s0) == ["source", "target", "weight", "delay", "synapse_type", "conductance",
                        "is_active", "conductance_state", "stdp_trace", "eligibility", "stp",
                        "plasticity_enabled", "contribution_avg"]
    assert net["time"] == 10.0 and net["region_names"] == ["Default"]
    nr = net["neurons"]
    assert nr["neuron_type"] == ["Excitatory"] * 4 + ["Inhibitory"] * 2
    assert nr["model_params"][1] == {"Izhikevich": {"a": float(np.float32(0.02)),
                                                    "b": float(np.float32(0.2)),
                                                    "c": -60.0, "d": 7.0}}
    assert nr["input_current"][3] == 1.5 and nr["just_spiked"] == [False, True, False, False,
                                                                   True, False]
    assert nr["last_spike_time"][0] == -1e9 and nr["last_spike_time"][1] == 9.0
    syn = net["synapses"]
    assert [s["target"] for s in syn] == net["adjacency_indices"] == csr[1].tolist()
    assert [s["source"] for s in syn] == [0, 0, 1, 2, 2, 4, 4, 5]
    assert [s["synapse_type"] for s in syn][:5] == ["AMPA"] * 5
    assert syn[5]["synapse_type"] == "GABA"
    assert [s["delay"] for s in syn] == [1.0, 2.0, 3.0, 1.0, 5.0, 20.0, 1.0, 2.0]
    # E->E synapses carry a StdpTrace (pair STDP on E->E only), others none
    ee = [s["plasticity_enabled"] for s in syn]
    assert ee == [True, False, True, True, False, False, False, False]
    assert syn[0]["stdp_trace"]["last_pre_spike"] == -1e9
    assert syn[2]["stdp_trace"]["last_pre_spike"] == 10.0  # spike at step 9, time after it
    assert syn[1]["stdp_trace"] is None and syn[1]["conductance_state"] is None
    back = checkpoint.read_network_json(path)
    assert np.array_equal(back["row_ptr"], csr[0]) and np.array_equal(back["col"], csr[1])
    assert np.array_equal(back["weight"].astype(np.float32), csr[2])
    assert np.array_equal(back["v"].astype(np.float32), state["v"])
    assert np.array_equal(back["spike_count"], state["spike_count"].astype(np.uint64))


def test_network_json_refuses_mismatched_adjacency(tmp_path):
    cfg, csr, state, cd = _small_network()
    path = tmp_path / "net.json"
    checkpoint.write_network_json(path, cfg, csr, state, 10, cd)
    net = json.loads(path.read_text())
    net["adjacency_indices"][0] = 5
    path.write_text(json.dumps(net))
    with pytest.raises(ValueError):
        checkpoint.read_network_json(path)
