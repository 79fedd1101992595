"""Checkpoints of device-resident simulations: CheckpointManager (src/io/checkpoint.rs:9-162)
with the reference's metadata file, the reference's `Network` JSON where it is feasible, and a
binary sidecar for the complete device state.

Files of checkpoint `name`:

* `{name}_meta.json` -- CheckpointMetadata with the reference's fields and names
  (checkpoint.rs:10-20), so listing and pruning behave alike.
* `{name}.json` -- the reference's pretty `Network` JSON (serde layout of network/mod.rs:36-48,
  NeuronArray neuron/mod.rs:127-143, SynapseState synapse/mod.rs:59-79), written when the
  network has at most NETWORK_JSON_MAX_SYNAPSES synapses (about 400 bytes of text per synapse:
  a 10^8-synapse network would be 40 GB of JSON).  CheckpointManager::load (checkpoint.rs:86-100)
  can read it; `load_network` reads it (or one the reference wrote) into a device engine.
* `{name}.nsx.json` -- the index of the exact checkpoint: configuration, sizes, step, array index.
* `{name}.nsx.npz` -- the arrays (numpy's uncompressed zip of .npy, loaded without pickle): CSR
  (u64 row pointers, u32 columns, f32 weights, u8 delays), every state field including the delay
  ring and the STDP trace records, the surrogate's parameters.  A checkpoint loaded from it
  resumes bit for bit.
"""
import dataclasses
import json
import math
import os
import time

import numpy as np

from . import abi
from .engine import SimulationEngine

FORMAT = "neuralsim_amd.checkpoint"
VERSION = 1
NETWORK_JSON_MAX_SYNAPSES = 2_000_000
NEVER_STEP = -(1 << 30)


@dataclasses.dataclass
class CheckpointMetadata:
    """CheckpointMetadata (checkpoint.rs:10-20)."""
    name: str
    sim_time_ms: float
    neuron_count: int
    synapse_count: int
    total_spikes: int
    timestamp: str
    seed: int

    def to_json(self):
        return json.dumps(dataclasses.asdict(self), indent=2)

    @classmethod
    def from_json(cls, text):
        d = json.loads(text)
        return cls(**{f.name: d[f.name] for f in dataclasses.fields(cls)})


def config_to_dict(cfg):
    out = {}
    for name, _ in abi.NsxConfig._fields_:
        v = getattr(cfg, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out


def config_from_dict(d):
    cfg = abi.NsxConfig()
    for name, _ in abi.NsxConfig._fields_:
        if name not in d:
            continue
        v = d[name]
        if isinstance(v, list):
            arr = getattr(cfg, name)
            for k, x in enumerate(v):
                arr[k] = x
        else:
            setattr(cfg, name, v)
    return cfg


def rust_duration_debug(ns):
    """Rust's `{:?}` of a Duration of `ns` nanoseconds, as chrono_now formats the time since the
    epoch (checkpoint.rs:164-171): seconds, the fraction without trailing zeros, then 's'."""
    secs, frac = divmod(int(ns), 1_000_000_000)
    if frac == 0:
        return f"{secs}s"
    return f"{secs}.{frac:09d}".rstrip("0") + "s"


def _timestamp():
    return rust_duration_debug(time.time_ns())


# ---- the reference's Network JSON ------------------------------------------------------------

def _f64(x):
    """A JSON number for an f64 as serde_json writes one (ryu: shortest round-trip digits, '.0'
    on integral values, exponents without '+' or leading zeros); non-finite values become null,
    as serde_json writes them."""
    x = float(x)
    if not math.isfinite(x):
        return "null"
    r = repr(x)
    if "e" in r:
        mant, ex = r.split("e")
        r = mant + "e" + str(int(ex))
    return r


def _model_params(cfg, c=None, indent=6):
    """NeuronModelParams of one neuron as serde's externally tagged enum (neuron/mod.rs:72-94),
    pretty-printed at `indent`."""
    if cfg.model == abi.MODEL_LIF:
        tag, fields = "Lif", (("resting", cfg.lif_v_rest), ("threshold", cfg.lif_v_th),
                              ("reset", cfg.lif_v_reset), ("tau_m", cfg.lif_tau_m),
                              ("refractory_period", cfg.lif_t_ref_ms),
                              ("input_resistance", cfg.lif_r))
    elif cfg.model == abi.MODEL_IZHIKEVICH:
        tag, fields = "Izhikevich", tuple(zip(("a", "b", "c", "d"), c))
    else:
        tag, fields = "HodgkinHuxley", (("g_na", cfg.hh_g_na), ("g_k", cfg.hh_g_k),
                                        ("g_l", cfg.hh_g_l), ("e_na", cfg.hh_e_na),
                                        ("e_k", cfg.hh_e_k), ("e_l", cfg.hh_e_l),
                                        ("c_m", cfg.hh_c_m))
    p0, p1, p2 = " " * indent, " " * (indent + 2), " " * (indent + 4)
    body = ",\n".join(f'{p2}"{k}": {_f64(v)}' for k, v in fields)
    return "{\n" + f'{p1}"{tag}": {{\n{body}\n{p1}}}\n{p0}}}'


def _array(f, key, values, fmt, indent, last=False):
    pad = " " * indent
    f.write(f'{pad}"{key}": [')
    if len(values) == 0:
        f.write("]" + ("" if last else ",") + "\n")
        return
    f.write("\n")
    inner = pad + "  "
    f.write(",\n".join(inner + fmt(v) for v in values))
    f.write(f"\n{pad}]" + ("" if last else ",") + "\n")


def write_network_json(path, cfg, csr, state, step, izh_cd=None, name="NeuralSim Network"):
    """Write the reference's `Network` (network/mod.rs:36-48) in serde_json's pretty layout
    (two-space indents, one element per line).  csr = (row_ptr, col, weight, delay_steps) in CSR slot order; state = the
    device state fields (v, u, m, h, n, refractory, last_spike_step, spike_count, pre/post trace
    records, ring); step = the next step index.  Synapses are written in CSR slot order, so
    synapses[k] is the synapse of adjacency_indices[k] (the order the reference's engine reads
    them in, engine.rs:420-428).  Mapping: time = step dt; last_spike_time = (spike step) dt, -1e9
    for none; input_current = the input arriving at the next step (ring slot `step`, both
    channels); just_spiked = spiked at step - 1; E->E synapses under pair STDP carry a StdpTrace
    whose values are the per-neuron traces decayed to `time` (every synapse of one pre / post
    neuron has the same trace, engine.rs:496-537), other synapses none; excitatory synapses are
    AMPA, inhibitory GABA."""
    rp, col, w, d = csr
    n = int(cfg.n_neurons)
    ne = int(cfg.n_exc)
    dt = float(cfg.dt_ms)
    t_now = step * dt
    v = state["v"].astype(np.float64)
    u = state.get("u", np.zeros(n, np.float32)).astype(np.float64)
    refr = state.get("refractory", np.zeros(n, np.int32))
    last = state.get("last_spike_step", np.full(n, NEVER_STEP, np.int32)).astype(np.int64)
    count = state.get("spike_count", np.zeros(n, np.uint32))
    is_hh = cfg.model in (abi.MODEL_HH, abi.MODEL_HH_SURROGATE)
    gm = state["m"].astype(np.float64) if is_hh else np.full(n, 0.05)
    gh = state["h"].astype(np.float64) if is_hh else np.full(n, 0.6)
    gn = state["n"].astype(np.float64) if is_hh else np.full(n, 0.32)
    last_t = np.where(last == NEVER_STEP, -1e9, last * dt)
    ring = state.get("ring")
    ch = abi.channels(cfg)
    if ring is not None:
        D = int(cfg.ring_slots)
        slot = ring.reshape(D, ch, n)[step % D].astype(np.float64) * 2.0 ** -int(cfg.frac_bits)
        inp = slot.sum(axis=0)
    else:
        inp = np.zeros(n)
    pair = cfg.plasticity == abi.PLAST_PAIR
    if pair:
        def now(vals, st, tau):
            vals = vals.astype(np.float64)
            st = st.astype(np.int64)
            return np.where(st == NEVER_STEP, 0.0, vals * np.exp(-(step - st) * dt / tau))
        pre_now = now(state["pre_trace"], state["pre_step"], cfg.tau_plus_ms)
        post_now = now(state["post_trace"], state["post_step"], cfg.tau_minus_ms)
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp.astype(np.int64)))
    with open(path, "w") as f:
        f.write("{\n")
        f.write(f'  "name": {json.dumps(name)},\n')
        f.write('  "neurons": {\n')
        _array(f, "membrane_potential", v, _f64, 4)
        _array(f, "recovery_variable", u, _f64, 4)
        _array(f, "refractory_counter", refr, lambda x: str(int(x)), 4)
        _array(f, "last_spike_time", last_t, _f64, 4)
        _array(f, "spike_count", count, lambda x: str(int(x)), 4)
        _array(f, "neuron_type", np.arange(n),
               lambda i: '"Excitatory"' if i < ne else '"Inhibitory"', 4)
        if cfg.model == abi.MODEL_IZHIKEVICH:
            cs, ds = izh_cd
            params = [_model_params(cfg, (cfg.izh_a[int(i >= ne)], cfg.izh_b[int(i >= ne)],
                                          cs[i], ds[i])) for i in range(n)]
        else:
            params = [_model_params(cfg)] * n
        _array(f, "model_params", params, lambda s: s, 4)
        _array(f, "input_current", inp, _f64, 4)
        _array(f, "is_output", np.zeros(n, bool), lambda x: "false", 4)
        _array(f, "hh_m", gm, _f64, 4)
        _array(f, "hh_h", gh, _f64, 4)
        _array(f, "hh_n", gn, _f64, 4)
        _array(f, "just_spiked", last == step - 1, lambda x: "true" if x else "false", 4,
               last=True)
        f.write("  },\n")
        _array(f, "adjacency_ptr", rp, lambda x: str(int(x)), 2)
        f.write('  "synapses": [')
        m = len(col)
        if m:
            f.write("\n")
        for k in range(m):
            i, j = int(src[k]), int(col[k])
            ee = pair and i < ne and j < ne
            trace = "null"
            cstate = "null"
            if ee:
                lp = -1e9 if last[i] == NEVER_STEP else (last[i] + 1) * dt
                lq = -1e9 if last[j] == NEVER_STEP else (last[j] + 1) * dt
                trace = ('{\n        "pre_trace": %s,\n        "post_trace": %s,\n'
                         '        "last_pre_spike": %s,\n        "last_post_spike": %s\n      }'
                         % (_f64(pre_now[i]), _f64(post_now[j]), _f64(lp), _f64(lq)))
                cstate = ('{\n        "g": 0.0,\n        "t_since_spike": 0.0,\n'
                          '        "pending_spike": false\n      }')
            f.write('    {\n      "source": %d,\n      "target": %d,\n      "weight": %s,\n'
                    '      "delay": %s,\n      "synapse_type": "%s",\n      "conductance": 0.0,\n'
                    '      "is_active": true,\n      "conductance_state": %s,\n'
                    '      "stdp_trace": %s,\n      "eligibility": null,\n      "stp": null,\n'
                    '      "plasticity_enabled": %s,\n      "contribution_avg": 0.0\n    }%s\n'
                    % (i, j, _f64(w[k]), _f64(float(d[k]) * dt), "AMPA" if i < ne else "GABA",
                       cstate, trace, "true" if ee else "false", "," if k + 1 < m else ""))
        f.write("  ],\n" if m else "],\n")
        _array(f, "adjacency_indices", col, lambda x: str(int(x)), 2)
        f.write(f'  "time": {_f64(t_now)},\n')
        _array(f, "neuron_region", np.zeros(n, np.int64), lambda x: "0", 2)
        f.write('  "region_names": [\n    "Default"\n  ]\n}')


def read_network_json(path):
    """Parse a `Network` JSON (the reference's serde layout) into arrays: the CSR in its slot
    order, per-neuron state and the model parameters of every neuron."""
    with open(path) as f:
        net = json.load(f)
    nr = net["neurons"]
    n = len(nr["membrane_potential"])
    syn = net["synapses"]
    rp = np.asarray(net["adjacency_ptr"], np.uint64)
    col = np.asarray(net["adjacency_indices"], np.uint32)
    if len(rp) != n + 1 or int(rp[-1]) != len(col) or len(syn) != len(col):
        raise ValueError(f"{path}: adjacency does not match the synapse list")
    w = np.array([s["weight"] for s in syn], np.float64)
    delay_ms = np.array([s["delay"] for s in syn], np.float64)
    tgt = np.array([s["target"] for s in syn], np.int64)
    if len(col) and not np.array_equal(tgt, col.astype(np.int64)):
        raise ValueError(f"{path}: synapses are not in CSR slot order")
    def none_f(a):
        return np.array([np.nan if x is None else x for x in a], np.float64)
    return {"n": n, "row_ptr": rp, "col": col, "weight": w, "delay_ms": delay_ms,
            "v": none_f(nr["membrane_potential"]), "u": none_f(nr["recovery_variable"]),
            "refractory": np.asarray(nr["refractory_counter"], np.int32),
            "last_spike_time": none_f(nr["last_spike_time"]),
            "spike_count": np.asarray(nr["spike_count"], np.uint64),
            "neuron_type": nr["neuron_type"], "model_params": nr["model_params"],
            "m": none_f(nr["hh_m"]), "h": none_f(nr["hh_h"]), "n_gate": none_f(nr["hh_n"]),
            "time": float(net["time"]), "name": net.get("name", "")}


def engine_from_network(netd, base_cfg):
    """A device engine holding a parsed `Network`: the network's CSR, weights, delays (ms ->
    steps of base_cfg.dt_ms, at least 1) and neuron state, with base_cfg's step, drive and
    plasticity settings.  The device keeps excitatory neurons first: a network whose types are
    interleaved is refused.  Izhikevich c/d come from each neuron's own parameters."""
    n = netd["n"]
    types = netd["neuron_type"]
    ne = sum(1 for t in types if t == "Excitatory")
    if any(t != "Excitatory" for t in types[:ne]):
        raise ValueError("the device engine needs the excitatory neurons first")
    cfg = base_cfg.copy()
    cfg.n_neurons = n
    cfg.n_exc = ne
    dt = float(cfg.dt_ms)
    steps = np.maximum(1, np.rint(netd["delay_ms"] / dt)).astype(np.int64)
    if len(steps) and int(steps.max()) >= int(cfg.ring_slots):
        cfg.ring_slots = int(steps.max()) + 1
    eng = SimulationEngine(cfg)
    try:
        eng.upload_csr(netd["row_ptr"], netd["col"], netd["weight"].astype(np.float32),
                       steps.astype(np.uint8))
        step = int(round(netd["time"] / dt))
        last = np.where(netd["last_spike_time"] <= -1e8, NEVER_STEP,
                        np.rint(netd["last_spike_time"] / dt)).astype(np.int32)
        state = {"v": netd["v"].astype(np.float32), "refractory": netd["refractory"],
                 "last_spike_step": last,
                 "spike_count": netd["spike_count"].astype(np.uint32)}
        if cfg.model == abi.MODEL_IZHIKEVICH:
            state["u"] = netd["u"].astype(np.float32)
            cd = [p["Izhikevich"] for p in netd["model_params"]]
            state["izh_c"] = np.array([p["c"] for p in cd], np.float32)
            state["izh_d"] = np.array([p["d"] for p in cd], np.float32)
        if cfg.model in (abi.MODEL_HH, abi.MODEL_HH_SURROGATE):
            state.update(m=netd["m"].astype(np.float32), h=netd["h"].astype(np.float32),
                         n=netd["n_gate"].astype(np.float32))
        eng.upload_state({k: v for k, v in state.items() if k in abi.STATE_FIELDS})
        eng.set_step(step)
    except Exception:
        eng.close()
        raise
    return eng


class CheckpointManager:
    """CheckpointManager (checkpoint.rs:22-162): a directory of checkpoints, an auto-save
    interval in simulated ms and a cap on the number kept (the oldest by simulated time go)."""

    def __init__(self, checkpoint_dir):
        self.checkpoint_dir = str(checkpoint_dir)
        os.makedirs(self.checkpoint_dir, exist_ok=True)
        self.auto_save_interval_ms = 1000.0
        self._last_save_time = 0.0
        self.max_checkpoints = 10
        self.network_json_max_synapses = NETWORK_JSON_MAX_SYNAPSES

    def _paths(self, name):
        d = self.checkpoint_dir
        return {"network": os.path.join(d, f"{name}.json"),
                "meta": os.path.join(d, f"{name}_meta.json"),
                "index": os.path.join(d, f"{name}.nsx.json"),
                "sidecar": os.path.join(d, f"{name}.nsx.npz")}

    def save(self, engine, name):
        """Save an engine's network and complete state (CheckpointManager::save / save_engine,
        checkpoint.rs:47-84); returns the path of the index `{name}.nsx.json`."""
        paths = self._paths(name)
        rp, col, w, d = engine.read_csr()
        state = engine.read_state(ring=True)
        step = engine.get_step()
        arrays = {"row_ptr": rp, "col": col, "weight": w, "delay": d,
                  "step": np.array([step], np.uint64)}
        for k, a in state.items():
            arrays["state_" + k] = a
        if engine.cfg.model == abi.MODEL_HH_SURROGATE:
            params, norm = engine.surrogate_params()
            arrays["surrogate_params"] = params
            arrays["surrogate_norm"] = norm
        totals = engine.step(0)
        meta = CheckpointMetadata(name=name, sim_time_ms=float(step) * float(engine.cfg.dt_ms),
                                  neuron_count=int(engine.cfg.n_neurons), synapse_count=len(col),
                                  total_spikes=int(totals.total_spikes), timestamp=_timestamp(),
                                  seed=int(engine.cfg.seed))
        with open(paths["meta"], "w") as f:
            f.write(meta.to_json())
        if len(col) <= self.network_json_max_synapses:
            izh = None
            if engine.cfg.model == abi.MODEL_IZHIKEVICH:
                cd = engine.read_state(["izh_c", "izh_d"])
                izh = (cd["izh_c"], cd["izh_d"])
            write_network_json(paths["network"], engine.cfg, (rp, col, w, d), state, step, izh)
        elif os.path.exists(paths["network"]):
            os.remove(paths["network"])  # a stale network of an earlier save under this name
        np.savez(paths["sidecar"], **arrays)
        doc = {"format": FORMAT, "version": VERSION, "name": name,
               "sidecar": os.path.basename(paths["sidecar"]), "step": int(step),
               "neuron_count": meta.neuron_count, "synapse_count": meta.synapse_count,
               "network_json": os.path.exists(paths["network"]),
               "config": config_to_dict(engine.cfg),
               "arrays": {k: {"dtype": str(a.dtype), "shape": list(a.shape)}
                          for k, a in arrays.items()}}
        with open(paths["index"], "w") as f:
            json.dump(doc, f, indent=2)
        self.prune()
        return paths["index"]

    save_engine = save

    def load(self, name):
        """Restore a checkpoint into a new engine, exactly (CheckpointManager::load,
        checkpoint.rs:86-100); returns (engine, metadata)."""
        paths = self._paths(name)
        if not os.path.exists(paths["index"]):
            raise ValueError(f"{paths['index']}: no exact checkpoint under this name (a "
                             f"reference Network JSON loads with load_network)")
        with open(paths["index"]) as f:
            doc = json.load(f)
        if doc.get("format") != FORMAT or doc.get("version") != VERSION:
            raise ValueError(f"{paths['index']}: not a {FORMAT} v{VERSION} checkpoint")
        with open(paths["meta"]) as f:
            meta = CheckpointMetadata.from_json(f.read())
        side = os.path.join(self.checkpoint_dir, doc["sidecar"])
        with np.load(side, allow_pickle=False) as z:
            arrays = {k: z[k] for k in z.files}
        for k, spec in doc["arrays"].items():
            a = arrays[k]
            if str(a.dtype) != spec["dtype"] or list(a.shape) != spec["shape"]:
                raise ValueError(f"{side}: array {k} does not match {paths['index']}")
        eng = SimulationEngine(config_from_dict(doc["config"]))
        try:
            eng.upload_csr(arrays["row_ptr"], arrays["col"], arrays["weight"], arrays["delay"])
            if "surrogate_params" in arrays:
                eng.set_surrogate_params(arrays["surrogate_params"], arrays["surrogate_norm"])
            eng.upload_state({k[len("state_"):]: a for k, a in arrays.items()
                              if k.startswith("state_")})
            eng.set_step(int(arrays["step"][0]))
        except Exception:
            eng.close()
            raise
        return eng, meta

    def load_network(self, name, base_cfg):
        """Load the `Network` JSON `{name}.json` (written here or by the reference) into a new
        engine with base_cfg's step, drive and plasticity; returns (engine, metadata).  The
        delay ring's in-flight events and the STDP records beyond the traces are not part of a
        Network, so the run continues from the network's state, not bit for bit."""
        paths = self._paths(name)
        with open(paths["meta"]) as f:
            meta = CheckpointMetadata.from_json(f.read())
        return engine_from_network(read_network_json(paths["network"]), base_cfg), meta

    def should_auto_save(self, current_time):
        return current_time - self._last_save_time >= self.auto_save_interval_ms

    def mark_saved(self, current_time):
        self._last_save_time = current_time

    def list_checkpoints(self):
        """Metadata of every checkpoint in the directory, latest simulated time first
        (checkpoint.rs:112-139; found by their metadata files, since a large checkpoint has no
        `{name}.json`)."""
        out = []
        for fn in os.listdir(self.checkpoint_dir):
            if not fn.endswith("_meta.json"):
                continue
            try:
                with open(os.path.join(self.checkpoint_dir, fn)) as f:
                    out.append(CheckpointMetadata.from_json(f.read()))
            except (OSError, ValueError, KeyError, TypeError):
                continue
        out.sort(key=lambda m: m.sim_time_ms, reverse=True)
        return out

    def prune(self):
        """Keep the max_checkpoints latest by simulated time (checkpoint.rs:141-161)."""
        cps = self.list_checkpoints()
        if len(cps) <= self.max_checkpoints:
            return
        cps.sort(key=lambda m: m.sim_time_ms)
        for m in cps[: len(cps) - self.max_checkpoints]:
            for p in self._paths(m.name).values():
                try:
                    os.remove(p)
                except OSError:
                    pass
