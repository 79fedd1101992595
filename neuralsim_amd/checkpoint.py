"""Checkpoints of device-resident simulations: CheckpointManager (src/io/checkpoint.rs:9-162)
with the reference's metadata file and a binary sidecar for the state.

The reference writes `{name}.json` (the whole Network through serde, pretty JSON) and
`{name}_meta.json` (CheckpointMetadata).  At 10^8 neurons and 10^10 synapses a JSON network is
out of the question, so here `{name}.json` is a small JSON document (configuration, sizes,
array index) and the arrays -- CSR (u64 row pointers, u32 columns, f32 weights, u8 delays), every
state field including the delay ring and the STDP trace records, the surrogate's parameters --
go to `{name}.nsx.npz` (numpy's uncompressed zip of .npy arrays, loaded without pickle).
`{name}_meta.json` has the reference's fields and names, so list_checkpoints / prune behave
alike.  A loaded checkpoint resumes bit for bit: the next steps equal those of the run that was
saved.
"""
import dataclasses
import json
import os
import time

import numpy as np

from . import abi
from .engine import SimulationEngine

FORMAT = "neuralsim_amd.checkpoint"
VERSION = 1


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


def _timestamp():
    # the reference formats a Duration since the epoch with {:?} (checkpoint.rs:164-171)
    ns = time.time_ns()
    return f"{ns // 1_000_000_000}.{ns % 1_000_000_000:09d}s"


class CheckpointManager:
    """CheckpointManager (checkpoint.rs:22-162): a directory of checkpoints, an auto-save
    interval in simulated ms and a cap on the number kept (the oldest by simulated time go)."""

    def __init__(self, checkpoint_dir):
        self.checkpoint_dir = str(checkpoint_dir)
        os.makedirs(self.checkpoint_dir, exist_ok=True)
        self.auto_save_interval_ms = 1000.0
        self._last_save_time = 0.0
        self.max_checkpoints = 10

    def _paths(self, name):
        d = self.checkpoint_dir
        return (os.path.join(d, f"{name}.json"), os.path.join(d, f"{name}_meta.json"),
                os.path.join(d, f"{name}.nsx.npz"))

    def save(self, engine, name):
        """Save an engine's network and complete state (CheckpointManager::save / save_engine,
        checkpoint.rs:47-84); returns the path of `{name}.json`."""
        path, meta_path, side_path = self._paths(name)
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
        with open(meta_path, "w") as f:
            f.write(meta.to_json())
        np.savez(side_path, **arrays)
        doc = {"format": FORMAT, "version": VERSION, "name": name,
               "sidecar": os.path.basename(side_path), "step": int(step),
               "neuron_count": meta.neuron_count, "synapse_count": meta.synapse_count,
               "config": config_to_dict(engine.cfg),
               "arrays": {k: {"dtype": str(a.dtype), "shape": list(a.shape)}
                          for k, a in arrays.items()}}
        with open(path, "w") as f:
            json.dump(doc, f, indent=2)
        self.prune()
        return path

    save_engine = save

    def load(self, name):
        """Restore a checkpoint into a new engine (CheckpointManager::load, checkpoint.rs:86-100);
        returns (engine, metadata)."""
        path, meta_path, _ = self._paths(name)
        with open(path) as f:
            doc = json.load(f)
        if doc.get("format") != FORMAT or doc.get("version") != VERSION:
            raise ValueError(f"{path}: not a {FORMAT} v{VERSION} checkpoint")
        with open(meta_path) as f:
            meta = CheckpointMetadata.from_json(f.read())
        side = os.path.join(self.checkpoint_dir, doc["sidecar"])
        with np.load(side, allow_pickle=False) as z:
            arrays = {k: z[k] for k in z.files}
        for k, spec in doc["arrays"].items():
            a = arrays[k]
            if str(a.dtype) != spec["dtype"] or list(a.shape) != spec["shape"]:
                raise ValueError(f"{side}: array {k} does not match {path}")
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

    def should_auto_save(self, current_time):
        return current_time - self._last_save_time >= self.auto_save_interval_ms

    def mark_saved(self, current_time):
        self._last_save_time = current_time

    def list_checkpoints(self):
        """Metadata of every checkpoint in the directory, latest simulated time first
        (checkpoint.rs:112-139)."""
        out = []
        for fn in os.listdir(self.checkpoint_dir):
            if not fn.endswith(".json") or "_meta." in fn:
                continue
            name = fn[: -len(".json")]
            mp = os.path.join(self.checkpoint_dir, f"{name}_meta.json")
            try:
                with open(mp) as f:
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
            for p in self._paths(m.name):
                try:
                    os.remove(p)
                except OSError:
                    pass
