"""Simulation statistics and their CSV recorder.

SimulationStats mirrors the reference's struct (src/simulation/engine.rs:21-33); the device
fills it (nsx_read_sim_stats) while statistics are enabled on an engine (set_stats).
StatsRecorder mirrors src/io/checkpoint.rs:164-254: a bounded ring of records taken every
`interval_steps` steps, exported as CSV with the reference's header and its number formatting
(Rust's `{}` Display: shortest round-trip digits, no exponent, no trailing ".0").
"""
import collections
import dataclasses
import decimal
import math


@dataclasses.dataclass
class SimulationStats:
    """SimulationStats (engine.rs:21-33) plus the engine's LFP (engine.rs:64, 804-806)."""
    total_spikes: int = 0
    output_spikes: int = 0
    active_neurons: int = 0
    sim_time_ms: float = 0.0
    weight_updates: int = 0
    mean_firing_rate: float = 0.0
    synchrony_index: float = 0.0
    weight_mean: float = 0.0
    weight_std: float = 0.0
    seed: int = 0
    lfp: float = 0.0
    steps: int = 0

    @classmethod
    def from_abi(cls, r):
        return cls(total_spikes=r.total_spikes, output_spikes=r.output_spikes,
                   active_neurons=r.active_neurons, sim_time_ms=r.sim_time_ms,
                   weight_updates=r.weight_updates, mean_firing_rate=r.mean_firing_rate,
                   synchrony_index=r.synchrony_index, weight_mean=r.weight_mean,
                   weight_std=r.weight_std, seed=r.seed, lfp=r.lfp, steps=r.steps)


def rust_display(x):
    """A number as Rust's `{}` prints it: integers as is; f64 with the shortest digits that
    round-trip, in plain notation, without a trailing '.0' ("1", "0.1", "-0", "NaN", "inf")."""
    if isinstance(x, int):
        return str(x)
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    s = format(decimal.Decimal(repr(x)), "f")
    if "." in s:
        s = s.rstrip("0").rstrip(".")
    return s


StatsRecord = collections.namedtuple(
    "StatsRecord", ["step", "sim_time_ms", "total_spikes", "output_spikes", "mean_firing_rate",
                    "synchrony_index", "weight_mean", "weight_std", "lfp"])

CSV_HEADER = ("step,sim_time_ms,total_spikes,output_spikes,mean_firing_rate,synchrony_index,"
              "weight_mean,weight_std,lfp\n")


class StatsRecorder:
    """StatsRecorder (checkpoint.rs:166-254): keeps at most `max_records` records, one every
    `interval_steps` steps, the oldest evicted first."""

    def __init__(self, max_records, interval_steps):
        self._records = collections.deque()
        self.max_records = int(max_records)
        self.record_interval_steps = int(interval_steps)

    def record(self, step, stats, lfp):
        """StatsRecorder::record (checkpoint.rs:199-218)."""
        if step % self.record_interval_steps != 0:
            return
        if len(self._records) >= self.max_records:
            self._records.popleft()
        self._records.append(StatsRecord(
            int(step), float(stats.sim_time_ms), int(stats.total_spikes),
            int(stats.output_spikes), float(stats.mean_firing_rate),
            float(stats.synchrony_index), float(stats.weight_mean), float(stats.weight_std),
            float(lfp)))

    def to_csv(self):
        """StatsRecorder::to_csv (checkpoint.rs:220-240)."""
        out = [CSV_HEADER]
        for r in self._records:
            out.append(",".join(rust_display(v) for v in r) + "\n")
        return "".join(out)

    def save_csv(self, path):
        with open(path, "w") as f:
            f.write(self.to_csv())

    def records(self):
        return self._records
