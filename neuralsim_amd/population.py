"""Population evaluation on the batched engine (src/evolution/population.rs, fitness.rs).

`evaluate` replaces the rayon loop of Population::evolve_generation (population.rs:116-124): all
genomes' networks run as one BatchedPopulation and each gets RateHomeostasis fitness
(fitness.rs:59-75).  The mean rate is the exact count rate of the lifetime, spikes / (n T),
where the reference reads its 100 ms moving average (engine.rs:633-648); the two agree in a
stationary regime and the count is exact and bit-reproducible.
"""
import dataclasses

import numpy as np

from .engine import BatchedPopulation


@dataclasses.dataclass
class RateHomeostasis:
    """RateHomeostasis (fitness.rs:24-80): 1 / (1 + |rate - target|), x0.1 if silent (< 1 Hz,
    when penalize_silence), x0.1 if exploding (> 500 Hz)."""
    target_rate: float
    lifetime_ms: float = 500.0
    penalize_silence: bool = True

    def fitness(self, mean_rate):
        r = np.asarray(mean_rate, np.float64)
        f = 1.0 / (1.0 + np.abs(r - self.target_rate))
        if self.penalize_silence:
            f = np.where(r < 1.0, f * 0.1, f)
        return np.where(r > 500.0, f * 0.1, f)


def lifetime_steps(lifetime_ms, dt_ms):
    """Steps of simulate_ms(lifetime_ms): the quotient truncated, as engine.rs:792-797 casts
    (duration_ms / dt) to usize (0.3 ms at dt 0.1 is 2.999... -> 2 steps)."""
    q = float(lifetime_ms) / float(dt_ms)
    return int(q) if q > 0 else 0


def evaluate(base, nets, evaluator, graph_steps=50):
    """Run every network for evaluator.lifetime_ms in one batch; returns a dict of per-network
    spike counts, mean rates (Hz), rate standard deviation across neurons (Hz) and fitness."""
    with BatchedPopulation(base, nets) as pop:
        steps = lifetime_steps(evaluator.lifetime_ms, base.dt_ms)
        if graph_steps:
            pop.set_graph_steps(graph_steps)
        st = pop.step(steps)
        spikes, sq = pop.net_stats()
    n = np.array([b.n_neurons for b in nets], np.float64)
    t_s = steps * base.dt_ms / 1000.0
    mean_count = spikes / n
    var = np.maximum(sq / n - mean_count**2, 0.0)
    rate = mean_count / t_s
    return {"spikes": spikes, "mean_rate": rate, "rate_std": np.sqrt(var) / t_s,
            "fitness": evaluator.fitness(rate), "steps": steps, "wall_ms": st.wall_ms,
            "events": st.events}
