"""Evolution host loop feeding the batched engine: the reference's genetic algorithm
(src/evolution/population.rs, selection.rs, crossover.rs, mutation.rs, genome.rs) with each
generation evaluated by the device.

The reference evolves a fixed topology ("only the weights differ between individuals",
genome.rs:46-49): a genome's heritable part is one weight per synapse, in the order the network
is built, and the topology is rebuilt from a fixed seed.  Here a genome is a weight vector in
the CSR slot order of one network specification (an NsxBatchNet: size, E/I split, in-degree,
delays, seed), shared by the whole population.

Evaluation, `Population.evaluate`:
* "batched" (default): the generation is one BatchedPopulation of identical networks, each
  given its genome's weights (nsx_upload_weights) and stepped together for the lifetime; the
  fitness is RateHomeostasis of each network's count rate (population.evaluate's convention).
  Every network sees the same drive (the spec's seed keys both topology and drive), so
  individuals differ only by their weights: common random numbers across the generation.  Under
  a plastic base configuration (pair or triplet STDP) each network learns as its own engine
  would, and with `lamarckian` the learned weights are read back into the genomes
  (nsx_read_weights of the batch, each network's slot range), where the reference re-runs a
  separate 500 ms simulation per genome for them (population.rs:126-138).
* "sequential": one SimulationEngine per genome with the base configuration -- the reference's
  per-genome build_engine + simulate_ms, the same results as "batched" one engine at a time.

Operators follow the reference's definitions with numpy's PCG64 in place of Rust's StdRng, so
runs are reproducible here but not draw-for-draw equal to the reference's.  Weights carry
Dale's sign: excitatory sources' weights stay >= 0 and inhibitory ones <= 0, and the reference's
[weight_min, weight_max] bound (mutation.rs:64-78) applies to the magnitude.  Neuron-parameter,
plasticity and connection mutations have no per-network counterpart in the batched engine (its
networks share one parameter set) and are not applied.
"""
import dataclasses
import enum
import json
import os

import numpy as np

from . import abi, configs
from .engine import BatchedPopulation, SimulationEngine
from .population import RateHomeostasis, lifetime_steps
from .stats import rust_display


class CrossoverMode(enum.Enum):
    """CrossoverMode (crossover.rs:12-16)."""
    UNIFORM = "Uniform"
    BLX_ALPHA = "BlxAlpha"


@dataclasses.dataclass
class MutationConfig:
    """MutationConfig with the reference's defaults (mutation.rs:16-55)."""
    weight_mutation_rate: float = 0.1
    weight_mutation_sigma: float = 0.3
    weight_reset_rate: float = 0.01
    param_mutation_rate: float = 0.2
    param_mutation_sigma: float = 0.1
    plasticity_mutation_rate: float = 0.3
    plasticity_mutation_sigma: float = 0.1
    connection_mutation_rate: float = 0.15
    weight_min: float = 0.0
    weight_max: float = 10.0


@dataclasses.dataclass
class EvolutionConfig:
    """EvolutionConfig with the reference's defaults (population.rs:37-66)."""
    elite_fraction: float = 0.1
    tournament_size: int = 3
    crossover_mode: CrossoverMode = CrossoverMode.BLX_ALPHA
    mutation: MutationConfig = dataclasses.field(default_factory=MutationConfig)
    lamarckian: bool = False
    seed: int = 42


@dataclasses.dataclass
class GenerationStats:
    """GenerationStats (population.rs:69-79)."""
    generation: int
    best_fitness: float
    mean_fitness: float
    worst_fitness: float
    fitness_std: float
    best_rate: float
    diversity: float


@dataclasses.dataclass
class Genome:
    """A heritable weight vector (CSR slot order of the population's network) and a name."""
    name: str
    weights: np.ndarray

    def synapse_count(self):
        return len(self.weights)

    def copy(self):
        return Genome(self.name, self.weights.copy())


# ---- selection (selection.rs) ---------------------------------------------------------------

def tournament_select(fitnesses, k, rng):
    """selection.rs:15-34: k uniform draws (with replacement), the first strictly fittest wins."""
    n = len(fitnesses)
    k = min(k, n)
    best = int(rng.integers(n))
    best_f = fitnesses[best]
    for _ in range(1, k):
        i = int(rng.integers(n))
        if fitnesses[i] > best_f:
            best, best_f = i, fitnesses[i]
    return best


def roulette_select(fitnesses, rng):
    """selection.rs:38-59: fitness-proportionate on fitnesses shifted to min + 1e-6."""
    f = np.asarray(fitnesses, np.float64)
    n = len(f)
    if n == 0:
        return 0
    shifted = f - f.min() + 1e-6
    total = shifted.sum()
    if total <= 0.0:
        return int(rng.integers(n))
    r = rng.random() * total
    c = 0.0
    for i, x in enumerate(shifted):
        c += x
        if r <= c:
            return i
    return n - 1


# ---- crossover (crossover.rs) ---------------------------------------------------------------

def _dale(w, exc):
    """Clip each weight to its source's sign: >= 0 excitatory, <= 0 inhibitory."""
    return np.where(exc, np.maximum(w, 0.0), np.minimum(w, 0.0))


def crossover_uniform(a, b, rng):
    """crossover.rs:21-80: each weight from either parent with probability 1/2."""
    m = min(len(a.weights), len(b.weights))
    pick = rng.random(m) < 0.5
    w = np.where(pick, a.weights[:m], b.weights[:m]).astype(np.float32)
    return Genome(f"{a.name}_x_{b.name}", w)


def crossover_blx_alpha(a, b, alpha, rng, exc=None):
    """crossover.rs:88-150: each weight uniform in [lo - alpha r, hi + alpha r] of its parents'
    values (r = hi - lo), then kept on its source's sign side."""
    m = min(len(a.weights), len(b.weights))
    wa = a.weights[:m].astype(np.float64)
    wb = b.weights[:m].astype(np.float64)
    lo, hi = np.minimum(wa, wb), np.maximum(wa, wb)
    r = hi - lo
    lo_e, hi_e = lo - alpha * r, hi + alpha * r
    w = rng.random(m) * (hi_e - lo_e) + lo_e
    if exc is not None:
        w = _dale(w, exc[:m])
    return Genome(f"{a.name}_x_{b.name}", w.astype(np.float32))


def crossover(mode, a, b, rng, exc=None):
    """crossover.rs:153-163 (BLX with alpha = 0.5)."""
    if mode == CrossoverMode.UNIFORM:
        return crossover_uniform(a, b, rng)
    return crossover_blx_alpha(a, b, 0.5, rng, exc)


# ---- mutation (mutation.rs) -----------------------------------------------------------------

def gaussian_sample(mean, sigma, rng, size=None):
    """mutation.rs:230-236: Box-Muller cosine branch, u1 floored at 1e-10."""
    u1 = np.maximum(rng.random(size), 1e-10)
    u2 = rng.random(size)
    return mean + sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def mutate_weights(genome, config, rng, exc):
    """mutation.rs:64-78 on magnitudes: with probability weight_reset_rate a weight is redrawn
    uniform in [weight_min, weight_max]; otherwise with probability weight_mutation_rate it gets
    N(0, weight_mutation_sigma) and is clamped to the bounds.  Signs follow the source."""
    m = len(genome.weights)
    sign = np.where(exc[:m], 1.0, -1.0)
    mag = np.abs(genome.weights.astype(np.float64))
    u_reset = rng.random(m)
    u_mut = rng.random(m)
    reset = u_reset < config.weight_reset_rate
    pert = ~reset & (u_mut < config.weight_mutation_rate)
    fresh = rng.random(m) * (config.weight_max - config.weight_min) + config.weight_min
    delta = gaussian_sample(0.0, config.weight_mutation_sigma, rng, m)
    mag = np.where(reset, fresh, mag)
    mag = np.where(pert, np.clip(mag + delta, config.weight_min, config.weight_max), mag)
    genome.weights = np.where(reset | pert, (sign * mag).astype(np.float32), genome.weights)


def mutate(genome, config, rng, exc):
    """mutation.rs:57-62 (weights; see the module docstring for the other operators)."""
    mutate_weights(genome, config, rng, exc)


# ---- population (population.rs) -------------------------------------------------------------

def _template(base, net):
    """The network's CSR (built once by the device generator) and, per slot, whether its source
    is excitatory."""
    cfg = configs.population_single(base, net)
    with SimulationEngine(cfg) as eng:
        eng.generate_fixed_indegree(net.k_in, net.w_exc_lo, net.w_exc_hi, net.w_inh_lo,
                                    net.w_inh_hi, net.d_lo, net.d_hi)
        rp, col, w, d = eng.read_csr()
    src = np.repeat(np.arange(len(rp) - 1, dtype=np.int64), np.diff(rp.astype(np.int64)))
    return (rp, col, w, d), src < net.n_exc


class Population:
    """Population (population.rs:82-200) of genomes over one network specification."""

    def __init__(self, genomes, config, base, net, mode="batched", graph_steps=50,
                 template=None):
        if mode not in ("batched", "sequential"):
            raise ValueError("mode must be 'batched' or 'sequential'")
        self.genomes = list(genomes)
        self.generation = 0
        self.config = config
        self.history = []
        self.best_genome = None
        self.best_fitness = -np.inf
        self.base, self.net = base, net
        self.mode = mode
        self.graph_steps = graph_steps
        self.rng = np.random.default_rng(config.seed)
        self.csr, self.exc = template if template is not None else _template(base, net)
        self.last_rates = None

    @classmethod
    def random(cls, size, config, base=None, net=None, **kw):
        """Population::random (population.rs:95-107): each genome's weights uniform in the
        network's excitatory / inhibitory ranges (the reference draws uniform [0.5, 2.5],
        genome.rs:127-131)."""
        base = base if base is not None else configs.population_base()
        if net is None:
            net = configs.population_nets(1, seed=config.seed)[0]
            net.n_neurons, net.n_exc, net.k_in = 2000, 1600, 100
        template = _template(base, net)
        exc = template[1]
        rng = np.random.default_rng(config.seed)
        genomes = []
        for _ in range(size):
            u = rng.random(len(exc))
            w = np.where(exc, net.w_exc_lo + u * (net.w_exc_hi - net.w_exc_lo),
                         net.w_inh_lo + u * (net.w_inh_hi - net.w_inh_lo))
            genomes.append(Genome(f"Genome_{int(rng.integers(1 << 32))}", w.astype(np.float32)))
        pop = cls(genomes, config, base, net, template=template, **kw)
        pop.rng = rng
        return pop

    # evaluation ---------------------------------------------------------------------------
    def evaluate(self, evaluator):
        """Fitness of every genome (population.rs:116-124); with `lamarckian`, the weights
        after the lifetime are written back into the genomes."""
        steps = lifetime_steps(evaluator.lifetime_ms, self.base.dt_ms)
        k = len(self.genomes)
        n = self.net.n_neurons
        if self.mode == "batched":
            nets = [self.net] * k
            with BatchedPopulation(self.base, nets) as pop:
                slots = pop.net_slots()
                w = pop.weights()
                for (a, b), g in zip(slots, self.genomes):
                    w[a:b] = g.weights
                pop.upload_weights(w)
                if self.graph_steps:
                    pop.set_graph_steps(self.graph_steps)
                pop.step(steps)
                spikes, _ = pop.net_stats()
                after = pop.weights() if self.config.lamarckian else None
            if after is not None:
                for (a, b), g in zip(slots, self.genomes):
                    g.weights = after[a:b].copy()
        else:
            spikes = np.zeros(k, np.uint64)
            cfg = configs.population_single(self.base, self.net)
            rp, col, _, d = self.csr
            for i, g in enumerate(self.genomes):
                with SimulationEngine(cfg) as eng:
                    eng.upload_csr(rp, col, g.weights, d)
                    if self.graph_steps:
                        eng.set_graph_steps(self.graph_steps)
                    st = eng.step(steps)
                    spikes[i] = st.total_spikes
                    if self.config.lamarckian:
                        g.weights = eng.weights()
        rate = spikes.astype(np.float64) / n / (steps * self.base.dt_ms / 1e3)
        self.last_rates = rate
        return np.asarray(evaluator.fitness(rate), np.float64)

    def evolve_generation(self, evaluator, fitness_fn=None):
        """Population::evolve_generation (population.rs:111-200): evaluate, record statistics
        and the hall of fame, keep round(elite_fraction * size) (at least 1) elites by fitness,
        fill with tournament-selected parents' crossover + mutation.  `fitness_fn(genomes)`
        replaces the device evaluation (host-only use).  Returns the generation's best."""
        fit = (np.asarray(fitness_fn(self.genomes), np.float64) if fitness_fn is not None
               else self.evaluate(evaluator))
        stats = self.compute_stats(fit)
        best = len(fit) - 1 - int(np.argmax(fit[::-1]))  # the last of the largest (max_by)
        if fit[best] > self.best_fitness:
            self.best_fitness = float(fit[best])
            self.best_genome = self.genomes[best].copy()
        self.history.append(stats)
        size = len(self.genomes)
        elites = min(max(int(np.floor(size * self.config.elite_fraction + 0.5)), 1), size)
        order = np.argsort(-fit, kind="stable")  # fitness descending, ties in index order
        nxt = [self.genomes[i].copy() for i in order[:elites]]
        cfg = self.config
        while len(nxt) < size:
            a = tournament_select(fit, cfg.tournament_size, self.rng)
            b = tournament_select(fit, cfg.tournament_size, self.rng)
            child = crossover(cfg.crossover_mode, self.genomes[a], self.genomes[b], self.rng,
                              self.exc)
            mutate(child, cfg.mutation, self.rng, self.exc)
            nxt.append(child)
        self.genomes = nxt
        self.generation += 1
        return stats.best_fitness

    def evolve(self, n, evaluator, fitness_fn=None):
        """population.rs:203-208."""
        for _ in range(n):
            self.evolve_generation(evaluator, fitness_fn)
        return self.best_fitness

    def last_mean_fitness(self):
        return self.history[-1].mean_fitness if self.history else 0.0

    def compute_stats(self, fit):
        """population.rs:255-282; best_rate is the reference's (the mean fitness when the best
        is positive)."""
        n = len(fit)
        best, worst = float(np.max(fit)), float(np.min(fit))
        mean = float(np.sum(fit) / n)
        std = float(np.sqrt(np.sum((fit - mean) ** 2) / n))
        return GenerationStats(self.generation, best, mean, worst, std,
                               mean if best > 0.0 else 0.0, self.compute_diversity())

    def compute_diversity(self):
        """population.rs:286-314: mean pairwise Euclidean distance of the first 50 genomes'
        weight vectors."""
        s = [g.weights.astype(np.float64) for g in self.genomes[:50]]
        if len(s) < 2:
            return 0.0
        tot, cnt = 0.0, 0
        for i in range(len(s)):
            for j in range(i + 1, len(s)):
                m = min(len(s[i]), len(s[j]))
                tot += float(np.sqrt(np.sum((s[i][:m] - s[j][:m]) ** 2)))
                cnt += 1
        return tot / cnt

    def history_csv(self):
        out = ["generation,best_fitness,mean_fitness,worst_fitness,fitness_std,best_rate,"
               "diversity\n"]
        for h in self.history:
            out.append(",".join(rust_display(x) for x in (
                h.generation, h.best_fitness, h.mean_fitness, h.worst_fitness, h.fitness_std,
                h.best_rate, h.diversity)) + "\n")
        return "".join(out)

    def save_history_csv(self, path):
        """population.rs:216-240."""
        d = os.path.dirname(str(path))
        if d:
            os.makedirs(d, exist_ok=True)
        with open(path, "w") as f:
            f.write(self.history_csv())

    def save_best_genome(self, path):
        """population.rs:243-251 (JSON; weights as a list)."""
        if self.best_genome is None:
            return
        with open(path, "w") as f:
            json.dump({"name": self.best_genome.name,
                       "initial_weights": [float(x) for x in self.best_genome.weights],
                       "network": {k: getattr(self.net, k) for k, _ in abi.NsxBatchNet._fields_},
                       "fitness": self.best_fitness}, f, indent=2)


__all__ = ["CrossoverMode", "MutationConfig", "EvolutionConfig", "GenerationStats", "Genome",
           "Population", "RateHomeostasis", "tournament_select", "roulette_select",
           "crossover_uniform", "crossover_blx_alpha", "crossover", "mutate", "mutate_weights",
           "gaussian_sample"]
