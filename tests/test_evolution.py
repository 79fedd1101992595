"""CPU checks of the evolution operators and loop (src/evolution/*.rs restated): the reference's
own tests (selection.rs:64-112, crossover.rs:166-221, mutation.rs:239-272,
population.rs:320-395) on weight-vector genomes, plus what the definitions fix exactly
(BLX's interval, elitism, the hall of fame, the diversity formula, the CSV format).  The loop
runs here with a host fitness function; the device evaluation is covered by
test_gpu_evolution.py."""
import numpy as np
import pytest

from neuralsim_amd import configs
from neuralsim_amd.evolution import (CrossoverMode, EvolutionConfig, Genome, MutationConfig,
                                     Population, crossover, crossover_blx_alpha,
                                     crossover_uniform, gaussian_sample, mutate,
                                     roulette_select, tournament_select)

M = 400
EXC = np.arange(M) % 5 != 4  # every fifth synapse from an inhibitory source


def _genome(rng, name="g"):
    u = rng.random(M)
    return Genome(name, np.where(EXC, 3.0 * u, -5.0 * u).astype(np.float32))


def _pop(size=20, **cfg):
    config = EvolutionConfig(**cfg)
    rng = np.random.default_rng(config.seed)
    genomes = [_genome(rng, f"g{i}") for i in range(size)]
    net = configs.population_nets(1)[0]
    return Population(genomes, config, configs.population_base(), net, template=(None, EXC))


def test_tournament_valid_and_prefers_fitter():
    """selection.rs:69-99: k = 5 of 20 with one far fitter picks it > 15 % of the time."""
    rng = np.random.default_rng(42)
    fit = np.arange(10, dtype=np.float64)
    assert 0 <= tournament_select(fit, 3, rng) < 10
    fit = np.where(np.arange(20) == 5, 100.0, 1.0)
    picks = [tournament_select(fit, 5, rng) for _ in range(1000)]
    assert picks.count(5) > 150
    # P(best in a 5-draw tournament) = 1 - (19/20)^5 = 0.226
    assert abs(picks.count(5) / 1000 - 0.226) < 0.05
    assert 0 <= tournament_select(fit, 100, rng) < 20  # k > n is clamped to n


def test_roulette_valid_and_proportional():
    rng = np.random.default_rng(42)
    fit = [0.1, 0.2, 0.3, 0.4, 0.5]
    picks = [roulette_select(fit, rng) for _ in range(20000)]
    assert all(0 <= p < 5 for p in picks)
    shifted = np.array(fit) - 0.1 + 1e-6
    want = shifted / shifted.sum()
    got = np.bincount(picks, minlength=5) / len(picks)
    assert np.allclose(got, want, atol=0.01)
    assert roulette_select([], rng) == 0


def test_uniform_crossover_takes_each_weight_from_a_parent():
    """crossover.rs:170-186."""
    rng = np.random.default_rng(42)
    a, b = _genome(rng, "a"), _genome(rng, "b")
    c = crossover_uniform(a, b, rng)
    assert len(c.weights) == M and c.name == "a_x_b"
    assert np.all((c.weights == a.weights) | (c.weights == b.weights))
    frac = np.mean(c.weights == a.weights)
    assert 0.4 < frac < 0.6


def test_blx_alpha_stays_in_extended_interval():
    """crossover.rs:189-200 plus the interval the definition fixes, and Dale's sign."""
    rng = np.random.default_rng(42)
    a, b = _genome(rng, "a"), _genome(rng, "b")
    c = crossover_blx_alpha(a, b, 0.5, rng)  # no sign constraint: the literal definition
    lo = np.minimum(a.weights, b.weights).astype(np.float64)
    hi = np.maximum(a.weights, b.weights).astype(np.float64)
    r = hi - lo
    assert np.all(np.isfinite(c.weights))
    assert np.all(c.weights >= lo - 0.5 * r - 1e-5) and np.all(c.weights <= hi + 0.5 * r + 1e-5)
    # the extended interval is filled: some children fall outside the parents' range
    assert np.any((c.weights < lo) | (c.weights > hi))
    d = crossover(CrossoverMode.BLX_ALPHA, a, b, rng, EXC)
    assert np.all(d.weights[EXC] >= 0) and np.all(d.weights[~EXC] <= 0)
    # identical parents give the parents' value
    e = crossover_blx_alpha(a, a, 0.5, rng)
    assert np.array_equal(e.weights, a.weights)


def test_mutation_changes_some_weights_and_keeps_bounds():
    """mutation.rs:243-289: defaults change some weights; extreme settings stay bounded."""
    rng = np.random.default_rng(42)
    g = _genome(rng)
    before = g.weights.copy()
    mutate(g, MutationConfig(), rng, EXC)
    assert np.any(g.weights != before)
    cfg = MutationConfig(weight_mutation_rate=1.0, weight_mutation_sigma=100.0,
                         weight_reset_rate=0.5)
    mutate(g, cfg, rng, EXC)
    mag = np.abs(g.weights)
    assert np.all(mag >= cfg.weight_min) and np.all(mag <= cfg.weight_max)
    assert np.all(g.weights[EXC] >= 0) and np.all(g.weights[~EXC] <= 0)


def test_no_mutation_when_rates_zero():
    """mutation.rs:258-272."""
    rng = np.random.default_rng(42)
    g = _genome(rng)
    before = g.weights.copy()
    cfg = MutationConfig(weight_mutation_rate=0.0, weight_reset_rate=0.0, param_mutation_rate=0.0,
                         plasticity_mutation_rate=0.0, connection_mutation_rate=0.0)
    mutate(g, cfg, rng, EXC)
    assert np.array_equal(g.weights.view(np.uint32), before.view(np.uint32))


def test_mutation_rates_and_sigma():
    """The fractions reset / perturbed and the perturbation's spread, from the definition."""
    rng = np.random.default_rng(7)
    n = 200_000
    exc = np.ones(n, bool)
    g = Genome("g", np.full(n, 5.0, np.float32))
    cfg = MutationConfig(weight_mutation_rate=0.1, weight_mutation_sigma=0.3, weight_reset_rate=0.01)
    mutate(g, cfg, rng, exc)
    changed = g.weights != 5.0
    assert abs(changed.mean() - (0.01 + 0.99 * 0.1)) < 0.003
    z = gaussian_sample(0.0, 1.0, rng, 200_000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01


def test_population_initializes():
    """population.rs:326-331."""
    pop = _pop(20)
    assert len(pop.genomes) == 20 and pop.generation == 0


def test_evolution_improves_a_host_fitness():
    """population.rs:333-364 with a host fitness (closeness to a target weight vector): the
    hall of fame never decreases, elites survive unchanged, and the best improves."""
    target = np.where(EXC, 2.0, -1.0)

    def fitness(genomes):
        return [1.0 / (1.0 + np.linalg.norm(g.weights - target)) for g in genomes]

    pop = _pop(30, seed=3)
    first = max(fitness(pop.genomes))
    hof = []
    for _ in range(15):
        fit = np.array(fitness(pop.genomes))
        elite = pop.genomes[int(np.argmax(fit))].weights.copy()
        pop.evolve_generation(None, fitness)
        hof.append(pop.best_fitness)
        # round(0.1 * 30) = 3 elites, the best of them first
        assert np.array_equal(pop.genomes[0].weights, elite)
    assert all(b >= a for a, b in zip(hof, hof[1:]))
    assert pop.best_fitness > first
    assert pop.generation == 15 and len(pop.history) == 15
    h = pop.history[-1]
    assert h.worst_fitness <= h.mean_fitness <= h.best_fitness and h.fitness_std >= 0


def test_diversity_and_history_csv(tmp_path):
    """population.rs:286-314 (mean pairwise distance of the first 50) and 216-240 (CSV)."""
    pop = _pop(5)
    w = [g.weights.astype(np.float64) for g in pop.genomes]
    want = np.mean([np.linalg.norm(w[i] - w[j]) for i in range(5) for j in range(i + 1, 5)])
    assert pop.compute_diversity() == pytest.approx(want, rel=1e-12)
    assert _pop(1).compute_diversity() == 0.0
    pop.evolve_generation(None, lambda gs: [0.5] * len(gs))
    path = tmp_path / "sub" / "history.csv"
    pop.save_history_csv(path)
    rows = path.read_text().splitlines()
    assert rows[0] == "generation,best_fitness,mean_fitness,worst_fitness,fitness_std,best_rate,diversity"
    assert rows[1].startswith("0,0.5,0.5,0.5,0,0.5,")
    pop.save_best_genome(tmp_path / "best.json")
    assert (tmp_path / "best.json").exists()


def test_ties_and_elite_count():
    """Elites: round(size * fraction), at least one; ties keep index order; the hall of fame
    takes the last of equal bests (Iterator::max_by)."""
    pop = _pop(4, elite_fraction=0.1)  # round(0.4) = 0 -> 1 elite
    fit = [1.0, 3.0, 3.0, 2.0]
    names = [g.name for g in pop.genomes]
    pop.evolve_generation(None, lambda gs: fit)
    assert pop.genomes[0].name == names[1]
    assert pop.best_genome.name == names[2]
