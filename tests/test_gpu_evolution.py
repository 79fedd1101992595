"""The evolution loop on the device: a generation evaluated as one batched launch equals the
same genomes run one network at a time (fitness, rates); weights uploaded into the batch land
in the right networks; the reference's evolve test (population.rs:333-364: the best never
decreases) on the device evaluation; Lamarckian extraction of STDP-learned weights, batched
and sequential alike."""
import numpy as np
import pytest

from neuralsim_amd import abi, configs
from neuralsim_amd.evolution import EvolutionConfig, Population, RateHomeostasis

pytestmark = pytest.mark.gpu


def _net(n=1500, k=60):
    net = configs.population_nets(1, seed=77)[0]
    net.n_neurons, net.n_exc, net.k_in = n, n * 4 // 5, k
    return net


def test_upload_weights_round_trip(nsx):
    spec = configs.izh_stdp_network(n=2000, seed=3, k_in=20)
    with nsx.SimulationEngine.from_spec(spec) as eng:
        w = eng.weights()
        w2 = (w * 0.5).astype(np.float32)
        eng.upload_weights(w2)
        assert np.array_equal(eng.weights().view(np.uint32), w2.view(np.uint32))
        with pytest.raises(Exception):
            eng.upload_weights(w2[:-1])


def test_batched_generation_equals_sequential(nsx):
    cfg = EvolutionConfig(seed=11)
    ev = RateHomeostasis(5.0, lifetime_ms=150.0)
    print("case evolution batched vs sequential", flush=True)
    pop = Population.random(6, cfg, net=_net())
    seq = Population([g.copy() for g in pop.genomes], cfg, pop.base, pop.net, mode="sequential",
                     template=(pop.csr, pop.exc))
    fb = pop.evaluate(ev)
    fs = seq.evaluate(ev)
    assert np.array_equal(pop.last_rates, seq.last_rates) and np.array_equal(fb, fs)
    assert len(set(pop.last_rates.tolist())) > 1  # the genomes' weights make a difference


def test_evolve_on_device_never_loses_the_best(nsx):
    cfg = EvolutionConfig(seed=5)
    ev = RateHomeostasis(5.0, lifetime_ms=100.0)
    print("case evolution 3 generations", flush=True)
    pop = Population.random(8, cfg, net=_net(1200, 40))
    best = []
    for _ in range(3):
        pop.evolve_generation(ev)
        best.append(pop.best_fitness)
    assert all(b >= a for a, b in zip(best, best[1:]))
    assert pop.generation == 3 and len(pop.history) == 3
    assert all(0.0 < h.best_fitness <= 1.0 for h in pop.history)


def test_lamarckian_extracts_learned_weights(nsx):
    base = configs.population_base()
    base.plasticity = abi.PLAST_PAIR
    base.a_plus, base.a_minus, base.tau_plus_ms, base.tau_minus_ms = 0.01, 0.012, 20.0, 20.0
    base.w_min, base.w_max = 0.0, 10.0
    cfg = EvolutionConfig(seed=9, lamarckian=True)
    ev = RateHomeostasis(5.0, lifetime_ms=200.0)
    print("case evolution lamarckian", flush=True)
    pop = Population.random(3, cfg, base=base, net=_net(1000, 40))  # batched, pair STDP
    seq = Population([g.copy() for g in pop.genomes], cfg, pop.base, pop.net, mode="sequential",
                     template=(pop.csr, pop.exc))
    before = [g.weights.copy() for g in pop.genomes]
    fb = pop.evaluate(ev)
    fs = seq.evaluate(ev)
    # the batch learns what each genome's own engine learns, bit for bit
    assert np.array_equal(fb, fs) and np.array_equal(pop.last_rates, seq.last_rates)
    for g, h in zip(pop.genomes, seq.genomes):
        assert np.array_equal(g.weights.view(np.uint32), h.weights.view(np.uint32))
    changed = [np.any(g.weights != b) for g, b in zip(pop.genomes, before)]
    assert all(changed)
    # only excitatory-to-excitatory synapses learn
    rp = pop.csr[0].astype(np.int64)
    col = pop.csr[1]
    ee = pop.exc & (col < pop.net.n_exc)
    for g, b in zip(pop.genomes, before):
        assert np.array_equal(g.weights[~ee], b[~ee])
    assert len(rp) == pop.net.n_neurons + 1
