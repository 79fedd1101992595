"""CPU checks of the population host logic: RateHomeostasis against the values its reference
documents (fitness.rs:26-31), the penalties, and the per-network seed mixer against the
published SplitMix64 output sequence."""
import numpy as np
import pytest

from neuralsim_amd import abi, configs
from neuralsim_amd.population import RateHomeostasis, lifetime_steps


def test_rate_homeostasis_documented_values():
    f = RateHomeostasis(5.0).fitness([5.0, 10.0, 55.0])
    # "Perfect match -> 1.0, 5 Hz off -> 0.17, 50 Hz off -> 0.02"
    assert f[0] == 1.0 and f[1] == pytest.approx(1 / 6) and f[2] == pytest.approx(1 / 51)
    assert round(f[1], 2) == 0.17 and round(f[2], 2) == 0.02


def test_rate_homeostasis_penalties():
    ev = RateHomeostasis(5.0)
    assert ev.fitness(0.5) == pytest.approx(0.1 / (1 + 4.5))
    assert ev.fitness(600.0) == pytest.approx(0.1 / (1 + 595.0))
    assert RateHomeostasis(5.0, penalize_silence=False).fitness(0.5) == pytest.approx(1 / 5.5)


def test_splitmix64_sequence():
    # SplitMix64 with state 0: the state advances by the golden gamma per output
    gamma = 0x9E3779B97F4A7C15
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    got = [configs._mix64((k * gamma) & 0xFFFFFFFFFFFFFFFF) for k in range(3)]
    assert got == want


def test_population_networks():
    nets = configs.population_nets(256, seed=2026)
    n = np.array([b.n_neurons for b in nets])
    k = np.array([b.k_in for b in nets])
    assert n.min() >= 1000 and n.max() <= 64000 and k.min() >= 20 and k.max() <= 200
    assert len({b.seed for b in nets}) == 256
    assert all(b.n_exc == b.n_neurons * 4 // 5 for b in nets)
    again = configs.population_nets(256, seed=2026)
    assert [(b.n_neurons, b.k_in, b.seed) for b in nets] == \
        [(b.n_neurons, b.k_in, b.seed) for b in again]
    base = configs.population_base()
    assert base.model == abi.MODEL_IZHIKEVICH and base.plasticity == abi.PLAST_NONE
    assert base.ring_slots == 11 and base.drive == abi.DRIVE_POISSON


def test_lifetime_steps_truncate_like_simulate_ms():
    """simulate_ms runs (duration / dt) as usize steps (engine.rs:792-797): truncation."""
    assert 0.3 / 0.1 < 3.0
    assert lifetime_steps(0.3, 0.1) == 2
    assert lifetime_steps(500.0, 1.0) == 500 and lifetime_steps(2000.0, 1.0) == 2000
    assert lifetime_steps(1.0, 0.3) == 3 and lifetime_steps(0.0, 1.0) == 0
    assert lifetime_steps(-5.0, 1.0) == 0
