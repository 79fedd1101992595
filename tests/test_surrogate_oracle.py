"""CPU checks of the surrogate's test oracle (oracle/mlp.py) and of the spike-displacement
metric: bf16 rounding against hand-worked ties and torch's conversion, the backward pass against
central finite differences, the forward pass against its closed form where every ReLU is
active, and the envelope selector on injected out-of-range states."""
import numpy as np
import pytest
import torch

from neuralsim_amd.metrics import spike_displacement
from oracle import mlp


def test_round_bf16_ties_to_even():
    one = 1.0
    cases = {
        one: one,
        one + 2.0**-8: one,                                # tie, 1.0 is even
        one + 3 * 2.0**-8: one + 2.0**-6,                  # tie, rounds to the even neighbour
        one + 2.0**-8 + 2.0**-20: one + 2.0**-7,           # above the tie
        -(one + 2.0**-8): -one,
        3.0: 3.0,
        0.0: 0.0,
    }
    x = np.array(list(cases), np.float32)
    assert np.array_equal(mlp.round_bf16(x), np.array(list(cases.values()), np.float32))


def test_round_bf16_matches_torch():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(100_000) * np.exp(rng.uniform(-20, 20, 100_000))).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(mlp.round_bf16(x), ref)


def test_split_bf16_two_words():
    """Layer-1 operands: x = hi + bf16(x - hi).  A hand-worked case, then the error bound of two
    bf16 words (each rounding is within 2^-8 relative, so 2^-16 |x|), which a split that
    dropped the low word (error up to 2^-8) fails."""
    x = np.float32(1.0 + 2.0**-8 + 2.0**-12)  # above the tie: hi = 1 + 2^-7
    assert mlp.round_bf16(x) == np.float32(1.0 + 2.0**-7)
    assert mlp.split_bf16(x) == 1.0 + 2.0**-8 + 2.0**-12  # lo = -(2^-8 - 2^-12), exact
    rng = np.random.default_rng(5)
    v = (rng.standard_normal(200_000) * np.exp(rng.uniform(-8, 8, 200_000))).astype(np.float32)
    rel = np.abs(mlp.split_bf16(v) - v.astype(np.float64)) / np.abs(v.astype(np.float64))
    assert rel.max() <= 2.0**-16 and rel.max() > 2.0**-19
    one_word = np.abs(mlp.round_bf16(v).astype(np.float64) - v) / np.abs(v.astype(np.float64))
    assert one_word.max() > 2.0**-10


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = mlp.init_params(3).astype(np.float64)
    xn = rng.standard_normal((128, 5))
    tn = rng.standard_normal((128, 4))
    _, g = mlp.loss_and_grad(p, xn, tn, emulate=False)
    checked = 0
    for name, (off, shape) in mlp.OFF.items():
        for k in rng.choice(int(np.prod(shape)), size=min(12, int(np.prod(shape))), replace=False):
            i = off + int(k)
            h = 1e-6 * max(1.0, abs(p[i]))
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (mlp.loss_only(pp, xn, tn) - mlp.loss_only(pm, xn, tn)) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])) + 1e-9, (name, k, fd, g[i])
            checked += 1
    assert checked > 50


def test_forward_closed_form_when_linear():
    """With every pre-activation positive the network is the affine map W3 (W2 (W1 x + b1) + b2)
    + b3; a transposed or mis-indexed weight block breaks this."""
    rng = np.random.default_rng(2)
    parts = {"w1": rng.uniform(0, 0.1, (64, 5)), "b1": np.full(64, 5.0),
             "w2": rng.uniform(0, 0.1, (64, 64)), "b2": np.full(64, 5.0),
             "w3": rng.standard_normal((4, 64)), "b3": rng.standard_normal(4)}
    p = mlp.pack(parts)
    xn = rng.uniform(-1, 1, (50, 5))
    y, _ = mlp.forward(p, xn, emulate=False)
    a = parts
    want = ((xn @ a["w1"].T + a["b1"]) @ a["w2"].T + a["b2"]) @ a["w3"].T + a["b3"]
    assert np.allclose(y, want, rtol=1e-12, atol=1e-12)


def test_envelope_selector_flags_injected_states():
    norm = np.zeros(32, np.float32)
    norm[5:10] = 1.0
    norm[14:18] = 1.0
    norm[18:23] = [-90, 0, 0, 0, -20]
    norm[23:28] = [50, 1, 1, 1, 40]
    x = np.array([[-65, 0.05, 0.6, 0.32, 6], [-65, 0.05, 0.6, 0.32, 41], [60, 0.5, 0.5, 0.5, 0],
                  [-65, 1.2, 0.6, 0.32, 6], [np.nan, 0.05, 0.6, 0.32, 6], [50, 1, 1, 1, 40]],
                 np.float32)
    assert mlp.inside(norm, x).tolist() == [True, False, False, False, False, True]


def test_displacement_metric():
    st = np.array([100, 300, 500, 120, 340], np.uint32)
    ids = np.array([0, 0, 0, 1, 1], np.uint32)
    assert spike_displacement(st, ids, st, ids, 0.05)["displaced_fraction"] == 0.0
    # within 0.5 ms (10 steps at 0.05 ms): not displaced; 11 steps: displaced both ways
    assert spike_displacement(st, ids, st + 10, ids, 0.05)["displaced_fraction"] == 0.0
    r = spike_displacement(st, ids, st + 11, ids, 0.05)
    assert (r["missing"], r["extra"], r["displaced_fraction"]) == (5, 5, 1.0)
    # same time, other neuron: displaced
    r = spike_displacement(st[:1], ids[:1], st[:1], ids[:1] + 1, 0.05)
    assert r["displaced_fraction"] == 1.0
    # one extra spike far from any reference spike
    r = spike_displacement(st, ids, np.append(st, 900), np.append(ids, 1), 0.05)
    assert (r["missing"], r["extra"]) == (0, 1)
    assert r["displaced_fraction"] == pytest.approx(1 / 11)
