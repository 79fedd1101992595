"""Learned HH surrogate on the device, through the C ABI.

* the MFMA forward pass and the hand-written backward pass against the FP64 MLP of oracle/mlp.py
  (which emulates the device's bf16 roundings and is itself pinned by finite differences);
* the frozen-rate sub-steps bit-exact against the FP32 mirror; the network step bit-exact
  against nsx_surrogate_forward, and the envelope fallback bit-exact against the plain step;
* training: deterministic, converging;
* accuracy: spikes displaced by more than 0.5 ms against the FP64 HH oracle (same drive), for
  the surrogate, for the same integrator without its learned correction, and for a plain single
  Euler step of the same length.
"""
import numpy as np
import pytest

from neuralsim_amd import abi, configs, surrogate
from neuralsim_amd.metrics import spike_displacement
from oracle import mlp

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _uncoupled(n, **kw):
    return configs.hh_surrogate_network(n=n, k_in=1, w_exc=(0.0, 0.0), w_inh=(0.0, 0.0), **kw)


@pytest.fixture(scope="module")
def small(nsx):
    """An engine with a 2M-sample training set and random (untrained) parameters."""
    print("fixture small: generate 2M samples", flush=True)
    eng = nsx.SimulationEngine.from_spec(_uncoupled(1024))
    n = eng.surrogate_generate(2_000_000, 3)
    data = eng.surrogate_data(0, 4096)
    norm = eng.surrogate_norm()
    p0 = mlp.init_params(7)
    eng.set_surrogate_params(p0, norm)
    yield eng, data, norm, p0, n
    eng.close()


def test_training_set(small):
    eng, data, norm, _, n = small
    assert n >= 1_800_000 and np.isfinite(data).all()
    nm = mlp.unpack_norm(norm)
    assert np.all(nm["lo"] < nm["hi"]) and np.all(nm["out_sd"] > 0)
    # states lie inside their own envelope; gates in [0, 1]
    assert mlp.inside(norm, data[:, :5]).all()
    assert data[:, 1:4].min() >= 0.0 and data[:, 1:4].max() <= 1.0


def test_forward_matches_oracle(small):
    eng, data, norm, p0, _ = small
    x = data[:, :5]
    y, inside = eng.surrogate_forward(x, 0)
    ref = mlp.predict(p0, norm, x, emulate=True)
    err = np.abs(y - ref) / mlp.unpack_norm(norm)["out_sd"]
    print("forward error (units of out_sd): median %.3g p99 %.3g max %.3g" %
          (np.median(err), np.percentile(err, 99), err.max()))
    # FP32-accumulated bf16 products vs FP64 sums; an activation rounding to the neighbouring
    # bf16 value now and then gives the tail
    assert np.median(err) < 1e-6 and np.percentile(err, 99) < 1e-5 and err.max() < 1e-3
    assert inside.all()


def test_gradient_matches_oracle(small):
    eng, data, norm, p0, _ = small
    xb, tb = data[:1024, :5], data[:1024, 5:]
    g, loss = eng.surrogate_gradient(xb, tb)
    lref, gref = mlp.loss_and_grad(p0, mlp.normalise_inputs(norm, xb),
                                   mlp.normalise_targets(norm, tb), emulate=True)
    assert loss == pytest.approx(lref, rel=1e-5)
    for k, (o, s) in mlp.OFF.items():
        m = int(np.prod(s))
        a, b = g[o:o + m].astype(np.float64), gref[o:o + m]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"gradient {k}: relative L2 error {rel:.3g}")
        # the device rounds dL/dy and dL/dz2 to bf16 for the MFMAs (2^-9 relative each)
        assert rel < 5e-3, k


def test_frozen_and_next_state_bitexact(small, orc):
    eng, data, norm, _, _ = small
    nm = mlp.unpack_norm(norm)
    x = data[:256, :5].copy()
    out = np.arange(0, 256, 9)  # inject states outside the envelope
    x[out[::2], 0] = nm["hi"][0] + 3.0
    x[out[1::2], 4] = nm["lo"][4] - 2.0
    frozen, _ = eng.surrogate_forward(x, 2)
    nxt, inside = eng.surrogate_forward(x, 1)
    corr, _ = eng.surrogate_forward(x, 0)
    assert inside.sum() == 256 - len(out) and not inside[out].any()
    cfg = eng.cfg
    want_frozen = orc.hh_step(cfg, x[:, :4], x[:, 4], 5, frozen=True)
    assert np.array_equal(_bits(frozen), _bits(want_frozen))
    # inside: frozen step + output * (|dV| + eps), gates clamped; outside: the plain step
    f32 = np.float32
    sc = (np.abs(frozen[:, 0] - x[:, 0]).astype(f32) + f32(nm["scale_eps"][0])).astype(f32)
    want = np.empty_like(nxt)
    want[:, 0] = (frozen[:, 0] + (corr[:, 0] * sc).astype(f32)).astype(f32)
    for k in (1, 2, 3):
        want[:, k] = np.clip((frozen[:, k] + (corr[:, k] * sc).astype(f32)).astype(f32), f32(0),
                             f32(1))
    want[out] = orc.hh_step(cfg, x[out, :4], x[out, 4], 5)
    assert np.array_equal(_bits(nxt), _bits(want))


def test_network_step_bitexact_with_fallback(nsx, orc, small):
    """One network step: every neuron equals nsx_surrogate_forward on its (V, m, h, n, I); the
    neurons pushed outside the envelope equal the plain step of the FP32 mirror."""
    _, _, norm, p0, _ = small
    spec = _uncoupled(600)
    hi_v = mlp.unpack_norm(norm)["hi"][0]
    with nsx.SimulationEngine.from_spec(spec) as eng:
        eng.set_surrogate_params(p0, norm)
        st = eng.read_state(["v", "m", "h", "n", "drive"])
        rng = np.random.default_rng(4)
        st["v"] = (-65.0 + 8.0 * rng.standard_normal(600)).astype(np.float32)
        st["h"] = rng.uniform(0.3, 0.7, 600).astype(np.float32)
        out = np.array([0, 5, 63, 64, 255, 256, 599])
        st["v"][out] = hi_v + 5.0
        eng.upload_state(st)
        x = np.stack([st["v"], st["m"], st["h"], st["n"], st["drive"]], 1)
        want, inside = eng.surrogate_forward(x, 1)
        eng.step(1)
        got = eng.read_state(["v", "m", "h", "n", "drive"])
        assert eng.surrogate_fallbacks() == len(out)
    assert not inside[out].any() and inside.sum() == 600 - len(out)
    for k, f in enumerate(["v", "m", "h", "n"]):
        assert np.array_equal(_bits(got[f]), _bits(want[:, k])), f
    cfg = spec.cfg.copy()
    cfg.model = abi.MODEL_HH
    mir = orc.MirrorNet(cfg)
    assert orc.lib().mir_generate_fixed_indegree(mir._h, *spec.generator_args()) == 0
    mir.write_state(st)
    mir.run(1)
    m_st = mir.read_state(["v", "m", "h", "n", "drive"])
    for f in ["v", "m", "h", "n", "drive"]:
        assert np.array_equal(_bits(got[f][out]), _bits(m_st[f][out])), f
    assert np.array_equal(_bits(got["drive"]), _bits(m_st["drive"]))


def test_training_deterministic_and_converging(small):
    eng, _, norm, p0, _ = small
    h1 = eng.surrogate_train(300, waves=256, chunks_per_wave=2, seed=9)
    a, _ = eng.surrogate_params()
    h2 = eng.surrogate_train(300, waves=256, chunks_per_wave=2, seed=9)
    b, _ = eng.surrogate_params()
    eng.set_surrogate_params(p0, norm)
    assert np.array_equal(a, b) and np.array_equal(h1, h2)
    assert h1[-20:].mean() < h1[:5].mean() / 5


@pytest.fixture(scope="module")
def trained(nsx):
    spec = _uncoupled(2000, record_raster=True, raster_capacity=1 << 22)
    print("fixture trained: full recipe", configs.SURROGATE_RECIPE, flush=True)
    eng = nsx.SimulationEngine.from_spec(spec)
    fit = surrogate.fit(eng)
    print("fit: %d samples, %.2f s training, final loss %.3g" %
          (fit["samples"], fit["train_s"], fit["loss"][-200:].mean()), flush=True)
    params, norm = eng.surrogate_params()
    eng.close()
    return spec, params, norm, fit


def _raster(nsx, spec, steps, params=None, norm=None):
    with nsx.SimulationEngine.from_spec(spec) as eng:
        if params is not None:
            eng.set_surrogate_params(params, norm)
        eng.step(steps)
        st, ids = eng.raster()
        fb = eng.surrogate_fallbacks() if params is not None else 0
        csr = eng.read_csr()
    return st, ids, fb, csr


def test_spike_displacement_and_gain(nsx, orc, trained):
    """500 ms of 2000 uncoupled neurons under the OU drive (6, 3, 5 ms).  Reference: the FP64
    oracle integrating five 0.01 ms Euler sub-steps per 0.05 ms step with the same drive."""
    spec, params, norm, fit = trained
    steps = 10_000
    s_st, s_id, fb, csr = _raster(nsx, spec, steps, params, norm)
    # the same integrator with the learned correction removed: zero network, zero offset
    zero_norm = norm.copy()
    zero_norm[10:14] = 0.0  # out_mu
    z_st, z_id, _, _ = _raster(nsx, spec, steps, np.zeros_like(params), zero_norm)
    plain = _uncoupled(2000, record_raster=True, raster_capacity=1 << 22)
    plain.cfg.model = abi.MODEL_HH
    p5_st, p5_id, _, _ = _raster(nsx, plain, steps)
    plain.cfg.hh_substeps = 1
    p1_st, p1_id, _, _ = _raster(nsx, plain, steps)
    ref_cfg = plain.cfg.copy()
    ref_cfg.hh_substeps = 5
    r_st, r_id = orc.RefNet(ref_cfg, *csr).run(steps)
    res = {}
    for name, (a, b) in {"surrogate": (s_st, s_id), "frozen_only": (z_st, z_id),
                         "plain_fp32": (p5_st, p5_id), "euler_0.05": (p1_st, p1_id)}.items():
        res[name] = spike_displacement(r_st, r_id, a, b, 0.05, 0.5)
        print(name, res[name])
    assert len(r_st) > 20_000
    assert fb < 0.001 * 2000 * steps
    assert res["plain_fp32"]["displaced_fraction"] < 0.001
    assert res["surrogate"]["displaced_fraction"] <= 0.01
    gain = res["frozen_only"]["displaced_fraction"] / max(res["surrogate"]["displaced_fraction"],
                                                           1e-9)
    assert gain >= 5.0
    assert res["euler_0.05"]["displaced_fraction"] >= 5 * res["surrogate"]["displaced_fraction"]


@pytest.mark.parametrize("plastic", [False, True])
def test_coupled_surrogate_network_same_in_every_delivery(nsx, trained, plastic):
    """A coupled 3000-neuron surrogate network (AMPA/GABA-A/NMDA input through the delay ring,
    optionally pair STDP): direct launches with atomic delivery, graph windows with atomic
    delivery and graph windows with binned delivery give the same raster, state, ring and
    weights bit for bit."""
    _, params, norm, _ = trained
    spec = configs.hh_surrogate_network(n=3000, seed=13, record_raster=True,
                                        raster_capacity=1 << 22, w_exc=(0.0, 0.02),
                                        w_inh=(-0.05, 0.0))
    if plastic:
        spec.cfg.plasticity = abi.PLAST_PAIR
        spec.cfg.w_min, spec.cfg.w_max = 0.0, 0.05
    fields = ["v", "m", "h", "n", "syn0", "syn1", "syn2", "drive", "spike_count"]
    out = []
    for mode, graph, te in [(abi.DELIVERY_ATOMIC, 0, 0), (abi.DELIVERY_ATOMIC, 50, 0),
                            (abi.DELIVERY_BINNED, 50, 4096)]:
        print(f"case coupled surrogate plastic={plastic} mode={mode} graph={graph}", flush=True)
        with nsx.SimulationEngine.from_spec(spec) as eng:
            eng.set_surrogate_params(params, norm)
            eng.set_delivery(mode, 0, te)
            eng.set_graph_steps(graph)
            st = eng.step(700)
            st = eng.step(1300)
            out.append((st, eng.raster(), eng.weights(), eng.read_state(fields, ring=True)))
    (st0, (s0, i0), w0, x0) = out[0]
    assert st0.total_spikes > 3000 and st0.events > 100_000
    assert np.any(x0["syn1"] != 0)  # inhibitory input arrived
    for st, (s, i), w, x in out[1:]:
        assert st.total_spikes == st0.total_spikes and st.total_events == st0.total_events
        assert np.array_equal(s, s0) and np.array_equal(i, i0), "raster differs"
        assert np.array_equal(w.view(np.uint32), w0.view(np.uint32)), "weights differ"
        for k in x0:
            assert np.array_equal(np.ascontiguousarray(x[k]).view(np.uint8),
                                  np.ascontiguousarray(x0[k]).view(np.uint8)), k
    if plastic:
        with nsx.SimulationEngine.from_spec(spec) as eng:
            w_init = eng.weights()
        assert np.any(w0 != w_init)  # STDP changed weights
