"""The allocation-free weight statistics (nsx_weight_stats_noalloc) through the C ABI: equal,
bit for bit, to nsx_weight_stats on the same engine, on two calls in a row -- the second on a
smaller network -- both sharing the library's static scratch array."""
import ctypes as C

import numpy as np
import pytest

from neuralsim_amd import configs

pytestmark = pytest.mark.gpu


def _both(eng):
    lib = eng._lib
    m0, s0, m1, s1 = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    assert lib.nsx_weight_stats(eng._h, C.byref(m0), C.byref(s0)) == 0
    assert lib.nsx_weight_stats_noalloc(eng._h, C.byref(m1), C.byref(s1)) == 0
    return (m0.value, s0.value), (m1.value, s1.value)


def test_weight_stats_noalloc_matches(nsx):
    out = []
    for n, k_in in ((20_000, 100), (3_000, 20)):  # second call: a smaller input
        spec = configs.izh_stdp_network(n=n, seed=5, k_in=k_in)
        print(f"case noalloc {spec.name} m={spec.m}", flush=True)
        with nsx.SimulationEngine.from_spec(spec) as eng:
            eng.step(120)  # pair STDP has changed the weights
            ref, got = _both(eng)
            w = eng.weights().astype(np.float64)
        assert np.array_equal(np.array(ref).view(np.uint64), np.array(got).view(np.uint64))
        assert abs(got[0] - w.mean()) < 1e-12 * max(1.0, abs(w.mean()))
        assert abs(got[1] - w.std()) < 1e-9
        out.append(got)
    assert out[0] != out[1]
