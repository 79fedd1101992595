"""C ABI checks that need no GPU: the library loads, exports every entry point the header
declares, and reports errors as statuses (never aborts) when there is no device or an argument
is bad."""
import ctypes as C
import os
import re

import pytest

import neuralsim_amd as ns
from neuralsim_amd import abi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "nsx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nsx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "nsx_create" in names and "nsx_backend_step_spmv" in names
    assert len(names) >= 35


def test_library_exports_every_declared_symbol(nsx):
    lib = nsx.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.nsx_abi_version() == abi.NSX_ABI_VERSION
    # every exported nsx_* prototype is declared in the ctypes loader
    from neuralsim_amd._lib import PROTOTYPES
    assert set(declared_functions()) <= set(PROTOTYPES)


def test_config_struct_layout_matches_library(nsx):
    lib = nsx.load()
    c = abi.NsxConfig()
    lib.nsx_config_init(C.byref(c), abi.MODEL_HH)
    assert c.struct_size == C.sizeof(abi.NsxConfig)
    # defaults from the reference (engine.rs:96-99, hodgkin_huxley.rs:98-114, plasticity.rs:17-28)
    assert c.dt_ms == 0.5 and c.seed == 42
    assert abs(c.hh_e_l + 54.387) < 1e-5 and c.hh_g_na == 120.0
    assert abs(c.a_plus - 0.01) < 1e-9 and abs(c.a_minus - 0.012) < 1e-9
    py = ns.configs.base_config(abi.MODEL_HH)
    for name, _ in abi.NsxConfig._fields_:
        if name in ("model", "_reserved"):
            continue
        a, b = getattr(c, name), getattr(py, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name


def _no_gpu(nsx):
    return nsx.device_count() == 0


def test_status_strings_and_thread_error(nsx):
    lib = nsx.load()
    for s in abi.STATUS_NAMES:
        assert lib.nsx_status_string(s)
    assert lib.nsx_last_error(None) is not None


def test_invalid_configs_rejected_before_device(nsx):
    lib = nsx.load()
    base = ns.configs.izh_stdp_network(n=100).cfg
    bad = []
    c = base.copy(); c.struct_size = 12; bad.append(c)
    c = base.copy(); c.n_neurons = 0; bad.append(c)
    c = base.copy(); c.n_exc = 101; bad.append(c)
    c = base.copy(); c.ring_slots = 1; bad.append(c)
    c = base.copy(); c.ring_slots = 300; bad.append(c)
    c = base.copy(); c.dt_ms = 0.0; bad.append(c)
    c = base.copy(); c.v_substeps = 0; bad.append(c)
    c = base.copy(); c.model = 9; bad.append(c)
    c = base.copy(); c.drive = abi.DRIVE_POISSON; c.drive_inputs = 10**6; c.drive_rate_hz = 100
    bad.append(c)
    for c in bad:
        h = C.c_void_p()
        st = lib.nsx_create(C.byref(c), C.byref(h))
        assert st == abi.NSX_E_INVALID, (st, lib.nsx_last_error(None))
        assert not h.value
        assert lib.nsx_last_error(None)


def test_null_arguments_return_status(nsx):
    lib = nsx.load()
    assert lib.nsx_create(None, None) == abi.NSX_E_INVALID
    assert lib.nsx_step(None, 1, None) == abi.NSX_E_INVALID
    assert lib.nsx_read_state(None, None) == abi.NSX_E_INVALID
    assert lib.nsx_backend_new(0, 0, None) == abi.NSX_E_INVALID
    n = C.c_int(-1)
    assert lib.nsx_device_count(None) == abi.NSX_E_INVALID
    assert lib.nsx_set_persistent(None, 1) == abi.NSX_E_INVALID
    assert lib.nsx_get_persistent(None, None, None) == abi.NSX_E_INVALID
    lib.nsx_destroy(None)  # no-op


def test_no_device_is_a_status_not_an_abort(nsx):
    if not _no_gpu(nsx):
        pytest.skip("a device is present")
    spec = ns.configs.izh_stdp_network(n=100)
    with pytest.raises(ns.NsxError) as e:
        ns.SimulationEngine(spec.cfg)
    assert e.value.status_name == "NSX_E_NODEVICE"
    with pytest.raises(ns.NsxError) as e:
        ns.GpuBackend(10, 10)
    assert e.value.status_name == "NSX_E_NODEVICE"


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import neuralsim_amd._lib as L
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "libnsx.so"))
    with pytest.raises(L.NativeLibraryMissing):
        L.load()


def test_workload_specs():
    c1 = ns.configs.izh_stdp_network()
    assert (c1.n, c1.m, c1.cfg.n_exc, c1.cfg.ring_slots) == (10**6, 10**8, 800_000, 21)
    assert c1.cfg.dt_ms == 1.0 and c1.cfg.v_substeps == 2 and c1.sim_ms == 10_000
    c0 = ns.configs.lif_network()
    assert (c0.n, c0.m, c0.cfg.ring_slots, c0.delay_steps) == (10_000, 10**6, 51, (10, 50))
    assert abs(c0.cfg.drive_inputs * c0.cfg.drive_rate_hz * c0.cfg.dt_ms / 1000 - 0.05) < 1e-12
    c2 = ns.configs.hh_network()
    assert (c2.n, c2.m, c2.cfg.dt_ms, c2.steps) == (200_000, 2 * 10**7, 0.01, 100_000)
    assert c2.delay_steps == (100, 200) and c2.cfg.ring_slots == 201
