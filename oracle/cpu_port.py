"""ctypes wrapper of oracle/libcpuport.so -- the CPU BASELINE (bench.py's cpu_baseline legs).

The reference's CPU step (SimulationEngine::step, src/simulation/engine.rs:254-697) restated in
C++ with its data layout, OpenMP where the reference uses rayon (cpu_port.cpp).  Timed beside the
device as a baseline; nothing in the product calls it.
"""
import ctypes as C
import os
import platform
import subprocess

import numpy as np

from neuralsim_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcpuport.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE, "libcpuport.so"], check=True)
        L = C.CDLL(LIB_PATH)
        L.port_threads.restype = C.c_int
        L.port_create.restype = C.c_void_p
        L.port_create.argtypes = [C.POINTER(abi.NsxConfig), C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_int]
        L.port_destroy.argtypes = [C.c_void_p]
        L.port_set_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.port_run.restype = C.c_double
        L.port_run.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint64),
                               C.POINTER(C.c_uint64)]
        L.port_run_many.restype = C.c_double
        L.port_run_many.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_uint32, C.c_void_p]
        _lib = L
    return _lib


def host_info():
    """CPU model and logical core count of this host."""
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "threads": lib().port_threads()}


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class PortNet:
    """One reference-shaped CPU engine on a CSR network."""

    def __init__(self, cfg, row_ptr, col, w, delay=None, conductance=False):
        self.cfg = cfg.copy()
        keep = [np.ascontiguousarray(row_ptr, np.uint64), np.ascontiguousarray(col, np.uint32),
                np.ascontiguousarray(w, np.float32)]
        d = None if delay is None else np.ascontiguousarray(delay, np.uint8)
        self.m = len(keep[1])
        self._h = lib().port_create(C.byref(self.cfg), *[_p(x) for x in keep], _p(d),
                                    1 if conductance else 0)

    def set_state(self, v=None, u=None):
        v = None if v is None else np.ascontiguousarray(v, np.float32)
        u = None if u is None else np.ascontiguousarray(u, np.float32)
        lib().port_set_state(self._h, _p(v), _p(u))

    def run(self, steps):
        """Wall seconds, spikes and events (sum of out-degrees of spiking rows) of `steps`."""
        sp, ev = C.c_uint64(), C.c_uint64()
        s = lib().port_run(self._h, steps, C.byref(sp), C.byref(ev))
        return s, sp.value, ev.value

    def close(self):
        if self._h:
            lib().port_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_many(nets, steps):
    """Engines evaluated in parallel (population.rs:116-124): wall seconds and per-engine
    seconds."""
    arr = (C.c_void_p * len(nets))(*[n._h for n in nets])
    secs = np.zeros(len(nets))
    wall = lib().port_run_many(arr, len(nets), steps, _p(secs))
    return wall, secs
