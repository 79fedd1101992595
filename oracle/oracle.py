"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and
only as the checker.  Build: `make -C oracle` (done by __graft_entry__.build()).
"""
import ctypes as C
import os
import subprocess

import numpy as np

from neuralsim_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_VP, _P = C.c_void_p, C.c_void_p
_u32, _u64, _i32, _f32, _f64, _int = C.c_uint32, C.c_uint64, C.c_int32, C.c_float, C.c_double, C.c_int
_dp = C.POINTER(C.c_double)

PROTOS = {
    "orc_philox4x32_10": (None, [_P, _P, _P]),
    "orc_expf": (_f32, [_f32]),
    "orc_logf": (_f32, [_f32]),
    "orc_cospi": (_f32, [_f32]),
    "orc_normal": (_f32, [_u32, _u32]),
    "orc_normal64": (_f64, [_u32, _u32]),
    "orc_drive_normal": (_f32, [_u64, _u64, _u32]),
    "orc_poisson_table": (_int, [_f64, _P, _int]),
    "orc_hh_plain": (None, [C.POINTER(abi.NsxConfig), _P, _P, _P, _P, _f32, _u32]),
    "orc_hh_frozen": (None, [C.POINTER(abi.NsxConfig), _P, _P, _P, _P, _f32, _u32]),
    "ref_lif_step": (_int, [_dp, C.POINTER(_i32), _f64, _f64, _f64, _f64, _f64, _f64, _f64, _f64]),
    "ref_izh_step": (_int, [_dp, _dp, _f64, _f64, _f64, _f64, _f64, _f64]),
    "ref_hh_alpha_m": (_f64, [_f64]), "ref_hh_beta_m": (_f64, [_f64]),
    "ref_hh_alpha_h": (_f64, [_f64]), "ref_hh_beta_h": (_f64, [_f64]),
    "ref_hh_alpha_n": (_f64, [_f64]), "ref_hh_beta_n": (_f64, [_f64]),
    "ref_hh_step": (_int, [_dp, _dp, _dp, _dp] + [_f64] * 9),
    "ref_izh_step_sub": (_int, [_dp, _dp, _f64, _f64, _f64, _f64, _f64, _int, _f64]),
    "ref_hh_integrate": (_f64, [_dp, _dp, _dp, _dp] + [_f64] * 9 + [_int]),
    "ref_ou_update": (_f64, [_f64] * 6),
    "ref_finalize": (None, [_u64, _u64, _P, _P, _P, _P, _P]),
    "ref_mg_block": (_f64, [_f64, _f64]),
    "ref_conductance": (_f64, [_f64, _f64, _f64, _f64]),
    "ref_trace_on_pre": (None, [_dp, _dp, _f64, _f64]),
    "ref_trace_on_post": (None, [_dp, _dp, _f64, _f64]),
    "ref_trace_decay": (None, [_dp, _dp, _f64, _f64, _f64]),
    "ref_stdp_weight_change": (_f64, [_f64] * 5),
    "ref_triplet_weight_change": (_f64, [_f64] * 8),
    "ref_rate_homeostasis": (_f64, [_f64, _f64, _int]),
    "ref_stdp_literal": (None, [_u64, _P, _P, _P, _P, _u32, _P, _P, _u64] + [_f64] * 7),
    "ref_net_create": (_VP, [C.POINTER(abi.NsxConfig), _P, _P, _P, _P]),
    "ref_net_destroy": (None, [_VP]),
    "ref_net_run": (_u64, [_VP, _u32, _P, _P, _u64]),
    "ref_net_state": (None, [_VP, _P, _P, _P, _P, _P]),
    "ref_net_weights": (None, [_VP, _P]),
    "ref_net_drive": (None, [_VP, _P]),
    "mir_create": (_VP, [C.POINTER(abi.NsxConfig)]),
    "mir_destroy": (None, [_VP]),
    "mir_build_csr": (_int, [_VP, _P, _P, _P, _P, _u64]),
    "mir_generate_fixed_indegree": (_int, [_VP, _u32, _f32, _f32, _f32, _f32, _u32, _u32]),
    "mir_generate_list": (None, [C.POINTER(abi.NsxConfig), _u32, _f32, _f32, _f32, _f32, _u32,
                                 _u32, _P, _P, _P, _P]),
    "mir_num_synapses": (_u64, [_VP]),
    "mir_read_csr": (None, [_VP, _P, _P, _P, _P]),
    "mir_run": (_u64, [_VP, _u32, _P, _P, _u64, C.POINTER(_u64)]),
    "mir_read_state": (None, [_VP, C.POINTER(abi.NsxStateView)]),
    "mir_write_state": (None, [_VP, C.POINTER(abi.NsxStateView)]),
    "mir_stimulate": (None, [_VP, _u64, _f32]),
    "mir_step_index": (_u64, [_VP]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for k, (r, a) in PROTOS.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data)


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def poisson_table(lam, cap=64):
    t = np.zeros(cap, np.uint32)
    n = lib().orc_poisson_table(lam, _p(t), cap)
    return t[:n]


def finalize(n, src, dst):
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    m = len(src)
    rp = np.zeros(n + 1, np.uint64)
    col = np.zeros(m, np.uint32)
    perm = np.zeros(m, np.uint64)
    lib().ref_finalize(n, m, _p(src), _p(dst), _p(rp), _p(col), _p(perm))
    return rp, col, perm


def generate_list(spec):
    m = spec.m
    src, dst = np.zeros(m, np.uint32), np.zeros(m, np.uint32)
    w, d = np.zeros(m, np.float32), np.zeros(m, np.uint8)
    cfg = spec.cfg.copy()
    lib().mir_generate_list(C.byref(cfg), *spec.generator_args(), _p(src), _p(dst), _p(w), _p(d))
    return src, dst, w, d


def hh_step(cfg, state, I, substeps, frozen=False):
    """One step of the FP32 HH membrane integration per neuron (loop over a few neurons):
    plain Euler sub-steps or the frozen-rate sub-steps.  state: (count, 4) V m h n."""
    fn = lib().orc_hh_frozen if frozen else lib().orc_hh_plain
    out = np.array(state, np.float32).reshape(-1, 4).copy()
    I = np.asarray(I, np.float32).reshape(-1)
    c = cfg.copy()
    cells = [np.zeros(1, np.float32) for _ in range(4)]
    for i in range(out.shape[0]):
        for k in range(4):
            cells[k][0] = out[i, k]
        fn(C.byref(c), *[_p(x) for x in cells], float(I[i]), int(substeps))
        for k in range(4):
            out[i, k] = cells[k][0]
    return out


class MirrorNet:
    """FP32 mirror of the device engine (bit-exact target)."""

    def __init__(self, cfg):
        self.cfg = cfg.copy()
        self._h = lib().mir_create(C.byref(self.cfg))

    @classmethod
    def from_spec(cls, spec):
        m = cls(spec.cfg)
        rc = lib().mir_generate_fixed_indegree(m._h, *spec.generator_args())
        assert rc == 0
        return m

    def close(self):
        if self._h:
            lib().mir_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_csr(self, src, dst, w, d):
        a = [np.ascontiguousarray(src, np.uint32), np.ascontiguousarray(dst, np.uint32),
             np.ascontiguousarray(w, np.float32), np.ascontiguousarray(d, np.uint8)]
        return lib().mir_build_csr(self._h, *[_p(x) for x in a], len(a[0]))

    def read_csr(self):
        n = int(self.cfg.n_neurons)
        m = lib().mir_num_synapses(self._h)
        rp, col = np.zeros(n + 1, np.uint64), np.zeros(m, np.uint32)
        w, d = np.zeros(m, np.float32), np.zeros(m, np.uint8)
        lib().mir_read_csr(self._h, _p(rp), _p(col), _p(w), _p(d))
        return rp, col, w, d

    def run(self, n_steps, cap=None):
        n = int(self.cfg.n_neurons)
        cap = cap if cap is not None else n * n_steps
        cap = max(cap, 1)
        st, ids = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
        ev = C.c_uint64()
        k = lib().mir_run(self._h, n_steps, _p(st), _p(ids), cap, C.byref(ev))
        k = min(k, cap)
        return st[:k], ids[:k], ev.value

    def read_state(self, fields=None, ring=False):
        n = int(self.cfg.n_neurons)
        fields = fields or list(abi.STATE_FIELDS)
        view, out = abi.NsxStateView(), {}
        for f in fields:
            a = np.zeros(n, abi.STATE_FIELDS[f])
            out[f] = a
            setattr(view, f, C.cast(a.ctypes.data, type(getattr(view, f))))
        if ring:
            r = np.zeros(int(self.cfg.ring_slots) * abi.channels(self.cfg) * n, np.int64)
            out["ring"] = r
            view.ring = C.cast(r.ctypes.data, C.POINTER(C.c_int64))
        lib().mir_read_state(self._h, C.byref(view))
        return out

    def write_state(self, state):
        view, keep = abi.NsxStateView(), []
        for f, a in state.items():
            if f == "ring":
                a = np.ascontiguousarray(a, np.int64)
                view.ring = C.cast(a.ctypes.data, C.POINTER(C.c_int64))
            else:
                a = np.ascontiguousarray(a, abi.STATE_FIELDS[f])
                setattr(view, f, C.cast(a.ctypes.data, type(getattr(view, f))))
            keep.append(a)
        lib().mir_write_state(self._h, C.byref(view))

    def stimulate(self, i, current):
        lib().mir_stimulate(self._h, i, current)

    def weights(self):
        return self.read_csr()[2]


class RefNet:
    """FP64 network run with the reference's step functions (tolerance target)."""

    def __init__(self, cfg, row_ptr, col, w, d):
        self.cfg = cfg.copy()
        self._keep = [np.ascontiguousarray(row_ptr, np.uint64), np.ascontiguousarray(col, np.uint32),
                      np.ascontiguousarray(w, np.float32), np.ascontiguousarray(d, np.uint8)]
        self.m = len(self._keep[1])
        self._h = lib().ref_net_create(C.byref(self.cfg), *[_p(x) for x in self._keep])

    def close(self):
        if self._h:
            lib().ref_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, n_steps, cap=None):
        n = int(self.cfg.n_neurons)
        cap = max(cap if cap is not None else n * n_steps, 1)
        st, ids = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
        k = min(lib().ref_net_run(self._h, n_steps, _p(st), _p(ids), cap), cap)
        return st[:k], ids[:k]

    def state(self):
        n = int(self.cfg.n_neurons)
        a = [np.zeros(n) for _ in range(5)]
        lib().ref_net_state(self._h, *[_p(x) for x in a])
        return dict(zip(["v", "u", "m", "h", "n"], a))

    def weights(self):
        w = np.zeros(self.m)
        lib().ref_net_weights(self._h, _p(w))
        return w

    def drive(self):
        x = np.zeros(int(self.cfg.n_neurons))
        lib().ref_net_drive(self._h, _p(x))
        return x
