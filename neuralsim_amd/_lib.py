"""Loader for the in-tree HIP library neuralsim_amd/lib/libnsx.so.

There is no CPU fallback: if the library is missing or cannot be loaded, every entry point
raises NativeLibraryMissing.  Build it with `python -c "import __graft_entry__ as g; g.build()"`
or `make -C neuralsim_amd/csrc`.
"""
import ctypes as C
import os

from . import abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libnsx.so")


class NativeLibraryMissing(RuntimeError):
    pass


class NsxError(RuntimeError):
    def __init__(self, status, message, call=""):
        self.status = status
        self.status_name = abi.STATUS_NAMES.get(status, str(status))
        super().__init__(f"{call}: {self.status_name}: {message}")


_lib = None

_VP = C.c_void_p
_P = C.c_void_p  # raw pointers (numpy .ctypes.data) are passed as void*
_u32, _u64, _i32, _f32, _f64, _int = C.c_uint32, C.c_uint64, C.c_int32, C.c_float, C.c_double, C.c_int

PROTOTYPES = {
    "nsx_abi_version": (_u32, []),
    "nsx_status_string": (C.c_char_p, [_int]),
    "nsx_last_error": (C.c_char_p, [_VP]),
    "nsx_device_count": (_int, [C.POINTER(_int)]),
    "nsx_config_init": (None, [C.POINTER(abi.NsxConfig), _u32]),
    "nsx_create": (_int, [C.POINTER(abi.NsxConfig), C.POINTER(_VP)]),
    "nsx_destroy": (None, [_VP]),
    "nsx_build_csr": (_int, [_VP, _P, _P, _P, _P, _u64]),
    "nsx_upload_csr": (_int, [_VP, _P, _P, _P, _P, _u64, _u64]),
    "nsx_generate_fixed_indegree": (_int, [_VP, _u32, _f32, _f32, _f32, _f32, _u32, _u32]),
    "nsx_read_csr": (_int, [_VP, _P, _P, _P, _P]),
    "nsx_network_size": (_int, [_VP, C.POINTER(_u64), C.POINTER(_u64)]),
    "nsx_reset_state": (_int, [_VP]),
    "nsx_upload_state": (_int, [_VP, C.POINTER(abi.NsxStateView)]),
    "nsx_read_state": (_int, [_VP, C.POINTER(abi.NsxStateView)]),
    "nsx_read_weights": (_int, [_VP, _P, _u64]),
    "nsx_upload_weights": (_int, [_VP, _P, _u64]),
    "nsx_stimulate": (_int, [_VP, _u64, _f32]),
    "nsx_get_step": (_int, [_VP, C.POINTER(_u64)]),
    "nsx_set_step": (_int, [_VP, _u64]),
    "nsx_step": (_int, [_VP, _u64, C.POINTER(abi.NsxStepStats)]),
    "nsx_simulate_ms": (_int, [_VP, _f64, C.POINTER(abi.NsxStepStats)]),
    "nsx_set_graph_steps": (_int, [_VP, _u32]),
    "nsx_set_prop_split": (_int, [_VP, _u32]),
    "nsx_set_update_width": (_int, [_VP, _u32]),
    "nsx_set_epoch": (_int, [_VP, _u32]),
    "nsx_set_row_split": (_int, [_VP, _u32]),
    "nsx_set_syn_block": (_int, [_VP, _u32]),
    "nsx_get_epoch": (_int, [_VP, C.POINTER(_u32)]),
    "nsx_set_persistent": (_int, [_VP, _int]),
    "nsx_get_persistent": (_int, [_VP, C.POINTER(_u32), C.POINTER(_u32)]),
    "nsx_set_delivery": (_int, [_VP, _u32, _u64, _u32]),
    "nsx_get_delivery": (_int, [_VP, C.POINTER(_u32)]),
    "nsx_read_raster": (_int, [_VP, _P, _P, _u64, C.POINTER(_u64), _int]),
    "nsx_read_step_counts": (_int, [_VP, _P, _u64, C.POINTER(_u64)]),
    "nsx_set_stats": (_int, [_VP, _int]),
    "nsx_read_sim_stats": (_int, [_VP, C.POINTER(abi.NsxSimStats)]),
    "nsx_weight_stats": (_int, [_VP, C.POINTER(_f64), C.POINTER(_f64)]),
    "nsx_weight_stats_noalloc": (_int, [_VP, C.POINTER(_f64), C.POINTER(_f64)]),
    "nsx_set_profiling": (_int, [_VP, _int]),
    "nsx_kernel_time": (_int, [_VP, _int, C.POINTER(_f64), C.POINTER(_u64), C.POINTER(_u64)]),
    "nsx_backend_new": (_int, [_u64, _u64, C.POINTER(_VP)]),
    "nsx_backend_upload_initial_state": (_int, [_VP, _P, _P, _P, _P, _P, _P, _P]),
    "nsx_backend_upload_hh_gates": (_int, [_VP, _P, _P, _P]),
    "nsx_backend_upload_csr": (_int, [_VP, _P, _P, _P]),
    "nsx_backend_upload_input_current": (_int, [_VP, _P]),
    "nsx_backend_write_uniforms": (_int, [_VP, _f32, _f32]),
    "nsx_backend_set_model": (_int, [_VP, _u32]),
    "nsx_backend_step_neurons": (_int, [_VP, _P, _P, _P, _P, _P]),
    "nsx_backend_step_spmv": (_int, [_VP, _P]),
    "nsx_autotune": (_int, [_VP, _u32, C.POINTER(abi.NsxTuneReport)]),
    "nsx_snapshot_save": (_int, [_VP]),
    "nsx_snapshot_restore": (_int, [_VP]),
    "nsx_batch_create": (_int, [C.POINTER(abi.NsxConfig), C.POINTER(abi.NsxBatchNet), _u32,
                                C.POINTER(_VP)]),
    "nsx_batch_layout": (_int, [_VP, _P]),
    "nsx_batch_stats": (_int, [_VP, _P, _P]),
    "nsx_surrogate_generate": (_int, [_VP, C.POINTER(abi.NsxGenConfig), C.POINTER(_u64)]),
    "nsx_surrogate_train": (_int, [_VP, C.POINTER(abi.NsxTrainConfig), _P]),
    "nsx_surrogate_get": (_int, [_VP, _P, _P]),
    "nsx_surrogate_set": (_int, [_VP, _P, _P]),
    "nsx_surrogate_forward": (_int, [_VP, _P, _u64, _u32, _P, _P]),
    "nsx_surrogate_gradient": (_int, [_VP, _P, _P, _u64, _P, C.POINTER(_f32)]),
    "nsx_surrogate_read_data": (_int, [_VP, _u64, _u64, _P]),
    "nsx_surrogate_fallbacks": (_int, [_VP, C.POINTER(_u64)]),
}


def load():
    """Load libnsx.so and declare every prototype; raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"{LIB_PATH} not found: the HIP library is not built (no CPU fallback exists)")
    try:
        lib = C.CDLL(LIB_PATH)
    except OSError as e:
        raise NativeLibraryMissing(f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.nsx_abi_version() != abi.NSX_ABI_VERSION:
        raise NativeLibraryMissing("libnsx.so ABI version mismatch; rebuild")
    _lib = lib
    return lib


def check(status, ctx=None, call=""):
    if status != abi.NSX_OK:
        lib = load()
        msg = lib.nsx_last_error(ctx)
        raise NsxError(status, msg.decode() if msg else "", call)
    return status


def device_count():
    lib = load()
    n = C.c_int(0)
    st = lib.nsx_device_count(C.byref(n))
    if st != abi.NSX_OK:
        return 0
    return n.value
