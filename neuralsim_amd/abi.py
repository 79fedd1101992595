"""ctypes mirror of include/nsx.h (types and constants only; loads nothing)."""
import ctypes as C

NSX_ABI_VERSION = 3

# nsx_status
NSX_OK = 0
NSX_E_NODEVICE = -1
NSX_E_OOM = -2
NSX_E_INVALID = -3
NSX_E_BADCSR = -4
NSX_E_HIP = -5
NSX_E_STATE = -6
NSX_E_OVERFLOW = -7

STATUS_NAMES = {
    NSX_OK: "NSX_OK", NSX_E_NODEVICE: "NSX_E_NODEVICE", NSX_E_OOM: "NSX_E_OOM",
    NSX_E_INVALID: "NSX_E_INVALID", NSX_E_BADCSR: "NSX_E_BADCSR", NSX_E_HIP: "NSX_E_HIP",
    NSX_E_STATE: "NSX_E_STATE", NSX_E_OVERFLOW: "NSX_E_OVERFLOW",
}

# nsx_model
MODEL_LIF = 0
MODEL_IZHIKEVICH = 1
MODEL_HH = 2
MODEL_HH_SURROGATE = 3
# nsx_drive_kind
DRIVE_NONE = 0
DRIVE_POISSON = 1
DRIVE_GAUSSIAN = 2
DRIVE_OU = 3
# nsx_plasticity_kind
PLAST_NONE = 0
PLAST_PAIR = 1
PLAST_TRIPLET = 2

DELIVERY_ATOMIC = 0  # nsx_set_delivery: one 64-bit ring atomic per event
DELIVERY_BINNED = 1  # events bucketed by (delay, target tile), summed in LDS at arrival
DELIVERY_PULL = 2    # no plasticity: each target tile gathers its input from recent spikes
# nsx_hh_spike_mode
HH_SPIKE_CROSSING = 0
HH_SPIKE_REFERENCE_RESET = 1

NEVER_STEP = -(1 << 30)


class NsxConfig(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("model", C.c_uint32),
        ("n_neurons", C.c_uint64),
        ("n_exc", C.c_uint64),
        ("dt_ms", C.c_double),
        ("v_substeps", C.c_uint32),
        ("ring_slots", C.c_uint32),
        ("frac_bits", C.c_uint32),
        ("record_raster", C.c_uint32),
        ("seed", C.c_uint64),
        ("raster_capacity", C.c_uint64),
        ("lif_v_rest", C.c_float), ("lif_v_reset", C.c_float), ("lif_v_th", C.c_float),
        ("lif_tau_m", C.c_float), ("lif_t_ref_ms", C.c_float), ("lif_r", C.c_float),
        ("tau_exc_ms", C.c_float), ("tau_inh_ms", C.c_float), ("tau_nmda_ms", C.c_float),
        ("izh_a", C.c_float * 2), ("izh_b", C.c_float * 2), ("izh_c", C.c_float * 2),
        ("izh_d", C.c_float * 2),
        ("izh_c_jitter", C.c_float), ("izh_d_jitter_frac", C.c_float),
        ("hh_g_na", C.c_float), ("hh_g_k", C.c_float), ("hh_g_l", C.c_float),
        ("hh_e_na", C.c_float), ("hh_e_k", C.c_float), ("hh_e_l", C.c_float),
        ("hh_c_m", C.c_float),
        ("hh_spike_v", C.c_float),
        ("hh_spike_mode", C.c_uint32),
        ("e_exc", C.c_float), ("e_inh", C.c_float),
        ("gmax_ampa", C.c_float), ("gmax_nmda", C.c_float), ("gmax_gaba", C.c_float),
        ("mg_mM", C.c_float),
        ("hh_substeps", C.c_uint32),
        ("drive", C.c_uint32),
        ("drive_inputs", C.c_uint32),
        ("drive_rate_hz", C.c_float),
        ("drive_weight", C.c_float),
        ("drive_mean", C.c_float), ("drive_sigma", C.c_float), ("drive_tau_ms", C.c_float),
        ("plasticity", C.c_uint32),
        ("a_plus", C.c_float), ("a_minus", C.c_float), ("tau_plus_ms", C.c_float),
        ("tau_minus_ms", C.c_float), ("w_min", C.c_float), ("w_max", C.c_float),
        ("tau_y_ms", C.c_float), ("a3_plus", C.c_float),
        ("_reserved", C.c_uint32 * 8),
    ]

    def copy(self):
        c = NsxConfig()
        C.memmove(C.byref(c), C.byref(self), C.sizeof(NsxConfig))
        return c

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            if name.startswith("_"):
                continue
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class NsxStepStats(C.Structure):
    _fields_ = [
        ("steps_done", C.c_uint64),
        ("step_index", C.c_uint64),
        ("spikes", C.c_uint64),
        ("events", C.c_uint64),
        ("total_spikes", C.c_uint64),
        ("total_events", C.c_uint64),
        ("sim_time_ms", C.c_double),
        ("wall_ms", C.c_double),
    ]


class NsxSimStats(C.Structure):
    """nsx_sim_stats (include/nsx.h): SimulationStats, src/simulation/engine.rs:21-33."""
    _fields_ = [
        ("total_spikes", C.c_uint64),
        ("output_spikes", C.c_uint64),
        ("active_neurons", C.c_uint64),
        ("weight_updates", C.c_uint64),
        ("sim_time_ms", C.c_double),
        ("mean_firing_rate", C.c_double),
        ("synchrony_index", C.c_double),
        ("weight_mean", C.c_double),
        ("weight_std", C.c_double),
        ("lfp", C.c_double),
        ("seed", C.c_uint64),
        ("steps", C.c_uint64),
    ]


class NsxTuneReport(C.Structure):
    """nsx_tune_report (include/nsx.h)."""
    _fields_ = [
        ("prop_split", C.c_uint32),
        ("graph_steps", C.c_uint32),
        ("n_trials", C.c_uint32),
        ("_pad", C.c_uint32),
        ("best_us_per_step", C.c_double),
        ("us_per_step", C.c_double * 48),
        ("trial_split", C.c_uint32 * 48),
        ("trial_graph", C.c_uint32 * 48),
        ("delivery", C.c_uint32),
        ("tile_entries", C.c_uint32),
        ("trial_tile", C.c_uint32 * 48),
        ("row_split", C.c_uint32),
        ("syn_block", C.c_uint32),
        ("update_width", C.c_uint32),
        ("persistent", C.c_uint32),
        ("trial_row_split", C.c_uint32 * 48),
        ("trial_syn_block", C.c_uint32 * 48),
        ("trial_width", C.c_uint32 * 48),
        ("trial_delivery", C.c_uint32 * 48),
        ("trial_persistent", C.c_uint32 * 48),
    ]


class NsxBatchNet(C.Structure):
    """nsx_batch_net (include/nsx.h)."""
    _fields_ = [
        ("n_neurons", C.c_uint32),
        ("n_exc", C.c_uint32),
        ("k_in", C.c_uint32),
        ("d_lo", C.c_uint32),
        ("d_hi", C.c_uint32),
        ("_pad", C.c_uint32),
        ("seed", C.c_uint64),
        ("w_exc_lo", C.c_float),
        ("w_exc_hi", C.c_float),
        ("w_inh_lo", C.c_float),
        ("w_inh_hi", C.c_float),
    ]


SURROGATE_PARAMS = 4804
SURROGATE_NORM = 32


class NsxGenConfig(C.Structure):
    """nsx_gen_config (include/nsx.h)."""
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("stride", C.c_uint32),
        ("n_samples", C.c_uint64),
        ("seed", C.c_uint64),
        ("random_fraction", C.c_float),
        ("scale_eps", C.c_float),
        ("warm_ms", C.c_float),
        ("mu_lo", C.c_float),
        ("mu_hi", C.c_float),
        ("sigma_hi", C.c_float),
        ("tau_lo", C.c_float),
        ("tau_hi", C.c_float),
        ("_reserved", C.c_uint32 * 8),
    ]


class NsxTrainConfig(C.Structure):
    """nsx_train_config (include/nsx.h)."""
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("steps", C.c_uint32),
        ("waves", C.c_uint32),
        ("chunks_per_wave", C.c_uint32),
        ("init", C.c_uint32),
        ("sample_offset", C.c_uint32),
        ("lr", C.c_float),
        ("lr_min", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("eps", C.c_float),
        ("_pad", C.c_uint32),
        ("seed", C.c_uint64),
        ("_reserved", C.c_uint32 * 8),
    ]


_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int32)
_UP = C.POINTER(C.c_uint32)


class NsxStateView(C.Structure):
    _fields_ = [
        ("v", _FP), ("u", _FP), ("m", _FP), ("h", _FP), ("n", _FP),
        ("syn0", _FP), ("syn1", _FP), ("syn2", _FP), ("drive", _FP),
        ("refractory", _IP), ("last_spike_step", _IP), ("spike_count", _UP),
        ("pre_trace", _FP), ("pre_step", _IP), ("post_trace", _FP), ("post_step", _IP),
        ("post_trace2", _FP),
        ("ring", C.POINTER(C.c_int64)),
        ("izh_c", _FP), ("izh_d", _FP),
    ]


# field -> (numpy dtype, per-neuron) for NsxStateView
STATE_FIELDS = {
    "v": "float32", "u": "float32", "m": "float32", "h": "float32", "n": "float32",
    "syn0": "float32", "syn1": "float32", "syn2": "float32", "drive": "float32",
    "refractory": "int32", "last_spike_step": "int32", "spike_count": "uint32",
    "pre_trace": "float32", "pre_step": "int32", "post_trace": "float32", "post_step": "int32",
    "izh_c": "float32", "izh_d": "float32",
}


def channels(cfg):
    return 1 if cfg.model == MODEL_IZHIKEVICH else 2
