"""ctypes binding of the engine's C ABI (include/neucor_amd.h).

The shared library is built in-tree by ``neurocorrelation_amd/build.py``.
Loading fails loudly when it is missing: there is no fallback path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libneucor_amd.so")

NC_OK = 0
NC_NEVER_STEP = -(1 << 30)
NC_NUM_COUNTERS = 19
STATUS = {
    0: "NC_OK", 1: "NC_ERR_ARG", 2: "NC_ERR_RANGE", 3: "NC_ERR_HIP", 4: "NC_ERR_OOM",
    5: "NC_ERR_CAPACITY", 6: "NC_ERR_STATE", 7: "NC_ERR_NO_DEVICE",
}


class NCError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (STATUS.get(code, code), msg))
        self.code = code


class WiringSpec(C.Structure):
    _fields_ = [
        ("n_neurons", C.c_uint32), ("spatial_order", C.c_uint32), ("seed", C.c_uint64),
        ("side", C.c_float), ("radius", C.c_float), ("delay_min", C.c_uint32),
        ("delay_max", C.c_uint32), ("w_min", C.c_float), ("w_max", C.c_float),
        ("inhib_frac", C.c_float), ("n_threads", C.c_uint32),
    ]


class Params(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("learning_rate", C.c_float), ("pre_factor", C.c_float),
        ("post_factor", C.c_float), ("pre_trace_decay", C.c_float), ("post_trace_decay", C.c_float),
    ]


class DriveSpec(C.Structure):
    _fields_ = [
        ("n_inputs", C.c_uint32), ("n_groups", C.c_uint32), ("n_correlated", C.c_uint32),
        ("period", C.c_uint32), ("p0", C.c_float), ("amp_corr", C.c_float),
        ("amp_ind", C.c_float), ("bg_prob", C.c_float), ("seed", C.c_uint64),
    ]


class LaunchCfg(C.Structure):
    _fields_ = [("waves_per_block", C.c_uint32), ("xcd_swizzle", C.c_uint32), ("graph_steps", C.c_uint32),
                ("stage_tables", C.c_uint32), ("fanout_waves", C.c_uint32), ("persistent", C.c_uint32)]


class ServerSpec(C.Structure):
    _fields_ = [
        ("device", C.c_int32), ("lanes", C.c_uint32), ("chunk", C.c_uint32), ("builders", C.c_uint32),
        ("device_wiring", C.c_uint32), ("max_nets", C.c_uint32), ("max_neurons", C.c_uint64),
        ("max_synapses", C.c_uint64), ("hbm_fraction", C.c_float), ("syn_per_neuron", C.c_float),
        ("evict_after_s", C.c_float), ("overflow_lanes", C.c_uint32),
    ]


class JobSpec(C.Structure):
    _fields_ = [("id", C.c_uint64), ("wiring", WiringSpec), ("drive", DriveSpec), ("params", Params),
                ("steps", C.c_uint32)]


class JobResult(C.Structure):
    _fields_ = [
        ("id", C.c_uint64), ("status", C.c_int32), ("evictions", C.c_uint32), ("n_neurons", C.c_uint32),
        ("steps", C.c_uint32), ("n_synapses", C.c_uint64), ("spikes", C.c_uint64), ("digest", C.c_uint64),
        ("t_submit", C.c_double), ("t_built", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("built", C.c_uint64), ("admitted", C.c_uint64), ("finished", C.c_uint64),
        ("evictions", C.c_uint64), ("chunks", C.c_uint64), ("running", C.c_uint32), ("waiting", C.c_uint32),
        ("lanes", C.c_uint32), ("gpu_wait_s", C.c_double), ("admit_s", C.c_double), ("retire_s", C.c_double),
        ("idle_s", C.c_double), ("mean_running", C.c_double),
    ]


class SimSpec(C.Structure):
    _fields_ = [
        ("device", C.c_int32), ("max_neurons", C.c_uint64), ("max_synapses", C.c_uint64),
        ("max_nets", C.c_uint32), ("raster_capacity", C.c_uint32), ("max_slots", C.c_uint64),
    ]


P = C.c_void_p
u8p, u32p, i32p, i64p, u64p, f32p = (C.POINTER(C.c_uint8), C.POINTER(C.c_uint32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int64), C.POINTER(C.c_uint64), C.POINTER(C.c_float))

# name -> (restype, argtypes); every symbol include/neucor_amd.h declares.
SIGNATURES = {
    "nc_net_generate": (C.c_int, [C.POINTER(WiringSpec), C.POINTER(P)]),
    "nc_net_from_edges": (C.c_int, [C.c_uint32, f32p, C.c_uint64, u32p, u32p, f32p, u8p, u8p, C.POINTER(P)]),
    "nc_net_info": (C.c_int, [P, u32p, u64p, u32p, u32p]),
    "nc_net_export": (C.c_int, [P, f32p, i64p, u32p, f32p, u8p, u8p]),
    "nc_net_free": (None, [P]),
    "nc_net_select_inputs": (C.c_int, [P, C.POINTER(DriveSpec), u32p, u32p]),
    "nc_sim_create": (C.c_int, [C.POINTER(SimSpec), C.POINTER(P)]),
    "nc_net_device_size": (C.c_int, [P, u64p, u64p]),
    "nc_net_prepare": (C.c_int, [P]),
    "nc_net_generate_device": (C.c_int, [C.POINTER(WiringSpec), C.c_int32, C.POINTER(P)]),
    "nc_sim_destroy": (None, [P]),
    "nc_sim_footprint": (C.c_int, [C.c_uint64, C.c_uint64, u64p]),
    "nc_sim_free_bytes": (C.c_int, [P, u64p]),
    "nc_sim_add": (C.c_int, [P, P, C.POINTER(Params), C.POINTER(DriveSpec), C.c_uint32, i32p]),
    "nc_sim_remove": (C.c_int, [P, C.c_int32]),
    "nc_sim_step": (C.c_int, [P, C.c_uint32]),
    "nc_sim_progress": (C.c_int, [P, C.c_int32, u32p, u32p]),
    "nc_get_time": (C.c_int, [P, C.c_int32, f32p]),
    "nc_set_params": (C.c_int, [P, C.c_int32, C.c_float, C.c_float, C.c_float]),
    "nc_push_forced": (C.c_int, [P, C.c_int32, C.c_uint32, u32p, u32p, u8p]),
    "nc_read_potential": (C.c_int, [P, C.c_int32, f32p, C.c_uint32]),
    "nc_read_last_fire": (C.c_int, [P, C.c_int32, i32p, C.c_uint32]),
    "nc_read_firings": (C.c_int, [P, C.c_int32, u32p, C.c_uint32]),
    "nc_read_weights": (C.c_int, [P, C.c_int32, f32p, C.c_uint64]),
    "nc_read_weights_noalloc": (C.c_int, [P, C.c_int32, f32p, C.c_uint64]),
    "nc_write_weights": (C.c_int, [P, C.c_int32, f32p, C.c_uint64]),
    "nc_set_weights": (C.c_int, [P, C.c_int32, C.c_uint64, u64p, f32p]),
    "nc_read_synapse_state": (C.c_int, [P, C.c_int32, i32p, i32p, f32p, C.c_uint64]),
    "nc_read_raster": (C.c_int, [P, C.c_int32, u32p, u32p, C.c_uint64, u64p]),
    "nc_reset_activities": (C.c_int, [P, C.c_int32]),
    "nc_read_counters": (C.c_int, [P, u64p, C.c_uint32]),
    "nc_set_counting": (C.c_int, [P, C.c_int]),
    "nc_sim_save": (C.c_int, [P, C.c_int32, C.c_void_p, C.c_uint64, u64p]),
    "nc_sim_restore": (C.c_int, [P, C.c_int32, C.c_void_p, C.c_uint64]),
    "nc_sim_set_launch": (C.c_int, [P, C.POINTER(LaunchCfg)]),
    "nc_sim_get_launch": (C.c_int, [P, C.POINTER(LaunchCfg)]),
    "nc_sim_tune": (C.c_int, [P, C.c_uint32, C.POINTER(LaunchCfg)]),
    "nc_sim_tune_list": (C.c_int, [P, C.POINTER(LaunchCfg), C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.POINTER(LaunchCfg), C.POINTER(C.c_float)]),
    "nc_sim_synchronize": (C.c_int, [P]),
    "nc_set_kernel_timing": (C.c_int, [P, C.c_int, C.c_uint32]),
    "nc_read_kernel_timing": (C.c_int, [P, C.c_uint32, u64p, C.POINTER(C.c_double)]),
    "nc_last_error": (C.c_char_p, []),
    "nc_server_create": (C.c_int, [C.POINTER(ServerSpec), C.POINTER(P)]),
    "nc_server_submit": (C.c_int, [P, C.POINTER(JobSpec)]),
    "nc_server_poll": (C.c_int, [P, C.POINTER(JobResult), C.c_uint32, u32p, C.c_double]),
    "nc_server_stats_get": (C.c_int, [P, C.POINTER(ServerStats)]),
    "nc_server_destroy": (C.c_int, [P]),
    "nc_digest": (C.c_uint64, [C.c_void_p, C.c_uint64, C.c_uint64]),
}

FNV_OFFSET = 14695981039346656037

_lib = None


def lib():
    """The loaded engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libneucor_amd.so is not built; run neurocorrelation_amd/build.py "
                              "(or __graft_entry__.build()) first: %s" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != NC_OK:
        msg = lib().nc_last_error()
        raise NCError(code, msg.decode() if msg else "")


def ptr(arr, ctype):
    """ctypes pointer into a contiguous numpy array (None passes NULL)."""
    if arr is None:
        return None
    assert arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(C.POINTER(ctype))
