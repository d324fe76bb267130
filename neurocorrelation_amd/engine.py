"""Python view of the engine: networks (host) and simulations (device).

This is a thin convenience layer over the C ABI for tests, the benchmark and
the serving loop; the product is the C ABI and the HIP kernels beneath it.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import LaunchCfg, Params, SimSpec, WiringSpec, check, lib, ptr

NEVER = _lib.NC_NEVER_STEP


def _cfg_dict(cfg):
    return {f: getattr(cfg, f) for f, _ in LaunchCfg._fields_}


def default_params(dt=1.0, learning_rate=1.0, pre_factor=0.13, post_factor=0.30,
                   pre_trace_decay=0.75, post_trace_decay=0.65):
    """NeuCor's defaults (src/NeuCor.cpp:18-29) with runSpeed = dt."""
    return Params(dt, learning_rate, pre_factor, post_factor, pre_trace_decay, post_trace_decay)


class Network:
    """A network on the host (nc_net)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def generate(cls, n_neurons, side, radius, delay_min=1, delay_max=20, w_min=0.2, w_max=0.8,
                 inhib_frac=0.0, seed=0, spatial_order=1, n_threads=0, device=None):
        """The workload wiring rule on host threads, or (device = a GPU
        index) on that GPU; both build the same network bit for bit."""
        spec = WiringSpec(n_neurons, spatial_order, seed, side, radius, delay_min, delay_max,
                          w_min, w_max, inhib_frac, n_threads)
        h = C.c_void_p()
        if device is None:
            check(lib().nc_net_generate(C.byref(spec), C.byref(h)))
        else:
            check(lib().nc_net_generate_device(C.byref(spec), int(device), C.byref(h)))
        return cls(h)

    @classmethod
    def from_edges(cls, n_neurons, to, frm, w, delay_steps, inhibitory=None, xyz=None):
        to = np.ascontiguousarray(to, dtype=np.uint32)
        frm = np.ascontiguousarray(frm, dtype=np.uint32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        d = np.ascontiguousarray(delay_steps, dtype=np.uint8)
        ih = None if inhibitory is None else np.ascontiguousarray(inhibitory, dtype=np.uint8)
        xyz = None if xyz is None else np.ascontiguousarray(xyz, dtype=np.float32)
        h = C.c_void_p()
        check(lib().nc_net_from_edges(n_neurons, ptr(xyz, C.c_float), len(to), ptr(to, C.c_uint32),
                                      ptr(frm, C.c_uint32), ptr(w, C.c_float), ptr(d, C.c_uint8),
                                      ptr(ih, C.c_uint8), C.byref(h)))
        return cls(h)

    def info(self):
        n, s, md, mdl = C.c_uint32(), C.c_uint64(), C.c_uint32(), C.c_uint32()
        check(lib().nc_net_info(self._h, C.byref(n), C.byref(s), C.byref(md), C.byref(mdl)))
        return {"n": n.value, "n_syn": s.value, "max_in_degree": md.value, "max_delay": mdl.value}

    @property
    def n(self):
        return self.info()["n"]

    @property
    def n_syn(self):
        return self.info()["n_syn"]

    def export(self):
        inf = self.info()
        n, S = inf["n"], inf["n_syn"]
        xyz = np.empty(3 * n, np.float32)
        in_ptr = np.empty(n + 1, np.int64)
        src = np.empty(S, np.uint32)
        w = np.empty(S, np.float32)
        d = np.empty(S, np.uint8)
        ih = np.empty(S, np.uint8)
        check(lib().nc_net_export(self._h, ptr(xyz, C.c_float), ptr(in_ptr, C.c_int64), ptr(src, C.c_uint32),
                                  ptr(w, C.c_float), ptr(d, C.c_uint8), ptr(ih, C.c_uint8)))
        return {"xyz": xyz.reshape(n, 3), "in_ptr": in_ptr, "in_src": src, "w": w, "delay": d, "inh": ih}

    def select_inputs(self, drive):
        ids = np.empty(drive.n_inputs, np.uint32)
        groups = np.empty(drive.n_inputs, np.uint32)
        check(lib().nc_net_select_inputs(self._h, C.byref(drive), ptr(ids, C.c_uint32), ptr(groups, C.c_uint32)))
        return ids, groups

    def device_size(self):
        p, s = C.c_uint64(), C.c_uint64()
        check(lib().nc_net_device_size(self._h, C.byref(p), C.byref(s)))
        return p.value, s.value

    def prepare(self):
        """Builds and keeps the device image (layout, slot records,
        out-edges), so that each Simulation.add of this network only
        relocates and uploads it; safe to call from a worker thread."""
        check(lib().nc_net_prepare(self._h))
        return self

    def close(self):
        if self._h:
            lib().nc_net_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Simulation:
    """A device-resident engine (nc_sim) holding one or more networks."""

    def __init__(self, max_neurons, max_synapses, max_nets=1, raster_capacity=0, device=0, max_slots=0):
        spec = SimSpec(device, max_neurons, max_synapses, max_nets, raster_capacity, max_slots)
        h = C.c_void_p()
        check(lib().nc_sim_create(C.byref(spec), C.byref(h)))
        self._h = h

    @classmethod
    def for_network(cls, net, raster_capacity=0, device=0):
        npos, nslots = net.device_size()
        return cls(npos, max(net.n_syn, 1), 1, raster_capacity, device, max_slots=nslots)

    def add(self, net, params=None, drive=None, steps=0):
        params = params or default_params()
        nid = C.c_int32()
        check(lib().nc_sim_add(self._h, net._h, C.byref(params), C.byref(drive) if drive is not None else None,
                               steps, C.byref(nid)))
        return nid.value

    def remove(self, net_id):
        check(lib().nc_sim_remove(self._h, net_id))

    def step(self, n_steps):
        check(lib().nc_sim_step(self._h, n_steps))

    def progress(self, net_id=0):
        d, l = C.c_uint32(), C.c_uint32()
        check(lib().nc_sim_progress(self._h, net_id, C.byref(d), C.byref(l)))
        return d.value, l.value

    def time(self, net_id=0):
        t = C.c_float()
        check(lib().nc_get_time(self._h, net_id, C.byref(t)))
        return t.value

    def set_params(self, net_id, learning_rate, pre_factor, post_factor):
        check(lib().nc_set_params(self._h, net_id, learning_rate, pre_factor, post_factor))

    def push_forced(self, net_id, step_off, ids, kind):
        step_off = np.ascontiguousarray(step_off, np.uint32)
        ids = np.ascontiguousarray(ids, np.uint32)
        kind = np.ascontiguousarray(kind, np.uint8)
        check(lib().nc_push_forced(self._h, net_id, len(ids), ptr(step_off, C.c_uint32), ptr(ids, C.c_uint32),
                                   ptr(kind, C.c_uint8)))

    def potential(self, net_id, n):
        out = np.empty(n, np.float32)
        check(lib().nc_read_potential(self._h, net_id, ptr(out, C.c_float), n))
        return out

    def last_fire(self, net_id, n):
        out = np.empty(n, np.int32)
        check(lib().nc_read_last_fire(self._h, net_id, ptr(out, C.c_int32), n))
        return out

    def firings(self, net_id, n):
        out = np.empty(n, np.uint32)
        check(lib().nc_read_firings(self._h, net_id, ptr(out, C.c_uint32), n))
        return out

    def weights(self, net_id, S):
        out = np.empty(S, np.float32)
        check(lib().nc_read_weights(self._h, net_id, ptr(out, C.c_float), S))
        return out

    def weights_noalloc(self, net_id, S):
        """weights() through the entry point that allocates no device memory
        (static per-GPU scratch; networks of at most 16,777,216 synapses)."""
        out = np.empty(S, np.float32)
        check(lib().nc_read_weights_noalloc(self._h, net_id, ptr(out, C.c_float), S))
        return out

    def write_weights(self, net_id, w):
        w = np.ascontiguousarray(w, np.float32)
        check(lib().nc_write_weights(self._h, net_id, ptr(w, C.c_float), len(w)))

    def set_weights(self, net_id, syn, w):
        """Synapse::setWeight on the listed synapses (export order)."""
        syn = np.ascontiguousarray(syn, np.uint64)
        w = np.ascontiguousarray(w, np.float32)
        check(lib().nc_set_weights(self._h, net_id, len(syn), ptr(syn, C.c_uint64), ptr(w, C.c_float)))

    def synapse_state(self, net_id, S):
        arr = np.empty(S, np.int32)
        la = np.empty(S, np.int32)
        dp = np.empty(S, np.float32)
        check(lib().nc_read_synapse_state(self._h, net_id, ptr(arr, C.c_int32), ptr(la, C.c_int32),
                                          ptr(dp, C.c_float), S))
        return arr, la, dp

    def raster(self, net_id, cap):
        steps = np.empty(cap, np.uint32)
        ids = np.empty(cap, np.uint32)
        n = C.c_uint64()
        check(lib().nc_read_raster(self._h, net_id, ptr(steps, C.c_uint32), ptr(ids, C.c_uint32), cap,
                                   C.byref(n)))
        return steps[:n.value].copy(), ids[:n.value].copy()

    def reset_activities(self, net_id):
        check(lib().nc_reset_activities(self._h, net_id))

    def counters(self):
        out = np.zeros(_lib.NC_NUM_COUNTERS, np.uint64)
        check(lib().nc_read_counters(self._h, ptr(out, C.c_uint64), len(out)))
        keys = ["neuron_steps", "synapse_steps", "active_slot_steps", "arrivals", "fired",
                "fired_in_degree", "fired_out_degree", "accepted_spikes", "pending_scanned",
                "mask_words_stepped", "mask_words_written", "fired_records", "mask_words_read",
                "bucket_entries", "records_read", "arrival_weight_writes", "cache_records_written",
                "arrival_records_read", "post_weight_writes"]
        return dict(zip(keys, (int(x) for x in out)))

    def set_launch(self, waves_per_block=4, xcd_swizzle=0, graph_steps=0, stage_tables=0, fanout_waves=0,
                   persistent=0):
        cfg = LaunchCfg(waves_per_block, xcd_swizzle, graph_steps, stage_tables, fanout_waves, persistent)
        check(lib().nc_sim_set_launch(self._h, C.byref(cfg)))

    def launch(self):
        cfg = LaunchCfg()
        check(lib().nc_sim_get_launch(self._h, C.byref(cfg)))
        return _cfg_dict(cfg)

    def tune(self, steps_per_trial=32):
        """Autotunes the launch shape over the engine's 38 candidate shapes,
        every trial from the same saved state; the simulation is left where
        it was.  Returns the chosen shape."""
        cfg = LaunchCfg()
        check(lib().nc_sim_tune(self._h, steps_per_trial, C.byref(cfg)))
        return _cfg_dict(cfg)

    def tune_list(self, candidates, steps_per_trial=16, rounds=3):
        """Autotunes over the given shapes (dicts of LaunchCfg fields); returns
        (chosen shape, median ms per candidate)."""
        arr = (LaunchCfg * len(candidates))()
        for i, c in enumerate(candidates):
            arr[i] = LaunchCfg(c.get("waves_per_block", 4), c.get("xcd_swizzle", 0), c.get("graph_steps", 0),
                               c.get("stage_tables", 0), c.get("fanout_waves", 0), c.get("persistent", 0))
        ms = np.zeros(len(candidates), np.float32)
        cfg = LaunchCfg()
        check(lib().nc_sim_tune_list(self._h, arr, len(candidates), steps_per_trial, rounds, C.byref(cfg),
                                     ptr(ms, C.c_float)))
        return _cfg_dict(cfg), ms

    def save(self, net_id):
        """The network's dynamic state as bytes (a checkpoint)."""
        size = C.c_uint64()
        check(lib().nc_sim_save(self._h, net_id, None, 0, C.byref(size)))
        buf = C.create_string_buffer(size.value)
        check(lib().nc_sim_save(self._h, net_id, buf, size.value, C.byref(size)))
        return buf.raw[:size.value]

    def restore(self, net_id, data):
        """Loads a checkpoint into a network just added from the same wiring."""
        buf = C.create_string_buffer(bytes(data), len(data))
        check(lib().nc_sim_restore(self._h, net_id, buf, len(data)))

    def set_counting(self, enable):
        check(lib().nc_set_counting(self._h, 1 if enable else 0))

    def synchronize(self):
        check(lib().nc_sim_synchronize(self._h))

    def set_kernel_timing(self, kernel_mask, every=1):
        check(lib().nc_set_kernel_timing(self._h, int(kernel_mask), int(every)))

    def kernel_timing(self, kernel):
        n, ms = C.c_uint64(), C.c_double()
        check(lib().nc_read_kernel_timing(self._h, kernel, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def free_bytes(self):
        f = C.c_uint64()
        check(lib().nc_sim_free_bytes(self._h, C.byref(f)))
        return f.value

    def close(self):
        if getattr(self, "_h", None):
            lib().nc_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def footprint(n_neurons, n_syn):
    b = C.c_uint64()
    check(lib().nc_sim_footprint(n_neurons, n_syn, C.byref(b)))
    return b.value
