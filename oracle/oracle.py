"""ctypes view of the C oracle (oracle/nc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class Wiring(C.Structure):
    _fields_ = [("n_neurons", C.c_uint32), ("spatial_order", C.c_uint32), ("seed", C.c_uint64),
                ("side", C.c_float), ("radius", C.c_float), ("delay_min", C.c_uint32),
                ("delay_max", C.c_uint32), ("w_min", C.c_float), ("w_max", C.c_float),
                ("inhib_frac", C.c_float)]


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.c_void_p
        L.or_philox4x32_10.argtypes = [P, P, P]
        L.or_positions.argtypes = [C.POINTER(Wiring), P]
        L.or_rows.restype = C.c_int64
        L.or_rows.argtypes = [C.POINTER(Wiring), P, C.c_uint32, P, P, P, P, P, P]
        L.or_select_inputs.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, P, P]
        L.or_sim_create.restype = P
        L.or_sim_create.argtypes = [C.c_uint32, C.c_uint64, P, P, P, P, P] + [C.c_float] * 6
        L.or_sim_free.argtypes = [P]
        L.or_sim_set_params.argtypes = [P, C.c_float, C.c_float, C.c_float]
        L.or_sim_set_weights.argtypes = [P, P, C.c_uint64]
        L.or_sim_set_drive.argtypes = [P, C.c_uint32, P, P, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                       C.c_float, C.c_float, C.c_uint64]
        L.or_sim_push_forced.argtypes = [P, C.c_uint32, P, P, P]
        L.or_sim_step.argtypes = [P, C.c_uint32, P, P, C.c_uint64, P]
        L.or_sim_read.argtypes = [P] * 9
        L.or_sim_counters.argtypes = [P, P]
        L.or_sim_reset_activities.argtypes = [P]
        L.or_trace.restype = C.c_float
        L.or_trace.argtypes = [C.c_float, C.c_float]
        L.or_ap_value.restype = C.c_float
        L.or_ap_value.argtypes = [C.c_float]
        L.or_bg_events.restype = C.c_uint64
        L.or_bg_events.argtypes = [C.c_uint64, C.c_float, C.c_uint32, C.c_uint32, C.c_uint32, P, P, C.c_uint64]
        _L = L
    return _L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def wiring(n_neurons, side, radius, delay_min=1, delay_max=20, w_min=0.2, w_max=0.8, inhib_frac=0.0, seed=0,
           spatial_order=1):
    return Wiring(n_neurons, spatial_order, seed, side, radius, delay_min, delay_max, w_min, w_max, inhib_frac)


def positions(spec):
    xyz = np.empty(3 * spec.n_neurons, np.float32)
    assert lib().or_positions(C.byref(spec), _p(xyz)) == 0
    return xyz.reshape(-1, 3)


def rows(spec, xyz, rows_):
    """In-rows of the given targets by brute force over all sources."""
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
    r = np.ascontiguousarray(rows_, np.uint32)
    S = lib().or_rows(C.byref(spec), _p(xyz), len(r), _p(r), None, None, None, None, None)
    ptr = np.empty(len(r) + 1, np.int64)
    src = np.empty(S, np.uint32)
    w = np.empty(S, np.float32)
    d = np.empty(S, np.uint8)
    ih = np.empty(S, np.uint8)
    lib().or_rows(C.byref(spec), _p(xyz), len(r), _p(r), _p(ptr), _p(src), _p(w), _p(d), _p(ih))
    return {"in_ptr": ptr, "in_src": src, "w": w, "delay": d, "inh": ih}


def network(spec):
    """Whole network by brute force (O(N^2): small N only)."""
    xyz = positions(spec)
    net = rows(spec, xyz, np.arange(spec.n_neurons, dtype=np.uint32))
    net["xyz"] = xyz
    return net


def select_inputs(n, seed, n_inputs, n_correlated, n_groups):
    ids = np.empty(n_inputs, np.uint32)
    groups = np.empty(n_inputs, np.uint32)
    assert lib().or_select_inputs(n, seed, n_inputs, n_correlated, n_groups, _p(ids), _p(groups)) == 0
    return ids, groups


class Sim:
    """The oracle's simulation of one network."""

    def __init__(self, net, dt=1.0, lr=1.0, pre_f=0.13, post_f=0.30, pre_decay=0.75, post_decay=0.65):
        self.n = len(net["in_ptr"]) - 1
        self.S = len(net["in_src"])
        self._keep = [np.ascontiguousarray(net[k]) for k in ("in_ptr", "in_src", "w", "delay")]
        ih = net.get("inh")
        self._keep.append(None if ih is None else np.ascontiguousarray(ih, np.uint8))
        a = self._keep
        self._h = lib().or_sim_create(self.n, self.S, _p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), _p(a[4]),
                                      dt, lr, pre_f, post_f, pre_decay, post_decay)
        assert self._h

    def set_drive(self, ids, groups, n_corr_groups, period, p0, amp_corr, amp_ind, bg_prob=0.0, seed=0):
        ids = np.ascontiguousarray(ids, np.uint32)
        groups = np.ascontiguousarray(groups, np.uint32)
        assert lib().or_sim_set_drive(self._h, len(ids), _p(ids), _p(groups), n_corr_groups, period, p0,
                                      amp_corr, amp_ind, bg_prob, seed) == 0

    def set_params(self, lr, pre_f, post_f):
        lib().or_sim_set_params(self._h, lr, pre_f, post_f)

    def push_forced(self, step_off, ids, kind):
        so = np.ascontiguousarray(step_off, np.uint32)
        ids = np.ascontiguousarray(ids, np.uint32)
        kind = np.ascontiguousarray(kind, np.uint8)
        assert lib().or_sim_push_forced(self._h, len(ids), _p(so), _p(ids), _p(kind)) == 0

    def step(self, n_steps, cap=None):
        """Run n_steps; returns the raster as (steps, ids), ordered by step then id."""
        cap = cap if cap is not None else max(1024, self.n * n_steps // 4)
        ids = np.empty(cap, np.uint32)
        steps = np.empty(cap, np.uint32)
        nf = C.c_uint64()
        lib().or_sim_step(self._h, n_steps, _p(ids), _p(steps), cap, C.byref(nf))
        if nf.value > cap:
            raise RuntimeError("raster buffer too small")
        return steps[:nf.value].copy(), ids[:nf.value].copy()

    def state(self):
        V = np.empty(self.n, np.float32)
        lf = np.empty(self.n, np.float32)
        fi = np.empty(self.n, np.uint32)
        w = np.empty(self.S, np.float32)
        ft = np.empty(self.S, np.float32)
        la = np.empty(self.S, np.float32)
        dp = np.empty(self.S, np.float32)
        t = C.c_float()
        lib().or_sim_read(self._h, _p(V), _p(lf), _p(fi), _p(w), _p(ft), _p(la), _p(dp), C.byref(t))
        return {"V": V, "last_fire": lf, "firings": fi, "w": w, "fire_t": ft, "last_arr": la, "depol": dp,
                "t": t.value}

    def counters(self):
        out = np.zeros(6, np.int64)
        lib().or_sim_counters(self._h, _p(out))
        return dict(zip(["ties_latch", "ties_weight", "arrivals", "fires", "accepted", "first_tie"],
                        (int(x) for x in out)))

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        lib().or_sim_set_weights(self._h, _p(w), C.c_uint64(len(w)))

    def reset_activities(self):
        lib().or_sim_reset_activities(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_sim_free(self._h)
            self._h = None


def drive_schedule(ids, groups, n_corr_groups, period, p0, amp_corr, amp_ind, seed, steps):
    """Per step 1..steps, the sorted ids of input neurons whose group fires:
    one Philox draw per (step, group) (stream 4) against floor(p * 2^32),
    p = p0 * (1 + amp * sin(2 pi (step mod period) / period)), amp_corr for
    groups below n_corr_groups, amp_ind for the others."""
    key = (seed & 0xFFFFFFFF, seed >> 32)
    members = {}
    for i, g in zip(list(ids), list(groups)):
        members.setdefault(int(g), []).append(int(i))
    per = []
    for s in range(1, steps + 1):
        k = s % period
        fired = []
        for g, mem in members.items():
            amp = amp_corr if g < n_corr_groups else amp_ind
            p = p0 * (1.0 + amp * math.sin(2.0 * math.pi * k / period))
            thr = 0 if p <= 0 else (0xFFFFFFFF if p >= 1 else int(math.floor(p * 4294967296.0)))
            if int(philox((s, g, 4, 0), key)[0]) < thr:
                fired.extend(mem)
        per.append(sorted(fired))
    return per


def bg_events(seed, bg_prob, n, s0, steps):
    """Raw background draws of steps s0+1 .. s0+steps (before the refractory
    check): (steps, ids) ordered by step then id."""
    cap = max(1024, int(n * steps * bg_prob * 1.5) + 1024)
    while True:
        st = np.empty(cap, np.uint32)
        ids = np.empty(cap, np.uint32)
        k = lib().or_bg_events(seed, bg_prob, n, s0, steps, _p(st), _p(ids), cap)
        if k <= cap:
            return st[:k].copy(), ids[:k].copy()
        cap = int(k)


def trace(decay, x):
    return lib().or_trace(decay, x)


def ap_value(x):
    return lib().or_ap_value(x)
