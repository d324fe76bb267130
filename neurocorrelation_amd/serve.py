"""The served form: simulation jobs over the engine's native request loop
(nc_server_*, include/neucor_amd.h).

Jobs (a wiring spec, a drive, a step count) arrive over time.  The native
server builds each one's network on builder threads (on the GPU with
device_wiring), admits it into a running simulation ("lane") as soon as the
lane has room, steps it together with every other admitted job (continuous
batching: one step kernel pair covers all of a lane's jobs; each job stops at
its own step count), and retires it at its step count, which frees its room.
A job's result is exactly what it gives when run alone: the step of one
network does not depend on the others.  Under overload, a job that has
waited `evict_after` seconds evicts the running job with the most steps left
(checkpoint, remove, resume later from the checkpoint, same result), and
overflow lanes are created while the device's free HBM, read at that moment,
holds one.

This module holds the job specs of the served workload, the JSON wire format
of jobs and results, and a thin wrapper over the native server.
"""
import ctypes as C
import json
import math
import time

import numpy as np

from . import workloads
from ._lib import (FNV_OFFSET, STATUS, JobResult, JobSpec, Params, ServerSpec, ServerStats, WiringSpec, check,
                   lib)
from .engine import Simulation, default_params


def served_specs(count, seed=0, n_range=(300, 100_000), k_range=(20, 100), steps_range=(1_000, 50_000),
                 weights=(0.05, 0.3, 0.2)):
    """Jobs of the served form: neuron count log-uniform in n_range, K
    (synapses per neuron in the bulk) uniform in k_range, step count uniform
    in steps_range; wiring (density 300 per unit volume, delays 1..32 steps)
    and drive (5% inputs in 4 correlated groups plus independent ones, p =
    0.05 (1 + sin(2 pi t / 200))) of the 20,000-neuron workload, seeded per
    job.  weights = (w_min, w_max, inhibitory fraction): the default is the
    10^6-neuron workload's rule, under which the jobs run at p ~ 0.0025 per
    step; (0.2, 0.8, 0.0) makes them run away to p ~ 0.27."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        k = float(rng.uniform(*k_range))
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        net_seed = int(rng.integers(0, 2**63 - 1))
        n_in = max(5, int(round(0.05 * n)))
        n_corr = 4 * ((4 * n_in // 5) // 4)
        out.append(dict(
            wiring=dict(n_neurons=n, side=workloads.side_for(n), radius=workloads.radius_for_degree(k),
                        delay_min=1, delay_max=32, w_min=weights[0], w_max=weights[1], inhib_frac=weights[2],
                        seed=net_seed),
            drive=dict(n_inputs=n_in, n_groups=4, n_correlated=n_corr, period=200, p0=0.05, amp_corr=1.0,
                       amp_ind=1.0, seed=net_seed),
            steps=steps, k=k))
    return out


# ---------------------------------------------------------------------------
# Wire format: one JSON object per job and per result (JSON lines).
#   job    {"id": 7, "steps": 5000,
#           "wiring": {"n_neurons", "side", "radius", "delay_min", "delay_max",
#                      "w_min", "w_max", "inhib_frac", "seed"},
#           "drive": {"n_inputs", "n_groups", "n_correlated", "period", "p0",
#                     "amp_corr", "amp_ind", "seed"[, "bg_prob"]},
#           "params": {"dt", "learning_rate", "pre_factor", "post_factor",
#                      "pre_trace_decay", "post_trace_decay"}}   (optional)
#   result {"id", "status" ("ok" or the error name), "spikes", "digest"
#           (16 hex digits), "evictions", "n_neurons", "n_synapses", "steps",
#           "t_submit", "t_built", "t_admit", "t_done"}
# ---------------------------------------------------------------------------
WIRING_KEYS = ("n_neurons", "side", "radius", "delay_min", "delay_max", "w_min", "w_max", "inhib_frac", "seed")
DRIVE_KEYS = ("n_inputs", "n_groups", "n_correlated", "period", "p0", "amp_corr", "amp_ind", "seed")
PARAM_KEYS = ("dt", "learning_rate", "pre_factor", "post_factor", "pre_trace_decay", "post_trace_decay")


def job_to_json(job_id, spec):
    out = {"id": int(job_id), "steps": int(spec["steps"]),
           "wiring": {k: spec["wiring"][k] for k in WIRING_KEYS},
           "drive": {k: spec["drive"][k] for k in DRIVE_KEYS + ("bg_prob",) if k in spec["drive"]}}
    if "params" in spec:
        out["params"] = {k: spec["params"][k] for k in PARAM_KEYS}
    return json.dumps(out, sort_keys=True)


def job_from_json(text):
    """(job id, spec) of a job line; raises ValueError on a malformed job."""
    d = json.loads(text)
    try:
        spec = {"steps": int(d["steps"]), "wiring": {k: d["wiring"][k] for k in WIRING_KEYS},
                "drive": {k: d["drive"][k] for k in DRIVE_KEYS}}
        if "bg_prob" in d["drive"]:
            spec["drive"]["bg_prob"] = float(d["drive"]["bg_prob"])
        if "params" in d:
            spec["params"] = {k: float(d["params"][k]) for k in PARAM_KEYS}
        job_id = int(d["id"])
    except (KeyError, TypeError) as e:
        raise ValueError("malformed job: %s" % e) from None
    if spec["steps"] <= 0 or job_id < 0:
        raise ValueError("malformed job: steps and id must be positive")
    return job_id, spec


def result_to_json(res):
    return json.dumps(res, sort_keys=True)


def job_struct(job_id, spec):
    """The nc_job of a spec (params default to the reference's)."""
    w = spec["wiring"]
    ws = WiringSpec(int(w["n_neurons"]), 1, int(w["seed"]), float(w["side"]), float(w["radius"]),
                    int(w["delay_min"]), int(w["delay_max"]), float(w["w_min"]), float(w["w_max"]),
                    float(w.get("inhib_frac", 0.0)), 1)
    ds = workloads.drive_spec(spec["drive"], seed=int(spec["drive"]["seed"]))
    p = spec.get("params")
    ps = default_params() if p is None else Params(*(float(p[k]) for k in PARAM_KEYS))
    return JobSpec(int(job_id), ws, ds, ps, int(spec["steps"]))


def digest_arrays(V, w, f):
    """The served result digest (nc_digest) of final potentials, weights and
    firing counts."""
    h = FNV_OFFSET
    for a in (np.ascontiguousarray(V, np.float32), np.ascontiguousarray(w, np.float32),
              np.ascontiguousarray(f, np.uint32)):
        h = lib().nc_digest(a.ctypes.data, a.nbytes, h)
    return h


def summarize(V, w, f):
    return {"spikes": int(np.asarray(f, np.uint64).sum()), "digest": "%016x" % digest_arrays(V, w, f)}


def _result_dict(r):
    out = {"id": int(r.id), "status": "ok" if r.status == 0 else STATUS.get(r.status, str(r.status)),
           "evictions": int(r.evictions), "n_neurons": int(r.n_neurons), "n_synapses": int(r.n_synapses),
           "steps": int(r.steps), "t_submit": r.t_submit, "t_built": r.t_built, "t_admit": r.t_admit,
           "t_done": r.t_done}
    if r.status == 0:
        out["spikes"] = int(r.spikes)
        out["digest"] = "%016x" % r.digest
    return out


class Job:
    """A job of a served run: its index (the job id), spec and arrival time
    (seconds from the start of the run); after the run, its build, admit and
    completion times (same clock) and result."""
    __slots__ = ("idx", "spec", "arrival", "t_built", "t_admit", "t_done", "result", "n", "n_syn", "steps",
                 "evictions")

    def __init__(self, idx, spec, arrival):
        self.idx, self.spec, self.arrival = idx, spec, arrival
        self.t_built = self.t_admit = self.t_done = self.result = None
        self.n = self.n_syn = self.evictions = 0
        self.steps = int(spec["steps"])


class Server:
    """The native server (nc_server_create).  max_neurons / max_synapses:
    each lane's room (0: the lanes share `hbm_fraction` of the free HBM);
    evict_after: seconds a job waits for room before it evicts (0: never);
    overflow_lanes: extra lanes made from free HBM while jobs wait."""

    def __init__(self, max_neurons=0, max_synapses=0, max_nets=2048, chunk=200, builders=4, device=0, lanes=2,
                 device_wiring=False, evict_after=0.0, overflow_lanes=0, hbm_fraction=0.5, syn_per_neuron=60.0):
        spec = ServerSpec(int(device), int(lanes), int(chunk), int(builders), int(bool(device_wiring)),
                          int(max_nets), int(max_neurons), int(max_synapses), float(hbm_fraction),
                          float(syn_per_neuron), float(evict_after), int(overflow_lanes))
        h = C.c_void_p()
        check(lib().nc_server_create(C.byref(spec), C.byref(h)))
        self._h = h
        self._created = time.perf_counter()
        self.spec = spec

    def submit(self, job_id, spec):
        job = job_struct(job_id, spec)
        check(lib().nc_server_submit(self._h, C.byref(job)))

    def poll(self, wait=0.0, cap=256):
        """Finished jobs' results (dicts of the wire format), waiting up to
        `wait` seconds for the first."""
        buf = (JobResult * cap)()
        n = C.c_uint32(0)
        check(lib().nc_server_poll(self._h, buf, cap, C.byref(n), float(wait)))
        return [_result_dict(buf[i]) for i in range(n.value)]

    def stats(self):
        s = ServerStats()
        check(lib().nc_server_stats_get(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def run(self, jobs, t_limit=None, log=None):
        """Serves `jobs` (Job objects with arrival times in seconds from now):
        submits each at its arrival time, collects results.  Returns
        (completed jobs, summary) where summary holds the unfinished count and
        the native loop's statistics."""
        t0 = time.perf_counter()
        off = t0 - self._created  # server clock at the start of the run
        pending = sorted(jobs, key=lambda j: j.arrival)
        by_id = {j.idx: j for j in jobs}
        done = []
        nxt = 0
        last_log = t0
        while len(done) < len(jobs):
            now = time.perf_counter() - t0
            if t_limit is not None and now > t_limit:
                break
            while nxt < len(pending) and pending[nxt].arrival <= now:
                self.submit(pending[nxt].idx, pending[nxt].spec)
                nxt += 1
            wait = 0.005
            if nxt < len(pending):
                wait = max(0.0, min(wait, pending[nxt].arrival - (time.perf_counter() - t0)))
            for r in self.poll(wait=wait):
                job = by_id[r["id"]]
                job.t_built, job.t_admit, job.t_done = r["t_built"] - off, r["t_admit"] - off, r["t_done"] - off
                job.n, job.n_syn, job.evictions = r["n_neurons"], r["n_synapses"], r["evictions"]
                job.result = ({"spikes": r["spikes"], "digest": r["digest"]} if r["status"] == "ok"
                              else {"error": r["status"]})
                done.append(job)
            if log and time.perf_counter() - last_log > 10.0:
                last_log = time.perf_counter()
                st = self.stats()
                log("served %d / %d jobs, %d running, %d waiting, %d evictions, %d lanes, %.1fs" %
                    (len(done), len(jobs), st["running"], st["waiting"], st["evictions"], st["lanes"],
                     last_log - t0))
        return done, {"unfinished": len(jobs) - len(done), "loop": self.stats()}

    def close(self):
        if self._h:
            h, self._h = self._h, None
            check(lib().nc_server_destroy(h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_alone(spec):
    """One job on the plain path: its own simulation, default launch.
    Returns (result, seconds including construction)."""
    t = time.perf_counter()
    net, drive, params, steps = workloads.build_spec(spec, n_threads=0)
    if "params" in spec:
        params = Params(*(float(spec["params"][k]) for k in PARAM_KEYS))
    sim = Simulation.for_network(net)
    nid = sim.add(net, params, drive)
    left = steps
    while left:
        k = min(500, left)
        sim.step(k)
        left -= k
    sim.synchronize()
    res = summarize(sim.potential(nid, net.n), sim.weights(nid, net.n_syn), sim.firings(nid, net.n))
    el = time.perf_counter() - t
    sim.close()
    net.close()
    return res, el
