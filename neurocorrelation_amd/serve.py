"""The served form: a request loop over one device-resident simulation.

Jobs (a network spec, a drive, a step count) arrive over time.  Each is
built on a host thread as it arrives, admitted into the running simulation
as soon as the simulation's device pools (sized from free HBM) have room for
it, stepped together with every other admitted job (continuous batching: one
step kernel pair covers all of them, each job stops at its own budget), and
retired when its budget is spent, which frees its room for the jobs waiting.
A job's result is exactly what it gives when run alone: the step of one
network does not depend on the others in the simulation.
"""
import collections
import concurrent.futures as cf
import hashlib
import math
import time

import numpy as np

from . import workloads
from .engine import Simulation, footprint
from ._lib import NCError


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


def read_result(sim, nid, n, n_syn):
    """A job's final state: potentials, weights, firing counts."""
    return sim.potential(nid, n), sim.weights(nid, n_syn), sim.firings(nid, n)


def summarize(V, w, f):
    """A job's result: its firing counts and a hash of its final potentials
    and weights (bit patterns)."""
    h = hashlib.sha256()
    h.update(V.tobytes())
    h.update(w.tobytes())
    h.update(f.tobytes())
    return {"spikes": int(f.sum()), "hash": h.hexdigest()}


def digest(sim, nid, n, n_syn):
    return summarize(*read_result(sim, nid, n, n_syn))


class Job:
    __slots__ = ("idx", "spec", "arrival", "built", "net", "drive", "params", "steps", "n", "n_syn", "nid",
                 "t_built", "t_admit", "t_done", "result")

    def __init__(self, idx, spec, arrival):
        self.idx, self.spec, self.arrival = idx, spec, arrival
        self.net = self.nid = self.t_built = self.t_admit = self.t_done = self.result = None


def _finish(job, state, t0):
    job.result = summarize(*state)
    job.t_done = time.perf_counter() - t0
    return job


def _build(job):
    net, drive, params, steps = workloads.build_spec(job.spec, n_threads=1)
    net.prepare()  # the device image, off the serving thread
    job.net, job.drive, job.params, job.steps = net, drive, params, steps
    job.n, job.n_syn = net.n, net.n_syn
    return job


class Server:
    """One simulation holding up to `max_nets` jobs within `max_neurons` /
    `max_synapses` of device room; `chunk` steps between admission and
    retirement checks."""

    def __init__(self, max_neurons, max_synapses, max_nets=4096, chunk=200, builders=8, device=0, max_ready=256):
        slots = max_synapses + max_synapses // 4 + (max_neurons + 64 * max_nets) * 4
        self.sim = Simulation(max_neurons, max_synapses, max_nets=max_nets, device=device, max_slots=slots)
        self.chunk = chunk
        self.max_ready = max_ready  # jobs built or being built but not admitted (bounds host memory)
        self.pool = cf.ThreadPoolExecutor(builders)
        self.finisher = cf.ThreadPoolExecutor(2)  # results, apart from the builders' queue

    @staticmethod
    def capacity_for(free_bytes, fraction=0.5, syn_per_neuron=60):
        """(max_neurons, max_synapses) whose footprint is about `fraction`
        of `free_bytes`, for networks of about `syn_per_neuron` synapses."""
        per_neuron = footprint(1_000_000, 1_000_000 * syn_per_neuron) / 1_000_000
        n = int(free_bytes * fraction / per_neuron)
        return n, n * syn_per_neuron

    def run(self, jobs, t_limit=None, log=None, kernel_every=0):
        """Serves `jobs` (each with an arrival time in seconds from now).
        Returns (completed jobs with build/admit/done times and results,
        summary) where summary holds the unfinished counts and where the
        loop's wall time went: stepping (launch + wait for the chunk),
        admission, retirement (result reads and removal) and idling.  With
        kernel_every > 0 every kernel_every-th step's kernels are timed with
        HIP events (mean microseconds in the summary)."""
        t0 = time.perf_counter()
        acc = collections.Counter()
        if kernel_every:
            self.sim.set_kernel_timing((1 << 1) | (1 << 3), kernel_every)
        arrivals = sorted(jobs, key=lambda j: j.arrival)
        nxt = 0
        building = set()
        waiting = collections.deque()
        running = {}
        finishing = []  # results being summarised on the pool
        done = []
        last_log = t0
        while nxt < len(arrivals) or building or waiting or running or finishing:
            now = time.perf_counter() - t0
            if t_limit is not None and now > t_limit:
                break
            while (nxt < len(arrivals) and arrivals[nxt].arrival <= now
                   and len(building) + len(waiting) < self.max_ready):
                building.add(self.pool.submit(_build, arrivals[nxt]))
                nxt += 1
            for fut in [f for f in finishing if f.done()]:
                finishing.remove(fut)
                done.append(fut.result())
            for fut in [f for f in building if f.done()]:
                building.discard(fut)
                job = fut.result()
                job.t_built = time.perf_counter() - t0
                waiting.append(job)
            # admission in arrival order while the device pools have room
            ta = time.perf_counter()
            while waiting:
                job = waiting[0]
                try:
                    job.nid = self.sim.add(job.net, job.params, job.drive, steps=job.steps)
                except NCError as e:
                    if e.code != 5:  # anything but NC_ERR_CAPACITY
                        raise
                    if running:  # wait for room
                        break
                    # nothing runs and it still does not fit: it never will
                    waiting.popleft()
                    job.result = {"error": "larger than the simulation's device room"}
                    job.t_done = time.perf_counter() - t0
                    job.net.close()
                    job.net = None
                    done.append(job)
                    continue
                waiting.popleft()
                job.t_admit = time.perf_counter() - t0
                job.net.close()  # the simulation keeps what it needs
                job.net = None
                running[job.nid] = job
            tb = time.perf_counter()
            acc["admit_s"] += tb - ta
            if running:
                self.sim.step(self.chunk)
                tc = time.perf_counter()
                self.sim.synchronize()
                td = time.perf_counter()
                acc["launch_s"] += tc - tb
                acc["gpu_wait_s"] += td - tc
                acc["chunks"] += 1
                acc["running_sum"] += len(running)
                for nid in list(running):
                    if self.sim.progress(nid)[1] == 0:
                        job = running.pop(nid)
                        state = read_result(self.sim, nid, job.n, job.n_syn)
                        self.sim.remove(nid)
                        finishing.append(self.finisher.submit(_finish, job, state, t0))
                acc["retire_s"] += time.perf_counter() - td
            else:
                wait = 0.001
                if nxt < len(arrivals) and not building:
                    wait = max(0.0, min(0.01, arrivals[nxt].arrival - (time.perf_counter() - t0)))
                time.sleep(wait)
                acc["idle_s"] += time.perf_counter() - tb
            if log and time.perf_counter() - last_log > 10.0:
                last_log = time.perf_counter()
                log("served %d / %d jobs, %d running, %d waiting, %d arrived, %.1fs" %
                    (len(done), len(arrivals), len(running), len(waiting),
                     sum(1 for j in arrivals if j.arrival <= last_log - t0), last_log - t0))
        summary = {"unfinished": len(arrivals) - len(done), "running": len(running), "waiting": len(waiting),
                   "loop": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in acc.items()}}
        if acc["chunks"]:
            summary["loop"]["mean_running"] = acc["running_sum"] / acc["chunks"]
        if kernel_every:
            for kind, name in ((1, "k_neuron_us"), (3, "k_fanout_us")):
                n, ms = self.sim.kernel_timing(kind)
                summary["loop"][name] = 1000.0 * ms / n if n else None
            self.sim.set_kernel_timing(0)
        return done, summary

    def close(self):
        self.pool.shutdown(wait=True)
        self.finisher.shutdown(wait=True)
        self.sim.close()


def run_alone(spec):
    """One job on the plain path: its own simulation, default launch.
    Returns (result, seconds including construction)."""
    t = time.perf_counter()
    net, drive, params, steps = workloads.build_spec(spec, n_threads=0)
    sim = Simulation.for_network(net)
    nid = sim.add(net, params, drive)
    left = steps
    while left:
        k = min(500, left)
        sim.step(k)
        left -= k
    sim.synchronize()
    res = digest(sim, nid, net.n, net.n_syn)
    el = time.perf_counter() - t
    sim.close()
    net.close()
    return res, el
