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


class _Lane:
    """One simulation of the server and the jobs it runs."""

    def __init__(self, sim):
        self.sim = sim
        self.running = {}  # net id -> job


class Server:
    """Simulations ("lanes", each with its own stream) sharing `max_neurons` /
    `max_synapses` of device room and up to `max_nets` jobs; `chunk` steps
    between admission and retirement checks.  With two lanes, the GPU runs
    one lane's chunk while the host retires and admits the other's jobs."""

    def __init__(self, max_neurons, max_synapses, max_nets=4096, chunk=200, builders=8, device=0, max_ready=256,
                 lanes=2, sim_factory=Simulation):
        n, syn, nets = max_neurons // lanes, max_synapses // lanes, max(1, max_nets // lanes)
        slots = syn + syn // 4 + (n + 64 * nets) * 4
        self.lanes = [_Lane(sim_factory(n, syn, max_nets=nets, device=device, max_slots=slots))
                      for _ in range(lanes)]
        self.chunk = chunk
        self.max_ready = max_ready  # jobs built or being built but not admitted (bounds host memory)
        self.pool = cf.ThreadPoolExecutor(builders)
        self.finisher = cf.ThreadPoolExecutor(2)  # results, apart from the builders' queue

    @property
    def sim(self):
        return self.lanes[0].sim

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
        loop's wall time went: launching chunks, waiting for the GPU,
        admission, retirement (result reads and removal) and idling.  With
        kernel_every > 0 every kernel_every-th step's kernels are timed with
        HIP events (mean microseconds in the summary)."""
        t0 = time.perf_counter()
        acc = collections.Counter()
        if kernel_every:
            for ln in self.lanes:
                ln.sim.set_kernel_timing((1 << 1) | (1 << 3), kernel_every)
        arrivals = sorted(jobs, key=lambda j: j.arrival)
        nxt = 0
        building = set()
        waiting = collections.deque()
        finishing = []  # results being summarised on the pool
        done = []
        in_flight = None  # the lane whose chunk the GPU is running
        last_log = t0

        def n_running():
            return sum(len(ln.running) for ln in self.lanes)

        while nxt < len(arrivals) or building or waiting or finishing or n_running():
            now = time.perf_counter() - t0
            if t_limit is not None and now > t_limit:
                break
            served_any = False
            for ln in self.lanes:
                while (nxt < len(arrivals) and arrivals[nxt].arrival <= time.perf_counter() - t0
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
                if ln is in_flight:
                    if len(self.lanes) > 1:
                        continue  # the other lanes are serviced while this one runs
                    tw = time.perf_counter()
                    ln.sim.synchronize()
                    in_flight = None
                    acc["gpu_wait_s"] += time.perf_counter() - tw
                # retirement: this lane has no chunk in flight
                tr = time.perf_counter()
                for nid in list(ln.running):
                    if ln.sim.progress(nid)[1] == 0:
                        job = ln.running.pop(nid)
                        state = read_result(ln.sim, nid, job.n, job.n_syn)
                        ln.sim.remove(nid)
                        finishing.append(self.finisher.submit(_finish, job, state, t0))
                ta = time.perf_counter()
                acc["retire_s"] += ta - tr
                # admission in arrival order while this lane has room
                while waiting:
                    job = waiting[0]
                    try:
                        job.nid = ln.sim.add(job.net, job.params, job.drive, steps=job.steps)
                    except NCError as e:
                        # no room in the lane's pools (NC_ERR_CAPACITY) or for
                        # its staging upload (NC_ERR_OOM): wait for room;
                        # anything else is an error
                        if e.code not in (4, 5):
                            raise
                        if ln.running:  # wait for room
                            break
                        # an empty lane has no room for it: no lane ever will
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
                    ln.running[job.nid] = job
                tb = time.perf_counter()
                acc["admit_s"] += tb - ta
                # the GPU finishes the other lane's chunk, then takes this one
                if in_flight is not None:
                    in_flight.sim.synchronize()
                    in_flight = None
                tc = time.perf_counter()
                acc["gpu_wait_s"] += tc - tb
                if ln.running:
                    ln.sim.step(self.chunk)
                    in_flight = ln
                    served_any = True
                    acc["chunks"] += 1
                    acc["running_sum"] += len(ln.running)
                acc["launch_s"] += time.perf_counter() - tc
            if not served_any and in_flight is None:
                tb = time.perf_counter()
                wait = 0.001
                if nxt < len(arrivals) and not building:
                    wait = max(0.0, min(0.01, arrivals[nxt].arrival - (time.perf_counter() - t0)))
                time.sleep(wait)
                acc["idle_s"] += time.perf_counter() - tb
            if log and time.perf_counter() - last_log > 10.0:
                last_log = time.perf_counter()
                log("served %d / %d jobs, %d running, %d waiting, %d arrived, %.1fs" %
                    (len(done), len(arrivals), n_running(), len(waiting),
                     sum(1 for j in arrivals if j.arrival <= last_log - t0), last_log - t0))
        if in_flight is not None:
            in_flight.sim.synchronize()
        summary = {"unfinished": len(arrivals) - len(done), "running": n_running(), "waiting": len(waiting),
                   "lanes": len(self.lanes),
                   "loop": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in acc.items()}}
        if acc["chunks"]:
            summary["loop"]["mean_running_per_chunk"] = acc["running_sum"] / acc["chunks"]
        if kernel_every:
            for kind, name in ((1, "k_neuron_us"), (3, "k_fanout_us")):
                n = ms = 0
                for ln in self.lanes:
                    a, b = ln.sim.kernel_timing(kind)
                    n, ms = n + a, ms + b
                summary["loop"][name] = 1000.0 * ms / n if n else None
            for ln in self.lanes:
                ln.sim.set_kernel_timing(0)
        return done, summary

    def close(self):
        self.pool.shutdown(wait=True)
        self.finisher.shutdown(wait=True)
        for ln in self.lanes:
            ln.sim.close()


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
