"""CPU: the served form's request loop (admission by room, retirement at the
job's budget, lanes alternating on the GPU, refusals other than "no room"
raised) over a stand-in simulation that counts steps and room the way nc_sim
does, so the scheduling is checked without a device.  The GPU tests check
the results themselves."""

import numpy as np
import pytest

from neurocorrelation_amd import serve as SV
from neurocorrelation_amd._lib import NCError


class Device:
    """Which stand-in simulation has a chunk the host has not waited for."""

    def __init__(self):
        self.pending = set()
        self.max_pending = 0


class FakeSim:
    def __init__(self, gpu, max_neurons, max_synapses, max_nets=1, **_):
        self.dev, self.max_n, self.max_s, self.max_nets = gpu, max_neurons, max_synapses, max_nets
        self.nets = {}
        self.next_id = 0

    def add(self, net, params, drive, steps=0):
        assert id(self) not in self.dev.pending, "add while this simulation's chunk is in flight"
        used_n = sum(r["n"] for r in self.nets.values())
        used_s = sum(r["S"] for r in self.nets.values())
        if (len(self.nets) >= self.max_nets or used_n + net.n > self.max_n or used_s + net.n_syn > self.max_s):
            raise NCError(5, "no room")
        nid = self.next_id
        self.next_id += 1
        self.nets[nid] = {"n": net.n, "S": net.n_syn, "done": 0, "left": steps}
        return nid

    def remove(self, nid):
        assert id(self) not in self.dev.pending
        del self.nets[nid]

    def step(self, k):
        for r in self.nets.values():
            run = min(k, r["left"])
            r["done"] += run
            r["left"] -= run
        self.dev.pending.add(id(self))
        self.dev.max_pending = max(self.dev.max_pending, len(self.dev.pending))

    def synchronize(self):
        self.dev.pending.discard(id(self))

    def progress(self, nid):
        assert id(self) not in self.dev.pending, "progress read before waiting for the chunk"
        r = self.nets[nid]
        return r["done"], r["left"]

    def _state(self, nid, n):
        r = self.nets[nid]
        return np.full(n, r["done"], np.float32)

    def potential(self, nid, n):
        return self._state(nid, n)

    def weights(self, nid, n):
        return self._state(nid, n)

    def firings(self, nid, n):
        return self._state(nid, n).astype(np.uint32)

    def set_kernel_timing(self, *a):
        pass

    def kernel_timing(self, kind):
        return 0, 0.0

    def close(self):
        pass


def expected(job):
    n, S, steps = job.n, job.n_syn, job.steps
    v = np.full(n, steps, np.float32)
    return SV.summarize(v, np.full(S, steps, np.float32), v.astype(np.uint32))


@pytest.mark.parametrize("lanes", [1, 2])
def test_every_job_runs_its_budget_and_lanes_alternate(lanes):
    specs = SV.served_specs(24, seed=4, n_range=(300, 1500), k_range=(20, 30), steps_range=(100, 900))
    rng = np.random.default_rng(1)
    arrivals = np.concatenate([[0.0] * 8, np.cumsum(rng.exponential(0.005, size=len(specs) - 8))])
    jobs = [SV.Job(i, sp, float(t)) for i, (sp, t) in enumerate(zip(specs, arrivals))]
    dev = Device()
    server = SV.Server(max_neurons=6000, max_synapses=200_000, max_nets=6, chunk=50, builders=4, lanes=lanes,
                       sim_factory=lambda *a, **k: FakeSim(dev, *a, **k))
    done, rest = server.run(jobs, t_limit=60)
    server.close()
    assert rest["unfinished"] == 0 and len(done) == len(jobs)
    assert dev.max_pending == 1  # one chunk in flight at a time
    assert any(j.t_admit > j.t_built + 1e-4 for j in done)  # room ran out, jobs waited
    for job in done:
        assert job.result == expected(job), job.idx
        assert job.arrival <= job.t_built <= job.t_admit <= job.t_done
    assert rest["loop"]["chunks"] > 0


def test_oversized_job_is_refused_and_the_rest_served():
    small = SV.served_specs(3, seed=5, n_range=(400, 400), k_range=(20, 20), steps_range=(60, 60))
    big = SV.served_specs(1, seed=6, n_range=(5000, 5000), k_range=(20, 20), steps_range=(60, 60))
    jobs = [SV.Job(0, small[0], 0.0), SV.Job(1, big[0], 0.0), SV.Job(2, small[1], 0.0), SV.Job(3, small[2], 0.0)]
    dev = Device()
    server = SV.Server(max_neurons=4000, max_synapses=100_000, max_nets=4, chunk=20, builders=2, lanes=2,
                       sim_factory=lambda *a, **k: FakeSim(dev, *a, **k))
    done, rest = server.run(jobs, t_limit=60)
    server.close()
    by = {j.idx: j for j in done}
    assert rest["unfinished"] == 0
    assert "error" in by[1].result
    for i in (0, 2, 3):
        assert by[i].result == expected(by[i])


def test_errors_other_than_room_stop_the_loop():
    """Only NC_ERR_CAPACITY makes a job wait; any other refusal is raised."""
    specs = SV.served_specs(2, seed=7, n_range=(300, 300), k_range=(20, 20), steps_range=(40, 40))
    jobs = [SV.Job(i, sp, 0.0) for i, sp in enumerate(specs)]
    dev = Device()

    class Refusing(FakeSim):
        def add(self, net, params, drive, steps=0):
            raise NCError(1, "bad argument")

    server = SV.Server(max_neurons=4000, max_synapses=100_000, max_nets=4, chunk=20, builders=2, lanes=2,
                       sim_factory=lambda *a, **k: Refusing(dev, *a, **k))
    with pytest.raises(NCError) as e:
        server.run(jobs, t_limit=30)
    server.close()
    assert e.value.code == 1
