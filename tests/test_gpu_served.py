"""GPU: the served form.  Jobs arriving over time into one running
simulation, admitted as room frees up and batched with whatever else is
running, give exactly the result each gives alone on the plain path."""
import sys

import numpy as np
import pytest

from neurocorrelation_amd import serve as SV
from neurocorrelation_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("lanes", [1, 2])
def test_served_jobs_equal_alone_under_queueing(lanes):
    specs = SV.served_specs(10, seed=11, n_range=(300, 2500), k_range=(20, 40), steps_range=(150, 700))
    rng = np.random.default_rng(5)
    arrivals = np.concatenate([[0.0] * 4, np.cumsum(rng.exponential(0.02, size=len(specs) - 4))])
    jobs = [SV.Job(i, sp, float(t)) for i, (sp, t) in enumerate(zip(specs, arrivals))]
    # room for two jobs at a time: four arrive together, so some wait for room
    server = SV.Server(max_neurons=6000, max_synapses=150_000, max_nets=2, chunk=64, builders=4, lanes=lanes)
    say("serving", len(jobs), "jobs on", lanes, "lane(s)")
    done, rest = server.run(jobs, t_limit=120)
    server.close()
    assert rest["unfinished"] == 0 and len(done) == len(jobs)
    assert any(j.t_admit > j.t_built + 1e-3 for j in done)  # someone waited for room
    for job in sorted(done, key=lambda j: j.idx):
        say("job", job.idx, "n", job.n, "steps", job.steps)
        res, _ = SV.run_alone(job.spec)
        assert res == job.result, "job %d differs alone and in the mix" % job.idx


def test_job_larger_than_the_room_is_refused():
    specs = SV.served_specs(2, seed=12, n_range=(400, 400), k_range=(20, 20), steps_range=(50, 50))
    big = SV.served_specs(1, seed=13, n_range=(20_000, 20_000), k_range=(30, 30), steps_range=(50, 50))
    jobs = [SV.Job(0, specs[0], 0.0), SV.Job(1, big[0], 0.0), SV.Job(2, specs[1], 0.01)]
    server = SV.Server(max_neurons=4000, max_synapses=60_000, max_nets=4, chunk=32, builders=2)
    done, rest = server.run(jobs, t_limit=60)
    server.close()
    by = {j.idx: j for j in done}
    assert rest["unfinished"] == 0
    assert "error" in by[1].result
    assert "hash" in by[0].result and "hash" in by[2].result


def test_served_jobs_match_the_oracle(oracle_mod):
    """Two jobs of the served stream (seeded drive, 20% inhibitory weights,
    delays up to 32 steps), run in the served mix, equal the oracle's run of
    each job: firings, final potentials and weights."""
    import hashlib

    specs = SV.served_specs(2, seed=21, n_range=(700, 1800), k_range=(40, 90), steps_range=(1500, 2500))
    jobs = [SV.Job(i, sp, 0.0) for i, sp in enumerate(specs)]
    server = SV.Server(max_neurons=8000, max_synapses=400_000, max_nets=4, chunk=100, builders=2, lanes=2)
    done, rest = server.run(jobs, t_limit=120)
    server.close()
    assert rest["unfinished"] == 0
    for job in sorted(done, key=lambda j: j.idx):
        net, drive, params, steps = W.build_spec(job.spec)
        ex = net.export()
        o = oracle_mod.Sim(ex, dt=params.dt, lr=params.learning_rate, pre_f=params.pre_factor,
                           post_f=params.post_factor)
        ids, groups = net.select_inputs(drive)
        ncorr = drive.n_groups if drive.n_correlated > 0 else 0
        o.set_drive(ids, groups, ncorr, drive.period, drive.p0, drive.amp_corr, drive.amp_ind, drive.bg_prob,
                    drive.seed)
        say("oracle: job", job.idx, "n", net.n, "steps", steps)
        o.step(steps, cap=net.n * steps)
        st = o.state()
        h = hashlib.sha256()
        h.update(st["V"].astype(np.float32).tobytes())
        h.update(st["w"].astype(np.float32).tobytes())
        h.update(st["firings"].astype(np.uint32).tobytes())
        assert job.result == {"spikes": int(st["firings"].sum()), "hash": h.hexdigest()}, job.idx
        net.close()
