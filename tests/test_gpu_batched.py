"""GPU: the batched form.  Many networks of uneven size and step budget in
one simulation, under every launch shape (waves per workgroup, XCD-ordered
blocks, hipGraph-captured step chunks) and after autotuning, give exactly
the results of the default launch, which the oracle pins elsewhere."""
import sys

import numpy as np
import pytest

import neurocorrelation_amd as nc
from neurocorrelation_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mix(seeds=(0, 1, 2, 3, 4, 5), budgets=(300, 900, 650, 1000, 120, 777)):
    nets = []
    for s in seeds:
        net, drive, params, _ = W.build("plain-300", seed=s)
        drive.bg_prob = 1.0 / 600 if s % 2 else 0.0
        nets.append((net, drive, params))
    big = W.build("plain-20k", seed=7, wiring={"n_neurons": 3000, "side": W.side_for(3000)})
    nets.append((big[0], big[1], big[2]))
    budgets = list(budgets) + [1000]
    return nets, budgets


def _run(nets, budgets, cfg=None, tune=0, calls=(250, 250, 250, 250)):
    pos = sum(n.device_size()[0] for n, _, _ in nets)
    slots = sum(n.device_size()[1] for n, _, _ in nets)
    syn = sum(n.n_syn for n, _, _ in nets)
    sim = nc.Simulation(pos, syn, max_nets=len(nets), raster_capacity=1 << 22, max_slots=slots)
    ids = [sim.add(n, p, d, steps=b) for (n, d, p), b in zip(nets, budgets)]
    if cfg is not None:
        sim.set_launch(*cfg)
    chosen = None
    if tune:
        chosen = sim.tune(tune)
    for k in calls:
        sim.step(k)
    out = []
    for (n, _, _), i in zip(nets, ids):
        out.append((sim.raster(i, 1 << 22), sim.potential(i, n.n), sim.weights(i, n.n_syn), sim.progress(i)))
    sim.close()
    return out, chosen


def _same(a, b):
    for (ra, va, wa, pa), (rb, vb, wb, pb) in zip(a, b):
        assert pa == pb
        assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1])
        assert np.array_equal(va.view(np.uint32), vb.view(np.uint32))
        assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))


@pytest.fixture(scope="module")
def mix_and_ref():
    nets, budgets = _mix()
    say("batched reference run: default launch")
    ref, _ = _run(nets, budgets)
    return nets, budgets, ref


@pytest.mark.parametrize("cfg", [(1, 0, 0), (2, 1, 0, 1, 64), (8, 0, 0), (4, 1, 16, 1, 100), (1, 0, 64, 0, 200),
                                 (8, 1, 200, 1), (4, 0, 0, 0, 0, 1), (2, 0, 16, 0, 0, 1), (8, 0, 0, 1, 0, 1),
                                 (1, 0, 64, 0, 0, 1)])
def test_launch_shapes_give_identical_results(mix_and_ref, cfg):
    nets, budgets, ref = mix_and_ref
    say("launch shape", cfg)
    got, _ = _run(nets, budgets, cfg=cfg)
    _same(got, ref)


def test_tuned_launch_gives_identical_results(mix_and_ref):
    """Tuning takes no simulated step: the tuned run's 1000 steps equal the
    default launch's."""
    nets, budgets, ref = mix_and_ref
    say("autotune, 8 steps per trial")
    got, chosen = _run(nets, budgets, tune=8)
    say("chosen", chosen)
    assert chosen["waves_per_block"] in (1, 2, 4, 8)
    _same(got, ref)


def test_tuner_times_every_candidate_from_the_same_state():
    """A deliberately slow first candidate (fan-out capped at 64 waves on a
    network firing at p ~ 0.27 per step, so each fan-out wave takes five
    slices' fired lanes in turn) loses to the default shape; the tuner leaves
    the simulation exactly where it was, and the run continues as an untuned
    one does."""
    net, drive, params, _ = W.build("plain-20k", seed=3)

    def fresh():
        sim = nc.Simulation.for_network(net, raster_capacity=1 << 24)
        i = sim.add(net, params, drive)
        sim.step(300)  # into the busy regime
        return sim, i

    sim, i = fresh()
    before = (sim.raster(i, 1 << 24), sim.potential(i, net.n), sim.weights(i, net.n_syn), sim.progress(i))
    cands = [dict(waves_per_block=1, xcd_swizzle=1, fanout_waves=64), dict(waves_per_block=4),
             dict(waves_per_block=2), dict(waves_per_block=4, graph_steps=16)]
    chosen, ms = sim.tune_list(cands, steps_per_trial=16, rounds=3)
    say("tuner medians (ms per 16 steps)", ms.tolist(), "chosen", chosen)
    assert chosen["fanout_waves"] != 64
    assert ms[0] > 1.2 * ms[1:].min()
    after = (sim.raster(i, 1 << 24), sim.potential(i, net.n), sim.weights(i, net.n_syn), sim.progress(i))
    assert before[3] == after[3] == (300, 0xFFFFFFFF)
    assert len(after[0][1]) == 0  # no step was simulated (the raster was drained before)
    assert np.array_equal(before[1].view(np.uint32), after[1].view(np.uint32))
    assert np.array_equal(before[2].view(np.uint32), after[2].view(np.uint32))
    sim.step(200)
    ref, j = fresh()
    ref.raster(j, 1 << 24)
    ref.step(200)
    _same([(sim.raster(i, 1 << 24), sim.potential(i, net.n), sim.weights(i, net.n_syn), sim.progress(i))],
          [(ref.raster(j, 1 << 24), ref.potential(j, net.n), ref.weights(j, net.n_syn), ref.progress(j))])


def test_graph_chunks_with_host_events():
    """Host-scheduled fires force direct launches for the steps that carry
    them; the captured chunks and the direct steps interleave exactly."""
    net, drive, params, _ = W.build("plain-300", seed=9)
    ev_steps = np.array([5, 40, 41, 333, 600], np.uint32)
    ids = np.array([3, 17, 17, 250, 0], np.uint32)
    kinds = np.array([0, 1, 0, 1, 0], np.uint8)

    def run(cfg):
        sim = nc.Simulation.for_network(net, raster_capacity=1 << 20)
        i = sim.add(net, params, drive)
        sim.set_launch(*cfg)
        sim.step(64)
        sim.push_forced(i, ev_steps, ids, kinds)
        for _ in range(5):
            sim.step(160)
        r = sim.raster(i, 1 << 20)
        v = sim.potential(i, net.n)
        w = sim.weights(i, net.n_syn)
        sim.close()
        return r, v, w

    a = run((4, 0, 0))
    b = run((2, 1, 32))
    assert np.array_equal(a[0][0], b[0][0]) and np.array_equal(a[0][1], b[0][1])
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32))


def test_prepared_network_adds_identically():
    """A network whose device image was built once (nc_net_prepare) and added
    twice, beside the other networks of a simulation, runs exactly as the same
    network added without a kept image."""
    nets, budgets = _mix(seeds=(11, 12), budgets=(700, 500))
    ref, _ = _run(nets, budgets)
    prepared = [(n.prepare(), d, p) for n, d, p in nets]
    got, _ = _run(prepared, budgets)
    _same(got, ref)
    twice = prepared + [prepared[0]]
    got2, _ = _run(twice, budgets + [700])
    _same(got2[:len(ref)], ref)
    _same(got2[-1:], ref[:1])
