"""GPU parity: the HIP path against the reference's own outputs (golden
fixtures) and against the oracle, bit for bit, through the C ABI."""
import sys

import numpy as np
import pytest

import neurocorrelation_amd as nc
from neurocorrelation_amd import workloads as W
from conftest import golden_files, load_golden

pytestmark = pytest.mark.gpu

NEVER = nc.NEVER


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def steps_to_time(steps, dt, never_value):
    t = steps.astype(np.float32) * np.float32(dt)
    return np.where(steps == NEVER, np.float32(never_value), t).astype(np.float32)


def gpu_state(sim, nid, n, S, dt):
    arr, la, dp = sim.synapse_state(nid, S)
    lf = sim.last_fire(nid, n)
    return {
        "V": sim.potential(nid, n),
        "w": sim.weights(nid, S),
        "fire_t": np.where(arr == 0, np.float32(0), arr.astype(np.float32) * np.float32(dt)).astype(np.float32),
        "last_arr": steps_to_time(la, dt, -np.inf),
        "last_fire": steps_to_time(lf, dt, np.nan),
        "depol": dp,
        "firings": sim.firings(nid, n),
    }


def assert_bits(a, b, what):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
    assert bad.size == 0, "%s differs at %s: %s vs %s" % (what, bad[:5], a[bad[:5]], b[bad[:5]])


def net_from_csr(in_ptr, in_src, w, delay, inh):
    n = len(in_ptr) - 1
    tgt = np.repeat(np.arange(n, dtype=np.uint32), np.diff(in_ptr))
    return nc.Network.from_edges(n, tgt, in_src, w, delay, inh)


@pytest.mark.parametrize("path", golden_files())
def test_gpu_matches_reference_fixture(path):
    """Raster, every potential after every step, and the final synapse state
    equal what the compiled reference's NeuCor::run() produced."""
    g = load_golden(path)
    say("fixture", path)
    n = len(g["in_ptr"]) - 1
    S = len(g["in_src"])
    dt = float(g["dt"])
    net = net_from_csr(g["in_ptr"], g["in_src"], g["w0"], g["delay"], g["inh"])
    sim = nc.Simulation.for_network(net, raster_capacity=1 << 16)
    nid = sim.add(net, nc.default_params(dt=dt))
    sim.push_forced(nid, g["ev_step"], g["ev_id"], g["ev_kind"])
    steps = int(g["steps"])
    for k in range(steps):
        sim.step(1)
        assert_bits(sim.potential(nid, n), g["ref_V"][k], "V at step %d" % (k + 1))
    rs, ri = sim.raster(nid, 1 << 16)
    ref_steps = np.repeat(np.arange(1, steps + 1), np.diff(g["ref_raster_ptr"]))
    assert np.array_equal(rs, ref_steps) and np.array_equal(ri, g["ref_raster_ids"])
    st = gpu_state(sim, nid, n, S, dt)
    assert_bits(st["w"], g["ref_w"], "w")
    assert np.array_equal(st["fire_t"], g["ref_fire_t"])
    assert np.array_equal(st["last_arr"], g["ref_last_arr"])
    assert np.array_equal(st["last_fire"], g["ref_last_fire"], equal_nan=True)


def run_pair(oracle_mod, net, drive, params, chunks, chunk=None, raster_cap=1 << 22, seed=0, launch=None):
    """Advance the oracle and the GPU together, comparing every chunk
    (`chunks` chunks of `chunk` steps, or the listed chunk lengths)."""
    sizes = list(chunks) if chunk is None else [chunk] * chunks
    ex = net.export()
    n, S = len(ex["in_ptr"]) - 1, len(ex["in_src"])
    o = oracle_mod.Sim(ex, dt=params.dt, lr=params.learning_rate, pre_f=params.pre_factor,
                       post_f=params.post_factor)
    ids, groups = net.select_inputs(drive)
    ncorr = drive.n_groups if drive.n_correlated > 0 else 0
    o.set_drive(ids, groups, ncorr, drive.period, drive.p0, drive.amp_corr, drive.amp_ind, drive.bg_prob,
                drive.seed)
    sim = nc.Simulation.for_network(net, raster_capacity=raster_cap)
    nid = sim.add(net, params, drive)
    if launch is not None:
        sim.set_launch(**launch)
    total = 0
    for c, chunk in enumerate(sizes):
        os_, oi = o.step(chunk, cap=n * chunk)
        sim.step(chunk)
        gs, gi = sim.raster(nid, raster_cap)
        assert np.array_equal(gs, os_) and np.array_equal(gi, oi), "raster differs in chunk %d" % c
        total += len(oi)
        st = gpu_state(sim, nid, n, S, params.dt)
        ost = o.state()
        assert_bits(st["V"], ost["V"], "V after chunk %d" % c)
        assert_bits(st["w"], ost["w"], "w after chunk %d" % c)
        assert np.array_equal(st["fire_t"], ost["fire_t"])
        assert np.array_equal(st["last_arr"], ost["last_arr"])
        assert np.array_equal(st["last_fire"], ost["last_fire"], equal_nan=True)
        assert np.array_equal(st["firings"], ost["firings"])
        assert_bits(st["depol"], ost["depol"], "depol")
    return total, sim, nid


def test_plain300_full_horizon(oracle_mod):
    """10^4 steps of the 300-neuron workload: identical rasters throughout,
    potentials and weights bit-equal."""
    net, drive, params, steps = W.build("plain-300", seed=0)
    say("plain-300 seed 0, %d steps" % steps)
    total, _, _ = run_pair(oracle_mod, net, drive, params, 10, steps // 10)
    assert total > 1000


@pytest.mark.parametrize("seed", [1, 2])
def test_plain300_other_seeds_with_background(oracle_mod, seed):
    net, drive, params, _ = W.build("plain-300", seed=seed)
    drive.bg_prob = 1.0 / 600.0  # the reference's background rate at dt = 1 ms
    say("plain-300 seed %d with background" % seed)
    total, _, _ = run_pair(oracle_mod, net, drive, params, 4, 500)
    assert total > 100


def test_plain20k(oracle_mod):
    """3000 steps of the 20,000-neuron workload (p ~ 0.27 per step)."""
    net, drive, params, _ = W.build("plain-20k", seed=0)
    say("plain-20k, 3000 steps")
    total, _, _ = run_pair(oracle_mod, net, drive, params, 6, 500, raster_cap=1 << 24)
    assert total > 300_000


@pytest.mark.parametrize("graph_steps", [0, 96])
def test_calls_longer_than_the_group_ring(oracle_mod, graph_steps):
    """One call of 2500 steps draws its group spike trains in sub-chunks of
    at most 1024 steps (960 with graph chunks of 96: whole chunks per
    sub-chunk, then direct launches for the rest); rasters and state equal
    the oracle's."""
    net, drive, params, _ = W.build("plain-300", seed=4)
    say("plain-300, one 2500-step call, graph_steps", graph_steps)
    total, _, _ = run_pair(oracle_mod, net, drive, params, [2500, 1100], launch=dict(graph_steps=graph_steps))
    assert total > 500


def test_plain1m_full_size(oracle_mod):
    """The 10^6-neuron, 10^8-synapse network at full size: 144 steps bit-equal
    to the oracle (delays up to 64 steps arrive; the pending lists' arrival
    windows flip nine times), then size-independent invariants over 400
    steps."""
    net, drive, params, _ = W.build("plain-1m", seed=0)
    say("plain-1m built", net.info())
    total, sim, nid = run_pair(oracle_mod, net, drive, params, [8, 8, 64, 64], raster_cap=1 << 24)
    assert total > 1000
    n, S = net.n, net.n_syn
    before = sim.firings(nid, n).astype(np.int64).sum()
    sim.step(400)
    gs, gi = sim.raster(nid, 1 << 24)
    fir = sim.firings(nid, n).astype(np.int64)
    assert fir.sum() - before == len(gi)                       # raster == firing counts
    assert np.all(np.diff(gs.astype(np.int64)) >= 0)
    w = sim.weights(nid, S)
    ex = net.export()
    inh = ex["inh"].astype(bool)
    assert np.all((w[~inh] >= 0) & (w[~inh] <= 1)) and np.all((w[inh] >= -1) & (w[inh] <= 0))
    V = sim.potential(nid, n)
    assert np.all(np.isfinite(V))
    lf = sim.last_fire(nid, n)
    # refractory: no non-input neuron fires twice within 2 ms
    ids, _ = net.select_inputs(drive)
    mask = np.ones(n, bool)
    mask[ids] = False
    order = np.lexsort((gs, gi))
    si, ss = gi[order], gs[order].astype(np.int64)
    same = (si[1:] == si[:-1]) & mask[si[1:]]
    assert np.all(ss[1:][same] - ss[:-1][same] > 2)
    assert lf.max() <= 544
    rate = len(gi) / (n * 400.0)
    say("plain-1m rate", rate)
    assert 0.001 < rate < 0.02


def test_batched_equals_alone(oracle_mod):
    """Networks of uneven size packed into one simulation give exactly the
    results each gives alone."""
    specs = [("plain-300", 0), ("plain-300", 3), ("plain-20k", 1), ("plain-300", 5)]
    nets = [W.build(name, seed=s) for name, s in specs]
    sizes = [net.device_size() for net, *_ in nets]
    total_pos = sum(p for p, _ in sizes)
    total_slots = sum(s for _, s in sizes)
    total_syn = sum(net.n_syn for net, *_ in nets)
    mix = nc.Simulation(total_pos, total_syn, max_nets=8, raster_capacity=1 << 24, max_slots=total_slots)
    ids = [mix.add(net, params, drive) for net, drive, params, _ in nets]
    mix.step(700)
    for (net, drive, params, _), nid in zip(nets, ids):
        alone = nc.Simulation.for_network(net, raster_capacity=1 << 24)
        a = alone.add(net, params, drive)
        alone.step(700)
        r1 = mix.raster(nid, 1 << 24)
        r2 = alone.raster(a, 1 << 24)
        assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])
        assert_bits(mix.potential(nid, net.n), alone.potential(a, net.n), "V")
        assert_bits(mix.weights(nid, net.n_syn), alone.weights(a, net.n_syn), "w")


def test_uneven_step_budgets_and_reuse(oracle_mod):
    """Networks with different step budgets stop on time; a removed
    network's room is reused and the newcomer starts from a clean state."""
    a_net, a_drive, p, _ = W.build("plain-300", seed=0)
    b_net, b_drive, _, _ = W.build("plain-300", seed=1)
    pos, slots = a_net.device_size()
    sim = nc.Simulation(2 * pos, 2 * a_net.n_syn + 2 * b_net.n_syn, max_nets=2, raster_capacity=1 << 20,
                        max_slots=4 * slots)
    a = sim.add(a_net, p, a_drive, steps=300)
    b = sim.add(b_net, p, b_drive, steps=700)
    sim.step(500)
    assert sim.progress(a) == (300, 0)
    assert sim.progress(b) == (500, 200)
    sim.step(500)
    assert sim.progress(b) == (700, 0)
    ra = sim.raster(a, 1 << 20)
    sim.remove(a)
    c = sim.add(a_net, p, a_drive, steps=300)
    sim.step(300)
    rc = sim.raster(c, 1 << 20)
    assert np.array_equal(ra[0], rc[0]) and np.array_equal(ra[1], rc[1])
    with pytest.raises(nc.NCError):
        sim.add(a_net, p, a_drive)  # no third network slot


def test_live_params_weights_and_reset(oracle_mod):
    """learningRate / factors changed between runs, setWeight and
    resetActivities behave as in the oracle."""
    net, drive, params, _ = W.build("plain-300", seed=2)
    ex = net.export()
    o = oracle_mod.Sim(ex)
    ids, groups = net.select_inputs(drive)
    o.set_drive(ids, groups, 2, drive.period, drive.p0, drive.amp_corr, drive.amp_ind, 0.0, drive.seed)
    sim = nc.Simulation.for_network(net, raster_capacity=1 << 20)
    nid = sim.add(net, params, drive)
    for lr, pf, qf in [(1.0, 0.13, 0.30), (0.5, 0.2, 0.1), (0.0, 0.13, 0.30), (2.0, 0.05, 0.4)]:
        sim.set_params(nid, lr, pf, qf)
        o.set_params(lr, pf, qf)
        sim.step(333)
        o.step(333)
        assert_bits(sim.weights(nid, net.n_syn), o.state()["w"], "w")
        assert_bits(sim.potential(nid, net.n), o.state()["V"], "V")
    # the weights are written with spikes in flight (their arrival entries
    # carry the old weights), fired by the eight widest sources one step
    # before; the run goes on as in the oracle with the same write
    top = np.argsort(-np.bincount(ex["in_src"], minlength=net.n), kind="stable")[:8]
    sim.push_forced(nid, [1] * 8, top, [0] * 8)
    o.push_forced([1] * 8, top, [0] * 8)
    sim.step(1)
    o.step(1)
    arr, _, _ = sim.synapse_state(nid, net.n_syn)
    assert (arr > sim.progress(nid)[0]).sum() >= 8
    w = sim.weights(nid, net.n_syn)
    assert_bits(w, o.state()["w"], "w before the write")
    w2 = np.where(np.arange(len(w)) % 3 == 0, -np.abs(w), w).astype(np.float32)
    sim.write_weights(nid, w2)
    o.set_weights(w2)
    assert_bits(sim.weights(nid, net.n_syn), w2, "written w")
    sim.step(333)
    o.step(333)
    assert_bits(sim.weights(nid, net.n_syn), o.state()["w"], "w after the write")
    assert_bits(sim.potential(nid, net.n), o.state()["V"], "V after the write")
    sim.reset_activities(nid)
    assert sim.firings(nid, net.n).sum() == 0


def test_host_events_fine_dt(oracle_mod):
    """FEW_NEURONS-style drive through host-scheduled fires at dt = 1/64 ms."""
    dt = 1.0 / 64
    net = nc.Network.from_edges(3, [1, 2], [0, 0], [0.5, 0.5], [54, 54])
    ex = net.export()
    o = oracle_mod.Sim(ex, dt=dt)
    sim = nc.Simulation.for_network(net, raster_capacity=1 << 16)
    nid = sim.add(net, nc.default_params(dt=dt))
    per = 20 * 64
    so, ids = [], []
    for k in range(1, 12):
        so += [k * per, k * per + 128, k * per - 128]
        ids += [0, 1, 2]
    o.push_forced(so, ids, [0] * len(ids))
    sim.push_forced(nid, so, ids, [0] * len(ids))
    o.step(12 * per)
    sim.step(12 * per)
    w = sim.weights(nid, 2)
    assert_bits(w, o.state()["w"], "w")
    assert w[0] == pytest.approx(1.0) and w[1] == pytest.approx(0.0)
    assert_bits(sim.potential(nid, 3), o.state()["V"], "V")


def test_edge_networks(oracle_mod):
    """No synapses at all; a neuron without inputs; the longest delay."""
    empty = nc.Network.from_edges(70, [], [], [], [])
    sim = nc.Simulation.for_network(empty, raster_capacity=1024)
    nid = sim.add(empty, nc.default_params())
    sim.push_forced(nid, [1, 5], [3, 69], [0, 1])
    sim.step(10)
    s, i = sim.raster(nid, 1024)
    assert s.tolist() == [1, 5] and i.tolist() == [3, 69]
    longd = nc.Network.from_edges(2, [1], [0], [0.8], [250])
    sim2 = nc.Simulation.for_network(longd, raster_capacity=1024)
    n2 = sim2.add(longd, nc.default_params())
    o = oracle_mod.Sim(longd.export())
    sim2.push_forced(n2, [1, 2], [0, 0], [0, 0])
    o.push_forced([1, 2], [0, 0], [0, 0])
    sim2.step(260)
    o.step(260)
    assert_bits(sim2.potential(n2, 2), o.state()["V"], "V")
    arr, la, _ = sim2.synapse_state(n2, 1)
    assert la[0] == 251
    with pytest.raises(nc.NCError):
        bad = nc.Network.from_edges(2, [1], [0], [0.8], [252])
        nc.Simulation.for_network(bad).add(bad, nc.default_params())


def test_largest_in_degree(oracle_mod):
    """A neuron with the largest supported in-degree (262140): a spike from
    its last source (the slot furthest into its slice) arrives and charges it
    exactly as in the oracle; one synapse more is refused."""
    k = 262_140
    src = np.arange(1, k + 1, dtype=np.uint32)
    net = nc.Network.from_edges(k + 1, np.zeros(k, np.uint32), src, np.full(k, 0.8, np.float32),
                                np.ones(k, np.uint8))
    sim = nc.Simulation.for_network(net, raster_capacity=1024)
    nid = sim.add(net, nc.default_params())
    o = oracle_mod.Sim(net.export())
    for x in (sim, o):
        if x is sim:
            x.push_forced(nid, [1, 1], [k, k - 1], [0, 0])
        else:
            x.push_forced([1, 1], [k, k - 1], [0, 0])
    sim.step(6)
    o.step(6)
    assert_bits(sim.potential(nid, k + 1), o.state()["V"], "V")
    assert_bits(sim.weights(nid, k), o.state()["w"], "w")
    too_many = nc.Network.from_edges(k + 2, np.zeros(k + 1, np.uint32), np.arange(1, k + 2, dtype=np.uint32),
                                     np.full(k + 1, 0.8, np.float32), np.ones(k + 1, np.uint8))
    sim2 = nc.Simulation.for_network(too_many)
    with pytest.raises(nc.NCError) as e:
        sim2.add(too_many, nc.default_params())
    assert e.value.code == 1


def test_arrival_bucket_overflow(oracle_mod):
    """Arrivals beyond a slice's bucket capacity for one step (256 spikes
    into one 64-neuron slice at the same step) and delays past the bucket
    ring (100 steps) take the pending lists; potentials, weights and latches
    equal the oracle's."""
    srcs, tg = [64, 65, 66, 67], np.arange(64)
    to = np.concatenate([tg for _ in srcs] + [tg[:8]])
    frm = np.concatenate([np.full(64, s) for s in srcs] + [np.full(8, 68)])
    w = np.full(len(to), 0.05, np.float32)
    delay = np.concatenate([np.full(256, 5), np.full(8, 100)]).astype(np.uint8)
    net = nc.Network.from_edges(69, to, frm, w, delay)
    ex = net.export()
    o = oracle_mod.Sim(ex)
    sim = nc.Simulation.for_network(net, raster_capacity=1 << 16)
    nid = sim.add(net, nc.default_params())
    so = [1, 1, 1, 1, 1, 9, 9, 9, 9]
    ids = [64, 65, 66, 67, 68, 64, 65, 66, 67]
    o.push_forced(so, ids, [0] * len(ids))
    sim.push_forced(nid, so, ids, [0] * len(ids))
    for k in (4, 3, 10, 100):
        o.step(k)
        sim.step(k)
        st = gpu_state(sim, nid, 69, len(to), 1.0)
        ost = o.state()
        assert_bits(st["V"], ost["V"], "V")
        assert_bits(st["w"], ost["w"], "w")
        assert np.array_equal(st["fire_t"], ost["fire_t"])
        assert np.array_equal(st["last_arr"], ost["last_arr"])
    far = ex["in_src"] == 68
    assert np.all(ost["last_arr"][~far] == 14.0) and np.all(ost["last_arr"][far] == 101.0)
