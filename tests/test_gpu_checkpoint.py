"""GPU: checkpoints.  A network saved mid-run and restored into a fresh
simulation (at a different place in device memory, next to another network)
continues exactly as the original does: same potentials, weights, firing
counts and spikes, including spikes in flight and queued host events."""
import sys

import numpy as np
import pytest

import neurocorrelation_amd as nc
from neurocorrelation_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _state(sim, i, net):
    return (sim.potential(i, net.n), sim.weights(i, net.n_syn), sim.firings(i, net.n),
            sim.synapse_state(i, net.n_syn))


def test_save_restore_continues_exactly():
    net, drive, params, _ = W.build("plain-300", seed=4)
    drive.bg_prob = 1.0 / 600
    other, odrive, oparams, _ = W.build("plain-300", seed=8)
    say("original run")
    a = nc.Simulation.for_network(net, raster_capacity=1 << 20)
    ia = a.add(net, params, drive)
    a.step(350)
    a.push_forced(ia, np.array([3, 60], np.uint32), np.array([7, 11], np.uint32), np.array([0, 1], np.uint8))
    ckpt = a.save(ia)
    a.raster(ia, 1 << 20)  # drop the spikes before the checkpoint
    a.step(250)
    ref = _state(a, ia, net)
    ref_r = a.raster(ia, 1 << 20)
    a.close()

    say("restored run")
    pos = net.device_size()[0] + other.device_size()[0]
    slots = net.device_size()[1] + other.device_size()[1]
    b = nc.Simulation(pos, net.n_syn + other.n_syn, max_nets=2, raster_capacity=1 << 20, max_slots=slots)
    b.add(other, oparams, odrive)  # another network first: different offsets
    ib = b.add(net, params, drive)
    b.restore(ib, ckpt)
    assert b.progress(ib)[0] == 350
    b.step(250)
    got = _state(b, ib, net)
    got_r = b.raster(ib, 1 << 20)
    for x, y in zip(ref[:3], got[:3]):
        assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))
    for x, y in zip(ref[3], got[3]):
        assert np.array_equal(x, y)
    assert np.array_equal(ref_r[0], got_r[0]) and np.array_equal(ref_r[1], got_r[1])
    b.close()


def test_restore_refuses_mismatches():
    net, drive, params, _ = W.build("plain-300", seed=4)
    other, _, _, _ = W.build("plain-300", seed=5)
    a = nc.Simulation.for_network(net)
    ia = a.add(net, params, drive)
    a.step(20)
    ckpt = a.save(ia)
    with pytest.raises(nc.NCError):
        a.restore(ia, ckpt)  # this network has stepped
    b = nc.Simulation.for_network(other)
    ib = b.add(other, params, drive)
    with pytest.raises(nc.NCError):
        b.restore(ib, ckpt)  # different wiring
    with pytest.raises(nc.NCError):
        b.restore(ib, ckpt[:-4])  # truncated
    a.close()
    b.close()


HEADER = 56  # magic, version, n, pad, n_syn (u64), steps, lr, pre_f, post_f, dt, n_events, digest (u64)


def test_restore_rejects_out_of_range_steps():
    """A checkpoint whose last fire or last arrival lies after its own step
    count, or whose arrival lies outside [1, steps + 255], or with an event of
    unknown kind, is refused with NC_ERR_ARG before anything is written (the
    step tables would be indexed out of range); the intact checkpoint then
    restores."""
    net, drive, params, _ = W.build("plain-300", seed=4)
    a = nc.Simulation.for_network(net)
    ia = a.add(net, params, drive)
    a.step(40)
    a.push_forced(ia, np.array([2], np.uint32), np.array([5], np.uint32), np.array([1], np.uint8))
    good = a.save(ia)
    a.close()
    n, S = net.n, net.n_syn
    steps = int(np.frombuffer(good, np.uint32, 1, 24)[0])
    assert steps == 40
    neu, syn, ev = HEADER, HEADER + 12 * n, HEADER + 12 * n + 16 * S

    def patched(off, value):
        b = bytearray(good)
        b[off:off + 4] = np.array([value], np.int32).tobytes()
        return bytes(b)

    bad = [patched(neu + 4, steps + 1),            # neuron 0 fired after the checkpoint's step
           patched(neu + 4, -(1 << 31) + 1),       # a last fire that would overflow a step difference
           patched(syn + 8, steps + 3),            # synapse 0 last arrival in the future
           patched(syn + 12, steps + 256),         # arrival beyond the longest delay
           patched(syn + 12, -5),                  # arrival before the first step, not "none"
           patched(ev + 8, 2)]                     # event kind neither input nor background
    b = nc.Simulation.for_network(net)
    ib = b.add(net, params, drive)
    for k, x in enumerate(bad):
        with pytest.raises(nc.NCError) as e:
            b.restore(ib, x)
        assert e.value.code == 1, k
        assert b.progress(ib)[0] == 0
    b.restore(ib, patched(syn + 12, -(1 << 30)))  # "no spike" is a valid arrival
    b.close()
    c = nc.Simulation.for_network(net)
    ic = c.add(net, params, drive)
    c.restore(ic, good)
    assert c.progress(ic)[0] == 40
    c.step(10)
    c.close()
