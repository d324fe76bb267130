"""GPU: nc_read_weights_noalloc, the read of a network's weights that
allocates no device memory (a static per-GPU scratch array), against
nc_read_weights: bit-equal on two calls in a row, the second on a smaller
network; free device memory unchanged by the calls; a network larger than
the scratch refused with NC_ERR_RANGE."""
import sys

import numpy as np
import pytest

import neurocorrelation_amd as nc
from neurocorrelation_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_noalloc_read_equals_the_allocating_read():
    big = W.build("plain-20k", seed=2)
    small = W.build("plain-300", seed=6)
    nets = [big, small]
    pos = sum(b[0].device_size()[0] for b in nets)
    slots = sum(b[0].device_size()[1] for b in nets)
    syn = sum(b[0].n_syn for b in nets)
    sim = nc.Simulation(pos, syn, max_nets=2, max_slots=slots)
    ids = [sim.add(net, params, drive) for net, drive, params, _ in nets]
    say("no-alloc read: stepping %d + %d synapses" % (big[0].n_syn, small[0].n_syn))
    sim.step(300)
    # warm every path once (the allocating read may grow its own buffer)
    for (net, *_), i in zip(nets, ids):
        sim.weights(i, net.n_syn)
    for rnd in range(2):
        for (net, *_), i in zip(nets, ids):  # the second call has the smaller input
            free0 = sim.free_bytes()
            got = sim.weights_noalloc(i, net.n_syn)
            free1 = sim.free_bytes()
            want = sim.weights(i, net.n_syn)
            assert free0 == free1, "device memory changed across the no-alloc read"
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "weights differ (net %d)" % i
            part = sim.weights_noalloc(i, 100)
            assert np.array_equal(part.view(np.uint32), want[:100].view(np.uint32))
        sim.step(50)
    sim.close()


def test_noalloc_read_refuses_networks_beyond_the_scratch():
    net, drive, params, _ = W.build("plain-1m", seed=1, wiring={"n_neurons": 180_000,
                                                               "side": W.side_for(180_000)})
    say("no-alloc read: %d synapses (more than the scratch holds)" % net.n_syn)
    assert net.n_syn > (1 << 24)
    sim = nc.Simulation.for_network(net)
    i = sim.add(net, params, drive)
    sim.step(5)
    with pytest.raises(nc.NCError) as e:
        sim.weights_noalloc(i, net.n_syn)
    assert e.value.code == 2  # NC_ERR_RANGE
    assert sim.weights(i, 10).shape == (10,)
    sim.close()
