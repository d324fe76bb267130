"""GPU: the workload wiring built on the device (nc_net_generate_device) is
the host builder's network bit for bit (positions, rows in ascending source
id, delays, weights, sign classes); the host builder is pinned to the
oracle's brute-force restatement in test_wiring.py."""
import sys
import time

import numpy as np
import pytest

import neurocorrelation_amd as nc
from neurocorrelation_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _same(a, b):
    ea, eb = a.export(), b.export()
    for k in ("xyz", "in_ptr", "in_src", "w", "delay", "inh"):
        assert ea[k].shape == eb[k].shape, k
        if ea[k].dtype == np.float32:
            assert np.array_equal(ea[k].view(np.uint32), eb[k].view(np.uint32)), k
        else:
            assert np.array_equal(ea[k], eb[k]), k
    assert a.info() == b.info()


CASES = [
    ("plain-300 seed 0", dict(W.WORKLOADS["plain-300"]["wiring"], seed=0)),
    ("plain-300 seed 7, draw order", dict(W.WORKLOADS["plain-300"]["wiring"], seed=7, spatial_order=0)),
    ("plain-20k seed 3", dict(W.WORKLOADS["plain-20k"]["wiring"], seed=3)),
    ("served-like, 20% inhibitory", dict(n_neurons=12_345, side=W.side_for(12_345), radius=W.radius_for_degree(77),
                                         delay_min=1, delay_max=32, w_min=0.05, w_max=0.3, inhib_frac=0.2, seed=99)),
    ("one cell per axis", dict(n_neurons=500, side=1.0, radius=0.9, delay_min=2, delay_max=9, seed=5)),
]


@pytest.mark.parametrize("name,spec", CASES, ids=[c[0] for c in CASES])
def test_device_wiring_equals_host(name, spec):
    say("wiring", name)
    host = nc.Network.generate(**spec)
    dev = nc.Network.generate(device=0, **spec)
    _same(host, dev)


def test_device_wiring_plain_1m():
    spec = dict(W.WORKLOADS["plain-1m"]["wiring"], seed=0)
    say("wiring plain-1m on the host")
    t = time.perf_counter()
    host = nc.Network.generate(**spec)
    th = time.perf_counter() - t
    say("wiring plain-1m on the device")
    t = time.perf_counter()
    dev = nc.Network.generate(device=0, **spec)
    td = time.perf_counter() - t
    say("host %.2fs, device %.2fs, %d synapses" % (th, td, host.n_syn))
    _same(host, dev)
