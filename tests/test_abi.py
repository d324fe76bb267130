"""The C-ABI library loads and exports every entry point include/neucor_amd.h
declares; host-only calls work without a GPU."""
import ctypes
import os
import re

import numpy as np

import neurocorrelation_amd as nc
from neurocorrelation_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "neucor_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nc_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name


def test_binding_covers_the_header():
    assert set(declared()) == set(_lib.SIGNATURES)


def test_host_only_calls():
    net = nc.Network.from_edges(4, [1, 2, 3], [0, 0, 0], [0.5, 0.6, 0.7], [1, 2, 3])
    info = net.info()
    assert info == {"n": 4, "n_syn": 3, "max_in_degree": 1, "max_delay": 3}
    npos, nslots = net.device_size()
    assert npos == 64 and nslots % 256 == 0 and nslots >= 3
    # the kept device image is host work and leaves the layout as it was
    assert net.prepare() is net and net.prepare().device_size() == (npos, nslots)
    assert nc.footprint(1000, 50_000) > 50_000 * 17
    assert nc.lib().nc_last_error() is not None


def _net_with_in_degrees(deg, seed=0):
    """A network whose neuron i has deg[i] incoming synapses (random sources)."""
    rng = np.random.default_rng(seed)
    n = len(deg)
    to = np.repeat(np.arange(n, dtype=np.uint32), deg)
    # distinct sources per target, none the target itself
    frm = np.concatenate([(i + 1 + rng.choice(n - 1, size=int(d), replace=False)) % n
                          for i, d in enumerate(deg)]).astype(np.uint32)
    w = rng.uniform(0.1, 0.5, size=to.size).astype(np.float32)
    return nc.Network.from_edges(n, to, frm, w, np.ones(to.size, dtype=np.uint8))


def test_slot_layout_accounting():
    """In-slots are stored in chunks of 8 per lane (one 128-byte line of slot
    records) across the 64 lanes of a slice: a slice takes 512 slots per
    chunk of its largest in-degree."""
    for d in (1, 7, 8, 9, 97):
        npos, nslots = _net_with_in_degrees(np.full(300, d)).device_size()
        assert npos == 320  # 5 slices of 64 positions
        assert nslots == 5 * 512 * ((d + 7) // 8), d
    # mixed degrees: every slice is whole chunks, and no lane gets less than
    # its own in-degree rounded up to a chunk
    deg = np.random.default_rng(1).integers(0, 40, size=1000)
    net = _net_with_in_degrees(deg, seed=2)
    assert net.info()["n_syn"] == int(deg.sum())
    npos, nslots = net.device_size()
    assert npos == 1024 and nslots % 512 == 0
    assert nslots >= int(((deg + 7) // 8 * 8).sum())
    assert nslots <= 16 * 512 * ((int(deg.max()) + 7) // 8)


def test_errors_cross_the_boundary_as_codes():
    L = nc.lib()
    assert L.nc_net_info(None, None, None, None, None) == 1  # NC_ERR_ARG
    assert b"null" in L.nc_last_error()
    assert L.nc_sim_step(None, 1) == 1
    assert L.nc_net_prepare(None) == 1
    assert L.nc_net_generate_device(None, 0, None) == 1
