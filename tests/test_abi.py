"""The C-ABI library loads and exports every entry point include/neucor_amd.h
declares; host-only calls work without a GPU."""
import ctypes
import os
import re

import pytest

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


def test_errors_cross_the_boundary_as_codes():
    L = nc.lib()
    assert L.nc_net_info(None, None, None, None, None) == 1  # NC_ERR_ARG
    assert b"null" in L.nc_last_error()
    assert L.nc_sim_step(None, 1) == 1
    assert L.nc_net_prepare(None) == 1
    assert L.nc_net_generate_device(None, 0, None) == 1
