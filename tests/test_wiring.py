"""The engine's network builder (spatial hash) against the oracle's brute
force restatement of the wiring rule, and its invariants."""

import numpy as np
import pytest

import neurocorrelation_amd as nc
from neurocorrelation_amd import workloads as W


def oracle_spec(O, name, seed):
    return O.wiring(seed=seed, **W.WORKLOADS[name]["wiring"])


@pytest.mark.parametrize("seed", [0, 1, 7])
def test_plain300_wiring_bit_exact(oracle_mod, seed):
    net, drive, _, _ = W.build("plain-300", seed=seed)
    ex = net.export()
    ref = oracle_mod.network(oracle_spec(oracle_mod, "plain-300", seed))
    for k in ("in_ptr", "in_src", "w", "delay", "inh"):
        assert np.array_equal(ex[k], ref[k]), k
    assert np.array_equal(ex["xyz"], ref["xyz"])
    ids, groups = net.select_inputs(drive)
    ids2, groups2 = oracle_mod.select_inputs(300, seed, 6, 4, 2)
    assert np.array_equal(ids, ids2) and np.array_equal(groups, groups2)
    assert sorted(np.bincount(groups).tolist()) == [1, 1, 2, 2]


def test_plain20k_sampled_rows(oracle_mod):
    net, drive, _, _ = W.build("plain-20k", seed=3)
    ex = net.export()
    spec = oracle_spec(oracle_mod, "plain-20k", 3)
    xyz = oracle_mod.positions(spec)
    assert np.array_equal(ex["xyz"], xyz)
    rng = np.random.default_rng(0)
    rows = np.sort(rng.choice(20_000, 60, replace=False)).astype(np.uint32)
    ref = oracle_mod.rows(spec, xyz, rows)
    for q, r in enumerate(rows):
        a, b = ex["in_ptr"][r], ex["in_ptr"][r + 1]
        c, d = ref["in_ptr"][q], ref["in_ptr"][q + 1]
        for k in ("in_src", "w", "delay", "inh"):
            assert np.array_equal(ex[k][a:b], ref[k][c:d]), (r, k)
    ids, groups = net.select_inputs(drive)
    ids2, groups2 = oracle_mod.select_inputs(20_000, 3, 1000, 800, 4)
    assert np.array_equal(ids, ids2) and np.array_equal(groups, groups2)


def test_wiring_invariants():
    n = 5000
    side, r = W.side_for(n), W.radius_for_degree(40)
    net = nc.Network.generate(n, side, r, 1, 32, 0.2, 0.8, 0.25, seed=11)
    ex = net.export()
    ptr, src = ex["in_ptr"], ex["in_src"]
    tgt = np.repeat(np.arange(n), np.diff(ptr))
    # rows strictly ascending in source, no self synapses
    for i in range(0, n, 97):
        row = src[ptr[i]:ptr[i + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0)
        assert i not in row
    assert np.all(src != tgt)
    # symmetric: i->j exists iff j->i exists (both directions are wired)
    fwd = set(zip(src.tolist(), tgt.tolist()))
    assert all((b, a) in fwd for a, b in list(fwd)[:20000])
    # delays and weights in range; inhibitory share near 0.25
    assert ex["delay"].min() >= 1 and ex["delay"].max() <= 32
    a = np.abs(ex["w"])
    assert a.min() >= 0.2 and a.max() < 0.8
    assert np.array_equal(ex["inh"].astype(bool), ex["w"] < 0)
    assert abs(ex["inh"].mean() - 0.25) < 0.02
    # distances below the radius and delay = ceil(32 d / r)
    xyz = ex["xyz"]
    d = np.linalg.norm(xyz[src].astype(np.float64) - xyz[tgt], axis=1)
    assert d.max() < r * (1 + 1e-6)
    assert np.all(np.abs(ex["delay"] - np.clip(np.ceil(32 * d / r), 1, 32)) <= 1)
    # bulk degree near the Poisson mean of the density rule
    xyzc = np.all((xyz > r) & (xyz < side - r), axis=1)
    deg = np.diff(ptr)[xyzc]
    assert abs(deg.mean() - 40) < 2.0


def test_threads_do_not_change_the_network():
    a = nc.Network.generate(6000, W.side_for(6000), 0.3, 1, 20, 0.2, 0.8, 0.1, seed=5, n_threads=1).export()
    b = nc.Network.generate(6000, W.side_for(6000), 0.3, 1, 20, 0.2, 0.8, 0.1, seed=5, n_threads=7).export()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_from_edges_dedupes_like_create_synapse():
    # (to, from): duplicates after the first are ignored (src/NeuCor.cpp:189-191)
    to = [1, 1, 2, 1]
    frm = [0, 0, 0, 2]
    w = [0.5, 0.9, -0.3, 0.2]
    d = [1, 2, 3, 4]
    net = nc.Network.from_edges(3, to, frm, w, d)
    ex = net.export()
    assert ex["in_ptr"].tolist() == [0, 0, 2, 3]
    assert ex["in_src"].tolist() == [0, 2, 0]
    assert ex["w"].tolist() == pytest.approx([0.5, 0.2, -0.3])
    assert ex["delay"].tolist() == [1, 4, 3]
    assert ex["inh"].tolist() == [0, 0, 1]


@pytest.mark.parametrize("bad", [
    dict(to=[3], frm=[0]),      # id out of range
    dict(to=[1], frm=[1]),      # self synapse
])
def test_from_edges_rejects(bad):
    with pytest.raises(nc.NCError):
        nc.Network.from_edges(3, bad["to"], bad["frm"], [0.5], [1])


def test_from_edges_rejects_zero_delay():
    with pytest.raises(nc.NCError):
        nc.Network.from_edges(3, [1], [0], [0.5], [0])


def test_generate_rejects_bad_spec():
    with pytest.raises(nc.NCError):
        nc.Network.generate(0, 1.0, 0.1)
    with pytest.raises(nc.NCError):
        nc.Network.generate(10, 1.0, 0.1, delay_min=3, delay_max=2)
