"""CPU: the workload generators behind bench.py (sizes, degrees, step budgets
and drives of the batched and served specs; deterministic per seed), the
benchmark's algorithmic byte model of the neuron-step kernel, and its refusal
of more than one GPU."""
import importlib.util
import math
import os

import numpy as np

from neurocorrelation_amd import serve as SV
from neurocorrelation_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_batched_specs_ranges_and_determinism():
    a = W.batched_specs(count=200, seed=3)
    b = W.batched_specs(count=200, seed=3)
    assert a == b
    ns = np.array([s["wiring"]["n_neurons"] for s in a])
    ks = np.array([s["k"] for s in a])
    assert ns.min() >= 200 and ns.max() <= 20_000
    assert ks.min() >= 10 and ks.max() <= 60
    # log-uniform sizes: the median sits near the geometric mean of the range
    assert abs(math.log(np.median(ns)) - 0.5 * (math.log(200) + math.log(20_000))) < 0.6
    for s in a:
        d, w = s["drive"], s["wiring"]
        assert d["n_inputs"] >= 6 and d["n_correlated"] % 2 == 0 and d["n_correlated"] <= d["n_inputs"]
        assert 0.0 <= d["amp_corr"] <= 1.0 and d["seed"] == w["seed"]
        assert w["side"] == W.side_for(w["n_neurons"]) and s["steps"] == 20_000


def test_served_specs_ranges_and_weights():
    a = SV.served_specs(300, seed=3)
    assert a == SV.served_specs(300, seed=3)
    ns = np.array([s["wiring"]["n_neurons"] for s in a])
    st = np.array([s["steps"] for s in a])
    assert ns.min() >= 300 and ns.max() <= 100_000
    assert st.min() >= 1_000 and st.max() <= 50_000
    assert all(s["wiring"]["inhib_frac"] == 0.2 and s["wiring"]["w_max"] == 0.3 for s in a)
    run = SV.served_specs(5, seed=1, weights=(0.2, 0.8, 0.0))
    assert all(s["wiring"]["inhib_frac"] == 0.0 and s["wiring"]["w_min"] == 0.2 for s in run)


def test_radius_gives_the_bulk_degree():
    # density x sphere volume at the radius is the wanted out-degree
    for k in (10, 50, 100):
        r = W.radius_for_degree(k)
        assert abs(W.DENSITY * 4.0 / 3.0 * math.pi * r ** 3 - k) < 1e-9


def test_neuron_kernel_byte_model():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    # bytes per unit, as the docstrings' tables state them
    # (an active slot read from its record instead of the cache: 16 - 8)
    per_unit = {"active_slot_steps": 8, "records_read": 8, "bucket_entries": 8, "arrival_records_read": 16,
                "arrival_weight_writes": 4, "cache_records_written": 16, "fired": 12, "fired_in_degree": 16,
                "post_weight_writes": 4, "fired_records": 16, "pending_scanned": 4, "mask_words_read": 4,
                "mask_words_written": 4, "arrivals": 0}
    zero = {k: 0 for k in per_unit}
    base, _ = bench.neuron_kernel_bytes(640, zero, zero, 10)
    assert base == 20 * 640 + 32 * 10  # per neuron and per 64-neuron slice
    for k, bpu in per_unit.items():
        c1 = dict(zero, **{k: 30})  # 3 units per step over 10 steps
        b, ev = bench.neuron_kernel_bytes(640, zero, c1, 10)
        assert abs(b - base - 3 * bpu) < 1e-9, k
    fan = {"fired_records": 16, "fired": 8, "fired_out_degree": 16, "accepted_spikes": 44}
    fz = {k: 0 for k in fan}
    assert bench.fanout_kernel_bytes(fz, fz, 10)[0] == 0
    for k, bpu in fan.items():
        b, _ = bench.fanout_kernel_bytes(fz, dict(fz, **{k: 50}), 10)
        assert abs(b - 5 * bpu) < 1e-9, k


def test_bench_measures_one_gpu_only():
    """bench.py --gpus N with N > 1 (or a launcher's WORLD_SIZE > 1) stops
    with a message, before touching a GPU."""
    import subprocess
    import sys
    for extra_env, args in (({}, ["--gpus", "2"]), ({"WORLD_SIZE": "2"}, [])):
        env = dict(os.environ, **extra_env)
        p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, env=env,
                           capture_output=True, text=True, timeout=120)
        assert p.returncode == 2, p.stderr
        assert "one GPU" in p.stderr and p.stdout == ""


def test_bench_tuned_shapes_and_event_spacing():
    """The benchmark's table of tuned launch shapes names real plain
    workloads and only options that Simulation.set_launch takes; kernel
    events ride on at least 5 launches per kernel and on at most every 10th
    step."""
    import inspect
    from neurocorrelation_amd.engine import Simulation
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    opts = set(inspect.signature(Simulation.set_launch).parameters) - {"self"}
    for name, shape in bench.TUNED_LAUNCH.items():
        assert name in W.WORKLOADS, name
        assert set(shape) <= opts, (name, shape)
        assert shape.get("waves_per_block", 4) in (1, 2, 4, 8)
    for steps, want in ((20, 4), (5, 1), (3, 1), (49, 9), (50, 10), (200, 10), (10_000, 10)):
        every = bench.timing_every(steps)
        assert every == want, (steps, every)
        assert steps < 5 or (steps + every - 1) // every >= 5
