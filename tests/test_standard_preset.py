"""CPU: why no fixed-step engine, this one included, can match the
reference's STANDARD preset (src/main.cpp:80-124: NeuCor(750), runAll =
false, runSpeed = 4) to 2% in firing rate and weight histogram, and one
rule of the preset's inputs that the facade must (and does) keep.

1. With runAll = false the reference runs a neuron only at its own events
   (an arrival, the re-run 2 ms after it, a scheduled background fire;
   src/NeuCor.cpp:583-617, :663-666, :718-726): between events its potential
   is frozen, and at the next event charge_insynapses and charge_passive
   (:677-700) take one step of dt = the whole time since its last run, over
   which a still-latched synapse is integrated as deltaT * depol * 0.9943 *
   exp(0.3702 deltaT).  A fixed-step engine updates every neuron every step
   (the runAll = true order, which the engine reproduces bit for bit when
   every event falls on its grid); the trajectories differ by construction.
2. Even with runAll = true, NeuCor(n)'s delays (2 * length ms) and the
   background fires' offsets (randomUnit() * runSpeed, :606) fall between
   grid times, and an arrival runs its target at the arrival time too: the
   PSP of a synapse depends on the sub-step phase of its arrival
   (deltaT * exp(0.3702 deltaT) is not additive over sub-steps).  An engine
   on a grid gives every spike the on-grid PSP.
3. The preset's outcome depends on the C library's rand() stream draw for
   draw: shifting it by one to four draws after construction (the
   background fires of :604-607 draw from it) moves the reference's own
   firing count by more than 2%.  A 2% bar against one reference run would
   need the stream replayed exactly; the reference also draws from it inside
   STDP whenever an excitatory weight sits at exactly 0 (:752, whose
   `&& false` comes after the draw), which a GPU step does not reproduce.
4. InputFirer::schedule (:333-345) drops a fire time that falls exactly on
   a frame boundary and then never advances: at 1 ms frames a 40 Hz input
   (25 ms period) never fires.  The facade restates the schedule with the
   same comparisons; tests/test_facade.py checks that it keeps this rule.

tests/golden/standard_ref.json (written by tests/golden/make_standard.py)
holds the reference's preset runs the statistical tests below use.
"""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
FIXTURE = os.path.join(ROOT, "tests", "golden", "standard_ref.json")


def _harness():
    if not os.path.exists(HARNESS):
        pytest.skip("oracle/_ref/ref_harness is not built (needs the reference checkout)")
    return HARNESS


def _gap(run_all):
    out = subprocess.run([_harness(), "--gap", str(run_all)], capture_output=True, text=True, check=True,
                         timeout=60).stdout
    rows = {}
    for line in out.splitlines():
        t, v, _ = line.split()
        rows[int(float(t))] = float.fromhex(v)
    return rows


def _f32(x):
    return float(np.float32(x))


def _step(V, deltaT, depol):
    """One Neuron::run of the reference with one latched synapse and no
    fire: charge_insynapses then charge_passive (src/NeuCor.cpp:677-700),
    float state with double terms as the reference evaluates them."""
    term = float(np.float32(deltaT) * np.float32(depol)) * 0.9943 * np.exp(0.3702 * deltaT)
    V = _f32(V + term)
    return _f32(_f32(_f32(V) - np.float32(-70.0)) * np.float32(0.5) ** np.float32(deltaT) + np.float32(-70.0))


def test_event_mode_runs_neurons_only_at_their_events():
    clocked, event = _gap(1), _gap(0)
    depol = _f32(np.float32(_f32(np.float32(0.2) * 52.0)) * np.float32(0.2))  # Synapse::fire (:730-732)
    # event mode: neuron 2 runs at the arrival (5 ms; the arriving synapse
    # itself does not count yet) and 2 ms later, one step of dt = 2
    v7 = _step(-70.0, 2.0, depol)
    assert event[5] == -70.0 and abs(event[7] - v7) < 1e-4
    # then nothing runs it until the next arrival: frozen, not decaying
    assert all(event[t] == event[7] for t in range(8, 44))
    # the next arrival (44 ms) runs it with dt = 37 over the synapse still
    # latched since 5 ms: exp(0.3702 * 37) ~ 9e5 times the per-ms term, times
    # 0.5^37 of passive decay
    assert abs(event[44] - _step(event[7], 37.0, depol)) < 1e-3
    # clocked: every neuron every step; the first PSP has decayed to rest
    # long before the second arrives, and both PSPs have the same shape
    assert abs(clocked[20] + 70.0) < 0.02 and clocked[43] == clocked[42]
    assert max(clocked[t] for t in range(4, 14)) > -68.0
    peak1 = max(clocked[t] for t in range(4, 14))
    peak2 = max(clocked[t] for t in range(43, 53))
    assert abs(peak1 - peak2) < 1e-3
    # the two orders give different trajectories
    assert max(abs(clocked[t] - event[t]) for t in range(1, 61)) > 1.0


def _psp(delay):
    out = subprocess.run([_harness(), "--psp", repr(delay)], capture_output=True, text=True, check=True,
                         timeout=60).stdout
    return [float.fromhex(l.split()[1]) for l in out.splitlines()]


def test_psp_depends_on_the_arrival_phase():
    """With runAll = true and runSpeed = 1 ms, a spike arriving between grid
    times runs its target at the arrival and 2 ms later as well (steps of
    fractional dt, over which deltaT * depol * 0.9943 * exp(0.3702 deltaT) is
    not additive), so the same synapse gives a different PSP depending on the
    fractional part of its delay (2 * length ms, continuous in NeuCor(n)).
    A step engine that rounds delays to its grid gives every synapse the
    on-grid PSP."""
    on = [_psp(d) for d in (1.0, 2.0)]
    off = [_psp(d) for d in (0.5, 0.7, 0.9, 1.5)]
    area = lambda v: sum(x + 70.0 for x in v)  # noqa: E731
    # on the grid, the PSP does not depend on the delay (shifted only)
    assert abs(area(on[0]) - area(on[1])) / area(on[0]) < 0.01
    assert max(on[0]) == max(on[1])
    # off the grid it depends on the phase, and is smaller
    for v in off:
        assert area(v) < 0.9 * area(on[0])
    assert max(area(v) for v in off) - min(area(v) for v in off) > 0.1 * area(on[0])


def _fixture():
    with open(FIXTURE) as f:
        return json.load(f)


def _fires(doc, mode):
    return np.array([r["fires"] for r in doc["runs"] if r["mode"] == mode], dtype=float)


def test_preset_outcome_moves_with_the_rand_stream():
    doc = _fixture()
    for mode in ("event", "clocked"):
        fires = _fires(doc, mode)
        assert len(fires) == 5
        # the reference against itself, one to four draws apart
        assert np.max(np.abs(fires[1:] / fires[0] - 1.0)) > 0.02, (mode, fires)


def test_event_and_clocked_orders_are_different_models():
    """Over the same five rand() shifts and 4 ms frames, the preset's event
    order (runAll = false) and the clocked order (runAll = true) end with
    mean weights 7% apart, more than four standard errors of the difference
    of the means and more than the 2% bar."""
    doc = _fixture()
    ev = np.array([r["mean_w"] for r in doc["runs"] if r["mode"] == "event"])
    cl = np.array([r["mean_w"] for r in doc["runs"] if r["mode"] == "clocked"])
    se = np.sqrt(ev.var(ddof=1) / len(ev) + cl.var(ddof=1) / len(cl))
    assert abs(cl.mean() - ev.mean()) > 4 * se, (ev, cl)
    assert abs(cl.mean() / ev.mean() - 1.0) > 0.02


def test_fixture_is_what_the_reference_gives():
    """The fixture's first event-mode run, recomputed by the compiled
    reference (needs oracle/_ref/ref_harness)."""
    doc = _fixture()
    m = doc["modes"]["event"]
    out = subprocess.run([_harness(), "--standard", str(doc["seed"]), str(doc["ms"]), str(m["run_speed"]),
                          str(m["run_all"]), str(m["walk_every"]), str(doc["rate_hz"]), "0"],
                         capture_output=True, text=True, check=True, timeout=600).stdout
    got = {l.split()[0]: l.split()[1:] for l in out.splitlines()}
    want = [r for r in doc["runs"] if r["mode"] == "event" and r["skip"] == 0][0]
    assert int(got["fires"][0]) == want["fires"]
    assert [int(x) for x in got["w_hist"]] == want["w_hist"]


def _standard(run_speed, run_all, rate, ms=300):
    out = subprocess.run([_harness(), "--standard", "1", str(ms), str(run_speed), str(run_all), "1", repr(rate),
                          "0"], capture_output=True, text=True, check=True, timeout=600).stdout
    return {l.split()[0]: l.split()[1:] for l in out.splitlines()}


def test_inputs_on_frame_boundaries_never_fire():
    """InputFirer::schedule (src/NeuCor.cpp:333-345) queues fire times t with
    currentT < t < currentT + runSpeed and only then advances lastFire: with
    1 ms frames every multiple of a 25 ms period is a frame boundary, so a
    40 Hz input never fires (the run equals one with no input); with 4 ms
    frames (24 < 25 < 28) it does."""
    none = 1e-9  # a period longer than the run
    a, b = _standard(1, 1, 40.0), _standard(1, 1, none)
    assert a == b
    c, d = _standard(4, 1, 40.0), _standard(4, 1, none)
    assert c["fires"] != d["fires"]
