"""Pins of the C oracle: known-answer vectors, closed forms, the reference's
own outputs (golden fixtures written by the compiled reference) and the
essay's FEW_NEURONS outcome."""
import math

import numpy as np
import pytest

from conftest import golden_files, load_golden

# Random123 known-answer vectors (kat_vectors, philox4x32 10 rounds).
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_kat(oracle_mod, ctr, key, expect):
    assert tuple(int(x) for x in oracle_mod.philox(ctr, key)) == expect


def test_ap_waveform_anchors(oracle_mod):
    # Neuron::AP (src/NeuCor.cpp:706-714) at 0, 1, 2 ms after a fire: values
    # the reference printed (survey probe), to its 8 significant digits.
    for x, v in [(0.0, -54.644085), (1.0, 26.914099), (2.0, -88.914787)]:
        assert oracle_mod.ap_value(x) == pytest.approx(v, abs=5e-6)
    # Closed form of the same waveform in double, rounded to float.
    for x in [0.0, 0.25, 0.5, 1.0, 1.5, 2.0]:
        e1 = math.exp(-((x - 1.0) ** 2) / (2 * 0.3 * 0.3))
        e2 = math.exp(-((x - 2.16) ** 2) / (2 * 0.6 * 0.6))
        v = 100 * (e1 - e2 * 0.2) - 70 + 15 * max(1 - x, 0)
        assert oracle_mod.ap_value(x) == pytest.approx(v, abs=2e-4)


def test_trace_special_cases(oracle_mod):
    assert oracle_mod.trace(0.75, 0.0) == 1.0          # -> treated as 0 by the rule "== 1 -> 0"
    assert oracle_mod.trace(0.75, float("inf")) == 0.0  # lastSpikeArrival = -inf
    assert math.isnan(oracle_mod.trace(0.65, float("nan")))  # lastFire = NaN -> getTrace returns 0
    assert oracle_mod.trace(0.65, 3.0) == pytest.approx(0.65 ** 3, rel=1e-6)


def _micro_net():
    # 0 -> 1 (w 0.5, delay 2 steps), 1 has no other input
    return {"in_ptr": np.array([0, 0, 1], np.int64), "in_src": np.array([0], np.uint32),
            "w": np.array([0.5], np.float32), "delay": np.array([2], np.uint8), "inh": np.zeros(1, np.uint8)}


def test_psp_and_passive_closed_form(oracle_mod):
    """One spike through one synapse: three contributions of
    dt*depol*0.9943*exp(0.3702*dt), each followed by the passive halving."""
    sim = oracle_mod.Sim(_micro_net(), dt=1.0)
    sim.push_forced([1], [0], [0])
    Vs = []
    for _ in range(8):
        sim.step(1)
        Vs.append(float(sim.state()["V"][1]))
    depol = np.float32(np.float32(0.2 * 52.0) * np.float32(0.5))
    term = float(np.float32(1.0) * depol) * 0.9943 * math.exp(0.3702)
    v = -70.0
    expect = []
    for s in range(1, 9):
        if 4 <= s <= 6:  # fired at 1, arrives at 3, contributes at 4, 5, 6
            v = float(np.float32(v + term))
        v = float((np.float32(v) - np.float32(-70.0)) * np.float32(0.5) + np.float32(-70.0))
        expect.append(v)
    assert Vs == expect


def test_stdp_closed_form(oracle_mod):
    """Arrival after the target fired: depression by 0.30*0.65^dt; target
    fire after an arrival: potentiation by 0.13*0.75^dt."""
    net = {"in_ptr": np.array([0, 0, 1], np.int64), "in_src": np.array([0], np.uint32),
           "w": np.array([0.5], np.float32), "delay": np.array([3], np.uint8), "inh": np.zeros(1, np.uint8)}
    sim = oracle_mod.Sim(net, dt=1.0)
    # target 1 forced at step 2; source 0 fires at step 1 -> arrival at 4
    sim.push_forced([2, 1], [1, 0], [0, 0])
    sim.step(4)
    w = float(sim.state()["w"][0])
    expect = np.float32(np.float32(0.5) + np.float32(-(np.float32(0.30) * np.float32(0.65 ** 2))))
    assert w == pytest.approx(float(expect), abs=1e-7)
    # now fire the target at step 6: potentiation 0.13 * 0.75^(6-4)
    sim.push_forced([2], [1], [0])
    sim.step(2)
    w2 = float(sim.state()["w"][0])
    assert w2 == pytest.approx(float(np.float32(w + np.float32(0.13) * np.float32(0.75 ** 2))), abs=1e-7)


def test_weight_clamp_and_sign_class(oracle_mod):
    net = {"in_ptr": np.array([0, 0, 1], np.int64), "in_src": np.array([0], np.uint32),
           "w": np.array([-0.05], np.float32), "delay": np.array([1], np.uint8), "inh": np.ones(1, np.uint8)}
    sim = oracle_mod.Sim(net, dt=1.0)
    # potentiation pushes an inhibitory weight up; it must stop at 0
    for k in range(1, 40, 4):
        sim.push_forced([k, k + 2], [0, 1], [0, 0])
    sim.step(44)
    w = float(sim.state()["w"][0])
    assert -1.0 <= w <= 0.0


def test_latch_drops_spikes_in_flight(oracle_mod):
    """A synapse carries one spike: re-firing the source before the latch
    clears (arrival + 3 steps at dt = 1) is dropped (src/NeuCor.cpp:728)."""
    net = _micro_net()
    sim = oracle_mod.Sim(net, dt=1.0)
    sim.push_forced([1, 2, 3, 6], [0, 0, 0, 0], [0, 0, 0, 0])
    sim.step(10)
    c = sim.counters()
    # fires at 1 (accepted: arrives 3, clears at 6), 2, 3 (dropped), 6 (accepted after release)
    assert c["accepted"] == 2


@pytest.mark.parametrize("path", golden_files())
def test_oracle_matches_reference_fixture(oracle_mod, path):
    """The oracle reproduces the compiled reference's NeuCor::run() bit for
    bit on the committed tie-free scenarios."""
    g = load_golden(path)
    net = {"in_ptr": g["in_ptr"], "in_src": g["in_src"], "w": g["w0"], "delay": g["delay"], "inh": g["inh"]}
    dt = float(g["dt"])
    steps = int(g["steps"])
    sim = oracle_mod.Sim(net, dt=dt)
    sim.push_forced(g["ev_step"], g["ev_id"], g["ev_kind"])
    for k in range(steps):
        _, ids = sim.step(1)
        ref_ids = g["ref_raster_ids"][g["ref_raster_ptr"][k]:g["ref_raster_ptr"][k + 1]]
        assert np.array_equal(ids, ref_ids), "raster differs at step %d" % (k + 1)
        V = sim.state()["V"]
        assert np.array_equal(V.view(np.uint32), g["ref_V"][k].view(np.uint32)), "V differs at step %d" % (k + 1)
    st = sim.state()
    assert np.array_equal(st["w"].view(np.uint32), g["ref_w"].view(np.uint32))
    assert np.array_equal(st["fire_t"], g["ref_fire_t"])
    assert np.array_equal(st["last_arr"], g["ref_last_arr"])
    assert np.array_equal(st["last_fire"], g["ref_last_fire"], equal_nan=True)
    assert sim.counters()["first_tie"] < 0


def test_few_neurons_essay_outcome(oracle_mod):
    """FEW_NEURONS (src/main.cpp:162-189): 0 -> 1 and 0 -> 2 at w = 0.5,
    inputs at 50 Hz with neuron 1 driven 2 ms after 0 and neuron 2 2 ms
    before.  The essay reports w(0->1) -> 1 and w(0->2) -> 0 within ~160 ms.
    Delay 2 * 0.4243 ms = 0.85 ms, i.e. 54 steps of dt = 1/64 ms."""
    dt = 1.0 / 64
    net = {"in_ptr": np.array([0, 0, 1, 2], np.int64), "in_src": np.array([0, 0], np.uint32),
           "w": np.array([0.5, 0.5], np.float32), "delay": np.array([54, 54], np.uint8),
           "inh": np.zeros(2, np.uint8)}
    sim = oracle_mod.Sim(net, dt=dt)
    per = 20 * 64  # 20 ms
    so, ids = [], []
    for k in range(1, 12):
        so += [k * per, k * per + 2 * 64, k * per - 2 * 64]
        ids += [0, 1, 2]
    sim.push_forced(so, ids, [0] * len(ids))
    sim.step(12 * per)
    w = sim.state()["w"]
    assert w[0] == pytest.approx(1.0)
    assert w[1] == pytest.approx(0.0)
