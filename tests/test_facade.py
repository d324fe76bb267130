"""class NeuCor over the engine (neurocorrelation_amd/facade).

The facade's constructor draws the C library's rand() like the reference's
(src/NeuCor.cpp:17-42, :154-195, :418-444, :463-486): after the same srand()
it must build the same network, checked against the reference's own
NeuCor(n) through the reference harness.  On the GPU, the FEW_NEURONS preset
(src/main.cpp:162-189) run through the facade must equal the oracle run on
the same network with the same input and background fires, which the test
derives independently: InputFirer::schedule (src/NeuCor.cpp:333-345) and the
background loop (:604-607) restated in float32, with the C library's rand()
called in the facade's order.
"""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "neurocorrelation_amd", "facade", "facade_check")


def _check_bin():
    if not os.path.exists(CHECK):
        pytest.fail("facade_check is not built (neurocorrelation_amd/build.py)")
    return CHECK


def _parse_wiring(text):
    pos, syn = [], []
    for line in text.splitlines():
        f = line.split()
        if f[0] == "p":
            pos.append([float.fromhex(x) for x in f[1:]])
        elif f[0] == "s":
            syn.append((int(f[1]), int(f[2]), float.fromhex(f[3]), float.fromhex(f[4])))
    return np.array(pos, np.float32), syn


@pytest.mark.parametrize("n,seed", [(40, 1), (300, 0), (750, 5)])
def test_constructor_matches_reference(n, seed):
    from oracle import ref as R
    if not R.available():
        pytest.skip("oracle/_ref/ref_harness is not built (needs the reference checkout)")
    mine = subprocess.run([_check_bin(), "wiring", str(n), str(seed)], capture_output=True, text=True, check=True)
    ref = subprocess.run([R.REF_HARNESS, "--wiring", str(n), str(seed)], capture_output=True, text=True,
                         check=True)
    assert mine.stdout == ref.stdout


def test_constructor_invariants():
    out = subprocess.run([_check_bin(), "wiring", "200", "3"], capture_output=True, text=True, check=True).stdout
    pos, syn = _parse_wiring(out)
    assert len(pos) == 200
    # a sphere of 8 neurons per unit volume: radius (200 / (4/3 pi 8))^(1/3)
    r = (200 / (1.3333 * 3.1459 * 8)) ** 0.33333
    assert np.all(np.sqrt((pos.astype(np.float64) ** 2).sum(1)) <= r + 1e-6)
    pairs = set()
    for a, b, w, length in syn:
        assert a != b and length < 1.0
        assert 0.2 <= abs(w) <= 1.0
        d = np.sqrt(((pos[a].astype(np.float64) - pos[b]) ** 2).sum())
        assert abs(d - length) < 1e-5
        pairs.add((a, b))
    assert len(pairs) == len(syn)
    assert all((b, a) in pairs for a, b in pairs)  # every neighbour pair, both directions


def _few_schedule(runs, dt):
    """Forced fires (1-based step, id, kind) of the FEW_NEURONS facade run."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)  # the C library's stream before any srand() call
    for _ in range(2 * 2):  # createSynapse draws a weight and a sign per synapse (src/NeuCor.cpp:471-473)
        libc.rand()
    RAND_MAX = 2147483647
    f32 = np.float32
    cur, rs = f32(0.0), f32(0.02)
    freq = f32(50.0)
    last = [f32(0.0), f32(2.0), f32(-2.0)]  # lastFire after addInputOffset(1, 2), (2, -2)
    ev = []
    period = max(1, int(f32(600.0) / rs))
    for _ in range(runs):
        for i in range(3):
            ft = f32(np.float64(last[i]) + 1000.0 / np.float64(freq))
            while ft < f32(cur + rs):
                if cur < ft:
                    ev.append((int(math.ceil(np.float64(ft) / dt)), i, 0))
                    last[i] = ft
                ft = f32(np.float64(ft) + 1000.0 / np.float64(freq))
        for _i in range(3):
            if libc.rand() % period == 0:
                nid = libc.rand() % 3
                off = f32(f32(libc.rand()) / f32(RAND_MAX)) * rs
                ev.append((int(math.ceil((np.float64(cur) + np.float64(f32(off))) / dt)), nid, 1))
        cur = f32(cur + rs)
    return ev, cur, libc.rand()


@pytest.mark.gpu
def test_few_neurons_through_the_facade(oracle_mod):
    runs = 12_000  # 240 ms at runSpeed 0.02
    print("facade_check few %d" % runs, flush=True)
    out = subprocess.run([_check_bin(), "few", str(runs)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    t = dt = r_next = None
    w, v, mir, lf = {}, {}, {}, {}
    for line in out.stdout.splitlines():
        f = line.split()
        if f[0] == "m":
            mir[(int(f[1]), int(f[2]))] = [np.float32(float.fromhex(x)) for x in f[3:6]]
        elif f[0] == "l":
            lf[int(f[1])] = (np.float32(float.fromhex(f[2])), np.float32(float.fromhex(f[3])))
        elif f[0] == "t":
            t = float.fromhex(f[1])
        elif f[0] == "dt":
            dt = float.fromhex(f[1])
        elif f[0] == "w":
            w[(int(f[1]), int(f[2]))] = np.float32(float.fromhex(f[3]))
        elif f[0] == "v":
            v[int(f[1])] = np.float32(float.fromhex(f[2]))
        elif f[0] == "r":
            r_next = int(f[1])
    assert dt == 1.0 / 64
    ev, t_expect, r_expect = _few_schedule(runs, dt)
    assert np.float32(t) == t_expect
    # the engine left the caller's rand() stream where the reference would
    assert r_next == r_expect
    steps = int(math.floor(t / dt))

    # the network: 0 -> 1, 0 -> 2 at w = 0.5; delay 2 * length ms in steps
    p = np.array([[0, 0, 0], [0.3, 0.3, 0], [0.3, -0.3, 0]], np.float32)
    length = [np.sqrt(np.float32(((p[0] - p[k]) ** 2).sum(dtype=np.float32))) for k in (1, 2)]
    delay = [int(round(2.0 * float(L) / dt)) for L in length]
    net = {"in_ptr": np.array([0, 0, 1, 2], np.int64), "in_src": np.array([0, 0], np.uint32),
           "w": np.array([0.5, 0.5], np.float32), "delay": np.array(delay, np.uint8), "inh": np.zeros(2, np.uint8)}
    sim = oracle_mod.Sim(net, dt=dt)
    ev = [e for e in ev if e[0] <= steps]
    sim.push_forced([e[0] for e in ev], [e[1] for e in ev], [e[2] for e in ev])
    sim.step(steps)
    st = sim.state()
    assert w[(0, 1)] == st["w"][0] and w[(0, 2)] == st["w"][1]
    assert all(v[i] == st["V"][i] for i in range(3))
    # the Neuron / Synapse mirror the renderer reads equals the oracle too:
    # weight, latched arrival (AP_fireTime, 0 = none), last arrival, and each
    # neuron's last fire and potential
    for k, key in enumerate([(0, 1), (0, 2)]):
        mw, mft, mla = mir[key]
        assert mw == st["w"][k] and mft == st["fire_t"][k] and mla == st["last_arr"][k], (key, mir[key])
    for i in range(3):
        assert (lf[i][0] == st["last_fire"][i]) or (np.isnan(lf[i][0]) and np.isnan(st["last_fire"][i]))
        assert lf[i][1] == st["V"][i]
    # the essay's outcome: the synapse whose target is driven after its
    # source strengthens to 1, the one driven before weakens to about 0
    # (each new pre-before-post pairing and background fire moves it a little)
    assert w[(0, 1)] == pytest.approx(1.0) and 0.0 <= w[(0, 2)] < 0.05


@pytest.mark.gpu
def test_renderer_mirror_matches_snapshots():
    """The Neuron / Synapse objects the renderer reads through its friend
    access (src/NeuCor_Renderer.cpp:655-699, :840-866) agree with the
    snapshots after runs, and Synapse::setWeight on the mirror reaches the
    engine."""
    print("facade_check mirror 300", flush=True)
    out = subprocess.run([_check_bin(), "mirror", "300"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    f = out.stdout.split()
    vals = dict(zip(f[0::2], f[1::2]))
    assert int(vals["synapses"]) > 0 and int(vals["mismatches"]) == 0
    assert int(vals["fired"]) > 0
    # the written weight, possibly moved by plasticity in the one run after it,
    # keeps the inhibitory sign class setWeight gave it
    assert -1.0 <= float.fromhex(vals["weight_after_set"]) <= 0.0


def test_renderer_accesses_compile_against_the_facade():
    """Every expression the reference's renderer (src/NeuCor_Renderer.cpp)
    and presets (src/main.cpp) apply to the simulation core, repeated in
    neurocorrelation_amd/facade/renderer_access.cpp (same member names, same
    result types, cited line by line), compiles against the facade's header.
    The renderer's own file needs OpenGL, GLFW, ImGui and glm, which this
    build does not have."""
    src = os.path.join(ROOT, "neurocorrelation_amd", "facade", "renderer_access.cpp")
    inc = os.path.join(ROOT, "neurocorrelation_amd", "facade")
    for extra in ([], ["-DNC_RENDERER_ACCESS_MAIN"]):
        r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", inc, src] + extra,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_renderer_accesses_run_on_the_facade():
    """The same accesses run on NeuCor(60) for five frames (alternating
    wall-clock-scaled and fixed run speeds)."""
    demo = os.path.join(ROOT, "neurocorrelation_amd", "facade", "renderer_access_demo")
    if not os.path.exists(demo):
        pytest.fail("renderer_access_demo is not built (neurocorrelation_amd/build.py)")
    print("renderer_access_demo", flush=True)
    out = subprocess.run([demo], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "renderer access ok 1" in out.stdout


@pytest.mark.gpu
def test_network_grown_after_running_keeps_its_state():
    """createNeuron / createSynapse after run() (which the reference allows at
    any time) rebuild the engine's network; the existing neurons keep their
    potential and last fire, the existing synapses their weight, latched
    spike and last arrival, as the reference's objects do, and the new
    synapse starts with none."""
    print("facade_check grow 200", flush=True)
    out = subprocess.run([_check_bin(), "grow", "200"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    sect, lines = None, {"before": [], "after": []}
    for line in out.stdout.splitlines():
        if line in ("before", "after"):
            sect = line
        elif line[0] in "ml" and sect:
            lines[sect].append(line)
    before, after = set(lines["before"]), lines["after"]
    new = [x for x in after if x not in before]
    assert len(lines["before"]) > 100 and len(after) == len(lines["before"]) + 2
    # the new neuron's line and the new synapse 0 -> 120 (no spike: AP_fireTime 0, last arrival -inf)
    assert sorted(new) == sorted(["m 0 120 0x1p-1 0x0p+0 -inf", "l 120 nan -0x1.18p+6"]), new
    assert any(x.startswith("m ") and x.split()[4] != "0x0p+0" for x in lines["before"])  # spikes were in flight


@pytest.mark.gpu
def test_standard_preset_inputs_keep_the_frame_boundary_rule():
    """The STANDARD preset (src/main.cpp:80-124, inputs held at 40 Hz)
    through the facade: InputFirer::schedule (src/NeuCor.cpp:333-345) drops a
    fire time on a frame boundary, so with 1 ms frames a 40 Hz input never
    fires and the run equals one with no input, bit for bit; with 4 ms
    frames it fires.  The reference does the same
    (tests/test_standard_preset.py); why the preset's statistics cannot be
    matched to 2% is stated and tested there."""
    def run(rs, rate):
        print("facade_check standard", rs, rate, flush=True)
        out = subprocess.run([_check_bin(), "standard", "1", "300", str(rs), "1", repr(rate), "0"],
                             capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stdout + out.stderr
        return {l.split()[0]: l.split()[1:] for l in out.stdout.splitlines()}
    none = 1e-9
    a, b = run(1, 40.0), run(1, none)
    assert a == b and int(a["fires"][0]) > 0
    c, d = run(4, 40.0), run(4, none)
    assert c["fires"] != d["fires"]
    # weights stay in their sign classes' ranges
    assert -1.0 <= float(c["mean_w"][0]) <= 1.0
