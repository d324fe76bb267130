"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Run in the build container (it needs oracle/_ref/ref_harness, which is built
from the reference's own src/NeuCor.cpp by oracle/Makefile):

    python tests/golden/make_golden.py

Each fixture is a micro-network (2 input neurons without incoming synapses,
4-6 other neurons, random weights of both sign classes, delays in steps),
per-step forced input and background fires, and what the reference's
NeuCor::run() produced on it: the raster and every potential after every
step, and the final weights, spike latches, last arrivals and last fires.
Only scenarios whose outcome does not depend on the reference's event-queue
tie order are kept (the oracle reports such ties; none may occur).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from oracle import ref as R  # noqa: E402


def micro(seed, steps, n=6, n_in=2, p_edge=0.5, p_in=0.04, p_bg=0.006, dmax=6):
    rng = np.random.default_rng(seed)
    edges = []
    for tg in range(n_in, n):
        for src in range(n):
            if src != tg and rng.random() < p_edge:
                edges.append((tg, src))
    in_ptr = np.zeros(n + 1, np.int64)
    for tg, _ in edges:
        in_ptr[tg + 1] += 1
    in_ptr = np.cumsum(in_ptr)
    src = np.array([s for _, s in edges], np.uint32)
    w = rng.uniform(0.2, 1.0, len(edges)).astype(np.float32)
    neg = rng.random(len(edges)) < 0.2
    w[neg] *= -1
    d = rng.integers(1, dmax + 1, len(edges)).astype(np.uint8)
    net = {"in_ptr": in_ptr, "in_src": src, "w": w, "delay": d, "inh": neg.astype(np.uint8)}
    inputs = [[i for i in range(n_in) if rng.random() < p_in] for _ in range(steps)]
    bg = [[i for i in range(n_in, n) if rng.random() < p_bg] for _ in range(steps)]
    return net, inputs, bg


def events(inputs, bg):
    so, ids, kd = [], [], []
    for k in range(len(inputs)):
        for i in inputs[k]:
            so.append(k + 1); ids.append(i); kd.append(0)
        for i in bg[k]:
            so.append(k + 1); ids.append(i); kd.append(1)
    return np.array(so, np.uint32), np.array(ids, np.uint32), np.array(kd, np.uint8)


def tie_free(net, inputs, bg, steps, dt):
    sim = O.Sim(net, dt=dt)
    sim.push_forced(*events(inputs, bg))
    sim.step(steps)
    c = sim.counters()
    return c["first_tie"] < 0 and c["fires"] > 10


def main(n_fixtures=12, steps=240):
    if not R.available():
        raise SystemExit("oracle/_ref/ref_harness is not built (make -C oracle)")
    made = 0
    seed = 0
    while made < n_fixtures:
        dt = [1.0, 0.5, 0.25][made % 3]
        seed += 1
        net, inputs, bg = micro(seed, steps)
        if len(net["in_src"]) == 0 or not tie_free(net, inputs, bg, steps, dt):
            continue
        ref = R.run_trace(net, steps, dt=dt, inputs=inputs, bg=bg)
        so, ids, kd = events(inputs, bg)
        r_ptr = np.cumsum([0] + [len(r) for r in ref["raster"]]).astype(np.int64)
        r_ids = np.concatenate(ref["raster"]).astype(np.uint32) if r_ptr[-1] else np.zeros(0, np.uint32)
        np.savez_compressed(
            os.path.join(HERE, "micro_%02d.npz" % made),
            dt=np.float32(dt), steps=np.int32(steps), seed=np.int64(seed),
            in_ptr=net["in_ptr"], in_src=net["in_src"], w0=net["w"], delay=net["delay"], inh=net["inh"],
            ev_step=so, ev_id=ids, ev_kind=kd,
            ref_raster_ptr=r_ptr, ref_raster_ids=r_ids, ref_V=ref["V"], ref_w=ref["w"],
            ref_fire_t=ref["fire_t"], ref_last_arr=ref["last_arr"], ref_last_fire=ref["last_fire"])
        made += 1
        print("fixture", made, "seed", seed, "dt", dt, "spikes", int(r_ptr[-1]))


if __name__ == "__main__":
    main()
