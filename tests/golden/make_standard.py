"""Writes tests/golden/standard_ref.json: what the compiled reference's
STANDARD preset (src/main.cpp:80-124: NeuCor(750), three inputs on a circle)
gives over 1000 ms with the inputs held at 40 Hz, for five shifts of the C
library's rand() stream (0-4 extra draws after construction: the background
fires of src/NeuCor.cpp:604-607 draw from that stream), in frames of 4 ms,
with every synapse's sign class repaired (setWeight(getWeight()); the
reference's Synapse copy loses it, :487-524), in two modes:

  event   runAll = false: the preset as it is configured; a neuron runs only
          at its own events
  clocked runAll = true: every neuron also runs at the start of every frame

Run in the build container (it needs oracle/_ref/ref_harness, built from
the reference's own src/NeuCor.cpp by oracle/Makefile):

    python tests/golden/make_standard.py
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

SEED, MS, RATE = 1, 1000, 40
MODES = {"event": (4, 0, 1), "clocked": (4, 1, 1)}  # run_speed, run_all, walk_every


def run(mode, skip):
    rs, ra, walk = MODES[mode]
    out = subprocess.run([HARNESS, "--standard", str(SEED), str(MS), str(rs), str(ra), str(walk), str(RATE),
                          str(skip)], capture_output=True, text=True, check=True, timeout=1800).stdout
    d = {}
    for line in out.splitlines():
        k, *v = line.split()
        d[k] = [int(x) for x in v] if k == "w_hist" else float(v[0])
    return {"mode": mode, "skip": skip, "fires": int(d["fires"]), "neurons_fired": int(d["neurons_fired"]),
            "mean_w": d["mean_w"], "w_hist": d["w_hist"]}


def main():
    if not os.path.exists(HARNESS):
        sys.exit("oracle/_ref/ref_harness is not built (make -C oracle ref)")
    runs = [run(m, k) for m in MODES for k in range(5)]
    doc = {"seed": SEED, "ms": MS, "rate_hz": RATE,
           "modes": {m: {"run_speed": v[0], "run_all": v[1], "walk_every": v[2]} for m, v in MODES.items()},
           "runs": runs}
    with open(os.path.join(HERE, "standard_ref.json"), "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    for r in runs:
        print(r["mode"], r["skip"], r["fires"], round(r["mean_w"], 5))


if __name__ == "__main__":
    main()
