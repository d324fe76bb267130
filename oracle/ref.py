"""Driver of the headless reference harness (oracle/_ref/ref_harness).

TEST INFRASTRUCTURE ONLY.  Writes a scenario (network + per-step forced
fires), runs the reference's own NeuCor::run() on it, and reads back the
per-step raster and potentials and the final synapse state.
"""
import json
import os
import struct
import subprocess
import tempfile

import numpy as np

from .oracle import REF_HARNESS


def available():
    return os.path.exists(REF_HARNESS)


def write_scenario(path, net, steps, dt=1.0, lr=1.0, pre_f=0.13, post_f=0.30, inputs=None, bg=None):
    """inputs / bg: list (len steps) of id lists forced at step k+1."""
    n = len(net["in_ptr"]) - 1
    S = len(net["in_src"])
    with open(path, "wb") as f:
        f.write(struct.pack("<IIQII", 0x3153434E, n, S, steps, 0))
        f.write(struct.pack("<ffff", dt, lr, pre_f, post_f))
        f.write(np.ascontiguousarray(net["in_ptr"], np.int64).tobytes())
        f.write(np.ascontiguousarray(net["in_src"], np.uint32).tobytes())
        f.write(np.ascontiguousarray(net["w"], np.float32).tobytes())
        f.write(np.ascontiguousarray(net["delay"], np.uint8).tobytes())
        for k in range(steps):
            for lst in (inputs, bg):
                ids = np.asarray(lst[k] if lst is not None else [], np.uint32)
                f.write(struct.pack("<I", len(ids)))
                f.write(ids.tobytes())


def run_trace(net, steps, **kw):
    with tempfile.TemporaryDirectory() as d:
        sp, op = os.path.join(d, "s.bin"), os.path.join(d, "o.bin")
        write_scenario(sp, net, steps, **kw)
        subprocess.run([REF_HARNESS, sp, op], check=True)
        buf = open(op, "rb").read()
    n = len(net["in_ptr"]) - 1
    S = len(net["in_src"])
    magic, n2, steps2, _ = struct.unpack_from("<IIII", buf, 0)
    assert magic == 0x314F434E and n2 == n and steps2 == steps
    off = 16
    raster, V = [], np.empty((steps, n), np.float32)
    for k in range(steps):
        (nf,) = struct.unpack_from("<I", buf, off)
        off += 4
        raster.append(np.frombuffer(buf, np.uint32, nf, off).copy())
        off += 4 * nf
        V[k] = np.frombuffer(buf, np.float32, n, off)
        off += 4 * n
    w = np.frombuffer(buf, np.float32, S, off).copy(); off += 4 * S
    ft = np.frombuffer(buf, np.float32, S, off).copy(); off += 4 * S
    la = np.frombuffer(buf, np.float32, S, off).copy(); off += 4 * S
    lf = np.frombuffer(buf, np.float32, n, off).copy()
    return {"raster": raster, "V": V, "w": w, "fire_t": ft, "last_arr": la, "last_fire": lf}


def run_bench(net, steps, dt=1.0, inputs=None, workdir=None, warmup=0, bg=None):
    """Times the reference's NeuCor::run() over `steps` steps (runAll,
    runSpeed = dt) after `warmup` untimed ones; inputs (a list of
    warmup + steps id lists) fired between runs.  Returns the harness JSON."""
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        sp, op = os.path.join(d, "s.bin"), os.path.join(d, "o.json")
        write_scenario(sp, net, warmup + steps, dt=dt, inputs=inputs, bg=bg)
        subprocess.run([REF_HARNESS, "--bench", "--warmup", str(warmup), sp, op], check=True)
        return json.loads(open(op).read())
