"""Build the engine's shared library in-tree (gfx950).

Host C++ sources are compiled with g++, HIP sources with hipcc
(--offload-arch=gfx950), and both are linked by hipcc into
neurocorrelation_amd/libneucor_amd.so.  Floating-point contraction is off
everywhere: the step reproduces the reference's float/double rounding
sequence operation by operation, and a fused multiply-add would change it.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libneucor_amd.so")
ARCH = os.environ.get("NC_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-pthread"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + err + "\n")
        raise RuntimeError("build step failed: " + cmd[0])
    return err


def build(verbose=False):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for src in sorted(os.listdir(CSRC)):
        path = os.path.join(CSRC, src)
        obj = os.path.join(bdir, src + ".o")
        if src.endswith(".cpp"):
            cmd = ["g++"] + CXX_FLAGS + ["-c", path, "-o", obj]
        elif src.endswith(".hip"):
            cmd = [hipcc] + HIP_FLAGS + ["-c", path, "-o", obj]
        else:
            continue
        msg = _run(cmd)
        if verbose and msg.strip():
            print(msg)
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-pthread"])
    os.replace(tmp, OUT)
    build_facade()
    return OUT


def build_facade():
    """class NeuCor over the C ABI: libneucor_facade.so and its test driver
    facade/facade_check, both linked against libneucor_amd.so by rpath."""
    fdir = os.path.join(HERE, "facade")
    obj = os.path.join(HERE, "_build", "NeuCor.o")
    _run(["g++"] + CXX_FLAGS + ["-c", os.path.join(fdir, "NeuCor.cpp"), "-o", obj])
    link = ["-L" + HERE, "-lneucor_amd", "-Wl,-rpath,$ORIGIN", "-pthread"]
    tmp = os.path.join(HERE, "libneucor_facade.so.tmp")
    _run(["g++", "-shared", "-o", tmp, obj] + link)
    os.replace(tmp, os.path.join(HERE, "libneucor_facade.so"))
    exe = os.path.join(fdir, "facade_check")
    _run(["g++"] + CXX_FLAGS + ["-o", exe + ".tmp", os.path.join(fdir, "facade_check.cpp"), obj,
                                "-L" + HERE, "-lneucor_amd", "-Wl,-rpath,$ORIGIN/..", "-pthread"])
    os.replace(exe + ".tmp", exe)
    demo = os.path.join(fdir, "renderer_access_demo")
    _run(["g++"] + CXX_FLAGS + ["-DNC_RENDERER_ACCESS_MAIN", "-o", demo + ".tmp",
                                os.path.join(fdir, "renderer_access.cpp"), obj, "-L" + HERE, "-lneucor_amd",
                                "-Wl,-rpath,$ORIGIN/..", "-pthread"])
    os.replace(demo + ".tmp", demo)


if __name__ == "__main__":
    print(build(verbose=True))
