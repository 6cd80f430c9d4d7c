"""Build the HIP extension libmjwave.so in-tree (gfx950).

    python -m montjoie_amd.build        # or montjoie_amd.build.build()

Every *.hip file of csrc/ is compiled by hipcc for gfx950 (packed-FP32
instructions disabled), the host-only *.cpp by g++, and the objects are
linked into montjoie_amd/libmjwave.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmjwave.so")
RUNNER = os.path.join(HERE, "mjwave_run")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
ARCH = os.environ.get("MJW_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose: bool = False, jobs: int = 0, incremental: bool = False) -> str:
    """Compile and link; with incremental=True an object newer than its source
    and than every header of csrc/ and include/ is kept."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))
    hdrs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    hdrs += [os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE)]
    newest_hdr = max(os.path.getmtime(h) for h in hdrs)
    cmds = []
    objs = []
    for f in srcs:
        src = os.path.join(CSRC, f)
        obj = os.path.join(OBJ, f + ".o")
        objs.append(obj)
        if (incremental and os.path.exists(obj) and
                os.path.getmtime(obj) > max(os.path.getmtime(src), newest_hdr)):
            continue
        if f.endswith(".hip"):
            cmds.append([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
        else:
            cmds.append(["g++", *CXX_FLAGS, "-c", src, "-o", obj])
    jobs = jobs or min(max(len(cmds), 1), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(jobs) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lgomp"])
    os.replace(tmp, LIB)
    # deck driver (host C++ over the C ABI)
    tmp = RUNNER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, os.path.join(HERE, "driver",
          "mjwave_run.cpp"), "-o", tmp, "-L", HERE, "-lmjwave", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, RUNNER)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="-i" in sys.argv))
