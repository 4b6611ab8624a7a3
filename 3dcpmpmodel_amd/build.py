"""Builds the HIP library libcpmp.so for gfx950 with hipcc.

    python 3dcpmpmodel_amd/build.py

Every source is compiled for gfx950 only, each translation unit in its own
hipcc process (they share no device code), then linked.  The host compiler
passes see the packed-FP32 feature flag as unknown and warn, which is harmless.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libcpmp.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["kernels.hip", "api.hip", "adjoint.hip", "observables.cpp", "plt.cpp", "noalloc.hip"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    # keep hipcc from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run_all(cmds, verbose):
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    failed = []
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            failed.append(cmd[-1])
    if failed:
        raise RuntimeError("hipcc failed: " + ", ".join(failed))


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    inc = ["-I", os.path.join(os.path.dirname(HERE), "include")]
    objs, cmds = [], []
    for s in SOURCES:
        o = os.path.join(OBJ, os.path.splitext(s)[0] + ".o")
        objs.append(o)
        cmds.append([HIPCC] + FLAGS + inc + ["-x", "hip", "-c", os.path.join(CSRC, s), "-o", o])
    _run_all(cmds, verbose)
    _run_all([[HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", OUT + ".tmp"]], verbose)
    os.replace(OUT + ".tmp", OUT)
    # the reference program's driver on the C ABI (host C++ only)
    drv = os.path.join(HERE, "cpmp_sweep")
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-I", os.path.join(os.path.dirname(HERE), "include"),
           os.path.join(HERE, "driver", "cpmp_sweep.cpp"), "-L", HERE, "-lcpmp", "-Wl,-rpath,$ORIGIN", "-o", drv]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed on cpmp_sweep.cpp")
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
