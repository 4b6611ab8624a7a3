"""Builds the HIP library lib/libcpmp.so for gfx950 with hipcc.

    python 3dcpmpmodel_amd/build.py

Every source is compiled for gfx950 only; the host compiler passes see the
packed-FP32 feature flag as unknown and warn, which is harmless.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcpmp.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["kernels.hip", "api.hip", "adjoint.hip", "observables.cpp", "plt.cpp"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",
    # keep hipcc from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(os.path.dirname(HERE), "include")]
    cmd += ["-x", "hip"] + srcs + ["-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(OUT + ".tmp", OUT)
    # reference kernels of known traffic (counter calibration, attainable bandwidth)
    probe = os.path.join(HERE, "probe_copy")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-Wno-unused-command-line-argument", os.path.join(CSRC, "probe_copy.hip"), "-o", probe]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed on probe_copy.hip")
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
