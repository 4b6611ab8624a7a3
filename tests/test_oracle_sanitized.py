"""The oracle under the host address and undefined-behaviour sanitizers: every
entry point (tables, EOS, Maxwell construction, init, the stages and steps of the
wall schemes, Laplacians, force terms and Qp, reductions, contact angle, noise,
the adjoint) on small lattices, built from oracle/asan_check.cpp (CPU only)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_entry_points_run_clean_under_sanitizers():
    if shutil.which(os.environ.get("CXX", "g++")) is None or shutil.which("make") is None:
        pytest.skip("no host C++ toolchain")
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "asan-check"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "ran clean" in r.stdout
