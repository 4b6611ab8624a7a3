"""The reference program's driver (cpmp_sweep, main of FluidCuda.cu:1255-1304) on the
C ABI: a two-value BasePt sweep with the state carried over; its SaveContactAngle
file against the same sequence run through the Python front end."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "3dcpmpmodel_amd", "cpmp_sweep")


def test_basept_sweep_matches_api(cpmp, tmp_path):
    assert os.path.exists(EXE), "build the package first (cpmp_sweep)"
    dims, radius, width, steps, show = (32, 24, 24), 8.0, 5.0, 60, 20
    args = [EXE, "--dims", *map(str, dims), "--eos", "PRW", "--tr", "0.7", "--radius", str(radius), "--width",
            str(width), "--basept", "-0.02", "-0.01", "0.01", "--steps", str(steps), "--show", str(show),
            "--out", str(tmp_path)]
    res = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(res.stdout[-2000:], res.stderr[-2000:], flush=True)
    assert res.returncode == 0
    files = os.listdir(tmp_path)
    assert len(files) == 1 and files[0].startswith("CPPRW_MRT_Tau=0.8_FIELD32x24x24_ContactAngle_R8.0_Tr0.7")
    lines = open(os.path.join(tmp_path, files[0])).read().split("\n")
    assert lines[0].split() == ["Droprad", "BasePt", "NowStep", "CaLeft", "CaRight", "CaCap", "XLeft", "XRight"]
    rows = [list(map(float, l.split())) for l in lines[1:] if l.strip()]
    assert len(rows) == 2 and rows[0][1] == -0.02 and rows[1][1] == pytest.approx(-0.01)
    # ShowData rows: 2 BasePt x (steps/show + 1) rows of 12 numbers
    table = [l for l in res.stdout.split("\n") if l.strip() and l.split()[0].isdigit()]
    assert len(table) == 2 * (steps // show + 1) and all(len(l.split()) == 12 for l in table)
    # the same sequence through the Python front end
    A, B, Tc, Rc = cpmp.set_multiphase("PRW", 0.7)
    g_, l_ = cpmp.coexistence("PRW", A, B, 0.7 * Tc)
    lat = cpmp.Lattice(cpmp.make_params(*dims, eos="PRW", tr=0.7))
    lat.init_reference(radius, width, g_, l_, -0.02)
    for k, bp in enumerate((-0.02, -0.01)):
        lat.set_wall_mu(bp)
        for _ in range(steps // show):
            lat.step(show)
        ca = lat.contact_angle(dims[1] // 2, g_, l_)
        got = rows[k][3:8]
        want = [ca["CaLeft"], ca["CaRight"], ca["CaCap"], ca["XLeft"], ca["XRight"]]
        assert np.allclose(got, want, rtol=1e-5, atol=1e-5, equal_nan=True), (k, got, want)
    lat.close()
