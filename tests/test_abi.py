"""The C ABI library loads without a GPU, exports every symbol include/cpmp.h
declares, and its host-side helpers (no device calls) agree with the oracle and
the reference's golden coexistence tables."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(cpmp_[a-z0-9_]+)\s*\(", text))
    return names


def test_exports_every_declared_symbol(cpmp):
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(cpmp.lib, n), n
    assert set(cpmp.EXPORTS) == names


def test_version_and_no_device(cpmp):
    assert b"gfx950" in cpmp.lib.cpmp_version()


@pytest.mark.parametrize("eos", ["VDW", "CSE", "RKE", "RKS", "PRW", "PRM"])
def test_set_multiphase_matches_oracle(cpmp, orc, eos):
    for tr in (0.6, 0.7, 0.85):
        assert cpmp.set_multiphase(eos, tr) == orc.set_multiphase(eos, tr)


@pytest.mark.parametrize("eos", ["VDW", "CSE", "RKE", "RKS", "PRW", "PRM"])
def test_eos_matches_oracle(cpmp, orc, eos):
    A, B, Tc, Rc = cpmp.set_multiphase(eos, 0.75)
    T = 0.75 * Tc
    for rho in np.linspace(0.01, 0.9, 12) / B:
        assert cpmp.eos_pressure(eos, A, B, T, rho) == pytest.approx(orc.pressure(eos, A, B, T, rho), rel=1e-13)
        for form in (0, 1):
            assert cpmp.eos_mu(eos, A, B, T, rho, form) == pytest.approx(orc.mu_eos(eos, A, B, T, rho, form),
                                                                         rel=1e-12, abs=1e-14)


def test_default_params_are_initialize(cpmp):
    p = cpmp.make_params(20, 10, 30)
    assert (p.K, p.Ka, p.K1, p.K2) == (0.2, 0.001, 0.0, 0.0)
    assert p.nwall_lo == p.nwall_hi == 3 and p.lap_points == 5 and p.wall_scheme == 1
    assert list(p.Ts) == cpmp.reference_ts(0.8)


@pytest.mark.parametrize("eos,tr", [("CSE", 0.85), ("PRW", 0.7), ("PRW", 0.75), ("VDW", 0.8), ("PRM", 0.6)])
def test_coexistence_solver_matches_tables(cpmp, cocurve, eos, tr):
    A, B, Tc, Rc = cpmp.set_multiphase(eos, tr)
    row = next(r for r in cocurve[eos] if abs(r[0] - tr) < 1e-12)
    g, l = cpmp.coexistence(eos, A, B, tr * Tc)           # no starting values
    assert g == pytest.approx(row[1] * Rc, rel=3e-5)
    assert l == pytest.approx(row[2] * Rc, rel=3e-6)


def test_cocurve_lookup(cpmp, cocurve, tmp_path):
    path = tmp_path / "CoCurve_PRW_2.txt"
    with open(path, "w") as fh:
        fh.write("T   DenGas   DenLiquid \n")
        for t, g, l in cocurve["PRW"]:
            fh.write(f"{t:.11f}    {g:.16f}    {l:.16f} \n")
    g, l = cpmp.cocurve_lookup(str(path), 0.7)
    assert (g, l) == pytest.approx((0.0209314, 3.0408454), rel=1e-6)
    with pytest.raises(cpmp.CPMPError) as e:
        cpmp.cocurve_lookup(str(path), 0.7123)
    assert e.value.code == -6
    with pytest.raises(cpmp.CPMPError) as e:
        cpmp.cocurve_lookup(str(tmp_path / "missing.txt"), 0.7)
    assert e.value.code == -4


def test_create_without_device_fails_cleanly(cpmp):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    p = cpmp.make_params(8, 8, 12)
    with pytest.raises(cpmp.CPMPError) as e:
        cpmp.Lattice(p)
    assert e.value.code == -7


def test_invalid_params_rejected(cpmp):
    p = cpmp.make_params(8, 8, 6)            # 3+3 wall layers leave no fluid
    with pytest.raises(cpmp.CPMPError) as e:
        cpmp.Lattice(p)
    assert e.value.code in (-1, -7)


def test_dim_conversion_is_dimconversion(cpmp):
    """DimConversion (FluidCuda.cu:474-486) with a lattice of 100 nodes for 1 mm,
    water-like density and viscosity; the factors obey their definitions"""
    L, ML, D, MD, V, MV = 100.0, 1e-3, 1.0, 1000.0, 0.1, 1e-6
    d = cpmp.dim_conversion(L, ML, D, MD, V, MV)
    dl = ML / L
    dt = V / MV * dl * dl
    dm = MD / D * dl ** 3
    want = dict(len=dl, time=dt, mass=dm, force=dm * dl / dt ** 2, torque=dm * dl * dl / dt ** 2,
                velocity=dl / dt, acceleration=dl / dt ** 2, omega=1 / dt, beta=1 / dt ** 2)
    for k, v in want.items():
        assert d[k] == pytest.approx(v, rel=1e-15), k
    # a lattice viscosity of 0.1 at dx = 10 um in water gives dt = 1e-5 s
    assert d["time"] == pytest.approx(1e-5, rel=1e-12)
    with pytest.raises(cpmp.CPMPError):
        cpmp.dim_conversion(0.0, ML, D, MD, V, MV)
