"""The Guo force term, ChemBoundaryComplex and the FP32-arithmetic form on the device.

FP64 parity of both (<= 1e-10 on rho) runs with every case of gpu_cases.CASES in
test_gpu_parity.py; here the FP32 and FP16 storage forms (fused density with
long chunks included) against the FP64 oracle, the stage fields of the wall
scheme, and the adjoint's refusal of the variants it is not derived for."""
import os

import numpy as np
import pytest

import gpu_cases as gc
import oracle_lib as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("xc", [None, 2, 32])
@pytest.mark.parametrize("form,tol", [("fp32", 2e-4), ("fp16", 2e-2)])
def test_guo_reduced_forms_match_oracle(cpmp, monkeypatch, form, tol, xc):
    case = gc.CASES["guo_rke_periodic"]
    if xc:
        monkeypatch.setenv("CPMP_XC", str(xc))
    print(f"[guo] {form} xc={xc} {case['dims']}", flush=True)
    o, g = gc.build(case, cpmp, form)
    fluid = o.get(orc.TYPE) == 1
    o.step(case["steps"])
    g.step(case["steps"])
    ro = o.get(orc.RHO)
    err = np.max(np.abs(g.get("rho")[fluid] - ro[fluid])) / np.max(ro[fluid])
    mv_o, mv_g = o.get(orc.MV).reshape(3, -1), g.get("mv").reshape(3, -1)
    mv_err = np.max(np.abs(mv_g - mv_o)) / np.max(np.abs(mv_o))
    print(f"[guo] {form} xc={xc}: rho err {err:.3e}, MV err {mv_err:.3e}", flush=True)
    assert err < tol
    assert mv_err < 50 * tol
    g.close()


def test_guo_and_exact_difference_differ_on_device(cpmp):
    """the variant is taken: the two force terms give different states, whose
    difference matches the oracle's"""
    base = dict(gc.CASES["guo_vdw_walls"])
    res = {}
    for forcing in (0, 1):
        case = dict(base, forcing=forcing, steps=50)
        o, g = gc.build(case, cpmp)
        o.step(50)
        g.step(50)
        res[forcing] = (o.get(orc.RHO), g.get("rho"))
        g.close()
    d_o = res[1][0] - res[0][0]
    d_g = res[1][1] - res[0][1]
    assert np.max(np.abs(d_o)) > 1e-8
    assert np.max(np.abs(d_g - d_o)) < 1e-9 * np.max(np.abs(res[0][0]))


@pytest.mark.parametrize("form,tol", [("fp32", 2e-4), ("fp16", 2e-2)])
def test_complex_wall_reduced_forms_match_oracle(cpmp, form, tol):
    case = gc.CASES["complex_wall"]
    o, g = gc.build(case, cpmp, form)
    fluid = o.get(orc.TYPE) == 1
    o.step(case["steps"])
    g.step(case["steps"])
    ro = o.get(orc.RHO)
    err = np.max(np.abs(g.get("rho")[fluid] - ro[fluid])) / np.max(ro[fluid])
    print(f"[complex-wall] {form}: rho err {err:.3e}", flush=True)
    assert err < tol
    g.close()


@pytest.mark.parametrize("lo,hi", [(3, 2), (1, 0), (4, 4)])
def test_complex_wall_stage_fields(cpmp, lo, hi):
    """the level fill of the wall layers (incl. one layer touching fluid on both sides
    through the periodic wrap, and walls deeper than the reference's three levels)
    and the wall chemical potential, after one step, against the oracle"""
    nx, ny, nz = 12, 10, 20
    kw = dict(eos="PRW", tr=0.7, nwall=lo, nwall_hi=hi, wall_scheme=2)
    rng = np.random.default_rng(3)
    rho0 = rng.uniform(1.0, 3.0, nx * ny * nz)
    o = orc.Oracle(orc.make_params(nx, ny, nz, **kw))
    g = cpmp.Lattice(cpmp.make_params(nx, ny, nz, **kw))
    mw = np.linspace(-0.05, 0.05, nx * ny)
    o.set_wall_mu(mw, 0.02)
    g.set_wall_mu(mw, 0.02)
    o.init_from_rho(rho0)
    g.init_from_rho(rho0)
    o.step(2)
    g.step(2)
    # rho incl. the wall layers filled in step 2, mu incl. the wall field
    assert gc.rel_err(g.get("rho"), o.get(orc.RHO)) < 1e-12
    assert np.max(np.abs(g.get("mu") - o.get(orc.MU))) < 1e-11 * np.max(np.abs(o.get(orc.MU)))
    g.close()


def test_adjoint_refuses_variants(cpmp):
    """the adjoint is derived for the exact-difference force term, the layered and
    simple wall schemes, and the FP64 and FP32 forms"""
    dims = (16, 12, 20)
    target = np.zeros(dims[0] * dims[1] * dims[2])
    for kw in (dict(forcing=1), dict(wall_scheme=2), dict(form="fp32a")):
        g = cpmp.Lattice(cpmp.make_params(*dims, eos="PRW", tr=0.75, **kw))
        g.set_wall_mu(0.0)
        g.init_reference(5, 4, 0.108623, 7.67892)
        with pytest.raises(cpmp.CPMPError) as e:
            g.adjoint(10, target, checkpoint_every=5)
        assert e.value.code == -1
        g.close()
