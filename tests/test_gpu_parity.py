"""FP64 HIP path against the FP64 oracle: <= 1e-10 relative on rho after N steps,
plus stage-level fields (mu, F, MV, u) and the first-step contract."""
import numpy as np
import pytest

import gpu_cases as gc
import oracle_lib as orc

pytestmark = pytest.mark.gpu

TOL = 1e-10   # relative, per node, on rho after N steps (FP64 form)


@pytest.mark.parametrize("collide", ["auto", "march"])
@pytest.mark.parametrize("name", list(gc.CASES))
def test_fp64_rho_parity_after_n_steps(cpmp, monkeypatch, name, collide):
    """every case with the collide the lattice size selects (the per-node kernel
    on these small FP64 lattices) and with the x march forced"""
    case = gc.CASES[name]
    print(f"[parity] {name} dims={case['dims']} steps={case['steps']} collide={collide}", flush=True)
    if collide == "march":
        monkeypatch.setenv("CPMP_COLLIDE", "march")
    o, g = gc.build(case, cpmp)
    fluid = o.get(orc.TYPE) == 1
    assert gc.rel_err(g.get("rho"), o.get(orc.RHO), fluid) < 1e-13   # device tanh vs glibc
    n = case["steps"]
    o.step(n)
    g.step(n)
    err = gc.rel_err(g.get("rho"), o.get(orc.RHO), fluid)
    print(f"[parity] {name}: max rel err rho after {n} steps = {err:.3e}", flush=True)
    assert err < TOL
    # populations and derived fields of the final state
    fo, fg = o.get(orc.F).reshape(19, -1), g.get("f").reshape(19, -1)
    assert np.max(np.abs(fg[:, fluid] - fo[:, fluid])) < TOL * np.max(np.abs(fo[:, fluid]))
    mvo, mvg = o.get(orc.MV).reshape(3, -1), g.get("mv").reshape(3, -1)
    scale = np.max(np.abs(mvo[:, fluid])) + 1e-30
    assert np.max(np.abs(mvg[:, fluid] - mvo[:, fluid])) < 1e-8 * scale
    vo, vg = o.get(orc.VEL).reshape(3, -1), g.get("vel").reshape(3, -1)
    assert np.max(np.abs(vg[:, fluid] - vo[:, fluid])) < 1e-8 * (np.max(np.abs(vo[:, fluid])) + 1e-30)


@pytest.mark.parametrize("name", ["ref_prw_walls", "vdw_variants", "prm_walls_2_4"])
def test_stage_fields_after_one_step(cpmp, name):
    """wall fill, mu and force of the same state (stage parity)."""
    case = gc.CASES[name]
    o, g = gc.build(case, cpmp)
    o.step(3)
    g.step(3)
    # force stage of the current state: both run wall fill + mu first
    o.stage(0); o.stage(1); o.stage(2)
    Fo = o.get(orc.FORCE).reshape(3, -1)
    Fg = g.get("force").reshape(3, -1)
    fluid = o.get(orc.TYPE) == 1
    assert np.max(np.abs(Fg[:, fluid] - Fo[:, fluid])) < 1e-9 * np.max(np.abs(Fo[:, fluid]))
    ro, rg = o.get(orc.RHO), g.get("rho")
    assert gc.rel_err(rg, ro) < 1e-12                      # incl. wall-filled solid layers
    mo, mg = o.get(orc.MU), g.get("mu")
    assert np.max(np.abs(mg - mo)) < 1e-11 * np.max(np.abs(mo))   # incl. solid mu = wall field


def test_type_field(cpmp):
    case = gc.CASES["prm_walls_2_4"]
    o, g = gc.build(case, cpmp)
    assert np.array_equal(g.get("type"), o.get(orc.TYPE))


def test_determinism_bitwise(cpmp):
    case = gc.CASES["ref_prw_walls"]
    outs = []
    for _ in range(2):
        _, g = gc.build(case, cpmp)
        g.step(50)
        outs.append(g.get("rho"))
        g.close()
    assert np.array_equal(outs[0], outs[1])


def test_reduce_matches_oracle(cpmp):
    case = gc.CASES["ref_prw_walls"]
    o, g = gc.build(case, cpmp)
    o.step(20)
    g.step(20)
    mass, bad, ms = o.reduce()
    s = g.reduce()
    assert s.mass == pytest.approx(mass, rel=1e-12)
    assert s.bad_count == 0 and s.first_bad == -1
    assert s.max_speed == pytest.approx(ms, rel=1e-8)
    # NaN / negative detection names the first bad node
    rho = g.get("rho")
    nx, ny, nz = case["dims"]
    I = (5 * ny + 7) * nz + 9
    rho[I] = np.nan
    rho[I + 1] = -1.0
    g.set("rho", rho)
    with pytest.raises(cpmp.CPMPError) as e:
        g.reduce()
    assert e.value.code == -5 and "(5,7,9)" in str(e.value)
    s = g.reduce(raise_on_nan=False)
    assert s.bad_count == 2 and s.first_bad == I


def test_set_get_roundtrip(cpmp):
    case = gc.CASES["rke_coalesce"]
    o, g = gc.build(case, cpmp)
    o.step(10)
    g.set("f", o.get(orc.F))
    g.set("rho", o.get(orc.RHO))
    assert np.array_equal(g.get("f"), o.get(orc.F))
    o.step(10)
    g.step(10)
    assert gc.rel_err(g.get("rho"), o.get(orc.RHO)) < TOL


def test_droplet_splash_init_and_parity(cpmp):
    """SetFlowField DROPLETSPLASH (FluidCuda.cu:219-233) against its formula with the
    single-precision radius, then 100 FP64 steps against the oracle from that state"""
    nx, ny, nz = 40, 36, 48
    dg, dl = gc.densities("PRW", 0.7)
    kw = dict(eos="PRW", tr=0.7, nwall=3)
    g = cpmp.Lattice(cpmp.make_params(nx, ny, nz, **kw))
    g.set_wall_mu(-0.02)
    g.init_splash(dg, dl, radius=12.0, width=5.0, film_z=10.0)
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    d2 = (x - nx // 2) ** 2.0 + (y - ny // 2) ** 2.0 + (z - nz // 2) ** 2.0
    r = np.sqrt(d2.astype(np.float32)).astype(np.float64)
    want = dl + (dl - dg) / 2 * (np.tanh((12.0 - r) * 2 / 5.0) + np.tanh((10.0 - z) * 2 / 5.0))
    want[:, :, :3] = 0
    want[:, :, nz - 3:] = 0
    rho0 = g.get("rho")
    assert np.allclose(rho0, want.ravel(), rtol=1e-13, atol=1e-13)
    o = orc.Oracle(orc.make_params(nx, ny, nz, **kw))
    o.set_wall_mu(np.full(nx * ny, -0.02), np.full(nx * ny, -0.02))
    o.init_from_rho(rho0)
    g.step(100)
    o.step(100)
    fl = g.get("type") == 1
    err = np.max(np.abs(g.get("rho")[fl] - o.get(orc.RHO)[fl]) / o.get(orc.RHO)[fl])
    print(f"[parity] splash: max rel err rho after 100 steps = {err:.3e}", flush=True)
    assert err < 1e-10
    g.close()


def test_64cubed_carnahan_starling_configuration(cpmp):
    """the 64^3 configuration in full: Carnahan-Starling at Tr = 0.85, centred sphere
    of radius 16 with tanh width 2.5, noise 1e-6 (seed 1), kappa 0.01, tau 1, 1000
    periodic FP64 steps, on the GPU and in the oracle: the density field, the Laplace
    jump K^2 [p(centre) - p(corner)] and the spurious current max |MV|"""
    case = dict(dims=(64, 64, 64), eos="CSE", tr=0.85, nwall=0, init="spheres", spheres=[(32, 32, 32, 16)],
                width=2.5, Ka=0.01, tau=1.0, noise=1e-6, seed=1)
    o, g = gc.build(case, cpmp)
    o.step(1000)
    g.step(1000)
    rho_o = o.get(orc.RHO)
    err = gc.rel_err(g.get("rho"), rho_o)
    mass, bad, ms = o.reduce()
    s = g.reduce()
    dp = g.laplace_jump((32, 32, 32), (0, 0, 0))
    A, B, Tc, Rc = cpmp.set_multiphase("CSE", 0.85)
    r3 = rho_o.reshape(64, 64, 64)
    dp_o = g.p.K ** 2 * (cpmp.eos_pressure("CSE", A, B, 0.85 * Tc, r3[32, 32, 32])
                         - cpmp.eos_pressure("CSE", A, B, 0.85 * Tc, r3[0, 0, 0]))
    print(f"[c1] rho rel err {err:.2e}; Laplace jump {dp:.6e} (oracle density {dp_o:.6e}); "
          f"max |MV| {s.max_speed:.6e} (oracle {ms:.6e})", flush=True)
    assert err < 1e-10
    assert s.mass == pytest.approx(mass, rel=1e-12)
    assert dp == pytest.approx(dp_o, rel=1e-9)
    assert s.max_speed == pytest.approx(ms, rel=1e-8)
    g.close()
