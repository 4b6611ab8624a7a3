"""The oracle's reverse mode (orc_adjoint) against central finite differences of
its own forward step: pins the FP64 reference gradient used to check the HIP
adjoint."""
import numpy as np
import pytest


def setup(orc, K1=0.0, lap=5, scheme=1, eos="PRW", tr=0.75):
    nx, ny, nz = 8, 6, 14
    p = orc.make_params(nx, ny, nz, eos=eos, tr=tr, nwall=3, K1=K1, K2=-K1, lap_points=lap,
                        wall_scheme=scheme, Ka=0.002)
    A, B, Tc, Rc = orc.set_multiphase(eos, tr)
    rng = np.random.default_rng(11)
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    r = np.sqrt((x - 4) ** 2 + (y - 3) ** 2 + (z - 3) ** 2)
    g, l = 0.108623, 7.67892
    rho0 = g + (l - g) * 0.5 * (1 - np.tanh(2 * (r - 4.5) / 4.0))
    rho0 += 1e-3 * rng.standard_normal(rho0.shape)
    mu_lo = -0.02 + 0.01 * rng.standard_normal(nx * ny)
    mu_hi = 0.01 * rng.standard_normal(nx * ny)
    target = rho0.ravel() * (1 + 0.01 * rng.standard_normal(rho0.size))
    return p, rho0.ravel(), mu_lo, mu_hi, target


def run_J(orc, p, rho0, mu_lo, mu_hi, target, nsteps, Ka=None):
    if Ka is not None:
        p.Ka = Ka
    o = orc.Oracle(p)
    o.set_wall_mu(mu_lo, mu_hi)
    o.init_from_rho(rho0)
    o.step(nsteps)
    rho = o.get(orc.RHO)
    fl = o.get(orc.TYPE) == 1
    return 0.5 * np.sum((rho[fl] - target[fl]) ** 2)


@pytest.mark.parametrize("variant", [dict(), dict(K1=0.01), dict(lap=7), dict(scheme=0)])
def test_oracle_adjoint_matches_finite_differences(orc, variant):
    p, rho0, mu_lo, mu_hi, target = setup(orc, **variant)
    nsteps = 6
    o = orc.Oracle(p)
    o.set_wall_mu(mu_lo, mu_hi)
    o.init_from_rho(rho0)
    J, glo, ghi, gka = o.adjoint(nsteps, target)
    assert J == pytest.approx(run_J(orc, p, rho0, mu_lo, mu_hi, target, nsteps), rel=1e-12)
    rng = np.random.default_rng(2)
    h = 1e-6
    for _ in range(3):
        d_lo, d_hi = rng.standard_normal(mu_lo.size), rng.standard_normal(mu_hi.size)
        jp = run_J(orc, p, rho0, mu_lo + h * d_lo, mu_hi + h * d_hi, target, nsteps)
        jm = run_J(orc, p, rho0, mu_lo - h * d_lo, mu_hi - h * d_hi, target, nsteps)
        fd = (jp - jm) / (2 * h)
        ad = glo @ d_lo + ghi @ d_hi
        assert ad == pytest.approx(fd, rel=1e-4, abs=1e-9 * abs(J))
    ka = p.Ka
    hk = 1e-7
    fd = (run_J(orc, p, rho0, mu_lo, mu_hi, target, nsteps, Ka=ka + hk)
          - run_J(orc, p, rho0, mu_lo, mu_hi, target, nsteps, Ka=ka - hk)) / (2 * hk)
    assert gka == pytest.approx(fd, rel=1e-4)
