"""Pins of the two scheme variants beyond the reference's default path (CPU only).

Guo forcing (forcing = 1): the device's moment-space collision (run on the CPU
through the self-test entry) against the dense population-space form of Guo,
Zheng and Shi (2002) in MRT form; mass kept exactly and momentum raised by F, per
node and over a periodic box in the oracle; agreement with the exact-difference
term of FluidCuda.cu:326-341 up to O(F^2).

ChemBoundaryComplex (wall_scheme = 2, FluidCuda.cu:403-429): the oracle's level
fill against the explicit stencils it reduces to on plane walls -- each solid
layer is (2 v(x,y) + v(x+-1,y) + v(x,y+-1)) / 6 of the layer nearer the fluid
(weights 1/18 : 1/36 of Alpha, Common.h:100, normalised), both fluid sides when a
single layer touches fluid through the periodic z wrap -- and constant fields
kept."""
import numpy as np
import pytest


def tables(orc):
    ex, ey, ez, re, w, M, R = orc.tables()
    return np.stack([ex, ey, ez], 1).astype(float), w, M, R


def feq(rho, u, e, w):
    eu = e @ u
    return w * rho * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * u @ u)


def guo_dense(f, rho, F, ts, qp, e, w, M, R):
    """m* = m - S(m - M feq(rho, g)) + (I - S/2) M Fbar + S Qp, g = (j + F/2) / rho"""
    S = np.diag(ts)
    g = (e.T @ f + F / 2) / rho
    Fbar = w * (3 * ((e - g) @ F) + 9 * (e @ g) * (e @ F))
    m = M @ f
    ms = m - S @ (m - M @ feq(rho, g, e, w)) + (np.eye(19) - S / 2) @ (M @ Fbar) + S @ qp
    return R @ ms


def test_guo_collide_matches_dense_form(orc, cpmp):
    e, w, M, R = tables(orc)
    rng = np.random.default_rng(17)
    for trial in range(40):
        ts = np.array(orc.reference_ts(rng.uniform(0.6, 1.5)))
        if trial % 3 == 2:
            ts = rng.uniform(0.5, 1.8, 19)        # general rates
        rho = rng.uniform(0.05, 8)
        u = rng.uniform(-0.1, 0.1, 3)
        F = rng.uniform(-0.02, 0.02, 3) * rho
        f = feq(rho, u, e, w) * (1 + 0.05 * rng.standard_normal(19))
        rho_s = f.sum()
        qp = np.zeros(19)
        if trial % 2:
            qp[[1, 9, 11, 13, 14, 15]] = rng.uniform(-1e-3, 1e-3, 6)
        got, gu = cpmp.selftest_collide(f, rho_s, F, ts, qp if trial % 2 else None, forcing=1)
        want = guo_dense(f, rho_s, F, ts, qp, e, w, M, R)
        assert np.allclose(gu, e.T @ f / rho_s, rtol=1e-13, atol=1e-16)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-13 * rho)
        # mass kept, momentum raised by F, whatever the rates (Qp acts on rows 1, 9-15 only)
        assert got.sum() == pytest.approx(rho_s, rel=1e-13)
        assert np.allclose(e.T @ got, e.T @ f + F, rtol=0, atol=1e-13 * rho)


def test_guo_agrees_with_exact_difference_to_second_order(orc, cpmp):
    """one node: the post-collision states of the two force terms differ by O(F^2)"""
    e, w, M, R = tables(orc)
    rng = np.random.default_rng(4)
    ts = np.array(orc.reference_ts(0.8))
    ratios = []
    for _ in range(10):
        rho = rng.uniform(0.5, 8)
        u = rng.uniform(-0.05, 0.05, 3)
        f = feq(rho, u, e, w) * (1 + 0.02 * rng.standard_normal(19))
        d0 = rng.standard_normal(3) * 0.01 * rho
        diffs = []
        for s in (1.0, 0.5, 0.25):
            a, _ = cpmp.selftest_collide(f, f.sum(), s * d0, ts, forcing=1)
            b, _ = cpmp.selftest_collide(f, f.sum(), s * d0, ts, forcing=0)
            diffs.append(np.max(np.abs(a - b)))
        ratios += [diffs[0] / diffs[1], diffs[1] / diffs[2]]
    assert 3.5 < min(ratios) and max(ratios) < 4.5


def test_guo_lattice_mass_and_momentum(orc):
    """periodic box in the oracle with Guo forcing: mass conserved over 20 steps;
    one step raises the total momentum by the total force"""
    n = 12
    p = orc.make_params(n, n, n, eos="CSE", tr=0.85, Ka=0.01, nwall=0, forcing=1)
    o = orc.Oracle(p)
    rng = np.random.default_rng(3)
    o.init_from_rho(0.15 + 0.02 * rng.standard_normal(n ** 3))
    m0 = o.reduce()[0]
    o.step(20)
    assert abs(o.reduce()[0] - m0) < 1e-12 * m0
    e, w, M, R = tables(orc)
    j0 = e.T @ o.get(orc.F).reshape(19, -1).sum(1)
    o.step(1)
    F = o.get(orc.FORCE).reshape(3, -1)
    j1 = e.T @ o.get(orc.F).reshape(19, -1).sum(1)
    assert np.allclose(j1, j0 + F.sum(1), atol=1e-12)
    # MV = j/rho + F/(2 rho), the velocity Guo's equilibrium uses
    mv = o.get(orc.MV).reshape(3, -1)
    assert np.all(np.isfinite(mv))


def test_guo_lattice_close_to_exact_difference(orc):
    """a weakly perturbed periodic box: 10 steps of both force terms differ by
    O(eps^2) in the density perturbation amplitude eps"""
    n = 10
    rng = np.random.default_rng(8)
    pert = rng.standard_normal(n ** 3)
    d = []
    for eps in (2e-3, 1e-3):
        out = []
        for forcing in (0, 1):
            o = orc.Oracle(orc.make_params(n, n, n, eos="CSE", tr=0.85, Ka=0.01, nwall=0, forcing=forcing))
            o.init_from_rho(0.2 + eps * pert)
            o.step(10)
            out.append(o.get(orc.RHO))
        d.append(np.max(np.abs(out[0] - out[1])))
    assert d[1] > 0 and 3.0 < d[0] / d[1] < 5.0


def wall_case(orc, lo, hi, nx=6, ny=5, nz=14, seed=2):
    p = orc.make_params(nx, ny, nz, eos="PRW", tr=0.7, nwall=lo, nwall_hi=hi, wall_scheme=2)
    o = orc.Oracle(p)
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.1, 8, (nx, ny, nz))
    o.init_from_rho(rho.ravel())
    o.set_wall_mu(np.linspace(-0.05, 0.05, nx * ny), 0.02)
    o.stage(0)
    return rho, o.get(orc.RHO).reshape(nx, ny, nz), o.get(orc.MU).reshape(nx, ny, nz)


def plus_avg(v):
    """(2 v + v(x+-1) + v(y+-1)) / 6 of one (x, y) layer, periodic"""
    return (2 * v + np.roll(v, 1, 0) + np.roll(v, -1, 0) + np.roll(v, 1, 1) + np.roll(v, -1, 1)) / 6


@pytest.mark.parametrize("lo,hi", [(3, 3), (3, 2), (1, 4), (2, 0)])
def test_complex_wall_levels(orc, lo, hi):
    """each solid layer of level d (its distance to the fluid) is the plus-average of
    the adjacent layers of level d - 1 (one, or two through the periodic z wrap,
    e.g. the bottom level-3 layer and a 2-layer top wall)"""
    rho, out, mu = wall_case(orc, lo, hi)
    nz = rho.shape[2]

    def level(k):
        return lo - k if k < lo else (k - (nz - hi) + 1 if k >= nz - hi else 0)
    fluid = slice(lo, nz - hi)
    assert np.array_equal(out[:, :, fluid], rho[:, :, fluid])
    for k in list(range(lo)) + list(range(nz - hi, nz)):
        srcs = [kk % nz for kk in (k - 1, k + 1) if level(kk % nz) == level(k) - 1]
        assert srcs
        want = sum(plus_avg(out[:, :, kk]) for kk in srcs) / len(srcs)
        assert np.allclose(out[:, :, k], want, rtol=1e-14), (lo, hi, k, srcs)
    # solid mu is the wall field
    assert np.allclose(mu[:, :, 0].ravel(), np.linspace(-0.05, 0.05, mu.shape[0] * mu.shape[1]))
    if hi:
        assert np.all(mu[:, :, nz - 1] == 0.02)


def test_complex_wall_single_layer_touching_fluid_on_both_sides(orc):
    """lo = 1, hi = 0: the one solid layer z = 0 has fluid at z = 1 and, through the
    periodic wrap, at z = nz - 1; it averages both (10 neighbours of level 0)"""
    rho, out, _ = wall_case(orc, 1, 0)
    want = (2 * rho[:, :, 1] + sum(np.roll(rho[:, :, 1], s, a) for s in (1, -1) for a in (0, 1))
            + 2 * rho[:, :, -1] + sum(np.roll(rho[:, :, -1], s, a) for s in (1, -1) for a in (0, 1))) / 12
    assert np.allclose(out[:, :, 0], want, rtol=1e-14)


def test_complex_wall_keeps_constant_fields(orc):
    nx, ny, nz = 7, 4, 12
    p = orc.make_params(nx, ny, nz, eos="PRW", tr=0.7, nwall=3, nwall_hi=3, wall_scheme=2)
    o = orc.Oracle(p)
    o.init_from_rho(np.full(nx * ny * nz, 2.5))
    o.stage(0)
    assert np.allclose(o.get(orc.RHO), 2.5, rtol=1e-15)
