"""Pins of the oracle's step stages (FluidCuda.cu:351-463) by closed forms,
conservation laws and brute force."""
import numpy as np
import pytest


def grid(nx, ny, nz):
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return x.astype(float), y.astype(float), z.astype(float)


def test_laplacian_exact_to_degree_5(orc):
    """GradD5 (Common.h:189-192) is exact for polynomials of degree <= 5.
    With K = 0 the mu stage returns -Ka * Laplacian(rho)."""
    n = 16
    p = orc.make_params(n, n, n, eos="PRW", tr=0.7, K=0.0, Ka=1.0, nwall=0)
    o = orc.Oracle(p)
    x, y, z = grid(n, n, n)
    x, y, z = x - 8, y - 8, z - 8
    poly = 1e-5 * (x ** 5 + 2 * y ** 4 * z - x * y * z ** 3 + 3 * x ** 2 * y ** 2 + z ** 2) + 2.0
    lap = 1e-5 * (20 * x ** 3 + 24 * y ** 2 * z - 6 * x * y * z + 6 * y ** 2 + 6 * x ** 2 + 2)
    o.init_from_rho(poly.ravel())
    o.stage(1)
    mu = o.get(orc.MU).reshape(n, n, n)
    inner = (slice(2, -2),) * 3          # away from the periodic wrap
    assert np.abs(-mu[inner] - lap[inner]).max() < 1e-12


def test_laplacian7_exact_to_degree_7(orc):
    n = 18
    p = orc.make_params(n, n, n, eos="PRW", tr=0.7, K=0.0, Ka=1.0, nwall=0, lap_points=7)
    o = orc.Oracle(p)
    x, y, z = grid(n, n, n)
    x, y, z = x - 9, y - 9, z - 9
    poly = 1e-7 * (x ** 7 + y ** 6 - z ** 4 * x ** 2) + 2.0
    lap = 1e-7 * (42 * x ** 5 + 30 * y ** 4 - 12 * z ** 2 * x ** 2 - 2 * z ** 4)
    o.init_from_rho(poly.ravel())
    o.stage(1)
    mu = o.get(orc.MU).reshape(n, n, n)
    inner = (slice(3, -3),) * 3
    assert np.abs(-mu[inner] - lap[inner]).max() < 1e-10


def test_gradient_exact_to_degree_6(orc):
    """Gradx7/y7/z7 (Common.h:222-241) are exact for polynomials of degree <= 6.
    With mu = 0 the force stage returns grad(rho)/3."""
    n = 18
    p = orc.make_params(n, n, n, eos="PRW", tr=0.7, nwall=0)
    o = orc.Oracle(p)
    x, y, z = grid(n, n, n)
    x, y, z = x - 9, y - 9, z - 9
    c = 1e-6
    rho = c * (x ** 6 + y ** 5 * z - 2 * x ** 3 * y ** 2 * z + z ** 4 + x * y) + 3.0
    gx = c * (6 * x ** 5 - 6 * x ** 2 * y ** 2 * z + y)
    gy = c * (5 * y ** 4 * z - 4 * x ** 3 * y * z + x)
    gz = c * (y ** 5 - 2 * x ** 3 * y ** 2 + 4 * z ** 3)
    o.init_from_rho(rho.ravel())
    o.set(orc.MU, np.zeros(n ** 3))
    o.stage(2)
    F = o.get(orc.FORCE).reshape(3, n, n, n) * 3
    inner = (slice(3, -3),) * 3
    for a, g in enumerate((gx, gy, gz)):
        assert np.abs(F[a][inner] - g[inner]).max() < 1e-11
    # the x^7 term is not reproduced: 6th order, error 36*h^6*d7/7! at the origin
    rho7 = 1e-6 * x ** 7 + 3.0
    o.init_from_rho(rho7.ravel())
    o.set(orc.MU, np.zeros(n ** 3))
    o.stage(2)
    F = o.get(orc.FORCE).reshape(3, n, n, n) * 3
    err = F[0][9, 9, 9] - 0.0
    assert abs(err / 1e-6 - 36.0) < 1e-6


def test_wall_fill_constant_and_linear(orc):
    """The layered wall fill (FluidCuda.cu:376-401) reproduces constant fields and
    fields linear in (x,y); solid mu is the wall field."""
    nx, ny, nz = 12, 10, 14
    p = orc.make_params(nx, ny, nz, eos="PRW", tr=0.7, nwall=3)
    o = orc.Oracle(p)
    x, y, z = grid(nx, ny, nz)
    rho = 1.5 + 0.01 * x + 0.02 * y + 0.3 * z
    o.init_from_rho(rho.ravel())
    lo = np.linspace(-0.1, 0.1, nx * ny)
    o.set_wall_mu(lo, -lo)
    o.stage(0)
    r = o.get(orc.RHO).reshape(nx, ny, nz)
    mu = o.get(orc.MU).reshape(nx, ny, nz)
    inner = (slice(3, -3), slice(3, -3))
    for k in range(3):
        assert np.allclose(r[inner + (k,)], rho[inner + (3,)], atol=1e-14)
        assert np.allclose(r[inner + (nz - 1 - k,)], rho[inner + (nz - 4,)], atol=1e-14)
        assert np.allclose(mu[:, :, k].ravel(), lo)
        assert np.allclose(mu[:, :, nz - 1 - k].ravel(), -lo)
    # layered: layer 0 is the third in-plane average of layer 3 (a 7x7 stencil)
    rng = np.random.default_rng(1)
    rho2 = rng.uniform(1, 2, (nx, ny, nz))
    o.init_from_rho(rho2.ravel())
    o.stage(0)
    r = o.get(orc.RHO).reshape(nx, ny, nz)

    def avg(a):
        return (2 * a + np.roll(a, (1, 1), (0, 1)) + np.roll(a, (1, -1), (0, 1)) + np.roll(a, (-1, 1), (0, 1))
                + np.roll(a, (-1, -1), (0, 1))) / 6
    a2 = avg(rho2[:, :, 3])
    a1 = avg(a2)
    a0 = avg(a1)
    assert np.allclose(r[:, :, 2], a2, atol=1e-14)
    assert np.allclose(r[:, :, 1], a1, atol=1e-14)
    assert np.allclose(r[:, :, 0], a0, atol=1e-14)
    # simple scheme: all layers from the first fluid layer
    p.wall_scheme = 0
    o2 = orc.Oracle(p)
    o2.init_from_rho(rho2.ravel())
    o2.stage(0)
    r2 = o2.get(orc.RHO).reshape(nx, ny, nz)
    for k in range(3):
        assert np.allclose(r2[:, :, k], a2, atol=1e-14)


def test_rest_state_is_fixed_point(orc):
    n = 10
    p = orc.make_params(n, n, n, eos="PRW", tr=0.7, nwall=0)
    o = orc.Oracle(p)
    o.init_from_rho(np.full(n ** 3, 2.5))
    f0 = o.get(orc.F)
    o.step(5)
    assert np.abs(o.get(orc.F) - f0).max() < 1e-14
    assert np.abs(o.get(orc.RHO) - 2.5).max() < 1e-14


@pytest.mark.parametrize("nwall,eos,tr,g,l", [(0, "CSE", 0.85, 0.0315818, 0.279262),
                                              (3, "PRW", 0.7, 0.0556211, 8.08045)])
def test_mass_conserved(orc, nwall, eos, tr, g, l):
    """Collision conserves mass, streaming with bounce-back permutes populations."""
    nx, ny, nz = 16, 12, 18
    p = orc.make_params(nx, ny, nz, eos=eos, tr=tr, nwall=nwall, Ka=0.01 if eos == "CSE" else 0.001)
    o = orc.Oracle(p)
    o.set_wall_mu(-0.02)
    o.init_reference(radius=5, width=4, den_g=g, den_l=l)
    m0 = o.reduce()[0]
    o.step(20)
    m1 = o.reduce()[0]
    assert abs(m1 - m0) < 1e-11 * m0


def test_momentum_balance_one_step(orc):
    """Periodic box: sum_x sum_q e f^{n+1} = sum_x (rho u + F) (EDM adds F)."""
    n = 12
    p = orc.make_params(n, n, n, eos="CSE", tr=0.85, Ka=0.01, nwall=0)
    o = orc.Oracle(p)
    rng = np.random.default_rng(3)
    rho = 0.15 + 0.02 * rng.standard_normal(n ** 3)
    o.init_from_rho(rho)
    o.step(3)
    ex, ey, ez, re, w, M, R = orc.tables()
    e = np.stack([ex, ey, ez], 1).astype(float)
    f = o.get(orc.F).reshape(19, -1)
    j0 = e.T @ f.sum(1)
    o.step(1)
    F = o.get(orc.FORCE).reshape(3, -1)
    f1 = o.get(orc.F).reshape(19, -1)
    j1 = e.T @ f1.sum(1)
    assert np.allclose(j1, j0 + F.sum(1), atol=1e-12)


def test_collide_matches_dense_brute_force(orc, cpmp):
    """The library's sparse moment-space collision (run on the CPU through the
    self-test entry) against a dense R (I-S) M f + S Meq + EDM written from the
    reference formula, on random states; and K1 != 0 through Qp."""
    ex, ey, ez, re, w, M, R = orc.tables()
    rng = np.random.default_rng(7)
    from test_oracle_tables import feq, meq
    ts = np.array(orc.reference_ts(0.7))
    S = np.diag(ts)
    for trial in range(50):
        rho = rng.uniform(0.05, 8)
        u = rng.uniform(-0.1, 0.1, 3)
        F = rng.uniform(-0.01, 0.01, 3)
        f = feq(rho, u, ex, ey, ez, w) * (1 + 0.05 * rng.standard_normal(19))
        qp = np.zeros(19)
        if trial % 2:
            qp[[1, 9, 11, 13, 14, 15]] = rng.uniform(-1e-3, 1e-3, 6)
        rho_s = rho * (1 + 1e-3 * rng.standard_normal())   # stored density (MacroCalculate)
        e = np.stack([ex, ey, ez], 1)
        u = e.T @ f / rho_s                                 # V of MacroCalculate, FluidCuda.cu:364-366
        m = M @ f
        dense = (R @ (m - S @ (m - meq(rho_s, u)) + S @ qp) + feq(rho_s, u + F / rho_s, ex, ey, ez, w)
                 - feq(rho_s, u, ex, ey, ez, w))
        got, gu = cpmp.selftest_collide(f, rho_s, F, ts, qp if trial % 2 else None)
        assert np.allclose(gu, u, rtol=1e-13, atol=1e-16)
        assert np.allclose(got, dense, rtol=1e-12, atol=1e-13 * rho)


def test_bounce_back_closed_box_conserves_mass(orc):
    """Walls on both z ends, random initial densities: mass is conserved exactly
    up to rounding (half-way bounce-back)."""
    nx, ny, nz = 8, 8, 12
    p = orc.make_params(nx, ny, nz, eos="CSE", tr=0.85, Ka=0.01, nwall=3)
    o = orc.Oracle(p)
    rng = np.random.default_rng(5)
    o.init_from_rho(0.15 + 0.01 * rng.standard_normal(nx * ny * nz))
    o.set_wall_mu(0.0)
    m0 = o.reduce()[0]
    o.step(30)
    assert abs(o.reduce()[0] - m0) < 1e-12 * m0


def test_reduce_exact_on_integers(orc):
    n = 6
    p = orc.make_params(n, n, n, eos="PRW", tr=0.7, nwall=1)
    o = orc.Oracle(p)
    rho = np.arange(n ** 3, dtype=float) % 7 + 1
    o.init_from_rho(rho)
    mass, bad, ms = o.reduce()
    r = rho.reshape(n, n, n)
    assert mass == r[:, :, 1:-1].sum() and bad == 0
    rho[10] = np.nan
    rho[13] = -1          # z = 13 % 6 = 1, a fluid node
    o.init_from_rho(rho)
    assert o.reduce()[1] == 2


def test_determinism(orc):
    n = 12
    p = orc.make_params(n, n, n + 6, eos="PRW", tr=0.7, nwall=3)
    outs = []
    for _ in range(2):
        o = orc.Oracle(p)
        o.set_wall_mu(-0.03)
        o.init_reference(4, 3, 0.0556211, 8.08045)
        o.step(10)
        outs.append(o.get(orc.RHO))
    assert np.array_equal(outs[0], outs[1])
