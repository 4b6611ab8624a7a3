"""Pins of the oracle's lattice tables (Common.h:94-146) and equilibrium."""
import numpy as np


def test_velocity_set_and_weights(orc):
    ex, ey, ez, re, w, M, R = orc.tables()
    e = np.stack([ex, ey, ez], 1)
    # 19 distinct velocities, opposite map is an involution with e[re] = -e
    assert len({tuple(v) for v in e}) == 19
    assert np.all(e[re] == -e)
    assert np.all(re[re] == np.arange(19))
    # Common.h:99 opposite map
    assert list(re) == [0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15]
    # isotropy of the weights (Common.h:100): sum w = 1, sum w e e = I/3, 4th moment 1/9
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose(np.einsum("q,qa,qb->ab", w, e, e), np.eye(3) / 3, atol=1e-15)
    assert abs(np.einsum("q,q,q->", w, ex ** 2, ey ** 2) - 1 / 9) < 1e-15


def test_moment_matrix_orthogonal_and_inverse(orc):
    ex, ey, ez, re, w, M, R = orc.tables()
    G = M @ M.T
    assert np.allclose(G - np.diag(np.diag(G)), 0)
    norms = [19, 2394, 252, 10, 40, 10, 40, 10, 40, 36, 72, 12, 24, 4, 4, 4, 8, 8, 8]
    assert np.array_equal(np.diag(G), norms)
    assert np.abs(R @ M - np.eye(19)).max() < 1e-15
    assert np.abs(M @ R - np.eye(19)).max() < 1e-14
    # spot entries as printed in Common.h:105-123 and :127-145
    assert list(M[1, :8]) == [-30, -11, -11, -11, -11, -11, -11, 8]
    assert list(M[16, 7:15]) == [1, -1, 1, -1, -1, 1, -1, 1]
    assert list(M[17, 15:19]) == [1, -1, 1, -1]
    assert R[0, 1] == -5.0 / 399 and R[1, 1] == -11.0 / 2394 and R[7, 2] == 1.0 / 252
    assert R[7, 13] == 1.0 / 4 and R[7, 16] == 1.0 / 8 and R[7, 17] == -1.0 / 8


def feq(rho, u, ex, ey, ez, w):
    dot = u[0] * ex + u[1] * ey + u[2] * ez
    return rho * w * (1 + 3 * dot + 4.5 * dot * dot - 1.5 * (u @ u))


def meq(rho, u):
    ux, uy, uz = u
    usq = u @ u
    g9 = 2 * ux * ux - uy * uy - uz * uz
    g11 = uy * uy - uz * uz
    return np.array([rho, rho * (-11 + 19 * usq), rho * (3 - 5.5 * usq), rho * ux, -2 / 3 * rho * ux,
                     rho * uy, -2 / 3 * rho * uy, rho * uz, -2 / 3 * rho * uz, rho * g9, -0.5 * rho * g9,
                     rho * g11, -0.5 * rho * g11, rho * ux * uy, rho * uy * uz, rho * ux * uz, 0, 0, 0])


def test_feq_moments_match_meq(orc):
    """M Feq(rho,u) = Meq(rho,u): the identity that lets the exact-difference force
    term (FluidCuda.cu:341) be applied in moment space."""
    ex, ey, ez, re, w, M, R = orc.tables()
    rng = np.random.default_rng(0)
    for _ in range(20):
        rho = rng.uniform(0.01, 9)
        u = rng.uniform(-0.2, 0.2, 3)
        f = feq(rho, u, ex, ey, ez, w)
        assert np.allclose(M @ f, meq(rho, u), rtol=1e-13, atol=1e-13)
        e = np.stack([ex, ey, ez], 1)
        assert abs(f.sum() - rho) < 1e-13 * rho
        assert np.allclose(e.T @ f, rho * u, atol=1e-14)
        assert np.allclose(np.einsum("q,qa,qb->ab", f, e, e), rho * (np.eye(3) / 3 + np.outer(u, u)), atol=1e-13)
