"""Closed-form pins of the oracle's contact-angle measurement, reference
initialisation and noise generator (CPU only).

* ContactAngle (FluidCuda.cu:557-689): synthetic density planes whose interface
  positions are known exactly give the angles, foot positions, base L and
  height H in closed form: wedges (two straight contact lines of angle theta,
  linear density profiles, so the scan's linear interpolation is exact) and the
  cap angle from H and L through tan(theta/2) = H / (L/2).  Rows outside the two
  the method uses are displaced, so a wrong row would show.  The L <= 0 branch
  and the DY/2 guard of FluidCuda.cu:670-672 are forced.  Spherical caps of known
  angle with a sharp and a tanh profile are measured within the method's own
  discretisation bound: the secant through rows 1 and 2 above the foot is the
  tangent at height ~1.5, i.e. theta - 1.5/(R sin theta) for a circle of radius R.
* SetFlowField (FluidCuda.cu:199-218, 236-239): a NumPy restatement with the
  radius rounded through float32 (the reference applies sqrtf).
* SplitMix64: the published output sequence of the generator (Steele, Lea and
  Flood 2014; the seed 1234567 sequence of the Rosetta Code task), and the
  uniformity of the [-1, 1) map built on it.
"""
import numpy as np
import pytest

import oracle_lib as orc

DG, DL = 0.1, 2.0
MID = (DG + DL) / 2


def plane_field(nx, ny, nz, fn):
    """rho(i, k) = fn(i, k) on every y plane"""
    i, k = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(nz, dtype=np.float64), indexing="ij")
    return np.repeat(fn(i, k)[:, None, :], ny, axis=1).ravel()


def wedge(xc, a, theta, ztop, base, g=0.37, rows=None):
    """two contact lines of angle theta through (xc -+ a, base), a flat top at ztop,
    linear profile of slope g inside; rows: {k: (dl, dr)} displacements of the
    interface positions on single rows"""
    cot = 1.0 / np.tan(np.radians(theta)) if theta != 90 else 0.0

    def fn(i, k):
        xl = xc - a + (k - base) * cot
        xr = xc + a - (k - base) * cot
        for kk, (dl_, dr_) in (rows or {}).items():
            xl = np.where(k == kk, xl + dl_, xl)
            xr = np.where(k == kk, xr + dr_, xr)
        return MID + g * np.minimum(np.minimum(i - xl, xr - i), ztop - k)
    return fn


@pytest.mark.parametrize("lo", [3, 2, 4])
@pytest.mark.parametrize("theta", [30, 60, 90, 120, 150])
def test_contact_angle_exact_on_wedges(theta, lo):
    """CaLeft = CaRight = theta, feet at xc -+ a, L = 2a, H = ztop - base and
    CaCap = 2 atan(H / a) to rounding; the reference geometry is lo = 3 (rows 2..5),
    other wall depths shift the rows with the wall"""
    nx, ny, nz = 64, 3, 72
    base = lo - 1                       # the top solid layer, the row the foot is extrapolated to
    xc, a = 31.3, 12.6
    H = a * np.tan(np.radians(theta) / 2)    # cap angle theta as well
    rows = {base: (0.41, -0.23), base + 3: (-0.17, 0.31)}   # rows the angle must not use
    p = orc.make_params(nx, ny, nz, eos="PRW", tr=0.7, nwall=lo)
    rho = plane_field(nx, ny, nz, wedge(xc, a, theta, base + H, base, rows=rows))
    out = orc.contact_angle(p, rho, DG, DL, 1)
    ca_l, ca_r, ca_cap, xl, xr, h, L, valid = out
    assert valid == 1
    assert ca_l == pytest.approx(theta, abs=1e-9)
    assert ca_r == pytest.approx(theta, abs=1e-9)
    assert xl == pytest.approx(xc - a, abs=1e-12)
    assert xr == pytest.approx(xc + a, abs=1e-12)
    assert L == pytest.approx(2 * a, abs=1e-12)
    assert h == pytest.approx(H, abs=1e-12)
    assert ca_cap == pytest.approx(np.degrees(2 * np.arctan(H / a)), abs=1e-9)
    assert ca_cap == pytest.approx(theta, abs=1e-9)


def test_contact_angle_crossed_feet_and_dy_guard():
    """extrapolated feet that cross (L <= 0): L falls back to the row-1 width and H
    loses half a cell (FluidCuda.cu:670); when the crossed feet straddle DY/2 they
    are both set to DY/2 (FluidCuda.cu:672, DY where DX is meant)"""
    theta, xc, a = 150.0, 20.2, -0.5
    cot = 1.0 / np.tan(np.radians(theta))
    ztop = 2 + 9.3
    fn = wedge(xc, a, theta, ztop, 2)

    def fn2(i, k):   # row 2 gets crossings of its own (the scan needs them)
        return np.where(k == 2, MID + 0.37 * np.minimum(i - (xc - 1), (xc + 1) - i), fn(i, k))
    nx, nz = 40, 40              # the scan assumes a droplet near DX/2
    for ny, guard in ((40, True), (3, False)):     # DY/2 = 20 inside (19.7, 20.7), or 1
        p = orc.make_params(nx, ny, nz, eos="PRW", tr=0.7, nwall=3)
        out = orc.contact_angle(p, plane_field(nx, ny, nz, fn2), DG, DL, 1)
        ca_l, ca_r, ca_cap, xl, xr, h, L, valid = out
        w1 = 2 * (a - cot)                          # row-1 width xr(3) - xl(3)
        assert valid == 1
        assert ca_l == pytest.approx(theta, abs=1e-9) and ca_r == pytest.approx(theta, abs=1e-9)
        assert L == pytest.approx(w1, abs=1e-12)
        assert h == pytest.approx(ztop - 2 - 0.5, abs=1e-12)
        assert ca_cap == pytest.approx(np.degrees(2 * np.arctan((ztop - 2.5) / (w1 / 2))), abs=1e-9)
        if guard:
            assert xl == 20.0 and xr == 20.0
        else:
            assert xl == pytest.approx(xc - a, abs=1e-12) and xr == pytest.approx(xc + a, abs=1e-12)


def test_contact_angle_without_interface():
    p = orc.make_params(40, 6, 24, eos="PRW", tr=0.7, nwall=3)
    out = orc.contact_angle(p, np.full(40 * 6 * 24, DG), DG, DL, 2)
    assert list(out[[0, 1, 2, 6, 7]]) == [0, 0, 0, 0, 0]
    assert out[3] == 3 and out[4] == 3          # DY/2
    assert np.isnan(out[5])                     # H: no interface above the foot
    assert orc.contact_angle(orc.make_params(40, 6, 24, nwall=0), np.full(40 * 6 * 24, DG), DG, DL, 2) is None


def secant_angle(R, zc, base):
    """angle of the chord of a circle (centre height zc) between rows base+1 and
    base+2 at its left edge, degrees"""
    x1 = -np.sqrt(R * R - (base + 1 - zc) ** 2)
    x2 = -np.sqrt(R * R - (base + 2 - zc) ** 2)
    a = np.degrees(np.arctan(1.0 / (x2 - x1)))
    return a + 180 if a < 0 else a


@pytest.mark.parametrize("R", [30, 60])
@pytest.mark.parametrize("theta", [30, 60, 90, 120, 150])
@pytest.mark.parametrize("profile", ["sharp", "tanh"])
def test_contact_angle_spherical_caps(theta, R, profile):
    """a sphere of radius R cut by the wall plane at angle theta: CaLeft/CaRight are
    the chord angle through rows 1 and 2 above the foot (exact for a sharp linear
    profile up to the interpolation of the curved interface) and within
    1.5 / (R sin theta) rad + a profile allowance of theta; CaCap within 1 degree
    at R = 30, 0.25 at R = 60"""
    nx, ny, nz = 2 * R + 40, 3, 2 * R + 20
    base = 2
    xc, zc = nx / 2 + 0.3, base - R * np.cos(np.radians(theta))
    x, z = np.meshgrid(np.arange(nx, dtype=float), np.arange(nz, dtype=float), indexing="ij")
    r = np.sqrt((x - xc) ** 2 + (z - zc) ** 2)
    if profile == "sharp":
        rho2 = MID + 0.3 * (R - r)
        allow = 0.0
    else:
        rho2 = MID - (DL - DG) / 2 * np.tanh(2 * (r - R) / 2.5)
        allow = 1.5
    rho = np.repeat(rho2[:, None, :], ny, axis=1).ravel()
    p = orc.make_params(nx, ny, nz, eos="PRW", tr=0.7, nwall=3)
    ca_l, ca_r, ca_cap, xl, xr, h, L, valid = orc.contact_angle(p, rho, DG, DL, 1)
    assert valid == 1
    bound = np.degrees(1.5 / (R * np.sin(np.radians(theta)))) * 1.25 + 0.2 + allow
    assert abs(ca_l - theta) < bound and abs(ca_r - theta) < bound
    assert ca_l < theta and ca_r < theta       # the chord sits above the foot
    if profile == "sharp":
        sec = secant_angle(R, zc, base)
        assert abs(ca_l - sec) < 0.1 and abs(ca_r - sec) < 0.1
    assert abs(ca_cap - theta) < (1.0 if R == 30 else 0.25)
    assert h == pytest.approx(zc + R - base, abs=0.6)


def test_init_reference_matches_closed_form():
    """SetFlowField CONTACTANGLE: integer DX/2, DY/2, centre height radius + 2, r
    through float32 sqrt, tanh profile; solid rho = mu = the initial wall value,
    fluid mu = 0, f = w_q rho"""
    nx, ny, nz, lo = 21, 17, 30, 3
    R, W, base_val = 7.3, 5.0, 0.013
    p = orc.make_params(nx, ny, nz, eos="PRW", tr=0.7, nwall=lo)
    o = orc.Oracle(p)
    o.init_reference(R, W, DG, DL, base_val)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    sq = (i - nx // 2) ** 2.0 + (j - ny // 2) ** 2.0 + (k - (R + 2)) ** 2.0
    r32 = np.sqrt(sq.astype(np.float32)).astype(np.float64)
    want = (DL + DG) / 2 - (DL - DG) / 2 * np.tanh((r32 - R) * 2 / W)
    solid = (k < lo) | (k >= nz - lo)
    want[solid] = base_val
    rho = o.get(orc.RHO).reshape(nx, ny, nz)
    assert np.allclose(rho, want, rtol=2e-15, atol=0)
    # the float32 radius matters at this tolerance: a double radius misses
    r64 = np.sqrt(sq)
    alt = (DL + DG) / 2 - (DL - DG) / 2 * np.tanh((r64 - R) * 2 / W)
    assert np.max(np.abs(alt - want)[~solid]) > 1e-9
    mu = o.get(orc.MU).reshape(nx, ny, nz)
    assert np.all(mu[solid] == base_val) and np.all(mu[~solid] == 0)
    _, _, _, _, w, _, _ = orc.tables()
    f = o.get(orc.F).reshape(19, -1)
    assert np.allclose(f, w[:, None] * rho.ravel()[None, :], rtol=1e-15, atol=0)


# published SplitMix64 outputs for the state 1234567 (Rosetta Code, "Pseudo-random
# numbers/Splitmix64")
SPLITMIX_1234567 = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                    4593380528125082431, 16408922859458223821]
GAMMA = 0x9E3779B97F4A7C15


def test_splitmix64_published_sequence():
    state = 1234567
    for want in SPLITMIX_1234567:
        assert orc.splitmix64(state) == want
        state = (state + GAMMA) % 2 ** 64


def test_noise_is_uniform_on_minus_one_one():
    """noise(seed, i) = 2 U - 1 with U the top 53 bits of splitmix64(seed ^
    splitmix64(i)) over 2^53: on the 2^-52 grid, in [-1, 1), mean 0, variance 1/3,
    flat histogram; seeds give independent streams"""
    n = 200_000
    for seed in (1, 3, 0xDEADBEEF):
        v = np.array([orc.noise(seed, i) for i in range(n)])
        assert np.all(v >= -1) and np.all(v < 1)
        assert np.all((v + 1) * 2.0 ** 52 == np.floor((v + 1) * 2.0 ** 52))
        assert abs(v.mean()) < 5 * np.sqrt(1 / 3 / n)
        assert abs(v.var() - 1 / 3) < 0.01
        hist, _ = np.histogram(v, bins=20, range=(-1, 1))
        chi2 = np.sum((hist - n / 20) ** 2 / (n / 20))
        assert chi2 < 60            # 19 degrees of freedom
        for i in (0, 1, 12345, n - 1):
            u = orc.splitmix64(seed ^ orc.splitmix64(i)) >> 11
            assert orc.noise(seed, i) == u * 2.0 ** -53 * 2 - 1
    a = np.array([orc.noise(1, i) for i in range(2000)])
    b = np.array([orc.noise(2, i) for i in range(2000)])
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.1
