"""The march configuration that large lattices run, against the oracle.

Lattices of 256^3 and 512x512x256 select x-march workgroups of 32 planes in every
storage form, 128x128x96 and similar lattices 4.  Here lattices with nx >= 64 run
with chunks of 4, 8, 16 and 32 planes (CPMP_XC), so the 8-slot LDS ring rotates
over long chunks, the chemical-potential march alternates its two register sets
over many planes and the chunk-end partial planes of the summed density are
exercised, in the FP64 form (<= 1e-10 relative on rho after 200 steps), the FP32
form with the fused density and the completion folded into the next
chemical-potential kernel, and the FP16 form; then every EOS in the FP32 form
(including the FP32 bulk chemical potential), and the large van der Waals
striped-wall wetting configuration at a size the oracle covers, with its
adjoint."""
import contextlib
import os

import numpy as np
import pytest

import gpu_cases as gc
import oracle_lib as orc

pytestmark = pytest.mark.gpu

TOL64 = 1e-10

GEOMS = {
    # periodic RK coalescence: two chunks of 32, 2 y tiles, 2 z tiles
    "periodic": dict(dims=(64, 16, 64), eos="RKE", tr=0.60, nwall=0, init="spheres",
                     spheres=[(20, 8, 31.5, 10), (43, 8, 33, 10)], width=5, noise=1e-4, seed=3, steps=200),
    # PR droplet on 3 + 3 wall layers (the reference geometry)
    "walled": dict(dims=(64, 16, 64), eos="PRW", tr=0.70, nwall=3, init="reference", radius=13, width=5,
                   mu_w=-0.02, steps=200),
    # the periodic case with the 7-point Laplacian (PTCALCSEVENPOINT, Common.h:243-246):
    # halo of 3 in the chemical-potential march and in the folded completion
    "periodic7": dict(dims=(64, 16, 64), eos="RKE", tr=0.60, nwall=0, init="spheres",
                      spheres=[(20, 8, 31.5, 10), (43, 8, 33, 10)], width=5, noise=1e-4, seed=3, steps=200,
                      lap_points=7),
}


@contextlib.contextmanager
def env(**kv):
    old = {k: os.environ.get(k) for k in kv}
    try:
        for k, v in kv.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = str(v)
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


_ORACLE = {}


def oracle_run(name):
    """oracle state of a geometry after its steps (computed once per session)"""
    if name not in _ORACLE:
        case = GEOMS[name]
        # gc.build sets up the oracle exactly as the device lattice
        o, _ = gc.build(case, _NoDevice())
        fluid = o.get(orc.TYPE) == 1
        rho0 = o.get(orc.RHO)
        o.step(case["steps"])
        _ORACLE[name] = (rho0, o.get(orc.RHO), o.get(orc.F), fluid)
    return _ORACLE[name]


class _NoDevice:
    """stands in for the device package when only the oracle side of gc.build is wanted"""

    @staticmethod
    def make_params(*a, **k):
        return None

    class Lattice:
        def __init__(self, p):
            pass

        def set_wall_mu(self, *a):
            pass

        def init_reference(self, *a):
            pass

        def init_spheres(self, *a, **k):
            pass


def device(name, cpmp, form, **envkv):
    case = GEOMS[name]
    with env(**envkv):
        _, g = gc.build(case, cpmp, form)
    return g


@pytest.mark.parametrize("xc", [8, 16, 32])
@pytest.mark.parametrize("name", list(GEOMS))
def test_fp64_march_long_chunks_match_oracle(cpmp, name, xc):
    case = GEOMS[name]
    print(f"[production] fp64 march {name} dims={case['dims']} xc={xc} steps={case['steps']}", flush=True)
    rho0, rho_o, f_o, fluid = oracle_run(name)
    g = device(name, cpmp, "fp64", CPMP_COLLIDE="march", CPMP_XC=xc)
    assert gc.rel_err(g.get("rho"), rho0, fluid) < 1e-13
    g.step(case["steps"])
    err = gc.rel_err(g.get("rho"), rho_o, fluid)
    print(f"[production] fp64 march {name} xc={xc}: max rel err rho {err:.3e}", flush=True)
    assert err < TOL64
    fo, fg = f_o.reshape(19, -1), g.get("f").reshape(19, -1)
    assert np.max(np.abs(fg[:, fluid] - fo[:, fluid])) < TOL64 * np.max(np.abs(fo[:, fluid]))
    g.close()


@pytest.mark.parametrize("xc", [4, 32])
@pytest.mark.parametrize("form,tol", [("fp32", 2e-4), ("fp16", 2e-2), ("fp32a", 2e-4)])
@pytest.mark.parametrize("name", list(GEOMS))
def test_reduced_forms_long_chunks_match_oracle(cpmp, name, form, tol, xc):
    """fused density with chunks of 4 and 32 planes (periodic FP32: the completion
    folded into the next step's chemical-potential kernel) against the FP64 oracle;
    the FP32-arithmetic form (FP64 storage) forced onto the march collide, which
    these lattices would otherwise run one thread per node as the FP64 form does"""
    case = GEOMS[name]
    print(f"[production] {form} {name} xc={xc} steps={case['steps']}", flush=True)
    rho0, rho_o, _, fluid = oracle_run(name)
    env = dict(CPMP_COLLIDE="march") if form == "fp32a" else {}
    g = device(name, cpmp, form, CPMP_XC=xc, **env)
    short = device(name, cpmp, form, CPMP_XC=2, **env)
    unfused = device(name, cpmp, form, CPMP_XC=xc, CPMP_FUSE=0, **env)
    if not gc.path_forced():
        assert g.memory() > unfused.memory()   # the fused path
    unfused.close()
    # rho is the sum of the pulled populations after one step and a run of five
    for n in (1, 5):
        g.step(n)
        short.step(n)
        rho, s = g.get("rho"), g.get("f").reshape(19, -1).sum(axis=0)
        assert np.max(np.abs(rho[fluid] - s[fluid])) < 2e-6 * np.max(np.abs(s[fluid]))
    g.step(case["steps"] - 6)
    short.step(case["steps"] - 6)
    rl, rs = g.get("rho"), short.get("rho")
    err = np.max(np.abs(rl[fluid] - rho_o[fluid])) / np.max(rho_o[fluid])
    d_short = np.max(np.abs(rl[fluid] - rs[fluid])) / np.max(rs[fluid])
    print(f"[production] {form} {name} xc={xc}: max |rho - rho_oracle| / max rho {err:.3e}; "
          f"vs chunks of 2 {d_short:.3e}", flush=True)
    assert err < tol
    assert d_short < tol / 4
    m = g.reduce().mass
    m_short = short.reduce().mass
    assert abs(m - m_short) <= 1e-6 * m_short
    g.close()
    short.close()


@pytest.mark.parametrize("name", ["periodic", "periodic7"])
def test_folded_completion_long_chunks_is_bitwise_the_same(cpmp, name):
    """periodic FP32 with chunks of 32 (5- and 7-point Laplacian): the completion
    folded into the next mu kernel leaves the same bits as the completion pass
    after every step"""
    print(f"[production] folded completion {name}", flush=True)
    a = device(name, cpmp, "fp32", CPMP_XC=32)
    b = device(name, cpmp, "fp32", CPMP_XC=32)
    for n in (1, 4, 9):
        with env(CPMP_DEFER=0):
            b.step(n)
        a.step(n)
        for field in ("rho", "mu", "f"):
            assert np.array_equal(a.get(field), b.get(field)), (n, field)
    a.close()
    b.close()


# ---------------------------------------------------------------------------
# every EOS in the FP32 form


def eos_term_scale(eos, A, B, T, d):
    """sum of the magnitudes of the additive terms of the bulk chemical potential
    (the scale of its FP32 rounding error, not its value)"""
    bd = B * d
    if eos in ("PRW", "PRM"):
        s2 = np.sqrt(2.0)
        return (np.abs(T * np.log(d / (1 - bd))) + np.abs(A / (2 * s2 * B) * np.log((s2 - 1 + bd) / (s2 + 1 - bd)))
                + T / (1 - bd) + A * d / np.abs(1 + 2 * bd - bd * bd))
    if eos == "VDW":
        return np.abs(T * np.log(d / (1 - bd))) + T / (1 - bd) + 2 * A * d
    if eos in ("RKE", "RKS"):
        return np.abs(T * np.log(d / (1 - bd))) + T / (1 - bd) + A / B * np.log(1 + bd) + A * d / (1 + bd)
    om = 1 - bd
    return T * (np.abs(np.log(d)) + (8 * bd + 9 * bd * bd + 3 * bd ** 3) / om ** 3) + 2 * A * d


EOS_TRS = [("VDW", 0.80), ("CSE", 0.85), ("RKE", 0.60), ("RKS", 0.70), ("PRW", 0.70), ("PRM", 0.70)]


@pytest.mark.parametrize("eos,tr", EOS_TRS)
def test_fp32_form_every_eos(cpmp, eos, tr):
    """each EOS with its own chemical potential on a walled droplet: the FP32 form
    against the FP64 oracle after 150 steps; the FP32 chemical potential of the
    bulk (hardware log and reciprocal, outside the interface band) and of the
    interface (FP64 EOS part) against the FP64 form's of the same FP32 density"""
    case = dict(dims=(32, 16, 32), eos=eos, tr=tr, nwall=3, init="reference", radius=9, width=5,
                mu_w=0.0, steps=150, Ka=0.01 if eos == "CSE" else 0.001)
    o, g = gc.build(case, cpmp, "fp32")
    fluid = o.get(orc.TYPE) == 1
    o.step(case["steps"])
    g.step(case["steps"])
    err = np.max(np.abs(g.get("rho")[fluid] - o.get(orc.RHO)[fluid])) / np.max(o.get(orc.RHO)[fluid])
    # chemical potential of the FP32 density field, FP32 form vs FP64 form
    r = g.get("rho")
    g.step(1)
    mu32 = g.get("mu")
    ref = cpmp.Lattice(cpmp.make_params(*case["dims"], eos=eos, tr=tr, nwall=3, Ka=case["Ka"]))
    ref.set_wall_mu(0.0)
    ref.init_from_rho(r)
    ref.step(1)
    mu64 = ref.get("mu")
    ref.close()
    p = g.p
    A, B, T, K2 = p.A, p.B, p.T, p.K ** 2
    dg, dl = cpmp.coexistence(eos, A, B, T)
    lo, hi = np.float32(dg + 0.02 * (dl - dg)), np.float32(dl - 0.02 * (dl - dg))   # the library's band
    z = np.tile(np.arange(case["dims"][2]), case["dims"][0] * case["dims"][1])
    inner = (z >= 5) & (z < case["dims"][2] - 5)
    rf = r.astype(np.float32)
    band = (rf > lo) & (rf < hi) & inner
    bulk = ~band & inner
    scale = K2 * eos_term_scale(eos, A, B, T, np.minimum(r, 9.7)) + 1e-3 * np.abs(mu64)
    bulk_err = np.max(np.abs(mu32 - mu64)[bulk] / scale[bulk])
    band_err = np.max(np.abs(mu32 - mu64)[band] / np.abs(mu64[band]).max()) if band.any() else 0.0
    print(f"[eos-fp32] {eos} Tr={tr}: rho err {err:.3e}; bulk mu err {bulk_err:.3e} of the term scale "
          f"({bulk.sum()} nodes); interface mu err {band_err:.3e} ({band.sum()} nodes)", flush=True)
    assert err < 2e-4
    assert bulk.sum() > 1000
    assert bulk_err < 5e-6
    assert band_err < 1e-6
    g.close()


# ---------------------------------------------------------------------------
# the large striped-wall wetting configuration (van der Waals, Tr 0.8, FP32) at a
# size the oracle covers: 60 / 120 degree stripes of period 16 instead of 64, a
# hemisphere of radius 14 instead of 100

MU60, MU120 = -0.0270, 0.0473   # wall potentials of 60 and 120 degrees for this EOS (calibrated on a sessile droplet)


def stripes(nx, ny, period):
    x = np.repeat(np.arange(nx), ny)
    return np.where((x // (period // 2)) % 2 == 0, MU60, MU120).astype(np.float64)


STRIPED = dict(dims=(64, 32, 64), eos="VDW", tr=0.80, nwall=3, init="spheres",
               spheres=[(32, 16, 2.5, 14)], width=5, steps=300)


def striped_pair(cpmp, form, dims=None, **envkv):
    case = dict(STRIPED)
    if dims:   # the hemisphere stays centred and clear of the periodic x, y boundaries
        case["dims"] = dims
        case["spheres"] = [(dims[0] / 2, dims[1] / 2, 2.5, min(dims[0], dims[1]) / 3)]
    nx, ny, nz = case["dims"]
    with env(**envkv):
        o, g = gc.build(dict(case, mu_w=0.0), cpmp, form)
    mw = stripes(nx, ny, 16)
    o.set_wall_mu(mw, mw)
    g.set_wall_mu(mw, mw)
    return case, o, g


def test_striped_wall_vdw_fp64_and_fp32_against_oracle(cpmp):
    case, o, g64 = striped_pair(cpmp, "fp64")
    _, _, g32 = striped_pair(cpmp, "fp32")
    _, _, g32l = striped_pair(cpmp, "fp32", CPMP_XC=32)
    fluid = o.get(orc.TYPE) == 1
    n = case["steps"]
    print(f"[striped] vdW {case['dims']} {n} steps: oracle, fp64, fp32 (auto), fp32 (xc=32)", flush=True)
    o.step(n)
    for g in (g64, g32, g32l):
        g.step(n)
    ro = o.get(orc.RHO)
    e64 = gc.rel_err(g64.get("rho"), ro, fluid)
    e32 = np.max(np.abs(g32.get("rho")[fluid] - ro[fluid])) / np.max(ro[fluid])
    e32l = np.max(np.abs(g32l.get("rho")[fluid] - ro[fluid])) / np.max(ro[fluid])
    mo = o.reduce()[0]
    drift = abs(g32l.reduce().mass - mo) / mo
    print(f"[striped] fp64 {e64:.3e}; fp32 {e32:.3e}; fp32 xc=32 {e32l:.3e}; fp32 mass vs oracle {drift:.3e}",
          flush=True)
    assert e64 < TOL64
    assert e32 < 2e-4 and e32l < 2e-4
    assert drift < 1e-5
    for g in (g64, g32, g32l):
        g.close()


def test_striped_wall_vdw_fp32_adjoint(cpmp):
    """the FP32 adjoint of the striped-wall run against the oracle's FP64 reverse mode
    and against central differences of the FP64 forward"""
    dims = (32, 24, 32)
    case, o, g = striped_pair(cpmp, "fp32", dims=dims)
    nx, ny, nz = dims
    nsteps = 60
    rng = np.random.default_rng(11)
    rho0 = o.get(orc.RHO)
    target = rho0 * (1 + 0.01 * rng.standard_normal(rho0.size))
    mw = stripes(nx, ny, 16)
    J32, g32, _, k32 = g.adjoint(nsteps, target, checkpoint_every=20)
    Jo, go, _, ko = o.adjoint(nsteps, target)
    rel = np.max(np.abs(g32 - go)) / np.max(np.abs(go))
    print(f"[striped-adjoint] J fp32 {J32:.8e} oracle {Jo:.8e}; grad max rel {rel:.3e}; "
          f"dJ/dKa {k32:.6e} vs {ko:.6e}", flush=True)
    assert J32 == pytest.approx(Jo, rel=1e-4)
    assert rel < 1e-3
    assert k32 == pytest.approx(ko, rel=1e-3)

    def J64(mu):
        lat = cpmp.Lattice(cpmp.make_params(*dims, eos="VDW", tr=0.8, nwall=3))
        lat.set_wall_mu(mu, mw)
        lat.init_from_rho(rho0)
        lat.step(nsteps)
        r = lat.get("rho")
        fl = lat.get("type") == 1
        lat.close()
        return 0.5 * np.sum((r[fl] - target[fl]) ** 2)

    h = 1e-6
    for i in range(3):
        d = rng.standard_normal(nx * ny)
        fd = (J64(mw + h * d) - J64(mw - h * d)) / (2 * h)
        ad = g32 @ d
        print(f"[striped-adjoint] dir {i}: fp32 adjoint {ad:.8e}, fp64 central difference {fd:.8e}", flush=True)
        assert abs(ad - fd) <= 1e-3 * abs(fd) + 1e-5 * np.linalg.norm(g32) * np.linalg.norm(d)
    g.close()
