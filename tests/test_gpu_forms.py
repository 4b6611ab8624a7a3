"""The FP32, FP32-arithmetic (FP64 storage) and FP16-storage forms against the FP64
form, conservation at the full 256^3 size, checkpoints, Tecplot output,
observables."""
import os

import numpy as np
import pytest

import gpu_cases as gc
import oracle_lib as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form,tol", [("fp32", 2e-4), ("fp32a", 2e-4), ("fp16", 2e-2)])
def test_reduced_precision_forms_track_fp64(cpmp, form, tol):
    case = gc.CASES["ref_prw_walls"]
    _, g64 = gc.build(case, cpmp, "fp64")
    _, gl = gc.build(case, cpmp, form)
    g64.step(200)
    gl.step(200)
    r64, rl = g64.get("rho"), gl.get("rho")
    fluid = g64.get("type") == 1
    err = np.max(np.abs(rl[fluid] - r64[fluid])) / np.max(r64[fluid])
    print(f"[forms] {form}: max |rho - rho_fp64| / max rho = {err:.3e}", flush=True)
    assert err < tol
    m64, ml = g64.reduce().mass, gl.reduce().mass
    assert abs(ml - m64) / m64 < tol


def _ulp32(x):
    a = np.abs(x).astype(np.float32)
    return (np.nextafter(a, np.float32(np.inf)) - a).astype(np.float64)


@pytest.mark.parametrize("form,geom", [("fp32", "periodic"), ("fp16", "periodic"), ("fp32", "walls"),
                                       ("fp32", "periodic-long"), ("fp16", "periodic-long")])
def test_interface_chemical_potential_in_fp64(cpmp, monkeypatch, form, geom):
    """The FP32 storage forms evaluate the EOS part of mu in FP64 at interface nodes
    (density inside the Maxwell gas-liquid gap less 2 % at each end) and round once:
    there mu matches the FP64 form's mu of the same (FP32) density field to 1.5 FP32
    ulp, both for the chemical-potential march that starts a run and for the one
    that completes the previous step's summed density (the second step of a run on a
    periodic tiled lattice).  The walled lattice has partial tiles (13 x 36 in y, z)
    and 3 + 3 wall layers; nodes whose Laplacian reaches the wall (within 2 layers)
    are left out there, since the wall values are stored in FP32.  periodic-long:
    chunks of 32 planes, where many waves hold more interface nodes than they finish
    themselves and pass the rest to the overflow kernel."""
    if geom.startswith("periodic"):
        dims, eos, tr, nwall, sphere = (64, 32, 64), "RKE", 0.6, 0, (32, 16, 32, 14)
        # the fused path with the folded completion
        monkeypatch.setenv("CPMP_XC", "32" if geom == "periodic-long" else "2")
    else:
        dims, eos, tr, nwall, sphere = (40, 13, 36), "PRW", 0.7, 3, (20, 6, 2.5, 10)
    A, B, Tc, Rc = cpmp.set_multiphase(eos, tr)
    g0, l0 = cpmp.coexistence(eos, A, B, tr * Tc)
    lo, hi = np.float32(g0 + 0.02 * (l0 - g0)), np.float32(l0 - 0.02 * (l0 - g0))
    z = np.tile(np.arange(dims[2]), dims[0] * dims[1])
    inner = (z >= nwall + 2) & (z < dims[2] - nwall - 2)

    def lat(f):
        return cpmp.Lattice(cpmp.make_params(*dims, eos=eos, tr=tr, nwall=nwall, form=f, rho0=g0))

    def start(g):
        g.init_spheres([sphere], 5.0, g0, l0, noise_amp=1e-3, seed=5)

    a = lat(form)
    start(a)
    r0 = a.get("rho")
    a.step(1)
    r1, mu_first = a.get("rho"), a.get("mu")
    a.close()
    b = lat(form)
    start(b)
    b.step(2)
    mu_second = b.get("mu")
    b.close()
    for name, r, mu in (("first", r0, mu_first), ("second", r1, mu_second)):
        ref = lat("fp64")
        ref.init_from_rho(r)
        ref.step(1)
        mu64 = ref.get("mu")
        ref.close()
        rf = r.astype(np.float32)
        band = (rf > lo) & (rf < hi) & inner
        ulps = np.abs(mu - mu64) / _ulp32(mu64)
        bulk = ~band & inner
        print(f"[interface-mu] {form} {geom} {name}: {band.sum()} interface nodes, max {ulps[band].max():.2f} ulp; "
              f"bulk max {np.nanmax(ulps[bulk]):.2f} ulp", flush=True)
        assert band.sum() > 500
        assert ulps[band].max() <= 1.5


@pytest.mark.parametrize("form", ["fp64", "fp32", "fp16", "fp32a"])
def test_mass_conservation_256cubed(cpmp, form):
    """Full-size property: a periodic 256^3 RK coalescence run conserves mass
    (FP64 sums on the device) and stays finite."""
    n = 256
    p = cpmp.make_params(n, n, n, eos="RKE", tr=0.6, nwall=0, form=form, rho0=0.0319804)
    g = cpmp.Lattice(p)
    g.init_spheres([(80, 128, 128, 48), (176, 128, 128, 48)], 5.0, 0.0319804, 8.25467, noise_amp=1e-4, seed=3)
    m0 = g.reduce().mass
    g.step(20)
    s = g.reduce()
    drift = abs(s.mass - m0) / m0
    print(f"[256^3 {form}] mass drift after 20 steps {drift:.3e}, max|u| {s.max_speed:.3e}", flush=True)
    assert s.bad_count == 0
    assert drift < {"fp64": 1e-12, "fp32": 1e-5, "fp16": 1e-3, "fp32a": 1e-5}[form]
    g.close()


@pytest.mark.parametrize("geom", ["walls", "periodic"])
def test_fp32_arithmetic_form_keeps_fp64_storage_and_mu(cpmp, monkeypatch, geom):
    """FP32A: populations, rho and mu stored in FP64 and mu evaluated in FP64 (the
    chemical-potential code of the FP64 form: bit-identical mu for the same density),
    the collision, force and streaming arithmetic in FP32 (one step differs from the
    FP64 form's by FP32 rounding, not by more).  Periodic: the fused density sum with
    chunks of 8 planes."""
    if geom == "walls":
        case = gc.CASES["ref_prw_walls"]
    else:
        case = dict(gc.CASES["cse_periodic"], dims=(64, 32, 64))
        monkeypatch.setenv("CPMP_XC", "8")
        monkeypatch.setenv("CPMP_COLLIDE", "march")
    _, g64 = gc.build(case, cpmp, "fp64")
    _, ga = gc.build(case, cpmp, "fp32a")
    assert ga.algorithmic_bytes_per_node() == 2 * 19 * 8    # FP64 populations
    r0 = ga.get("rho")
    assert np.array_equal(r0, g64.get("rho"))
    g64.step(1)
    ga.step(1)
    # mu of the same initial density, FP64 in both
    assert np.array_equal(ga.get("mu"), g64.get("mu"))
    fluid = g64.get("type") == 1
    r64, ra = g64.get("rho")[fluid], ga.get("rho")[fluid]
    d = np.max(np.abs(ra - r64) / r64)
    print(f"[fp32a] {geom}: one step, max rel |rho - rho_fp64| = {d:.3e}", flush=True)
    assert 0 < d < 5e-6
    g64.close()
    ga.close()


def test_checkpoint_roundtrip_and_corruption(cpmp, tmp_path):
    case = gc.CASES["ref_prw_walls"]
    _, g = gc.build(case, cpmp)
    g.step(30)
    path = str(tmp_path / "state.ckp")
    g.checkpoint_save(path)
    g.step(20)
    ref = g.get("rho")
    _, g2 = gc.build(case, cpmp)
    g2.checkpoint_load(path)
    assert g2.step_count == 30
    g2.step(20)
    assert np.array_equal(g2.get("rho"), ref)
    bad = str(tmp_path / "bad.ckp")
    with open(path, "rb") as fh:
        data = bytearray(fh.read())
    data[0] ^= 0xFF
    with open(bad, "wb") as fh:
        fh.write(data)
    with pytest.raises(cpmp.CPMPError) as e:
        g2.checkpoint_load(bad)
    assert e.value.code == -4
    with open(bad, "wb") as fh:
        fh.write(bytes(data[:1000]))
    with pytest.raises(cpmp.CPMPError):
        g2.checkpoint_load(bad)


def test_contact_angle_matches_oracle(cpmp):
    case = gc.CASES["prw_2d"]
    o, g = gc.build(case, cpmp)
    o.step(300)
    g.step(300)
    dg, dl = gc.densities("PRW", 0.7)
    a = g.contact_angle(0, dg, dl)
    b = orc.contact_angle(o.p, o.get(orc.RHO), dg, dl, 0)
    assert a["valid"] == b[7] == 1
    got = np.array([a[k] for k in ("CaLeft", "CaRight", "CaCap", "XLeft", "XRight", "H", "L")])
    assert np.allclose(got, b[:7], rtol=1e-8, atol=1e-8, equal_nan=True)


@pytest.mark.parametrize("lo", [3, 2, 4])
def test_contact_angle_on_pinned_fields_matches_oracle(cpmp, lo):
    """the device measurement (the XZ plane extracted on the device) on the wedge
    and spherical-cap fields whose angles the oracle tests pin in closed form, with
    the scan rows following the bottom wall depth; no bottom wall is an error"""
    import test_oracle_pins as pins
    nx, ny, nz = 64, 3, 72
    base = lo - 1
    for theta in (30, 60, 90, 120, 150):
        H = 12.6 * np.tan(np.radians(theta) / 2)
        fields = [pins.plane_field(nx, ny, nz, pins.wedge(31.3, 12.6, theta, base + H, base,
                                                          rows={base: (0.41, -0.23), base + 3: (-0.17, 0.31)}))]
        x, z = np.meshgrid(np.arange(nx, dtype=float), np.arange(nz, dtype=float), indexing="ij")
        r = np.sqrt((x - 32.3) ** 2 + (z - (base - 25 * np.cos(np.radians(theta)))) ** 2)
        cap = pins.MID - (pins.DL - pins.DG) / 2 * np.tanh(2 * (r - 25) / 2.5)
        fields.append(np.repeat(cap[:, None, :], ny, axis=1).ravel())
        p = orc.make_params(nx, ny, nz, eos="PRW", tr=0.7, nwall=lo)
        g = cpmp.Lattice(cpmp.make_params(nx, ny, nz, eos="PRW", tr=0.7, nwall=lo))
        for rho in fields:
            g.set("rho", rho)
            a = g.contact_angle(1, pins.DG, pins.DL)
            b = orc.contact_angle(p, rho, pins.DG, pins.DL, 1)
            got = np.array([a[k] for k in ("CaLeft", "CaRight", "CaCap", "XLeft", "XRight", "H", "L", "valid")])
            assert np.array_equal(got, b, equal_nan=True), (theta, got, b)
        g.close()
    g = cpmp.Lattice(cpmp.make_params(nx, ny, nz, eos="PRW", tr=0.7, nwall=0))
    with pytest.raises(cpmp.CPMPError):
        g.contact_angle(1, pins.DG, pins.DL)
    g.close()


def _plt_zone_values(raw, n):
    """the Density values (last variable, FP64) of a one-zone PLT of n nodes"""
    return np.frombuffer(raw[-8 * n:], np.float64)


def test_write_plt_layout(cpmp, tmp_path):
    """full field: header magic and the density block in reference order; the
    decimated forms keep every 5th node per axis, or (SaveTask's CompressField,
    FluidCuda.cu:466-472) the linear index times 5"""
    case = gc.CASES["vdw_variants"]
    _, g = gc.build(case, cpmp)
    g.step(5)
    nx, ny, nz = case["dims"]
    rho = g.get("rho")
    path = str(tmp_path / "full.plt")
    g.write_plt(path)
    raw = open(path, "rb").read()
    assert raw[:8] == b"#!TDV112"
    assert np.array_equal(_plt_zone_values(raw, nx * ny * nz), rho)
    lx, ly, lz = nx // 5, ny // 5, nz // 5
    i, j, k = np.meshgrid(np.arange(lx), np.arange(ly), np.arange(lz), indexing="ij")
    for linear, src in ((False, (i * 5 * ny + j * 5) * nz + k * 5), (True, ((i * ny + j) * nz + k) * 5)):
        path = str(tmp_path / f"dec{int(linear)}.plt")
        g.write_plt(path, decimate=5, decimate_linear=linear)
        raw = open(path, "rb").read()
        assert np.array_equal(_plt_zone_values(raw, lx * ly * lz), rho[src.ravel()])
    g.close()


@pytest.mark.parametrize("decimate", [1, 5])
def test_write_plt_matches_reference_writer(cpmp, tmp_path, decimate):
    """byte for byte against the reference's own ordered_tec writer (oracle/_ref/plt_ref,
    built from the reference sources by oracle/ref.mk) driven as SaveTask drives it
    (FluidCuda.cu:901-927): the full field and SaveTask's DR = 5 linear-stride
    compression (FluidCuda.cu:466-472)"""
    ref_bin = os.path.join(os.path.dirname(orc.LIB_PATH), "..", "_ref", "plt_ref")
    if not os.path.exists(ref_bin):
        pytest.skip("oracle/_ref/plt_ref (the reference's Tecplot writer) is not built")
    import subprocess
    case = gc.CASES["vdw_variants"]
    _, g = gc.build(case, cpmp)
    g.step(5)
    nx, ny, nz = case["dims"]
    name = f"FlowField_{decimate}"
    ours = tmp_path / "ours"
    ours.mkdir()
    g.write_plt(str(ours / f"{name}.plt"), decimate=decimate, decimate_linear=decimate > 1)
    rho_path = str(tmp_path / "rho.bin")
    g.get("rho").tofile(rho_path)
    subprocess.check_call([ref_bin, rho_path, str(nx), str(ny), str(nz), str(decimate), str(tmp_path), name])
    assert open(str(tmp_path / f"{name}.plt"), "rb").read() == open(str(ours / f"{name}.plt"), "rb").read()
    g.close()


@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_device_observables_match_host_restatement(cpmp, form):
    """cpmp_probe, cpmp_cap_fit (column reduction on the device) and
    cpmp_laplace_jump against a NumPy restatement on the same density field"""
    case = gc.CASES["ref_prw_walls"]
    _, g = gc.build(case, cpmp, form)
    g.step(60)
    nx, ny, nz = case["dims"]
    rho = g.get("rho").reshape(nx, ny, nz)
    mu = g.get("mu").reshape(nx, ny, nz)
    for x, y, z in [(0, 0, 0), (5, 7, 11), (nx - 1, ny - 1, nz - 1), (nx // 2, ny // 2, 3)]:
        assert g.probe("rho", x, y, z) == rho[x, y, z]
        assert g.probe("mu", x, y, z) == mu[x, y, z]
    dg, dl = gc.densities("PRW", 0.7)
    mid, lo, hi = (dg + dl) / 2, 3, 3
    area = float(np.sum(rho[:, :, lo] >= mid))
    a, b = rho[:, :, lo:nz - hi - 1], rho[:, :, lo + 1:nz - hi]
    cross = (a >= mid) & (b < mid)
    k = np.arange(lo, nz - hi - 1)[None, None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        zc = np.where(cross, k + (a - mid) / (a - b), -np.inf)
    H = max(0.0, float(np.max(zc - (lo - 0.5))))
    cap = g.cap_fit(dg, dl)
    assert cap["height"] == pytest.approx(H, rel=1e-14)
    assert cap["base_radius"] == pytest.approx(np.sqrt(area / np.pi), rel=1e-14)
    p = g.p
    dp = g.laplace_jump((nx // 2, ny // 2, 8), (0, 0, nz // 2))
    want = p.K ** 2 * (cpmp.eos_pressure("PRW", p.A, p.B, p.T, rho[nx // 2, ny // 2, 8])
                       - cpmp.eos_pressure("PRW", p.A, p.B, p.T, rho[0, 0, nz // 2]))
    assert dp == pytest.approx(want, rel=1e-13, abs=1e-15)
    with pytest.raises(cpmp.CPMPError):
        g.probe("rho", nx, 0, 0)
    g.close()


@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_neck_radius_matches_host_restatement(cpmp, form):
    """cpmp_neck_radius (device reduction) against the same sum over the downloaded
    density: a sharp disc along x gives its node count, a smooth column after a few
    steps the clamped liquid fraction summed on the host"""
    nx, ny, nz = 8, 48, 40
    g, l = 0.0319804, 8.25467            # RKE Tr = 0.6 coexistence (CoCurve_RKE_2.txt x Rc)
    tol = 1e-12 if form == "fp64" else 1e-6
    y, z = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    r = np.sqrt((y - 23.5) ** 2 + (z - 19.5) ** 2)
    disc = r < 12.3
    lat = cpmp.Lattice(cpmp.make_params(nx, ny, nz, eos="RKE", tr=0.6, nwall=0, form=form))
    lat.init_from_rho(np.repeat(np.where(disc, l, g)[None], nx, 0).ravel())
    R, A = lat.neck_radius(3, g, l)
    assert A == pytest.approx(disc.sum(), rel=tol)
    assert R == pytest.approx(np.sqrt(disc.sum() / np.pi), rel=tol)
    lat.close()
    lat = cpmp.Lattice(cpmp.make_params(nx, ny, nz, eos="RKE", tr=0.6, nwall=0, form=form))
    prof = g + (l - g) * 0.5 * (1 - np.tanh(2 * (r - 12.3) / 5.0))
    lat.init_from_rho(np.repeat(prof[None], nx, 0).ravel())
    lat.step(20)
    rho = lat.get("rho").reshape(nx, ny, nz)
    for x in (0, 5):
        R, A = lat.neck_radius(x, g, l)
        ref = np.clip((rho[x] - g) / (l - g), 0, 1).sum()
        print(f"[neck] {form} x={x}: A {A:.6f} host {ref:.6f}, r {R:.4f}", flush=True)
        assert A == pytest.approx(ref, rel=tol)
    with pytest.raises(cpmp.CPMPError):
        lat.neck_radius(nx, g, l)
    lat.close()
