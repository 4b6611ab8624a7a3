"""HIP discrete adjoint: against the oracle's reverse mode (FP64 reference
gradient), central finite differences of the HIP forward (h = 1e-6, FP64), and
checkpoint-interval invariance; then the 96x96x64 adjoint configuration."""
import time

import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.gpu


def fourier_field(nx, ny, seed, modes=4, amp=0.2, mean=0.0):
    """smooth random wall field: sum of `modes` low Fourier modes, max |.| = amp"""
    rng = np.random.default_rng(seed)
    x, y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    f = np.zeros((nx, ny))
    for _ in range(modes):
        k, l = rng.integers(0, 3, 2)
        if k == 0 and l == 0:
            k = 1
        f += rng.standard_normal() * np.cos(2 * np.pi * (k * x / nx + l * y / ny) + rng.uniform(0, 2 * np.pi))
    return (mean + amp * f / np.max(np.abs(f))).ravel()


def small_case(cpmp, nx=16, ny=12, nz=20, K1=0.0):
    g, l = 0.108623, 7.67892     # PRW Tr = 0.75 (CoCurve_PRW_2.txt)
    kw = dict(eos="PRW", tr=0.75, nwall=3, Ka=0.002, K1=K1, K2=-K1)
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    r = np.sqrt((x - nx / 2) ** 2 + (y - ny / 2) ** 2 + (z - 2.5) ** 2)
    rho0 = g + (l - g) * 0.5 * (1 - np.tanh(2 * (r - 6) / 4.0))
    rho0[:, :, :3] = 0
    rho0[:, :, nz - 3:] = 0
    mu_lo = fourier_field(nx, ny, 4, amp=0.02, mean=0.0)
    mu_hi = np.full(nx * ny, 0.01)
    rng = np.random.default_rng(9)
    target = rho0.ravel() * (1 + 0.02 * rng.standard_normal(rho0.size))
    return kw, rho0.ravel(), mu_lo, mu_hi, target


def gpu_J(cpmp, kw, dims, rho0, mu_lo, mu_hi, target, nsteps, Ka=None):
    k = dict(kw)
    if Ka is not None:
        k["Ka"] = Ka
    g = cpmp.Lattice(cpmp.make_params(*dims, **k))
    g.set_wall_mu(mu_lo, mu_hi)
    g.init_from_rho(rho0)
    g.step(nsteps)
    rho = g.get("rho")
    fl = g.get("type") == 1
    g.close()
    return 0.5 * np.sum((rho[fl] - target[fl]) ** 2)


@pytest.mark.parametrize("K1", [0.0, 0.01])
def test_adjoint_matches_oracle_reverse_mode(cpmp, K1):
    dims = (16, 12, 20)
    kw, rho0, mu_lo, mu_hi, target = small_case(cpmp, *dims, K1=K1)
    nsteps = 30
    g = cpmp.Lattice(cpmp.make_params(*dims, **kw))
    g.set_wall_mu(mu_lo, mu_hi)
    g.init_from_rho(rho0)
    J, glo, ghi, gka = g.adjoint(nsteps, target, checkpoint_every=7)
    o = orc.Oracle(orc.make_params(*dims, **kw))
    o.set_wall_mu(mu_lo, mu_hi)
    o.init_from_rho(rho0)
    Jo, glo_o, ghi_o, gka_o = o.adjoint(nsteps, target)
    print(f"[adjoint] K1={K1}: J {J:.12e} vs {Jo:.12e}; |g_lo| {np.linalg.norm(glo):.6e}", flush=True)
    assert J == pytest.approx(Jo, rel=1e-10)
    assert np.max(np.abs(glo - glo_o)) < 1e-7 * np.max(np.abs(glo_o))
    assert np.max(np.abs(ghi - ghi_o)) < 1e-7 * np.max(np.abs(ghi_o)) + 1e-12
    assert gka == pytest.approx(gka_o, rel=1e-7)
    # the state is left at the end of the forward run
    assert g.step_count == nsteps
    assert np.allclose(g.get("rho"), o.get(orc.RHO), rtol=1e-10)


def test_adjoint_checkpoint_interval_invariance(cpmp):
    dims = (16, 12, 20)
    kw, rho0, mu_lo, mu_hi, target = small_case(cpmp, *dims)
    res = []
    for ck in (1, 5, 40):
        g = cpmp.Lattice(cpmp.make_params(*dims, **kw))
        g.set_wall_mu(mu_lo, mu_hi)
        g.init_from_rho(rho0)
        res.append(g.adjoint(40, target, checkpoint_every=ck))
        g.close()
    for r in res[1:]:
        assert r[0] == pytest.approx(res[0][0], rel=1e-14)
        assert np.allclose(r[1], res[0][1], rtol=1e-12, atol=1e-14 * np.max(np.abs(res[0][1])))
        assert r[3] == pytest.approx(res[0][3], rel=1e-12)


def test_adjoint_matches_central_differences(cpmp):
    dims = (16, 12, 20)
    kw, rho0, mu_lo, mu_hi, target = small_case(cpmp, *dims)
    nsteps = 60
    g = cpmp.Lattice(cpmp.make_params(*dims, **kw))
    g.set_wall_mu(mu_lo, mu_hi)
    g.init_from_rho(rho0)
    J, glo, ghi, gka = g.adjoint(nsteps, target, checkpoint_every=20)
    rng = np.random.default_rng(6)
    h = 1e-6
    for i in range(8):
        d = rng.standard_normal(mu_lo.size)
        fd = (gpu_J(cpmp, kw, dims, rho0, mu_lo + h * d, mu_hi, target, nsteps)
              - gpu_J(cpmp, kw, dims, rho0, mu_lo - h * d, mu_hi, target, nsteps)) / (2 * h)
        ad = glo @ d
        print(f"[adjoint-fd] dir {i}: adjoint {ad:.10e} fd {fd:.10e}", flush=True)
        assert ad == pytest.approx(fd, rel=1e-4)
    hk = 1e-7
    fd = (gpu_J(cpmp, kw, dims, rho0, mu_lo, mu_hi, target, nsteps, Ka=kw["Ka"] + hk)
          - gpu_J(cpmp, kw, dims, rho0, mu_lo, mu_hi, target, nsteps, Ka=kw["Ka"] - hk)) / (2 * hk)
    assert gka == pytest.approx(fd, rel=1e-4)


def test_autograd_wrapper(cpmp):
    """the checking wrapper: backward returns the adjoint gradient; a directional
    derivative of the objective matches central differences."""
    import importlib
    import torch
    ag = importlib.import_module("3dcpmpmodel_amd.autograd")
    dims = (16, 12, 20)
    kw, rho0, mu_lo, mu_hi, target = small_case(cpmp, *dims)
    kw = dict(kw)
    ka0 = kw.pop("Ka")
    setup = dict(dims=dims, params=kw, rho0=rho0, target=target, nsteps=40, checkpoint_every=10, mu_hi=mu_hi)
    mu = torch.tensor(mu_lo, dtype=torch.float64, requires_grad=True)
    ka = torch.tensor(ka0, dtype=torch.float64, requires_grad=True)
    J = ag.wall_objective(mu, ka, setup)
    J.backward()
    d = torch.from_numpy(np.random.default_rng(12).standard_normal(mu_lo.size))
    h = 1e-6
    with torch.no_grad():
        fd = (ag.wall_objective(mu + h * d, ka, setup) - ag.wall_objective(mu - h * d, ka, setup)) / (2 * h)
    assert float(mu.grad @ d) == pytest.approx(float(fd), rel=1e-4)
    with torch.no_grad():
        fdk = (ag.wall_objective(mu, ka + 1e-7, setup) - ag.wall_objective(mu, ka - 1e-7, setup)) / 2e-7
    assert float(ka.grad) == pytest.approx(float(fdk), rel=1e-4)


def c4_setup(cpmp, nsteps_target):
    """96x96x64 sessile PR droplet at Tr = 0.75; wall field seed 4, target from a
    forward run with the wall field of seed 5."""
    nx, ny, nz = 96, 96, 64
    kw = dict(eos="PRW", tr=0.75, nwall=3)
    g, l = 0.108623, 7.67892
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    r = np.sqrt((x - nx / 2) ** 2 + (y - ny / 2) ** 2 + (z - 2.5) ** 2)
    rho0 = g + (l - g) * 0.5 * (1 - np.tanh(2 * (r - 24) / 5.0))
    rho0[:, :, :3] = 0
    rho0[:, :, nz - 3:] = 0
    rho0 = rho0.ravel()
    mu_design = fourier_field(nx, ny, 4, amp=0.2)
    mu_target = fourier_field(nx, ny, 5, amp=0.2)
    lat = cpmp.Lattice(cpmp.make_params(nx, ny, nz, **kw))
    lat.set_wall_mu(mu_target, 0.0)
    lat.init_from_rho(rho0)
    lat.step(nsteps_target)
    target = lat.get("rho")
    lat.close()
    return (nx, ny, nz), kw, rho0, mu_design, target


def test_adjoint_96x96x64_configuration(cpmp):
    """2000 steps, checkpoints every 100, FD (h = 1e-6, FP64) on 16 random
    directions of the wall field; adjoint/forward time ratio reported."""
    nsteps = 2000
    dims, kw, rho0, mu, target = c4_setup(cpmp, nsteps)
    hi = np.zeros_like(mu)

    def run(m):
        lat = cpmp.Lattice(cpmp.make_params(*dims, **kw))
        lat.set_wall_mu(m, hi)
        lat.init_from_rho(rho0)
        lat.step(nsteps)
        rho = lat.get("rho")
        fl = lat.get("type") == 1
        lat.close()
        return 0.5 * np.sum((rho[fl] - target[fl]) ** 2)

    lat = cpmp.Lattice(cpmp.make_params(*dims, **kw))
    lat.set_wall_mu(mu, hi)
    lat.init_from_rho(rho0)
    t0 = time.perf_counter()
    lat.step(nsteps)
    t_fwd = time.perf_counter() - t0
    lat.init_from_rho(rho0)
    t0 = time.perf_counter()
    J, glo, ghi, gka = lat.adjoint(nsteps, target, checkpoint_every=100)
    t_adj = time.perf_counter() - t0
    lat.init_from_rho(rho0)
    t0 = time.perf_counter()
    Jt, glot, _, gkat = lat.adjoint(nsteps, target, checkpoint_every=100, tape_budget_bytes=-1)
    t_tape = time.perf_counter() - t0
    print(f"[c4] J={J:.6e} |g|={np.linalg.norm(glo):.4e} adjoint/forward time = {t_adj / t_fwd:.2f} "
          f"(checkpoints + recompute), {t_tape / t_fwd:.2f} (tape in device memory)", flush=True)
    assert Jt == J and np.array_equal(glot, glo) and gkat == gka
    assert np.isfinite(J) and J > 0
    rng = np.random.default_rng(6)
    h = 1e-6
    worst = 0.0
    for i in range(16):
        d = rng.standard_normal(mu.size)
        fd = (run(mu + h * d) - run(mu - h * d)) / (2 * h)
        ad = glo @ d
        worst = max(worst, abs(ad - fd) / abs(fd))
        print(f"[c4-fd] dir {i}: adjoint {ad:.8e} fd {fd:.8e} rel {abs(ad - fd) / abs(fd):.2e}", flush=True)
    assert worst < 1e-4


def test_adjoint_tape_budget_schedules_agree(cpmp):
    """taping all steps, taping the trailing 13 of 40 (the rest recomputed from
    checkpoints every 6 steps) and recomputing everything run the same FP64
    forward steps: identical J and gradients"""
    dims = (16, 12, 20)
    kw, rho0, mu_lo, mu_hi, target = small_case(cpmp, *dims)
    n = dims[0] * dims[1] * dims[2]
    per_step = 10 * n * 8          # rho, u, grad mu, grad rho in FP64
    res = []
    for budget in (0, 13 * per_step, -1):
        g = cpmp.Lattice(cpmp.make_params(*dims, **kw))
        g.set_wall_mu(mu_lo, mu_hi)
        g.init_from_rho(rho0)
        res.append(g.adjoint(40, target, checkpoint_every=6, tape_budget_bytes=budget))
        assert g.step_count == 40
        g.close()
    for r in res[1:]:
        assert r[0] == res[0][0]
        assert np.array_equal(r[1], res[0][1])
        assert np.array_equal(r[2], res[0][2])
        assert r[3] == res[0][3]


@pytest.mark.parametrize("form,dims", [("fp64", (16, 12, 20)), ("fp32", (32, 16, 32))])
def test_adjoint_overlapped_schedule_matches_serial(cpmp, monkeypatch, form, dims):
    """segments recomputed on a second stream while the reverse sweep of the
    segment after them runs (two tapes), with and without trailing taped steps:
    the same J and gradients, bit for bit, as recompute and sweep in turn"""
    n = dims[0] * dims[1] * dims[2]
    per_step = 10 * n * (8 if form == "fp64" else 4)
    for budget in (0, 13 * per_step):
        res = []
        for pipe in ("1", "0"):
            monkeypatch.setenv("CPMP_ADJ_PIPE", pipe)
            g, target = _small_lattice(cpmp, dims, form)
            plan = g.adjoint_plan(40, 6, tape_budget_bytes=budget)
            assert plan["overlapped"] == (1 if pipe == "1" else 0)
            res.append(g.adjoint(40, target, checkpoint_every=6, tape_budget_bytes=budget))
            assert g.step_count == 40
            rho_end = g.get("rho")
            g.close()
            res[-1] = res[-1] + (rho_end,)
        print(f"[adjoint-overlap] {form} budget {budget}: J {res[0][0]:.12e}", flush=True)
        a, b = res
        assert a[0] == b[0] and a[3] == b[3]
        for i in (1, 2, 4):
            assert np.array_equal(a[i], b[i])


def _small_lattice(cpmp, dims, form="fp64", **extra):
    kw, rho0, mu_lo, mu_hi, target = small_case(cpmp, *dims)
    kw.update(extra)
    g = cpmp.Lattice(cpmp.make_params(*dims, form=form, **kw))
    g.set_wall_mu(mu_lo, mu_hi)
    g.init_from_rho(rho0)
    return g, target


def test_adjoint_memory_limit_sub_segments_agree(cpmp):
    """a memory limit that the segment tape does not fit: the segment is
    recomputed in sub-segments (a first pass stores sub-checkpoints, then each
    sub-segment is taped) with identical results; a limit below every schedule
    is refused before anything runs"""
    dims = (16, 12, 20)
    n = dims[0] * dims[1] * dims[2]
    per_tape = 80 * n           # 10 doubles per node
    res, plans = [], []
    for extra in (None, 1280 * n, 1120 * n):
        g, target = _small_lattice(cpmp, dims)
        if extra is None:
            plan = g.adjoint_plan(40, 40)
            base = plan["device_bytes"] - 40 * per_tape   # everything but the tape
            limit = 0
        else:
            limit = base + extra
            plan = g.adjoint_plan(40, 40, memory_limit_bytes=limit)
        plans.append(plan)
        res.append(g.adjoint(40, target, checkpoint_every=40, memory_limit_bytes=limit))
        assert g.step_count == 40
        g.close()
    print(f"[adjoint-plan] {plans}", flush=True)
    # total(S) = base + S * per_tape + max(ceil(40 / S) - 2, 0) * (160 B per node)
    # without a limit the one segment fits the working tape: taped in the first run
    assert (plans[0]["sub_length"], plans[0]["sub_checkpoints"], plans[0]["recomputed_steps"]) == (40, 0, 0)
    assert (plans[0]["taped_steps"], plans[0]["segments"]) == (40, 0)
    assert (plans[1]["sub_length"], plans[1]["sub_checkpoints"], plans[1]["recomputed_steps"]) == (14, 1, 68)
    assert (plans[2]["sub_length"], plans[2]["sub_checkpoints"], plans[2]["recomputed_steps"]) == (10, 2, 70)
    assert plans[1]["device_bytes"] == base + 1280 * n and plans[2]["device_bytes"] == base + 1120 * n
    for r in res[1:]:
        assert r[0] == res[0][0]
        assert np.array_equal(r[1], res[0][1])
        assert np.array_equal(r[2], res[0][2])
        assert r[3] == res[0][3]
    g, target = _small_lattice(cpmp, dims)
    with pytest.raises(cpmp.CPMPError):
        g.adjoint_plan(40, 40, memory_limit_bytes=base + 400 * n)
    with pytest.raises(cpmp.CPMPError):
        g.adjoint(40, target, checkpoint_every=40, memory_limit_bytes=base + 400 * n)
    assert g.step_count == 0
    g.close()


@pytest.mark.parametrize("dims", [(16, 12, 20), (32, 16, 32)])
def test_adjoint_fp32_form(cpmp, monkeypatch, dims):
    """FP32 form (32 x 16 x 32 takes the fused-density path): the reverse sweep
    runs in FP64 along the FP32 trajectory.  J and the gradients follow the FP64
    form's to the FP32 rounding of the trajectory, and every schedule (checkpoint
    interval, tape budget, sub-segments) gives identical results."""
    n = dims[0] * dims[1] * dims[2]
    g, target = _small_lattice(cpmp, dims, "fp64")
    ref = g.adjoint(40, target, checkpoint_every=10)
    g.close()
    res = []
    per_tape = 40 * n           # 10 FP32 values per node
    if dims == (32, 16, 32):
        monkeypatch.setenv("CPMP_XC", "2")   # the fused path (the lattice alone would select single planes)
    for ck, budget, sub in ((10, 0, False), (40, 0, False), (7, 13 * per_tape, False), (40, 0, True)):
        g, target = _small_lattice(cpmp, dims, "fp32")
        limit = 0
        if sub:
            plan = g.adjoint_plan(40, 40)
            assert plan["sub_length"] == 40
            base = plan["device_bytes"] - 40 * per_tape
            limit = base + 16 * per_tape
            assert g.adjoint_plan(40, 40, memory_limit_bytes=limit)["sub_length"] < 40
        res.append(g.adjoint(40, target, checkpoint_every=ck, tape_budget_bytes=budget, memory_limit_bytes=limit))
        assert g.step_count == 40
        g.close()
    J, glo, ghi, gka = res[0]
    eg = np.max(np.abs(glo - ref[1])) / np.max(np.abs(ref[1]))
    print(f"[adjoint-fp32] {dims}: J {J:.10e} vs FP64 {ref[0]:.10e}; gradient rel {eg:.2e}; "
          f"dJ/dKa {gka:.6e} vs {ref[3]:.6e}", flush=True)
    assert J == pytest.approx(ref[0], rel=1e-5)
    assert eg < 1e-3
    assert np.max(np.abs(ghi - ref[2])) <= 1e-3 * np.max(np.abs(ref[2]))
    assert gka == pytest.approx(ref[3], rel=1e-3)
    for r in res[1:]:
        assert r[0] == J and np.array_equal(r[1], glo) and np.array_equal(r[2], ghi) and r[3] == gka


@pytest.mark.parametrize("dims,xc,lap", [((16, 12, 20), "2", 5), ((15, 17, 40), "2", 5), ((15, 17, 40), "", 7),
                                          ((24, 12, 20), "", 7)])
def test_adjoint_stencil_march_matches_per_node(cpmp, dims, xc, lap):
    """the transposed force stencils and chemical potential as x marches
    (partial y / z tiles, chunks of two planes with a last chunk of one plane for
    nx = 15, chunks of 8 with a short last one, 5- and 7-point Laplacians)
    against the one-thread-per-node kernels: the same sums in another order"""
    import os
    res = {}
    for mode in ("v1", "march"):
        os.environ["CPMP_ADJ_STENCIL"] = mode
        if xc:
            os.environ["CPMP_ADJ_XC"] = xc
        try:
            g, target = _small_lattice(cpmp, dims, lap_points=lap)
            res[mode] = g.adjoint(30, target, checkpoint_every=10)
            g.close()
        finally:
            os.environ.pop("CPMP_ADJ_STENCIL", None)
            os.environ.pop("CPMP_ADJ_XC", None)
    a, b = res["v1"], res["march"]
    assert a[0] == b[0]                       # the forward run is the same
    for k in (1, 2):
        err = np.max(np.abs(a[k] - b[k])) / np.max(np.abs(a[k]))
        print(f"[adjoint-march] {dims} gradient {k}: rel {err:.2e}", flush=True)
        assert err < 1e-11
    assert b[3] == pytest.approx(a[3], rel=1e-10)
