"""Benchmark: MLUPS of the chemical-potential multiphase D3Q19 MRT step on one
MI355X at 256^3 (two-droplet Redlich-Kwong coalescence, Tr = 0.6, periodic),
FP32 form, with the FP64 and FP16-storage forms timed beside it.

Prints one JSON line on stdout; progress goes to stderr.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0     # MI355X HBM3E, AMD published
G_RKE, L_RKE = 0.0319804, 8.25467   # CoCurve_RKE_2.txt Tr=0.60 row times Rc


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_lattice(cpmp, n, form):
    p = cpmp.make_params(n, n, n, eos="RKE", tr=0.6, nwall=0, form=form, rho0=G_RKE)
    g = cpmp.Lattice(p)
    g.init_spheres([(80 * n / 256, n / 2, n / 2, 48 * n / 256), (176 * n / 256, n / 2, n / 2, 48 * n / 256)],
                   5.0, G_RKE, L_RKE, noise_amp=1e-4, seed=3)
    return g


def time_form(cpmp, n, form, steps, warmup):
    g = make_lattice(cpmp, n, form)
    g.step(warmup)
    m0 = g.reduce().mass
    g.step(1)
    t0 = time.perf_counter()
    g.step(steps)            # cpmp_step synchronises the stream before returning
    dt = time.perf_counter() - t0
    s = g.reduce()
    ktimes, per = g.kernel_times(max(5, min(20, steps // 10)))
    bpn = g.algorithmic_bytes_per_node()
    g.close()
    nodes = n ** 3
    return dict(ms_per_step=dt / steps * 1e3, mlups=nodes * steps / dt / 1e6, kernel_ms=ktimes, launches=per,
                bytes_per_node=bpn, mass_drift=abs(s.mass - m0) / m0, max_speed=s.max_speed, bad=s.bad_count)


G_PRW70, L_PRW70 = 0.0556211, 8.08045    # CoCurve_PRW_2.txt Tr=0.70 row times Rc
G_PRW75, L_PRW75 = 0.108623, 7.67892     # Tr=0.75
G_CSE85, L_CSE85 = 0.0315818, 0.279262   # CoCurve_CSE_2.txt Tr=0.85
# target contact angles of the sessile-droplet configuration; the wall chemical
# potentials are calibrated for them (calibrate_wall_mu)
TARGET_ANGLES = [30.0, 60.0, 90.0, 120.0, 150.0]


def sessile(cpmp, form, mu_w, steps):
    """128x128x96 sessile PR droplet (Tr = 0.7): hemisphere of radius 30 on the
    bottom wall, 3+3 wall layers; returns the contact angles and phase densities,
    and whether the droplet still has a contact line on the wall (`valid`: the
    reference method found both feet and a height, and the 3-D cap is attached)."""
    import math
    import numpy as np
    nx, ny, nz = 128, 128, 96
    g = cpmp.Lattice(cpmp.make_params(nx, ny, nz, eos="PRW", tr=0.7, form=form, rho0=G_PRW70))
    g.set_wall_mu(mu_w)
    g.init_spheres([(nx / 2, ny / 2, 2.5, 30.0)], 5.0, G_PRW70, L_PRW70, solid_rho=0.0)
    g.step(steps)
    ca = g.contact_angle(ny // 2, G_PRW70, L_PRW70)
    cap = g.cap_fit(G_PRW70, L_PRW70)
    # bulk phases: the gas far from the droplet, the liquid at the middle of the cap's
    # height (only for an attached cap at least 8 nodes high: a wetting film or a
    # droplet that has left the wall has no bulk liquid at a known node)
    H = cap["height"]
    valid = bool(ca["valid"] == 1 and math.isfinite(ca["H"]) and ca["H"] > 0 and math.isfinite(cap["angle"])
                 and cap["angle"] < 179 and ca["CaLeft"] > 0 and ca["CaRight"] > 0)
    attached = bool(np.isfinite(H) and H >= 8 and cap["angle"] < 179)
    out = dict(ca_left=ca["CaLeft"], ca_right=ca["CaRight"], ca_cap=ca["CaCap"], cap3d=cap["angle"], cap_height=H,
               valid=valid, rho_gas=g.probe("rho", 2, 2, nz // 2),
               rho_liq=g.probe("rho", nx // 2, ny // 2, int(round(3 + H / 2))) if attached else None)
    g.close()
    return out


# the forms compared with the FP64 form: FP32 arithmetic on FP64 storage (the EOS mu in
# FP64 everywhere), FP32 storage and arithmetic (FP64 EOS mu at the interface), FP16
# storage with FP32 arithmetic
LOWER = ("fp32a", "fp32", "fp16")


def accuracy(cpmp, steps, coalescence_steps=5000):
    """The reduced-precision forms against the FP64 form on the same inputs: the
    sessile droplet at wall chemical potentials calibrated for 30..150 degrees,
    the Laplace jump and spurious current of a periodic droplet, the neck of two
    coalescing droplets.  A wetting case without a contact line in any form is
    reported invalid and makes the worst-case errors None (never a zero error)."""
    import math
    res = {"sessile_steps": steps, "cases": []}
    A, B, Tc, Rc = cpmp.set_multiphase("PRW", 0.7)
    p0 = cpmp.make_params(8, 8, 8, eos="PRW", tr=0.7)
    mu_b = p0.K ** 2 * cpmp.eos_mu("PRW", A, B, 0.7 * Tc, L_PRW70)
    t0 = time.perf_counter()
    cal = {a: calibrate_wall_mu(cpmp, "PRW", 0.7, G_PRW70, L_PRW70, a, mu_b) for a in TARGET_ANGLES}
    log(f"[bench] wall mu for {TARGET_ANGLES} deg: {[round(v, 6) for v in cal.values()]} "
        f"(neutral {mu_b:.6f}), {time.perf_counter() - t0:.1f} s")
    res["wall_mu"] = {str(a): v for a, v in cal.items()}
    worst = {f: 0.0 for f in LOWER}
    worst_rho = {f: 0.0 for f in LOWER}
    per_angle = {f: {} for f in LOWER}
    invalid = []
    for angle in TARGET_ANGLES:
        mu_w = cal[angle]
        r64 = sessile(cpmp, "fp64", mu_w, steps)
        case = {"target_angle": angle, "mu_w": mu_w, "fp64": r64}
        for form in LOWER:
            r = sessile(cpmp, form, mu_w, steps)
            case[form] = r
            if not (r64["valid"] and r["valid"]):
                per_angle[form][str(angle)] = None
                continue
            err = max(abs(r[k] - r64[k]) for k in ("ca_left", "ca_right", "ca_cap", "cap3d"))
            rerr = abs(r["rho_gas"] / r64["rho_gas"] - 1)
            if r64["rho_liq"] is not None and r["rho_liq"] is not None:
                rerr = max(rerr, abs(r["rho_liq"] / r64["rho_liq"] - 1))
            if math.isnan(err) or math.isnan(rerr):
                per_angle[form][str(angle)] = None
                continue
            per_angle[form][str(angle)] = err
            worst[form] = max(worst[form], err)
            worst_rho[form] = max(worst_rho[form], rerr)
        if not all(case[f]["valid"] for f in ("fp64",) + LOWER):
            invalid.append(angle)
        res["cases"].append(case)
        log(f"[bench] wetting {angle} deg (mu_w={mu_w:.6f}): fp64 CaLeft {r64['ca_left']:.2f} CaCap "
            f"{r64['ca_cap']:.2f} cap3d {r64['cap3d']:.2f} valid {r64['valid']}; err "
            + ", ".join(f"{f} {per_angle[f][str(angle)]}" for f in LOWER) + " deg")
    res["contact_angle_err_deg_per_angle"] = per_angle
    res["invalid_cases"] = invalid
    res["contact_angle_err_deg"] = None if invalid else worst
    res["coexistence_rel_err"] = None if invalid else worst_rho
    # Laplace jump and spurious current: 64^3 Carnahan-Starling droplet, 1000 steps
    lap = {}
    for form in ("fp64",) + LOWER:
        n = 64
        p = cpmp.make_params(n, n, n, eos="CSE", tr=0.85, nwall=0, Ka=0.01, tau=1.0, form=form, rho0=G_CSE85)
        g = cpmp.Lattice(p)
        g.init_spheres([(32, 32, 32, 16)], 2.5, G_CSE85, L_CSE85, noise_amp=1e-6, seed=1)
        g.step(1000)
        s = g.reduce()                                   # max |MV| on the device
        dp = g.laplace_jump((32, 32, 32), (0, 0, 0))     # K^2 [p(centre) - p(corner)] on the device
        lap[form] = {"laplace_dp": dp, "spurious_max_u": s.max_speed}
        g.close()
    res["laplace"] = lap
    res["laplace_rel_err"] = {f: abs(lap[f]["laplace_dp"] / lap["fp64"]["laplace_dp"] - 1) for f in LOWER}
    # two-droplet coalescence, 256^3 RK at Tr = 0.6: neck radius in the mid plane and mass drift
    coal = {}
    for form in ("fp64",) + LOWER:
        g = make_lattice(cpmp, 256, form)
        m0 = g.reduce().mass
        g.step(coalescence_steps)
        r, _ = g.neck_radius(128, G_RKE, L_RKE)
        coal[form] = {"neck_radius": r, "mass_drift": abs(g.reduce().mass - m0) / m0}
        g.close()
        log(f"[bench] coalescence {form}: neck radius {r:.4f} after {coalescence_steps} steps, "
            f"mass drift {coal[form]['mass_drift']:.2e}")
    res["coalescence"] = {"steps": coalescence_steps, **coal}
    res["neck_radius_rel_err"] = {f: abs(coal[f]["neck_radius"] / coal["fp64"]["neck_radius"] - 1) for f in LOWER}
    return res


def measured_traffic(kernel_key):
    """Bytes per launch that left the L2 for one kernel of the 256^3 workload, from
    the committed rocprofv3 summary (FETCH_SIZE and WRITE_SIZE passes calibrated with a
    copy kernel of known size); None if the summary does not list the kernel."""
    import re
    path = os.path.join(ROOT, "profiles", "collide_traffic_256cubed.txt")
    try:
        with open(path) as fh:
            for line in fh:
                m = re.search(r"traffic_bytes_per_launch (\S+) (\d+)", line)
                if m and m.group(1) == kernel_key:
                    return int(m.group(2))
    except OSError:
        pass
    return None


def adjoint_cost(cpmp, nsteps=2000, every=100):
    """adjoint time over forward time, 96x96x64 sessile PR droplet, Tr = 0.75,
    FP64, two schedules: checkpoints every `every` steps with the segments
    recomputed (the last one is taped during the first forward run, which its
    working tape holds), and the tape of all steps held in device memory (80 B per
    node and step: 94 GB for 2000 steps) so that nothing is recomputed."""
    import numpy as np
    nx, ny, nz = 96, 96, 64
    p = cpmp.make_params(nx, ny, nz, eos="PRW", tr=0.75)
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    r = np.sqrt((x - nx / 2) ** 2 + (y - ny / 2) ** 2 + (z - 2.5) ** 2)
    rho0 = G_PRW75 + (L_PRW75 - G_PRW75) * 0.5 * (1 - np.tanh(2 * (r - 24) / 5.0))
    rho0[:, :, :3] = 0
    rho0[:, :, nz - 3:] = 0
    g = cpmp.Lattice(p)
    g.set_wall_mu(0.0)
    g.init_from_rho(rho0.ravel())
    g.step(nsteps)
    target = g.get("rho")
    g.set_wall_mu(0.01)
    g.init_from_rho(rho0.ravel())
    t0 = time.perf_counter()
    g.step(nsteps)
    t_fwd = time.perf_counter() - t0
    res = {}
    for name, budget in (("recompute", 0), ("taped", -1)):
        g.init_from_rho(rho0.ravel())
        t0 = time.perf_counter()
        J, glo, ghi, gka = g.adjoint(nsteps, target, checkpoint_every=every, tape_budget_bytes=budget)
        res[name] = (time.perf_counter() - t0, J)
    g.close()
    return {"config": f"{nx}x{ny}x{nz} PRW Tr=0.75, {nsteps} steps, checkpoints every {every}",
            "forward_s": t_fwd, "adjoint_s": res["recompute"][0], "adjoint_over_forward": res["recompute"][0] / t_fwd,
            "adjoint_taped_s": res["taped"][0], "adjoint_taped_over_forward": res["taped"][0] / t_fwd,
            "J": res["recompute"][1], "J_taped": res["taped"][1]}


def calibrate_wall_mu(cpmp, eos, tr, rho_g, rho_l, angle, mu_b, steps=6000, iters=12):
    """Wall chemical potential giving a target contact angle: bisection on the
    3-D cap-fit angle of a hemisphere of radius 16 on the bottom wall of a
    64x64x48 lattice after `steps` FP64 steps (the angle grows with mu_w).  The
    bracket starts at mu_b -/+ 0.05 (mu_b: the bulk chemical potential at
    coexistence, a neutral wall) and widens until it holds the angle."""
    nx, ny, nz = 64, 64, 48

    def measure(mu_w):
        g = cpmp.Lattice(cpmp.make_params(nx, ny, nz, eos=eos, tr=tr))
        g.set_wall_mu(mu_w, mu_b)
        g.init_spheres([(nx / 2, ny / 2, 2.5, 16.0)], 5.0, rho_g, rho_l, solid_rho=0.0)
        g.step(steps)
        a = g.cap_fit(rho_g, rho_l)["angle"]
        g.close()
        return a

    lo, hi, w = mu_b - 0.05, mu_b + 0.05, 0.05
    for _ in range(6):
        if measure(lo) > angle:
            lo -= 2 * w
        elif measure(hi) < angle:
            hi += 2 * w
        else:
            break
        w *= 2
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        if measure(mid) < angle:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def large_run(cpmp, fwd_steps=10000, adj_steps=2000, every=100):
    """Large FP32 wetting run with adjoint: 512x512x256 van der Waals fluid at
    Tr = 0.8, bottom wall striped along x (period 64: 32 columns at the wall
    chemical potential of 60 degrees, 32 at that of 120 degrees, both calibrated
    with calibrate_wall_mu), top wall neutral, hemisphere of radius 100.  The
    forward run takes fwd_steps steps; the adjoint covers the last adj_steps with
    checkpoints every `every` steps (20 checkpoints), objective the mismatch of
    the final density against the density at the start of the adjoint window.
    Reports MLUPS, roofline fraction, adjoint/forward time and device memory."""
    import numpy as np
    eos, tr = "VDW", 0.8
    A, B, Tc, Rc = cpmp.set_multiphase(eos, tr)
    rho_g, rho_l = cpmp.coexistence(eos, A, B, tr * Tc)
    p0 = cpmp.make_params(8, 8, 8, eos=eos, tr=tr)
    mu_b = p0.K ** 2 * cpmp.eos_mu(eos, A, B, tr * Tc, rho_l)
    t0 = time.perf_counter()
    mu60 = calibrate_wall_mu(cpmp, eos, tr, rho_g, rho_l, 60.0, mu_b)
    mu120 = calibrate_wall_mu(cpmp, eos, tr, rho_g, rho_l, 120.0, mu_b)
    log(f"[large] calibrated wall mu: 60 deg {mu60:.6f}, 120 deg {mu120:.6f} (neutral {mu_b:.6f}), "
        f"{time.perf_counter() - t0:.1f} s")
    nx, ny, nz = 512, 512, 256
    nodes = nx * ny * nz
    lat = cpmp.Lattice(cpmp.make_params(nx, ny, nz, eos=eos, tr=tr, form="fp32"))
    xs = np.repeat(np.arange(nx), ny)
    lat.set_wall_mu(np.where((xs // 32) % 2 == 0, mu60, mu120), mu_b)
    lat.init_spheres([(nx / 2, ny / 2, 2.5, 100.0)], 5.0, rho_g, rho_l, solid_rho=0.0)
    n_plain = fwd_steps - adj_steps
    kt_steps = 20
    log(f"[large] forward {n_plain} steps ...")
    t0 = time.perf_counter()
    lat.step(n_plain - kt_steps)
    t_plain = time.perf_counter() - t0
    ktimes, _ = lat.kernel_times(kt_steps)
    mlups = nodes * (n_plain - kt_steps) / t_plain / 1e6
    bpn = lat.algorithmic_bytes_per_node()
    s = lat.reduce()
    log(f"[large] {mlups:.1f} MLUPS, kernels {ktimes}, mass {s.mass:.6e}")
    target = lat.get("rho")
    ctx_bytes = lat.memory()
    plan = lat.adjoint_plan(adj_steps, every)
    log(f"[large] adjoint plan {plan}, context {ctx_bytes / 2**30:.1f} GiB")
    t0 = time.perf_counter()
    J, glo, ghi, gka = lat.adjoint(adj_steps, target, checkpoint_every=every)
    t_adj = time.perf_counter() - t0
    lat.close()
    t_fwd = t_plain / (n_plain - kt_steps) * adj_steps
    kc = ktimes["collide"]
    out = {"config": f"{nx}x{ny}x{nz} VDW Tr=0.8 FP32, stripes of period 64 (60/120 deg), hemisphere R=100, "
                     f"{fwd_steps} steps, adjoint over the last {adj_steps} with checkpoints every {every}",
           "mu_w_60": mu60, "mu_w_120": mu120, "mu_neutral": mu_b,
           "mlups": mlups, "ms_per_step": t_plain / (n_plain - kt_steps) * 1e3,
           "kernel_ms": ktimes,
           "collide_roofline_frac": bpn * nodes / (kc * 1e-3) / 1e9 / PEAK_HBM_GBS,
           "step_roofline_frac": mlups * 1e6 * bpn / 1e9 / PEAK_HBM_GBS,
           "forward_s": t_fwd, "adjoint_s": t_adj, "adjoint_over_forward": t_adj / t_fwd,
           "J": J, "grad_norm": float(np.linalg.norm(glo)), "dJ_dKa": gka,
           "context_bytes": ctx_bytes, "adjoint_plan": plan,
           "device_bytes_high_water": ctx_bytes + plan["device_bytes"]}
    log(f"[large] adjoint {t_adj:.1f} s over forward {t_fwd:.1f} s = {t_adj / t_fwd:.2f}; "
        f"device memory {(ctx_bytes + plan['device_bytes']) / 2**30:.1f} GiB")
    return out


XGMI_LINK_GBS = 153.0     # one MI355X xGMI link, one direction


def shard_step(cpmp, n=256, gpus=8, steps=200, warmup=20):
    """One GPU's share of the 256^3 FP32 step split into x slabs over `gpus` GPUs:
    the slab x in [64, 64 + n/gpus) of the two-droplet field (it cuts the first
    droplet through its centre), run as a periodic lattice of that shape, which
    stands in for the halo planes the neighbours would send.  Reports the slab's
    step time and kernels beside the bytes one step would exchange per face (the 5
    populations with c_x = +-1 of the boundary plane and 5 planes of rho: radius 2
    for mu and 3 for the force) and their time on one xGMI link."""
    import numpy as np
    nxs = n // gpus
    x0 = 64
    x = np.arange(x0, x0 + nxs, dtype=np.float64)
    y = np.arange(n, dtype=np.float64)
    X, Y, Z = np.meshgrid(x, y, y, indexing="ij")
    phi = np.zeros((nxs, n, n))
    for cx in (80 * n / 256, 176 * n / 256):
        r = np.sqrt((X - cx) ** 2 + (Y - n / 2) ** 2 + (Z - n / 2) ** 2)
        phi = np.maximum(phi, 0.5 * (1 - np.tanh(2 * (r - 48 * n / 256) / 5.0)))
    rho = G_RKE + (L_RKE - G_RKE) * phi
    g = cpmp.Lattice(cpmp.make_params(nxs, n, n, eos="RKE", tr=0.6, nwall=0, form="fp32"))
    g.init_from_rho(rho.ravel())
    g.step(warmup)
    t0 = time.perf_counter()
    g.step(steps)
    dt = time.perf_counter() - t0
    ktimes, _ = g.kernel_times(10)
    g.close()
    nodes = nxs * n * n
    face_bytes = (5 * 4 + 5 * 4) * n * n
    return {"config": f"{nxs}x{n}x{n} FP32 x slab (planes {x0}..{x0 + nxs - 1} of the {n}^3 two-droplet field), "
                      f"periodic in place of the halo exchange, one GPU of {gpus}",
            "ms_per_step": dt / steps * 1e3, "mlups": nodes * steps / dt / 1e6, "kernel_ms": ktimes,
            "exchange_bytes_per_face": face_bytes, "exchange_us_one_link": face_bytes / (XGMI_LINK_GBS * 1e9) * 1e6}


def cpu_baseline(n, steps):
    """The FP64 OpenMP oracle (CPU restatement of the reference step) on the host
    cores, on a bounded sample: the same 256^3 workload for `steps` steps."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    p = orc.make_params(n, n, n, eos="RKE", tr=0.6, nwall=0)
    o = orc.Oracle(p)
    x = np.arange(n, dtype=np.float64)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    phi = np.zeros((n, n, n))
    for cx in (80 * n / 256, 176 * n / 256):
        r = np.sqrt((X - cx) ** 2 + (Y - n / 2) ** 2 + (Z - n / 2) ** 2)
        phi = np.maximum(phi, 0.5 * (1 - np.tanh(2 * (r - 48 * n / 256) / 5.0)))
    rho = G_RKE + (L_RKE - G_RKE) * phi
    del X, Y, Z, phi
    o.init_from_rho(rho.ravel())
    o.step(1)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return dict(value=n ** 3 * steps / dt / 1e6, unit="MLUPS", cores=threads, kind="port", cpu=cpu_model(),
                sample=f"{n}^3 RK coalescence, FP64 oracle, {steps} steps after 1 warm-up step ({dt:.1f} s)")


def cpu_model():
    """the host CPU's model name (/proc/cpuinfo) and its logical CPU count"""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name} ({os.cpu_count()} logical CPUs on the host)"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp32,fp64,fp16,fp32a")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--sessile-steps", type=int, default=20000)
    ap.add_argument("--coalescence-steps", type=int, default=5000)
    ap.add_argument("--adjoint", action="store_true", help="run the adjoint cost leg also with --no-accuracy")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 512x512x256 FP32 wetting run with adjoint (about 75 s, ~270 GB)")
    ap.add_argument("--no-shard", action="store_true",
                    help="skip the 32x256x256 x slab (one GPU's share of the 256^3 lattice over 8 GPUs)")
    ap.add_argument("--large-fwd", type=int, default=10000, help="forward steps of the large run")
    ap.add_argument("--large-adj", type=int, default=2000, help="steps the large run's adjoint covers")
    ap.add_argument("--large-every", type=int, default=100, help="checkpoint interval of the large run's adjoint")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only save a checkpoint of the 256^3 FP32 lattice: its population relayout kernel "
                         "(k_pop_relayout) is a plain copy of 19*N*4 bytes read and written, the known-size "
                         "kernel that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (the lattice "
              "would be split into x slabs, one per GPU, with halo exchange of the "
              "streamed populations and of rho between steps)", file=sys.stderr)
        sys.exit(2)
    cpmp = importlib.import_module("3dcpmpmodel_amd")
    n = args.size
    if args.copy_calibration:
        import tempfile
        g = make_lattice(cpmp, n, "fp32")
        with tempfile.TemporaryDirectory() as d:
            g.checkpoint_save(os.path.join(d, "calibration.ckp"))
        g.close()
        log(f"[bench] copy calibration: k_pop_relayout moved {19 * n ** 3 * 4} bytes each way")
        return
    res = {}
    for form in args.forms.split(","):
        steps = args.steps if form == "fp32" else max(20, args.steps // 5)
        log(f"[bench] {form}: {n}^3, warmup {args.warmup}, timing {steps} steps")
        res[form] = time_form(cpmp, n, form, steps, args.warmup if form == "fp32" else 10)
        log(f"[bench] {form}: {res[form]['mlups']:.1f} MLUPS, {res[form]['ms_per_step']:.3f} ms/step, "
            f"kernels {res[form]['kernel_ms']}")
    main_form = res["fp32"]
    nodes = n ** 3
    kc = main_form["kernel_ms"]["collide"]
    achieved = main_form["bytes_per_node"] * nodes / (kc * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                    frac=round(achieved / PEAK_HBM_GBS, 4),
                    traffic=measured_traffic("k_collide_march_fp32") if n == 256 else None,
                    kernel="k_collide_march (fp32)", bytes_per_node=main_form["bytes_per_node"])
    step_bw = main_form["bytes_per_node"] * nodes / (main_form["ms_per_step"] * 1e-3) / 1e9
    acc = adj = None
    if not args.no_accuracy:
        log("[bench] accuracy of the FP32 / FP16 forms against FP64 ...")
        acc = accuracy(cpmp, args.sessile_steps, args.coalescence_steps)
        log(f"[bench] contact angle err {acc['contact_angle_err_deg']}, coexistence {acc['coexistence_rel_err']}")
    if not args.no_accuracy or args.adjoint:
        log("[bench] adjoint cost ...")
        adj = adjoint_cost(cpmp)
        log(f"[bench] adjoint/forward = {adj['adjoint_over_forward']:.2f} (checkpoints + recompute), "
            f"{adj['adjoint_taped_over_forward']:.2f} (tape in device memory)")
    shard = None
    if n == 256 and "fp32" in res and not args.no_shard:
        shard = shard_step(cpmp, n)
        log(f"[bench] shard {shard['config']}: {shard['mlups']:.1f} MLUPS, {shard['ms_per_step']:.4f} ms/step, "
            f"kernels {shard['kernel_ms']}; exchange {shard['exchange_bytes_per_face'] / 1e6:.2f} MB per face, "
            f"{shard['exchange_us_one_link']:.1f} us on one link")
    large = None
    if not args.no_large:
        large = large_run(cpmp, args.large_fwd, args.large_adj, args.large_every)
    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle) ...")
        cpu = cpu_baseline(n, 18)   # about 10 s of CPU work at 256^3 on 16 threads
        log(f"[bench] cpu: {cpu['value']:.2f} MLUPS on {cpu['cores']} threads")
    out = {
        "metric": "MLUPS 256^3 D3Q19 FP32 vs FP64 at <=0.5deg contact-angle error; adjoint/fwd cost",
        "value": round(main_form["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_form["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"two-droplet Redlich-Kwong coalescence, {n}^3 periodic, Tr=0.6, "
                               f"FP32 populations, FP64 mu and mass sums", "nodes": nodes,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "step_roofline_frac": round(step_bw / PEAK_HBM_GBS, 4),
        # per form: the collide's and the step's fraction of the 8 TB/s roofline at the
        # form's algorithmic population bytes (2 x 19 x storage size per node)
        "forms": {k: {"mlups": round(v["mlups"], 1), "ms_per_step": round(v["ms_per_step"], 4),
                      "kernel_ms": {a: round(b, 4) for a, b in v["kernel_ms"].items()},
                      "bytes_per_node": v["bytes_per_node"], "mass_drift": v["mass_drift"],
                      "collide_roofline_frac": round(v["bytes_per_node"] * nodes / (v["kernel_ms"]["collide"] * 1e-3)
                                                     / 1e9 / PEAK_HBM_GBS, 4),
                      "step_roofline_frac": round(v["bytes_per_node"] * nodes / (v["ms_per_step"] * 1e-3)
                                                  / 1e9 / PEAK_HBM_GBS, 4)}
                  for k, v in res.items()},
        "fp32_over_fp64": round(res["fp32"]["mlups"] / res["fp64"]["mlups"], 3) if "fp64" in res else None,
        "contact_angle_err_deg": acc["contact_angle_err_deg"] if acc else None,
        "coexistence_rel_err": acc["coexistence_rel_err"] if acc else None,
        "laplace_rel_err": acc["laplace_rel_err"] if acc else None,
        "neck_radius_rel_err": acc["neck_radius_rel_err"] if acc else None,
        "adjoint_over_forward": round(adj["adjoint_over_forward"], 3) if adj else None,
        "adjoint_taped_over_forward": round(adj["adjoint_taped_over_forward"], 3) if adj else None,
        "accuracy": acc,
        "adjoint": adj,
        "shard": shard,
        "large": large,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
