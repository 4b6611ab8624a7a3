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
    return dict(value=n ** 3 * steps / dt / 1e6, unit="MLUPS", cores=threads, kind="port",
                sample=f"{n}^3 RK coalescence, FP64 oracle, {steps} steps after 1 warm-up step ({dt:.1f} s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp32,fp64,fp16")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the domain decomposition design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    cpmp = importlib.import_module("3dcpmpmodel_amd")
    n = args.size
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
                    frac=round(achieved / PEAK_HBM_GBS, 4), traffic=None,
                    kernel="k_collide (fp32)", bytes_per_node=main_form["bytes_per_node"])
    step_bw = main_form["bytes_per_node"] * nodes / (main_form["ms_per_step"] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle) ...")
        cpu = cpu_baseline(n, 3)
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
        "forms": {k: {"mlups": round(v["mlups"], 1), "ms_per_step": round(v["ms_per_step"], 4),
                      "kernel_ms": {a: round(b, 4) for a, b in v["kernel_ms"].items()},
                      "bytes_per_node": v["bytes_per_node"], "mass_drift": v["mass_drift"]}
                  for k, v in res.items()},
        "fp32_over_fp64": round(res["fp32"]["mlups"] / res["fp64"]["mlups"], 3) if "fp64" in res else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
