"""ctypes front end of libcpmp.so, the MI355X chemical-potential multiphase
D3Q19 MRT lattice-Boltzmann library (C ABI in include/cpmp.h).

The package directory name starts with a digit, so import it with
``importlib.import_module("3dcpmpmodel_amd")``.  Loading fails with an
ImportError when the HIP library has not been built: there is no CPU
fallback for any compute entry point.

Host arrays use the reference layout I(x,y,z) = (x*DY + y)*DZ + z
(Common.h:164) and FP64 values.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcpmp.so")

EOS = {"VDW": 0, "CSE": 1, "RKE": 2, "RKS": 3, "PRW": 4, "PRM": 5}
FORM = {"fp64": 0, "fp32": 1, "fp16": 2, "fp32a": 3}
FIELD = {"type": 0, "rho": 1, "mu": 2, "f": 3, "vel": 4, "mv": 5, "force": 6}
Q = 19

# every symbol include/cpmp.h declares
EXPORTS = [
    "cpmp_set_multiphase", "cpmp_default_params", "cpmp_dim_conversion", "cpmp_eos_pressure", "cpmp_eos_mu",
    "cpmp_coexistence", "cpmp_cocurve_lookup", "cpmp_create", "cpmp_destroy", "cpmp_get_params",
    "cpmp_memory", "cpmp_init_reference", "cpmp_init_splash", "cpmp_init_spheres", "cpmp_init_from_rho",
    "cpmp_set_wall_mu", "cpmp_step", "cpmp_sync", "cpmp_step_count", "cpmp_get_field",
    "cpmp_set_field", "cpmp_reduce", "cpmp_reduce_noalloc", "cpmp_contact_angle", "cpmp_cap_fit", "cpmp_probe",
    "cpmp_laplace_jump",
    "cpmp_checkpoint_save", "cpmp_checkpoint_load", "cpmp_write_plt", "cpmp_adjoint", "cpmp_adjoint_plan", "cpmp_neck_radius",
    "cpmp_kernel_times", "cpmp_time_steps", "cpmp_algorithmic_bytes_per_node",
    "cpmp_last_error", "cpmp_version", "cpmp_selftest_collide", "cpmp_selftest_collide_guo",
]


class Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("nwall_lo", C.c_int), ("nwall_hi", C.c_int),
                ("eos", C.c_int), ("mu_form", C.c_int), ("lap_points", C.c_int),
                ("wall_scheme", C.c_int), ("form", C.c_int), ("device", C.c_int),
                ("A", C.c_double), ("B", C.c_double), ("T", C.c_double),
                ("K", C.c_double), ("Ka", C.c_double), ("K1", C.c_double), ("K2", C.c_double),
                ("Ts", C.c_double * 19), ("rho0", C.c_double), ("forcing", C.c_int)]


class InitDesc(C.Structure):
    _fields_ = [("nspheres", C.c_int),
                ("cx", C.c_double * 8), ("cy", C.c_double * 8), ("cz", C.c_double * 8),
                ("radius", C.c_double * 8), ("width", C.c_double),
                ("rho_g", C.c_double), ("rho_l", C.c_double), ("invert", C.c_int),
                ("noise_amp", C.c_double), ("seed", C.c_uint64), ("solid_rho", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("mass", C.c_double), ("max_speed", C.c_double), ("bad_count", C.c_int64),
                ("first_bad", C.c_int64), ("rho_min", C.c_double), ("rho_max", C.c_double),
                ("momentum", C.c_double * 3), ("kinetic", C.c_double)]


class AdjDesc(C.Structure):
    _fields_ = [("nsteps", C.c_int), ("checkpoint_every", C.c_int),
                ("target", C.POINTER(C.c_double)), ("tape_budget_bytes", C.c_int64),
                ("memory_limit_bytes", C.c_int64)]


class AdjPlan(C.Structure):
    _fields_ = [("device_bytes", C.c_int64), ("taped_steps", C.c_int), ("segments", C.c_int),
                ("sub_length", C.c_int), ("sub_checkpoints", C.c_int), ("recomputed_steps", C.c_int64),
                ("overlapped", C.c_int)]


class CPMPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cpmp error {code}: {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    for name in EXPORTS:
        getattr(L, name)          # raises AttributeError if a symbol is missing
    d, i, p = C.c_double, C.c_int, C.c_void_p
    dp = C.POINTER(C.c_double)
    L.cpmp_set_multiphase.argtypes = [i, d, dp, dp, dp, dp]
    L.cpmp_default_params.argtypes = [C.POINTER(Params), i, i, i, i, d]
    L.cpmp_dim_conversion.argtypes = [d, d, d, d, d, d, dp]
    L.cpmp_eos_pressure.argtypes = [i, d, d, d, d]
    L.cpmp_eos_pressure.restype = d
    L.cpmp_eos_mu.argtypes = [i, i, d, d, d, d]
    L.cpmp_eos_mu.restype = d
    L.cpmp_coexistence.argtypes = [i, d, d, d, dp, dp]
    L.cpmp_cocurve_lookup.argtypes = [C.c_char_p, d, dp, dp]
    L.cpmp_create.argtypes = [C.POINTER(Params), C.POINTER(p)]
    L.cpmp_destroy.argtypes = [p]
    L.cpmp_get_params.argtypes = [p, C.POINTER(Params)]
    L.cpmp_memory.argtypes = [p, C.POINTER(C.c_int64)]
    L.cpmp_init_reference.argtypes = [p, d, d, d, d, d]
    L.cpmp_init_splash.argtypes = [p, d, d, d, d, d]
    L.cpmp_init_spheres.argtypes = [p, C.POINTER(InitDesc)]
    L.cpmp_init_from_rho.argtypes = [p, p]
    L.cpmp_set_wall_mu.argtypes = [p, p, p, C.c_int64]
    L.cpmp_step.argtypes = [p, i]
    L.cpmp_sync.argtypes = [p]
    L.cpmp_step_count.argtypes = [p]
    L.cpmp_step_count.restype = C.c_int64
    L.cpmp_get_field.argtypes = [p, i, p]
    L.cpmp_set_field.argtypes = [p, i, p]
    L.cpmp_reduce.argtypes = [p, C.POINTER(Stats)]
    L.cpmp_reduce_noalloc.argtypes = [p, C.POINTER(Stats)]
    L.cpmp_contact_angle.argtypes = [p, i, d, d, p]
    L.cpmp_cap_fit.argtypes = [p, d, d, p]
    L.cpmp_checkpoint_save.argtypes = [p, C.c_char_p]
    L.cpmp_checkpoint_load.argtypes = [p, C.c_char_p]
    L.cpmp_write_plt.argtypes = [p, C.c_char_p, C.c_char_p, i, i]
    L.cpmp_adjoint.argtypes = [p, C.POINTER(AdjDesc), p, p, p, p]
    L.cpmp_adjoint_plan.argtypes = [p, C.POINTER(AdjDesc), C.POINTER(AdjPlan)]
    L.cpmp_neck_radius.argtypes = [p, C.c_int, C.c_double, C.c_double, p, p]
    L.cpmp_probe.argtypes = [p, i, i, i, i, p]
    L.cpmp_laplace_jump.argtypes = [p, p, p, p]
    L.cpmp_kernel_times.argtypes = [p, i, p, p]
    L.cpmp_time_steps.argtypes = [p, i, dp]
    L.cpmp_algorithmic_bytes_per_node.argtypes = [p]
    L.cpmp_algorithmic_bytes_per_node.restype = d
    L.cpmp_last_error.restype = C.c_char_p
    L.cpmp_version.restype = C.c_char_p
    L.cpmp_selftest_collide.argtypes = [p, d, p, p, p, p, p]
    L.cpmp_selftest_collide_guo.argtypes = [p, d, p, p, p, p, p]
    return L


lib = _load()


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def check(rc):
    if rc != 0:
        raise CPMPError(rc, lib.cpmp_last_error().decode())
    return rc


def set_multiphase(eos, tr):
    """SetMultiphase constants (FluidCuda.cu:1318-1344): A, B, Tc, Rc."""
    A, B, Tc, Rc = (C.c_double() for _ in range(4))
    check(lib.cpmp_set_multiphase(EOS[eos], tr, C.byref(A), C.byref(B), C.byref(Tc), C.byref(Rc)))
    return A.value, B.value, Tc.value, Rc.value


def reference_ts(tau):
    """Relaxation rates of FluidCuda.cu:154-157."""
    ts = [1.0] * 19
    for q in (2, 4, 6, 8):
        ts[q] = 1.1
    for q in (9, 11, 13, 14, 15):
        ts[q] = 1.0 / tau
    return ts


def make_params(nx, ny, nz, eos="PRW", tr=0.7, tau=0.8, K=0.2, Ka=0.001, K1=0.0, K2=0.0,
                nwall=3, nwall_hi=None, mu_form=0, lap_points=5, wall_scheme=1, form="fp64",
                rho0=0.0, device=0, ts=None, forcing=0):
    p = Params()
    check(lib.cpmp_default_params(C.byref(p), nx, ny, nz, EOS[eos], tr))
    p.nwall_lo = nwall
    p.nwall_hi = nwall if nwall_hi is None else nwall_hi
    p.mu_form, p.lap_points, p.wall_scheme = mu_form, lap_points, wall_scheme
    p.form = FORM[form] if isinstance(form, str) else form
    p.device = device
    p.K, p.Ka, p.K1, p.K2 = K, Ka, K1, K2
    for q, v in enumerate(ts if ts is not None else reference_ts(tau)):
        p.Ts[q] = v
    p.rho0 = rho0
    p.forcing = forcing
    return p


def coexistence(eos, A, B, T, g0=0.0, l0=0.0):
    g, l = C.c_double(g0), C.c_double(l0)
    check(lib.cpmp_coexistence(EOS[eos], A, B, T, C.byref(g), C.byref(l)))
    return g.value, l.value


def cocurve_lookup(path, tr):
    g, l = C.c_double(), C.c_double()
    check(lib.cpmp_cocurve_lookup(path.encode(), tr, C.byref(g), C.byref(l)))
    return g.value, l.value


def dim_conversion(length, macro_length, density, macro_density, viscosity, macro_viscosity):
    """DimConversion (FluidCuda.cu:474-486): dict of the lattice-to-physical factors"""
    out = (C.c_double * 9)()
    check(lib.cpmp_dim_conversion(length, macro_length, density, macro_density, viscosity, macro_viscosity, out))
    return dict(zip(["len", "time", "mass", "force", "torque", "velocity", "acceleration", "omega", "beta"], list(out)))


def eos_pressure(eos, A, B, T, rho):
    return lib.cpmp_eos_pressure(EOS[eos], A, B, T, rho)


def eos_mu(eos, A, B, T, rho, mu_form=0):
    return lib.cpmp_eos_mu(EOS[eos], mu_form, A, B, T, rho)


def selftest_collide(f, rho, F, ts, qp=None, forcing=0):
    """The device collision of one node evaluated on the CPU: (out, u); forcing 1:
    the Guo force term instead of the exact-difference one."""
    f = np.ascontiguousarray(f, np.float64)
    F = np.ascontiguousarray(F, np.float64)
    ts = np.ascontiguousarray(ts, np.float64)
    out, u = np.zeros(19), np.zeros(3)
    qpp = None if qp is None else np.ascontiguousarray(qp, np.float64)
    fn = lib.cpmp_selftest_collide_guo if forcing == 1 else lib.cpmp_selftest_collide
    check(fn(_ptr(f), rho, _ptr(F), _ptr(ts), None if qpp is None else _ptr(qpp), _ptr(out), _ptr(u)))
    return out, u


class Lattice:
    """One device lattice (a cpmp_ctx)."""

    def __init__(self, params):
        self.p = params
        self.n = params.nx * params.ny * params.nz
        self.h = C.c_void_p()
        check(lib.cpmp_create(C.byref(params), C.byref(self.h)))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib.cpmp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # initial conditions (SetFlowField, FluidCuda.cu:199-241)
    def init_reference(self, radius, width, rho_g, rho_l, solid_value=0.0):
        check(lib.cpmp_init_reference(self.h, radius, width, rho_g, rho_l, solid_value))

    def init_splash(self, rho_g, rho_l, radius=50.0, width=10.0, film_z=32.0):
        """SetFlowField DROPLETSPLASH (FluidCuda.cu:219-233)"""
        check(lib.cpmp_init_splash(self.h, radius, width, film_z, rho_g, rho_l))

    def init_spheres(self, spheres, width, rho_g, rho_l, invert=False, noise_amp=0.0, seed=0,
                     solid_rho=0.0):
        d = InitDesc()
        d.nspheres = len(spheres)
        for s, (cx, cy, cz, r) in enumerate(spheres):
            d.cx[s], d.cy[s], d.cz[s], d.radius[s] = cx, cy, cz, r
        d.width, d.rho_g, d.rho_l = width, rho_g, rho_l
        d.invert, d.noise_amp, d.seed, d.solid_rho = int(invert), noise_amp, seed, solid_rho
        check(lib.cpmp_init_spheres(self.h, C.byref(d)))

    def init_from_rho(self, rho):
        a = np.ascontiguousarray(rho, np.float64).ravel()
        assert a.size == self.n
        check(lib.cpmp_init_from_rho(self.h, _ptr(a)))

    def set_wall_mu(self, lo, hi=None):
        lo = np.ascontiguousarray(np.atleast_1d(lo), np.float64).ravel()
        hi_a = None if hi is None else np.ascontiguousarray(np.atleast_1d(hi), np.float64).ravel()
        if hi_a is not None and hi_a.size != lo.size:
            hi_a = np.ascontiguousarray(np.broadcast_to(hi_a, lo.shape))
        check(lib.cpmp_set_wall_mu(self.h, _ptr(lo), None if hi_a is None else _ptr(hi_a), lo.size))

    def step(self, n=1):
        check(lib.cpmp_step(self.h, n))

    @property
    def step_count(self):
        return lib.cpmp_step_count(self.h)

    def get(self, field):
        fid = FIELD[field] if isinstance(field, str) else field
        size = {3: Q * self.n, 4: 3 * self.n, 5: 3 * self.n, 6: 3 * self.n}.get(fid, self.n)
        out = np.empty(size, np.float64)
        check(lib.cpmp_get_field(self.h, fid, _ptr(out)))
        return out

    def set(self, field, arr):
        fid = FIELD[field] if isinstance(field, str) else field
        a = np.ascontiguousarray(arr, np.float64).ravel()
        check(lib.cpmp_set_field(self.h, fid, _ptr(a)))

    def reduce(self, raise_on_nan=True):
        """CalcMacroCPU (FluidCuda.cu:1626-1657) on the device, FP64."""
        s = Stats()
        rc = lib.cpmp_reduce(self.h, C.byref(s))
        if rc != 0 and (raise_on_nan or rc != -5):
            check(rc)
        return s

    def contact_angle(self, tier, rho_g, rho_l):
        out = np.zeros(8)
        check(lib.cpmp_contact_angle(self.h, tier, rho_g, rho_l, _ptr(out)))
        keys = ["CaLeft", "CaRight", "CaCap", "XLeft", "XRight", "H", "L", "valid"]
        return dict(zip(keys, out))

    def cap_fit(self, rho_g, rho_l):
        out = np.zeros(4)
        check(lib.cpmp_cap_fit(self.h, rho_g, rho_l, _ptr(out)))
        return dict(zip(["angle", "height", "base_radius", "radius"], out))

    def checkpoint_save(self, path):
        check(lib.cpmp_checkpoint_save(self.h, path.encode()))

    def checkpoint_load(self, path):
        check(lib.cpmp_checkpoint_load(self.h, path.encode()))

    def write_plt(self, path, title="Model3D", decimate=1, decimate_linear=False):
        check(lib.cpmp_write_plt(self.h, path.encode(), title.encode(), decimate, int(decimate_linear)))

    def _adj_desc(self, nsteps, checkpoint_every, tape_budget_bytes, memory_limit_bytes, target=None):
        d = AdjDesc()
        d.nsteps, d.checkpoint_every = nsteps, checkpoint_every
        d.tape_budget_bytes = int(tape_budget_bytes)
        d.memory_limit_bytes = int(memory_limit_bytes)
        if target is not None:
            d.target = target.ctypes.data_as(C.POINTER(C.c_double))
        return d

    def adjoint(self, nsteps, target, checkpoint_every=100, want_hi=True, tape_budget_bytes=0,
                memory_limit_bytes=0):
        """cpmp_adjoint.  tape_budget_bytes: trailing steps taped in the first
        forward run instead of recomputed (0: recompute everything, -1: half of
        the free device memory).  memory_limit_bytes: device bytes the call may
        allocate (0: the free device memory less 2 GiB); segments whose tape does
        not fit are recomputed in sub-segments."""
        nxy = self.p.nx * self.p.ny
        t = np.ascontiguousarray(target, np.float64).ravel()
        d = self._adj_desc(nsteps, checkpoint_every, tape_budget_bytes, memory_limit_bytes, t)
        J = np.zeros(1)
        glo, ghi, gka = np.zeros(nxy), np.zeros(nxy), np.zeros(1)
        check(lib.cpmp_adjoint(self.h, C.byref(d), _ptr(J), _ptr(glo), _ptr(ghi) if want_hi else None,
                               _ptr(gka)))
        return J[0], glo, ghi, gka[0]

    def adjoint_plan(self, nsteps, checkpoint_every=100, tape_budget_bytes=0, memory_limit_bytes=0):
        """cpmp_adjoint_plan: the schedule and device bytes of an adjoint call"""
        d = self._adj_desc(nsteps, checkpoint_every, tape_budget_bytes, memory_limit_bytes)
        q = AdjPlan()
        check(lib.cpmp_adjoint_plan(self.h, C.byref(d), C.byref(q)))
        return {k: getattr(q, k) for k, _ in AdjPlan._fields_}

    def probe(self, field, x, y, z):
        """one node's rho or mu (cpmp_probe)"""
        fid = FIELD[field] if isinstance(field, str) else field
        v = np.zeros(1)
        check(lib.cpmp_probe(self.h, fid, int(x), int(y), int(z), _ptr(v)))
        return float(v[0])

    def neck_radius(self, x, rho_g, rho_l):
        """cpmp_neck_radius: (sqrt(A / pi), A) of the liquid fraction in plane x"""
        r, a = np.zeros(1), np.zeros(1)
        check(lib.cpmp_neck_radius(self.h, int(x), rho_g, rho_l, _ptr(r), _ptr(a)))
        return float(r[0]), float(a[0])

    def laplace_jump(self, inside, outside):
        """K^2 [p(rho(inside)) - p(rho(outside))] (cpmp_laplace_jump)"""
        a = np.ascontiguousarray(inside, np.int32)
        b = np.ascontiguousarray(outside, np.int32)
        v = np.zeros(1)
        check(lib.cpmp_laplace_jump(self.h, a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p), _ptr(v)))
        return float(v[0])

    def kernel_times(self, nsteps):
        ms = np.zeros(4)
        per = np.zeros(4, np.int32)
        check(lib.cpmp_kernel_times(self.h, nsteps, _ptr(ms), _ptr(per)))
        return dict(zip(["wall", "mu", "collide", "macro"], ms)), dict(zip(["wall", "mu", "collide", "macro"], per))

    def time_steps(self, nsteps):
        ms = C.c_double()
        check(lib.cpmp_time_steps(self.h, nsteps, C.byref(ms)))
        return ms.value

    def algorithmic_bytes_per_node(self):
        return lib.cpmp_algorithmic_bytes_per_node(self.h)

    def memory(self):
        b = C.c_int64()
        check(lib.cpmp_memory(self.h, C.byref(b)))
        return b.value
