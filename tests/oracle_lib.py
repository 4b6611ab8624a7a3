"""ctypes binding of the CPU FP64 oracle (oracle/libcpmp_oracle.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg only, as the checker.  Arrays use the reference layout
I(x,y,z) = (x*DY + y)*DZ + z (Common.h:164).
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "libcpmp_oracle.so")

EOS = {"VDW": 0, "CSE": 1, "RKE": 2, "RKS": 3, "PRW": 4, "PRM": 5}
TYPE, RHO, MU, F, VEL, MV, FORCE = range(7)
Q = 19


class OrcParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("nwall_lo", C.c_int), ("nwall_hi", C.c_int),
                ("eos", C.c_int), ("mu_form", C.c_int), ("lap_points", C.c_int),
                ("wall_scheme", C.c_int),
                ("A", C.c_double), ("B", C.c_double), ("T", C.c_double),
                ("K", C.c_double), ("Ka", C.c_double), ("K1", C.c_double), ("K2", C.c_double),
                ("Ts", C.c_double * 19), ("forcing", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle")],
                                  stdout=subprocess.DEVNULL)
        L = C.CDLL(LIB_PATH)
        d, i, p = C.c_double, C.c_int, C.c_void_p
        dp = C.POINTER(C.c_double)
        L.orc_set_multiphase.argtypes = [i, d, dp, dp, dp, dp]
        L.orc_pressure.argtypes = [i, d, d, d, d]
        L.orc_pressure.restype = d
        L.orc_mu_eos.argtypes = [i, i, d, d, d, d]
        L.orc_mu_eos.restype = d
        L.orc_maxwell.argtypes = [i, d, d, d, dp, dp]
        L.orc_tables.argtypes = [p] * 7
        L.orc_create.argtypes = [C.POINTER(OrcParams)]
        L.orc_create.restype = p
        L.orc_destroy.argtypes = [p]
        L.orc_get.argtypes = [p, i, p]
        L.orc_set.argtypes = [p, i, p]
        L.orc_set_wall_mu.argtypes = [p, p, p]
        L.orc_init_reference.argtypes = [p, d, d, d, d, d]
        L.orc_init_from_rho.argtypes = [p, p]
        L.orc_step.argtypes = [p, i]
        L.orc_stage.argtypes = [p, i]
        L.orc_reduce.argtypes = [p, p]
        L.orc_contact_angle.argtypes = [C.POINTER(OrcParams), p, d, d, i, p]
        L.orc_noise.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_noise.restype = d
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_adjoint.argtypes = [p, i, p, p, p, p, p]
        L.orc_adjoint.restype = i
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def set_multiphase(eos, tr):
    A, B, Tc, Rc = (C.c_double() for _ in range(4))
    rc = lib().orc_set_multiphase(EOS[eos], tr, C.byref(A), C.byref(B), C.byref(Tc), C.byref(Rc))
    assert rc == 0
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
                nwall=3, nwall_hi=None, mu_form=0, lap_points=5, wall_scheme=1, ts=None, forcing=0):
    A, B, Tc, Rc = set_multiphase(eos, tr)
    p = OrcParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.nwall_lo = nwall
    p.nwall_hi = nwall if nwall_hi is None else nwall_hi
    p.eos, p.mu_form, p.lap_points, p.wall_scheme = EOS[eos], mu_form, lap_points, wall_scheme
    p.A, p.B, p.T = A, B, tr * Tc
    p.K, p.Ka, p.K1, p.K2 = K, Ka, K1, K2
    for q, v in enumerate(ts if ts is not None else reference_ts(tau)):
        p.Ts[q] = v
    p.forcing = forcing
    return p


class Oracle:
    """One oracle lattice state."""

    def __init__(self, params):
        self.p = params
        self.n = params.nx * params.ny * params.nz
        self.h = lib().orc_create(C.byref(params))
        if not self.h:
            raise ValueError("orc_create rejected the parameters")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def get(self, field):
        size = {F: Q * self.n, VEL: 3 * self.n, MV: 3 * self.n, FORCE: 3 * self.n}.get(field, self.n)
        out = np.empty(size, dtype=np.float64)
        assert lib().orc_get(self.h, field, _ptr(out)) == 0
        return out

    def set(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert lib().orc_set(self.h, field, _ptr(a)) == 0

    def set_wall_mu(self, lo, hi=None):
        nxy = self.p.nx * self.p.ny
        lo = np.ascontiguousarray(np.broadcast_to(np.asarray(lo, np.float64), (nxy,)))
        hi = lo if hi is None else np.ascontiguousarray(np.broadcast_to(np.asarray(hi, np.float64), (nxy,)))
        lib().orc_set_wall_mu(self.h, _ptr(lo), _ptr(hi))
        self._keep = (lo, hi)

    def init_reference(self, radius, width, den_g, den_l, basept_init=0.0):
        lib().orc_init_reference(self.h, radius, width, den_g, den_l, basept_init)

    def init_from_rho(self, rho):
        a = np.ascontiguousarray(rho, dtype=np.float64)
        lib().orc_init_from_rho(self.h, _ptr(a))

    def step(self, n=1):
        lib().orc_step(self.h, n)

    def stage(self, s):
        lib().orc_stage(self.h, s)

    def reduce(self):
        out = np.zeros(3)
        lib().orc_reduce(self.h, _ptr(out))
        return out

    def adjoint(self, nsteps, target):
        nxy = self.p.nx * self.p.ny
        t = np.ascontiguousarray(target, dtype=np.float64)
        J = np.zeros(1)
        glo, ghi, gka = np.zeros(nxy), np.zeros(nxy), np.zeros(1)
        if lib().orc_adjoint(self.h, nsteps, _ptr(t), _ptr(J), _ptr(glo), _ptr(ghi), _ptr(gka)) != 0:
            raise ValueError("the oracle adjoint does not cover these parameters")
        return J[0], glo, ghi, gka[0]


def tables():
    ex, ey, ez, re = (np.zeros(19, np.int32) for _ in range(4))
    w = np.zeros(19)
    M, R = np.zeros((19, 19)), np.zeros((19, 19))
    lib().orc_tables(_ptr(ex), _ptr(ey), _ptr(ez), _ptr(re), _ptr(w), _ptr(M), _ptr(R))
    return ex, ey, ez, re, w, M, R


def pressure(eos, A, B, T, rho):
    return lib().orc_pressure(EOS[eos], A, B, T, rho)


def mu_eos(eos, A, B, T, rho, form=0):
    return lib().orc_mu_eos(EOS[eos], form, A, B, T, rho)


def maxwell(eos, A, B, T, g0, l0):
    g, l = C.c_double(g0), C.c_double(l0)
    rc = lib().orc_maxwell(EOS[eos], A, B, T, C.byref(g), C.byref(l))
    return rc, g.value, l.value


def contact_angle(params, rho, den_g, den_l, tier):
    """CaLeft, CaRight, CaCap, XLeft, XRight, H, L, valid; None when the wall is
    too thin for the scan"""
    out = np.zeros(8)
    a = np.ascontiguousarray(rho, dtype=np.float64)
    if lib().orc_contact_angle(C.byref(params), _ptr(a), den_g, den_l, tier, _ptr(out)) != 0:
        return None
    return out


def splitmix64(x):
    return lib().orc_splitmix64(x)


def noise(seed, index):
    return lib().orc_noise(seed, index)
