"""ctypes binding of the CPU oracle (oracle/dnsbox_oracle.c) for the tests.

The oracle is the checker: tests, ``__graft_entry__.smoke()`` and the
cpu_baseline leg of ``bench.py`` load it; the product never does.

Array conventions (the reference's, one rank, p3dfft STRIDE1):
  physical  numpy (nz, ny, nx) float64     == Fortran u(Nx, Ny, Nz)
  spectral  numpy (nx//2+1, ny, nz) complex128 == Fortran uhat(Nz, Ny, Nx/2+1)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

EKIN, DISP, PROD, COURANT, DIVMAX, EZERO, EONE, ESQRT2, EBAND, NDIAG = range(10)


class OrcParams(C.Structure):
    _fields_ = [
        ("nu", C.c_double), ("Q", C.c_double), ("Pin", C.c_double), ("kcut", C.c_double),
        ("kolmo_amp", C.c_double), ("deltak", C.c_double),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("kolmo_k", C.c_int),
        ("bandlim", C.c_int), ("spherical", C.c_int), ("parity_mode", C.c_int),
        ("nthreads", C.c_int), ("alpha_x", C.c_double), ("alpha_y", C.c_double),
        ("alpha_z", C.c_double),
    ]


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        _lib = C.CDLL(ORACLE_SO)
        d = C.POINTER(C.c_double)
        P = C.POINTER(OrcParams)
        sig = {
            "orc_nspec": (C.c_int, [P]),
            "orc_ktable": (None, [C.c_int, d]),
            "orc_r2c": (None, [P, d, d]),
            "orc_c2r": (None, [P, d, d]),
            "orc_fft1d": (None, [d, C.c_int, C.c_int]),
            "orc_masked": (C.c_int, [P, C.c_int, C.c_int, C.c_int]),
            "orc_eband": (C.c_double, [P, d, d, d]),
            "orc_rhs_nonlinear": (None, [P, d, d, d, d, d, d]),
            "orc_int_fact": (None, [P, C.c_double, d]),
            "orc_heun": (None, [P, d, d, d, C.c_int, C.c_double]),
            "orc_ab2": (None, [P, d, d, d, C.c_int, C.c_double]),
            "orc_ab2_var": (None, [P, d, d, d, C.c_int, d]),
            "orc_ab2_adaptive": (None, [P, d, d, d, C.c_int, d, C.c_double, C.c_double,
                                        C.c_double, d]),
            "orc_project": (None, [P, d, d, d]),
            "orc_dealias": (None, [P, d, d, d]),
            "orc_diagnostics": (None, [P, d, d, d, C.c_double, d, d]),
            "orc_courant": (C.c_double, [P, d, d, d, C.c_double]),
            "orc_cfl_dt": (C.c_double, [C.c_double] * 5),
            "orc_heun_adaptive": (None, [P, d, d, d, C.c_int, d, C.c_double, C.c_double,
                                         C.c_double, d]),
            "orc_init_solenoidal": (None, [P, C.c_uint64, C.c_int, C.c_double, C.c_double,
                                           d, d, d]),
            "orc_init_shapiro": (None, [P, C.c_double, d, d, d]),
            "orc_shapiro_error": (C.c_double, [P, C.c_double, d, d, d]),
            "orc_hash": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
            "orc_rhs_vjp": (None, [P, d, d, d, d, d, d, d, d, d]),
            "orc_heun_vjp": (None, [P, d, d, d, C.c_double, d, d, d]),
            "orc_rhs_adj": (None, [P, d, d, d, d, d, d]),
            "orc_rhs_adj_sign": (None, [P, d, d, d, C.c_double, d, d, d]),
            "orc_adj_descent": (None, [P, d, d, d, C.c_double, C.c_int, C.c_double]),
        }
        for name, (res, args) in sig.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One box of the oracle with the reference's parameters."""

    def __init__(self, n, nu=0.05, Q=0.0, Pin=0.0, kcut=0.0, kolmo_amp=0.0, kolmo_k=0,
                 bandlim=False, spherical=False, parity_mode=True, deltak=1.0, nthreads=0,
                 ny=None, nz=None, alpha=(1.0, 1.0, 1.0)):
        self.nx = n
        self.ny = ny or n
        self.nz = nz or n
        self.alpha = tuple(float(a) for a in alpha)
        self.p = OrcParams(nu, Q, Pin, kcut, kolmo_amp, deltak, self.nx, self.ny, self.nz,
                           kolmo_k, int(bandlim), int(spherical), int(parity_mode), nthreads,
                           *self.alpha)
        self.L = lib()

    # shapes
    @property
    def spec_shape(self):
        return (self.nx // 2 + 1, self.ny, self.nz)

    @property
    def phys_shape(self):
        return (self.nz, self.ny, self.nx)

    def zeros_spec(self, ncomp=3):
        return [np.zeros(self.spec_shape, np.complex128) for _ in range(ncomp)]

    def nspec(self):
        return self.L.orc_nspec(C.byref(self.p))

    def ktable(self, n):
        k = np.zeros(n)
        self.L.orc_ktable(n, _p(k))
        return k

    def r2c(self, phys):
        phys = np.ascontiguousarray(phys, np.float64)
        out = np.zeros(self.spec_shape, np.complex128)
        self.L.orc_r2c(C.byref(self.p), _p(phys), _p(out))
        return out

    def c2r(self, spec):
        spec = np.ascontiguousarray(spec, np.complex128)
        out = np.zeros(self.phys_shape, np.float64)
        self.L.orc_c2r(C.byref(self.p), _p(spec), _p(out))
        return out

    def masked(self, ix, iy, iz):
        return bool(self.L.orc_masked(C.byref(self.p), ix, iy, iz))

    def mask_array(self):
        hx = self.nx // 2 + 1
        m = np.zeros(self.spec_shape, bool)
        for ix in range(hx):
            for iy in range(self.ny):
                for iz in range(self.nz):
                    m[ix, iy, iz] = self.masked(ix, iy, iz)
        return m

    def eband(self, f):
        f = [np.ascontiguousarray(a, np.complex128) for a in f]
        return self.L.orc_eband(C.byref(self.p), *[_p(a) for a in f])

    def rhs(self, f):
        f = [np.ascontiguousarray(a, np.complex128) for a in f]
        out = self.zeros_spec()
        self.L.orc_rhs_nonlinear(C.byref(self.p), *[_p(a) for a in f], *[_p(a) for a in out])
        return out

    def int_fact(self, dt):
        out = np.zeros(self.spec_shape, np.float64)
        self.L.orc_int_fact(C.byref(self.p), dt, _p(out))
        return out

    def heun(self, f, nsteps, dt):
        f = [np.array(a, np.complex128, order="C") for a in f]
        self.L.orc_heun(C.byref(self.p), *[_p(a) for a in f], nsteps, dt)
        return f

    def ab2(self, f, nsteps, dt):
        f = [np.array(a, np.complex128, order="C") for a in f]
        self.L.orc_ab2(C.byref(self.p), *[_p(a) for a in f], nsteps, dt)
        return f

    def ab2_var(self, f, dts):
        f = [np.array(a, np.complex128, order="C") for a in f]
        d = np.ascontiguousarray(dts, np.float64)
        self.L.orc_ab2_var(C.byref(self.p), *[_p(a) for a in f], len(d), _p(d))
        return f

    def ab2_adaptive(self, f, nsteps, dt, cmin, cmax, dtmax, t=0.0):
        f = [np.array(a, np.complex128, order="C") for a in f]
        dtv = np.array([dt])
        tv = np.array([t])
        self.L.orc_ab2_adaptive(C.byref(self.p), *[_p(a) for a in f], nsteps, _p(dtv), cmin,
                                cmax, dtmax, _p(tv))
        return f, float(dtv[0]), float(tv[0])

    def heun_adaptive(self, f, nsteps, dt, cmin, cmax, dtmax, t=0.0):
        f = [np.array(a, np.complex128, order="C") for a in f]
        dtv = np.array([dt])
        tv = np.array([t])
        self.L.orc_heun_adaptive(C.byref(self.p), *[_p(a) for a in f], nsteps, _p(dtv), cmin,
                                 cmax, dtmax, _p(tv))
        return f, float(dtv[0]), float(tv[0])

    def project(self, f):
        f = [np.array(a, np.complex128, order="C") for a in f]
        self.L.orc_project(C.byref(self.p), *[_p(a) for a in f])
        return f

    def dealias(self, f):
        f = [np.array(a, np.complex128, order="C") for a in f]
        self.L.orc_dealias(C.byref(self.p), *[_p(a) for a in f])
        return f

    def diagnostics(self, f, dt):
        f = [np.ascontiguousarray(a, np.complex128) for a in f]
        diag = np.zeros(NDIAG)
        espec = np.zeros(max(self.nspec(), 1))
        self.L.orc_diagnostics(C.byref(self.p), *[_p(a) for a in f], dt, _p(diag), _p(espec))
        return diag, espec

    def courant(self, f, dt):
        f = [np.ascontiguousarray(a, np.complex128) for a in f]
        return self.L.orc_courant(C.byref(self.p), *[_p(a) for a in f], dt)

    def cfl_dt(self, courant, dt, cmin, cmax, dtmax):
        return self.L.orc_cfl_dt(courant, dt, cmin, cmax, dtmax)

    def init_solenoidal(self, seed, box=0, k0=4.0, energy=0.5):
        out = self.zeros_spec()
        self.L.orc_init_solenoidal(C.byref(self.p), seed, box, k0, energy, *[_p(a) for a in out])
        return out

    def init_shapiro(self, t=0.0):
        out = [np.zeros(self.phys_shape) for _ in range(3)]
        self.L.orc_init_shapiro(C.byref(self.p), t, *[_p(a) for a in out])
        return out

    def shapiro_error(self, t, u):
        u = [np.ascontiguousarray(a, np.float64) for a in u]
        return self.L.orc_shapiro_error(C.byref(self.p), t, *[_p(a) for a in u])

    def rhs_vjp(self, f, lam):
        f = [np.ascontiguousarray(a, np.complex128) for a in f]
        lam = [np.ascontiguousarray(a, np.complex128) for a in lam]
        out = self.zeros_spec()
        self.L.orc_rhs_vjp(C.byref(self.p), *[_p(a) for a in f], *[_p(a) for a in lam],
                           *[_p(a) for a in out])
        return out

    def heun_vjp(self, f, lam, dt):
        f = [np.ascontiguousarray(a, np.complex128) for a in f]
        lam = [np.array(a, np.complex128, order="C") for a in lam]
        self.L.orc_heun_vjp(C.byref(self.p), *[_p(a) for a in f], dt, *[_p(a) for a in lam])
        return lam

    def rhs_adj(self, f):
        f = [np.ascontiguousarray(a, np.complex128) for a in f]
        out = self.zeros_spec()
        self.L.orc_rhs_adj(C.byref(self.p), *[_p(a) for a in f], *[_p(a) for a in out])
        return out

    def rhs_adj_sign(self, f, lin_sign):
        f = [np.ascontiguousarray(a, np.complex128) for a in f]
        out = self.zeros_spec()
        self.L.orc_rhs_adj_sign(C.byref(self.p), *[_p(a) for a in f], float(lin_sign),
                                *[_p(a) for a in out])
        return out

    def adj_descent(self, f, dtau, nsteps, lin_sign=1):
        f = [np.array(a, np.complex128, order="C") for a in f]
        self.L.orc_adj_descent(C.byref(self.p), *[_p(a) for a in f], dtau, nsteps, float(lin_sign))
        return f

    def hermitian_weights(self):
        w = np.full(self.spec_shape, 2.0)
        w[0] = 1.0
        w[self.nx // 2] = 1.0
        return w

    def hash(self, seed, box, idx):
        return self.L.orc_hash(seed, box, idx)


def to_torch_layout(spec):
    """oracle spectral (nxh, ny, nz) -> (nz, ny, nxh) == numpy/torch rfftn of (nz, ny, nx)."""
    return np.ascontiguousarray(np.transpose(spec, (2, 1, 0)))


def from_torch_layout(spec):
    return np.ascontiguousarray(np.transpose(spec, (2, 1, 0)))
