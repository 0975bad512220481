"""ctypes front end of the CPU oracle (oracle/libql_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker of the HIP path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module; the product package never does.

`Oracle` mirrors the reference objects it restates: ``rhs``/``linear_ops_init``
are ``Model::rhs``/``Model::linearOPs_Init`` (Models/Model.h:46-50), ``step`` is
``Integrator::Step`` (Integrators/Integrator.h:23), ``steps`` is the time loop
of main.cpp:108-123, ``remap`` is ``ShearingBox_Remap`` and ``diagnostics`` is
``Calc_Energy_AM_Diss`` with every save flag on.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MODEL_MHD_FULLQLIN, MODEL_MHD_BQLIN, MODEL_CONSTANT_DAMPING, MODEL_LINEAR_ONLY, MODEL_HD_FULLU = 0, 1, 2, 3, 4


class _Params(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int), ("integrator", ctypes.c_int),
                ("NX", ctypes.c_int), ("NY", ctypes.c_int), ("NZ", ctypes.c_int),
                ("L", ctypes.c_double * 3), ("q", ctypes.c_double), ("omega", ctypes.c_double),
                ("nu", ctypes.c_double), ("eta", ctypes.c_double), ("B0z", ctypes.c_double),
                ("viscosity_order", ctypes.c_int), ("quasilinear", ctypes.c_int),
                ("remap", ctypes.c_int), ("dt", ctypes.c_double), ("nthreads", ctypes.c_int),
                ("CFL", ctypes.c_double)]


def build(native: bool = False) -> str:
    target = "native" if native else "all"
    subprocess.run(["make", "-s", "-C", _HERE, target], check=True)
    return os.path.join(_HERE, "libql_oracle_native.so" if native else "libql_oracle.so")


def lib(native: bool = False):
    global _LIB
    name = "libql_oracle_native.so" if native else "libql_oracle.so"
    path = os.path.join(_HERE, name)
    if not native and _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        build(native)
    L = bind(ctypes.CDLL(path))
    if not native:
        _LIB = L
    return L


def bind(L):
    """Declares the argument and result types of the oracle's C entry points
    on a loaded library (the default build, the timing build, or a variant
    compiled by a test)."""
    P, D, I = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    L.orc_create.restype = P
    L.orc_create.argtypes = [ctypes.POINTER(_Params)]
    L.orc_destroy.argtypes = [P]
    L.orc_dims.argtypes = [P, I]
    L.orc_set_state.argtypes = [P, D, D, D]
    L.orc_get_state.argtypes = [P, D, D, D]
    L.orc_tables.argtypes = [P, D, D, I, I, D]
    L.orc_linear_ops_init.argtypes = [P, ctypes.c_double, D, D]
    L.orc_rhs.argtypes = [P, ctypes.c_double, ctypes.c_double, D, D, D]
    U = ctypes.POINTER(ctypes.c_uint64)
    L.orc_set_noise.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                ctypes.c_uint64]
    L.orc_driving_noise.argtypes = [P, ctypes.c_double, ctypes.c_double]
    L.orc_noise_normals.argtypes = [P, ctypes.c_double, ctypes.c_int, ctypes.c_int, D]
    L.orc_philox4x64.argtypes = [U, U, U]
    L.orc_calculate_cfl.restype = ctypes.c_double
    L.orc_calculate_cfl.argtypes = [P]
    L.orc_check_sums.argtypes = [P, D]
    L.orc_check_solution.argtypes = [P]
    L.orc_step.restype = ctypes.c_double
    L.orc_step.argtypes = [P, ctypes.c_double]
    L.orc_set_dt.argtypes = [P, ctypes.c_double]
    L.orc_steps.restype = ctypes.c_double
    L.orc_steps.argtypes = [P, ctypes.c_int, ctypes.c_double]
    L.orc_remap.argtypes = [P, ctypes.c_double]
    L.orc_diagnostics.argtypes = [P, ctypes.c_double, D]
    L.orc_get_drive.argtypes = [P, D]
    L.orc_mf_real.argtypes = [P, D]
    L.orc_fft_inverse.argtypes = [ctypes.c_int, ctypes.c_int, D, D]
    L.orc_fft_forward.argtypes = [ctypes.c_int, ctypes.c_int, D, D]
    L.orc_fft_plain.argtypes = [ctypes.c_int, ctypes.c_int, D, D]
    return L


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def philox4x64(ctr, key) -> np.ndarray:
    """Philox4x64-10 block of counter ctr[4] under key[2]."""
    c = np.ascontiguousarray(ctr, dtype=np.uint64)
    k = np.ascontiguousarray(key, dtype=np.uint64)
    out = np.zeros(4, dtype=np.uint64)
    P = ctypes.POINTER(ctypes.c_uint64)
    lib().orc_philox4x64(c.ctypes.data_as(P), k.ctypes.data_as(P), out.ctypes.data_as(P))
    return out


def fft_inverse(NZ: int, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros(NZ, dtype=np.complex128)
    lib().orc_fft_inverse(NZ, x.size, _dp(x), _dp(out))
    return out


def fft_forward(nz: int, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros(nz, dtype=np.complex128)
    lib().orc_fft_forward(x.size, nz, _dp(x), _dp(out))
    return out


def fft_plain(x: np.ndarray, sign: int) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    lib().orc_fft_plain(x.size, sign, _dp(x), _dp(out))
    return out


class Oracle:
    def __init__(self, model=0, integrator=1, NX=16, NY=16, NZ=32, L=(1.0, 4.0, 1.0), q=1.5,
                 omega=1.0, nu=5e-4, eta=5e-4, B0z=0.0, viscosity_order=2, quasilinear=1,
                 remap=1, dt=0.01, CFL=1.0, nthreads=1, native=False, library=None):
        self._lib = library if library is not None else lib(native)
        p = _Params(model, integrator, NX, NY, NZ, (ctypes.c_double * 3)(*L), q, omega, nu, eta,
                    B0z, viscosity_order, quasilinear, remap, dt, nthreads, CFL)
        self._h = self._lib.orc_create(ctypes.byref(p))
        d = (ctypes.c_int * 6)()
        self._lib.orc_dims(self._h, d)
        self.nxy, self.nz, self.nMF, self.nLin, self.half, self.NZ = list(d)

    @classmethod
    def from_params(cls, p, nthreads=1, native=False, library=None):
        return cls(nthreads=nthreads, native=native, library=library, **p.oracle_dict())

    def close(self):
        if self._h:
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, lin=None, mf=None, kx0=None):
        def c(a, dt):
            return None if a is None else np.ascontiguousarray(a, dtype=dt)
        lin, mf, kx0 = c(lin, np.complex128), c(mf, np.complex128), c(kx0, np.float64)
        if lin is not None:
            assert lin.size == self.nxy * self.nLin * self.nz
        if mf is not None:
            assert mf.size == self.nMF * self.nz
        self._lib.orc_set_state(self._h, _dp(lin) if lin is not None else None,
                                _dp(mf) if mf is not None else None,
                                _dp(kx0) if kx0 is not None else None)

    def get_state(self):
        lin = np.zeros((self.nxy, self.nLin, self.nz), dtype=np.complex128)
        mf = np.zeros((self.nMF, self.nz), dtype=np.complex128)
        kx0 = np.zeros(self.nxy)
        self._lib.orc_get_state(self._h, _dp(lin), _dp(mf), _dp(kx0))
        return lin, mf, kx0

    def tables(self):
        kx0, ky = np.zeros(self.nxy), np.zeros(self.nxy)
        kxi, kyi = np.zeros(self.nxy, dtype=np.int32), np.zeros(self.nxy, dtype=np.int32)
        kz = np.zeros(self.nz)
        ip = ctypes.POINTER(ctypes.c_int)
        self._lib.orc_tables(self._h, _dp(kx0), _dp(ky), kxi.ctypes.data_as(ip),
                             kyi.ctypes.data_as(ip), _dp(kz))
        return kx0, ky, kxi, kyi, kz

    def linear_ops_init(self, t0=0.0):
        L = np.zeros((self.nxy, self.nLin, self.nz))
        Lmf = np.zeros((self.nMF, self.nz))
        self._lib.orc_linear_ops_init(self._h, t0, _dp(L), _dp(Lmf))
        return L, Lmf

    def rhs(self, t, dt_lin):
        out = np.zeros((self.nxy, self.nLin, self.nz), dtype=np.complex128)
        omf = np.zeros((self.nMF, self.nz), dtype=np.complex128)
        L = np.zeros((self.nxy, self.nLin, self.nz))
        self._lib.orc_rhs(self._h, t, dt_lin, _dp(out), _dp(omf), _dp(L))
        return out, omf, L

    def step(self, t):
        return self._lib.orc_step(self._h, t)

    def set_dt(self, dt):
        """The time step of the following steps: > 0 fixed, < 0 variable."""
        self._lib.orc_set_dt(self._h, dt)

    def set_noise(self, f_noise, range_low=0.0, range_high=1e16, drive_only_velocity=0, drive_only_magnetic=0,
                  seed=0x5EED0004):
        self._lib.orc_set_noise(self._h, f_noise, range_low, range_high, drive_only_velocity,
                                drive_only_magnetic, seed)

    def driving_noise(self, t, dt):
        self._lib.orc_driving_noise(self._h, t, dt)

    def noise_normals(self, t, i, jj):
        z = np.zeros(8)
        self._lib.orc_noise_normals(self._h, t, i, jj, _dp(z))
        return z

    def calculate_cfl(self):
        """Model::Calculate_CFL: CFL / dt of the next variable step."""
        return self._lib.orc_calculate_cfl(self._h)

    def check_sums(self):
        """Check_Solution's reductions: sum|Re|, sum|Im| of ifft(pad(By)),
        max|By|, max|Bx|, sum|By|."""
        out = np.zeros(5)
        self._lib.orc_check_sums(self._h, _dp(out))
        return out

    def check_solution(self):
        return self._lib.orc_check_solution(self._h)

    def steps(self, n, t=0.0):
        return self._lib.orc_steps(self._h, n, t)

    def remap(self, qt):
        self._lib.orc_remap(self._h, qt)

    def diagnostics(self, t):
        out = np.zeros(18)
        self._lib.orc_diagnostics(self._h, t, _dp(out))
        return out

    def drive(self):
        d = np.zeros((self.nMF, self.nz), dtype=np.complex128)
        self._lib.orc_get_drive(self._h, _dp(d))
        return d

    def mf_real(self):
        out = np.zeros((self.nMF, self.NZ))
        self._lib.orc_mf_real(self._h, _dp(out))
        return out
