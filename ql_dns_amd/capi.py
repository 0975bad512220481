"""ctypes binding of the C ABI in include/ql_dns.h (libql_dns.so).

`QLContext` is a thin object over one ``ql_ctx`` and `QLBatch` over one
``ql_batch`` (a set of runs stepped together): every method is one C-ABI
call.  Loading fails loudly when the library is missing; there is no fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import QLParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libql_dns.so")

# every symbol include/ql_dns.h declares
EXPORTS = ("ql_create", "ql_destroy", "ql_dims", "ql_set_state", "ql_get_state",
           "ql_linear_ops_init", "ql_rhs", "ql_step", "ql_steps", "ql_remap", "ql_diagnostics",
           "ql_get_drive", "ql_mf_real", "ql_synchronize", "ql_profile_kernels", "ql_last_error",
           "ql_batch_create", "ql_batch_destroy", "ql_batch_of", "ql_batch_size", "ql_batch_dims",
           "ql_batch_set_state", "ql_batch_get_state", "ql_batch_diagnostics", "ql_batch_get_drive",
           "ql_batch_linear_ops_init", "ql_batch_steps", "ql_batch_tune", "ql_batch_get_config",
           "ql_batch_set_config", "ql_adjoint_forward", "ql_adjoint_backward", "ql_calculate_cfl",
           "ql_check_solution", "ql_batch_calculate_cfl", "ql_batch_check_solution", "ql_set_noise",
           "ql_driving_noise", "ql_batch_set_noise", "ql_batch_driving_noise", "ql_drive_real",
           "ql_rhs_noalloc", "ql_batch_mixed_split", "ql_batch_set_transform", "ql_batch_steps_runs",
           "ql_batch_set_tail")

QL_ERR_STATE = -4


class QLError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ql error {code}: {msg}")
        self.code = code


class ql_params(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int), ("integrator", ctypes.c_int), ("precision", ctypes.c_int),
                ("NX", ctypes.c_int), ("NY", ctypes.c_int), ("NZ", ctypes.c_int),
                ("L", ctypes.c_double * 3), ("q", ctypes.c_double), ("omega", ctypes.c_double),
                ("nu", ctypes.c_double), ("eta", ctypes.c_double), ("B0z", ctypes.c_double),
                ("viscosity_order", ctypes.c_int), ("quasilinear", ctypes.c_int),
                ("remap", ctypes.c_int), ("dt", ctypes.c_double), ("device", ctypes.c_int),
                ("CFL", ctypes.c_double)]


class ql_noise(ctypes.Structure):
    _fields_ = [("f_noise", ctypes.c_double), ("range_low", ctypes.c_double), ("range_high", ctypes.c_double),
                ("drive_only_velocity", ctypes.c_int), ("drive_only_magnetic", ctypes.c_int),
                ("seed", ctypes.c_ulonglong)]


def noise_config(f_noise: float, range_low: float = 0.0, range_high: float = 1e16, drive_only_velocity: int = 0,
                 drive_only_magnetic: int = 0, seed: int = 0x5EED0004) -> ql_noise:
    """The reference's noise inputs (Input_parameters.cpp:107-111), defaults included."""
    return ql_noise(f_noise, range_low, range_high, drive_only_velocity, drive_only_magnetic, seed)


_LIB = None


def load(path: str = LIB_PATH):
    """Load libql_dns.so and declare its signatures (raises if absent)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `python -m ql_dns_amd.build` (hipcc, gfx950)")
    L = ctypes.CDLL(path)
    P, D, I = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    sig = {
        "ql_create": ([ctypes.POINTER(ql_params), ctypes.POINTER(P)], ctypes.c_int),
        "ql_destroy": ([P], ctypes.c_int),
        "ql_dims": ([P, I], ctypes.c_int),
        "ql_set_state": ([P, D, D, D], ctypes.c_int),
        "ql_get_state": ([P, D, D, D], ctypes.c_int),
        "ql_linear_ops_init": ([P, ctypes.c_double, D, D], ctypes.c_int),
        "ql_rhs": ([P, ctypes.c_double, ctypes.c_double, D, D, D], ctypes.c_int),
        "ql_rhs_noalloc": ([P, ctypes.c_double, ctypes.c_double, P, P, P], ctypes.c_int),
        "ql_step": ([P, ctypes.c_double, D], ctypes.c_int),
        "ql_steps": ([P, ctypes.c_int, D], ctypes.c_int),
        "ql_remap": ([P, ctypes.c_double], ctypes.c_int),
        "ql_diagnostics": ([P, ctypes.c_double, D], ctypes.c_int),
        "ql_get_drive": ([P, D], ctypes.c_int),
        "ql_mf_real": ([P, D], ctypes.c_int),
        "ql_synchronize": ([P], ctypes.c_int),
        "ql_profile_kernels": ([P, ctypes.c_int, D, D, I], ctypes.c_int),
        "ql_last_error": ([], ctypes.c_char_p),
        "ql_batch_create": ([ctypes.POINTER(ql_params), ctypes.c_int, ctypes.POINTER(P)], ctypes.c_int),
        "ql_batch_destroy": ([P], ctypes.c_int),
        "ql_batch_of": ([P], P),
        "ql_batch_size": ([P, I], ctypes.c_int),
        "ql_batch_dims": ([P, ctypes.c_int, I], ctypes.c_int),
        "ql_batch_set_state": ([P, ctypes.c_int, D, D, D], ctypes.c_int),
        "ql_batch_get_state": ([P, ctypes.c_int, D, D, D], ctypes.c_int),
        "ql_batch_diagnostics": ([P, ctypes.c_int, ctypes.c_double, D], ctypes.c_int),
        "ql_batch_get_drive": ([P, ctypes.c_int, D], ctypes.c_int),
        "ql_batch_linear_ops_init": ([P, ctypes.c_double], ctypes.c_int),
        "ql_batch_steps": ([P, ctypes.c_int, D], ctypes.c_int),
        "ql_batch_steps_runs": ([P, ctypes.c_int, D], ctypes.c_int),
        "ql_batch_tune": ([P, ctypes.c_int, D], ctypes.c_int),
        "ql_batch_get_config": ([P, ctypes.c_int, I, I], ctypes.c_int),
        "ql_batch_set_config": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "ql_batch_mixed_split": ([P, ctypes.c_double, I], ctypes.c_int),
        "ql_batch_set_transform": ([P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "ql_batch_set_tail": ([P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "ql_adjoint_forward": ([P, ctypes.c_int, ctypes.c_int, D], ctypes.c_int),
        "ql_adjoint_backward": ([P, D, D, D, D, D], ctypes.c_int),
        "ql_calculate_cfl": ([P, D], ctypes.c_int),
        "ql_drive_real": ([P, D], ctypes.c_int),
        "ql_check_solution": ([P, D, I], ctypes.c_int),
        "ql_batch_calculate_cfl": ([P, ctypes.c_int, D], ctypes.c_int),
        "ql_batch_check_solution": ([P, ctypes.c_int, D, I], ctypes.c_int),
        "ql_set_noise": ([P, ctypes.POINTER(ql_noise)], ctypes.c_int),
        "ql_driving_noise": ([P, ctypes.c_double, ctypes.c_double], ctypes.c_int),
        "ql_batch_set_noise": ([P, ctypes.c_int, ctypes.POINTER(ql_noise)], ctypes.c_int),
        "ql_batch_driving_noise": ([P, ctypes.c_double, ctypes.c_double], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = L
    return L


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _checked(a, dt, n):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=dt)
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _raise(L, rc: int):
    if rc != 0:
        raise QLError(rc, L.ql_last_error().decode())


def _check_solution(L, fn, args, strict: bool):
    """Check_Solution: (verdict, sums); a failed check raises when strict
    (the reference aborts, solution.cpp:169-203)."""
    sums = np.zeros(5)
    v = ctypes.c_int(0)
    rc = fn(*args, _dp(sums), ctypes.byref(v))
    if rc != 0 and (strict or rc != QL_ERR_STATE):
        _raise(L, rc)
    return v.value, sums


def _get_config(L, h):
    n = ctypes.c_int(0)
    _raise(L, L.ql_batch_get_config(h, 0, ctypes.byref(n), None))
    keys = ("NZ", "line", "target", "modes_per_pass", "workgroups", "transform", "tail")
    cfg = (ctypes.c_int * (len(keys) * max(1, n.value)))()
    _raise(L, L.ql_batch_get_config(h, n.value, ctypes.byref(n), cfg))
    return [dict(zip(keys, cfg[len(keys) * k:len(keys) * (k + 1)])) for k in range(n.value)]


def _tune(L, h, trial_steps: int) -> float:
    ms = ctypes.c_double(0.0)
    _raise(L, L.ql_batch_tune(h, trial_steps, ctypes.byref(ms)))
    return ms.value


def to_c_params(p: QLParams, device: int = 0) -> ql_params:
    return ql_params(p.model, p.integrator, p.precision, p.NX, p.NY, p.NZ,
                     (ctypes.c_double * 3)(p.Lx, p.Ly, p.Lz), p.q, p.omega, p.nu, p.eta, p.B0z,
                     p.viscosity_order, p.quasilinear, p.remap, p.dt, device, p.CFL)


class QLContext:
    """One run on one GPU, driven through the C ABI."""

    def __init__(self, p: QLParams, device: int = 0):
        self._L = load()
        self.params = p
        self._h = ctypes.c_void_p()
        cp = to_c_params(p, device)
        self._check(self._L.ql_create(ctypes.byref(cp), ctypes.byref(self._h)))
        d = (ctypes.c_int * 6)()
        self._check(self._L.ql_dims(self._h, d))
        self.nxy, self.nz, self.nMF, self.nLin, self.half, self.NZ = list(d)

    def _check(self, rc: int):
        if rc != 0:
            raise QLError(rc, self._L.ql_last_error().decode())

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.ql_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_state(self, lin=None, mf=None, kx0=None):
        lin = _checked(lin, np.complex128, self.nxy * self.nLin * self.nz)
        mf = _checked(mf, np.complex128, self.nMF * self.nz)
        kx0 = _checked(kx0, np.float64, self.nxy)
        self._check(self._L.ql_set_state(self._h, _dp(lin), _dp(mf), _dp(kx0)))

    def get_state(self):
        lin = np.zeros((self.nxy, self.nLin, self.nz), dtype=np.complex128)
        mf = np.zeros((self.nMF, self.nz), dtype=np.complex128)
        kx0 = np.zeros(self.nxy)
        self._check(self._L.ql_get_state(self._h, _dp(lin), _dp(mf), _dp(kx0)))
        return lin, mf, kx0

    def linear_ops_init(self, t0: float = 0.0):
        L = np.zeros((self.nxy, self.nLin, self.nz))
        Lmf = np.zeros((self.nMF, self.nz))
        self._check(self._L.ql_linear_ops_init(self._h, t0, _dp(L), _dp(Lmf)))
        return L, Lmf

    def rhs(self, t: float, dt_lin: float):
        out = np.zeros((self.nxy, self.nLin, self.nz), dtype=np.complex128)
        omf = np.zeros((self.nMF, self.nz), dtype=np.complex128)
        L = np.zeros((self.nxy, self.nLin, self.nz))
        self._check(self._L.ql_rhs(self._h, t, dt_lin, _dp(out), _dp(omf), _dp(L)))
        return out, omf, L

    def step(self, t: float) -> float:
        dt = ctypes.c_double(0.0)
        self._check(self._L.ql_step(self._h, t, ctypes.byref(dt)))
        return dt.value

    def steps(self, n: int, t: float = 0.0) -> float:
        tt = ctypes.c_double(t)
        self._check(self._L.ql_steps(self._h, n, ctypes.byref(tt)))
        return tt.value

    def remap(self, qt: float):
        self._check(self._L.ql_remap(self._h, qt))

    def diagnostics(self, t: float):
        out = np.zeros(18)
        self._check(self._L.ql_diagnostics(self._h, t, _dp(out)))
        return out

    def drive(self):
        d = np.zeros((self.nMF, self.nz), dtype=np.complex128)
        self._check(self._L.ql_get_drive(self._h, _dp(d)))
        return d

    def mf_real(self):
        out = np.zeros((self.nMF, self.NZ))
        self._check(self._L.ql_mf_real(self._h, _dp(out)))
        return out

    def drive_real(self):
        """fac1D Re(ifft(pad(drive))) of the last rhs: [nMF][NZ]."""
        out = np.zeros((self.nMF, self.NZ))
        self._check(self._L.ql_drive_real(self._h, _dp(out)))
        return out

    def calculate_cfl(self) -> float:
        """Model::Calculate_CFL: CFL / dt of the next variable step."""
        v = ctypes.c_double(0.0)
        self._check(self._L.ql_calculate_cfl(self._h, ctypes.byref(v)))
        return v.value

    def check_solution(self, strict: bool = False):
        return _check_solution(self._L, self._L.ql_check_solution, (self._h,), strict)

    def set_noise(self, f_noise: float, **kw):
        """Model::ChangeNoiseRange plus the noise inputs; see noise_config."""
        cfg = noise_config(f_noise, **kw)
        self._check(self._L.ql_set_noise(self._h, ctypes.byref(cfg)))

    def driving_noise(self, t: float, dt: float):
        self._check(self._L.ql_driving_noise(self._h, t, dt))

    def synchronize(self):
        self._check(self._L.ql_synchronize(self._h))

    def profile_kernels(self, n: int, t: float = 0.0):
        tt = ctypes.c_double(t)
        avg = np.zeros(2)
        cnt = (ctypes.c_int * 2)()
        self._check(self._L.ql_profile_kernels(self._h, n, ctypes.byref(tt), _dp(avg), cnt))
        return tt.value, avg, list(cnt)

    # adjoint of `steps` (checkpointed)
    def adjoint_forward(self, n: int, every: int, t: float = 0.0) -> float:
        tt = ctypes.c_double(t)
        self._check(self._L.ql_adjoint_forward(self._h, n, every, ctypes.byref(tt)))
        return tt.value

    def adjoint_backward(self, g_lin=None, g_mf=None):
        """dL/d(initial lin, mf) and dL/d(q, omega, nu, eta, B0z) for the given
        dL/d(final lin, mf) (complex arrays; None = zero)."""
        g_lin = _checked(g_lin, np.complex128, self.nxy * self.nLin * self.nz)
        g_mf = _checked(g_mf, np.complex128, self.nMF * self.nz)
        gl0 = np.zeros((self.nxy, self.nLin, self.nz), dtype=np.complex128)
        gm0 = np.zeros((self.nMF, self.nz), dtype=np.complex128)
        gp = np.zeros(5)
        self._check(self._L.ql_adjoint_backward(self._h, _dp(g_lin), _dp(g_mf), _dp(gl0), _dp(gm0), _dp(gp)))
        return gl0, gm0, gp

    # launch shape (the context's run set of one)
    def tune(self, trial_steps: int = 5) -> float:
        return _tune(self._L, self._L.ql_batch_of(self._h), trial_steps)

    def config(self):
        return _get_config(self._L, self._L.ql_batch_of(self._h))

    def set_config(self, NZ: int, line: int, target: int):
        self._check(self._L.ql_batch_set_config(self._L.ql_batch_of(self._h), NZ, line, target))

    def set_transform(self, NZ: int, transform: int):
        """QL_TRANSFORM_FFT (0) or QL_TRANSFORM_DFT_MFMA (1) for the NZ group."""
        self._check(self._L.ql_batch_set_transform(self._L.ql_batch_of(self._h), NZ, transform))

    def set_tail(self, NZ: int, tail: int):
        """QL_TAIL_TWO_KERNELS (0) or QL_TAIL_FUSED (1) for the NZ group."""
        self._check(self._L.ql_batch_set_tail(self._L.ql_batch_of(self._h), NZ, tail))

    def mixed_split(self, tau: float = 0.0) -> int:
        """Mixed precision: sets the FP64 amplitude fraction when tau > 0 and
        returns how many modes currently reach it (ql_batch_mixed_split)."""
        n = ctypes.c_int()
        self._check(self._L.ql_batch_mixed_split(self._L.ql_batch_of(self._h), tau, ctypes.byref(n)))
        return n.value


class QLBatch:
    """A set of runs (one model, integrator, precision and dt) stepped together
    on one GPU.  Run r has the sizes of ``params[r]``."""

    def __init__(self, params, device: int = 0):
        self._L = load()
        self.params = list(params)
        n = len(self.params)
        arr = (ql_params * n)(*[to_c_params(p, device) for p in self.params])
        self._h = ctypes.c_void_p()
        self._check(self._L.ql_batch_create(arr, n, ctypes.byref(self._h)))
        self.dims = []
        for r in range(n):
            d = (ctypes.c_int * 6)()
            self._check(self._L.ql_batch_dims(self._h, r, d))
            self.dims.append(dict(zip(("nxy", "nz", "nMF", "nLin", "half", "NZ"), list(d))))

    def __len__(self):
        return len(self.params)

    def _check(self, rc: int):
        _raise(self._L, rc)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.ql_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_state(self, run: int, lin=None, mf=None, kx0=None):
        d = self.dims[run]
        lin = _checked(lin, np.complex128, d["nxy"] * d["nLin"] * d["nz"])
        mf = _checked(mf, np.complex128, d["nMF"] * d["nz"])
        kx0 = _checked(kx0, np.float64, d["nxy"])
        self._check(self._L.ql_batch_set_state(self._h, run, _dp(lin), _dp(mf), _dp(kx0)))

    def get_state(self, run: int):
        d = self.dims[run]
        lin = np.zeros((d["nxy"], d["nLin"], d["nz"]), dtype=np.complex128)
        mf = np.zeros((d["nMF"], d["nz"]), dtype=np.complex128)
        kx0 = np.zeros(d["nxy"])
        self._check(self._L.ql_batch_get_state(self._h, run, _dp(lin), _dp(mf), _dp(kx0)))
        return lin, mf, kx0

    def linear_ops_init(self, t0: float = 0.0):
        self._check(self._L.ql_batch_linear_ops_init(self._h, t0))

    def steps_runs(self, n: int, t_runs) -> np.ndarray:
        """n steps with each run's own time (ql_batch_steps_runs); returns the
        runs' times after them."""
        tt = np.ascontiguousarray(t_runs, dtype=np.float64).copy()
        self._check(self._L.ql_batch_steps_runs(self._h, n, _dp(tt)))
        return tt

    def steps(self, n: int, t: float = 0.0) -> float:
        tt = ctypes.c_double(t)
        self._check(self._L.ql_batch_steps(self._h, n, ctypes.byref(tt)))
        return tt.value

    def diagnostics(self, run: int, t: float):
        out = np.zeros(18)
        self._check(self._L.ql_batch_diagnostics(self._h, run, t, _dp(out)))
        return out

    def drive(self, run: int):
        d = self.dims[run]
        out = np.zeros((d["nMF"], d["nz"]), dtype=np.complex128)
        self._check(self._L.ql_batch_get_drive(self._h, run, _dp(out)))
        return out

    def calculate_cfl(self, run: int) -> float:
        v = ctypes.c_double(0.0)
        self._check(self._L.ql_batch_calculate_cfl(self._h, run, ctypes.byref(v)))
        return v.value

    def check_solution(self, run: int, strict: bool = False):
        return _check_solution(self._L, self._L.ql_batch_check_solution, (self._h, run), strict)

    def set_noise(self, run: int, f_noise: float, **kw):
        cfg = noise_config(f_noise, **kw)
        self._check(self._L.ql_batch_set_noise(self._h, run, ctypes.byref(cfg)))

    def driving_noise(self, t: float, dt: float):
        self._check(self._L.ql_batch_driving_noise(self._h, t, dt))

    def tune(self, trial_steps: int = 5) -> float:
        return _tune(self._L, self._h, trial_steps)

    def config(self):
        return _get_config(self._L, self._h)

    def set_config(self, NZ: int, line: int, target: int):
        self._check(self._L.ql_batch_set_config(self._h, NZ, line, target))

    def set_transform(self, NZ: int, transform: int):
        self._check(self._L.ql_batch_set_transform(self._h, NZ, transform))

    def set_tail(self, NZ: int, tail: int):
        self._check(self._L.ql_batch_set_tail(self._h, NZ, tail))
