"""ctypes binding of include/ql_io.h (libql_io.so): the reference's HDF5
full-solution dump / restart (FullSolution.h5) and the .dat time series.

Arrays use the layouts of ql_dns.h: lin [nxy][4][nz] complex128, mf [nMF][nz]
complex128, kx0 [nxy] float64.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .capi import ql_params, to_c_params
from .params import QLParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libql_io.so")

EXPORTS = ("qlio_layout_from_params", "qlio_h5_save", "qlio_h5_load_latest", "qlio_h5_count", "qlio_dat_init",
           "qlio_dat_append", "qlio_dat_append_values", "qlio_last_error")


class qlio_layout(ctypes.Structure):
    _fields_ = [("NX", ctypes.c_int), ("NY", ctypes.c_int), ("NZ", ctypes.c_int), ("nxy", ctypes.c_int),
                ("nz", ctypes.c_int), ("nlin", ctypes.c_int), ("nmf", ctypes.c_int), ("L", ctypes.c_double * 3)]


class QLIOError(RuntimeError):
    pass


_LIB = None


def load(path: str = LIB_PATH):
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with `python -m ql_dns_amd.build` (needs HDF5)")
    L = ctypes.CDLL(path)
    P, D, I, S = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), ctypes.c_char_p
    LP = ctypes.POINTER(qlio_layout)
    sig = {
        "qlio_layout_from_params": ([ctypes.POINTER(ql_params), LP], ctypes.c_int),
        "qlio_h5_save": ([S, ctypes.c_int, LP, ctypes.c_double, ctypes.c_double, D, D, D, ctypes.c_int], ctypes.c_int),
        "qlio_h5_load_latest": ([S, LP, ctypes.c_double, D, D, D, D], ctypes.c_int),
        "qlio_h5_count": ([S, I], ctypes.c_int),
        "qlio_dat_init": ([S], ctypes.c_int),
        "qlio_dat_append": ([S, ctypes.c_double, D, ctypes.c_int, D, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "qlio_dat_append_values": ([S, ctypes.c_double, D, D, D, D, ctypes.c_int, D, ctypes.c_int, ctypes.c_int],
                                   ctypes.c_int),
        "qlio_last_error": ([], ctypes.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = L
    return L


def _check(L, rc):
    if rc != 0:
        raise QLIOError(L.qlio_last_error().decode())


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def layout(p: QLParams) -> qlio_layout:
    L = load()
    out = qlio_layout()
    cp = to_c_params(p)
    _check(L, L.qlio_layout_from_params(ctypes.byref(cp), ctypes.byref(out)))
    return out


def h5_save(path: str, p: QLParams, t: float, lin, mf, kx0, create: bool = False, fp64: bool = False):
    """Append the time-t group to FullSolution-style file `path` (create: new file)."""
    L = load()
    lay = layout(p)
    lin = np.ascontiguousarray(lin, dtype=np.complex128)
    mf = np.ascontiguousarray(mf, dtype=np.complex128)
    kx0 = np.ascontiguousarray(kx0, dtype=np.float64)
    if lin.size != lay.nxy * 4 * lay.nz or mf.size != lay.nmf * lay.nz or kx0.size != lay.nxy:
        raise ValueError("array sizes do not match the layout")
    _check(L, L.qlio_h5_save(path.encode(), int(create), ctypes.byref(lay), t, p.q, _dp(lin), _dp(mf), _dp(kx0),
                             int(fp64)))


def h5_load_latest(path: str, p: QLParams):
    """(t, lin, mf, kx0) of the latest save in `path`."""
    L = load()
    lay = layout(p)
    lin = np.zeros((lay.nxy, 4, lay.nz), dtype=np.complex128)
    mf = np.zeros((lay.nmf, lay.nz), dtype=np.complex128)
    kx0 = np.zeros(lay.nxy)
    t = ctypes.c_double(0.0)
    _check(L, L.qlio_h5_load_latest(path.encode(), ctypes.byref(lay), p.q, ctypes.byref(t), _dp(lin), _dp(mf),
                                    _dp(kx0)))
    return t.value, lin, mf, kx0


def h5_count(path: str) -> int:
    L = load()
    n = ctypes.c_int(0)
    _check(L, L.qlio_h5_count(path.encode(), ctypes.byref(n)))
    return n.value


def dat_init(directory: str):
    L = load()
    _check(L, L.qlio_dat_init(directory.encode()))


def dat_append(directory: str, t: float, diag, nrey: int = 4, mf_real=None):
    L = load()
    diag = np.ascontiguousarray(diag, dtype=np.float64)
    if diag.size != 18:
        raise ValueError("diag must hold the 18 ql_diagnostics values")
    nmf = NZ = 0
    if mf_real is not None:
        mf_real = np.ascontiguousarray(mf_real, dtype=np.float64)
        nmf, NZ = mf_real.shape
    _check(L, L.qlio_dat_append(directory.encode(), t, _dp(diag), nrey, _dp(mf_real), nmf, NZ))
