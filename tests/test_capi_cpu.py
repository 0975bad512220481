"""The C-ABI library without a GPU: it loads, exports every symbol the header
declares, and rejects invalid parameters before touching a device."""
import os
import re

import pytest

from conftest import ROOT


def _declared():
    with open(os.path.join(ROOT, "include", "ql_dns.h")) as f:
        src = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|ql_batch \*)\s*(ql_\w+)\s*\(", src, re.M)))


def test_header_matches_binding():
    from ql_dns_amd.capi import EXPORTS
    assert sorted(EXPORTS) == _declared()


def test_library_exports_all_symbols():
    from ql_dns_amd import capi
    L = capi.load()
    for name in _declared():
        assert hasattr(L, name), name


@pytest.mark.parametrize("over,code", [
    (dict(NZ=48), -1), (dict(NZ=512), -1), (dict(NX=7), -1), (dict(NY=0), -1), (dict(model=5), -1),
    (dict(integrator=3), -1), (dict(precision=7), -3), (dict(dt=0.0), -1), (dict(dt=float("nan")), -1),
    (dict(dt=-0.01, CFL=0.0), -1), (dict(dt=-0.01, CFL=float("inf")), -1),
    (dict(viscosity_order=-2), -1)])
def test_create_rejects_invalid_params(over, code):
    import ctypes
    from ql_dns_amd import capi
    from ql_dns_amd.params import QLParams
    L = capi.load()
    p = QLParams().replace(**over)
    cp = capi.to_c_params(p)
    h = ctypes.c_void_p()
    rc = L.ql_create(ctypes.byref(cp), ctypes.byref(h))
    assert rc == code
    assert h.value is None
    assert len(L.ql_last_error()) > 0


def test_null_arguments_are_errors():
    import ctypes
    from ql_dns_amd import capi
    L = capi.load()
    assert L.ql_create(None, None) == -1
    assert L.ql_destroy(None) == 0
    assert L.ql_steps(None, 1, None) == -1
    assert L.ql_dims(None, None) == -1


def _create_set(params):
    import ctypes
    from ql_dns_amd import capi
    L = capi.load()
    arr = (capi.ql_params * len(params))(*[capi.to_c_params(p) for p in params])
    h = ctypes.c_void_p()
    rc = L.ql_batch_create(arr, len(params), ctypes.byref(h))
    return rc, h, L


@pytest.mark.parametrize("second,code", [
    (dict(dt=0.02), -1),                 # runs of a set share dt
    (dict(model=1), -1),                 # ... and model
    (dict(integrator=0), -1),            # ... and integrator
    (dict(precision=1), -1),             # ... and precision
    (dict(Lz=2.0), -1),                  # runs of one NZ share Lz
    (dict(NZ=96), -1),                   # each run is checked like ql_create
    (dict(dt=-1.0), -1)])
def test_set_create_rejects_inconsistent_runs(second, code):
    from ql_dns_amd.params import QLParams
    p = QLParams()
    rc, h, L = _create_set([p, p.replace(**second)])
    assert rc == code
    assert h.value is None
    assert len(L.ql_last_error()) > 0


def test_variable_step_sets_pass_validation():
    """Variable steps in a set of several runs are accepted (each run keeps its
    own clock); without a device creation stops at the device check."""
    from ql_dns_amd.params import QLParams
    p = QLParams(dt=-0.01)
    rc, h, L = _create_set([p, p.replace(NZ=64)])
    assert rc == -2 and b"device" in L.ql_last_error()
    assert L.ql_batch_steps_runs(None, 1, None) == -1


def test_set_null_arguments_are_errors():
    import ctypes
    from ql_dns_amd import capi
    L = capi.load()
    h = ctypes.c_void_p()
    assert L.ql_batch_create(None, 1, ctypes.byref(h)) == -1
    rc, h, L = _create_set([])
    assert rc == -1
    assert L.ql_batch_destroy(None) == 0
    assert L.ql_batch_of(None) is None
    assert L.ql_batch_steps(None, 1, None) == -1
    assert L.ql_batch_tune(None, 1, None) == -1
    n = ctypes.c_int(0)
    assert L.ql_batch_size(None, ctypes.byref(n)) == -1
    assert L.ql_adjoint_forward(None, 1, 1, None) == -1
    assert L.ql_adjoint_backward(None, None, None, None, None, None) == -1
    v = ctypes.c_double(0.0)
    assert L.ql_calculate_cfl(None, ctypes.byref(v)) == -1
    assert L.ql_check_solution(None, None, None) == -1
    assert L.ql_batch_calculate_cfl(None, 0, ctypes.byref(v)) == -1
    assert L.ql_batch_check_solution(None, 0, None, None) == -1
