"""ql_rhs_noalloc (the entry point that allocates nothing: its scratch is a
fixed per-device array) against ql_rhs, through the C ABI: two calls in a row,
the second on a smaller box, each equal to ql_rhs of the same state (RHS,
mean-field RHS, dissipation operator and drive)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _DeviceArray:
    """n doubles of device memory from the HIP runtime the library uses,
    filled with NaN (the test allocates; the entry point under test does not).
    The runtime is opened by its soname, which resolves to the copy already
    loaded with libql_dns.so (another copy would be a second runtime)."""

    def __init__(self, n):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.n = n
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(8 * n)) == 0
        nan = np.full(n, np.nan)
        assert self.hip.hipMemcpy(self.ptr, nan.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(8 * n), 1) == 0

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(8 * self.n), 2) == 0
        return out

    def free(self):
        assert self.hip.hipFree(self.ptr) == 0


@pytest.mark.parametrize("model", [0, 1])
def test_rhs_noalloc_equals_rhs(model):
    from ql_dns_amd import capi
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import QLParams
    L = capi.load()
    for k, (NX, NY, NZ) in enumerate([(16, 16, 128), (8, 8, 32)]):   # second call smaller
        p = QLParams(model=model, NX=NX, NY=NY, NZ=NZ, B0z=0.02)
        lin, mf = initial_state(p, seed=77 + k)
        c = QLContext(p)
        c.set_state(lin, mf)
        c.linear_ops_init(0.0)
        t, dtl = 0.31 + k, 0.0123
        r_lin, r_mf, r_L = c.rhs(t, dtl)
        drive = c.drive()
        nxy, nlin, nz = lin.shape
        dl, dm, dL = (_DeviceArray(n) for n in (2 * lin.size, 2 * mf.size, lin.size))
        rc = L.ql_rhs_noalloc(c._h, t, dtl, dl.ptr, dm.ptr, dL.ptr)
        assert rc == 0, L.ql_last_error()
        c.synchronize()
        g_lin = dl.get().view(np.complex128).reshape(lin.shape)
        g_mf = dm.get().view(np.complex128).reshape(mf.shape)
        g_L = dL.get().reshape(lin.shape)
        for x in (dl, dm, dL):
            x.free()
        # same kernels, same launch shape, same summation order: bit for bit
        assert np.array_equal(g_lin, r_lin)
        assert np.array_equal(g_mf, r_mf)
        assert np.array_equal(g_L, r_L)
        assert np.array_equal(c.drive(), drive)
        c.close()


def test_rhs_noalloc_refuses_oversize_and_hd():
    from ql_dns_amd import capi
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.params import QLParams
    L = capi.load()
    c = QLContext(QLParams(NX=64, NY=64, NZ=256))
    c.set_config(256, 256, 1 << 30)  # one workgroup per mode pass: beyond the scratch
    assert L.ql_rhs_noalloc(c._h, 0.0, 0.0, None, None, None) == -1
    assert b"scratch" in L.ql_last_error()
    c.close()
    h = QLContext(QLParams(model=2, NX=8, NY=8, NZ=32))
    assert L.ql_rhs_noalloc(h._h, 0.0, 0.0, None, None, None) == -3
    h.close()
