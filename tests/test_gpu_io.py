"""Dump and restart on the GPU through the reference-shaped file classes.

A run that saves FullSolution.h5 after n1 steps and continues n2 steps is
compared with a run restarted from that save (LoadSolution_forRestart, a new
integrator at t_start, RK3CN_Elin): bit for bit with the float64 dump, to
float rounding with the reference's float32 dump.  Both are anchored to the
oracle: the continuous run against the oracle's n1 + n2 steps from the same
initial state, the restarted run against the oracle's n2 steps from the state
read back from the file, each within 1e-10 (kx0 and t exactly).  The .dat time
series hold one record per save.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_err

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cpp", "test_io_restart")


def _read(path, nxy, nz, nmf):
    raw = np.fromfile(path, dtype=np.float64)
    per = nxy * 4 * nz * 2 + nmf * nz * 2 + nxy + 1
    assert raw.size == 2 * per
    out = []
    for k in range(2):
        r = raw[k * per:(k + 1) * per]
        o = 0
        lin = r[o:o + nxy * 4 * nz * 2].view(np.complex128).reshape(nxy, 4, nz)
        o += nxy * 4 * nz * 2
        mf = r[o:o + nmf * nz * 2].view(np.complex128).reshape(nmf, nz)
        o += nmf * nz * 2
        kx = r[o:o + nxy]
        t = r[o + nxy]
        out.append((lin, mf, kx, t))
    return out


def _anchor(p, lin, mf, n, got, t0=0.0, kx0=None):
    """The oracle's n steps of main.cpp's loop from (lin, mf, kx0) at t0 (its
    integrator constructed at t0, as a restart constructs it) against got."""
    from oracle.oracle import Oracle
    o = Oracle.from_params(p, nthreads=8)
    o.set_state(lin, mf, kx0)
    o.linear_ops_init(t0)
    t = o.steps(n, t0)
    lo, mo, ko = o.get_state()
    lg, mg, kg, tg = got
    assert tg == t and np.array_equal(kg, ko)
    assert rel_err(mg, mo) < 1e-10 and rel_err(lg, lo) < 1e-10
    o.close()


@pytest.mark.parametrize("fp64", [1, 0])
def test_cpp_dump_and_restart(tmp_path, fp64):
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import workload
    p = workload("box16x16x32")
    lin, mf = initial_state(p)
    state = str(tmp_path / "state.bin")
    with open(state, "wb") as f:
        f.write(np.ascontiguousarray(lin).tobytes())
        f.write(np.ascontiguousarray(mf).tobytes())
    out = str(tmp_path / "out.bin")
    d = str(tmp_path) + "/"
    n1, n2 = 37, 41
    r = subprocess.run([BIN, d, state, str(n1), str(n2), str(fp64), out], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    (la, ma, ka, ta), (lb, mb, kb, tb) = _read(out, p.nxy, p.nz, p.nMF)
    if fp64:
        assert tb == ta and np.array_equal(kb, ka)
        assert np.array_equal(lb, la) and np.array_equal(mb, ma)
    else:
        # the group name keeps t to 15 significant digits and kx is stored
        # as float32 (FullSave_Load.cpp:183-188, 231-240)
        assert abs(tb - ta) < 1e-12 and rel_err(kb, ka) < 1e-6
        assert rel_err(mb, ma) < 1e-5 and rel_err(lb, la) < 1e-5
    _anchor(p, lin, mf, n1 + n2, (la, ma, ka, ta))
    from ql_dns_amd import qlio
    t0, l0, m0, k0 = qlio.h5_load_latest(os.path.join(d, "FullSolution.h5"), p)
    _anchor(p, l0, m0, n2, (lb, mb, kb, tb), t0, k0)
    n = n1 + n2 + 1
    rd = lambda name: np.fromfile(os.path.join(d, name), dtype="<f8")
    assert rd("time_vec.dat").size == n
    assert rd("energy.dat").size == 4 * n and rd("reynolds_stress.dat").size == 4 * n
    assert rd("mean_fields.dat").size == n * p.nMF * p.NZ
    assert abs(rd("time_vec.dat")[-1] - ta) < 1e-12
    from ql_dns_amd import qlio
    assert qlio.h5_count(os.path.join(d, "FullSolution.h5")) == 1


def test_python_dump_and_restart(tmp_path):
    from ql_dns_amd import qlio
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import workload
    p = workload("box32x32x64")
    lin, mf = initial_state(p)
    path = str(tmp_path / "FullSolution.h5")
    a = QLContext(p)
    a.set_state(lin, mf)
    a.linear_ops_init(0.0)
    t = a.steps(30, 0.0)
    qlio.h5_save(path, p, t, *a.get_state(), create=True, fp64=True)
    ta = a.steps(30, t)
    b = QLContext(p)
    t0, l0, m0, k0 = qlio.h5_load_latest(path, p)
    assert t0 == t
    b.set_state(l0, m0, k0)
    b.linear_ops_init(t0)
    tb = b.steps(30, t0)
    assert tb == ta
    for x, y in zip(a.get_state(), b.get_state()):
        assert np.array_equal(x, y)
    lb, mb, kb = b.get_state()
    _anchor(p, lin, mf, 60, (lb, mb, kb, tb))
    _anchor(p, l0, m0, 30, (lb, mb, kb, tb), t0, k0)
