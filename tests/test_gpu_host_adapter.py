"""The C++ Model / Integrator / solution mirror (ql_dns_amd/csrc/host) driven
like the reference's main loop, against the oracle (GPU)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_err

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cpp", "test_host_adapter")


@pytest.mark.parametrize("model,integrator,dt,cfl", [
    ("MHD_FullQlin", "RK3CN_Elin", 0.01, 1.0), ("MHD_FullQlin", "RK3CN", 0.01, 1.0),
    ("MHD_BQlin", "RK3CN_Elin", 0.01, 1.0),
    ("MHD_FullQlin", "RK3CN_Elin", -0.05, 0.5), ("MHD_FullQlin", "RK3CN", -0.05, 0.5),
    ("MHD_BQlin", "RK3CN_Elin", -0.05, 0.5), ("HD_fullU", "RK3CN_Elin", 0.01, 1.0), ("HD_fullU", "RK3CN", -0.05, 0.5)])
def test_main_loop_through_cpp_adapter(tmp_path, model, integrator, dt, cfl):
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import INTEGRATOR_NAMES, MODEL_NAMES, workload
    from oracle.oracle import Oracle
    p = workload("box16x16x32", model=MODEL_NAMES[model], integrator=INTEGRATOR_NAMES[integrator], dt=dt,
                 CFL=cfl)
    lin, mf = initial_state(p)
    src = tmp_path / "state.bin"
    out = tmp_path / "out.bin"
    with open(src, "wb") as f:
        f.write(np.ascontiguousarray(lin).tobytes())
        f.write(np.ascontiguousarray(mf).tobytes())
    nsteps = 50
    r = subprocess.run([BIN, model, integrator, str(nsteps), str(src), str(out), repr(dt), repr(cfl)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    nl, nm = p.nxy * p.nLin * p.nz, p.nMF * p.nz
    o_ = 0

    def take(n, cplx=True):
        nonlocal o_
        k = 2 * n if cplx else n
        a = raw[o_:o_ + k]
        o_ += k
        return a.view(np.complex128) if cplx else a
    rhs_lin, rhs_mf, linop = take(nl), take(nm), take(nl, False)
    fl, fm, kx, t, diag = take(nl), take(nm), take(p.nxy, False), take(1, False)[0], take(18, False)
    nrey = int(take(1, False)[0])
    rey = take(nrey, False)
    ktab = take(3 * p.nxy, False).reshape(p.nxy, 3)
    kzt = take(2 * p.nz, False).reshape(p.nz, 2)
    laps = take(2 * (p.NY // 2) * p.nz, False).reshape(p.NY // 2, 2, p.nz)
    assert o_ == raw.size

    o = Oracle.from_params(p, nthreads=8)
    o.set_state(lin, mf)
    o.linear_ops_init(0.0)
    ro = o.rhs(0.0, 0.0053)
    assert rel_err(rhs_lin, ro[0].ravel()) < 1e-10
    assert rel_err(rhs_mf, ro[1].ravel()) < 1e-10
    assert rel_err(linop, ro[2].ravel()) < 1e-10
    # Model::K: the reference's tables (Kdata.cpp:28-125) against the oracle's
    kx0_t, ky_t, kxi_t, kyi_t, kz_t = o.tables()
    assert np.array_equal(ktab[:, 0], ky_t) and np.array_equal(ktab[:, 1], kxi_t) and np.array_equal(ktab[:, 2], kyi_t)
    assert np.array_equal(kzt[:, 0], kz_t) and np.array_equal(kzt[:, 1], -kz_t * kz_t)
    for j in range(p.NY // 2):
        lap2 = -ky_t[j] ** 2 - kz_t ** 2
        ilap2 = 1 / np.where(lap2 == 0, 1, lap2)
        if ky_t[j] == 0:
            lap2[0], ilap2[0] = 0.0, 1.0
        assert np.array_equal(laps[j, 0], lap2) and np.array_equal(laps[j, 1], ilap2)
    # the K->kx write of the C++ program before its loop
    kx_w = kx0_t.copy()
    kx_w[-1] += 0.5 * 2 * np.pi / p.Lx
    o.set_state(kx0=kx_w)
    o.linear_ops_init(0.0)  # the integrator is constructed after the write
    to = o.steps(nsteps, 0.0)
    lo, mo, ko = o.get_state()
    if dt > 0:
        assert t == to
    else:
        # each step's dt comes from transformed profiles: equal to rounding
        assert abs(t - to) < 1e-12 * to and t < nsteps * -dt
    assert np.array_equal(kx, ko)
    assert rel_err(fl, lo.ravel()) < 1e-10
    assert rel_err(fm, mo.ravel()) < 1e-10
    do = o.diagnostics(to)
    if model == "HD_fullU":
        # TimeVariables of HD_fullU (HD_fullU.cpp:466-535): (MU, u) pairs, then
        # the Ux and Uy stress profiles fac1D Re(ifft(pad(drive)))
        from oracle.oracle import fft_inverse
        d = o.drive()
        prof = np.concatenate([fft_inverse(p.NZ, d[f]).real / p.NZ for f in range(2)])
        assert nrey == 2 * p.NZ and rel_err(rey, prof) < 1e-10
        hd = np.zeros(18)
        for k, src in enumerate((0, 2)):
            hd[k], hd[4 + k], hd[8 + k] = do[src], do[4 + src], do[8 + src]
        hd[12:16] = prof[:4]
        hd[16] = do[16]
        do = hd
    else:
        assert nrey == 4
    fin = np.isfinite(do)
    assert np.all(np.abs(diag[fin] - do[fin]) <= 1e-10 * np.abs(do[fin]) + 1e-300)
