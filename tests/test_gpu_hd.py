"""HD_fullU on the GPU (the MHD_FullQlin kernels with the magnetic fields held
at zero) against the oracle's own restatement of HD_fullU (Models/HD_fullU.cpp).

Bar as for the MHD models: 1e-10 relative (norm-wise) on states, drives,
diagnostics and profiles; remapped kx0 and fixed-step times exactly equal.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _pair(**kw):
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import MODEL_HD_FULLU, QLParams
    from oracle.oracle import Oracle
    p = QLParams(model=MODEL_HD_FULLU, **kw)
    lin, mf = initial_state(p)
    o = Oracle.from_params(p, nthreads=8)
    o.set_state(lin, mf)
    o.linear_ops_init(0.0)
    g = QLContext(p)
    g.set_state(lin, mf)
    g.linear_ops_init(0.0)
    return p, o, g


def _compare(o, g, t):
    lo, mo, ko = o.get_state()
    lg, mg, kg = g.get_state()
    assert lg.shape == lo.shape and mg.shape == mo.shape == (2, o.nz)
    assert np.array_equal(kg, ko)
    assert rel_err(mg, mo) < TOL and rel_err(lg, lo) < TOL
    assert rel_err(g.drive(), o.drive()) < TOL
    do, dg = o.diagnostics(t), g.diagnostics(t)
    assert np.all(np.abs(dg - do) <= TOL * np.abs(do) + 1e-300)
    assert rel_err(g.mf_real(), o.mf_real()) < TOL


def test_dims_and_state_round_trip():
    p, o, g = _pair()
    assert (g.nLin, g.nMF) == (2, 2)
    lo, mo, ko = o.get_state()
    lg, mg, kg = g.get_state()
    assert np.array_equal(lg, lo) and np.array_equal(mg, mo) and np.array_equal(kg, ko)


@pytest.mark.parametrize("NX,NY,NZ,t", [(16, 16, 32, 0.0), (8, 8, 64, 0.21), (6, 4, 128, 0.4)])
def test_rhs_matches_oracle(NX, NY, NZ, t):
    p, o, g = _pair(NX=NX, NY=NY, NZ=NZ)
    ro, rg = o.rhs(t, 0.0123), g.rhs(t, 0.0123)
    for a, b in zip(rg, ro):
        assert a.shape == b.shape
        assert rel_err(a, b) < TOL
    assert rel_err(g.drive(), o.drive()) < TOL
    Lo, Lmo = o.linear_ops_init(0.3)
    Lg, Lmg = g.linear_ops_init(0.3)
    assert rel_err(Lg, Lo) < 1e-14 and rel_err(Lmg, Lmo) < 1e-14


@pytest.mark.parametrize("integrator,NX,NY,NZ", [(1, 16, 16, 32), (0, 16, 16, 32), (1, 8, 8, 64), (1, 4, 4, 256)])
def test_steps_match_oracle(integrator, NX, NY, NZ):
    p, o, g = _pair(integrator=integrator, NX=NX, NY=NY, NZ=NZ)
    to, tg = o.steps(40, 0.0), g.steps(40, 0.0)
    assert tg == to
    _compare(o, g, to)


@pytest.mark.parametrize("integrator,line", [(1, 256), (0, 256), (1, 128)])
def test_workgroup_shapes_match_oracle(integrator, line):
    """The velocity-only policy in the workgroup kernel at NZ = 128, where the
    default shape is the wavefront kernel."""
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import MODEL_HD_FULLU, QLParams
    from oracle.oracle import Oracle
    p = QLParams(model=MODEL_HD_FULLU, integrator=integrator, NX=6, NY=4, NZ=128)
    lin, mf = initial_state(p)
    o = Oracle.from_params(p, nthreads=8)
    g = QLContext(p)
    g.set_config(128, line, 5)
    assert g.config()[0]["line"] == line
    for x in (o, g):
        x.set_state(lin, mf)
        x.linear_ops_init(0.0)
    to, tg = o.steps(30, 0.0), g.steps(30, 0.0)
    assert tg == to
    _compare(o, g, to)


@pytest.mark.parametrize("line", [64, 256])
def test_mixed_path_close_to_fp64(line):
    """NZ = 256 in the mixed path: the velocity-only policy in FP32 transforms,
    in the wavefront kernel (the default shape there) and in the workgroup
    kernel; bar 1e-6 relative of the FP64 run."""
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import MODEL_HD_FULLU, PREC_MIXED, QLParams
    p = QLParams(model=MODEL_HD_FULLU, NX=8, NY=8, NZ=256, dt=0.005, nu=1 / 8000)
    lin, mf = initial_state(p)
    outs = []
    for q in (p, p.replace(precision=PREC_MIXED)):
        g = QLContext(q)
        if q.precision == PREC_MIXED:
            assert g.config()[0]["line"] == 64
            g.set_config(256, line, 256 if line == 64 else 512)
        g.set_state(lin, mf)
        g.linear_ops_init(0.0)
        t = g.steps(50, 0.0)
        outs.append((t, g.get_state(), g.drive()))
    (ta, sa, da), (tb, sb, db) = outs
    assert ta == tb and np.array_equal(sa[2], sb[2])
    assert rel_err(sb[1], sa[1]) < 1e-6 and rel_err(sb[0], sa[0]) < 1e-6 and rel_err(db, da) < 1e-6


def test_stress_profiles_match_oracle():
    from oracle.oracle import fft_inverse
    p, o, g = _pair()
    o.steps(5, 0.0)
    g.steps(5, 0.0)
    d = o.drive()
    want = np.array([fft_inverse(p.NZ, d[f]).real / p.NZ for f in range(2)])
    assert rel_err(g.drive_real(), want) < TOL


def test_variable_steps_match_oracle():
    p, o, g = _pair(dt=-0.05, CFL=0.6)
    o.rhs(0.0, 0.0)
    g.rhs(0.0, 0.0)
    to = tg = 0.0
    for n in range(6):
        vo, vg = o.calculate_cfl(), g.calculate_cfl()
        assert abs(vg - vo) <= 1e-12 * vo
        dto, dtg = o.step(to), g.step(tg)
        assert dtg == min(0.6 / vg, 0.05)
        to += dto
        tg += dtg
        o.remap(p.q * to)
        g.remap(p.q * tg)
    to, tg = o.steps(20, to), g.steps(20, tg)
    assert abs(tg - to) <= 1e-12 * to
    lo, mo, _ = o.get_state()
    lg, mg, _ = g.get_state()
    assert rel_err(mg, mo) < TOL and rel_err(lg, lo) < TOL


def test_noise_matches_oracle():
    p, o, g = _pair()
    for x in (o, g):
        x.set_noise(0.6, seed=12)
    to, tg = o.steps(10, 0.0), g.steps(10, 0.0)
    assert tg == to
    lo, mo, _ = o.get_state()
    lg, mg, _ = g.get_state()
    assert rel_err(mg, mo) < TOL and rel_err(lg, lo) < TOL


def test_check_solution():
    p, o, g = _pair()
    v, s = g.check_solution()
    so = o.check_sums()
    assert v == 0
    for k in (0, 2, 3, 4):
        assert abs(s[k] - so[k]) <= 1e-12 * so[k]
    z = np.zeros((2, p.nz), complex)
    z[1, 1] = 1.0
    g.set_state(None, z)
    assert g.check_solution()[0] == 1


def test_hd_run_set_matches_single_contexts():
    from ql_dns_amd.capi import QLBatch, QLContext
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import MODEL_HD_FULLU, QLParams
    ps = [QLParams(model=MODEL_HD_FULLU), QLParams(model=MODEL_HD_FULLU, NX=8, NY=8, NZ=64, q=1.2)]
    b = QLBatch(ps)
    st = [initial_state(p) for p in ps]
    for r, (lin, mf) in enumerate(st):
        b.set_state(r, lin, mf)
    b.linear_ops_init(0.0)
    tb = b.steps(20, 0.0)
    for r, (p, (lin, mf)) in enumerate(zip(ps, st)):
        g = QLContext(p)
        g.set_state(lin, mf)
        g.linear_ops_init(0.0)
        assert g.steps(20, 0.0) == tb
        (lg, mg, kg), (lb, mb, kb) = g.get_state(), b.get_state(r)
        assert np.array_equal(kg, kb)
        assert rel_err(lb, lg) < TOL and rel_err(mb, mg) < TOL
