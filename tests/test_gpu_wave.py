"""The mode kernel with one wavefront per mode (launch line 64) against the
oracle and against the workgroup kernel.

Bars as for the other shapes: states, drives and diagnostics within 1e-10
relative of the oracle (kx0 after remaps and fixed-step times exactly equal);
the same run under line 64 and under (256, 512), which sum the stress
partials in different orders, within 1e-12; one shape bitwise reproducible;
the mixed path within 1e-6 of FP64.  HD_fullU runs its velocity-only policy
(ModelHD) there, MHD_BQlin its own (ModelBQ).  In mixed precision the
wavefront kernel runs the flagged modes in FP64 itself (k_wave_mix).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10
WAVE = 64


def _params(**kw):
    from ql_dns_amd.params import QLParams
    return QLParams(**kw)


def _pair(p, target=512, seed=None):
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from oracle.oracle import Oracle
    lin, mf = initial_state(p, seed=seed)
    o = Oracle.from_params(p, nthreads=8)
    o.set_state(lin, mf)
    o.linear_ops_init(0.0)
    g = QLContext(p)
    g.set_config(p.NZ, WAVE, target)
    g.set_state(lin, mf)
    g.linear_ops_init(0.0)
    return o, g


@pytest.mark.parametrize("integrator", [1, 0])
@pytest.mark.parametrize("NX,NY,NZ,target", [(16, 16, 32, 512), (8, 8, 64, 7), (6, 6, 128, 3), (4, 4, 128, 1 << 30),
                                               (6, 4, 256, 5)])
def test_steps_match_oracle(integrator, NX, NY, NZ, target):
    p = _params(integrator=integrator, NX=NX, NY=NY, NZ=NZ)
    o, g = _pair(p, target)
    assert g.config()[0]["line"] == WAVE and g.config()[0]["modes_per_pass"] == 1
    to, tg = o.steps(40, 0.0), g.steps(40, 0.0)
    assert tg == to
    lo, mo, ko = o.get_state()
    lg, mg, kg = g.get_state()
    assert np.array_equal(kg, ko)
    assert rel_err(mg, mo) < TOL and rel_err(lg, lo) < TOL
    assert rel_err(g.drive(), o.drive()) < TOL
    do, dg = o.diagnostics(to), g.diagnostics(tg)
    assert np.all(np.abs(dg - do) <= TOL * np.abs(do) + 1e-300)


@pytest.mark.parametrize("NZ", [32, 64, 128])
def test_rhs_matches_oracle(NZ):
    p = _params(NX=8, NY=8, NZ=NZ, B0z=0.03)
    o, g = _pair(p)
    ro, rg = o.rhs(0.37, 0.011), g.rhs(0.37, 0.011)
    for a, b in zip(rg, ro):
        assert a.shape == b.shape
        assert rel_err(a, b) < TOL
    assert rel_err(g.drive(), o.drive()) < TOL


def test_hd_and_variable_steps_match_oracle():
    from ql_dns_amd.params import MODEL_HD_FULLU
    for p in (_params(model=MODEL_HD_FULLU), _params(model=MODEL_HD_FULLU, integrator=0, NX=8, NY=8, NZ=64),
              _params(model=MODEL_HD_FULLU, NX=6, NY=4, NZ=128, dt=-0.05, CFL=0.6), _params(dt=-0.05, CFL=0.5)):
        o, g = _pair(p)
        o.rhs(0.0, 0.0)
        g.rhs(0.0, 0.0)
        to, tg = o.steps(25, 0.0), g.steps(25, 0.0)
        assert abs(tg - to) <= 1e-12 * to
        lo, mo, _ = o.get_state()
        lg, mg, _ = g.get_state()
        assert rel_err(mg, mo) < TOL and rel_err(lg, lo) < TOL


def test_wave_and_workgroup_shapes_agree():
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    p = _params(NX=32, NY=32, NZ=128)
    lin, mf = initial_state(p)
    res = {}
    for shape in [(256, 512), (WAVE, 512), (WAVE, 1024), (WAVE, 512)]:
        c = QLContext(p)
        c.set_config(128, *shape)
        c.set_state(lin, mf)
        c.linear_ops_init(0.0)
        c.steps(20, 0.0)
        out = c.get_state()
        if shape in res:
            for x, y in zip(res[shape], out):
                assert np.array_equal(x, y), f"shape {shape} not reproducible"
        res[shape] = out
    ref = res[(256, 512)]
    for shape, out in res.items():
        assert rel_err(out[1], ref[1]) < 1e-12, shape
        assert rel_err(out[0], ref[0]) < 1e-12, shape
        assert np.array_equal(out[2], ref[2]), shape


def test_mixed_path_close_to_fp64():
    from ql_dns_amd.params import PREC_MIXED
    p = _params(NX=32, NY=32, NZ=128, dt=0.005, nu=1 / 8000, eta=1 / 8000)
    outs = []
    for q in (p, p.replace(precision=PREC_MIXED)):
        _, g = _pair(q)
        t = g.steps(50, 0.0)
        outs.append((t, g.get_state(), g.drive()))
    (ta, sa, da), (tb, sb, db) = outs
    assert ta == tb and np.array_equal(sa[2], sb[2])
    assert rel_err(sb[1], sa[1]) < 1e-6 and rel_err(db, da) < 1e-6


def test_run_set_matches_oracle_per_run():
    from ql_dns_amd.capi import QLBatch
    from ql_dns_amd.ic import initial_state
    from oracle.oracle import Oracle
    runs = [_params(NX=8, NY=8, NZ=32, q=1.5), _params(NX=6, NY=4, NZ=32, q=1.1, Lx=2.0),
            _params(NX=8, NY=6, NZ=64, q=1.7, B0z=0.05), _params(NX=4, NY=4, NZ=128, viscosity_order=4)]
    st = [initial_state(p, seed=70 + r) for r, p in enumerate(runs)]
    b = QLBatch(runs)
    for NZ in (32, 64, 128):
        b.set_config(NZ, WAVE, 5)
    for r, (lin, mf) in enumerate(st):
        b.set_state(r, lin, mf)
    b.linear_ops_init(0.0)
    tb = b.steps(30, 0.0)
    for r, (p, (lin, mf)) in enumerate(zip(runs, st)):
        o = Oracle.from_params(p, nthreads=8)
        o.set_state(lin, mf)
        o.linear_ops_init(0.0)
        assert o.steps(30, 0.0) == tb
        (lo, mo, ko), (lg, mg, kg) = o.get_state(), b.get_state(r)
        assert np.array_equal(kg, ko)
        assert rel_err(mg, mo) < TOL and rel_err(lg, lo) < TOL


@pytest.mark.parametrize("model", [0, 1])
def test_mixed_at_256_takes_the_wave_shape(model):
    """Mixed precision at NZ = 256 defaults to the wavefront kernel (flagged
    modes in FP64 in the same launch, k_wave_mix); it agrees with the
    workgroup shape (companion + FP32 launch) and with FP64 to the FP32
    bar, and the workgroup shape stays available."""
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import PREC_MIXED
    p = _params(model=model, NX=8, NY=8, NZ=256, precision=PREC_MIXED)
    lin, mf = initial_state(p)
    out = []
    for line in (None, 256):
        c = QLContext(p)
        if line is None:
            assert c.config()[0]["line"] == WAVE
        else:
            c.set_config(256, line, 512)
        c.set_state(lin, mf)
        c.linear_ops_init(0.0)
        c.steps(20, 0.0)
        out.append(c.get_state())
        c.close()
    f = QLContext(p.replace(precision=0))
    f.set_state(lin, mf)
    f.linear_ops_init(0.0)
    f.steps(20, 0.0)
    ref = f.get_state()
    f.close()
    for o in out:
        assert np.array_equal(o[2], ref[2])
        assert rel_err(o[1], ref[1]) < 1e-6 and rel_err(o[0], ref[0]) < 1e-6


@pytest.mark.parametrize("model", [0, 1, 2])
@pytest.mark.parametrize("NZ", [128, 256])
def test_default_line_fp64(model, NZ):
    """Without set_config, every model takes the wavefront kernel at NZ >= 128 in FP64."""
    from ql_dns_amd.capi import QLContext
    c = QLContext(_params(model=model, NX=4, NY=4, NZ=NZ))
    assert c.config()[0]["line"] == WAVE
    c.close()


@pytest.mark.parametrize("integrator", [1, 0])
@pytest.mark.parametrize("NX,NY,NZ,target", [(16, 16, 32, 512), (8, 8, 64, 7), (6, 6, 128, 3), (6, 4, 256, 5)])
def test_bqlin_steps_match_oracle(integrator, NX, NY, NZ, target):
    """MHD_BQlin in the wavefront kernel (its dz uz from the ux and uy lines)."""
    p = _params(model=1, integrator=integrator, NX=NX, NY=NY, NZ=NZ)
    o, g = _pair(p, target)
    assert g.config()[0]["line"] == WAVE
    to, tg = o.steps(40, 0.0), g.steps(40, 0.0)
    assert tg == to
    lo, mo, ko = o.get_state()
    lg, mg, kg = g.get_state()
    assert np.array_equal(kg, ko)
    assert rel_err(mg, mo) < TOL and rel_err(lg, lo) < TOL
    assert rel_err(g.drive(), o.drive()) < TOL
    ro, rg = o.rhs(0.37, 0.011), g.rhs(0.37, 0.011)
    for a, b in zip(rg, ro):
        assert rel_err(a, b) < TOL


def test_noise_and_remap_match_oracle():
    """Driven steps (the noise kernel remaps after the step) with the wavefront
    kernel, remaps inside the window."""
    p = _params(NX=8, NY=8, NZ=128, q=1.5, Ly=1.0)
    o, g = _pair(p)
    k_init = g.get_state()[2]
    for x in (o, g):
        x.set_noise(0.8, seed=5)
    to, tg = o.steps(30, 0.0), g.steps(30, 0.0)
    assert tg == to
    lo, mo, ko = o.get_state()
    lg, mg, kg = g.get_state()
    assert np.array_equal(kg, ko) and not np.array_equal(kg, k_init)
    assert rel_err(mg, mo) < TOL and rel_err(lg, lo) < TOL
