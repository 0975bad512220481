"""Run sets on the GPU: many runs of different sizes and physics stepped in the
same launches, graph replay, and the launch-shape tuner.

Bars: every run of a set within 1e-10 relative of the oracle stepping that run
alone (kx0 after remaps exactly equal); graph replay (ql_steps) bitwise equal
to host-driven stepping (ql_step + ql_remap), which runs the same kernels
eagerly; tuning leaves the state bitwise unchanged; two launch shapes, which
sum the stress partials in different orders, agree within 1e-12 and one shape
is bitwise reproducible.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _p(**kw):
    from ql_dns_amd.params import QLParams
    return QLParams(**kw)


# ragged sizes, three NZ groups, per-run physics
def _runs(model=0, integrator=1, precision=0):
    base = dict(model=model, integrator=integrator, precision=precision, dt=0.01)
    # box sizes chosen so that every run with remap on is remapped within 40 steps
    return [
        _p(NX=8, NY=8, NZ=32, q=1.5, nu=1 / 2000, eta=1 / 1500, Ly=1.0, **base),
        _p(NX=6, NY=4, NZ=32, q=1.1, nu=1 / 800, eta=1 / 900, Lx=2.0, Ly=1.0, **base),
        _p(NX=8, NY=6, NZ=64, q=1.7, nu=1 / 3000, eta=1 / 5000, B0z=0.05, Ly=1.0, **base),
        _p(NX=4, NY=4, NZ=128, q=1.3, nu=1 / 1000, eta=1 / 1000, viscosity_order=4, Ly=0.5, **base),
        _p(NX=10, NY=8, NZ=32, q=1.5, nu=1 / 4000, eta=1 / 4000, remap=0, **base),
        _p(NX=6, NY=6, NZ=64, q=1.6, nu=1 / 2500, eta=1 / 2000, quasilinear=0, Ly=1.0, **base),
    ]


def _states(runs):
    from ql_dns_amd.ic import initial_state
    return [initial_state(p, seed=1000 + r) for r, p in enumerate(runs)]


def _batch(runs, states, t0=0.0):
    from ql_dns_amd.capi import QLBatch
    b = QLBatch(runs)
    for r, (lin, mf) in enumerate(states):
        b.set_state(r, lin, mf)
    b.linear_ops_init(t0)
    return b


def _oracle(p, lin, mf, t0=0.0):
    from oracle.oracle import Oracle
    o = Oracle.from_params(p, nthreads=8)
    o.set_state(lin, mf)
    o.linear_ops_init(t0)
    return o


@pytest.mark.parametrize("model,integrator", [(0, 1), (0, 0), (1, 1), (1, 0)])
def test_batch_matches_oracle_per_run(model, integrator):
    runs = _runs(model, integrator)
    states = _states(runs)
    b = _batch(runs, states)
    n = 40  # graph of 16 steps twice + 8 single-step replays
    tb = b.steps(n, 0.0)
    for r, p in enumerate(runs):
        o = _oracle(p, *states[r])
        to = o.steps(n, 0.0)
        assert tb == to
        lo, mo, ko = o.get_state()
        lg, mg, kg = b.get_state(r)
        assert np.array_equal(kg, ko), f"run {r}: remapped kx0 differ"
        assert rel_err(mg, mo) < TOL, f"run {r}"
        assert rel_err(lg, lo) < TOL, f"run {r}"
        assert rel_err(b.drive(r), o.drive()) < TOL, f"run {r}"
        do, dg = o.diagnostics(to), b.diagnostics(r, tb)
        fin = np.isfinite(do)
        assert np.array_equal(fin, np.isfinite(dg))
        assert np.all(np.abs(dg[fin] - do[fin]) <= TOL * np.abs(do[fin]) + 1e-300), f"run {r}"
    # every run was remapped at least once where remap is on (the test covers it)
    for r, p in enumerate(runs):
        from ql_dns_amd.ic import mode_tables
        kx0_init = mode_tables(p)[0]
        moved = not np.array_equal(b.get_state(r)[2], kx0_init)
        assert moved == bool(p.remap), f"run {r}"


def test_graph_replay_equals_eager_steps():
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    p = _p(NX=16, NY=16, NZ=32)
    lin, mf = initial_state(p)
    a, e = QLContext(p), QLContext(p)
    for c in (a, e):
        c.set_state(lin, mf)
        c.linear_ops_init(0.0)
    n = 37
    ta = a.steps(n, 0.0)
    t = 0.0
    for _ in range(n):
        e.step(t)
        t = t + p.dt
        e.remap(p.q * t)
    assert ta == t
    la, ma, ka = a.get_state()
    le, me, ke = e.get_state()
    assert np.array_equal(ka, ke)
    assert np.array_equal(ma, me)
    assert np.array_equal(la, le)


def test_tune_keeps_state_and_picks_a_valid_shape():
    runs = _runs()
    states = _states(runs)
    b = _batch(runs, states)
    b.steps(3, 0.0)
    before = [b.get_state(r) for r in range(len(runs))]
    drive = [b.drive(r) for r in range(len(runs))]
    ms = b.tune(2)
    assert ms > 0
    for r in range(len(runs)):
        after = b.get_state(r)
        for x, y in zip(before[r], after):
            assert np.array_equal(x, y), f"run {r}: tuning changed the state"
        assert np.array_equal(drive[r], b.drive(r))
    cfg = b.config()
    assert [c["NZ"] for c in cfg] == [32, 64, 128]
    for c in cfg:
        assert c["line"] in (64, 128, 256)  # 64: one wavefront per mode
        assert c["modes_per_pass"] == (1 if c["line"] == 64 else max(1, c["line"] // c["NZ"]))
        assert c["workgroups"] >= 1


def test_shapes_agree_to_rounding_and_repeat_bitwise():
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    p = _p(NX=32, NY=32, NZ=64)
    lin, mf = initial_state(p)
    res = {}
    for shape in [(128, 256), (256, 512), (256, 1 << 30), (128, 4096), (256, 512)]:
        c = QLContext(p)
        c.set_config(64, *shape)
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


def test_batch_mixed_precision_close_to_fp64():
    runs64 = _runs(precision=0)
    runs32 = _runs(precision=1)
    states = _states(runs64)
    b64, b32 = _batch(runs64, states), _batch(runs32, states)
    b64.steps(50, 0.0)
    b32.steps(50, 0.0)
    for r in range(len(runs64)):
        l64, m64, k64 = b64.get_state(r)
        l32, m32, k32 = b32.get_state(r)
        assert np.array_equal(k64, k32)
        assert rel_err(m32, m64) < 1e-6, f"run {r}"
        assert rel_err(l32, l64) < 1e-6, f"run {r}"


def test_batch_bad_run_index():
    from ql_dns_amd.capi import QLError
    runs = _runs()[:2]
    b = _batch(runs, _states(runs))
    with pytest.raises(QLError):
        b.diagnostics(2, 0.0)
    with pytest.raises(QLError):
        b.get_state(-1)
    with pytest.raises(QLError):
        b.set_config(256, 256, 512)  # no NZ=256 run in this set
