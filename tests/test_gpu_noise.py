"""Driving noise on the GPU against the oracle (Models/MHD_FullQlin.cpp:786-1053,
MHD_BQlin.cpp:558-740).

Both draw the same Philox4x64-10 normals, so the GPU's noise agrees with the
oracle's to the rounding of log, sqrt, sin and cos (1e-12 relative); entries
the noise must not touch stay bit for bit; ky = 0 pairs are exact conjugates.
Inside ql_steps the noise comes after each step and before the remap
(main.cpp:111-123), from the replayed graph as from eager launches.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _params(**kw):
    from ql_dns_amd.params import QLParams
    return QLParams(**kw)


def _pair(p, noise):
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from oracle.oracle import Oracle
    lin, mf = initial_state(p)
    o = Oracle.from_params(p, nthreads=8)
    o.set_state(lin, mf)
    o.linear_ops_init(0.0)
    o.set_noise(**noise)
    g = QLContext(p)
    g.set_state(lin, mf)
    g.linear_ops_init(0.0)
    g.set_noise(**noise)
    return o, g, lin


NOISES = [dict(f_noise=0.8, seed=5), dict(f_noise=0.3, range_low=8.0, range_high=30.0, seed=6),
          dict(f_noise=1.0, drive_only_velocity=1, seed=7), dict(f_noise=1.0, drive_only_magnetic=1, seed=8)]


@pytest.mark.parametrize("model,NX,NY,NZ", [(0, 16, 16, 32), (1, 16, 16, 32), (0, 8, 8, 64), (0, 8, 4, 128),
                                            (1, 4, 4, 256)])
@pytest.mark.parametrize("k", range(len(NOISES)))
def test_driving_noise_matches_oracle(model, NX, NY, NZ, k):
    p = _params(model=model, NX=NX, NY=NY, NZ=NZ)
    o, g, lin0 = _pair(p, NOISES[k])
    t, dt = 0.37, 0.013
    o.driving_noise(t, dt)
    g.driving_noise(t, dt)
    lo = o.get_state()[0]
    lg = g.get_state()[0]
    untouched = lo == lin0
    assert np.array_equal(lg[untouched], lo[untouched])
    assert rel_err(lg - lin0, lo - lin0) < 1e-12
    # ky = 0 pairs: exact conjugates of each other's noise
    _, ky, kxi, kyi, _ = o.tables()
    d = lg - lin0
    nz = g.nz
    for i in range(g.nxy):
        if kyi[i] == 0 and kxi[i] > 0:
            j = int(np.nonzero((kyi == 0) & (kxi == -kxi[i]))[0][0])
            assert np.array_equal(d[i, :, 1:], np.conj(d[j, :, :0:-1]))


@pytest.mark.parametrize("model,integrator", [(0, 1), (0, 0), (1, 1)])
def test_steps_with_noise_match_oracle(model, integrator):
    p = _params(model=model, integrator=integrator)
    o, g, _ = _pair(p, dict(f_noise=0.5, seed=31))
    to = o.steps(21, 0.0)
    tg = g.steps(21, 0.0)  # one 16-step graph, then 5 one-step graphs
    assert tg == to
    lo, mo, ko = o.get_state()
    lg, mg, kg = g.get_state()
    assert np.array_equal(kg, ko)
    assert rel_err(mg, mo) < 1e-10
    assert rel_err(lg, lo) < 1e-10


@pytest.mark.parametrize("model,integrator,NZ", [(0, 1, 32), (0, 0, 128), (2, 1, 64), (1, 1, 256), (2, 0, 256)])
def test_variable_steps_with_noise_match_oracle(model, integrator, NZ):
    """Driven variable steps.  The noise is keyed on the bits of the step's
    start time, and a variable step size agrees with the oracle's only to
    rounding (1e-12), so after a step whose size differs in the last bit the
    two draw different noise.  The oracle therefore replays the GPU's step
    sizes (each checked against the oracle's own CFL step to 1e-12); states
    then agree to 1e-10 at every step, and the GPU's replayed graph equals
    its step loop bit for bit."""
    p = _params(model=model, integrator=integrator, NX=8, NY=8, NZ=NZ, dt=-0.05, CFL=0.5)
    o, g, lin0 = _pair(p, dict(f_noise=0.5, seed=41))
    mf0 = o.get_state()[1]
    t = 0.0
    for n in range(12):
        v = o.calculate_cfl()
        dt = g.step(t)
        assert abs(dt - min(0.5 / v, 0.05)) <= 1e-12 * dt
        o.set_dt(dt)
        assert o.step(t) == dt
        for x in (o, g):
            x.driving_noise(t, dt)
        t += dt
        for x in (o, g):
            x.remap(p.q * t)
        lo, mo, ko = o.get_state()
        lg, mg, kg = g.get_state()
        assert np.array_equal(kg, ko)
        assert rel_err(mg, mo) < 1e-10 and rel_err(lg, lo) < 1e-10, n
    # the same 12 driven steps from the replayed graph
    from ql_dns_amd.capi import QLContext
    h = QLContext(p)
    h.set_state(lin0, mf0)
    h.linear_ops_init(0.0)
    h.set_noise(f_noise=0.5, seed=41)
    assert h.steps(12, 0.0) == t
    for a, b in zip(h.get_state(), g.get_state()):
        assert np.array_equal(a, b)


def test_noise_graph_replay_equals_eager_and_step_loop():
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    p = _params(dt=-0.02, CFL=0.5)  # variable steps too
    lin, mf = initial_state(p)
    res = []
    for how in ("graph", "eager", "loop"):
        g = QLContext(p)
        g.set_state(lin, mf)
        g.linear_ops_init(0.0)
        g.set_noise(0.4, seed=77)
        if how == "graph":
            t = g.steps(20, 0.0)
        elif how == "eager":
            t = g.profile_kernels(20, 0.0)[0]
        else:
            # the driver loop through the single-step entry points (main.cpp:111-123)
            t = 0.0
            for n in range(20):
                dt = g.step(t)
                g.driving_noise(t, dt)
                t += dt
                g.remap(p.q * t)
        res.append((t,) + tuple(g.get_state()))
    for r in res[1:]:
        assert r[0] == res[0][0]
        for a, b in zip(r[1:], res[0][1:]):
            assert np.array_equal(a, b)


def test_noise_off_again_restores_the_plain_step():
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    p = _params()
    lin, mf = initial_state(p)
    res = []
    for toggle in (False, True):
        g = QLContext(p)
        g.set_state(lin, mf)
        g.linear_ops_init(0.0)
        if toggle:
            g.set_noise(1.0, seed=1)
            g.set_noise(0.0)
        t = g.steps(17, 0.0)
        res.append((t,) + tuple(g.get_state()))
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)


def test_noise_in_run_sets():
    """Per-run noise in a set: each run as its own context (the sets' launch
    shapes sum the stresses in another order, so to rounding)."""
    from ql_dns_amd.capi import QLBatch, QLContext
    from ql_dns_amd.ic import initial_state
    ps = [_params(), _params(NX=8, NY=8, NZ=64, q=1.2), _params(NX=8, NY=8, q=1.7)]
    noises = [dict(f_noise=0.5, seed=1), None, dict(f_noise=0.9, range_high=25.0, seed=2)]
    b = QLBatch(ps)
    states = [initial_state(p) for p in ps]
    for r, (p, (lin, mf)) in enumerate(zip(ps, states)):
        b.set_state(r, lin, mf)
        if noises[r]:
            b.set_noise(r, **noises[r])
    b.linear_ops_init(0.0)
    tb = b.steps(20, 0.0)
    for r, (p, (lin, mf)) in enumerate(zip(ps, states)):
        g = QLContext(p)
        g.set_state(lin, mf)
        g.linear_ops_init(0.0)
        if noises[r]:
            g.set_noise(**noises[r])
        assert g.steps(20, 0.0) == tb
        (lg, mg, kg), (lb, mb, kb) = g.get_state(), b.get_state(r)
        assert np.array_equal(kg, kb)
        assert rel_err(lb, lg) < 1e-10 and rel_err(mb, mg) < 1e-10

