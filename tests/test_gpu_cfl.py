"""Variable time steps and Check_Solution on the GPU, against the oracle.

Variable steps (dt < 0, Integrators/RK3CN_Elin.cpp:76-80, RK3CN.cpp:79-83) are
chosen on the device from profile maxima that the mean-field kernels record;
the bar is the oracle's step sizes and states to 1e-12 / 1e-10 relative (each
dt is a function of transformed profiles, so it agrees to rounding), every
device dt equal bit for bit to CFL / ql_calculate_cfl capped at -dt, capped
variable steps equal to fixed steps bit for bit, and graph replay equal to
eager launches bit for bit.  Check_Solution (solution.cpp:169-203) sums agree
with the oracle's to 1e-12 and its verdicts match on planted states.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _params(**kw):
    from ql_dns_amd.params import QLParams
    return QLParams(**kw)


def _pair(p, seed=None):
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from oracle.oracle import Oracle
    lin, mf = initial_state(p, seed=seed)
    o = Oracle.from_params(p, nthreads=8)
    o.set_state(lin, mf)
    o.linear_ops_init(0.0)
    g = QLContext(p)
    g.set_state(lin, mf)
    g.linear_ops_init(0.0)
    return o, g


CASES = [(0, 1, 16, 16, 32), (0, 0, 8, 8, 64), (1, 1, 16, 16, 32), (0, 1, 8, 4, 128), (1, 0, 4, 4, 256)]


@pytest.mark.parametrize("model,integrator,NX,NY,NZ", CASES)
def test_variable_steps_match_oracle(model, integrator, NX, NY, NZ):
    cfl, dtmax = 0.5, 0.05
    p = _params(model=model, integrator=integrator, NX=NX, NY=NY, NZ=NZ, dt=-dtmax, CFL=cfl)
    o, g = _pair(p)
    to = tg = 0.0
    limited = 0
    for n in range(6):
        vo, vg = o.calculate_cfl(), g.calculate_cfl()
        assert abs(vg - vo) <= 1e-12 * vo
        dto, dtg = o.step(to), g.step(tg)
        assert dtg == min(cfl / vg, dtmax)
        assert abs(dtg - dto) <= 1e-12 * dto
        limited += dtg < dtmax
        to += dto
        tg += dtg
        o.remap(p.q * to)
        g.remap(p.q * tg)
    assert limited >= 4
    # the loop of main.cpp:108-123 from a replayed graph (16 + 4 steps)
    to = o.steps(20, to)
    tg = g.steps(20, tg)
    assert abs(tg - to) <= 1e-12 * to
    lo, mo, ko = o.get_state()
    lg, mg, kg = g.get_state()
    assert np.array_equal(kg, ko)
    assert rel_err(mg, mo) < TOL
    assert rel_err(lg, lo) < TOL
    assert abs(g.calculate_cfl() - o.calculate_cfl()) <= 1e-12 * o.calculate_cfl()


@pytest.mark.parametrize("model,integrator", [(0, 1), (0, 0), (1, 1)])
def test_capped_variable_steps_equal_fixed_steps(model, integrator):
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    base = _params(model=model, integrator=integrator)
    lin, mf = initial_state(base)
    out = []
    for p in (base.replace(dt=0.01), base.replace(dt=-0.01, CFL=1e6)):
        g = QLContext(p)
        g.set_state(lin, mf)
        g.linear_ops_init(0.0)
        t = g.step(0.0)
        g.remap(p.q * t)
        t = g.steps(37, t)
        out.append((t,) + tuple(g.get_state()))
    (ta, la, ma, ka), (tb, lb, mb, kb) = out
    assert ta == tb
    assert np.array_equal(ka, kb) and np.array_equal(ma, mb) and np.array_equal(la, lb)


def test_variable_graph_replay_equals_eager():
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    p = _params(integrator=0, dt=-0.05, CFL=0.4)
    lin, mf = initial_state(p)
    res = []
    for eager in (False, True):
        g = QLContext(p)
        g.set_state(lin, mf)
        g.linear_ops_init(0.0)
        t = g.profile_kernels(40, 0.0)[0] if eager else g.steps(40, 0.0)
        res.append((t, g.calculate_cfl()) + tuple(g.get_state()))
    assert res[0][0] == res[1][0] and res[0][1] == res[1][1]
    for a, b in zip(res[0][2:], res[1][2:]):
        assert np.array_equal(a, b)


def test_tuner_keeps_the_variable_step_state():
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    p = _params(dt=-0.05, CFL=0.5)
    lin, mf = initial_state(p)
    res = []
    for tune in (False, True):
        g = QLContext(p)
        g.set_state(lin, mf)
        g.linear_ops_init(0.0)
        t = g.steps(5, 0.0)
        v = g.calculate_cfl()
        if tune:
            g.tune(2)
            assert g.calculate_cfl() == v
        t = g.steps(10, t)
        res.append((t,) + tuple(g.get_state()))
    assert abs(res[1][0] - res[0][0]) <= 1e-12 * res[0][0]
    assert rel_err(res[1][1], res[0][1]) < TOL and rel_err(res[1][2], res[0][2]) < TOL


# ----------------------------------------------------------------------------
# Check_Solution
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("model,NX,NY,NZ", [(0, 16, 16, 32), (0, 8, 8, 64), (1, 8, 4, 128), (0, 4, 4, 256)])
def test_check_sums_match_oracle(model, NX, NY, NZ):
    p = _params(model=model, NX=NX, NY=NY, NZ=NZ)
    o, g = _pair(p)
    for n in range(2):
        so = o.check_sums()
        v, sg = g.check_solution()
        assert v == 0 == o.check_solution()
        for k in (0, 2, 3, 4):
            assert abs(sg[k] - so[k]) <= 1e-12 * so[k]
        assert sg[1] < 1e-12 * sg[0] and so[1] < 1e-12 * so[0]
        t = o.steps(3, 0.0)
        assert g.steps(3, 0.0) == t


def test_check_solution_verdicts_match_oracle():
    from ql_dns_amd.capi import QLError
    p = _params()
    o, g = _pair(p)
    lin, _ = o.get_state()[:2]
    planted = []
    z = np.zeros((o.nMF, o.nz), complex)
    z[1, 1] = 1.0  # one-sided By mode: complex profile
    planted.append((z, 1))
    z = np.zeros((o.nMF, o.nz), complex)
    z[1, 1] = 1e-6  # ... below the size threshold
    planted.append((z, 0))
    for f in (0, 1):
        z = np.zeros((o.nMF, o.nz), complex)
        z[f, 0] = 2e20
        planted.append((z, 2))
    z = np.zeros((o.nMF, o.nz), complex)
    z[1, 2] = np.nan
    planted.append((z, 3))
    for mf, want in planted:
        o.set_state(lin, mf)
        g.set_state(lin, mf)
        assert o.check_solution() == want
        v, s = g.check_solution()
        assert v == want
        if want:
            with pytest.raises(QLError) as e:
                g.check_solution(strict=True)
            assert e.value.code == -4 and "ERROR" in str(e.value)


@pytest.mark.parametrize("model", [0, 1, 2])
def test_fixed_step_cfl_is_the_current_state_value(model):
    """With fixed steps nothing is recorded while stepping; ql_calculate_cfl
    evaluates the current state, i.e. the oracle's value right after an rhs."""
    p = _params(model=model)
    o, g = _pair(p)
    to = o.steps(7, 0.0)
    g.steps(7, 0.0)
    o.rhs(to, 0.0)
    g.rhs(to, 0.0)
    vo, vg = o.calculate_cfl(), g.calculate_cfl()
    assert abs(vg - vo) <= 1e-12 * vo


@pytest.mark.parametrize("integrator", [1, 0])
def test_variable_steps_in_a_run_set(integrator):
    """A set of runs of different sizes and physics with variable steps: each
    run keeps its own clock and chooses its own steps (ql_batch_steps_runs);
    every run's times and state match the oracle stepping that run alone."""
    from ql_dns_amd.capi import QLBatch, QLError
    from ql_dns_amd.ic import initial_state
    from oracle.oracle import Oracle
    base = dict(integrator=integrator, dt=-0.05, CFL=0.5)
    runs = [_params(NX=16, NY=16, NZ=32, q=1.5, **base),
            _params(NX=8, NY=8, NZ=32, q=1.2, nu=1 / 600, eta=1 / 700, Ly=1.0, **base),
            _params(NX=8, NY=6, NZ=64, q=1.7, B0z=0.05, Ly=1.0, **base),
            _params(NX=4, NY=4, NZ=128, q=1.3, Ly=0.5, **base)]
    states = [initial_state(p, seed=900 + r, mf_amp=0.3 + 0.4 * r) for r, p in enumerate(runs)]
    b = QLBatch(runs)
    for r, (lin, mf) in enumerate(states):
        b.set_state(r, lin, mf)
    b.linear_ops_init(0.0)
    with pytest.raises(QLError):
        b.steps(1, 0.0)  # one shared time does not fit runs with their own steps
    tg = b.steps_runs(12, np.zeros(len(runs)))
    tg = b.steps_runs(8, tg)
    assert len(set(tg.tolist())) == len(runs)  # the runs took different steps
    for r, (p, (lin, mf)) in enumerate(zip(runs, states)):
        o = Oracle.from_params(p, nthreads=8)
        o.set_state(lin, mf)
        o.linear_ops_init(0.0)
        to = o.steps(20, 0.0)
        assert abs(tg[r] - to) <= 1e-12 * to, r
        (lo, mo, ko), (lg, mg, kg) = o.get_state(), b.get_state(r)
        assert np.array_equal(kg, ko)
        assert rel_err(mg, mo) < TOL and rel_err(lg, lo) < TOL, r
        assert abs(b.calculate_cfl(r) - o.calculate_cfl()) <= 1e-12 * o.calculate_cfl()
        o.close()
    b.close()
