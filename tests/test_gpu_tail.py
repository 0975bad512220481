"""The substep tail as one launch (k_mf_fused, QL_TAIL_FUSED) against the
column reduction and mean-field kernel as two launches.

Both sum the stress / EMF partial rows in the same order and update the mean
fields with the same arithmetic.  For MHD_FullQlin and HD_fullU every result
is bitwise equal: states, mean fields, kx0, drives and times, both
integrators, NZ 32 / 64 / 128, both mode-kernel shapes, mixed precision,
driving noise, run sets, graph replay and eager steps, and the adjoint (whose
recomputation reads the mean fields between substeps from the buffers the
one-launch tail writes them to).  MHD_BQlin's results differ in the last
bits (1e-16 relative after 24 steps; times and kx0 equal), which the test
bounds at 1e-13, as it bounds two launch shapes.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _p(**kw):
    from ql_dns_amd.params import QLParams
    return QLParams(**kw)


def _run(p, tail, n=24, shape=None, noise=None, eager=False):
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    lin, mf = initial_state(p)
    c = QLContext(p)
    if shape:
        c.set_config(p.NZ, *shape)
    c.set_tail(p.NZ, tail)
    if noise:
        c.set_noise(**noise)
    c.set_state(lin, mf)
    c.linear_ops_init(0.0)
    if eager:
        t = 0.0
        for _ in range(n):
            t = t + c.step(t)
            c.remap(p.q * t)
    else:
        t = c.steps(n, 0.0)
    out = (t,) + tuple(c.get_state()) + (c.drive(),)
    c.close()
    return out


def _equal(a, b, tol=0.0):
    assert a[0] == b[0] and np.array_equal(a[3], b[3])  # times, kx0
    for x, y in zip(a[1:], b[1:]):
        if tol:
            assert rel_err(x, y) < tol
        else:
            assert np.array_equal(x, y)


@pytest.mark.parametrize("model,integrator,NX,NY,NZ,shape", [
    (0, 1, 64, 128, 128, None), (0, 0, 16, 16, 128, None), (1, 1, 16, 16, 128, None), (2, 1, 16, 16, 128, None),
    (0, 1, 16, 16, 64, None), (0, 0, 16, 16, 32, None), (1, 0, 8, 8, 64, None),
    (0, 1, 16, 16, 128, (256, 512)), (0, 1, 16, 16, 64, (64, 512))])
def test_fused_tail_equals_two_kernel_tail(model, integrator, NX, NY, NZ, shape):
    p = _p(model=model, integrator=integrator, NX=NX, NY=NY, NZ=NZ, Ly=1.0, B0z=0.02)
    _equal(_run(p, 1, shape=shape), _run(p, 0, shape=shape), 1e-13 if model == 1 else 0.0)


def test_fused_tail_mixed_noise_and_eager():
    from ql_dns_amd.params import PREC_MIXED
    p = _p(NX=32, NY=32, NZ=128, Ly=1.0, precision=PREC_MIXED)
    _equal(_run(p, 1), _run(p, 0))
    # the workgroup shape: the FP64 companion's partial rows follow the FP32 launch's
    _equal(_run(p, 1, shape=(256, 512)), _run(p, 0, shape=(256, 512)))
    noise = dict(f_noise=2.0, seed=9)
    q = _p(NX=16, NY=16, NZ=64, Ly=1.0)
    _equal(_run(q, 1, noise=noise), _run(q, 0, noise=noise))
    # eager steps (ql_step + ql_remap) against graph replay, across tails
    _equal(_run(q, 1, n=7, eager=True), _run(q, 0, n=7))


def test_fused_tail_run_set():
    from ql_dns_amd.capi import QLBatch
    from ql_dns_amd.ic import initial_state
    runs = [_p(NX=8, NY=8, NZ=32, q=1.5), _p(NX=6, NY=4, NZ=32, q=1.1, Lx=2.0),
            _p(NX=8, NY=6, NZ=64, q=1.7, B0z=0.05), _p(NX=4, NY=4, NZ=128, viscosity_order=4),
            _p(NX=8, NY=8, NZ=128, q=1.3)]
    st = [initial_state(p, seed=40 + r) for r, p in enumerate(runs)]
    res = []
    for tail in (1, 0):
        b = QLBatch(runs)
        for NZ in (32, 64, 128):
            b.set_tail(NZ, tail)
        for r, (lin, mf) in enumerate(st):
            b.set_state(r, lin, mf)
        b.linear_ops_init(0.0)
        t = b.steps(21, 0.0)
        res.append([t] + [b.get_state(r) for r in range(len(runs))])
        b.close()
    assert res[0][0] == res[1][0]
    for a, c in zip(res[0][1:], res[1][1:]):
        for x, y in zip(a, c):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("model,integrator", [(0, 1), (1, 0)])
def test_fused_tail_adjoint(model, integrator):
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    p = _p(model=model, integrator=integrator, NX=8, NY=8, NZ=32, Ly=1.0)
    lin, mf = initial_state(p)
    out = []
    for tail in (1, 0):
        c = QLContext(p)
        c.set_tail(p.NZ, tail)
        c.set_state(lin, mf)
        c.linear_ops_init(0.0)
        c.adjoint_forward(9, 4, 0.0)
        lT, mT, _ = c.get_state()
        out.append(c.adjoint_backward(1e-3 * lT, mT))
        c.close()
    for x, y in zip(*out):
        if model == 1:
            assert rel_err(np.asarray(x), np.asarray(y)) < 1e-12
        else:
            assert np.array_equal(np.asarray(x), np.asarray(y))


def test_tuner_times_both_tails():
    """The tuner times the one-launch tail where it applies and reports the
    tail it keeps; results agree with the two-kernel tail."""
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    p = _p(NX=16, NY=16, NZ=32, Ly=1.0)
    lin, mf = initial_state(p)
    c = QLContext(p)
    c.set_state(lin, mf)
    c.tune(2)
    cfg = c.config()[0]
    assert cfg["tail"] in (0, 1)
    c.set_config(p.NZ, cfg["line"], cfg["target"])
    c.set_tail(p.NZ, 1 - cfg["tail"])
    assert c.config()[0]["tail"] == 1 - cfg["tail"]
    c.close()
    q = _p(NX=4, NY=4, NZ=32, dt=-0.01)
    c = QLContext(q)
    c.set_state(*initial_state(q))
    c.tune(1)
    assert c.config()[0]["tail"] == 0  # variable steps: the two-kernel tail only
    c.close()


def test_fused_tail_refusals():
    from ql_dns_amd.capi import QLContext, QLError
    for kw in (dict(NZ=32, dt=-0.01), dict(NZ=256)):
        c = QLContext(_p(NX=4, NY=4, **kw))
        with pytest.raises(QLError):
            c.set_tail(kw["NZ"], 1)
        c.close()
    c = QLContext(_p(NX=4, NY=4, NZ=32))
    with pytest.raises(QLError):
        c.set_tail(32, 2)
    with pytest.raises(QLError):
        c.set_tail(64, 1)  # no such group
    c.close()
