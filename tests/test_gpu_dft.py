"""The workgroup mode kernel with its dealiased z transforms as matrix-core
products (QL_TRANSFORM_DFT_MFMA, v_mfma_f64_16x16x4f64; NZ 32 and 64)
against the oracle: states, drives and diagnostics within 1e-10 relative
(kx0 after remaps and the times exact), the right-hand side of one call, and
agreement with the radix transforms of the same launch shape to 1e-12."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10
DFT, FFT = 1, 0


def _pair(p, line, target=512, transform=DFT):
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from oracle.oracle import Oracle
    lin, mf = initial_state(p)
    o = Oracle.from_params(p, nthreads=8)
    g = QLContext(p)
    g.set_config(p.NZ, line, target)
    g.set_transform(p.NZ, transform)
    for x in (o, g):
        x.set_state(lin, mf)
        x.linear_ops_init(0.0)
    return o, g


@pytest.mark.parametrize("model", [0, 1, 2])
@pytest.mark.parametrize("integrator", [1, 0])
@pytest.mark.parametrize("NZ,line", [(32, 128), (32, 256), (64, 128), (64, 256)])
def test_dft_steps_match_oracle(model, integrator, NZ, line):
    from ql_dns_amd.params import QLParams
    p = QLParams(model=model, integrator=integrator, NX=8, NY=8, NZ=NZ, B0z=0.02, Ly=1.0)
    o, g = _pair(p, line)
    assert g.config()[0]["transform"] == DFT
    to, tg = o.steps(30, 0.0), g.steps(30, 0.0)
    assert tg == to
    (lo, mo, ko), (lg, mg, kg) = o.get_state(), g.get_state()
    assert np.array_equal(kg, ko)
    assert rel_err(mg, mo) < TOL and rel_err(lg, lo) < TOL
    assert rel_err(g.drive(), o.drive()) < TOL
    do, dg = o.diagnostics(to), g.diagnostics(tg)
    assert np.all(np.abs(dg - do) <= TOL * np.abs(do) + 1e-300)


@pytest.mark.parametrize("NZ", [32, 64])
def test_dft_rhs_matches_oracle_and_radix(NZ):
    from ql_dns_amd.params import QLParams
    p = QLParams(NX=16, NY=16, NZ=NZ, B0z=0.03)
    o, g = _pair(p, 128)
    _, f = _pair(p, 128, transform=FFT)
    ro, rg, rf = o.rhs(0.37, 0.011), g.rhs(0.37, 0.011), f.rhs(0.37, 0.011)
    for a, b, c in zip(rg, ro, rf):
        assert rel_err(a, b) < TOL
        assert rel_err(a, c) < 1e-12
    assert rel_err(g.drive(), o.drive()) < TOL and rel_err(g.drive(), f.drive()) < 1e-12


def test_dft_refused_where_it_does_not_exist():
    from ql_dns_amd.capi import QLContext, QLError
    from ql_dns_amd.params import PREC_MIXED, QLParams
    for p, line in ((QLParams(NX=8, NY=8, NZ=128), 256), (QLParams(NX=8, NY=8, NZ=32, precision=PREC_MIXED), 128),
                    (QLParams(NX=8, NY=8, NZ=64), 64)):
        c = QLContext(p)
        c.set_config(p.NZ, line, 512)
        with pytest.raises(QLError) as e:
            c.set_transform(p.NZ, DFT)
        assert e.value.code == -1
        assert c.config()[0]["transform"] == FFT
        c.close()


def test_tuner_times_the_matrix_core_candidate():
    """The tuner picks among both transforms at NZ 32 and 64 and leaves the
    state bitwise unchanged; stepping with what it picked matches the oracle."""
    from ql_dns_amd.capi import QLBatch
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import QLParams
    from oracle.oracle import Oracle
    runs = [QLParams(NX=16, NY=16, NZ=32), QLParams(NX=16, NY=8, NZ=64, q=1.3)]
    states = [initial_state(p, seed=40 + r) for r, p in enumerate(runs)]
    b = QLBatch(runs)
    for r, (lin, mf) in enumerate(states):
        b.set_state(r, lin, mf)
    b.tune(2)
    for r, (lin, mf) in enumerate(states):
        lg, mg, _ = b.get_state(r)
        assert np.array_equal(lg, lin) and np.array_equal(mg, mf)
    assert all(c["transform"] in (FFT, DFT) for c in b.config())
    b.linear_ops_init(0.0)
    tb = b.steps(10, 0.0)
    for r, (p, (lin, mf)) in enumerate(zip(runs, states)):
        o = Oracle.from_params(p, nthreads=8)
        o.set_state(lin, mf)
        o.linear_ops_init(0.0)
        assert o.steps(10, 0.0) == tb
        lo, mo, _ = o.get_state()
        lg, mg, _ = b.get_state(r)
        assert rel_err(mg, mo) < TOL and rel_err(lg, lo) < TOL
    b.close()
