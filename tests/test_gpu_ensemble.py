"""The 256-run parameter-study ensemble (Rm log-uniform in [500, 20000], q
uniform in [1.0, 1.8], NZ from {32, 64, 128}, NX and NY from {8, 16, 32},
the 16x16x32 box's other physics) as one run set: launch shapes tuned first,
then 20 steps of graph replay; every run within 1e-10 relative of the oracle
stepping that run alone (kx0 after remaps and the accumulated time exactly
equal)."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10
STEPS = 20


def test_ensemble_256_runs_match_oracle():
    from ql_dns_amd.capi import QLBatch
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import ensemble
    from oracle.oracle import Oracle
    runs = ensemble()
    assert len(runs) == 256 and len({p.NZ for p in runs}) == 3
    states = [initial_state(p) for p in runs]
    b = QLBatch(runs)
    for r, (lin, mf) in enumerate(states):
        b.set_state(r, lin, mf)
    b.tune(2)
    b.linear_ops_init(0.0)
    tb = b.steps(STEPS, 0.0)
    worst = 0.0
    for r, (p, (lin, mf)) in enumerate(zip(runs, states)):
        o = Oracle.from_params(p, nthreads=8)
        o.set_state(lin, mf)
        o.linear_ops_init(0.0)
        assert o.steps(STEPS, 0.0) == tb
        (lo, mo, ko), (lg, mg, kg) = o.get_state(), b.get_state(r)
        assert np.array_equal(kg, ko), r
        e = max(rel_err(mg, mo), rel_err(lg, lo), rel_err(b.drive(r), o.drive()))
        assert e < TOL, (r, p.NX, p.NY, p.NZ, e)
        worst = max(worst, e)
        o.close()
    b.close()
    print(f"256-run ensemble, {STEPS} steps: largest relative difference from the oracle {worst:.2e}")
