"""Golden fixtures of the 16x16x32 box (tests/golden/make_golden.py).

CPU: the oracle and the initial-state generator reproduce the fixtures
(regression anchor).  GPU: the HIP path, started from the fixture's initial
state, matches the fixture values to 1e-10 relative."""
import os

import numpy as np
import pytest

from conftest import ROOT, rel_err

GOLD = os.path.join(ROOT, "tests", "golden")
VARIANTS = ["fullqlin_elin", "fullqlin_cn", "bqlin_elin"]


def _load(name):
    return np.load(os.path.join(GOLD, f"box16x16x32_{name}.npz"), allow_pickle=False)


def _params(g):
    from ql_dns_amd.params import workload
    return workload("box16x16x32", model=int(g["model"]), integrator=int(g["integrator"]))


def test_initial_state_generator_reproduces_fixture():
    from ql_dns_amd.ic import initial_state
    g = _load("fullqlin_elin")
    lin, mf = initial_state(_params(g))
    assert np.array_equal(lin, g["ic_lin"]) and np.array_equal(mf, g["ic_mf"])


def _check(g, mf, drive, diag, kx, n, tol):
    assert rel_err(mf, g[f"mf{n}"]) < tol
    assert rel_err(drive, g[f"drive{n}"]) < tol
    assert np.array_equal(kx, g[f"kx{n}"])
    d0 = g[f"diag{n}"]
    fin = np.isfinite(d0)
    assert np.array_equal(fin, np.isfinite(diag))
    assert np.all(np.abs(diag[fin] - d0[fin]) <= tol * np.abs(d0[fin]) + 1e-300)


@pytest.mark.parametrize("name", VARIANTS)
def test_oracle_reproduces_fixture(name):
    from ql_dns_amd.ic import initial_state
    from oracle.oracle import Oracle
    g = _load(name)
    p = _params(g)
    lin, mf = initial_state(p)
    o = Oracle.from_params(p, nthreads=3)
    o.set_state(lin, mf)
    o.linear_ops_init(0.0)
    _, rmf, _ = o.rhs(0.0, 0.0053)
    assert rel_err(rmf, g["rhs_mf"]) < 1e-13
    t, done = 0.0, 0
    for n in (1, 10, 100):
        t = o.steps(n - done, t)
        done = n
        assert t == float(g[f"t{n}"])
        l_, m_, k_ = o.get_state()
        _check(g, m_, o.drive(), o.diagnostics(t), k_, n, 1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("name", VARIANTS)
def test_gpu_matches_fixture(name):
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    g = _load(name)
    p = _params(g)
    lin, mf = initial_state(p)
    c = QLContext(p)
    c.set_state(lin, mf)
    c.linear_ops_init(0.0)
    r_lin, r_mf, r_L = c.rhs(0.0, 0.0053)
    assert rel_err(r_mf, g["rhs_mf"]) < 1e-10
    assert rel_err(c.drive(), g["rhs_drive"]) < 1e-10
    if "rhs_lin" in g:
        assert rel_err(r_lin, g["rhs_lin"]) < 1e-10
        assert rel_err(r_L, g["rhs_linop"]) < 1e-10
    t, done = 0.0, 0
    for n in (1, 10, 100):
        t = c.steps(n - done, t)
        done = n
        assert t == float(g[f"t{n}"])
        l_, m_, k_ = c.get_state()
        _check(g, m_, c.drive(), c.diagnostics(t), k_, n, 1e-10)
        assert rel_err(np.abs(l_).sum(axis=(1, 2)), g[f"linnorm{n}"]) < 1e-10
    if "lin100" in g:
        assert rel_err(l_, g["lin100"]) < 1e-10
        assert rel_err(c.mf_real(), g["mfreal100"]) < 1e-10


# ----------------------------------------------------------------------------
# the 64x128x128 production box over its configured 1000 steps
# ----------------------------------------------------------------------------
BOX64 = "box64x128x128_fullqlin_elin_1000.npz"


def test_box64_fixture_initial_state():
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import workload
    g = np.load(os.path.join(GOLD, BOX64), allow_pickle=False)
    lin, mf = initial_state(workload("box64x128x128"))
    assert np.array_equal(mf, g["ic_mf"])
    assert rel_err(np.abs(lin).sum(axis=2), g["ic_linnorm"]) < 1e-15


@pytest.mark.gpu
def test_gpu_box64_1000_steps_matches_fixture():
    """The HIP path (graph replay, default launch shape) over the full 1000
    steps of the 64x128x128 box: mean fields, drives, diagnostics, per-mode
    norms and the full state of 16 modes within 1e-10 of the oracle; kx0 after
    the remaps and the accumulated time bit for bit."""
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import workload
    g = np.load(os.path.join(GOLD, BOX64), allow_pickle=False)
    p = workload("box64x128x128")
    lin, mf = initial_state(p)
    c = QLContext(p)
    c.set_state(lin, mf)
    c.linear_ops_init(0.0)
    t, done = 0.0, 0
    for n in (1, 100, 1000):
        t = c.steps(n - done, t)
        done = n
        assert t == float(g[f"t{n}"])
        l_, m_, k_ = c.get_state()
        _check(g, m_, c.drive(), c.diagnostics(t), k_, n, 1e-10)
        assert rel_err(np.abs(l_).sum(axis=2), g[f"linnorm{n}"]) < 1e-10
    assert rel_err(l_[g["keep"]], g["lin_keep"]) < 1e-10
    c.close()
