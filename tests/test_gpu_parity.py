"""GPU parity: the HIP path (called through the C ABI) against the CPU oracle.

Bar: FP64 results within 1e-10 relative (norm-wise, max|a-b|/max|b|) of the
oracle on identical seeded inputs; integer/index data (kx0 after remaps, step
times) exactly equal.  Sizes cover every NZ the kernels are instantiated for
(32, 64, 128, 256), both models, both integrators, ragged mode counts, and
non-default physics (B0z, hyperviscosity, quasilinear off, t0 != 0).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _ctx(p):
    from ql_dns_amd.capi import QLContext
    return QLContext(p)


def _pair(p, t0=0.0, nthreads=8, seed=None):
    from ql_dns_amd.ic import initial_state
    from oracle.oracle import Oracle
    lin, mf = initial_state(p, seed=seed)
    o = Oracle.from_params(p, nthreads=nthreads)
    o.set_state(lin, mf)
    o.linear_ops_init(t0)
    g = _ctx(p)
    g.set_state(lin, mf)
    g.linear_ops_init(t0)
    return o, g


def _params(**kw):
    from ql_dns_amd.params import QLParams
    return QLParams(**kw)


def _compare_state(o, g, to, tg):
    assert tg == to
    lo, mo, ko = o.get_state()
    lg, mg, kg = g.get_state()
    assert np.array_equal(kg, ko), "remapped kx0 differ"
    assert rel_err(mg, mo) < TOL
    assert rel_err(lg, lo) < TOL
    assert rel_err(g.drive(), o.drive()) < TOL
    do, dg = o.diagnostics(to), g.diagnostics(tg)
    fin = np.isfinite(do)
    assert np.array_equal(fin, np.isfinite(dg))
    assert np.all(np.abs(dg[fin] - do[fin]) <= TOL * np.maximum(np.abs(do[fin]), 1e-300) + 1e-300)
    assert rel_err(g.mf_real(), o.mf_real()) < TOL


SHAPES = [(16, 16, 32), (6, 6, 32), (8, 8, 64), (10, 6, 64), (8, 4, 128), (4, 4, 256)]


@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("shape", SHAPES)
def test_rhs_parity(model, shape):
    NX, NY, NZ = shape
    p = _params(NX=NX, NY=NY, NZ=NZ, model=model)
    o, g = _pair(p)
    for t, dtl in [(0.0, 0.0053), (0.37, 0.002)]:
        ro = o.rhs(t, dtl)
        rg = g.rhs(t, dtl)
        assert rel_err(rg[0], ro[0]) < TOL
        assert rel_err(rg[1], ro[1]) < TOL
        assert rel_err(rg[2], ro[2]) < TOL
        assert rel_err(g.drive(), o.drive()) < TOL
    # rhs must not modify the state
    lg, mg, _ = g.get_state()
    lo, mo, _ = o.get_state()
    assert np.array_equal(lg, lo) and np.array_equal(mg, mo)


@pytest.mark.parametrize("model,integrator", [(0, 1), (0, 0), (1, 1), (1, 0)])
def test_steps_box16(model, integrator):
    from ql_dns_amd.params import workload
    p = workload("box16x16x32", model=model, integrator=integrator)
    o, g = _pair(p)
    to = o.steps(p.nsteps, 0.0)
    tg = g.steps(p.nsteps, 0.0)
    _compare_state(o, g, to, tg)


def test_steps_box32_200():
    from ql_dns_amd.params import workload
    p = workload("box32x32x64")
    o, g = _pair(p)
    to = o.steps(200, 0.0)
    tg = g.steps(200, 0.0)
    _compare_state(o, g, to, tg)


@pytest.mark.parametrize("integrator", [0, 1])
@pytest.mark.parametrize("NZ", [128, 256])
def test_steps_large_nz(integrator, NZ):
    p = _params(NX=8, NY=8, NZ=NZ, integrator=integrator, dt=0.005, nu=1 / 8000, eta=1 / 8000)
    o, g = _pair(p)
    to = o.steps(25, 0.0)
    tg = g.steps(25, 0.0)
    _compare_state(o, g, to, tg)


@pytest.mark.parametrize("case", [
    dict(B0z=0.3), dict(viscosity_order=4, nu=1e-7, eta=2e-7), dict(quasilinear=0),
    dict(omega=0.8, q=1.2), dict(remap=0), dict(NX=2, NY=4), dict(viscosity_order=3)])
@pytest.mark.parametrize("model", [0, 1])
def test_steps_physics_variants(case, model):
    kw = dict(NX=12, NY=8, NZ=64, dt=0.01, model=model)
    kw.update(case)
    p = _params(**kw)
    t0 = 0.37
    o, g = _pair(p, t0=t0)
    to = o.steps(20, t0)
    tg = g.steps(20, t0)
    _compare_state(o, g, to, tg)


@pytest.mark.parametrize("model,integrator", [(0, 1), (0, 0), (1, 1), (2, 1)])
@pytest.mark.parametrize("NZ", [32, 256])
def test_single_zero_mode_box(model, integrator, NZ):
    """NX = NY = 2: the only mode is (kx, ky) = (0, 0), which stays zero, so
    the drive is zero and the mean fields evolve under their linear terms
    alone (one workgroup, one mode)."""
    p = _params(model=model, integrator=integrator, NX=2, NY=2, NZ=NZ)
    o, g = _pair(p)
    assert g.nxy == 1
    to, tg = o.steps(30, 0.0), g.steps(30, 0.0)
    _compare_state(o, g, to, tg)
    lg, mg, _ = g.get_state()
    assert np.all(lg == 0) and np.all(g.drive() == 0)
    assert np.abs(mg).max() > 0


def test_single_step_and_remap_entry_points():
    from ql_dns_amd.params import workload
    p = workload("box16x16x32")
    o, g = _pair(p)
    t = 0.0
    for _ in range(7):
        dto, dtg = o.step(t), g.step(t)
        assert dto == dtg == p.dt
        t += dto
    # a remap with a large shear displacement hits many modes
    o.remap(p.q * 3.0)
    g.remap(p.q * 3.0)
    _compare_state(o, g, t, t)
    from ql_dns_amd.ic import mode_tables
    lg, _, kg = g.get_state()
    assert np.sum(kg != mode_tables(p)[0]) > 1
    assert np.sum(np.all(lg == 0, axis=(1, 2))) > 1


FULL = ["box64x128x128", "box128x128x256"]


@pytest.mark.parametrize("name", FULL)
def test_full_size_two_steps(name):
    """64x128x128 (4032 modes) and 128x128x256 (8128 modes): two steps
    against the oracle."""
    from ql_dns_amd.params import workload
    p = workload(name)
    o, g = _pair(p, nthreads=16)
    to = o.steps(2, 0.0)
    tg = g.steps(2, 0.0)
    _compare_state(o, g, to, tg)


@pytest.mark.parametrize("name", FULL)
def test_full_size_properties_and_determinism(name):
    """The full workloads over 200 steps: size-independent properties.

    * bitwise determinism of the fixed-order reduction (two fresh runs);
    * the ky = 0 plane stays Hermitian: f(-kx, 0, -kz) = conj f(kx, 0, kz);
    * mean-field spectra stay Hermitian (real profiles), kz = 0 slot real;
    * the (kx, ky) = (0, 0) mode stays zero;
    * remapped modes are exactly zero and their kx0 moved by (NX-1) 2 pi / Lx.
    """
    from ql_dns_amd.ic import initial_state, mode_tables
    from ql_dns_amd.params import workload
    p = workload(name)
    lin, mf = initial_state(p)
    runs = []
    for _ in range(2):
        g = _ctx(p)
        g.set_state(lin, mf)
        g.linear_ops_init(0.0)
        t = g.steps(200, 0.0)
        runs.append(g.get_state() + (g.drive(),))
        g.close()
    (l1, m1, k1, d1), (l2, m2, k2, d2) = runs
    assert np.array_equal(l1, l2) and np.array_equal(m1, m2) and np.array_equal(d1, d2)
    assert np.array_equal(k1, k2)
    kx0, ky, kxi, kyi, kz, _ = mode_tables(p)
    neg = (-np.arange(p.nz)) % p.nz
    idx = {(int(a), int(b)): n for n, (a, b) in enumerate(zip(kxi, kyi))}
    scale = np.abs(l1).max()
    for n in range(p.nxy):
        if kyi[n] == 0 and kxi[n] > 0:
            partner = idx[(-int(kxi[n]), 0)]
            assert np.abs(l1[partner] - np.conj(l1[n][:, neg])).max() <= 1e-12 * scale
    assert np.all(l1[0] == 0)
    assert np.abs(m1[:, neg] - np.conj(m1)).max() <= 1e-13 * np.abs(m1).max()
    moved = k1 != kx0
    assert moved.sum() > 0
    assert np.all(np.abs((kx0 - k1)[moved] - (p.NX - 1) * 2 * np.pi / p.Lx) < 1e-9)
    assert np.all(np.isfinite(l1)) and np.all(np.isfinite(m1))


def test_create_rejects_bad_device():
    from ql_dns_amd.capi import QLContext, QLError
    p = _params()
    with pytest.raises(QLError) as e:
        QLContext(p, device=97)
    assert e.value.code == -1
