"""A seeded sweep of configurations against the oracle: model, integrator,
box, physics, fixed or variable steps, driving noise and the library's own
choices (launch shape, z transform, substep tail) drawn together, so that
combinations no targeted test names are checked too.

Bar as for the other parity tests: states, mean fields, drives and
diagnostics within 1e-10 relative of the oracle after the steps, kx0 after
remaps exactly equal, the step times exactly equal with fixed steps and to
1e-12 with variable (CFL) steps.  A choice the library refuses for a
configuration (QLError from set_config / set_transform / set_tail) is
skipped, the rest of the case still runs.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _draw(seed):
    rng = np.random.default_rng(1000 + seed)
    model = int(rng.integers(0, 3))
    NZ = int(rng.choice([32, 64, 128, 256]))
    NX = int(rng.choice([4, 6, 8, 10, 16]))
    NY = int(rng.choice([4, 6, 8, 16]))
    kw = dict(model=model, integrator=int(rng.integers(0, 2)), NX=NX, NY=NY, NZ=NZ,
              Lx=float(rng.choice([1.0, 2.0])), Ly=float(rng.choice([1.0, 4.0])),
              q=float(rng.uniform(1.0, 1.8)), viscosity_order=int(rng.choice([2, 2, 4])),
              nu=float(1.0 / rng.uniform(500, 3000)), eta=float(1.0 / rng.uniform(500, 3000)))
    if model == 0:
        kw["B0z"] = float(rng.choice([0.0, 0.03]))
    vdt = rng.random() < 0.3
    if vdt:
        kw.update(dt=-0.02, CFL=0.5)
    noise = float(rng.choice([0.0, 0.0, 2.0]))
    if vdt:
        # a variable step agrees with the oracle's to rounding only, and the
        # noise is keyed on the bits of the step's start time: driven variable
        # steps are compared with the oracle replaying the steps instead
        # (test_gpu_noise.py::test_variable_steps_with_noise_match_oracle)
        noise = 0.0
    line = int(rng.choice([64, 128, 256]))
    target = int(rng.choice([256, 512, 1024]))
    transform = int(rng.integers(0, 2))
    tail = int(rng.integers(0, 2))
    return kw, noise, (line, target), transform, tail


@pytest.mark.parametrize("seed", range(120))
def test_sweep_matches_oracle(seed):
    from ql_dns_amd.capi import QLContext, QLError
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import QLParams
    from oracle.oracle import Oracle
    kw, noise, shape, transform, tail = _draw(seed)
    p = QLParams(**kw)
    print(f"seed {seed}: {kw} noise {noise} shape {shape} transform {transform} tail {tail}", flush=True)
    lin, mf = initial_state(p, seed=seed)
    o = Oracle.from_params(p, nthreads=8)
    g = QLContext(p)
    for call in (lambda: g.set_config(p.NZ, *shape), lambda: g.set_transform(p.NZ, transform),
                 lambda: g.set_tail(p.NZ, tail)):
        try:
            call()
        except QLError:
            pass
    for x in (o, g):
        if noise:
            x.set_noise(noise, seed=77 + seed)
        x.set_state(lin, mf)
        x.linear_ops_init(0.0)
    to, tg = o.steps(8, 0.0), g.steps(8, 0.0)
    if p.dt > 0:
        assert tg == to
    else:
        assert abs(tg - to) <= 1e-12 * to  # CFL steps agree to rounding
    lo, mo, ko = o.get_state()
    lg, mg, kg = g.get_state()
    assert np.array_equal(kg, ko)
    assert rel_err(mg, mo) < TOL and rel_err(lg, lo) < TOL
    assert rel_err(g.drive(), o.drive()) < TOL
    do, dg = o.diagnostics(to), g.diagnostics(tg)
    fin = np.isfinite(do)
    assert np.array_equal(fin, np.isfinite(dg))
    assert np.all(np.abs(dg[fin] - do[fin]) <= TOL * np.maximum(np.abs(do[fin]), 1e-300) + 1e-300)
    g.close()


@pytest.mark.parametrize("seed", range(24))
def test_run_set_sweep_matches_oracle(seed):
    """Seeded run sets: 2 to 7 runs of one model, integrator and step, each with
    its own box, NZ and physics, the library's choices drawn per NZ group;
    every run against its own oracle."""
    from ql_dns_amd.capi import QLBatch, QLError
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import QLParams
    from oracle.oracle import Oracle
    rng = np.random.default_rng(5000 + seed)
    model, integrator = int(rng.integers(0, 3)), int(rng.integers(0, 2))
    runs = []
    for _ in range(int(rng.integers(2, 8))):
        kw = dict(model=model, integrator=integrator, NX=int(rng.choice([4, 6, 8, 16])),
                  NY=int(rng.choice([4, 6, 8, 16])), NZ=int(rng.choice([32, 64, 128])),
                  Lx=float(rng.choice([1.0, 2.0])), Ly=float(rng.choice([1.0, 4.0])),
                  q=float(rng.uniform(1.0, 1.8)), viscosity_order=int(rng.choice([2, 2, 4])),
                  nu=float(1.0 / rng.uniform(500, 3000)), eta=float(1.0 / rng.uniform(500, 3000)))
        if model == 0:
            kw["B0z"] = float(rng.choice([0.0, 0.03]))
        runs.append(QLParams(**kw))
    print(f"seed {seed}: model {model} integrator {integrator} runs "
          f"{[(p.NX, p.NY, p.NZ) for p in runs]}", flush=True)
    st = [initial_state(p, seed=100 * seed + r) for r, p in enumerate(runs)]
    b = QLBatch(runs)
    for NZ in sorted({p.NZ for p in runs}):
        for call in (lambda: b.set_config(NZ, int(rng.choice([64, 128, 256])), int(rng.choice([256, 512, 1024]))),
                     lambda: b.set_transform(NZ, int(rng.integers(0, 2))),
                     lambda: b.set_tail(NZ, int(rng.integers(0, 2)))):
            try:
                call()
            except QLError:
                pass
    for r, (lin, mf) in enumerate(st):
        b.set_state(r, lin, mf)
    b.linear_ops_init(0.0)
    tb = b.steps(7, 0.0)
    for r, (p, (lin, mf)) in enumerate(zip(runs, st)):
        o = Oracle.from_params(p, nthreads=8)
        o.set_state(lin, mf)
        o.linear_ops_init(0.0)
        assert o.steps(7, 0.0) == tb
        (lo, mo, ko), (lg, mg, kg) = o.get_state(), b.get_state(r)
        assert np.array_equal(kg, ko), r
        assert rel_err(mg, mo) < TOL and rel_err(lg, lo) < TOL, r
        assert rel_err(b.drive(r), o.drive()) < TOL, r
    b.close()


@pytest.mark.parametrize("seed", range(24))
def test_mixed_sweep_close_to_fp64(seed):
    """Seeded mixed-precision draws (model, integrator, box, NZ, shape, the
    FP64 split on or off) against the FP64 run of the same configuration
    after 20 steps: mean fields, drives and states within 1e-6, kx0 and times
    equal."""
    from ql_dns_amd.capi import QLContext, QLError
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import PREC_MIXED, QLParams
    rng = np.random.default_rng(9000 + seed)
    model = int(rng.integers(0, 3))
    kw = dict(model=model, integrator=int(rng.integers(0, 2)), NX=int(rng.choice([4, 8, 16])),
              NY=int(rng.choice([4, 8, 16])), NZ=int(rng.choice([32, 64, 128, 256])),
              Ly=float(rng.choice([1.0, 4.0])), q=float(rng.uniform(1.0, 1.8)), dt=0.005,
              nu=1 / 4000, eta=1 / 4000)
    if model == 0:
        kw["B0z"] = float(rng.choice([0.0, 0.03]))
    p = QLParams(**kw)
    shape = (int(rng.choice([64, 128, 256])), int(rng.choice([256, 512, 1024])))
    tau = float(rng.choice([0.03, 0.1, 2.0]))
    print(f"seed {seed}: {kw} shape {shape} tau {tau}", flush=True)
    lin, mf = initial_state(p, seed=seed)
    out = []
    for q in (p, p.replace(precision=PREC_MIXED)):
        c = QLContext(q)
        if q.precision == PREC_MIXED:
            try:
                c.set_config(p.NZ, *shape)
            except QLError:
                pass
            c.mixed_split(tau)
        c.set_state(lin, mf)
        c.linear_ops_init(0.0)
        t = c.steps(20, 0.0)
        out.append((t,) + tuple(c.get_state()) + (c.drive(),))
        c.close()
    (ta, la, ma, ka, da), (tb, lb, mb, kb, db) = out
    assert ta == tb and np.array_equal(ka, kb)
    assert rel_err(mb, ma) < 1e-6 and rel_err(db, da) < 1e-6 and rel_err(lb, la) < 1e-6


@pytest.mark.parametrize("seed", range(12))
def test_adjoint_sweep_matches_finite_differences(seed):
    """Seeded adjoint draws (model, integrator, NZ, box, noise, checkpoint
    spacing): directional derivatives of L = 0.5/NZ^2 |mf|^2 + alpha |u|^2
    along two state directions and dL/dq, dL/dnu within 1e-6 of central
    differences of the GPU forward."""
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import QLParams
    rng = np.random.default_rng(7000 + seed)
    model = int(rng.integers(0, 3))
    kw = dict(model=model, integrator=int(rng.integers(0, 2)), NX=int(rng.choice([4, 6, 8])),
              NY=int(rng.choice([4, 6, 8])), NZ=int(rng.choice([32, 64, 128])), Ly=float(rng.choice([1.0, 4.0])),
              q=float(rng.uniform(1.0, 1.8)), nu=1 / 1500, eta=1 / 1200)
    if model == 0:
        kw["B0z"] = float(rng.choice([0.0, 0.03]))
    p = QLParams(**kw)
    noise = dict(f_noise=5.0, seed=300 + seed) if rng.random() < 0.4 else None
    n, every = 10, int(rng.choice([1, 3, 5, 10]))
    print(f"seed {seed}: {kw} noise {noise} every {every}", flush=True)
    lin0, mf0 = initial_state(p, seed=seed)
    alpha = 0.01 / float((np.abs(lin0) ** 2).sum())

    def run(q, lin, mf):
        c = QLContext(q)
        if noise:
            c.set_noise(**noise)
        c.set_state(lin, mf)
        c.linear_ops_init(0.0)
        c.steps(n, 0.0)
        lT, mT, _ = c.get_state()
        c.close()
        return lT, mT

    def loss(lin, mf):
        return 0.5 * float((np.abs(mf) ** 2).sum()) / p.NZ ** 2 + alpha * float((np.abs(lin) ** 2).sum())

    c = QLContext(p)
    if noise:
        c.set_noise(**noise)
    c.set_state(lin0, mf0)
    c.linear_ops_init(0.0)
    c.adjoint_forward(n, every, 0.0)
    lT, mT, _ = c.get_state()
    gl0, gm0, gpar = c.adjoint_backward(2.0 * alpha * lT, mT / p.NZ ** 2)
    c.close()
    eps = 1e-6
    rms_l = np.sqrt(np.mean(np.abs(lin0) ** 2))
    rms_m = np.sqrt(np.mean(np.abs(mf0) ** 2))
    for d in range(2):
        vl = (rng.standard_normal(lin0.shape) + 1j * rng.standard_normal(lin0.shape)) * rms_l
        if model != 1:
            vl[::p.NY // 2, 0::2, 0] = 0  # divergence-free directions (ux, bx at ky = kz = 0)
        vm = (rng.standard_normal(mf0.shape) + 1j * rng.standard_normal(mf0.shape)) * rms_m
        ad = float(np.sum(gl0.real * vl.real + gl0.imag * vl.imag) + np.sum(gm0.real * vm.real + gm0.imag * vm.imag))
        fd = (loss(*run(p, lin0 + eps * vl, mf0 + eps * vm)) - loss(*run(p, lin0 - eps * vl, mf0 - eps * vm))) / (2 * eps)
        assert abs(ad - fd) <= 1e-6 * abs(fd), f"state direction {d}: adjoint {ad!r} fd {fd!r}"
    for k, name in ((0, "q"), (2, "nu")):
        base = getattr(p, name)
        h = eps * abs(base)
        fd = (loss(*run(p.replace(**{name: base + h}), lin0, mf0)) -
              loss(*run(p.replace(**{name: base - h}), lin0, mf0))) / (2 * h)
        assert abs(gpar[k] - fd) <= 1e-6 * abs(fd), f"d/d{name}: adjoint {gpar[k]!r} fd {fd!r}"
