"""The hand-written adjoint of the QL step against central finite differences
of the FP64 GPU forward.

Bars: directional derivatives from the adjoint within 1e-6 (relative) of
central differences; gradients bitwise independent of the checkpoint spacing
(the recomputation reproduces the forward's bits); the harness's forward equal
to plain ql_steps.  Differences use eps = 1e-6 along directions scaled to each
input's magnitude (state: its rms; q: 1; Rm: Rm), so every input moves by a
relative 1e-6.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-6
EPS = 1e-6


def _p(**kw):
    from ql_dns_amd.params import QLParams
    return QLParams(**kw)


def _loss_parts(p, lin, mf, alpha):
    # L = 0.5/NZ^2 sum_{Bx,By}|M|^2 + alpha sum |u|^2
    from ql_dns_amd.adjoint import mf_energy
    L = mf_energy(p, mf) + alpha * float((np.abs(lin) ** 2).sum())
    g_mf = np.zeros_like(mf)
    g_mf[:2] = mf[:2] / p.NZ ** 2
    return L, 2.0 * alpha * lin, g_mf


def _run(p, lin, mf, n, noise=None):
    from ql_dns_amd.capi import QLContext
    c = QLContext(p)
    if noise:
        c.set_noise(**noise)
    c.set_state(lin, mf)
    c.linear_ops_init(0.0)
    c.steps(n, 0.0)
    out = c.get_state()
    c.close()
    return out[0], out[1]


PARAMS = ("q", "omega", "nu", "eta", "B0z")


@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("integrator,NZ", [(1, 32), (0, 32), (1, 64), (1, 128), (0, 128), (0, 256)])
def test_adjoint_matches_finite_differences(model, integrator, NZ):
    """MHD_FullQlin (model 0) and MHD_BQlin (model 1); MHD_BQlin has no B0z
    terms and rotates at 2 q / 3, so its B0z and omega derivatives are zero on
    both sides and its q derivative carries the rotation's.  MHD_BQlin's state
    directions include the kz = 0 slot of the ky = 0 modes' ux, which enters
    its dz uz as a z-constant."""
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state, mode_tables
    # remaps happen within these steps (Ly = 1); B0z != 0 exercises its terms;
    # at NZ = 128 the forward runs the wavefront kernel (the default shape there)
    p = _p(model=model, NX=8, NY=8, NZ=NZ, Ly=1.0, B0z=0.03, integrator=integrator, nu=1 / 1500, eta=1 / 1200)
    n, every = 14, 5
    lin0, mf0 = initial_state(p)
    alpha = 0.01 / float((np.abs(lin0) ** 2).sum())
    c = QLContext(p)
    c.set_state(lin0, mf0)
    c.linear_ops_init(0.0)
    c.adjoint_forward(n, every, 0.0)
    linT, mfT, kxT = c.get_state()
    assert not np.array_equal(kxT, mode_tables(p)[0]), "no remap happened"
    _, g_lin, g_mf = _loss_parts(p, linT, mfT, alpha)
    gl0, gm0, gpar = c.adjoint_backward(g_lin, g_mf)
    c.close()

    rng = np.random.default_rng(7)
    rms_l = np.sqrt(np.mean(np.abs(lin0) ** 2))
    rms_m = np.sqrt(np.mean(np.abs(mf0) ** 2))
    for d in range(4):
        vl = (rng.standard_normal(lin0.shape) + 1j * rng.standard_normal(lin0.shape)) * rms_l
        if model == 0:
            vl[::p.NY // 2, 0::2, 0] = 0  # directions within divergence-free states (ux, bx at ky = kz = 0)
        vm = (rng.standard_normal(mf0.shape) + 1j * rng.standard_normal(mf0.shape)) * rms_m
        ad = float(np.sum(gl0.real * vl.real + gl0.imag * vl.imag) + np.sum(gm0.real * vm.real + gm0.imag * vm.imag))
        Lp = _loss_parts(p, *_run(p, lin0 + EPS * vl, mf0 + EPS * vm, n), alpha)[0]
        Lm = _loss_parts(p, *_run(p, lin0 - EPS * vl, mf0 - EPS * vm, n), alpha)[0]
        fd = (Lp - Lm) / (2 * EPS)
        assert abs(ad - fd) <= TOL * abs(fd), f"state direction {d}: adjoint {ad!r} fd {fd!r}"
    for k, name in enumerate(PARAMS):
        base = getattr(p, name)
        h = EPS * abs(base)
        Lp = _loss_parts(p, *_run(p.replace(**{name: base + h}), lin0, mf0, n), alpha)[0]
        Lm = _loss_parts(p, *_run(p.replace(**{name: base - h}), lin0, mf0, n), alpha)[0]
        fd = (Lp - Lm) / (2 * h)
        if model == 1 and name in ("omega", "B0z"):
            # MHD_BQlin has no B0z terms and rotates at 2 q / 3 whatever omega is
            assert fd == 0.0 and gpar[k] == 0.0
            continue
        assert abs(gpar[k] - fd) <= TOL * abs(fd), f"d/d{name}: adjoint {gpar[k]!r} fd {fd!r}"


@pytest.mark.parametrize("model", [0, 1])
def test_adjoint_independent_of_checkpoint_spacing(model):
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    p = _p(model=model, NX=8, NY=6, NZ=64, Ly=1.0)
    n = 9
    lin0, mf0 = initial_state(p)
    res = []
    for every in (1, 4, 9, 20):
        c = QLContext(p)
        c.set_state(lin0, mf0)
        c.linear_ops_init(0.0)
        c.adjoint_forward(n, every, 0.0)
        linT, mfT, _ = c.get_state()
        _, g_lin, g_mf = _loss_parts(p, linT, mfT, 1e-9)
        res.append(c.adjoint_backward(g_lin, g_mf))
        # the context is left at the forward's final state
        l2, m2, _ = c.get_state()
        assert np.array_equal(l2, linT) and np.array_equal(m2, mfT)
        c.close()
    for r in res[1:]:
        for x, y in zip(res[0], r):
            assert np.array_equal(x, y)


def test_adjoint_rejects_unsupported():
    from ql_dns_amd.capi import QLContext, QLError
    for kw in (dict(precision=1),):
        c = QLContext(_p(NX=4, NY=4, NZ=32, **kw))
        with pytest.raises(QLError):
            c.adjoint_forward(2, 1, 0.0)
        c.close()
    c = QLContext(_p(NX=4, NY=4, NZ=32))
    with pytest.raises(QLError):
        c.adjoint_backward(None, None)  # no forward yet
    with pytest.raises(QLError):
        c.adjoint_forward(2, 0, 0.0)
    # the backward recomputes the forward: the noise may not change in between
    c.set_noise(5.0)
    c.adjoint_forward(2, 1, 0.0)
    c.set_noise(0.0)
    with pytest.raises(QLError):
        c.adjoint_backward(None, None)
    c.close()


def test_sensitivity_box_harness_matches_finite_differences():
    """32x32x64, Rm = 4000, 200 steps, checkpoint every 20; loss = final
    0.5 <Bx^2 + By^2>_z; gradients w.r.t. the initial Bx(z), By(z), q and Rm
    through the torch.autograd.Function harness, checked on 16 seeded
    directions."""
    import torch
    from ql_dns_amd.adjoint import final_mf_energy, kept_index, mf_energy
    from ql_dns_amd.ic import initial_state, mf_profile_real
    from ql_dns_amd.params import workload
    p = workload("box32x32x64")
    n, every = 200, 20
    lin0, mf0 = initial_state(p)
    prof = mf_profile_real(p, mf0)[:2]  # real profiles whose stored spectrum is mf0
    bx = torch.tensor(prof[0], dtype=torch.float64, requires_grad=True)
    by = torch.tensor(prof[1], dtype=torch.float64, requires_grad=True)
    q = torch.tensor(p.q, dtype=torch.float64, requires_grad=True)
    Rm = torch.tensor(1.0 / p.eta, dtype=torch.float64, requires_grad=True)
    L = final_mf_energy(bx, by, q, Rm, p, lin0, n, every)
    L.backward()
    # the harness's forward is the plain forward from the same coefficients
    kept = torch.from_numpy(kept_index(p))
    spec = torch.fft.fft(torch.stack([bx, by]).detach().to(torch.complex128), dim=-1)[:, kept].numpy()
    assert np.abs(spec - mf0[:2]).max() < 1e-12 * np.abs(mf0).max()
    mf_h = np.concatenate([spec, np.zeros((p.nMF - 2, p.nz), dtype=np.complex128)])
    _, mfT = _run(p, lin0, mf_h, n)
    assert float(L.detach()) == mf_energy(p, mfT)

    def loss(bx_, by_, q_, Rm_):
        with torch.no_grad():
            return float(final_mf_energy(bx_, by_, torch.tensor(q_, dtype=torch.float64),
                                         torch.tensor(Rm_, dtype=torch.float64), p, lin0, n, every))

    rng = np.random.default_rng(16)
    scale_prof = float(np.sqrt(np.mean(prof ** 2)))
    for d in range(16):
        vx = rng.standard_normal(p.NZ) * scale_prof
        vy = rng.standard_normal(p.NZ) * scale_prof
        vq = rng.standard_normal() * 1.0
        vR = rng.standard_normal() * float(Rm)
        ad = float(bx.grad @ torch.from_numpy(vx) + by.grad @ torch.from_numpy(vy)) + float(q.grad) * vq + \
            float(Rm.grad) * vR
        b0, b1 = bx.detach(), by.detach()
        Lp = loss(b0 + EPS * torch.from_numpy(vx), b1 + EPS * torch.from_numpy(vy), p.q + EPS * vq,
                  float(Rm) + EPS * vR)
        Lm = loss(b0 - EPS * torch.from_numpy(vx), b1 - EPS * torch.from_numpy(vy), p.q - EPS * vq,
                  float(Rm) - EPS * vR)
        fd = (Lp - Lm) / (2 * EPS)
        assert abs(ad - fd) <= TOL * abs(fd), f"direction {d}: adjoint {ad!r} fd {fd!r}"


@pytest.mark.parametrize("model,integrator", [(0, 1), (0, 0), (1, 1), (2, 1)])
def test_adjoint_with_driving_noise_matches_finite_differences(model, integrator):
    """Driven runs (DrivingNoise after every step, then the remap): the noise
    is a function of t, the mode and the slot only, so the same inputs redraw
    the same noise and finite differences apply.  Its ux and bx multipliers
    depend on q through kx(t); the noise is large here (f_noise 40) so that a
    missing q term would show well above the 1e-6 bar, and the test also
    checks that the noise changes dL/dq.  HD_fullU (model 2): the loss reads
    its mean flows (Ux, Uy), and dL/deta is zero on both sides."""
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state, mode_tables
    p = _p(model=model, NX=8, NY=8, NZ=32, Ly=1.0, B0z=0.0 if model == 2 else 0.03, integrator=integrator,
           nu=1 / 1500, eta=1 / 1200)
    noise = dict(f_noise=40.0, seed=0x5EED0011)
    n, every = 14, 5
    lin0, mf0 = initial_state(p)
    alpha = 0.01 / float((np.abs(lin0) ** 2).sum())
    grads = {}
    for nz in (None, noise):
        c = QLContext(p)
        if nz:
            c.set_noise(**nz)
        c.set_state(lin0, mf0)
        c.linear_ops_init(0.0)
        c.adjoint_forward(n, every, 0.0)
        linT, mfT, kxT = c.get_state()
        assert not np.array_equal(kxT, mode_tables(p)[0]), "no remap happened"
        _, g_lin, g_mf = _loss_parts(p, linT, mfT, alpha)
        grads[bool(nz)] = c.adjoint_backward(g_lin, g_mf)
        c.close()
    gl0, gm0, gpar = grads[True]
    assert abs(gpar[0] - grads[False][2][0]) > 1e-3 * abs(gpar[0])  # the noise matters to dL/dq
    rng = np.random.default_rng(11)
    rms_l = np.sqrt(np.mean(np.abs(lin0) ** 2))
    rms_m = np.sqrt(np.mean(np.abs(mf0) ** 2))
    for d in range(2):
        vl = (rng.standard_normal(lin0.shape) + 1j * rng.standard_normal(lin0.shape)) * rms_l
        if model in (0, 2):
            vl[::p.NY // 2, 0::2, 0] = 0
        vm = (rng.standard_normal(mf0.shape) + 1j * rng.standard_normal(mf0.shape)) * rms_m
        ad = float(np.sum(gl0.real * vl.real + gl0.imag * vl.imag) + np.sum(gm0.real * vm.real + gm0.imag * vm.imag))
        Lp = _loss_parts(p, *_run(p, lin0 + EPS * vl, mf0 + EPS * vm, n, noise), alpha)[0]
        Lm = _loss_parts(p, *_run(p, lin0 - EPS * vl, mf0 - EPS * vm, n, noise), alpha)[0]
        fd = (Lp - Lm) / (2 * EPS)
        assert abs(ad - fd) <= TOL * abs(fd), f"state direction {d}: adjoint {ad!r} fd {fd!r}"
    for k, name in enumerate(PARAMS[:4]):
        if model == 1 and name == "omega":
            continue
        base = getattr(p, name)
        h = EPS * abs(base)
        Lp = _loss_parts(p, *_run(p.replace(**{name: base + h}), lin0, mf0, n, noise), alpha)[0]
        Lm = _loss_parts(p, *_run(p.replace(**{name: base - h}), lin0, mf0, n, noise), alpha)[0]
        fd = (Lp - Lm) / (2 * h)
        if model == 2 and name == "eta":
            assert fd == 0.0 and gpar[k] == 0.0
            continue
        assert abs(gpar[k] - fd) <= TOL * abs(fd), f"d/d{name}: adjoint {gpar[k]!r} fd {fd!r}"


@pytest.mark.parametrize("integrator,NZ", [(1, 32), (0, 32), (1, 128)])
def test_hd_adjoint_matches_finite_differences(integrator, NZ):
    """HD_fullU (two fluctuation fields, mean flows Ux, Uy in its own layout):
    the MHD_FullQlin reverse kernels with b = B = 0; dL/deta and dL/dB0z are
    zero.  Loss: 0.5/NZ^2 sum |Ux|^2 + |Uy|^2 + alpha sum |u|^2."""
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state, mode_tables
    p = _p(model=2, NX=8, NY=8, NZ=NZ, Ly=1.0, integrator=integrator, nu=1 / 1500)
    n, every = 14, 5
    lin0, mf0 = initial_state(p)
    assert lin0.shape[1] == 2 and mf0.shape[0] == 2
    alpha = 0.01 / float((np.abs(lin0) ** 2).sum())

    def loss(lin, mf):
        return 0.5 * float((np.abs(mf) ** 2).sum()) / p.NZ ** 2 + alpha * float((np.abs(lin) ** 2).sum())

    c = QLContext(p)
    c.set_state(lin0, mf0)
    c.linear_ops_init(0.0)
    c.adjoint_forward(n, every, 0.0)
    linT, mfT, kxT = c.get_state()
    assert not np.array_equal(kxT, mode_tables(p)[0]), "no remap happened"
    gl0, gm0, gpar = c.adjoint_backward(2.0 * alpha * linT, mfT / p.NZ ** 2)
    c.close()
    assert gl0.shape == lin0.shape and gm0.shape == mf0.shape
    assert gpar[3] == 0.0 and gpar[4] == 0.0
    rng = np.random.default_rng(5)
    rms_l = np.sqrt(np.mean(np.abs(lin0) ** 2))
    rms_m = np.sqrt(np.mean(np.abs(mf0) ** 2))
    for d in range(3):
        vl = (rng.standard_normal(lin0.shape) + 1j * rng.standard_normal(lin0.shape)) * rms_l
        vl[::p.NY // 2, 0, 0] = 0  # ux at ky = kz = 0: states stay divergence free
        vm = (rng.standard_normal(mf0.shape) + 1j * rng.standard_normal(mf0.shape)) * rms_m
        ad = float(np.sum(gl0.real * vl.real + gl0.imag * vl.imag) + np.sum(gm0.real * vm.real + gm0.imag * vm.imag))
        fd = (loss(*_run(p, lin0 + EPS * vl, mf0 + EPS * vm, n)) - loss(*_run(p, lin0 - EPS * vl, mf0 - EPS * vm, n))) / (2 * EPS)
        assert abs(ad - fd) <= TOL * abs(fd), f"state direction {d}: adjoint {ad!r} fd {fd!r}"
    for k, name in enumerate(("q", "omega", "nu")):
        base = getattr(p, name)
        h = EPS * abs(base)
        fd = (loss(*_run(p.replace(**{name: base + h}), lin0, mf0, n)) -
              loss(*_run(p.replace(**{name: base - h}), lin0, mf0, n))) / (2 * h)
        assert abs(gpar[k] - fd) <= TOL * abs(fd), f"d/d{name}: adjoint {gpar[k]!r} fd {fd!r}"



@pytest.mark.parametrize("model,integrator", [(0, 1), (0, 0), (1, 1), (2, 1)])
def test_variable_steps_at_the_cap_equal_fixed_steps(model, integrator):
    """Variable steps (dt < 0) whose CFL step always exceeds the cap |dt| take
    the cap every step: forward and adjoint then equal those of fixed steps
    of that size (the step sizes come from the device clock, the RK3CN old
    operator from the step before, the checkpoints restore the profile
    maxima the next step's size comes from)."""
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    p = _p(model=model, NX=8, NY=8, NZ=32, Ly=1.0, integrator=integrator, nu=1 / 1500, eta=1 / 1200, dt=0.004)
    n, every = 14, 5
    lin0, mf0 = initial_state(p)
    out = []
    for q in (p, p.replace(dt=-0.004, CFL=1e3)):
        c = QLContext(q)
        c.set_state(lin0, mf0)
        c.linear_ops_init(0.0)
        t = c.adjoint_forward(n, every, 0.0)
        linT, mfT, kxT = c.get_state()
        g_lin = 2e-3 * linT
        g_mf = mfT / p.NZ ** 2
        out.append((t, linT, mfT, kxT, c.adjoint_backward(g_lin, g_mf)))
        c.close()
    (ta, la, ma, ka, ga), (tb, lb, mb, kb, gb) = out
    assert ta == tb and np.array_equal(ka, kb)
    assert rel_err(lb, la) < 1e-13 and rel_err(mb, ma) < 1e-13
    for x, y in zip(gb, ga):
        assert rel_err(np.asarray(x), np.asarray(y)) < 1e-12


def _frozen_run(p, lin, mf, dts):
    """The forward with its step sizes held fixed: step k is one fixed step of
    size dts[k] (ql_step), then the remap at q (t + dt), as main.cpp:111-123."""
    from ql_dns_amd.capi import QLContext
    t, kx0 = 0.0, None
    for dt in dts:
        c = QLContext(p.replace(dt=float(dt)))
        c.set_state(lin, mf, kx0)
        c.linear_ops_init(t)
        assert c.step(t) == dt
        t = t + dt
        c.remap(p.q * t)
        lin, mf, kx0 = c.get_state()
        c.close()
    return lin, mf


def test_variable_steps_adjoint_matches_frozen_step_differences():
    """Varying CFL steps (RK3CN_Elin): the adjoint differentiates the forward
    with its step sizes frozen at the ones the forward chose (the step size's
    own dependence on the state through the CFL condition is not
    differentiated); central differences of that frozen forward agree within
    the 1e-6 bar, for state and parameter directions."""
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    p = _p(NX=8, NY=8, NZ=32, Ly=1.0, B0z=0.03, nu=1 / 1500, eta=1 / 1200, dt=-0.05, CFL=1.0)
    n, every = 14, 5
    lin0, mf0 = initial_state(p)
    alpha = 0.01 / float((np.abs(lin0) ** 2).sum())
    # the step sizes the device chooses (ql_step equals ql_steps step by step)
    c = QLContext(p)
    c.set_state(lin0, mf0)
    c.linear_ops_init(0.0)
    t, dts = 0.0, []
    for _ in range(n):
        dt = c.step(t)
        dts.append(dt)
        t = t + dt
        c.remap(p.q * t)
    lin_s, mf_s, _ = c.get_state()
    c.close()
    assert len(set(dts)) > 5, "the steps do not vary"
    c = QLContext(p)
    c.set_state(lin0, mf0)
    c.linear_ops_init(0.0)
    assert c.adjoint_forward(n, every, 0.0) == t
    linT, mfT, _ = c.get_state()
    assert np.array_equal(linT, lin_s) and np.array_equal(mfT, mf_s)
    _, g_lin, g_mf = _loss_parts(p, linT, mfT, alpha)
    gl0, gm0, gpar = c.adjoint_backward(g_lin, g_mf)
    c.close()
    lf, mff = _frozen_run(p, lin0, mf0, dts)
    assert rel_err(lf, linT) < 1e-13 and rel_err(mff, mfT) < 1e-13

    rng = np.random.default_rng(23)
    rms_l = np.sqrt(np.mean(np.abs(lin0) ** 2))
    rms_m = np.sqrt(np.mean(np.abs(mf0) ** 2))
    for d in range(2):
        vl = (rng.standard_normal(lin0.shape) + 1j * rng.standard_normal(lin0.shape)) * rms_l
        vl[::p.NY // 2, 0::2, 0] = 0
        vm = (rng.standard_normal(mf0.shape) + 1j * rng.standard_normal(mf0.shape)) * rms_m
        ad = float(np.sum(gl0.real * vl.real + gl0.imag * vl.imag) + np.sum(gm0.real * vm.real + gm0.imag * vm.imag))
        Lp = _loss_parts(p, *_frozen_run(p, lin0 + EPS * vl, mf0 + EPS * vm, dts), alpha)[0]
        Lm = _loss_parts(p, *_frozen_run(p, lin0 - EPS * vl, mf0 - EPS * vm, dts), alpha)[0]
        fd = (Lp - Lm) / (2 * EPS)
        assert abs(ad - fd) <= TOL * abs(fd), f"state direction {d}: adjoint {ad!r} fd {fd!r}"
    for k, name in enumerate(PARAMS):
        base = getattr(p, name)
        h = EPS * abs(base)
        Lp = _loss_parts(p, *_frozen_run(p.replace(**{name: base + h}), lin0, mf0, dts), alpha)[0]
        Lm = _loss_parts(p, *_frozen_run(p.replace(**{name: base - h}), lin0, mf0, dts), alpha)[0]
        fd = (Lp - Lm) / (2 * h)
        assert abs(gpar[k] - fd) <= TOL * abs(fd), f"d/d{name}: adjoint {gpar[k]!r} fd {fd!r}"
