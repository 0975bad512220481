"""Pins of the CPU oracle (oracle/ql_oracle.c) against checks that do not come
from the reference's code: numpy and brute-force transforms, closed-form
integrator factors, epicyclic / shearing-wave solutions, a Cartesian statement
of the QL equations, and physical-space averages of 3-D fields.  A plausible
slip in the restatement (a dropped term, a sign, a transposed index, a wrong
RK coefficient or substep time) fails at least one of these.
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

import physics_ref as ph
from conftest import rel_err
from oracle.oracle import (MODEL_CONSTANT_DAMPING, MODEL_LINEAR_ONLY, Oracle, fft_forward,
                           fft_inverse, fft_plain)

A = (8 / 15, 5 / 12, 3 / 4)
B = (0.0, -17 / 60, -5 / 12)
P = (0.0, 8 / 15, 2 / 3)


# ----------------------------------------------------------------------------
# transforms (Auxiliary/fftwPlans.h:33-97)
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("N", [4, 8, 32, 64, 128, 256])
def test_fft_matches_numpy(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    assert np.abs(fft_plain(x, -1) - np.fft.fft(x)).max() < 1e-13 * N
    assert np.abs(fft_plain(x, +1) - N * np.fft.ifft(x)).max() < 1e-13 * N


def test_fft_matches_bruteforce_dft():
    N = 32
    rng = np.random.default_rng(1)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    j = np.arange(N)
    W = np.exp(-2j * np.pi * np.outer(j, j) / N)
    assert np.abs(fft_plain(x, -1) - W @ x).max() < 1e-12
    assert np.abs(fft_plain(x, +1) - W.conj() @ x).max() < 1e-12


@pytest.mark.parametrize("NZ", [32, 64, 128])
def test_dealiased_transforms(NZ):
    nz = 1 + 2 * int(NZ / 3.0)
    rng = np.random.default_rng(NZ)
    c = rng.standard_normal(nz) + 1j * rng.standard_normal(nz)
    k = ph.kz_slots(nz)
    zj = np.arange(NZ)
    # inverse: sum_m c_m exp(+2 pi i k_m j / NZ), unnormalised
    expect = np.exp(2j * np.pi * np.outer(zj, k) / NZ) @ c
    got = fft_inverse(NZ, c)
    assert np.abs(got - expect).max() < 1e-12 * np.abs(expect).max()
    # forward keeps exactly the retained band
    y = rng.standard_normal(NZ) + 1j * rng.standard_normal(NZ)
    expect_f = np.exp(-2j * np.pi * np.outer(k, zj) / NZ) @ y
    assert np.abs(fft_forward(nz, y) - expect_f).max() < 1e-12 * np.abs(expect_f).max()
    # round trip is the identity on the band
    assert np.abs(fft_forward(nz, fft_inverse(NZ, c)) / NZ - c).max() < 1e-14 * NZ
    # 2/3 rule: the retained band of a grid product is the exact convolution
    d = rng.standard_normal(nz) + 1j * rng.standard_normal(nz)
    prod = fft_forward(nz, fft_inverse(NZ, c) * fft_inverse(NZ, d)) / NZ
    conv = np.zeros(nz, dtype=complex)
    for a, ka in enumerate(k):
        for b, kb in enumerate(k):
            s = ka + kb
            hit = np.nonzero(k == s)[0]
            if hit.size:
                conv[hit[0]] += c[a] * d[b]
    assert np.abs(prod - conv).max() < 1e-12 * np.abs(conv).max()


# ----------------------------------------------------------------------------
# wavenumber tables (Auxiliary/Kdata.cpp:28-80), written out by hand
# ----------------------------------------------------------------------------
def test_tables_small_box():
    o = Oracle(NX=4, NY=4, NZ=8, L=(1.0, 2.0, 0.5))
    kx0, ky, kxi, kyi, kz = o.tables()
    assert (o.nxy, o.nz, o.half) == (6, 5, 3)
    assert list(kxi) == [0, 0, 1, 1, -1, -1]
    assert list(kyi) == [0, 1, 0, 1, 0, 1]
    assert np.allclose(kx0, 2 * np.pi * np.array([0, 0, 1, 1, -1, -1]), rtol=0, atol=1e-15)
    assert np.allclose(ky, np.pi * np.array([0, 1, 0, 1, 0, 1]), rtol=0, atol=1e-15)
    assert np.allclose(kz, 4 * np.pi * np.array([0, 1, 2, -2, -1]), rtol=0, atol=1e-14)


def test_tables_match_generator():
    from ql_dns_amd.ic import mode_tables
    from ql_dns_amd.params import workload
    p = workload("box16x16x32")
    o = Oracle.from_params(p)
    kx0, ky, kxi, kyi, kz = o.tables()
    gx0, gky, gxi, gyi, gkz, _ = mode_tables(p)
    assert np.array_equal(kxi, gxi) and np.array_equal(kyi, gyi)
    assert np.abs(kx0 - gx0).max() < 1e-12 and np.abs(ky - gky).max() < 1e-12
    assert np.abs(kz - gkz).max() < 1e-12


# ----------------------------------------------------------------------------
# integrators (Integrators/RK3CN.cpp, RK3CN_Elin.cpp)
# ----------------------------------------------------------------------------
def _random_state(o, seed=3):
    rng = np.random.default_rng(seed)
    lin = rng.standard_normal((o.nxy, 4, o.nz)) + 1j * rng.standard_normal((o.nxy, 4, o.nz))
    mf = rng.standard_normal((o.nMF, o.nz)) + 1j * rng.standard_normal((o.nMF, o.nz))
    return lin, mf


@pytest.mark.parametrize("integrator", [0, 1])
def test_rk3_stability_polynomial(integrator):
    """du/dt = -u/2 (ConstantDamping, L = 0): one step multiplies by
    1 + z + z^2/2 + z^3/6 with z = -dt/2, exactly for a third-order 3-stage RK."""
    dt = 0.37
    o = Oracle(model=MODEL_CONSTANT_DAMPING, integrator=integrator, NX=4, NY=4, NZ=8, dt=dt)
    lin, mf = _random_state(o)
    o.set_state(lin, mf)
    o.linear_ops_init(0.0)
    o.step(0.0)
    z = -0.5 * dt
    R = 1 + z + z * z / 2 + z ** 3 / 6
    got, gmf, _ = o.get_state()
    assert np.abs(got - R * lin).max() < 1e-14 * np.abs(lin).max()
    assert np.abs(gmf - mf).max() == 0  # mean-field RHS and operator are zero


def _linear_only(integrator, dt, q=1.5, nu=0.02, vo=2):
    o = Oracle(model=MODEL_LINEAR_ONLY, integrator=integrator, NX=8, NY=8, NZ=32, dt=dt, q=q,
               nu=nu, eta=2 * nu, viscosity_order=vo, remap=0, L=(1.0, 2.0, 1.0))
    lin, mf = _random_state(o)
    o.set_state(lin, np.zeros_like(mf))
    o.linear_ops_init(0.0)
    return o, lin


def _Lphys(o, kx0, ky, kz, t, coef, vo=2, q=1.5):
    """-coef |k(t)|^vo with k(t) = (kx0 + q t ky, ky, kz)"""
    kx = kx0 + q * t * ky
    k2 = kx[:, None] ** 2 + ky[:, None] ** 2 + kz[None, :] ** 2
    return -coef * k2 ** (vo // 2)


def test_elin_linear_factor():
    """RK3CN_Elin with no RHS: each substep multiplies by exp(dt (a+b) L(t_end)),
    t_end = (8/15, 2/3, 1) dt; the weights sum to 1, so ky = 0 modes get exactly
    exp(dt L)."""
    dt, nu = 0.05, 0.02
    o, lin = _linear_only(1, dt, nu=nu)
    kx0, ky, kxi, kyi, kz = o.tables()
    o.step(0.0)
    got = o.get_state()[0]
    w = np.array([A[0] + B[0], A[1] + B[1], A[2] + B[2]])
    tend = np.array([P[1], P[2], 1.0]) * dt
    for f, coef in ((0, nu), (2, 2 * nu)):
        expo = sum(w[s] * dt * _Lphys(o, kx0, ky, kz, tend[s], coef) for s in range(3))
        expect = np.exp(expo) * lin[:, f]
        assert np.abs(got[:, f] - expect).max() < 1e-13 * np.abs(lin).max()
        m0 = kyi == 0
        assert np.abs(got[m0, f] - np.exp(dt * _Lphys(o, kx0, ky, kz, 0, coef)[m0]) * lin[m0, f]).max() \
            < 1e-13 * np.abs(lin).max()
    assert abs(w.sum() - 1) < 1e-15


def test_cn_linear_factor():
    """RK3CN with no RHS on ky = 0 modes (constant L): the product of the
    trapezoidal factors (1 + h L)/(1 - h L), h = dt (a+b)/2, which agrees with
    exp(dt L) to O((dt L)^3)."""
    dt, nu = 0.05, 0.02
    o, lin = _linear_only(0, dt, nu=nu)
    kx0, ky, kxi, kyi, kz = o.tables()
    o.step(0.0)
    got = o.get_state()[0]
    m0 = kyi == 0
    L = _Lphys(o, kx0, ky, kz, 0.0, nu)[m0]
    fac = np.ones_like(L)
    for s in range(3):
        h = dt * (A[s] + B[s]) / 2
        fac *= (1 + h * L) / (1 - h * L)
    assert np.abs(got[m0, 0] - fac * lin[m0, 0]).max() < 1e-13 * np.abs(lin).max()
    x = dt * L
    assert np.all(np.abs(fac - np.exp(x)) <= 0.1 * np.abs(x) ** 3 + 1e-15)


def _mf_only(dt, integrator=1, eta=0.01, nu=0.02, q=1.5, Om=1.0):
    o = Oracle(model=0, integrator=integrator, NX=4, NY=4, NZ=32, dt=dt, q=q, omega=Om, nu=nu, eta=eta,
               remap=0)
    rng = np.random.default_rng(7)
    mf = np.zeros((4, o.nz), dtype=complex)
    k = ph.kz_slots(o.nz)
    for f in range(4):
        for m in (1, 2, 3):
            c = rng.standard_normal() + 1j * rng.standard_normal()
            mf[f, m] = c
            mf[f, (-m) % o.nz] = np.conj(c)
    o.set_state(np.zeros((o.nxy, 4, o.nz)), mf)
    o.linear_ops_init(0.0)
    return o, mf, 2 * np.pi * k


def _mf_exact(mf, kz, t, eta, nu, q, Om):
    Bx = mf[0] * np.exp(-eta * kz ** 2 * t)
    By = (mf[1] - q * t * mf[0]) * np.exp(-eta * kz ** 2 * t)
    Amat = np.array([[0.0, 2 * Om], [q - 2 * Om, 0.0]])
    w, V = np.linalg.eig(Amat)
    E = (V @ np.diag(np.exp(w * t)) @ np.linalg.inv(V)).real
    U = E @ np.stack([mf[2], mf[3]]) * np.exp(-nu * kz ** 2 * t)
    return np.stack([Bx, By, U[0], U[1]])


@pytest.mark.parametrize("integrator,diffusive,ratio", [
    (1, False, 8.0), (0, False, 8.0), (1, True, 2.0), (0, True, 4.0)])
def test_mean_fields_alone(integrator, diffusive, ratio):
    """No fluctuations: Bx decays, By is sheared from Bx, (Ux, Uy) is an
    epicyclic pair.  Without diffusion both integrators converge at third order
    to the closed form.  With diffusion the order is that of the reference's
    schemes: RK3CN (trapezoidal factors) is second order; RK3CN_Elin multiplies
    each stage's RHS by the exponential of its own substep only, i.e. the
    weights (8/15, 2/15, 1/3) at exponents (1, 7/15, 1/3) dt L, whose first
    moment is 0.707 instead of 1/2 -- first order.  Bx alone (no RHS) decays
    exactly under RK3CN_Elin."""
    T, q, Om = 1.0, 1.5, 1.0
    eta, nu = (0.01, 0.02) if diffusive else (0.0, 0.0)
    errs = []
    for n in (40, 80):
        o, mf0, kz = _mf_only(T / n, integrator, eta, nu, q, Om)
        o.steps(n, 0.0)
        got = o.get_state()[1]
        errs.append(np.abs(got - _mf_exact(mf0, kz, T, eta, nu, q, Om)).max())
        if integrator == 1:
            assert np.abs(got[0] - mf0[0] * np.exp(-eta * kz ** 2 * T)).max() < 1e-13
    assert errs[0] < (5e-3 if diffusive else 1e-5)
    assert 0.8 * ratio < errs[0] / errs[1] < 1.25 * ratio


def _single_mode_run(p, mode, vec_u, vec_b, n, dt):
    from ql_dns_amd.params import QLParams
    o = Oracle(model=0, integrator=1, NX=p["NX"], NY=p["NY"], NZ=p["NZ"], L=p["L"], q=p["q"],
               omega=p["Om"], nu=0.0, eta=0.0, quasilinear=0, remap=0, dt=dt)
    kx0, ky, kxi, kyi, kz = o.tables()
    lin = np.zeros((o.nxy, 4, o.nz), dtype=complex)
    lin[mode, 0] = vec_u[:, 0]
    lin[mode, 1] = ph.zeta_of(ky[mode], kz, vec_u)
    lin[mode, 2] = vec_b[:, 0]
    lin[mode, 3] = ph.zeta_of(ky[mode], kz, vec_b)
    o.set_state(lin, np.zeros((4, o.nz)))
    o.linear_ops_init(0.0)
    o.steps(n, 0.0)
    return o.get_state()[0][mode], (kx0[mode], ky[mode], kz)


def _divfree(rng, K):
    v = rng.standard_normal(K.shape) + 1j * rng.standard_normal(K.shape)
    k2 = np.sum(K * K, axis=-1)
    k2s = np.where(k2 > 0, k2, 1.0)
    return v - K * (np.sum(K * v, axis=-1) / k2s)[:, None]


def test_epicyclic_mode():
    """ky = 0, zero mean field, no dissipation: (ux, zeta) oscillates at
    omega^2 = 2 Omega (2 Omega - q) kz^2/|k|^2 and bx is constant while
    by grows as -q t bx (zeta, eta as defined by zeta = -(curl u)_x)."""
    prm = dict(NX=8, NY=8, NZ=32, L=(1.0, 1.0, 1.0), q=1.5, Om=1.0)
    T = 1.0
    rng = np.random.default_rng(11)
    o = Oracle(NX=8, NY=8, NZ=32, L=(1.0, 1.0, 1.0))
    kx0, ky, kxi, kyi, kz = o.tables()
    mode = int(np.nonzero((kxi == 2) & (kyi == 0))[0][0])
    K = np.stack([np.full(o.nz, kx0[mode]), np.zeros(o.nz), kz], axis=-1)
    u0, b0 = _divfree(rng, K), _divfree(rng, K)
    ze0 = ph.zeta_of(0.0, kz, u0)
    k2 = np.sum(K * K, axis=-1)
    w = np.sqrt(2 * 1.0 * (2 * 1.0 - 1.5) * kz ** 2 / k2)
    # ux'' = -w^2 ux with ux'(0) = -2 Omega i kz zeta / |k|^2
    dux0 = -2 * 1.0 * 1j * kz * ze0 / k2
    with np.errstate(invalid="ignore", divide="ignore"):
        ux_T = u0[:, 0] * np.cos(w * T) + np.where(w > 0, dux0 / np.where(w > 0, w, 1) * np.sin(w * T), dux0 * T)
    by_T = b0[:, 1] - 1.5 * T * b0[:, 0]
    errs = []
    for n in (25, 50):
        st, _ = _single_mode_run(prm, mode, u0, b0, n, T / n)
        ok = kz != 0
        errs.append(np.abs(st[0][ok] - ux_T[ok]).max())
        assert np.abs(st[2] - b0[:, 0]).max() < 1e-13
        assert np.abs(st[3][ok] - ph.zeta_of(0.0, kz, np.stack([b0[:, 0], by_T, b0[:, 2]], -1))[ok]).max() < 1e-12
    assert errs[0] < 1e-5 and 6.0 < errs[0] / errs[1] < 10.0


def test_shearing_wave_vs_ivp():
    """ky != 0, zero mean field, no dissipation: the sheared mode against
    scipy's solve_ivp of the Cartesian equations with explicit pressure."""
    prm = dict(NX=8, NY=8, NZ=32, L=(1.0, 2.0, 1.0), q=1.5, Om=1.0)
    o = Oracle(NX=8, NY=8, NZ=32, L=(1.0, 2.0, 1.0))
    kx0, ky, kxi, kyi, kz = o.tables()
    mode = int(np.nonzero((kxi == -1) & (kyi == 2))[0][0])
    rng = np.random.default_rng(5)
    K0 = np.stack([np.full(o.nz, kx0[mode]), np.full(o.nz, ky[mode]), kz], axis=-1)
    u0, b0 = _divfree(rng, K0), _divfree(rng, K0)
    q, Om, kyv = 1.5, 1.0, ky[mode]
    T = 0.5
    zero_mf = np.zeros((4, o.nz), dtype=complex)

    def f(t, y):
        u = (y[:6 * o.nz:2] + 1j * y[1:6 * o.nz:2]).reshape(o.nz, 3)
        b = (y[6 * o.nz::2] + 1j * y[6 * o.nz + 1::2]).reshape(o.nz, 3)
        du, db = ph.cartesian_rhs(o.NZ, kx0[mode] + q * t * kyv, kyv, kz, u, b, zero_mf, q, Om, 0.0)
        out = np.concatenate([du.ravel(), db.ravel()])
        r = np.empty(2 * out.size)
        r[0::2], r[1::2] = out.real, out.imag
        return r

    y0c = np.concatenate([u0.ravel(), b0.ravel()])
    y0 = np.empty(2 * y0c.size)
    y0[0::2], y0[1::2] = y0c.real, y0c.imag
    sol = solve_ivp(f, (0, T), y0, rtol=1e-12, atol=1e-14, method="DOP853")
    yT = sol.y[0::2, -1] + 1j * sol.y[1::2, -1]
    uT, bT = yT[:3 * o.nz].reshape(o.nz, 3), yT[3 * o.nz:].reshape(o.nz, 3)
    errs = []
    for n in (25, 50):
        st, _ = _single_mode_run(prm, mode, u0, b0, n, T / n)
        ex = np.stack([uT[:, 0], ph.zeta_of(kyv, kz, uT), bT[:, 0], ph.zeta_of(kyv, kz, bT)])
        errs.append(np.abs(st - ex).max() / np.abs(ex).max())
    assert errs[0] < 1e-4 and 6.0 < errs[0] / errs[1] < 10.0


# ----------------------------------------------------------------------------
# the RHS itself against the Cartesian equations (MHD_FullQlin.cpp:171-427,
# MHD_BQlin.cpp:117-284)
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("model,B0z,t", [(0, 0.0, 0.0), (0, 0.4, 0.23), (1, 0.0, 0.31)])
def test_rhs_vs_cartesian_equations(model, B0z, t):
    q, Om = 1.5, 1.0
    o = Oracle(model=model, NX=8, NY=8, NZ=32, L=(1.0, 2.0, 1.0), q=q, omega=Om, B0z=B0z, nu=0.003,
               eta=0.005)
    Omega = 2.0 / 3.0 * q if model == 1 else Om
    kx0, ky, kxi, kyi, kz = o.tables()
    rng = np.random.default_rng(23 + model)
    lin = np.zeros((o.nxy, 4, o.nz), dtype=complex)
    cart = {}
    for n in range(o.nxy):
        if kxi[n] == 0 and kyi[n] == 0:
            continue
        kx = kx0[n] + q * t * ky[n]
        K = np.stack([np.full(o.nz, kx), np.full(o.nz, ky[n]), kz], axis=-1)
        u, b = _divfree(rng, K), _divfree(rng, K)
        if kyi[n] == 0:  # ky = kz = 0 slot is not representable
            u[0], b[0] = 0, 0
        cart[n] = (u, b)
        lin[n] = np.stack([u[:, 0], ph.zeta_of(ky[n], kz, u), b[:, 0], ph.zeta_of(ky[n], kz, b)])
    mf = np.zeros((o.nMF, o.nz), dtype=complex)
    for f in range(o.nMF):
        for m in range(1, o.half):
            c = (rng.standard_normal() + 1j * rng.standard_normal()) / m
            mf[f, m], mf[f, (-m) % o.nz] = c, np.conj(c)
    o.set_state(lin, mf)
    rl, _, L = o.rhs(t, 0.0123)
    scale = np.abs(rl).max()
    for n, (u, b) in cart.items():
        kx = kx0[n] + q * t * ky[n]
        du, db = ph.cartesian_rhs(o.NZ, kx, ky[n], kz, u, b, mf, q, Omega, B0z, model)
        ex = np.stack([du[:, 0], ph.zeta_of(ky[n], kz, du), db[:, 0], ph.zeta_of(ky[n], kz, db)])
        ok = ~((kyi[n] == 0) & (np.arange(o.nz) == 0))
        assert np.abs(rl[n][:, ok] - ex[:, ok]).max() < 1e-11 * scale, n
        # dissipation operator at t + dt_lin: -nu |k|^2 (viscosity order 2)
        kxn = kx0[n] + q * (t + 0.0123) * ky[n]
        k2 = kxn ** 2 + ky[n] ** 2 + kz ** 2
        assert np.abs(L[n, 0] + 0.003 * k2).max() < 1e-12 * np.abs(L[n, 0]).max()
        assert np.abs(L[n, 2] + 0.005 * k2).max() < 1e-12 * np.abs(L[n, 2]).max()


# ----------------------------------------------------------------------------
# drives and diagnostics against physical-space averages of the 3-D fields
# ----------------------------------------------------------------------------
def _state_3d(o, p, lin):
    kx0, ky, kxi, kyi, kz = o.tables()
    scale = p.NX * p.NY * p.NZ
    U, Bf = [], []
    cu = np.zeros((o.nxy, o.nz, 3), dtype=complex)
    cb = np.zeros_like(cu)
    for n in range(o.nxy):
        cu[n] = ph.cartesian_from_uzeta(kx0[n], ky[n], kz, lin[n, 0], lin[n, 1])
        cb[n] = ph.cartesian_from_uzeta(kx0[n], ky[n], kz, lin[n, 2], lin[n, 3])
    for c in range(3):
        U.append(ph.real_field(ph.spectrum_3d(p.NX, p.NY, p.NZ, kxi, kyi, o.nz, cu[:, :, c] / scale)))
        Bf.append(ph.real_field(ph.spectrum_3d(p.NX, p.NY, p.NZ, kxi, kyi, o.nz, cb[:, :, c] / scale)))
    return np.array(U), np.array(Bf)


def test_drive_and_diagnostics_vs_physical_space():
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import QLParams
    p = QLParams(NX=8, NY=8, NZ=32, Lx=1.0, Ly=2.0, Lz=1.0, nu=0.003, eta=0.004)
    lin, mf = initial_state(p, amp=1.0, kdecay=400.0)
    o = Oracle.from_params(p)
    o.set_state(lin, mf)
    o.linear_ops_init(0.0)
    o.rhs(0.0, 0.0)
    drive = o.drive()
    U, Bf = _state_3d(o, p, lin)
    assert np.abs(U.imag).max() < 1e-12 * np.abs(U).max()
    U, Bf = U.real, Bf.real
    avg = lambda f: f.mean(axis=(0, 1))  # xy average -> z profile
    kzs = 2 * np.pi / p.Lz * ph.kz_slots(o.nz)
    dz = lambda prof: 1j * kzs * np.fft.fft(prof)[ph.pad_index(p.NZ, o.nz)]
    ux, uy, uz = U
    bx, by, bz = Bf
    exp_drive = np.stack([dz(avg(ux * bz - uz * bx)), dz(avg(uy * bz - uz * by)),
                          -dz(avg(uz * ux - bz * bx)), -dz(avg(uz * uy - bz * by))])
    assert rel_err(drive, exp_drive) < 1e-12
    # Calc_Energy_AM_Diss: E = Lz <|u|^2>/2, AM = Lz <ux uy>, diss = nu Lz <|grad u|^2>
    d = o.diagnostics(0.0)
    Lz = p.Lz
    assert abs(d[2] - Lz * 0.5 * np.mean(ux ** 2 + uy ** 2 + uz ** 2)) < 1e-12 * d[2]
    assert abs(d[3] - Lz * 0.5 * np.mean(bx ** 2 + by ** 2 + bz ** 2)) < 1e-12 * d[3]
    assert abs(d[6] - Lz * np.mean(ux * uy)) < 1e-12 * abs(d[2])
    assert abs(d[7] - Lz * np.mean(bx * by)) < 1e-12 * abs(d[3])

    def grad2(F, L):
        ks = [2 * np.pi / L[i] * np.fft.fftfreq(F.shape[i + 1], 1.0 / F.shape[i + 1]) for i in range(3)]
        tot = 0.0
        for c in range(3):
            Fh = np.fft.fftn(F[c])
            for i in range(3):
                sh = [1, 1, 1]
                sh[i] = -1
                tot = tot + np.mean(np.fft.ifftn(1j * ks[i].reshape(sh) * Fh).real ** 2)
        return tot
    assert abs(d[10] - p.nu * Lz * grad2(U, (p.Lx, p.Ly, p.Lz))) < 1e-10 * d[10]
    assert abs(d[11] - p.eta * Lz * grad2(Bf, (p.Lx, p.Ly, p.Lz))) < 1e-10 * d[11]
    prof = ph.to_grid(p.NZ, mf).real
    assert abs(d[1] - Lz * 0.5 * np.mean(prof[0] ** 2 + prof[1] ** 2)) < 1e-12 * d[1]
    assert abs(d[5] - Lz * np.mean(prof[0] * prof[1])) < 1e-12 * abs(d[1])
    # Save_Mean_Fields: the real-space profiles themselves
    assert rel_err(o.mf_real(), prof) < 1e-13


def _budget_mismatch(dt, integrator):
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import QLParams
    p = QLParams(NX=8, NY=8, NZ=32, Lx=1.0, Ly=2.0, Lz=1.0, nu=0.003, eta=0.004, remap=0, dt=dt,
                 integrator=integrator)
    lin, mf = initial_state(p, amp=1.0, kdecay=400.0)
    mf[2, 1] = mf[0, 2]  # give the mean flow a profile too
    mf[2, -1] = np.conj(mf[0, 2])
    o = Oracle.from_params(p)
    o.set_state(lin, mf)
    o.linear_ops_init(0.0)
    E, t = [], 0.0
    for k in range(3):
        d = o.diagnostics(t)
        E.append(d[:4].sum())
        if k == 1:
            budget = p.q * (d[6] + d[4] - d[7] - d[5]) - d[8:12].sum()
            scale = p.q * np.abs(d[4:8]).sum() + d[8:12].sum()
        if k < 2:
            t = o.steps(1, t)
    return (E[2] - E[0]) / (2 * dt) - budget, scale


@pytest.mark.parametrize("integrator", [0, 1])
def test_energy_budget(integrator):
    """Total energy of the QL system changes only by shear work and
    dissipation, dE/dt = q (AM_u + AM_MU - AM_b - AM_MB) - sum(diss): the
    quasilinear exchange terms cancel.  Checked with a centred difference over
    two steps; the residual is time-discretisation error and falls ~dt^3."""
    m1, s1 = _budget_mismatch(1e-3, integrator)
    m2, s2 = _budget_mismatch(2.5e-4, integrator)
    assert abs(m2) < 2e-4 * s2
    assert abs(m1) / abs(m2) > 30
