"""Pins of the oracle's DrivingNoise (Models/MHD_FullQlin.cpp:786-1053,
MHD_BQlin.cpp:558-740).

The reference draws from a clock-seeded mt19937; the oracle (and the GPU) use
Philox4x64-10 with a Box-Muller transform.  The generator is pinned bit for
bit against numpy's Philox; the normals by their distribution; the
reference's multipliers by what they are built to do (each drawn entry
injects the same expected energy 4 nf f^2 into u and into b, nf = dt (NX NY
NZ)^2 / (16 32 32)^2, split evenly between the ux and zeta parts, whatever
its |k|), measured through the energy of Calc_Energy_AM_Diss; the |k| range,
the drive_only flags, the ky = 0 Hermitian pairs and MHD_BQlin's undriven
ky = 0 plane by exact masks.
"""
import math

import numpy as np
import pytest
from scipy import stats

from oracle.oracle import Oracle, philox4x64
from ql_dns_amd.params import MODEL_MHD_BQLIN, MODEL_MHD_FULLQLIN, QLParams


@pytest.mark.parametrize("ctr,key", [((1, 0, 0, 0), (1, 2)), ((7, 3, 0xFFFFFFFFFFFFFFFF, 9), (0xDEADBEEF, 5)),
                                     ((0x243F6A8885A308D3, 0x13198A2E03707344, 1, 2), (0xA4093822299F31D0, 0))])
def test_philox_matches_numpy(ctr, key):
    bg = np.random.Philox(key=np.array(key, dtype=np.uint64),
                          counter=np.array((ctr[0] - 1,) + tuple(ctr[1:]), dtype=np.uint64))
    assert np.array_equal(philox4x64(ctr, key), bg.random_raw(4))


def _orc(model=MODEL_MHD_FULLQLIN, **kw):
    p = QLParams(model=model, **kw)
    o = Oracle.from_params(p, nthreads=2)
    o.set_state(np.zeros((o.nxy, 4, o.nz), complex), np.zeros((o.nMF, o.nz), complex))
    return p, o


def test_normals_are_standard_and_independent():
    p, o = _orc()
    z = np.array([o.noise_normals(0.37, i, j) for i in range(o.nxy) for j in range(o.nz)])
    assert z.shape == (o.nxy * o.nz, 8)
    for c in range(8):
        assert stats.kstest(z[:, c], "norm").pvalue > 1e-4
    C = np.corrcoef(z.T)
    assert np.abs(C - np.eye(8)).max() < 5 / math.sqrt(z.shape[0])
    # a function of (seed, t, mode, slot) only
    assert np.array_equal(o.noise_normals(0.37, 5, 3), o.noise_normals(0.37, 5, 3))
    assert not np.array_equal(o.noise_normals(0.37, 5, 3), o.noise_normals(0.38, 5, 3))
    assert not np.array_equal(o.noise_normals(0.37, 5, 3), o.noise_normals(0.37, 5, 4))
    o.set_noise(1.0, seed=12345)
    assert not np.array_equal(o.noise_normals(0.37, 5, 3), z[5 * o.nz + 3])


def _k2(o, p, t):
    """|k|^2 per (mode, slot) at time t from the wavenumber tables."""
    kx0, ky, _, _, kz = o.tables()
    kx = kx0 + p.q * t * ky
    return (kx ** 2 + ky ** 2)[:, None] + (kz ** 2)[None, :], ky


def test_energy_injection_is_white():
    p, o = _orc(NX=32, NY=32, NZ=64)
    f, t, dt = 0.7, 0.3, 0.02
    o.set_noise(f, seed=99)
    o.driving_noise(t, dt)
    lin = o.get_state()[0]
    k2, ky = _k2(o, p, t)
    mult = np.where(ky == 0, 1.0, 2.0)[:, None]
    # the two parts of E_u (MHD_FullQlin.cpp:530, 572): ux part lapF/lap2 |ux|^2,
    # zeta part -|zeta|^2/lap2; lap2 = -(ky^2 + kz^2), lapF = -|k|^2
    _, kyv, _, _, kz = o.tables()
    lap2 = -(kyv[:, None] ** 2 + kz[None, :] ** 2)
    lapF = -k2
    ok = lap2 != 0
    ux, ze, bx, et = lin[:, 0], lin[:, 1], lin[:, 2], lin[:, 3]
    parts = []
    for a, b in ((ux, ze), (bx, et)):
        pu = np.where(ok, mult * lapF / np.where(ok, lap2, 1) * np.abs(a) ** 2, 0)
        pz = np.where(ok, -mult * np.abs(b) ** 2 / np.where(ok, lap2, 1), 0)
        parts.append((pu, pz))
    # entries drawn: every slot of ky != 0 modes, slots >= 1 of the kx < 0, ky = 0 modes
    ngen = int(np.sum(kyv != 0)) * o.nz + (p.NX // 2 - 1) * (o.nz - 1)
    nf = dt * (p.NX * p.NY * p.NZ) ** 2 * (1.0 / (16 * 32 * 32)) ** 2
    want = 4 * nf * f * f * ngen
    sd = math.sqrt(8.0) / 4 / math.sqrt(ngen)  # relative spread of a sum of chi^2_4 terms
    for pu, pz in parts:
        assert abs((pu.sum() + pz.sum()) / want - 1) < 6 * sd
        assert abs(pu.sum() / (want / 2) - 1) < 6 * sd * math.sqrt(2)
        assert abs(pz.sum() / (want / 2) - 1) < 6 * sd * math.sqrt(2)
        # flat in |k|: low and high halves of the drawn entries alike
        e = (pu + pz)[ok]
        kk = k2[ok]
        med = np.median(kk)
        for sel in (kk <= med, kk > med):
            m = e[sel].mean() / (4 * nf * f * f)
            assert abs(m - 1) < 8 * math.sqrt(8.0) / 4 / math.sqrt(sel.sum())
    # the diagnostics agree (Calc_Energy_AM_Diss, E_u and E_b)
    d = o.diagnostics(t)
    Lz, N2 = p.Lz, float(p.NX * p.NY * p.NZ) ** 2
    assert abs(d[2] - (parts[0][0].sum() + parts[0][1].sum()) * Lz / N2 / 2) < 1e-10 * d[2]
    assert abs(d[3] - (parts[1][0].sum() + parts[1][1].sum()) * Lz / N2 / 2) < 1e-10 * d[3]


def test_ky0_pairs_are_hermitian_and_slot0_undriven():
    p, o = _orc()
    o.set_noise(1.3, seed=7)
    o.driving_noise(0.21, 0.01)
    lin = o.get_state()[0]
    _, ky, kxi, kyi, _ = o.tables()
    nz = o.nz
    for i in range(o.nxy):
        if kyi[i] != 0 or kxi[i] <= 0:
            continue
        j = [k for k in range(o.nxy) if kyi[k] == 0 and kxi[k] == -kxi[i]][0]
        for f in range(4):
            assert lin[i, f, 0] == 0 and lin[j, f, 0] == 0
            for s in range(1, nz):
                assert lin[i, f, s] == np.conj(lin[j, f, nz - s])
                assert lin[i, f, s] != 0
    assert np.all(lin[0] == 0)  # kx = ky = 0


def test_masks_range_and_flags():
    p, o = _orc()
    t, lo, hi = 0.4, 12.0, 40.0
    k2, ky = _k2(o, p, t)
    inside = (k2 > lo * lo) & (k2 < hi * hi)
    near = (np.abs(k2 - lo * lo) < 1e-9 * lo * lo) | (np.abs(k2 - hi * hi) < 1e-9 * hi * hi)
    for dv, dm in ((0, 0), (1, 0), (0, 1)):
        o.set_state(np.zeros((o.nxy, 4, o.nz), complex), np.zeros((o.nMF, o.nz), complex))
        o.set_noise(1.0, range_low=lo, range_high=hi, drive_only_velocity=dv, drive_only_magnetic=dm, seed=3)
        o.driving_noise(t, 0.01)
        lin = o.get_state()[0]
        nonzero = np.abs(lin) > 0
        for f in range(4):
            driven = not ((f < 2 and dm) or (f >= 2 and dv))
            if not driven:
                assert not nonzero[:, f].any()
                continue
            want = inside & (ky[:, None] != 0)
            got = nonzero[:, f] & (ky[:, None] != 0)
            assert np.array_equal(got[~near], want[~near])


def test_bqlin_leaves_ky0_undriven_and_ignores_flags():
    p, o = _orc(model=MODEL_MHD_BQLIN)
    o.set_noise(1.0, drive_only_velocity=1, seed=11)
    o.driving_noise(0.1, 0.01)
    lin = o.get_state()[0]
    _, ky, _, _, _ = o.tables()
    assert np.all(lin[ky == 0] == 0)
    assert np.all(np.abs(lin[ky != 0]) > 0)  # b and eta driven too


def test_steps_with_noise_draw_at_the_step_start_time():
    from ql_dns_amd.ic import initial_state
    p = QLParams()
    lin, mf = initial_state(p)
    res = []
    for manual in (False, True):
        o = Oracle.from_params(p, nthreads=2)
        o.set_state(lin, mf)
        o.linear_ops_init(0.0)
        o.set_noise(0.5, seed=21)
        if manual:
            t = 0.0
            for n in range(6):
                dt = o.step(t)
                o.driving_noise(t, dt)
                t += dt
                o.remap(p.q * t)
        else:
            t = o.steps(6, 0.0)
        res.append((t,) + tuple(o.get_state()))
    assert res[0][0] == res[1][0]
    for a, b in zip(res[0][1:], res[1][1:]):
        assert np.array_equal(a, b)
