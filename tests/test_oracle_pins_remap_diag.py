"""Pins of three oracle functions that the physics pins in test_oracle_pins.py
leave open, each checked from its definition rather than from the oracle's
expression:

* the shearing-box remap (Models/MHD_FullQlin.cpp:464-481): a mode whose
  sheared wavenumber kx0 + q t ky has left the band (NX-1)/2 * 2 pi/Lx is moved
  back by (NX-1) * 2 pi/Lx and zeroed; every other mode is untouched.  Checked
  with hand-placed wavenumbers at the threshold, one ulp above it, between
  (NX-1)/2 and NX/2 band widths, and with the sheared term carrying the event;
* the Reynolds projections of Calc_Energy_AM_Diss (:669-676): the z-average of
  drive(z) * B(z) over the rms of B(z), formed in physical space from numpy
  transforms of the spectra;
* eta_K * kmax (:629-630): the Kolmogorov scale (nu^3/epsilon)^(1/4) of the
  dissipation rate epsilon (itself checked here against <|dB/dz|^2>,
  <|dU/dz|^2> of the mean-field profiles) times the largest retained kz of the
  wavenumber table.

test_pins_catch_planted_mistakes compiles the oracle with one plausible slip
at a time (>= for >, NX/2 for (NX-1)/2, a swapped field index, a dropped
square root or power, the wrong kmax) and requires each to fail a pin.
"""
import os
import shutil
import subprocess

import ctypes
import numpy as np
import pytest

import physics_ref as ph
from oracle.oracle import Oracle, bind

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def _herm_profile_spectrum(rng, NZ, nz, scale):
    """Spectrum (NZ x coefficient, kz slot order) of a random real profile."""
    prof = rng.standard_normal(NZ)
    spec = np.fft.fft(prof)[ph.pad_index(NZ, nz)]
    return scale * spec


# ----------------------------------------------------------------------------
# remap
# ----------------------------------------------------------------------------
def check_remap(library=None):
    NX, NY, NZ, Lx, Ly = 16, 16, 32, 1.0, 4.0
    o = Oracle(NX=NX, NY=NY, NZ=NZ, L=(Lx, Ly, 1.0), library=library)
    kx0, ky, kxi, kyi, kz = o.tables()
    rng = np.random.default_rng(7)
    lin = rng.standard_normal((o.nxy, 4, o.nz)) + 1j * rng.standard_normal((o.nxy, 4, o.nz))
    mf = rng.standard_normal((4, o.nz)) + 1j * rng.standard_normal((4, o.nz))
    kfac = 2 * np.pi / Lx
    T = (NX - 1) / 2.0 * kfac          # edge of the band
    shift = (NX - 1) * kfac

    # (1) t = 0: kx(t) = kx0 exactly, planted at and around the edge
    planted = kx0.copy()
    cases = {0: (T, False),                               # at the edge: stays
             1: (np.nextafter(T, np.inf), True),          # one ulp out: moves
             2: (np.nextafter(T, -np.inf), False),        # one ulp in: stays
             3: (7.75 * kfac, True),                      # between (NX-1)/2 and NX/2 widths
             4: (NX / 2.0 * kfac, True),                  # at the NX/2 width
             5: (-T - kfac, False),                       # far negative: stays
             6: (20.0 * kfac, True)}                      # far out
    for n, (v, _) in cases.items():
        planted[n] = v
    o.set_state(lin, mf, planted)
    o.remap(0.0)
    lin1, mf1, kx1 = o.get_state()
    for n, (v, moves) in cases.items():
        if moves:
            assert kx1[n] == v - shift, (n, v, kx1[n])
            assert not lin1[n].any(), n
        else:
            assert kx1[n] == v, (n, v, kx1[n])
            assert np.array_equal(lin1[n], lin[n]), n
    rest = np.arange(len(cases), o.nxy)
    moved = planted[rest] > T
    assert np.array_equal(kx1[rest][~moved], planted[rest][~moved])
    assert np.array_equal(lin1[rest][~moved], lin[rest][~moved])
    assert np.array_equal(mf1, mf)

    # (2) the sheared term decides: q t such that modes with ky > 0 cross
    o.set_state(lin, mf, kx0)
    qt = 0.61
    o.remap(qt)
    lin2, _, kx2 = o.get_state()
    kxt = kx0 + qt * ky
    out = kxt > T
    assert out.any() and (~out).any() and (ky[out] > 0).all()
    assert np.array_equal(kx2[out], kx0[out] - shift)
    assert np.array_equal(kx2[~out], kx0[~out])
    assert not lin2[out].any()
    assert np.array_equal(lin2[~out], lin[~out])
    # after the remap every mode's sheared wavenumber lies in the band
    assert ((kx2 + qt * ky) <= T).all()
    assert ((kx2 + qt * ky) > T - shift - 1e-9).all()
    o.close()


def test_remap_predicate_and_shift():
    check_remap()


# ----------------------------------------------------------------------------
# Reynolds projections and eta_K kmax
# ----------------------------------------------------------------------------
def _diag_case(library=None, model=0):
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import QLParams
    p = QLParams(NX=8, NY=8, NZ=32, Lx=1.0, Ly=2.0, Lz=1.3, nu=0.003, eta=0.0045, model=model)
    lin, mf = initial_state(p, amp=1.0, kdecay=400.0)
    rng = np.random.default_rng(11)
    nz, NZ = p.nz, p.NZ
    if mf.shape[0] == 4:
        # a mean flow of its own, so that every projection has a distinct field
        mf[2] = _herm_profile_spectrum(rng, NZ, nz, 0.07)
        mf[3] = _herm_profile_spectrum(rng, NZ, nz, 0.05)
    o = Oracle.from_params(p, library=library)
    o.set_state(lin, mf)
    o.linear_ops_init(0.0)
    o.rhs(0.0, 0.0)
    return p, o, mf


def check_projections(library=None):
    p, o, mf = _diag_case(library)
    d = o.diagnostics(0.0)
    drive = o.drive()
    # physical-space profiles from numpy transforms (NZ x storage, 1/NZ transform)
    Dz = ph.to_grid(p.NZ, drive).real
    Mz = ph.to_grid(p.NZ, mf).real
    assert np.abs(ph.to_grid(p.NZ, mf).imag).max() < 1e-14 * np.abs(Mz).max()
    for f in range(4):
        # Parseval with numpy's ifft: sum_k Re(X_k conj Y_k) = NZ sum_z x(z) y(z)
        num = p.NZ * np.sum(Dz[f] * Mz[f])
        rms = np.sqrt(p.NZ * np.sum(Mz[f] ** 2))
        expect = num / rms
        assert abs(d[12 + f] - expect) < 1e-12 * (abs(expect) + np.abs(Dz[f]).max() * np.sqrt(p.NZ)), \
            (f, d[12 + f], expect)
    o.close()


def check_etak(library=None):
    p, o, mf = _diag_case(library)
    d = o.diagnostics(0.0)
    kz = o.tables()[4]
    # mean-field dissipation = nu or eta * Lz * <(dF/dz)^2> over z, from numpy
    kzs = 2 * np.pi / p.Lz * ph.kz_slots(o.nz)
    dprof = ph.to_grid(p.NZ, 1j * kzs[None, :] * mf).real
    diss_MB = p.eta * p.Lz * np.mean(dprof[0] ** 2 + dprof[1] ** 2)
    diss_MU = p.nu * p.Lz * np.mean(dprof[2] ** 2 + dprof[3] ** 2)
    assert abs(d[9] - diss_MB) < 1e-12 * diss_MB
    assert abs(d[8] - diss_MU) < 1e-12 * diss_MU
    # Kolmogorov scale of the total dissipation times the largest retained kz
    kmax = np.abs(kz).max()
    assert kmax == pytest.approx(2 * np.pi / p.Lz * ((o.nz - 1) // 2), rel=1e-15)
    eps_u, eps_b = d[8] + d[10], d[9] + d[11]
    assert d[16] == pytest.approx((p.nu ** 3 / eps_u) ** 0.25 * kmax, rel=1e-13)
    assert d[17] == pytest.approx((p.eta ** 3 / eps_b) ** 0.25 * kmax, rel=1e-13)
    o.close()


def test_reynolds_projections_vs_physical_space():
    check_projections()


def test_etak_kmax_from_dissipation():
    check_etak()


# ----------------------------------------------------------------------------
# each pin fails on a planted slip
# ----------------------------------------------------------------------------
MUTANTS = [
    ("remap >=", "if (kxt > nx / 2.0 * kxLfac) {", "if (kxt >= nx / 2.0 * kxLfac) {", check_remap),
    ("remap NX/2", "if (kxt > nx / 2.0 * kxLfac) {", "if (kxt > o->p.NX / 2.0 * kxLfac) {", check_remap),
    ("remap shift NX", "o->kx0[i] = o->kx0[i] - nx * kxLfac;", "o->kx0[i] = o->kx0[i] - o->p.NX * kxLfac;",
     check_remap),
    ("remap unsheared", "kxt = o->kx0[i] + qt * o->ky[i];", "kxt = o->kx0[i];", check_remap),
    ("projection field", "v = M[(size_t)f * nz + m];", "v = M[(size_t)(f ^ 1) * nz + m];", check_projections),
    ("projection no sqrt", "out[12 + f] = num / sqrt(den);", "out[12 + f] = num / den;", check_projections),
    ("projection conj", "num += creal(d * conj(v));", "num += creal(d * v);", check_projections),
    ("etak kmax", "double kzfac = 2 * QL_PI / o->p.L[2] * (o->NZ / 3);",
     "double kzfac = 2 * QL_PI / o->p.L[2] * (o->NZ / 2);", check_etak),
    ("etak power", "out[16] = pow(nu * nu * nu / (diss_MU + diss_u), 0.25) * kzfac;",
     "out[16] = pow(nu * nu / (diss_MU + diss_u), 0.25) * kzfac;", check_etak),
    ("etak fields", "out[17] = o->p.model == ORC_MODEL_HD_FULLU ? 0.0 : pow(eta * eta * eta / (diss_MB + diss_b), "
     "0.25) * kzfac;",
     "out[17] = o->p.model == ORC_MODEL_HD_FULLU ? 0.0 : pow(eta * eta * eta / (diss_MU + diss_b), 0.25) * kzfac;",
     check_etak),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build oracle variants")
@pytest.mark.parametrize("name,old,new,check", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pins_catch_planted_mistakes(tmp_path, name, old, new, check):
    src = open(os.path.join(ORACLE_DIR, "ql_oracle.c")).read()
    assert src.count(old) == 1, name
    mut = tmp_path / "ql_oracle_mutant.c"
    mut.write_text(src.replace(old, new))
    so = tmp_path / "libql_oracle_mutant.so"
    subprocess.run(["gcc", "-O1", "-std=gnu11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fcx-fortran-rules",
                    "-I", ORACLE_DIR, "-shared", "-o", str(so), str(mut), "-lm"], check=True)
    L = bind(ctypes.CDLL(str(so)))
    with pytest.raises(AssertionError):
        check(L)
