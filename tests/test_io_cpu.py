"""The reference's file formats (include/ql_io.h) on the CPU: FullSolution.h5
dump / restart round trips, the choice of the latest save, the file's
structure as h5dump reports it (dataset names, types, shapes, kx = kx0 + q t ky,
the /ky and /kz tables), and the .dat time series byte for byte."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT


def _p(**kw):
    from ql_dns_amd.params import QLParams
    return QLParams(**kw)


def _state(p, seed=3):
    rng = np.random.default_rng(seed)
    lin = rng.standard_normal((p.nxy, 4, p.nz)) + 1j * rng.standard_normal((p.nxy, 4, p.nz))
    mf = rng.standard_normal((p.nMF, p.nz)) + 1j * rng.standard_normal((p.nMF, p.nz))
    from ql_dns_amd.ic import mode_tables
    kx0 = mode_tables(p)[0].copy()
    kx0[3] -= (p.NX - 1) * 2 * np.pi / p.Lx  # one remapped mode
    return lin * 1e3, mf, kx0


def _declared():
    with open(os.path.join(ROOT, "include", "ql_io.h")) as f:
        src = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(qlio_\w+)\s*\(", src, re.M)))


def test_io_header_matches_binding_and_library():
    from ql_dns_amd import qlio
    assert sorted(qlio.EXPORTS) == _declared()
    L = qlio.load()
    for name in _declared():
        assert hasattr(L, name), name


@pytest.mark.parametrize("model", [0, 1])
def test_h5_roundtrip_fp64_exact_and_fp32_rounded(tmp_path, model):
    from ql_dns_amd import qlio
    p = _p(NX=8, NY=6, NZ=32, model=model)
    lin, mf, kx0 = _state(p)
    f64, f32 = str(tmp_path / "a64.h5"), str(tmp_path / "a32.h5")
    t = 0.1 + 0.2  # not representable in 15 digits: the attribute keeps it
    qlio.h5_save(f64, p, t, lin, mf, kx0, create=True, fp64=True)
    qlio.h5_save(f32, p, t, lin, mf, kx0, create=True, fp64=False)
    t2, l2, m2, k2 = qlio.h5_load_latest(f64, p)
    assert t2 == t
    assert np.array_equal(l2, lin) and np.array_equal(m2, mf) and np.array_equal(k2, kx0)
    t3, l3, m3, k3 = qlio.h5_load_latest(f32, p)
    assert t3 == float("%.15g" % t)
    c64 = lambda a: a.astype(np.complex64).astype(np.complex128)
    assert np.array_equal(l3, c64(lin)) and np.array_equal(m3, c64(mf))
    # without kx0 the reference's kx - q t ky (kx stored as float)
    from ql_dns_amd.ic import mode_tables
    ky = mode_tables(p)[1]
    kx_f = (kx0 + p.q * t3 * ky).astype(np.float32).astype(np.float64)
    assert np.array_equal(k3, kx_f - p.q * t3 * ky)


def test_h5_latest_of_several_saves(tmp_path):
    from ql_dns_amd import qlio
    p = _p(NX=4, NY=4, NZ=32)
    path = str(tmp_path / "FullSolution.h5")
    states = {}
    for k, t in enumerate([0.5, 2.25, 1.0]):
        st = _state(p, seed=10 + k)
        states[t] = st
        qlio.h5_save(path, p, t, *st, create=(k == 0), fp64=True)
    assert qlio.h5_count(path) == 3
    t, lin, mf, kx0 = qlio.h5_load_latest(path, p)
    assert t == 2.25
    assert np.array_equal(lin, states[2.25][0])
    with pytest.raises(qlio.QLIOError):
        qlio.h5_save(path, p, 2.25, *states[2.25], fp64=True)  # same time twice


@pytest.mark.skipif(shutil.which("h5dump") is None and not os.path.exists("/opt/conda/bin/h5dump"),
                    reason="h5dump not available")
def test_h5_structure_matches_the_reference_layout(tmp_path):
    from ql_dns_amd import qlio
    from ql_dns_amd.ic import mode_tables
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    p = _p(NX=4, NY=4, NZ=32)
    lin, mf, kx0 = _state(p)
    path = str(tmp_path / "FullSolution.h5")
    t = 0.75
    qlio.h5_save(path, p, t, lin, mf, kx0, create=True)
    hdr = subprocess.run([h5dump, "-H", path], capture_output=True, text=True, check=True).stdout
    kx0_t, ky, kxi, kyi, kz, _ = mode_tables(p)
    assert 'GROUP "t=0.75"' in hdr
    for jj in range(p.nxy):  # FullSave_Load.cpp:191-199
        name = f"Ind {jj}: kx {kxi[jj]} ky {kyi[jj]}"
        m = re.search(r'DATASET "%s" \{\s*DATATYPE\s+(\S+)\s*DATASPACE\s+SIMPLE \{ \( (\d+), (\d+) \)' % re.escape(name), hdr)
        assert m, name
        assert m.group(1) == "H5T_IEEE_F32LE" and int(m.group(2)) == 4 * p.nz and int(m.group(3)) == 2
    assert re.search(r'DATASET "Mean" \{\s*DATATYPE\s+H5T_IEEE_F32LE\s*DATASPACE\s+SIMPLE \{ \( %d, 2 \)' % (4 * p.nz), hdr)
    assert re.search(r'DATASET "ky" \{\s*DATATYPE\s+H5T_IEEE_F32LE', hdr)
    assert re.search(r'DATASET "kz" \{\s*DATATYPE\s+H5T_IEEE_F64LE', hdr)

    def values(obj):
        out = subprocess.run([h5dump, "-d", obj, "-m", "%.17g", "-y", "-w", "0", path], capture_output=True,
                             text=True, check=True).stdout
        body = out[out.index("DATA {") + 6:out.rindex("}")]
        return np.array([float(x) for x in re.findall(r"[-+0-9.eE]+(?:inf|nan)?", body.split("}")[0])])
    assert np.array_equal(values("/kz"), kz)
    assert np.array_equal(values("/ky"), ky.astype(np.float32).astype(np.float64))
    assert np.array_equal(values("/t=0.75/kx"), (kx0 + p.q * t * ky).astype(np.float32).astype(np.float64))
    first = values(f"/t=0.75/Ind 1: kx {kxi[1]} ky {kyi[1]}")
    ref = np.stack([lin[1].reshape(-1).real, lin[1].reshape(-1).imag], axis=-1).reshape(-1)
    assert np.array_equal(first, ref.astype(np.float32).astype(np.float64))


def test_dat_series_bytes(tmp_path):
    from ql_dns_amd import qlio
    d = str(tmp_path) + "/"
    qlio.dat_init(d)
    rng = np.random.default_rng(1)
    saves = [(0.0, rng.standard_normal(18), rng.standard_normal((4, 32))),
             (0.01, rng.standard_normal(18), rng.standard_normal((4, 32)))]
    for t, diag, mfr in saves:
        qlio.dat_append(d, t, diag, nrey=4, mf_real=mfr)
    rd = lambda n: np.fromfile(os.path.join(d, n), dtype="<f8")
    assert np.array_equal(rd("time_vec.dat"), [0.0, 0.01])
    assert np.array_equal(rd("energy.dat"), np.concatenate([s[1][0:4] for s in saves]))
    assert np.array_equal(rd("angular_momentum.dat"), np.concatenate([s[1][4:8] for s in saves]))
    assert np.array_equal(rd("dissipation.dat"), np.concatenate([s[1][8:12] for s in saves]))
    assert np.array_equal(rd("reynolds_stress.dat"), np.concatenate([s[1][12:16] for s in saves]))
    assert np.array_equal(rd("mean_fields.dat"), np.concatenate([s[2].reshape(-1) for s in saves]))
    # MHD_BQlin declares 5 Reynolds values and fills none: zeros past the 4 held
    qlio.dat_init(d)
    qlio.dat_append(d, 0.0, saves[0][1], nrey=5)
    assert np.array_equal(rd("reynolds_stress.dat"), np.concatenate([saves[0][1][12:16], [0.0]]))
