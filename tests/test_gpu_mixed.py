"""Mixed precision (FP32 transforms and products; FP64 state, accumulation,
reduction, mean-field update and dissipation factors) against the FP64 GPU
path on identical inputs.  Bar: mean fields, EMF / stress drives and
fluctuation energies within 1e-6 relative of FP64."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _run(p, nsteps):
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    lin, mf = initial_state(p)
    c = QLContext(p)
    c.set_state(lin, mf)
    c.linear_ops_init(0.0)
    t = c.steps(nsteps, 0.0)
    lin2, mf2, kx = c.get_state()
    out = dict(t=t, lin=lin2, mf=mf2, kx=kx, drive=c.drive(), diag=c.diagnostics(t))
    c.close()
    return out


@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("shape,nsteps", [((16, 16, 32), 100), ((16, 16, 64), 100), ((32, 32, 128), 50),
                                          ((8, 8, 256), 50)])
def test_mixed_vs_fp64(model, shape, nsteps):
    from ql_dns_amd.params import PREC_MIXED, QLParams
    NX, NY, NZ = shape
    p = QLParams(NX=NX, NY=NY, NZ=NZ, model=model, dt=0.005, nu=1 / 8000, eta=1 / 8000)
    a = _run(p, nsteps)
    b = _run(p.replace(precision=PREC_MIXED), nsteps)
    assert a["t"] == b["t"]
    assert np.array_equal(a["kx"], b["kx"])
    assert rel_err(b["mf"], a["mf"]) < 1e-6
    assert rel_err(b["drive"], a["drive"]) < 1e-6
    assert abs(b["diag"][2] - a["diag"][2]) < 1e-6 * abs(a["diag"][2])
    assert abs(b["diag"][3] - a["diag"][3]) < 1e-6 * abs(a["diag"][3])
    # the fluctuations themselves carry FP32 rounding of the RHS only
    assert rel_err(b["lin"], a["lin"]) < 1e-5


def _ctx(p, lin, mf):
    from ql_dns_amd.capi import QLContext
    c = QLContext(p)
    c.set_state(lin, mf)
    c.linear_ops_init(0.0)
    return c


def test_mixed_large_box_2000_steps():
    """128x128x256, Re = Rm = 16000, 2000 steps: mean fields, EMF / stress
    drives (each drive and all together) and fluctuation energies within 1e-6
    of the FP64 run at steps 100, 500 and 2000."""
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import PREC_MIXED, workload
    p = workload("box128x128x256")
    lin, mf = initial_state(p)
    a, b = _ctx(p, lin, mf), _ctx(p.replace(precision=PREC_MIXED), lin, mf)
    ta = tb = 0.0
    done = 0
    for n in (100, 500, 2000):
        ta, tb = a.steps(n - done, ta), b.steps(n - done, tb)
        done = n
        assert ta == tb
        (la, ma, ka), (lb, mb, kb) = a.get_state(), b.get_state()
        da, db = a.drive(), b.drive()
        ea, eb = a.diagnostics(ta), b.diagnostics(tb)
        assert np.array_equal(ka, kb)
        errs = dict(mf=rel_err(mb, ma), drive=rel_err(db, da),
                    drive_fields=max(rel_err(db[f], da[f]) for f in range(4)),
                    eu=abs(eb[2] - ea[2]) / abs(ea[2]), eb=abs(eb[3] - ea[3]) / abs(ea[3]), state=rel_err(lb, la))
        print(f"step {n}: " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
        for k in ("mf", "drive", "drive_fields", "eu", "eb"):
            assert errs[k] < 1e-6, (n, k, errs[k])
    a.close()
    b.close()


def test_mixed_runs_comparable_modes_in_fp64():
    """With the FP64 fraction below every mode's share of the largest amplitude
    (a flat spectrum), every nonzero mode takes the FP64 companion: the mixed
    run equals FP64 to rounding."""
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import PREC_MIXED, QLParams
    p = QLParams(NX=16, NY=16, NZ=64, dt=0.005, nu=1 / 8000, eta=1 / 8000)
    lin, mf = initial_state(p, kdecay=1e9)
    a, b = _ctx(p, lin, mf), _ctx(p.replace(precision=PREC_MIXED), lin, mf)
    assert b.mixed_split(1e-6) == p.nxy - 1  # all but the (0, 0) mode
    ta, tb = a.steps(30, 0.0), b.steps(30, 0.0)
    assert ta == tb
    (la, ma, _), (lb, mb, _) = a.get_state(), b.get_state()
    assert rel_err(mb, ma) < 1e-12 and rel_err(lb, la) < 1e-12
    assert rel_err(b.drive(), a.drive()) < 1e-12


@pytest.mark.parametrize("NZ", [64, 256])
def test_mixed_repeatable_and_graph_equals_eager(NZ):
    """One launch shape repeats bit for bit, and graph replay equals host-driven
    steps (ql_step + ql_remap), with the precision split deciding per substep."""
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import PREC_MIXED, QLParams
    p = QLParams(NX=16, NY=16, NZ=NZ, dt=0.005, nu=1 / 8000, eta=1 / 8000, Ly=1.0, precision=PREC_MIXED)
    lin, mf = initial_state(p)
    runs = []
    for _ in range(2):
        c = _ctx(p, lin, mf)
        t = c.steps(20, 0.0)
        runs.append((t, *c.get_state(), c.drive()))
        c.close()
    for x, y in zip(runs[0], runs[1]):
        assert np.array_equal(x, y)
    c = _ctx(p, lin, mf)
    t = 0.0
    for _ in range(20):
        dt = c.step(t)
        t += dt
        c.remap(p.q * t)
    assert t == runs[0][0]
    for x, y in zip(c.get_state(), runs[0][1:4]):
        assert np.array_equal(x, y)
    c.close()


def test_mixed_split_count_and_arguments():
    from ql_dns_amd.capi import QLContext, QLError
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import PREC_MIXED, QLParams
    p = QLParams(NX=16, NY=16, NZ=64, dt=0.005, precision=PREC_MIXED)
    lin, mf = initial_state(p)
    c = _ctx(p, lin, mf)
    n_default = c.mixed_split()
    assert 1 < n_default < p.nxy - 1          # a strict subset
    assert c.mixed_split(1.0) == 1            # the largest mode only
    assert c.mixed_split(1e-300) == p.nxy - 1
    assert c.mixed_split(1.5) == 0            # no mode reaches it: all FP32
    with pytest.raises(QLError):
        c.mixed_split(float("inf"))
    c.close()
    f = QLContext(p.replace(precision=0))
    with pytest.raises(QLError):
        f.mixed_split(0.1)
    f.close()


def test_mixed_without_split_runs_all_fp32():
    """tau > 1 switches the FP64 companion off: the steps then differ from
    the split run's, stay close to FP64 over a short run, and repeat bitwise;
    tau back at its default restores the split run's bits."""
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import PREC_MIXED, QLParams
    p = QLParams(NX=16, NY=16, NZ=64, dt=0.005, precision=PREC_MIXED)
    lin, mf = initial_state(p)

    def run(tau):
        c = _ctx(p, lin, mf)
        if tau:
            c.mixed_split(tau)
        t = c.steps(40, 0.0)
        out = c.get_state()
        c.close()
        return t, out

    t_split, s_split = run(None)
    t_off, s_off = run(2.0)
    t_off2, s_off2 = run(2.0)
    t_back, s_back = run(0.03)
    f = _ctx(p.replace(precision=0), lin, mf)
    f.steps(40, 0.0)
    s64 = f.get_state()
    f.close()
    assert t_split == t_off == t_back
    assert all(np.array_equal(x, y) for x, y in zip(s_off, s_off2))
    assert all(np.array_equal(x, y) for x, y in zip(s_split, s_back))
    assert not np.array_equal(s_off[0], s_split[0])
    assert rel_err(s_off[1], s64[1]) < 1e-6 and rel_err(s_off[0], s64[0]) < 1e-6

