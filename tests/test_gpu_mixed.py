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
