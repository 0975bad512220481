"""The kz = 0 slot of a ky = 0 mode, which the (ux, zeta) representation leaves
unconstrained (uy = uz = 0 there, Models/MHD_FullQlin.cpp:220-221): a nonzero
ux or bx in it is a compressive component that the reference keeps frozen (its
RHS is zero, :283).

* MHD_BQlin forms dz uz on the z grid from the ux and uy lines; the kernels
  add back the z-constant i kx ux(kz = 0) that the reference's own dz uz line
  (MHD_BQlin.cpp:195) does not hold, so a state with that slot filled matches
  the oracle to the usual 1e-10, in both kernel shapes.
* MHD_FullQlin and HD_fullU form their stresses as z-derivatives, exact for a
  divergence-free state only, so ql_set_state refuses a state with more than
  rounding (1e-12 of the largest field value) in that slot.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _planted(p, level=0.3, seed=21):
    from ql_dns_amd.ic import initial_state
    lin, mf = initial_state(p)
    ny = p.NY // 2
    rng = np.random.default_rng(seed)
    big = np.abs(lin).max()
    for i in range(0, lin.shape[0], ny):   # ky = 0 modes: mode index j + ny*i' with j = 0
        for f in range(0, lin.shape[1], 2):  # ux and bx
            lin[i, f, 0] = level * big * (rng.standard_normal() + 1j * rng.standard_normal())
    return lin, mf


@pytest.mark.parametrize("line,NZ", [(64, 128), (256, 128), (64, 32), (128, 64)])
@pytest.mark.parametrize("integrator", [1, 0])
def test_bqlin_compressive_slot_matches_oracle(line, NZ, integrator):
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.params import QLParams
    from oracle.oracle import Oracle
    p = QLParams(model=1, integrator=integrator, NX=8, NY=8, NZ=NZ)
    lin, mf = _planted(p)
    o = Oracle.from_params(p, nthreads=8)
    g = QLContext(p)
    g.set_config(NZ, line, 512)
    for x in (o, g):
        x.set_state(lin, mf)
        x.linear_ops_init(0.0)
    ro, rg = o.rhs(0.21, 0.013), g.rhs(0.21, 0.013)
    for a, b in zip(rg, ro):
        assert rel_err(a, b) < TOL
    assert rel_err(g.drive(), o.drive()) < TOL
    to, tg = o.steps(20, 0.0), g.steps(20, 0.0)
    assert tg == to
    (lo, mo, ko), (lg, mg, kg) = o.get_state(), g.get_state()
    assert np.array_equal(kg, ko)
    assert rel_err(mg, mo) < TOL and rel_err(lg, lo) < TOL
    # the slot is still filled: the comparison above covered it
    assert np.abs(lo[0, 0, 0]) > 1e-3 * np.abs(lo).max()


@pytest.mark.parametrize("model", [0, 2])
def test_div_form_models_refuse_a_compressive_state(model):
    from ql_dns_amd.capi import QLContext, QLError
    from ql_dns_amd.params import QLParams
    p = QLParams(model=model, NX=8, NY=8, NZ=32)
    lin, mf = _planted(p)
    g = QLContext(p)
    with pytest.raises(QLError) as e:
        g.set_state(lin, mf)
    assert e.value.code == -1 and "kz = 0" in str(e.value)
    # rounding-level content is accepted
    lin2, _ = _planted(p, level=1e-14)
    g.set_state(lin2, mf)
    g.close()
