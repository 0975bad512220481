"""Writes the golden fixtures of the 16x16x32 box and of the 64x128x128 box
from the CPU oracle.

    python tests/golden/make_golden.py           # all fixtures
    python tests/golden/make_golden.py box64     # only the 64x128x128 one

Fixtures (numpy .npz, no pickles): the seeded initial state, one Model::rhs
evaluation, and the state / drives / diagnostics after 1, 10 and 100 steps of
the reference main loop, for MHD_FullQlin with RK3CN_Elin (full state kept)
and, with reduced content, MHD_FullQlin with RK3CN and MHD_BQlin with
RK3CN_Elin.  The oracle itself is pinned by tests/test_oracle_pins.py.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ql_dns_amd.ic import initial_state  # noqa: E402
from ql_dns_amd.params import workload  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

VARIANTS = {"fullqlin_elin": (0, 1, True), "fullqlin_cn": (0, 0, False), "bqlin_elin": (1, 1, False)}


def make(name, model, integrator, full):
    p = workload("box16x16x32", model=model, integrator=integrator)
    lin, mf = initial_state(p)
    o = Oracle.from_params(p, nthreads=4)
    o.set_state(lin, mf)
    o.linear_ops_init(0.0)
    out = dict(NX=p.NX, NY=p.NY, NZ=p.NZ, seed=p.seed, model=model, integrator=integrator, dt=p.dt)
    r_lin, r_mf, r_L = o.rhs(0.0, 0.0053)
    if full:
        out.update(ic_lin=lin, ic_mf=mf, rhs_lin=r_lin, rhs_linop=r_L)
    out["rhs_mf"], out["rhs_drive"] = r_mf, o.drive()
    t, done = 0.0, 0
    for n in (1, 10, 100):
        t = o.steps(n - done, t)
        done = n
        l_, m_, k_ = o.get_state()
        out[f"t{n}"] = t
        out[f"mf{n}"], out[f"drive{n}"], out[f"diag{n}"], out[f"kx{n}"] = m_, o.drive(), o.diagnostics(t), k_
        out[f"linnorm{n}"] = np.abs(l_).sum(axis=(1, 2))
        if full and n == 100:
            out["lin100"], out["mfreal100"] = l_, o.mf_real()
    path = os.path.join(HERE, f"box16x16x32_{name}.npz")
    np.savez_compressed(path, **out)
    return path


BOX64_MARKS = (1, 100, 1000)
BOX64_MODES = 16  # modes kept in full: the largest at the start


def make_box64(nthreads=8):
    """The 64x128x128 production box (MHD_FullQlin, RK3CN_Elin, dt 0.005, remap
    on) over its configured 1000 steps: at steps 1, 100 and 1000 the mean
    fields, drives, the 18 diagnostics, kx0 and the per-mode, per-field sum of
    |value| over kz; at step 1000 also the full state of the 16 modes that are
    largest at the start.  Checksums of the initial state pin the generator."""
    p = workload("box64x128x128")
    lin, mf = initial_state(p)
    o = Oracle.from_params(p, nthreads=nthreads)
    o.set_state(lin, mf)
    o.linear_ops_init(0.0)
    keep = np.argsort(-np.abs(lin).max(axis=(1, 2)), kind="stable")[:BOX64_MODES]
    out = dict(NX=p.NX, NY=p.NY, NZ=p.NZ, seed=p.seed, dt=p.dt, ic_mf=mf, ic_linnorm=np.abs(lin).sum(axis=2),
               keep=keep)
    t, done = 0.0, 0
    for n in BOX64_MARKS:
        t = o.steps(n - done, t)
        done = n
        l_, m_, k_ = o.get_state()
        out[f"t{n}"] = t
        out[f"mf{n}"], out[f"drive{n}"], out[f"diag{n}"], out[f"kx{n}"] = m_, o.drive(), o.diagnostics(t), k_
        out[f"linnorm{n}"] = np.abs(l_).sum(axis=2)
    out["lin_keep"] = l_[keep]
    path = os.path.join(HERE, "box64x128x128_fullqlin_elin_1000.npz")
    np.savez_compressed(path, **out)
    return path


if __name__ == "__main__":
    if sys.argv[1:] != ["box64"]:
        for name, (m, i, full) in VARIANTS.items():
            print(make(name, m, i, full))
    print(make_box64())
