"""Run parameters of the quasilinear (QL) shearing-box step and the named workloads.

`QLParams` carries what the reference reads from its ``.DSSinput`` file for this
path (``Auxiliary/Input_parameters.cpp:68-138``) and what ``main.cpp`` hard-codes
(the model factory ``main.cpp:58-74`` and the integrator ``main.cpp:95``).
Re and Rm map to nu = 1/Re and eta = 1/Rm as the reference's input writer does
(``Python/WriteMRIDNSinput.py:22-23``).
"""
from __future__ import annotations

import dataclasses
import math

MODEL_MHD_FULLQLIN = 0
MODEL_MHD_BQLIN = 1
MODEL_HD_FULLU = 2
INT_RK3CN = 0
INT_RK3CN_ELIN = 1
PREC_FP64 = 0
PREC_MIXED = 1

MODEL_NAMES = {"MHD_FullQlin": MODEL_MHD_FULLQLIN, "MHD_BQlin": MODEL_MHD_BQLIN, "HD_fullU": MODEL_HD_FULLU}
# the oracle's model ids (oracle/ql_oracle.h)
_ORACLE_MODEL = {MODEL_MHD_FULLQLIN: 0, MODEL_MHD_BQLIN: 1, MODEL_HD_FULLU: 4}
INTEGRATOR_NAMES = {"RK3CN": INT_RK3CN, "RK3CN_Elin": INT_RK3CN_ELIN}


@dataclasses.dataclass
class QLParams:
    NX: int = 16
    NY: int = 16
    NZ: int = 32
    Lx: float = 1.0
    Ly: float = 4.0
    Lz: float = 1.0
    q: float = 1.5
    omega: float = 1.0
    nu: float = 1.0 / 2000
    eta: float = 1.0 / 2000
    B0z: float = 0.0
    viscosity_order: int = 2
    quasilinear: int = 1
    remap: int = 1
    dt: float = 0.01      # < 0: variable steps of at most -dt
    CFL: float = 1.0      # CFL number of variable steps (Input_parameters.cpp:77)
    model: int = MODEL_MHD_FULLQLIN
    integrator: int = INT_RK3CN_ELIN
    precision: int = PREC_FP64
    seed: int = 0x5EED0000
    nsteps: int = 100

    # sizes fixed by the reference's conventions (Models/Model.h:27-38)
    @property
    def nxy(self) -> int:
        return (self.NX - 1) * (self.NY // 2)

    @property
    def nz(self) -> int:
        return 1 + 2 * int(self.NZ / 3.0)

    @property
    def half(self) -> int:
        return (self.nz - 1) // 2 + 1

    @property
    def nMF(self) -> int:
        return 4 if self.model == MODEL_MHD_FULLQLIN else 2

    @property
    def nLin(self) -> int:
        return 2 if self.model == MODEL_HD_FULLU else 4

    @property
    def state_bytes(self) -> int:
        """FP64 fluctuation state size S = nxy * nLin fields * nz * 16 B."""
        return self.nxy * self.nLin * self.nz * 16

    def replace(self, **kw) -> "QLParams":
        return dataclasses.replace(self, **kw)

    def oracle_dict(self) -> dict:
        return dict(model=_ORACLE_MODEL[self.model], integrator=self.integrator, NX=self.NX, NY=self.NY,
                    NZ=self.NZ, L=(self.Lx, self.Ly, self.Lz), q=self.q, omega=self.omega,
                    nu=self.nu, eta=self.eta, B0z=self.B0z,
                    viscosity_order=self.viscosity_order, quasilinear=self.quasilinear,
                    remap=self.remap, dt=self.dt, CFL=self.CFL)


def _re(Re: float, Rm: float) -> dict:
    return dict(nu=1.0 / Re, eta=1.0 / Rm)


# The named workloads: shearing box Lx, Ly, Lz = 1, 4, 1, Omega = 1, q = 1.5.
WORKLOADS = {
    # CPU-runnable plain QL MHD box, 16x16x32, Re = Rm = 2000, dt = 0.01, 100 steps
    "box16x16x32": QLParams(NX=16, NY=16, NZ=32, **_re(2000, 2000), dt=0.01, nsteps=100,
                            seed=0x5EED0000),
    # production-size single-GPU box, 64x128x128, Re = Rm = 8000, dt = 0.005, 1000 steps
    "box64x128x128": QLParams(NX=64, NY=128, NZ=128, **_re(8000, 8000), dt=0.005,
                              nsteps=1000, seed=0x5EED0001),
    # sensitivity box, 32x32x64, Re = Rm = 4000, 200 steps
    "box32x32x64": QLParams(NX=32, NY=32, NZ=64, **_re(4000, 4000), dt=0.01, nsteps=200,
                            seed=0x5EED0002),
    # large mixed-precision box, 128x128x256, Re = Rm = 16000, 2000 steps
    "box128x128x256": QLParams(NX=128, NY=128, NZ=256, **_re(16000, 16000), dt=0.0025,
                               nsteps=2000, seed=0x5EED0004),
}


def workload(name: str, **overrides) -> QLParams:
    return WORKLOADS[name].replace(**overrides)


def ensemble(n: int = 256, seed: int = 0x5EED0003, **overrides) -> list:
    """A parameter-study ensemble of n independent runs of the 16x16x32 box's
    physics (Lx, Ly, Lz = 1, 4, 1, Omega = 1, Re = 2000, dt = 0.01): Rm
    log-uniform in [500, 20000], q uniform in [1.0, 1.8], NZ from {32, 64, 128},
    NX and NY each from {8, 16, 32}; run r's initial state uses seed + 1 + r."""
    import numpy as np
    rng = np.random.default_rng(seed)
    base = WORKLOADS["box16x16x32"].replace(nsteps=500, **overrides)
    runs = []
    for r in range(n):
        Rm = float(np.exp(rng.uniform(np.log(500.0), np.log(20000.0))))
        q = float(rng.uniform(1.0, 1.8))
        NZ = int(rng.choice([32, 64, 128]))
        NX = int(rng.choice([8, 16, 32]))
        NY = int(rng.choice([8, 16, 32]))
        runs.append(base.replace(NX=NX, NY=NY, NZ=NZ, q=q, eta=1.0 / Rm, seed=seed + 1 + r))
    return runs


def kmax(p: QLParams) -> float:
    """MHD_FullQlin.cpp:92-97: largest dealiased wavenumber over x, y, z."""
    L = (p.Lx, p.Ly, p.Lz)
    N = (p.NX, p.NY, p.NZ)
    fac = (2.0, 2.0, 3.0)
    km = 0.0
    for i in range(3):
        v = 2 * math.pi / L[i] * (N[i] / fac[i])
        km = max(km, v)
    return km
