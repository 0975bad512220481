"""Seeded initial states in the reference's storage conventions.

The reference draws its initial fluctuations from a clock-seeded Boost stream
(``solution.cpp:60-77`` via ``DrivingNoise``), which cannot be reproduced, so
workloads here use a counter-based generator instead: splitmix64 hashing of
(seed, stream, counter) followed by Box-Muller.  Every value depends only on
its own (seed, stream, counter), so any subset can be regenerated.

Conventions (all from the reference):
  * mode order ``j + ny*i'`` with kx indices 0..nx-1 then -(nx-1)..-1 and
    ky indices 0..NY/2-1 (``Auxiliary/Kdata.cpp:28-46``);
  * kz order (0..nzl, -nzl..-1) (``Kdata.cpp:74-78``);
  * fluctuation variables (ux, zeta, bx, eta) with
    uy = (ky kx ux + i kz zeta)/(-(ky^2+kz^2)), uz = (kx kz ux - i ky zeta)/(-(ky^2+kz^2))
    (``Models/MHD_FullQlin.cpp:220-221`` with k stored as i*k), which inverts to
    zeta = i (kz uy - ky uz) for a divergence-free u;
  * fluctuations stored at NX*NY*NZ times the Fourier coefficient and mean
    fields at NZ times it (``MHD_FullQlin.cpp:63-67,572,587``,
    ``solution.cpp:159-161``).
"""
from __future__ import annotations

import numpy as np

from .params import QLParams

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def _uniform(seed: int, stream: int, counter: np.ndarray) -> np.ndarray:
    """Uniform in (0, 1], 53 bits, from (seed, stream, counter)."""
    key = _splitmix64(np.uint64(seed & 0xFFFFFFFFFFFFFFFF) ^ _splitmix64(np.uint64(stream)))
    z = _splitmix64(key ^ np.asarray(counter, dtype=np.uint64))
    return ((z >> np.uint64(11)).astype(np.float64) + 1.0) * (2.0 ** -53)


def normals(seed: int, stream: int, counter: np.ndarray) -> np.ndarray:
    """Standard normal deviates: Box-Muller on counters 2c and 2c+1."""
    c = np.asarray(counter, dtype=np.uint64)
    u1 = _uniform(seed, stream, c * np.uint64(2))
    u2 = _uniform(seed, stream, c * np.uint64(2) + np.uint64(1))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def mode_tables(p: QLParams):
    """kx0, ky (physical), kx_index, ky_index per mode; kz per kz slot."""
    nx, ny = p.NX // 2, p.NY // 2
    kxL, kyL, kzL = 2 * np.pi / p.Lx, 2 * np.pi / p.Ly, 2 * np.pi / p.Lz
    ix = np.concatenate([np.arange(nx), np.arange(nx + 1, 2 * nx) - 2 * nx])
    kx_index = np.repeat(ix, ny)
    ky_index = np.tile(np.arange(ny), ix.size)
    kx0 = kx_index * kxL
    ky = ky_index * kyL
    nzl = (p.nz - 1) // 2
    kz_index = np.concatenate([np.arange(nzl + 1), np.arange(nzl + 1, 2 * nzl + 1) - 2 * nzl - 1])
    kz = kz_index * kzL
    return kx0, ky, kx_index, ky_index, kz, kz_index


def initial_state(p: QLParams, seed: int | None = None, amp: float = 1e-2,
                  mf_amp: float = 0.1, kdecay: float = 50.0):
    """Return (lin, mf) complex128 arrays of shapes (nxy, nLin, nz) and (nMF, nz).

    Fluctuations: each Cartesian component of u and b is complex N(0,1) times
    amp*exp(-|k|^2/kdecay) (zero where |k|^2/kdecay > 300) with k = (kx0, ky, kz),
    projected divergence free,
    converted to (ux, zeta, bx, eta), scaled to the storage convention, made
    Hermitian across the ky = 0 plane (the (kx, 0, kz) mode is the conjugate of
    (-kx, 0, -kz)), with the (kx, ky) = (0, 0) mode zero.
    Mean fields: Bx, By = mf_amp * sum_{m=1..4} (a_m cos 2 pi m z/Lz + b_m sin 2 pi m z/Lz)
    with a, b ~ N(0,1); Ux = Uy = 0.  HD_fullU (no magnetic field) keeps (ux,
    zeta) and puts those two profiles in its mean fields Ux, Uy.
    """
    seed = p.seed if seed is None else seed
    nxy, nz = p.nxy, p.nz
    kx0, ky, kxi, kyi, kz, kzi = mode_tables(p)
    K = np.stack(np.broadcast_arrays(kx0[:, None], ky[:, None], kz[None, :]), axis=-1)  # (nxy,nz,3)
    k2 = np.sum(K * K, axis=-1)
    # envelopes below amp*e^-300 are set to exact zero: such amplitudes carry no
    # physical content and would only produce subnormal numbers
    env = np.where(k2 / kdecay > 300.0, 0.0, amp * np.exp(-np.minimum(k2 / kdecay, 300.0)))
    cnt = (np.arange(nxy)[:, None] * nz + np.arange(nz)[None, :]).astype(np.uint64)

    def field(stream0: int) -> np.ndarray:
        v = np.empty((nxy, nz, 3), dtype=np.complex128)
        for c in range(3):
            s = stream0 + c
            re = normals(seed, s, cnt * np.uint64(2))
            im = normals(seed, s, cnt * np.uint64(2) + np.uint64(1))
            v[..., c] = (re + 1j * im) * env
        kk = np.where(k2 > 0, k2, 1.0)
        kdotv = np.sum(K * v, axis=-1)
        v = v - K * (kdotv / kk)[..., None]
        v[k2 == 0] = 0.0
        return v

    scale = float(p.NX * p.NY * p.NZ)
    u = field(0)
    b = field(3)
    lin = np.zeros((nxy, 4, nz), dtype=np.complex128)
    kyb, kzb = ky[:, None], kz[None, :]
    lin[:, 0] = u[..., 0]
    lin[:, 1] = 1j * (kzb * u[..., 1] - kyb * u[..., 2])
    lin[:, 2] = b[..., 0]
    lin[:, 3] = 1j * (kzb * b[..., 1] - kyb * b[..., 2])
    lin *= scale
    # Hermitian pairing in the ky = 0 plane, (0, 0) mode zero
    idx = {(int(a), int(c)): n for n, (a, c) in enumerate(zip(kxi, kyi))}
    neg_kz = (-np.arange(nz)) % nz
    for n in range(nxy):
        if kyi[n] != 0:
            continue
        if kxi[n] == 0:
            lin[n] = 0.0
        elif kxi[n] < 0:
            partner = idx[(-int(kxi[n]), 0)]
            lin[n] = np.conj(lin[partner][:, neg_kz])

    lin = lin[:, :p.nLin].copy()
    mf = np.zeros((p.nMF, nz), dtype=np.complex128)
    for f in range(2):
        coeff = normals(seed, 10 + f, np.arange(2, 10, dtype=np.uint64))  # a1,b1,...,a4,b4
        for m in range(1, 5):
            a, bb = coeff[2 * (m - 1)], coeff[2 * (m - 1) + 1]
            c = p.NZ * mf_amp * (a - 1j * bb) / 2.0
            mf[f, m] = c
            mf[f, (-m) % nz] = np.conj(c)
    return lin, mf


def mf_profile_real(p: QLParams, mf: np.ndarray) -> np.ndarray:
    """Real-space mean-field profiles on the NZ grid from the stored spectrum."""
    nz, NZ = p.nz, p.NZ
    nzl = (nz - 1) // 2
    full = np.zeros((mf.shape[0], NZ), dtype=np.complex128)
    full[:, :nzl + 1] = mf[:, :nzl + 1]
    full[:, NZ - nzl:] = mf[:, nzl + 1:]
    return np.fft.ifft(full, axis=-1).real
