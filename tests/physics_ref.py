"""Independent restatements used to pin the oracle (CPU, numpy).

Nothing here follows the reference's code: each function states the physics or
mathematics directly.

* Cartesian shearing-box QL equations for one (kx, ky) Fourier mode, with
  explicit pressure: for background shear U_y = -q x (so k_x(t) = kx0 + q t ky),
  Coriolis 2 Omega z-hat, mean fields U(z) = (Ux, Uy, 0), B(z) = (Bx, By, 0),
  vertical field B0z and incompressible fluctuations u, b:
      du/dt = F - k (k.F + q ky u_x)/|k|^2,
      F     = -(U.grad)u - u_z dU/dz + (B.grad)b + b_z dB/dz
              + (2 Omega u_y, (q - 2 Omega) u_x, 0) + i kz B0z b
      db/dt = -(U.grad)b - u_z dB/dz + (B.grad)u + b_z dU/dz
              + (0, -q b_x, 0) + i kz B0z u
  Products are formed on the NZ grid from band-limited profiles and truncated
  back to |kz| <= nzl (the 2/3 rule; exact in the retained band).
* The (ux, zeta) representation: zeta = i (kz u_y - ky u_z) = -(curl u)_x,
  and its inverse for divergence-free u, obtained by solving
  ky u_y + kz u_z = -kx u_x,  kz u_y - ky u_z = -i zeta.
* Full 3-D real fields on an NX x NY x NZ grid assembled from the stored
  half-plane spectrum (ky >= 0; the ky < 0 half is the complex conjugate).
"""
import numpy as np


def kz_slots(nz):
    nzl = (nz - 1) // 2
    return np.concatenate([np.arange(nzl + 1), np.arange(-nzl, 0)])


def pad_index(NZ, nz):
    """grid index of each retained kz slot"""
    return kz_slots(nz) % NZ


def to_grid(NZ, spec):
    """band-limited complex profile on the NZ grid: sum_m spec_m e^{i k_m z}/NZ"""
    nz = spec.shape[-1]
    full = np.zeros(spec.shape[:-1] + (NZ,), dtype=np.complex128)
    full[..., pad_index(NZ, nz)] = spec
    return np.fft.ifft(full, axis=-1)


def from_grid(nz, prof):
    """retained spectrum of a grid profile (inverse of to_grid on the band)"""
    NZ = prof.shape[-1]
    return np.fft.fft(prof, axis=-1)[..., pad_index(NZ, nz)]


def cartesian_from_uzeta(kx, ky, kz, ux, ze):
    """(u_x, u_y, u_z) from (u_x, zeta) for divergence-free u (per kz slot)."""
    k2p = ky * ky + kz * kz
    with np.errstate(invalid="ignore", divide="ignore"):
        uy = np.where(k2p > 0, -(kx * ky * ux + 1j * kz * ze) / np.where(k2p > 0, k2p, 1), 0)
        uz = np.where(k2p > 0, (1j * ky * ze - kx * kz * ux) / np.where(k2p > 0, k2p, 1), 0)
    return np.stack([ux, uy, uz], axis=-1)


def zeta_of(ky, kz, v):
    return 1j * (kz * v[..., 1] - ky * v[..., 2])


def mean_profiles(NZ, mf, kz):
    """real-space mean fields and their z-derivatives on the grid"""
    prof = to_grid(NZ, mf).real
    dprof = to_grid(NZ, 1j * kz[None, :] * mf).real
    return prof, dprof


def cartesian_rhs(NZ, kx, ky, kz, u, b, mf, q, Omega, B0z, model=0):
    """du/dt, db/dt (Cartesian, no dissipation) of one mode; mf = (Bx, By, Ux, Uy)
    spectra (model 0) or (Bx, By) (model 1: fluctuations see By only)."""
    nz = kz.size
    prof, dprof = mean_profiles(NZ, mf, kz)
    Bx, By = prof[0], prof[1]
    dBx, dBy = dprof[0], dprof[1]
    if model == 0:
        Ux, Uy, dUx, dUy = prof[2], prof[3], dprof[2], dprof[3]
    else:
        Bx = dBx = np.zeros(NZ)
        Ux = Uy = dUx = dUy = np.zeros(NZ)
    ug = to_grid(NZ, u.T)  # (3, NZ)
    bg = to_grid(NZ, b.T)
    Ugrad = 1j * (kx * Ux + ky * Uy)
    Bgrad = 1j * (kx * Bx + ky * By)
    Au = -Ugrad * ug + Bgrad * bg
    Au[0] += -ug[2] * dUx + bg[2] * dBx
    Au[1] += -ug[2] * dUy + bg[2] * dBy
    Ab = -Ugrad * bg + Bgrad * ug
    Ab[0] += -ug[2] * dBx + bg[2] * dUx
    Ab[1] += -ug[2] * dBy + bg[2] * dUy
    Fu = from_grid(nz, Au).T  # (nz, 3)
    Fb = from_grid(nz, Ab).T
    Fu[:, 0] += 2 * Omega * u[:, 1]
    Fu[:, 1] += (q - 2 * Omega) * u[:, 0]
    Fu += 1j * kz[:, None] * B0z * b
    Fb[:, 1] += -q * b[:, 0]
    Fb += 1j * kz[:, None] * B0z * u
    K = np.stack([np.full(nz, kx), np.full(nz, ky), kz], axis=-1)
    k2 = np.sum(K * K, axis=-1)
    kF = np.sum(K * Fu, axis=-1) + q * ky * u[:, 0]
    with np.errstate(invalid="ignore", divide="ignore"):
        du = Fu - K * (kF / np.where(k2 > 0, k2, 1))[:, None]
    return du, Fb


def spectrum_3d(NX, NY, NZ, kxi, kyi, nz, comp):
    """complex 3-D coefficient array from stored half-plane values comp[mode, kz slot]
    (already divided by the storage scale)."""
    C = np.zeros((NX, NY, NZ), dtype=np.complex128)
    kzs = kz_slots(nz)
    for n in range(len(kxi)):
        a, j = int(kxi[n]), int(kyi[n])
        C[a % NX, j % NY, kzs % NZ] += comp[n]
        if j > 0:
            C[(-a) % NX, (-j) % NY, (-kzs) % NZ] += np.conj(comp[n])
    return C


def real_field(C):
    N = C.size
    return np.fft.ifftn(C) * N
