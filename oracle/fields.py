"""Host field solve of the oracle (numpy, FP64).  TEST INFRASTRUCTURE ONLY.

Restates, for one periodic rank in S-units (eps0 = 1, dr = 1):
  * the periodic spectral Poisson solve of src/field/poisson/poissn.F90:360-410
    with the 7-point Laplacian eigenvalues kmod of
    src/field/poisson/inifft.F90:390-480 (zero mode dropped);
  * E = -grad(phi) on the staggered Yee points, src/field/esfild.F90:351-367;
  * the Yee updates bfield.F90:38-52 (half step) and efield.F90:55-72.

FFTW's r2r round-off is not reproducible here (parity unpinned against FFTW);
the solve is pinned by the single-Fourier-mode closed form and by the
reference's own poichk residual (poissn.F90:936-944) in tests.

Arrays: compact periodic fields are numpy arrays indexed [k, j, i] (Fortran
(i, j, k) column-major order); halo arrays are eb[k+1, j+1, i+1, c] for the
Fortran eb(6, -1:n+1, -1:n+1, -1:n+1).
"""
import numpy as np


def kmod(n):
    """-Laplacian eigenvalue per FFT index: (2 sin(pi i / n))^2, inifft.F90:394-404."""
    i = np.arange(n)
    return (2.0 * np.sin(np.pi * i / n)) ** 2


def poisson_periodic(rho):
    """phi with -lap7(phi) = rho on a periodic grid (rho[k, j, i], nodes 0..n-1)."""
    nz, ny, nx = rho.shape
    div = (kmod(nz)[:, None, None] + kmod(ny)[None, :, None] + kmod(nx)[None, None, :])
    rh = np.fft.fftn(rho)
    with np.errstate(divide="ignore", invalid="ignore"):
        ph = np.where(div != 0.0, rh / np.where(div != 0.0, div, 1.0), 0.0)
    return np.real(np.fft.ifftn(ph))


def lap7_residual(rho, phi):
    """poichk, poissn.F90:936-944: rho + (sum of 6 neighbours - 6 phi)."""
    lap = (np.roll(phi, 1, 2) + np.roll(phi, -1, 2) + np.roll(phi, 1, 1) + np.roll(phi, -1, 1)
           + np.roll(phi, 1, 0) + np.roll(phi, -1, 0) - 6.0 * phi)
    return rho + lap


def halo_wrap(a, lo=1, hi=1):
    """Periodic halo extension of a compact [k, j, i, ...] array."""
    pad = [(lo, hi)] * 3 + [(0, 0)] * (a.ndim - 3)
    return np.pad(a, pad, mode="wrap")


def eb_from_phi(phi, bfield=None):
    """esfld2 (esfild.F90:351-367): eb(EX,i) = phi(i) - phi(i+1) etc., returned as
    the halo array eb[-1:n+1]^3 with E zero where esfld2 leaves it zero.
    bfield: optional compact [k, j, i, 3] staggered B copied into the halo array."""
    nz, ny, nx = phi.shape
    ph = halo_wrap(phi, 1, 2)  # phi(-1 : n+1)
    eb = np.zeros((nz + 3, ny + 3, nx + 3, 6))
    # EX over i = -1..n, j = 0..n, k = 0..n ; arrays offset by +1
    eb[1:nz + 2, 1:ny + 2, 0:nx + 2, 0] = ph[1:nz + 2, 1:ny + 2, 0:nx + 2] - ph[1:nz + 2, 1:ny + 2, 1:nx + 3]
    eb[1:nz + 2, 0:ny + 2, 1:nx + 2, 1] = ph[1:nz + 2, 0:ny + 2, 1:nx + 2] - ph[1:nz + 2, 1:ny + 3, 1:nx + 2]
    eb[0:nz + 2, 1:ny + 2, 1:nx + 2, 2] = ph[0:nz + 2, 1:ny + 2, 1:nx + 2] - ph[1:nz + 3, 1:ny + 2, 1:nx + 2]
    if bfield is not None:
        eb[..., 3:6] = halo_wrap(bfield, 1, 2)
    return eb


def compact_from_halo(eb):
    """Interior [0:n) part of a halo array eb[-1:n+1]."""
    return eb[1:-2, 1:-2, 1:-2]


def bfield_half(ebc):
    """bfield.F90:38-52 on a compact periodic staggered array ebc[k, j, i, 6]."""
    ex, ey, ez = ebc[..., 0], ebc[..., 1], ebc[..., 2]
    out = ebc.copy()
    out[..., 3] = ebc[..., 3] + np.roll(ey, -1, 0) - ey - np.roll(ez, -1, 1) + ez
    out[..., 4] = ebc[..., 4] + np.roll(ez, -1, 2) - ez - np.roll(ex, -1, 0) + ex
    out[..., 5] = ebc[..., 5] + np.roll(ex, -1, 1) - ex - np.roll(ey, -1, 2) + ey
    return out


def efield_full(ebc, aj, tcs, dlt=2.0):
    """efield.F90:55-72 (compact periodic): E += tcs curl B - dlt J."""
    bx, by, bz = ebc[..., 3], ebc[..., 4], ebc[..., 5]
    out = ebc.copy()
    out[..., 0] = ebc[..., 0] + tcs * (bz - np.roll(bz, 1, 1) - by + np.roll(by, 1, 0)) - dlt * aj[..., 0]
    out[..., 1] = ebc[..., 1] + tcs * (bx - np.roll(bx, 1, 0) - bz + np.roll(bz, 1, 2)) - dlt * aj[..., 1]
    out[..., 2] = ebc[..., 2] + tcs * (by - np.roll(by, 1, 2) - bx + np.roll(bx, 1, 1)) - dlt * aj[..., 2]
    return out
