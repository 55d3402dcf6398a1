"""Host field solve that runs between the GPU particle steps (numpy, FP64).

The reference keeps its field solver on the host (FFTW + MPI,
src/field/poisson/poissn.F90); the GPU path exchanges only grid arrays with
it.  This module is that host side for one periodic rank, in S-units:

  * periodic Poisson -lap7(phi) = rho with the 7-point Laplacian eigenvalues
    (2 sin(pi k / n))^2 (inifft.F90:390-480), zero mode dropped
    (poissn.F90:398-410);
  * E = -grad(phi) on the staggered points (esfild.F90:351-367);
  * Yee half-step B and full-step E (bfield.F90:38-52, efield.F90:55-72);
  * the uniform neutralising background of chgntr (chgntr.F90:69-80).

Compact periodic arrays are [k, j, i] numpy arrays of nodes / cells 0..n-1.
"""
import numpy as np


def _eig(n):
    return (2.0 * np.sin(np.pi * np.arange(n) / n)) ** 2


def poisson(rho):
    nz, ny, nx = rho.shape
    kz = _eig(nz)[:, None, None]
    ky = _eig(ny)[None, :, None]
    kx = _eig(nx)[: nx // 2 + 1][None, None, :]
    div = kz + ky + kx
    spec = np.fft.rfftn(rho)
    div[0, 0, 0] = 1.0
    spec /= div
    spec[0, 0, 0] = 0.0
    return np.fft.irfftn(spec, s=rho.shape, axes=(0, 1, 2))


def efield(phi):
    """Staggered E components [k, j, i, 3]: E_x(i+1/2) = phi(i) - phi(i+1)."""
    e = np.empty(phi.shape + (3,))
    e[..., 0] = phi - np.roll(phi, -1, 2)
    e[..., 1] = phi - np.roll(phi, -1, 1)
    e[..., 2] = phi - np.roll(phi, -1, 0)
    return e


def eb_halo(e, b=None):
    """Fortran eb(6, -1:n+1)^3 as numpy [k+1, j+1, i+1, 6] from compact staggered
    E (and B) with periodic halos."""
    ebc = np.zeros(e.shape[:3] + (6,))
    ebc[..., 0:3] = e
    if b is not None:
        ebc[..., 3:6] = b
    return np.pad(ebc, [(1, 2), (1, 2), (1, 2), (0, 0)], mode="wrap")


def background(q, npin, n):
    """chgntr's uniform neutraliser per node: -sum(q npin) / (nx ny nz)
    (zero when below 1e-10 after the 1e-3 sub-point scaling, chgntr.F90:79)."""
    rion = -float(np.sum(np.asarray(q) * np.asarray(npin, dtype=np.float64)))
    rion = rion / (n[0] * n[1] * n[2]) * 1.0e-3
    if abs(rion) < 1.0e-10:
        return 0.0
    return rion * 1000.0


def b_half(ebc):
    ex, ey, ez = ebc[..., 0], ebc[..., 1], ebc[..., 2]
    out = ebc.copy()
    out[..., 3] += np.roll(ey, -1, 0) - ey - np.roll(ez, -1, 1) + ez
    out[..., 4] += np.roll(ez, -1, 2) - ez - np.roll(ex, -1, 0) + ex
    out[..., 5] += np.roll(ex, -1, 1) - ex - np.roll(ey, -1, 2) + ey
    return out


def e_full(ebc, aj, tcs):
    bx, by, bz = ebc[..., 3], ebc[..., 4], ebc[..., 5]
    out = ebc.copy()
    out[..., 0] += tcs * (bz - np.roll(bz, 1, 1) - by + np.roll(by, 1, 0)) - 2.0 * aj[..., 0]
    out[..., 1] += tcs * (bx - np.roll(bx, 1, 0) - bz + np.roll(bz, 1, 2)) - 2.0 * aj[..., 1]
    out[..., 2] += tcs * (by - np.roll(by, 1, 2) - bx + np.roll(bx, 1, 1)) - 2.0 * aj[..., 2]
    return out


def rhoupd(rho, aj):
    """rhoupd (src/field/rhoupd.F90:40-49): rho -= 2 div aj on the nodes
    0..n of a rho[k, j, i] over -1..n+1 (the ghost planes are left as they
    are), aj[k, j, i, 3] the folded current on the periodic nodes."""
    jx, jy, jz = aj[..., 0], aj[..., 1], aj[..., 2]
    div = (jx - np.roll(jx, 1, 2)) + (jy - np.roll(jy, 1, 1)) + (jz - np.roll(jz, 1, 0))
    nz, ny, nx = div.shape
    d = np.pad(div, ((0, 1), (0, 1), (0, 1)), mode="wrap")   # nodes 0..n, node n = node 0
    out = rho.copy()
    out[1:nz + 2, 1:ny + 2, 1:nx + 2] -= d * 2.0
    return out


def field_energy(ebc, cs):
    """energy.F90:52-70 over the staggered cells 0..n-1."""
    vol = float(np.prod(ebc.shape[:3]))
    te = float(np.sum(ebc[..., 0:3] ** 2))
    tb = float(np.sum(ebc[..., 3:6] ** 2))
    return 0.5 * te / vol, 0.5 * cs * tb / vol
