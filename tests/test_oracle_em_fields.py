"""The oracle's psolve1 zigzag current and current fold pinned by the discrete
continuity equation rho_new - rho_old = -ustep div J -- the identity the
reference itself relies on in rhoupd (src/field/rhoupd.F90:40-49) -- and the
oracle's host field solve pinned by single-mode closed forms and the
reference's poichk residual (src/field/poisson/poissn.F90:936-944)."""
import numpy as np
import pytest

import oracle
from oracle import fields

PART = oracle.PARTICLE


def rand_state(rng, n, count, vth):
    p = np.zeros(count, dtype=PART)
    for d, f in enumerate(("x", "y", "z")):
        p[f] = rng.uniform(0, n[d], count)
    for f in ("vx", "vy", "vz"):
        p[f] = rng.normal(0, vth, count)
    p["spec"] = 1
    return p


def rand_eb(rng, n, amp):
    return amp * rng.normal(size=(n[2] + 3, n[1] + 3, n[0] + 3, 6))


def _continuity_residual(p, counts, q, qm, eb, n, b0=(0, 0, 0)):
    sl = np.array(n, dtype=float)
    wrk = oracle.grdcrt(eb, n, (0, 0, 0), b0)
    rho0 = oracle.charge(p, counts, q, n)
    oracle.fold_rho(rho0, n)
    aj = oracle.psolve1(p, counts, qm, q, wrk, n, 2, sl)
    oracle.fold_current(aj, n)
    rho1 = oracle.charge(p, counts, q, n)
    oracle.fold_rho(rho1, n)
    upd = rho0.copy()
    oracle.rhoupd(upd, aj, n)
    return oracle.rho_nodes(upd) - oracle.rho_nodes(rho1), oracle.rho_nodes(rho1)


@pytest.mark.parametrize("plane", ["xy", "xz", "yz"])
@pytest.mark.parametrize("vth", [0.05, 0.3])
def test_zigzag_continuity_planar(plane, vth):
    # For motion inside a coordinate plane the zigzag current is exactly
    # charge conserving (cell crossings and periodic wraps included) as long
    # as a step crosses at most one cell face per axis (|v| ustep < 1).
    n = (6, 5, 7)
    rng = np.random.default_rng(11)
    counts = [3000, 1000]
    p = rand_state(rng, n, sum(counts), vth)
    for f in ("vx", "vy", "vz"):
        p[f] = np.clip(p[f], -0.35, 0.35)
    off = {"xy": 2, "xz": 1, "yz": 0}[plane]
    p[("vx", "vy", "vz")[off]] = 0.0
    eb = rand_eb(rng, n, 0.01)    # kicks stay below 0.15: |v| ustep < 1 after the push
    eb[..., off] = 0.0            # no acceleration out of the plane
    eb[..., 3:6] = 0.0
    b0 = [0.0, 0.0, 0.0]
    b0[off] = 0.1                 # gyration inside the plane
    res, rho1 = _continuity_residual(p, counts, [0.3, -0.7], [1.0, -0.05], eb, n, tuple(b0))
    assert np.abs(res).max() < 1e-13 * np.abs(rho1).max()


def test_zigzag_cubic_residual_single_particle():
    # A 3D move inside one cell is split at the relay point (the midpoint,
    # psolve1.F90:399-415) into two half moves; each half's midpoint-rule
    # current misses the cubic term of the trilinear weight, so
    # (rho0 - ustep div J) - rho1 = s q dx dy dz / 16 at the cell's 8 nodes,
    # s = +1 at the lower corner and alternating with every unit offset.
    n = (4, 4, 4)
    p = np.zeros(1, dtype=PART)
    p["x"], p["y"], p["z"] = 1.3, 2.2, 0.4
    p["vx"], p["vy"], p["vz"] = 0.1, 0.15, 0.05   # moves 0.2, 0.3, 0.1
    res, _ = _continuity_residual(p, [1], [1.0], [1.0], np.zeros((7, 7, 7, 6)), n)
    d = 0.2 * 0.3 * 0.1 / 16.0
    for (i, j, k) in [(1, 2, 0), (2, 2, 0), (1, 3, 0), (2, 3, 0), (1, 2, 1), (2, 2, 1), (1, 3, 1), (2, 3, 1)]:
        sign = (-1) ** ((i - 1) + (j - 2) + (k - 0))
        assert abs(res[k, j, i] - sign * d) < 1e-16, (i, j, k)
    mask = np.ones_like(res, dtype=bool)
    mask[0:2, 2:4, 1:3] = False
    assert np.abs(res[mask]).max() < 1e-16


def test_psolve1_push_equals_psolve2_push():
    n = (5, 5, 5)
    rng = np.random.default_rng(12)
    p = rand_state(rng, n, 800, 0.2)
    wrk = oracle.grdcrt(rand_eb(rng, n, 0.05), n, (0.01, 0, 0), (0, 0.05, 0))
    sl = np.array(n, dtype=float)
    a = p.copy()
    b = p.copy()
    oracle.psolve2(a, [800], [1.0], wrk, n, 2, sl)
    oracle.psolve1(b, [800], [1.0], [1.0], wrk, n, 2, sl)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(a[f], b[f]), f


def test_current_of_uniform_drift():
    # a uniform beam moving +v_x: J_x = q n v_x on every x-face, J_y = J_z = 0
    n = (4, 4, 4)
    k = 4
    g = (np.arange(4 * k) + 0.5) / k
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    p = np.zeros(X.size, dtype=PART)
    p["x"], p["y"], p["z"] = X.ravel(), Y.ravel(), Z.ravel()
    p["vx"] = 0.1
    wrk = oracle.grdcrt(np.zeros((7, 7, 7, 6)), n, (0, 0, 0), (0, 0, 0))
    aj = oracle.psolve1(p, [p.size], [1.0], [1.0], wrk, n, 2, np.array(n, dtype=float))
    oracle.fold_current(aj, n)
    J = oracle.aj_cells(aj)
    assert np.allclose(J[..., 0], 0.1 * k ** 3, rtol=1e-12)
    assert np.allclose(J[..., 1:], 0.0, atol=1e-12)


def test_poisson_single_mode():
    n = 16
    x = np.arange(n)
    kx = 3
    rho = np.cos(2 * np.pi * kx * x / n)[None, None, :] * np.ones((8, 12, 1))
    phi = fields.poisson_periodic(rho)
    kappa = (2 * np.sin(np.pi * kx / n)) ** 2
    assert np.allclose(phi, rho / kappa, atol=1e-12)


def test_poichk_residual_and_gauge():
    rng = np.random.default_rng(13)
    rho = rng.normal(size=(8, 6, 10))
    rho -= rho.mean()
    phi = fields.poisson_periodic(rho)
    assert np.abs(fields.lap7_residual(rho, phi)).max() < 1e-12
    assert abs(phi.mean()) < 1e-13


def test_e_is_minus_grad_phi():
    rng = np.random.default_rng(14)
    phi = rng.normal(size=(5, 6, 7))
    eb = fields.eb_from_phi(phi)
    e = fields.compact_from_halo(eb)
    assert np.allclose(e[..., 0], phi - np.roll(phi, -1, 2))
    assert np.allclose(e[..., 1], phi - np.roll(phi, -1, 1))
    assert np.allclose(e[..., 2], phi - np.roll(phi, -1, 0))
    # the halo holds the periodic images the relocation reads
    assert np.array_equal(eb[1:-2, 1:-2, 0, 0], eb[1:-2, 1:-2, -3, 0])


def test_yee_divergence_free_b_and_gauss_law():
    # discrete Yee: B update keeps div B, E update with J keeps
    # div E - rho consistent when continuity holds (div of curl = 0)
    rng = np.random.default_rng(15)
    ebc = rng.normal(size=(6, 5, 4, 6))
    def divb(a):  # at cell centres: forward differences of the face-centred B
        return (np.roll(a[..., 3], -1, 2) - a[..., 3] + np.roll(a[..., 4], -1, 1) - a[..., 4]
                + np.roll(a[..., 5], -1, 0) - a[..., 5])
    def dive(a):
        return (a[..., 0] - np.roll(a[..., 0], 1, 2) + a[..., 1] - np.roll(a[..., 1], 1, 1)
                + a[..., 2] - np.roll(a[..., 2], 1, 0))
    b1 = fields.bfield_half(ebc)
    assert np.allclose(divb(b1), divb(ebc), atol=1e-12)
    e1 = fields.efield_full(b1, np.zeros(ebc.shape[:3] + (3,)), tcs=3.0)
    assert np.allclose(dive(e1), dive(b1), atol=1e-12)


@pytest.mark.parametrize("m", [1, 3])
def test_yee_vacuum_wave_follows_the_discrete_dispersion(m):
    """A transverse plane wave E_y = cos(k x), B = 0, in vacuum: the leapfrog
    of bfield (half step, twice) and efield (tcs = 2 cv^2, no current) must
    oscillate with sin(w / 2) = 2 cv sin(k / 2) -- the Yee dispersion for
    c dt / dx = 2 cv.  The mode amplitude a_n obeys a_{n+1} + a_{n-1} =
    2 cos(w) a_n; a wrong sign in either update makes it grow instead."""
    n = (16, 4, 4)
    cv = 0.2
    tcs = 2.0 * cv * cv
    k = 2.0 * np.pi * m / n[0]
    x = np.arange(n[0])
    ebc = np.zeros((n[2], n[1], n[0], 6))
    ebc[..., 1] = np.cos(k * x)[None, None, :]
    mode = np.cos(k * x) / (n[0] / 2.0)
    amp = [float((ebc[0, 0, :, 1] * mode).sum())]
    zero = np.zeros(ebc.shape[:3] + (3,))
    for _ in range(40):
        ebc = fields.bfield_half(ebc)
        ebc = fields.bfield_half(ebc)
        ebc = fields.efield_full(ebc, zero, tcs)
        amp.append(float((ebc[0, 0, :, 1] * mode).sum()))
    a = np.array(amp)
    cosw = 1.0 - 8.0 * cv * cv * np.sin(k / 2.0) ** 2
    assert np.allclose(a[2:] + a[:-2], 2.0 * cosw * a[1:-1], atol=1e-12)
    # a stable oscillation (the B = 0 start adds a small sine part): bounded
    assert np.abs(a).max() <= 1.05 and a.min() < -0.5


def test_field_energy_single_mode_closed_form():
    """energy.F90:52-70: 0.5 sum E^2 / V and 0.5 cs sum B^2 / V of single modes
    A cos(k x): A^2 / 4 and cs B^2 / 4."""
    n = (16, 8, 4)
    x = np.arange(n[0])
    ebc = np.zeros((n[2], n[1], n[0], 6))
    ebc[..., 0] = 0.3 * np.cos(2 * np.pi * 2 * x / n[0])[None, None, :]
    ebc[..., 5] = 0.7 * np.sin(2 * np.pi * 3 * x / n[0])[None, None, :]
    eb = np.pad(ebc, [(1, 2), (1, 2), (1, 2), (0, 0)], mode="wrap")
    cs = 0.04
    out = oracle.field_energy(eb, n, cs, np.array(n, dtype=float))
    assert out[0] == pytest.approx(0.3 ** 2 / 4.0, rel=1e-13)
    assert out[1] == pytest.approx(cs * 0.7 ** 2 / 4.0, rel=1e-13)
