"""The field updates on the device against the oracle: the EM Yee update
(emg_bfield, emg_efield, emg_grdcrt, emg_field_energy; bfield.F90:38-52,
efield.F90:55-72, energy.F90:52-70) with a device-resident EM loop in the
reference's order (emses.F90:365-434: bfield, grdcrt, psolve1 + current fold,
bfield, efield), and the ES field solve (emg_solve_es: periodic spectral
Poisson, E = -grad phi, grdcrt) with a device-resident ES loop.

Bar: the B and E updates are bit-identical to the oracle for the same inputs
(both without FP contraction, same expression order; E takes the device's own
folded current, read back through emg_get_current); ghosts equal their
periodic images; field energy to 1e-12; three loop steps to 1e-9 (FP64 form,
whose current sums are order dependent) and 1e-4 (FP32 form).  The Poisson
solve (hipFFT) agrees with numpy's FFT to 1e-12 of the potential's maximum
(the FFT round-off is not FFTW's: parity with the reference's FFTW is
unpinned), and its 7-point residual is the dropped mean."""
import numpy as np
import pytest

import oracle
from ctca_amd import emg

pytestmark = pytest.mark.gpu

N = (16, 12, 8)
TCS = 2.0 * 0.4 ** 2   # 2 cv^2, cv = c dt / dx = 0.4


def random_eb(n, rng, amp_e=0.02, amp_b=0.01):
    ebc = np.zeros((n[2], n[1], n[0], 6))
    ebc[..., 0:3] = rng.normal(size=(n[2], n[1], n[0], 3)) * amp_e
    ebc[..., 3:6] = rng.normal(size=(n[2], n[1], n[0], 3)) * amp_b
    return ebc


def halo(ebc):
    return np.pad(ebc, [(1, 2), (1, 2), (1, 2), (0, 0)], mode="wrap")


def particles(n, counts, vth, seed):
    spec = [dict(npin=int(c), peth=v, path=v) for c, v in zip(counts, vth)]
    p = emg.host_inipcl(spec, n, np.eye(3), seed)
    p["pid"] = np.arange(len(p))
    return p


def interior(eb):
    return eb[1:-2, 1:-2, 1:-2]


def test_device_yee_update_is_bitwise_the_oracles():
    rng = np.random.default_rng(41)
    ebc = random_eb(N, rng)
    counts = (30000, 8000)
    q, qm, rm = (-0.002, 0.002), (-1.0, 0.01), (0.002, 0.2)
    p = particles(N, counts, (0.1, 0.01), 5)
    sl = np.array(N, dtype=float)
    with emg.Context(N, q, qm, rm, counts, b0=(0.0, 0.0, 0.02)) as ctx:
        ctx.upload(p, counts)
        ctx.set_fields(halo(ebc))
        e0 = ctx.field_energy(TCS / 2.0)
        ctx.bfield()
        eb1 = ctx.fields()
        with pytest.raises(emg.EmgError):
            ctx.push_em(2)                # nodes older than the field update
        ctx.grdcrt()
        ctx.push_em(2)
        aj = ctx.current()                # folded, periodic cells 0..n-1
        ctx.efield(TCS, 2.0)
        eb2 = ctx.fields()
        e2 = ctx.field_energy(TCS / 2.0)
    ob1 = oracle.fields.bfield_half(ebc)
    assert np.array_equal(interior(eb1), ob1)
    assert np.array_equal(eb1, halo(ob1))          # ghosts = periodic images
    ob2 = oracle.fields.efield_full(ob1, aj, TCS, 2.0)
    assert np.array_equal(interior(eb2), ob2)
    assert np.array_equal(eb2, halo(ob2))
    for got, ref in ((e0, oracle.field_energy(halo(ebc), N, TCS / 2.0, sl)),
                     (e2, oracle.field_energy(halo(ob2), N, TCS / 2.0, sl))):
        assert np.allclose(got, ref, rtol=1e-12, atol=0.0)


@pytest.mark.parametrize("precision,tol", [(emg.PLAIN_FP64, 1e-9), (emg.FP32_FIXED, 1e-4)])
def test_device_resident_em_loop(precision, tol):
    """Three EM steps with the fields on the device, in the reference's order,
    against the oracle's host loop started from the same (rounded) state."""
    rng = np.random.default_rng(43)
    ebc = random_eb(N, rng, 0.01, 0.005)
    counts = (40000, 10000)
    q, qm, rm = (-0.002, 0.002), (-1.0, 0.01), (0.002, 0.2)
    b0 = (0.0, 0.0, 0.02)
    vth = (0.1, 0.01)
    p = particles(N, counts, vth, 7)
    sl = np.array(N, dtype=float)
    nsteps = 3
    kw = dict(track_pid=True, tile=4) if precision == emg.FP32_FIXED else {}
    with emg.Context(N, q, qm, rm, counts, precision=precision, b0=b0, **kw) as ctx:
        ctx.upload(p, counts)
        start, _ = ctx.download()
        ctx.set_fields(halo(ebc))
        for _ in range(nsteps):
            ctx.bfield()
            ctx.grdcrt()
            ctx.push_em(2)
            ctx.bfield()
            ctx.efield(TCS, 2.0)
        g, cnt = ctx.download()
        eb = ctx.fields()
    assert list(cnt) == list(counts)
    o = start[np.argsort(start["pid"], kind="stable")]
    g = g[np.argsort(g["pid"], kind="stable")]
    e = ebc.copy()
    for _ in range(nsteps):
        e = oracle.fields.bfield_half(e)
        wrk = oracle.grdcrt(halo(e), N, (0, 0, 0), b0)
        aj = oracle.psolve1(o, counts, qm, q, wrk, N, 2, sl)
        oracle.fold_current(aj, N)
        e = oracle.fields.bfield_half(e)
        e = oracle.fields.efield_full(e, oracle.aj_cells(aj), TCS, 2.0)
    for f, L in zip(("x", "y", "z"), N):
        d = np.abs(g[f] - o[f])
        assert np.minimum(d, L - d).max() <= tol, f
    vt = np.array(vth)[np.repeat([0, 1], counts)]
    for f in ("vx", "vy", "vz"):
        assert (np.abs(g[f] - o[f]) / np.maximum(np.abs(o[f]), vt)).max() <= tol, f
    for c in range(6):
        ref = e[..., c]
        assert np.abs(interior(eb)[..., c] - ref).max() <= tol * np.abs(ref).max(), c


def test_device_poisson_solve_matches_the_host_solve():
    """emg_solve_es against an independent FFT (numpy) of the same charge, and
    the reference's own check: the 7-point residual rho + lap7 phi of the
    device potential is the dropped mean (poissn.F90:936-944)."""
    counts = (60000, 20000)
    q, qm, rm = (-0.002, 0.002), (-1.0, 0.01), (0.002, 0.2)
    p = particles(N, counts, (0.1, 0.01), 9)
    bk = 0.001
    with emg.Context(N, q, qm, rm, counts) as ctx:
        ctx.upload(p, counts)
        ctx.deposit_rho(0)
        rho = ctx.rho() + bk
        ctx.solve_es(bk)
        phi = ctx.potential()
        eb = ctx.fields()
    ref = oracle.fields.poisson_periodic(rho)
    assert np.abs(phi - ref).max() <= 1e-12 * np.abs(ref).max()
    res = oracle.fields.lap7_residual(rho, phi)
    assert np.abs(res - rho.mean()).max() <= 1e-12 * np.abs(rho).max()
    e_ref = interior(oracle.fields.eb_from_phi(ref))[..., 0:3]
    assert np.abs(interior(eb)[..., 0:3] - e_ref).max() <= 1e-12 * np.abs(e_ref).max()
    assert np.array_equal(interior(eb)[..., 3:6], np.zeros_like(e_ref))


@pytest.mark.parametrize("precision,tol", [(emg.PLAIN_FP64, 1e-9), (emg.FP32_FIXED, 1e-4)])
def test_device_resident_es_loop(precision, tol):
    """Four ES steps with the field solve on the device (deposit, solve,
    relocate, push: esses.F90:300-478 for one periodic rank) against the
    oracle's host loop from the same (rounded) state."""
    counts = (60000, 20000)
    q, qm, rm = (-0.002, 0.002), (-1.0, 0.01), (0.002, 0.2)
    b0 = (0.0, 0.0, 0.02)
    vth = (0.1, 0.01)
    p = particles(N, counts, vth, 11)
    sl = np.array(N, dtype=float)
    bk = 0.0
    nsteps = 4
    kw = dict(track_pid=True, tile=4) if precision == emg.FP32_FIXED else {}
    with emg.Context(N, q, qm, rm, counts, precision=precision, b0=b0, **kw) as ctx:
        ctx.upload(p, counts)
        start, _ = ctx.download()
        ctx.deposit_rho(0)
        for _ in range(nsteps):
            if precision == emg.PLAIN_FP64:
                ctx.deposit_rho(0)
            ctx.solve_es(bk)
            ctx.push_es(2)      # the FP32 push deposits for the next solve
        g, cnt = ctx.download()
    assert list(cnt) == list(counts)
    o = start[np.argsort(start["pid"], kind="stable")]
    g = g[np.argsort(g["pid"], kind="stable")]
    for _ in range(nsteps):
        r = oracle.charge(o, counts, q, N)
        oracle.fold_rho(r, N)
        phi = oracle.fields.poisson_periodic(oracle.rho_nodes(r) + bk)
        wrk = oracle.grdcrt(oracle.fields.eb_from_phi(phi), N, (0, 0, 0), b0)
        oracle.psolve2(o, counts, qm, wrk, N, 2, sl)
    for f, L in zip(("x", "y", "z"), N):
        d = np.abs(g[f] - o[f])
        assert np.minimum(d, L - d).max() <= tol, f
    vt = np.array(vth)[np.repeat([0, 1], counts)]
    for f in ("vx", "vy", "vz"):
        assert (np.abs(g[f] - o[f]) / np.maximum(np.abs(o[f]), vt)).max() <= tol, f
