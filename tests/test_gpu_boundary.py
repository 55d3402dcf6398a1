"""The boundary's charge semantics against the reference, on the GPU:

  * charge(ps, func = 1) fills only rhodg: in the EM loop's order
    (emses.F90:474-476) rhoupd updates rho by continuity and the following
    charge(1, 1) leaves that rho untouched, while rhodg gets the per-species
    density of the new positions, folded with its ghost planes as the
    oracle's charge + oh3_exchange_borders + add_boundary_charge leave them;
  * the FP32 form's fused kinetic energy accumulates in FP64 per lane, as
    energy.F90:38-43 sums it: with thousands of particles per lane it equals
    the oracle's FP64 sum over the same velocities to 1e-9 (an FP32 running
    sum of that length is off by ~1e-6)."""
import numpy as np
import pytest

import oracle
from ctca_amd import driver, emg, fieldsolve, units

pytestmark = pytest.mark.gpu


def em_small():
    nc = 16 * 12 * 8
    return units.Workload("em_small", (16, 12, 8), 0.02, 4, [
        units.Species("electron", -1.0, 1.0, 16 * nc, 1.0),
        units.Species("beam", -1.0, float(np.sqrt(0.1)), 2 * nc, 1.0, drift=(5.0, 0.0, 0.0)),
    ], cv=20.0, em=True, seed=31)


@pytest.mark.parametrize("precision", [emg.PLAIN_FP64, emg.FP32_FIXED], ids=["fp64", "fp32"])
def test_em_loop_rhoupd_then_charge_func1(precision):
    w = em_small()
    su = units.to_sunits(w)
    n = w.n
    nsp = len(w.species)
    run = driver.EMRun(w, precision=precision)
    try:
        tol = 1e-12 if precision == emg.PLAIN_FP64 else 1e-6
        for _ in range(3):
            rho_prev = run.rho.copy()
            aj = run.step(2)
            # rho is exactly the continuity update of the previous rho:
            # charge(1, 1) did not overwrite it with a fresh deposit
            assert np.array_equal(run.rho, fieldsolve.rhoupd(rho_prev, aj))
            p, counts = run.ctx.download()
            _, ref = oracle.charge(p, counts, su.q, n, with_species=True)
            oracle.fold_rho(ref, n, ncomp=nsp)
            for k in range(nsp):
                scale = np.abs(ref[..., k]).max()
                assert np.abs(run.rhodg[k] - ref[..., k]).max() <= tol * scale, k
            assert not np.any(run.rhodg[nsp:])
            # the continuity-updated rho is not the fresh deposit (the zigzag's
            # in-cell residual), so overwriting it would have been visible
            fresh = ref.sum(axis=-1)
            assert np.abs(run.rho - fresh)[1:-1, 1:-1, 1:-1].max() > 1e-9 * np.abs(fresh).max()
    finally:
        run.close()


def test_charge_func_arrays_and_errors():
    n = (8, 6, 4)
    counts = (3000, 2000)
    q, qm, rm = (0.3, -0.2), (1.0, -0.01), (1.0, 100.0)
    spec = [dict(npin=counts[0], peth=0.3, path=0.3), dict(npin=counts[1], peth=0.01, path=0.01)]
    p = emg.host_inipcl(spec, n, np.eye(3), 12)
    shape = (n[2] + 3, n[1] + 3, n[0] + 3)
    with emg.Context(n, q, qm, rm, counts) as ctx:
        ctx.upload(p, counts)
        rho = np.full(shape, 7.0)
        rhodg = np.full((5,) + shape, 7.0)
        ctx.charge(1, rho=rho, rhodg=rhodg)
        assert np.all(rho == 7.0)                       # func 1 leaves rho alone
        ref, refdg = oracle.charge(p, counts, q, n, with_species=True)
        oracle.fold_rho(ref, n)
        oracle.fold_rho(refdg, n, ncomp=2)
        for k in range(2):
            assert np.abs(rhodg[k] - refdg[..., k]).max() <= 1e-12 * np.abs(refdg[..., k]).max()
        assert not np.any(rhodg[2:])                     # zeroed like the reference
        rhodg[:] = 7.0
        ctx.charge(0, rho=rho, rhodg=rhodg)
        assert np.abs(rho - ref).max() <= 1e-12 * np.abs(ref).max()
        assert np.all(rhodg == 7.0)                     # func 0 leaves rhodg alone
        rho[:] = 0.0
        ctx.charge(2, rho=rho, rhodg=rhodg)
        assert np.abs(rho - ref).max() <= 1e-12 * np.abs(ref).max()
        assert np.abs(rhodg[0] - refdg[..., 0]).max() <= 1e-12 * np.abs(refdg[..., 0]).max()
        with pytest.raises(ValueError):
            ctx.charge(3, rho=rho, rhodg=rhodg)
        with pytest.raises(emg.EmgError):
            ctx.deposit_rho(3)


def test_fp32_kinetic_energy_fp64_lanes_dense_buckets():
    # one 8^3 tile per species, 2,048 particles per lane of its 512-thread
    # block; thermal speeds of 0.04 cell per step keep the particles that
    # cross the box faces (the tile's movers) within the spill capacity
    n = (8, 8, 8)
    counts = (1_048_576, 524_288)
    q, qm, rm = (-0.002, 0.002), (-1.0, 0.01), (0.002, 0.2)
    spec = [dict(npin=counts[0], peth=0.02, path=0.02, vdx=0.005),
            dict(npin=counts[1], peth=0.002, path=0.002)]
    p = emg.host_inipcl(spec, n, np.eye(3), 21)
    rng = np.random.default_rng(4)
    phi = rng.normal(size=(n[2], n[1], n[0])) * 0.02
    eb = oracle.fields.eb_from_phi(phi, rng.normal(size=(n[2], n[1], n[0], 3)) * 0.01)
    with emg.Context(n, q, qm, rm, counts, precision=emg.FP32_FIXED, tile=8) as ctx:
        ctx.upload(p, counts)
        ctx.set_fields(eb)
        ctx.push_es(2)
        ek = ctx.energy()                  # the push's fused FP64 lane sums
        g, cnt = ctx.download()
    assert list(cnt) == list(counts)
    ok = oracle.energy(g, cnt, rm, np.array(n, dtype=float))
    assert np.abs(ek - ok).max() <= 1e-9 * np.abs(ok).max(), (ek, ok)
