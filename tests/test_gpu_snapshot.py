"""Checkpoint / resume through the snapshot format on the GPU.

A device-resident ES run of 2k steps against the same run stopped after k
steps, written to SNAPSHOT1/esdat0000.h5 (save_esdat, snap.F90:203-356),
loaded into a fresh context (inipcl.F90:359-426, inifld.F90:59-74) and run k
more steps.  The FP32 form is bit-identical (its charge is int64 fixed point,
so the particles' new order does not change any sum); the FP64 form's charge
is summed in floating point in an order that may change, so it agrees to
1e-12.  The run uses a non-uniform background charge (emg_set_background),
which travels through the file's 'rhobk'."""
import numpy as np
import pytest

import oracle
from ctca_amd import emg, snapshot

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not snapshot.hdf5_available(), reason="libhdf5 not found")]

N = (16, 12, 8)
Q, QM, RM = (-0.002, 0.002), (-1.0, 0.01), (0.002, 0.2)
B0 = (0.0, 0.0, 0.02)


def particles(counts, vth, seed):
    spec = [dict(npin=int(c), peth=v, path=v) for c, v in zip(counts, vth)]
    return emg.host_inipcl(spec, N, np.eye(3), seed)


def background(seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(N[2], N[1], N[0])) * 1e-3


def test_background_array_enters_the_field_solve():
    counts = (60000, 20000)
    p = particles(counts, (0.1, 0.01), 21)
    bk = background(22)
    with emg.Context(N, Q, QM, RM, counts) as ctx:
        ctx.upload(p, counts)
        ctx.deposit_rho(0)
        rho = ctx.rho()
        ctx.set_background(bk)
        ctx.solve_es(0.001)
        phi = ctx.potential()
        ctx.set_background(None)
        ctx.solve_es(0.001)
        phi0 = ctx.potential()
        with pytest.raises(ValueError):
            ctx.set_background(bk[:, :, :-1])
    ref = oracle.fields.poisson_periodic(rho + bk + 0.001)
    assert np.abs(phi - ref).max() <= 1e-12 * np.abs(ref).max()
    ref0 = oracle.fields.poisson_periodic(rho + 0.001)
    assert np.abs(phi0 - ref0).max() <= 1e-12 * np.abs(ref0).max()


def es_steps(ctx, precision, nsteps):
    """deposit, solve, push (esses.F90:300-478); the FP32 push deposits the
    charge of the next step itself"""
    ctx.deposit_rho(0)
    for s in range(nsteps):
        if s and precision == emg.PLAIN_FP64:
            ctx.deposit_rho(0)
        ctx.solve_es(0.0)
        ctx.push_es(2)


def canonical(pbuf):
    """The records as a set: sorted by species, then position and velocity."""
    keys = [pbuf[f] for f in ("vz", "vy", "vx", "z", "y", "x", "spec")]
    return pbuf[np.lexsort(keys)]


@pytest.mark.parametrize("precision", [emg.FP32_FIXED, emg.PLAIN_FP64])
def test_resume_equals_uninterrupted(precision, tmp_path):
    counts = (60000, 20000)
    vth = (0.1, 0.01)
    p = particles(counts, vth, 23)
    bk = background(24)
    k = 3
    kw = dict(tile=4) if precision == emg.FP32_FIXED else {}

    def context():
        return emg.Context(N, Q, QM, RM, counts, precision=precision, b0=B0, **kw)

    with context() as ctx:
        ctx.upload(p, counts)
        ctx.set_background(bk)
        es_steps(ctx, precision, 2 * k)
        whole, cnt = ctx.download()
    assert list(cnt) == list(counts)

    path = snapshot.snapshot_file(str(tmp_path / snapshot.SAVE_DIR), 0)
    with context() as ctx:
        ctx.upload(p, counts)
        ctx.set_background(bk)
        es_steps(ctx, precision, k)
        snapshot.save(ctx, path, rhobk=bk)
    with context() as ctx:
        snap = snapshot.load(ctx, path)
        assert snap.np.tolist() == list(counts) and snap.n == N
        es_steps(ctx, precision, k)
        resumed, cnt = ctx.download()
    assert list(cnt) == list(counts)

    if precision == emg.FP32_FIXED:
        # the re-bucketing reorders the records: compare them as sets
        a, b = canonical(whole), canonical(resumed)
        for f in ("spec", "x", "y", "z", "vx", "vy", "vz"):
            assert np.array_equal(a[f], b[f]), f
    else:
        # the FP64 form keeps the upload order, which the file keeps too
        assert np.array_equal(whole["spec"], resumed["spec"])
        for f, L in zip(("x", "y", "z"), N):
            d = np.abs(whole[f] - resumed[f])
            assert np.minimum(d, L - d).max() <= 1e-12 * L, f
        for f in ("vx", "vy", "vz"):
            assert np.abs(whole[f] - resumed[f]).max() <= 1e-12, f


def test_driver_job_continuation(tmp_path):
    """The host loop's jobnum: a job with jobnum = (0, 1) runs and writes
    SNAPSHOT1 at its end; moved to SNAPSHOT0, a job with jobnum = (1, 0)
    starts from it with the half push positn(1, 1) of a continued job
    (esses.F90:198-201): its particles are the first job's last ones moved by
    the oracle's positn, bit for bit, and its background is the file's."""
    import os
    import oracle
    from ctca_amd import driver, units
    w = units.es32()
    su = units.to_sunits(w)
    a = driver.ESRun(w, jobnum=(0, 1), root=str(tmp_path))
    a.initialize()
    a.run(3)
    last, counts = a.ctx.download()
    a.finalize()
    a.close()
    os.rename(tmp_path / snapshot.SAVE_DIR, tmp_path / snapshot.LOAD_DIR)
    b = driver.ESRun(w, jobnum=(1, 0), root=str(tmp_path))
    assert np.all(b.rhobk == a.rhobk)
    b.initialize()
    got, cnt = b.ctx.download()
    b.close()
    assert list(cnt) == list(counts)
    ref = last.copy()
    oracle.positn(ref, 1, np.array(w.n, dtype=float))
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(got[f], ref[f]), f
