"""Edge cases of the particle path on the GPU against the oracle: an empty
species, a grid of one tile per axis (every neighbouring tile is the tile
itself through the periodic wrap), particles on the box faces, the capacity
limit of the FP32 form (refused before anything is allocated), and a device
allocation that fails (an error code, and the library keeps working)."""
import numpy as np
import pytest

import oracle
from ctca_amd import emg

pytestmark = pytest.mark.gpu

Q, QM, RM = (-0.002, 0.002), (-1.0, 0.01), (0.002, 0.2)


def particles(n, counts, vth, seed):
    spec = [dict(npin=int(c), peth=v, path=v) for c, v in zip(counts, vth)]
    p = emg.host_inipcl(spec, n, np.eye(3), seed)
    p["pid"] = np.arange(len(p))
    return p


def eb_random(n, seed, amp=0.01):
    rng = np.random.default_rng(seed)
    return oracle.fields.eb_from_phi(rng.normal(size=(n[2], n[1], n[0])) * amp)


def run_both(n, counts, p, eb, steps, tile=None, b0=(0.0, 0.0, 0.02), sort_every=0):
    """(FP64 result, FP32 result, oracle result, FP32 start) after `steps` ES pushes;
    sort_every applies to the FP64 form (2: sorted before the first push, and the
    re-sort before the third carried by that push)."""
    out = {}
    for prec in (emg.PLAIN_FP64, emg.FP32_FIXED):
        kw = (dict(track_pid=True, tile=tile) if prec == emg.FP32_FIXED
              else dict(sort_every=sort_every))
        with emg.Context(n, Q, QM, RM, counts, precision=prec, b0=b0, **kw) as ctx:
            ctx.upload(p, counts)
            start, _ = ctx.download()
            ctx.set_fields(eb)
            for _ in range(steps):
                ctx.push_es(2)
            if prec == emg.PLAIN_FP64:
                ctx.deposit_rho(0)
            rho = [ctx.rho(species=k) for k in range(len(counts))]
            g, cnt = ctx.download()
        assert list(cnt) == list(counts)
        out[prec] = (g[np.argsort(g["pid"], kind="stable")], rho,
                     start[np.argsort(start["pid"], kind="stable")])
    return out


def oracle_run(n, counts, start, eb, steps, b0=(0.0, 0.0, 0.02)):
    o = start.copy()
    wrk = oracle.grdcrt(eb, n, (0, 0, 0), b0)
    for _ in range(steps):
        oracle.psolve2(o, counts, QM, wrk, n, 2, np.array(n, dtype=float))
    _, rhodg = oracle.charge(o, counts, Q, n, with_species=True)
    oracle.fold_rho(rhodg, n, ncomp=len(counts))
    return o, rhodg


def check(n, counts, res, o, rhodg, bitwise):
    g, rho, _ = res
    for f, L in zip(("x", "y", "z"), n):
        if bitwise:
            assert np.array_equal(g[f], o[f]), f
        else:
            d = np.abs(g[f] - o[f])
            assert np.minimum(d, L - d).max() <= 1e-4, f
    for k in range(len(counts)):
        ok = oracle.rho_nodes(rhodg)[..., k]
        scale = max(np.abs(ok).max(), 1e-300)
        assert np.abs(rho[k] - ok).max() <= (1e-12 if bitwise else 1e-5) * scale, k


@pytest.mark.parametrize("sort_every", [0, 2])
def test_empty_species(sort_every):
    n = (16, 8, 8)
    counts = (20000, 0)
    p = particles(n, (20000, 1), (0.1, 0.01), 3)[:20000]
    eb = eb_random(n, 1)
    res = run_both(n, counts, p, eb, 3, tile=8, sort_every=sort_every)
    o, rhodg = oracle_run(n, counts, res[emg.PLAIN_FP64][2], eb, 3)
    check(n, counts, res[emg.PLAIN_FP64], o, rhodg, True)
    o, rhodg = oracle_run(n, counts, res[emg.FP32_FIXED][2], eb, 3)
    check(n, counts, res[emg.FP32_FIXED], o, rhodg, False)
    assert not np.any(res[emg.FP32_FIXED][1][1])   # the empty species deposits nothing


@pytest.mark.parametrize("sort_every", [0, 2])
def test_one_tile_per_axis(sort_every):
    """n = 4^3 with 4^3-cell tiles: each of the 26 neighbouring tiles is the
    tile itself, so every mover leaves through its own outbox and is pulled
    back by its own block.  With sort_every = 2 the FP64 form's runs hold
    over a hundred particles per cell and a re-sort is carried by a push."""
    n = (4, 4, 4)
    counts = (8000, 4000)
    p = particles(n, counts, (0.3, 0.05), 5)
    eb = eb_random(n, 2)
    res = run_both(n, counts, p, eb, 4, tile=4, sort_every=sort_every)
    o, rhodg = oracle_run(n, counts, res[emg.FP32_FIXED][2], eb, 4)
    check(n, counts, res[emg.FP32_FIXED], o, rhodg, False)
    o, rhodg = oracle_run(n, counts, res[emg.PLAIN_FP64][2], eb, 4)
    check(n, counts, res[emg.PLAIN_FP64], o, rhodg, True)


def test_particles_on_box_faces():
    """Particles at x = 0 and just below L on every axis, moving out through
    the faces in both directions."""
    n = (8, 8, 8)
    m = 512
    p = particles(n, (m, 8), (0.1, 0.01), 7)
    rng = np.random.default_rng(8)
    L = float(n[0])
    for f in ("x", "y", "z"):
        side = rng.integers(0, 3, size=m)
        p[f][:m] = np.where(side == 0, 0.0, np.where(side == 1, np.nextafter(L, 0.0), p[f][:m]))
    counts = (m, 8)
    eb = eb_random(n, 3)
    res = run_both(n, counts, p, eb, 2, tile=4)
    o, rhodg = oracle_run(n, counts, res[emg.PLAIN_FP64][2], eb, 2)
    check(n, counts, res[emg.PLAIN_FP64], o, rhodg, True)
    o, rhodg = oracle_run(n, counts, res[emg.FP32_FIXED][2], eb, 2)
    check(n, counts, res[emg.FP32_FIXED], o, rhodg, False)


def test_fp32_capacity_limit_is_refused():
    # the hot loop addresses a bucket's slots from the bucket's first slot, so
    # more than 2^29 slots per context work (the 1G ramp runs in
    # test_gpu_fullsize); the step's capacity scan bounds a context at 2^32
    with pytest.raises(emg.EmgError) as e:
        emg.Context((64, 64, 64), Q, QM, RM, (1 << 31, 1 << 31), precision=emg.FP32_FIXED)
    assert "2^32" in str(e.value)


def test_failed_device_allocation_returns_an_error():
    """A context far larger than the GPU's memory (2^37 FP64 particles, over
    6 TB) fails with EMG_ERR_NOMEM, not an abort; contexts created after it
    push and deposit normally (the failed allocation leaves no error behind
    for the later launch checks)."""
    with pytest.raises(emg.EmgError) as e:
        emg.Context((64, 64, 64), Q, QM, RM, (1 << 37, 1), precision=emg.PLAIN_FP64)
    assert e.value.code == -3, str(e.value)   # EMG_ERR_NOMEM
    n = (8, 8, 8)
    counts = (4000, 1000)
    p = particles(n, counts, (0.1, 0.01), 9)
    eb = eb_random(n, 4)
    res = run_both(n, counts, p, eb, 2, tile=4)
    o, rhodg = oracle_run(n, counts, res[emg.PLAIN_FP64][2], eb, 2)
    check(n, counts, res[emg.PLAIN_FP64], o, rhodg, True)
    o, rhodg = oracle_run(n, counts, res[emg.FP32_FIXED][2], eb, 2)
    check(n, counts, res[emg.FP32_FIXED], o, rhodg, False)
