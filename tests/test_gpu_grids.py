"""The FP32 form on grids of any size (the reference takes any nx, ny, nz,
src/wrapper/ohinit.F90:131-159): grids that are not multiples of the tile
edge end in partial tiles, and the packed cell index takes ceil(log2 n) bits
per axis, so axes beyond 1024 cells work.  Each case is checked against the
oracle from the same (FP32-rounded) state: positions 1e-5 cells and
velocities 1e-5 relative per step, the folded density (ghost planes
included) and the zigzag current 1e-5 of their maximum over several steps,
1e-6 after one; every particle is kept.  Grids the form cannot pack are
refused with EMG_ERR_UNSUPPORTED (the plain FP64 form takes them)."""
import numpy as np
import pytest

import oracle
from ctca_amd import emg

pytestmark = pytest.mark.gpu

Q, QM, RM = (0.002, -0.002), (1.0, -0.01), (0.002, 0.2)


def particles(n, counts, vth, seed=10, drift=None):
    spec = []
    for k, c in enumerate(counts):
        d = dict(npin=int(c), peth=vth[k], path=vth[k])
        if drift is not None:
            d["vdx"] = drift[k]
        spec.append(d)
    p = emg.host_inipcl(spec, n, np.eye(3), seed)
    p["pid"] = np.arange(len(p))
    return p


def by_pid(p):
    return p[np.argsort(p["pid"], kind="stable")]


def random_eb(n, seed, amp_e=0.03, amp_b=0.01):
    rng = np.random.default_rng(seed)
    phi = rng.normal(size=(n[2], n[1], n[0])) * amp_e
    b = rng.normal(size=(n[2], n[1], n[0], 3)) * amp_b
    return oracle.fields.eb_from_phi(phi, b)


def check_state(g, o, n, vth, counts, tol):
    for f, L in zip(("x", "y", "z"), n):
        d = np.abs(g[f] - o[f])
        assert np.minimum(d, L - d).max() <= tol, f
    vt = np.array(vth)[np.repeat(np.arange(len(counts)), counts)]
    for f in ("vx", "vy", "vz"):
        assert (np.abs(g[f] - o[f]) / np.maximum(np.abs(o[f]), vt)).max() <= tol, f


@pytest.mark.parametrize("tile,n,steps", [(4, (30, 30, 30), 1), (8, (30, 30, 30), 1),
                                          (8, (30, 21, 13), 5), (4, (18, 13, 6), 5),
                                          (8, (2050, 9, 8), 1)],
                         ids=["30^3-t4", "30^3-t8", "30x21x13-t8-5steps", "18x13x6-t4-5steps",
                              "2050x9x8-t8"])
def test_es_push_on_any_grid_vs_oracle(tile, n, steps):
    ncell = n[0] * n[1] * n[2]
    counts = (24 * ncell // 3, 8 * ncell // 3)
    vth = (0.25, 0.02)        # ~0.5 cell per step: particles cross the partial tiles
    eb = random_eb(n, 7)
    b0 = (0.0, 0.01, 0.02)
    p = particles(n, counts, vth)
    lo, hi = (-1, -1, -1), tuple(v + 1 for v in n)
    with emg.Context(n, Q, QM, RM, counts, precision=emg.FP32_FIXED, b0=b0, tile=tile,
                     track_pid=True) as ctx:
        ctx.upload(p, counts)
        start, _ = ctx.download()
        ctx.set_fields(eb)
        for _ in range(steps):
            ctx.push_es(2)
        rho = [ctx.rho(lo=lo, hi=hi, species=k) for k in range(2)]
        g, cnt = ctx.download()
    assert list(cnt) == list(counts)
    o = by_pid(start)
    g = by_pid(g)
    assert np.array_equal(g["pid"], o["pid"])
    wrk = oracle.grdcrt(eb, n, (0, 0, 0), b0)
    for _ in range(steps):
        oracle.psolve2(o, counts, QM, wrk, n, 2, np.array(n, dtype=float))
    tol = 1e-5 if steps == 1 else 1e-4
    check_state(g, o, n, vth, counts, tol)
    _, rhodg = oracle.charge(o, counts, Q, n, with_species=True)
    oracle.fold_rho(rhodg, n, ncomp=2)
    rtol = 1e-6 if steps == 1 else 1e-5
    for k in range(2):
        ref = rhodg[..., k]
        assert np.abs(rho[k] - ref).max() <= rtol * np.abs(ref).max(), k


@pytest.mark.parametrize("tile,n", [(4, (30, 30, 30)), (8, (30, 21, 13))], ids=["30^3-t4", "30x21x13-t8"])
def test_em_push_on_any_grid_vs_oracle(tile, n):
    ncell = n[0] * n[1] * n[2]
    counts = (12 * ncell, 2 * ncell)
    q, qm = (0.002, 0.002), (1.0, 1.0)
    eb = random_eb(n, 8, 0.02, 0.01)
    b0 = (0.0, 0.0, 0.02)
    p = particles(n, counts, (0.15, 0.02), drift=(0.0, 0.15))
    with emg.Context(n, q, qm, (0.002, 0.002), counts, precision=emg.FP32_FIXED, b0=b0,
                     tile=tile, track_pid=True) as ctx:
        ctx.upload(p, counts)
        start, _ = ctx.download()
        ctx.set_fields(eb)
        ctx.push_em(2)
        J = ctx.current(lo=(-1, -1, -1), hi=tuple(v + 1 for v in n))
        g, cnt = ctx.download()
    assert list(cnt) == list(counts)
    o = by_pid(start)
    g = by_pid(g)
    wrk = oracle.grdcrt(eb, n, (0, 0, 0), b0)
    aj = oracle.psolve1(o, counts, qm, q, wrk, n, 2, np.array(n, dtype=float))
    oracle.fold_current(aj, n)
    check_state(g, o, n, (0.15, 0.02), counts, 1e-5)
    oj = aj[3:n[2] + 6, 3:n[1] + 6, 3:n[0] + 6]
    assert np.abs(J - oj).max() <= 1e-5 * np.abs(oj).max()


@pytest.mark.parametrize("n,tile", [((3, 8, 8), 4), ((8, 8, 6), 8), ((4096, 4096, 512), 8)],
                         ids=["axis-below-tile-4", "axis-below-tile-8", "cell-index-over-32-bits"])
def test_grids_the_fp32_form_refuses(n, tile):
    with pytest.raises(emg.EmgError) as e:
        emg.Context(n, Q, QM, RM, (1000, 1000), precision=emg.FP32_FIXED, tile=tile)
    assert e.value.code == -6, str(e.value)   # EMG_ERR_UNSUPPORTED
    assert "EMG_PLAIN_FP64" in str(e.value)
