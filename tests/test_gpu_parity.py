"""GPU parity of the particle path against the oracle (CPU FP64 restatement).

Bar: the plain FP64 push / positn / zigzag push are bit-identical to the
oracle (both built without FP contraction); FP64 deposits (atomics, order
dependent) agree to 1e-12 of the grid maximum; the FP32 cell+offset form
agrees with the FP64 oracle per step to 1e-5 (positions in cells, velocities
relative to max(|v|, v_th)) and its per-species density to 1e-6 of the
species maximum; fixed-point densities are bitwise independent of particle
order and tile size.
"""
import numpy as np
import pytest

import oracle
from ctca_amd import emg, fieldsolve, units

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz")


def mt_particles(n, counts, vth, seed=10, drift=None):
    spec = []
    for k, c in enumerate(counts):
        d = dict(npin=int(c), peth=vth[k], path=vth[k])
        if drift is not None:
            d["vdx"] = drift[k]
        spec.append(d)
    p = emg.host_inipcl(spec, n, np.eye(3), seed)
    p["pid"] = np.arange(len(p))
    return p


def periodic_eb(n, rng, amp_e, amp_b):
    """Staggered E from a random potential plus random B, periodic halos."""
    phi = rng.normal(size=(n[2], n[1], n[0])) * amp_e
    b = rng.normal(size=(n[2], n[1], n[0], 3)) * amp_b
    return oracle.fields.eb_from_phi(phi, b)


def ctx_for(n, q, qm, rm, counts, precision=emg.PLAIN_FP64, b0=(0, 0, 0), e0=(0, 0, 0), **kw):
    return emg.Context(n, q, qm, rm, counts, precision=precision, e0=e0, b0=b0, **kw)


def by_pid(p):
    return p[np.argsort(p["pid"], kind="stable")]


SMALL = dict(n=(16, 12, 8), counts=(20000, 12000), vth=(0.08, 0.01), q=(0.002, -0.002),
             qm=(1.0, -0.01), rm=(0.002, 0.2))


def small_state():
    s = SMALL
    return mt_particles(s["n"], s["counts"], s["vth"])


def test_push_es_fp64_bitwise():
    s = SMALL
    n = s["n"]
    rng = np.random.default_rng(1)
    eb = periodic_eb(n, rng, 0.05, 0.02)
    b0, e0 = (0.0, 0.01, 0.05), (0.001, 0.0, -0.002)
    p = small_state()
    with ctx_for(n, s["q"], s["qm"], s["rm"], s["counts"], b0=b0, e0=e0) as ctx:
        ctx.upload(p, s["counts"])
        ctx.set_fields(eb)
        ctx.push_es(2)
        ctx.push_es(2)
        ctx.push_es(1)
        g, cnt = ctx.download()
    o = p.copy()
    wrk = oracle.grdcrt(eb, n, e0, b0)
    sl = np.array(n, dtype=float)
    for us in (2, 2, 1):
        oracle.psolve2(o, s["counts"], s["qm"], wrk, n, us, sl)
    assert list(cnt) == list(s["counts"])
    for f in FIELDS:
        assert np.array_equal(g[f], o[f]), f
    assert np.array_equal(g["pid"], o["pid"])
    assert (g["spec"] == np.repeat([1, 2], s["counts"])).all()


def test_positn_fp64_bitwise():
    s = SMALL
    n = s["n"]
    p = small_state()
    with ctx_for(n, s["q"], s["qm"], s["rm"], s["counts"]) as ctx:
        ctx.upload(p, s["counts"])
        ctx.positn(1)
        g, _ = ctx.download()
    o = p.copy()
    oracle.positn(o, 1, np.array(n, dtype=float))
    for f in FIELDS:
        assert np.array_equal(g[f], o[f]), f


def test_fp32_positn_and_rebucket():
    """FP32 form: positn (the push kernel without gather and deposit) against
    the oracle's positn from the same rounded state; emg_sort (a re-bucketing
    of the tiles) is a permutation: the particles (by pid) and the fixed-point
    charge are bit-identical before and after it, and a push after it agrees
    with the oracle as any push does."""
    s = SMALL
    n = s["n"]
    p = small_state()
    rng = np.random.default_rng(6)
    eb = periodic_eb(n, rng, 0.05, 0.02)
    b0 = (0.0, 0.0, 0.02)
    sl = np.array(n, dtype=float)
    with ctx_for(n, s["q"], s["qm"], s["rm"], s["counts"], precision=emg.FP32_FIXED, b0=b0,
                 tile=4, track_pid=True) as ctx:
        ctx.upload(p, s["counts"])
        start, _ = ctx.download()
        ctx.positn(1)
        after_positn, _ = ctx.download()
        ctx.deposit_rho(0)
        rho_a = [ctx.rho(species=k) for k in range(2)]
        ctx.sort()
        ctx.deposit_rho(0)
        rho_b = [ctx.rho(species=k) for k in range(2)]
        after_sort, cnt = ctx.download()
        ctx.set_fields(eb)
        ctx.push_es(2)
        g, _ = ctx.download()
    assert list(cnt) == list(s["counts"])
    o = by_pid(start)
    oracle.positn(o, 1, sl)
    a = by_pid(after_positn)
    for f, L in zip(("x", "y", "z"), n):
        d = np.abs(a[f] - o[f])
        assert np.minimum(d, L - d).max() <= 1e-5, f
    b = by_pid(after_sort)
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f
    for k in range(2):
        assert np.array_equal(rho_a[k], rho_b[k]), k
    o = b.copy()
    oracle.psolve2(o, s["counts"], s["qm"], oracle.grdcrt(eb, n, (0, 0, 0), b0), n, 2, sl)
    g = by_pid(g)
    for f, L in zip(("x", "y", "z"), n):
        d = np.abs(g[f] - o[f])
        assert np.minimum(d, L - d).max() <= 1e-5, f


def test_deposit_fp64_vs_oracle():
    s = SMALL
    n = s["n"]
    p = small_state()
    lo, hi = (-1, -1, -1), tuple(v + 1 for v in n)
    with ctx_for(n, s["q"], s["qm"], s["rm"], s["counts"]) as ctx:
        ctx.upload(p, s["counts"])
        ctx.deposit_rho(2)
        full = ctx.rho(lo=lo, hi=hi)
        per = [ctx.rho(species=k) for k in range(2)]
        per_full = [ctx.rho(lo=lo, hi=hi, species=k) for k in range(2)]
    rho, rhodg = oracle.charge(p, s["counts"], s["q"], n, with_species=True)
    oracle.fold_rho(rho, n)
    oracle.fold_rho(rhodg, n, ncomp=2)
    # the whole rho(1, -1:n+1)^3: nodes 0..n folded, and the ghost planes as
    # the border exchange + add_boundary_charge leave them (raw node n in
    # ghost -1, raw node 0 in ghost n + 1), which differ from periodic images
    scale = np.abs(rho).max()
    assert np.abs(full - rho).max() <= 1e-12 * scale
    assert np.abs(rho[0] - rho[n[2]]).max() > 1e-3 * scale
    for k in range(2):
        ok = oracle.rho_nodes(rhodg)[..., k]
        assert np.abs(per[k] - ok).max() <= 1e-12 * np.abs(ok).max()
        assert np.abs(per_full[k] - rhodg[..., k]).max() <= 1e-12 * np.abs(rhodg[..., k]).max()


def test_energy_fp64_vs_oracle():
    s = SMALL
    n = s["n"]
    p = small_state()
    with ctx_for(n, s["q"], s["qm"], s["rm"], s["counts"]) as ctx:
        ctx.upload(p, s["counts"])
        e = ctx.energy()
    o = oracle.energy(p, s["counts"], s["rm"], np.array(n, dtype=float))
    assert np.allclose(e, o, rtol=1e-13, atol=0)


def test_es32_loop_fp64_matches_oracle():
    """The CPU-runnable workload (32^3, 2 x 524,288 particles, dt 0.1): the
    GPU path and the oracle run the reference's ES loop side by side for 100
    steps with their own host field solves."""
    w = units.es32()
    su = units.to_sunits(w)
    n = w.n
    p0 = emg.host_inipcl(units.host_species(w), n, su.t, w.seed)
    units.scale_to_s(p0, su)
    counts = su.npin
    from ctca_amd import driver
    run = driver.ESRun(w, particles=(p0, counts))
    run.initialize()           # a fresh job: no positn (esses.F90:198-201)
    o = p0.copy()
    sl = np.array(n, dtype=float)

    def o_rho(o):
        r = oracle.charge(o, counts, su.q, n)
        oracle.fold_rho(r, n)
        return oracle.rho_nodes(r).copy() + fieldsolve.background(su.q, su.npin, n)

    rho = o_rho(o)
    nsteps = w.steps
    max_rho_err = 0.0
    for k in range(1, nsteps + 1):
        ustep = 2 if k != nsteps else 1
        eb = oracle.fields.eb_from_phi(oracle.fields.poisson_periodic(rho))
        wrk = oracle.grdcrt(eb, n, su.e0, su.b0)
        oracle.psolve2(o, counts, su.qm, wrk, n, ustep, sl)
        rho = o_rho(o)
        run.step(ustep)
        max_rho_err = max(max_rho_err, np.abs(run.rho - rho).max() / np.abs(rho).max())
    g, _ = run.ctx.download()
    ek = run.ctx.energy()
    run.close()
    ok = oracle.energy(o, counts, su.rm, sl)
    dx = max(np.abs(g[f] - o[f]).max() for f in ("x", "y", "z"))
    dv = max(np.abs(g[f] - o[f]).max() for f in ("vx", "vy", "vz")) / np.abs(o["vx"]).max()
    # trajectories are chaotic: round-off of atomics/FFT grows; 100 steps stay far below 1e-6
    assert dx < 1e-6 and dv < 1e-6, (dx, dv)
    assert max_rho_err < 1e-6
    assert np.allclose(ek, ok, rtol=1e-9)


@pytest.mark.parametrize("tile", [4, 8])
def test_fp32_fixed_one_step_vs_fp64_oracle(tile):
    n = (16, 16, 8)
    counts = (40000, 30000)
    q, qm, rm = (0.002, -0.002), (1.0, -0.01), (0.002, 0.2)
    rng = np.random.default_rng(3)
    eb = periodic_eb(n, rng, 0.05, 0.01)
    b0 = (0.0, 0.0, 0.0125)
    p = mt_particles(n, counts, (0.05, 0.005))
    lo, hi = (-1, -1, -1), tuple(v + 1 for v in n)
    with ctx_for(n, q, qm, rm, counts, precision=emg.FP32_FIXED, b0=b0, tile=tile,
                 track_pid=True) as ctx:
        ctx.upload(p, counts)
        # the FP32 state is the FP64 state rounded: compare against the oracle
        # started from exactly that rounded state
        start, _ = ctx.download()
        ctx.set_fields(eb)
        ctx.push_es(2)
        rho32 = [ctx.rho(species=k) for k in range(2)]
        full32 = [ctx.rho(lo=lo, hi=hi, species=k) for k in range(2)]
        g, cnt = ctx.download()
    o = by_pid(start)
    g = by_pid(g)
    assert list(cnt) == list(counts)
    wrk = oracle.grdcrt(eb, n, (0, 0, 0), b0)
    oracle.psolve2(o, counts, qm, wrk, n, 2, np.array(n, dtype=float))
    _, rhodg = oracle.charge(o, counts, q, n, with_species=True)
    oracle.fold_rho(rhodg, n, ncomp=2)
    # positions: periodic distance in cells
    for f, L in zip(("x", "y", "z"), n):
        d = np.abs(g[f] - o[f])
        d = np.minimum(d, L - d)
        assert d.max() <= 1e-5, f
    vth = np.array([0.05, 0.005])[np.repeat([0, 1], counts)]
    for f in ("vx", "vy", "vz"):
        rel = np.abs(g[f] - o[f]) / np.maximum(np.abs(o[f]), vth)
        assert rel.max() <= 1e-5, f
    for k in range(2):
        ok = oracle.rho_nodes(rhodg)[..., k]
        assert np.abs(rho32[k] - ok).max() <= 1e-6 * np.abs(ok).max(), k
        # ghost planes included (upper-face fixed-point deposit)
        assert np.abs(full32[k] - rhodg[..., k]).max() <= 1e-6 * np.abs(rhodg[..., k]).max(), k


def test_fp32_fixed_multistep_with_tile_crossings():
    """Six FP32 steps on a grid with at least three tiles per axis, fast enough
    that a few percent of the particles change tile every step: movers travel
    through the outboxes and are pulled by their new tiles.  Every particle is
    kept, the trajectories follow the FP64 oracle, and the standalone deposit
    and energy of the final state (which includes the arrivals still waiting
    in outboxes) match the oracle's."""
    n = (32, 24, 16)
    counts = (60000, 30000)
    q, qm, rm = (0.002, -0.002), (1.0, -0.01), (0.002, 0.2)
    rng = np.random.default_rng(11)
    eb = periodic_eb(n, rng, 0.02, 0.005)
    b0 = (0.0, 0.01, 0.02)
    vth = (0.3, 0.02)
    p = mt_particles(n, counts, vth)
    nsteps = 6
    with ctx_for(n, q, qm, rm, counts, precision=emg.FP32_FIXED, b0=b0, tile=4,
                 track_pid=True) as ctx:
        ctx.upload(p, counts)
        start, _ = ctx.download()
        ctx.set_fields(eb)
        for _ in range(nsteps):
            ctx.push_es(2)
        ctx.deposit_rho(0)
        rho32 = [ctx.rho(species=k) for k in range(2)]
        full32 = [ctx.rho(lo=(-1, -1, -1), hi=tuple(v + 1 for v in n), species=k)
                  for k in range(2)]
        ek32 = ctx.energy()
        g, cnt = ctx.download()
    assert list(cnt) == list(counts)
    o = by_pid(start)
    g = by_pid(g)
    assert np.array_equal(g["pid"], o["pid"])
    moved = 0
    wrk = oracle.grdcrt(eb, n, (0, 0, 0), b0)
    for _ in range(nsteps):
        before = np.floor(o["x"] / 4) + 100 * np.floor(o["y"] / 4) + 1e4 * np.floor(o["z"] / 4)
        oracle.psolve2(o, counts, qm, wrk, n, 2, np.array(n, dtype=float))
        after = np.floor(o["x"] / 4) + 100 * np.floor(o["y"] / 4) + 1e4 * np.floor(o["z"] / 4)
        moved += int(np.count_nonzero(before != after))
    assert moved > 0.02 * nsteps * sum(counts)   # the outbox path carries real traffic
    for f, L in zip(("x", "y", "z"), n):
        d = np.abs(g[f] - o[f])
        assert np.minimum(d, L - d).max() <= 1e-4, f
    vt = np.array(vth)[np.repeat([0, 1], counts)]
    for f in ("vx", "vy", "vz"):
        assert (np.abs(g[f] - o[f]) / np.maximum(np.abs(o[f]), vt)).max() <= 1e-4, f
    _, rhodg = oracle.charge(o, counts, q, n, with_species=True)
    oracle.fold_rho(rhodg, n, ncomp=2)
    for k in range(2):
        ok = oracle.rho_nodes(rhodg)[..., k]
        assert np.abs(rho32[k] - ok).max() <= 1e-5 * np.abs(ok).max(), k
        assert np.abs(full32[k] - rhodg[..., k]).max() <= 1e-5 * np.abs(ok).max(), k
    ok = oracle.energy(o, counts, rm, np.array(n, dtype=float))
    assert np.allclose(ek32, ok, rtol=1e-5)


def test_fp32_ramp_density_with_b_along_x_vs_oracle():
    """Strongly non-uniform buckets: an exp(z) density ramp of 100 across the
    box (the 1G configuration's profile) generated on the device, B along x,
    8^3 tiles.  Three FP32 steps follow the oracle's trajectories to 1e-4 and
    the per-species density to 1e-5 of its maximum; every particle is kept."""
    n = (32, 32, 16)
    counts = (300000, 300000)
    q, qm, rm = (-0.002, 0.002), (-1.0, 0.01), (0.002, 0.2)
    t = units.rotation(90.0, 0.0)
    spec = [dict(count=counts[0], vth_perp=0.1, vth_para=0.1, t=t, ramp_ratio=100.0),
            dict(count=counts[1], vth_perp=0.01, vth_para=0.01, t=t, ramp_ratio=100.0)]
    with ctx_for(n, q, qm, rm, counts) as c64:
        c64.init_device(spec, 77)
        p, _ = c64.download()
    p["pid"] = np.arange(len(p))
    rng = np.random.default_rng(21)
    eb = periodic_eb(n, rng, 0.02, 0.005)
    b0 = (0.02, 0.0, 0.0)
    nsteps = 3
    with ctx_for(n, q, qm, rm, counts, precision=emg.FP32_FIXED, b0=b0, tile=8,
                 track_pid=True) as ctx:
        ctx.upload(p, counts)
        start, _ = ctx.download()
        ctx.set_fields(eb)
        for _ in range(nsteps):
            ctx.push_es(2)
        rho32 = [ctx.rho(species=k) for k in range(2)]
        g, cnt = ctx.download()
    assert list(cnt) == list(counts)
    o = by_pid(start)
    g = by_pid(g)
    assert np.array_equal(g["pid"], o["pid"])
    zc = np.histogram(o["z"][:counts[0]], bins=2, range=(0, n[2]))[0]
    assert zc[1] > 5 * zc[0]   # the buckets differ several-fold in size
    wrk = oracle.grdcrt(eb, n, (0, 0, 0), b0)
    for _ in range(nsteps):
        oracle.psolve2(o, counts, qm, wrk, n, 2, np.array(n, dtype=float))
    for f, L in zip(("x", "y", "z"), n):
        d = np.abs(g[f] - o[f])
        assert np.minimum(d, L - d).max() <= 1e-4, f
    vt = np.array([0.1, 0.01])[np.repeat([0, 1], counts)]
    for f in ("vx", "vy", "vz"):
        assert (np.abs(g[f] - o[f]) / np.maximum(np.abs(o[f]), vt)).max() <= 1e-4, f
    _, rhodg = oracle.charge(o, counts, q, n, with_species=True)
    oracle.fold_rho(rhodg, n, ncomp=2)
    for k in range(2):
        ok = oracle.rho_nodes(rhodg)[..., k]
        assert np.abs(rho32[k] - ok).max() <= 1e-5 * np.abs(ok).max(), k


def test_fp32_sustained_outbox_overflow():
    """Thousands of particles leave each 8^3 tile every step, far more than an
    outbox holds: the excess is placed directly or spilled, and each step's
    spill joins the lists of the buckets it lies in on the next step.  Over
    six steps nothing is lost and the trajectories and density follow the
    oracle."""
    n = (16, 16, 16)
    counts = (400000, 100000)
    q, qm, rm = (-0.002, 0.002), (-1.0, 0.01), (0.002, 0.2)
    rng = np.random.default_rng(31)
    eb = periodic_eb(n, rng, 0.02, 0.005)
    b0 = (0.0, 0.0, 0.02)
    vth = (0.1, 0.01)
    p = mt_particles(n, counts, vth, seed=12)
    nsteps = 6
    with ctx_for(n, q, qm, rm, counts, precision=emg.FP32_FIXED, b0=b0, tile=8,
                 track_pid=True) as ctx:
        ctx.upload(p, counts)
        start, _ = ctx.download()
        ctx.set_fields(eb)
        for _ in range(nsteps):
            ctx.push_es(2)
        rho32 = [ctx.rho(species=k) for k in range(2)]
        ek32 = ctx.energy()
        g, cnt = ctx.download()
    assert list(cnt) == list(counts)
    o = by_pid(start)
    g = by_pid(g)
    assert np.array_equal(g["pid"], o["pid"])
    wrk = oracle.grdcrt(eb, n, (0, 0, 0), b0)
    most = 0
    for _ in range(nsteps):
        e = o[:counts[0]]
        before = np.floor(e["x"] / 8) + 2 * np.floor(e["y"] / 8) + 4 * np.floor(e["z"] / 8)
        oracle.psolve2(o, counts, qm, wrk, n, 2, np.array(n, dtype=float))
        e = o[:counts[0]]
        after = np.floor(e["x"] / 8) + 2 * np.floor(e["y"] / 8) + 4 * np.floor(e["z"] / 8)
        left = np.bincount(before[before != after].astype(int), minlength=8)
        most = max(most, int(left.max()))
    assert most > 4 * 512   # every step overflows the outboxes several times over
    for f, L in zip(("x", "y", "z"), n):
        d = np.abs(g[f] - o[f])
        assert np.minimum(d, L - d).max() <= 1e-4, f
    vt = np.array(vth)[np.repeat([0, 1], counts)]
    for f in ("vx", "vy", "vz"):
        assert (np.abs(g[f] - o[f]) / np.maximum(np.abs(o[f]), vt)).max() <= 1e-4, f
    _, rhodg = oracle.charge(o, counts, q, n, with_species=True)
    oracle.fold_rho(rhodg, n, ncomp=2)
    for k in range(2):
        ok = oracle.rho_nodes(rhodg)[..., k]
        assert np.abs(rho32[k] - ok).max() <= 1e-5 * np.abs(ok).max(), k
    assert np.allclose(ek32, oracle.energy(o, counts, rm, np.array(n, dtype=float)), rtol=1e-5)


def test_fixed_point_rho_is_order_and_tile_independent():
    n = (16, 16, 16)
    counts = (50000, 20000)
    q, qm, rm = (0.001, -0.001), (1.0, -0.01), (0.001, 0.1)
    p = mt_particles(n, counts, (0.05, 0.005))
    rng = np.random.default_rng(4)
    perm = np.concatenate([rng.permutation(counts[0]), counts[0] + rng.permutation(counts[1])])
    eb = periodic_eb(n, rng, 0.05, 0.0)
    outs = []
    for tile, order in ((4, None), (4, perm), (8, perm)):
        pp = p if order is None else p[order]
        with ctx_for(n, q, qm, rm, counts, precision=emg.FP32_FIXED, tile=tile) as ctx:
            pp = pp.copy()
            pp["pid"] = 0
            ctx.upload(pp, counts)
            ctx.deposit_rho(0)
            r0 = [ctx.rho(species=k) for k in range(2)]
            ctx.set_fields(eb)
            ctx.push_es(2)
            r1 = [ctx.rho(species=k) for k in range(2)]
            outs.append((r0, r1))
    for a, b in ((0, 1), (0, 2)):
        for st in range(2):
            for k in range(2):
                assert np.array_equal(outs[a][st][k], outs[b][st][k]), (a, b, st, k)
    # exact charge conservation of the fixed-point weights
    tot = outs[0][1][0].sum()
    assert tot == pytest.approx(counts[0] * q[0], rel=1e-12)


def _cell_key(seg, n):
    return (np.floor(seg["z"]) * n[1] + np.floor(seg["y"])) * n[0] + np.floor(seg["x"])


def test_sort_fp64_is_a_z_slab_ordered_permutation():
    """A sort of a random order (the record pass) and a re-sort five pushes
    later (the direct pass, whose histogram groups each wave's lanes by cell):
    both leave every species segment in cell order (z-major, so floor(z) never
    decreases: an order the reference's comparator accepts) and permute the
    particles without changing a bit of them."""
    s = SMALL
    n = s["n"]
    p = small_state()
    rng = np.random.default_rng(5)
    eb = periodic_eb(n, rng, 0.05, 0.02)
    with ctx_for(n, s["q"], s["qm"], s["rm"], s["counts"], b0=(0.0, 0.0, 0.03)) as ctx:
        ctx.upload(p, s["counts"])
        ctx.deposit_rho(0)
        r0 = ctx.rho()
        ctx.sort()
        ctx.deposit_rho(0)
        r1 = ctx.rho()
        g, cnt = ctx.download()
        ctx.set_fields(eb)
        for _ in range(5):
            ctx.push_es(2)
        moved, _ = ctx.download()
        ctx.sort()
        g2, _ = ctx.download()
    for before, after in ((p, g), (moved, g2)):
        off = 0
        for c in s["counts"]:
            seg = after[off:off + c]
            assert np.all(np.diff(_cell_key(seg, n)) >= 0)
            assert np.all(np.diff(np.floor(seg["z"])) >= 0)
            # the reference comparator's property (psort-struct.c:15-20)
            assert np.all(np.diff(seg["z"]) > -1.0)
            # each species keeps its own particles
            assert np.array_equal(np.sort(seg["pid"]), np.sort(before[off:off + c]["pid"]))
            off += c
        gp, pp = by_pid(after), by_pid(before)
        for f in FIELDS:
            assert np.array_equal(gp[f], pp[f])
    # the pushes moved particles across cells, so the re-sort had work to do
    assert not np.array_equal(_cell_key(moved, n), _cell_key(g2, n))
    assert np.abs(r1 - r0).max() <= 1e-13 * np.abs(r0).max()


def test_sort_every_only_permutes():
    """sort_every = 3: the FP64 form sorts itself before the first push and
    before every third push after that; positions and velocities stay bit-identical to a
    run that never sorts (matched by pid), the densities and currents agree
    to round-off, and every species segment ends in the sort's order."""
    s = SMALL
    n = s["n"]
    rng = np.random.default_rng(3)
    eb = periodic_eb(n, rng, 0.05, 0.02)
    p = small_state()
    out = {}
    for every in (0, 3):
        with ctx_for(n, s["q"], s["qm"], s["rm"], s["counts"], b0=(0.0, 0.0, 0.03),
                     sort_every=every) as ctx:
            ctx.upload(p, s["counts"])
            ctx.set_fields(eb)
            for _ in range(6):
                ctx.push_es(2)  # every = 3: sorted before the 1st and 4th
            ctx.deposit_rho(0)
            rho = ctx.rho()
            ctx.push_em(2)      # the 7th push: sorted before it
            j = ctx.current()
            g, cnt = ctx.download()
        out[every] = (by_pid(g), rho, j, g)
    for f in FIELDS:
        assert np.array_equal(out[0][0][f], out[3][0][f]), f
    assert np.abs(out[3][1] - out[0][1]).max() <= 1e-13 * np.abs(out[0][1]).max()
    assert np.abs(out[3][2] - out[0][2]).max() <= 1e-12 * np.abs(out[0][2]).max()
    g = out[3][3]
    off = 0
    L = float(n[2])
    for c in s["counts"]:
        # sorted (floor z non-decreasing) before the 7th push, which moves a
        # particle by vz ustep: neighbours in the order differ by more than
        # -1 - 2 max|vz| ustep (a random order reaches -L/2)
        seg = g[off:off + c]
        d = np.diff(seg["z"])
        d = (d + L / 2) % L - L / 2     # across the periodic face
        assert np.all(d > -1.0 - 4.0 * np.abs(seg["vz"]).max() - 1e-12)
        off += c
    with pytest.raises(emg.EmgError):
        ctx_for(n, s["q"], s["qm"], s["rm"], s["counts"], precision=emg.FP32_FIXED,
                sort_every=3)
    with pytest.raises(emg.EmgError):
        ctx_for(n, s["q"], s["qm"], s["rm"], s["counts"], sort_every=-1)


@pytest.mark.parametrize("every", [1, 2])
def test_resorts_carried_by_pushes_only_permute(every):
    """sort_every = 1 or 2 over ES pushes, a positn and EM pushes: every
    re-sort after the first is carried by the push that follows it (the push
    writes each particle to its slot in the new order; with 1, each carried
    push runs on the order and run tables the previous one left), except the
    one before the positn, which sorts on its own.  Positions and velocities
    stay bit-identical to a run that never sorts (matched by pid), rho and J
    agree to round-off after every push, and each species keeps its own
    particles."""
    s = SMALL
    n = s["n"]
    rng = np.random.default_rng(7)
    eb = periodic_eb(n, rng, 0.05, 0.02)
    p = small_state()
    # with 2: sorts before pushes 1 (a sort of its own: no earlier order), 3
    # (carried), 5 (the positn: its own), 7 and 9 (carried by zigzag pushes)
    seq = ["es"] * 4 + ["positn"] + ["em"] * 4
    out = {}
    for se in (0, every):
        with ctx_for(n, s["q"], s["qm"], s["rm"], s["counts"], b0=(0.0, 0.0, 0.03),
                     sort_every=se) as ctx:
            ctx.upload(p, s["counts"])
            ctx.set_fields(eb)
            grids, states = [], []
            for op in seq:
                if op == "es":
                    ctx.push_es(2)
                    grids.append(ctx.rho())
                elif op == "em":
                    ctx.push_em(2)
                    grids.append(ctx.current())
                else:
                    ctx.positn(1)
                    grids.append(None)
                states.append(ctx.download()[0])
        out[se] = (grids, states)
    for k, op in enumerate(seq):
        a, b = out[0][1][k], out[every][1][k]
        ap, bp = by_pid(a), by_pid(b)
        for f in FIELDS:
            assert np.array_equal(ap[f], bp[f]), (k, op, f)
        off = 0
        for c in s["counts"]:
            assert np.array_equal(np.sort(a[off:off + c]["pid"]), np.sort(b[off:off + c]["pid"]))
            off += c
        if op != "positn":
            ga, gb = out[0][0][k], out[every][0][k]
            assert np.abs(gb - ga).max() <= 1e-12 * np.abs(ga).max(), (k, op)


def test_sorted_runs_on_a_nonuniform_grid():
    """After a sort the FP64 kernels take their blocks from the sort's run
    table (chunks of consecutive cells, each cut into runs of nearly equal
    length).  On a 30x26x22 grid (no multiple of the chunk sizes) whose
    density falls steeply in z -- chunks of thousands of particles, chunks of
    a few and empty ones -- sorting before every push leaves positions and
    velocities bit-identical to a run that never sorts (matched by pid), the
    densities and currents equal to round-off, and every particle's charge
    in rho."""
    n = (30, 26, 22)
    counts = (300_000, 60_000)
    q, qm, rm = (0.002, -0.002), (1.0, -0.01), (0.002, 0.2)
    p = mt_particles(n, counts, (0.08, 0.01), seed=12)
    # z = 0.8 nz u^4: hundreds per cell near z = 0, a few per cell near
    # z = 0.8 nz and none above
    p["z"] = 0.8 * n[2] * (p["z"] / n[2]) ** 4
    rng = np.random.default_rng(8)
    eb = periodic_eb(n, rng, 0.05, 0.02)
    out = {}
    for every in (0, 1):
        with ctx_for(n, q, qm, rm, counts, b0=(0.0, 0.0, 0.03), sort_every=every) as ctx:
            ctx.upload(p, counts)
            ctx.set_fields(eb)
            for _ in range(3):
                ctx.push_es(2)
            rho = [ctx.rho(species=k) for k in range(2)]
            ctx.push_em(2)
            j = ctx.current()
            g, cnt = ctx.download()
        out[every] = (by_pid(g), rho, j)
    for f in FIELDS:
        assert np.array_equal(out[0][0][f], out[1][0][f]), f
    for k in range(2):
        r0, r1 = out[0][1][k], out[1][1][k]
        assert np.abs(r1 - r0).max() <= 1e-13 * np.abs(r0).max(), k
        assert float(r1.sum()) == pytest.approx(q[k] * counts[k], rel=1e-12)
    assert np.abs(out[1][2] - out[0][2]).max() <= 1e-12 * np.abs(out[0][2]).max()


@pytest.mark.parametrize("count", [150_000, 400_000], ids=["es_runs_only", "no_runs"])
def test_sorted_runs_fall_back_on_crowded_grids(count):
    """More blocks of a run length than cells (over 1,024 or 4,096 particles
    per cell) leave that run table unbuilt and the kernels on their plain
    slices: 150,000 particles on 4x4x4 cells build the charge table only,
    400,000 build neither.  Sorting before every push still changes nothing
    but the order."""
    n = (4, 4, 4)
    counts = (count, count // 5)
    q, qm, rm = (0.002, -0.002), (1.0, -0.01), (0.002, 0.2)
    p = mt_particles(n, counts, (0.08, 0.01), seed=13)
    eb = periodic_eb(n, np.random.default_rng(9), 0.05, 0.02)
    out = {}
    for every in (0, 1):
        with ctx_for(n, q, qm, rm, counts, b0=(0.0, 0.0, 0.03), sort_every=every) as ctx:
            ctx.upload(p, counts)
            ctx.set_fields(eb)
            for _ in range(2):
                ctx.push_es(2)
            rho = ctx.rho()
            ctx.push_em(2)
            j = ctx.current()
            g, _ = ctx.download()
        out[every] = (by_pid(g), rho, j)
    for f in FIELDS:
        assert np.array_equal(out[0][0][f], out[1][0][f]), f
    assert np.abs(out[1][1] - out[0][1]).max() <= 1e-13 * np.abs(out[0][1]).max()
    assert np.abs(out[1][2] - out[0][2]).max() <= 1e-12 * np.abs(out[0][2]).max()


def test_push_em_fp64_vs_oracle():
    n = (12, 10, 8)
    counts = (20000, 5000)
    q, qm, rm = (0.002, 0.002), (1.0, 1.0), (0.002, 0.002)
    rng = np.random.default_rng(5)
    eb = periodic_eb(n, rng, 0.02, 0.01)
    p = mt_particles(n, counts, (0.1, 0.02), drift=(0.0, 0.15))
    with ctx_for(n, q, qm, rm, counts, b0=(0.0, 0.0, 0.02)) as ctx:
        ctx.upload(p, counts)
        ctx.set_fields(eb)
        ctx.push_em(2)
        J = ctx.current(lo=(-1, -1, -1), hi=tuple(v + 1 for v in n))
        g, _ = ctx.download()
    o = p.copy()
    wrk = oracle.grdcrt(eb, n, (0, 0, 0), (0.0, 0.0, 0.02))
    aj = oracle.psolve1(o, counts, qm, q, wrk, n, 2, np.array(n, dtype=float))
    oracle.fold_current(aj, n)
    for f in FIELDS:
        assert np.array_equal(g[f], o[f]), f
    oj = aj[3:n[2] + 6, 3:n[1] + 6, 3:n[0] + 6]   # indices -1 .. n+1
    assert np.abs(J - oj).max() <= 1e-12 * np.abs(oj).max()


@pytest.mark.parametrize("precision", [emg.PLAIN_FP64, emg.FP32_FIXED])
def test_ajdg_per_species_current(precision):
    """emg_get_ajdg (psolve1 func = 1, psolve1.F90:803-817): each species'
    q J alone; the species add up to aj, and in the FP64 form each matches the
    oracle's current with the other species' charge set to zero."""
    n = (12, 8, 8)
    counts = (20000, 5000)
    q, qm, rm = (0.002, -0.003), (1.0, -0.5), (0.002, 0.006)
    rng = np.random.default_rng(9)
    eb = periodic_eb(n, rng, 0.02, 0.01)
    b0 = (0.0, 0.0, 0.02)
    p = mt_particles(n, counts, (0.1, 0.05), drift=(0.0, 0.1))
    kw = dict(tile=4, track_pid=True) if precision == emg.FP32_FIXED else {}
    lo, hi = (-1, -1, -1), tuple(v + 1 for v in n)
    with ctx_for(n, q, qm, rm, counts, precision=precision, b0=b0, **kw) as ctx:
        ctx.upload(p, counts)
        start, _ = ctx.download()
        ctx.set_fields(eb)
        ctx.push_em(2)
        J = ctx.current(lo=lo, hi=hi)
        Js = [ctx.current(lo=lo, hi=hi, species=k) for k in range(2)]
        with pytest.raises(emg.EmgError):
            ctx.current(species=2)
    scale = np.abs(J).max()
    assert np.abs(Js[0] + Js[1] - J).max() <= 1e-15 * scale
    assert np.abs(Js[1]).max() > 0.1 * scale and np.abs(Js[0]).max() > 0.1 * scale
    if precision == emg.PLAIN_FP64:
        wrk = oracle.grdcrt(eb, n, (0, 0, 0), b0)
        for k in range(2):
            qk = tuple(q[s] if s == k else 0.0 for s in range(2))
            o = start.copy()
            aj = oracle.psolve1(o, counts, qm, qk, wrk, n, 2, np.array(n, dtype=float))
            oracle.fold_current(aj, n)
            oj = aj[3:n[2] + 6, 3:n[1] + 6, 3:n[0] + 6]
            assert np.abs(Js[k] - oj).max() <= 1e-12 * np.abs(oj).max(), k


@pytest.mark.parametrize("tile,n", [(4, (16, 12, 12)), (4, (12, 8, 8)), (8, (16, 16, 8))])
def test_push_em_fp32_fixed_vs_fp64_oracle(tile, n):
    """FP32 zigzag push: one step against the oracle's psolve1 from the same
    (FP32-rounded) state.  Positions 1e-5 cells, velocities 1e-5 relative, the
    folded fixed-point current 1e-5 of its maximum.  The grids have three
    or more tiles per axis, and two (the two neighbours along such an axis
    are the same tile).  A few tail particles move two cells in the step, some
    of them through a box face, which the reference treats by its own rule."""
    counts = (30000, 8000)
    q, qm, rm = (0.002, 0.002), (1.0, 1.0), (0.002, 0.002)
    rng = np.random.default_rng(5)
    eb = periodic_eb(n, rng, 0.02, 0.01)
    b0 = (0.0, 0.0, 0.02)
    p = mt_particles(n, counts, (0.15, 0.02), drift=(0.0, 0.15))
    with ctx_for(n, q, qm, rm, counts, precision=emg.FP32_FIXED, b0=b0, tile=tile,
                 track_pid=True) as ctx:
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
    for f, L in zip(("x", "y", "z"), n):
        d = np.abs(g[f] - o[f])
        assert np.minimum(d, L - d).max() <= 1e-5, f
    vt = np.array([0.15, 0.02])[np.repeat([0, 1], counts)]
    for f in ("vx", "vy", "vz"):
        assert (np.abs(g[f] - o[f]) / np.maximum(np.abs(o[f]), vt)).max() <= 1e-5, f
    oj = aj[3:n[2] + 6, 3:n[1] + 6, 3:n[0] + 6]
    assert np.abs(J - oj).max() <= 1e-5 * np.abs(oj).max()


def test_push_em_fp32_fixed_multistep_current():
    """Five FP32 zigzag steps with tile crossings (outboxes carry the movers);
    the current of the last step matches the oracle's psolve1 run on the same
    trajectory to 1e-4 of its maximum, and the kinetic energy to 1e-5."""
    n = (24, 16, 16)
    counts = (50000, 10000)
    q, qm, rm = (0.002, 0.002), (1.0, 1.0), (0.002, 0.002)
    rng = np.random.default_rng(9)
    eb = periodic_eb(n, rng, 0.01, 0.005)
    b0 = (0.01, 0.0, 0.02)
    p = mt_particles(n, counts, (0.3, 0.05), drift=(0.1, 0.0))
    with ctx_for(n, q, qm, rm, counts, precision=emg.FP32_FIXED, b0=b0, tile=4,
                 track_pid=True) as ctx:
        ctx.upload(p, counts)
        start, _ = ctx.download()
        ctx.set_fields(eb)
        for _ in range(5):
            ctx.push_em(2)
        J = ctx.current(lo=(0, 0, 0), hi=tuple(v - 1 for v in n))
        ek = ctx.energy()
        _, cnt = ctx.download()
    assert list(cnt) == list(counts)
    o = by_pid(start)
    wrk = oracle.grdcrt(eb, n, (0, 0, 0), b0)
    sl = np.array(n, dtype=float)
    for _ in range(5):
        aj = oracle.psolve1(o, counts, qm, q, wrk, n, 2, sl)
    oracle.fold_current(aj, n)
    oj = oracle.aj_cells(aj)
    assert np.abs(J - oj).max() <= 1e-4 * np.abs(oj).max()
    assert np.allclose(ek, oracle.energy(o, counts, rm, sl), rtol=1e-5)


def test_em_gpu_current_satisfies_planar_continuity():
    n = (12, 10, 8)
    counts = (30000,)
    rng = np.random.default_rng(6)
    p = mt_particles(n, counts, (0.12,))
    p["vz"] = 0.0
    for f in ("vx", "vy"):
        p[f] = np.clip(p[f], -0.4, 0.4)
    with ctx_for(n, (0.01,), (1.0,), (0.01,), counts, b0=(0, 0, 0.05)) as ctx:
        ctx.upload(p, counts)
        ctx.deposit_rho(0)
        r0 = ctx.rho()
        ctx.set_fields(np.zeros((n[2] + 3, n[1] + 3, n[0] + 3, 6)))
        ctx.push_em(2)
        J = ctx.current()
        ctx.deposit_rho(0)
        r1 = ctx.rho()
    div = (J[..., 0] - np.roll(J[..., 0], 1, 2) + J[..., 1] - np.roll(J[..., 1], 1, 1)
           + J[..., 2] - np.roll(J[..., 2], 1, 0))
    assert np.abs(r0 - 2.0 * div - r1).max() <= 1e-12 * np.abs(r1).max()


def test_upload_compaction_and_errors():
    n = (8, 8, 8)
    p = mt_particles(n, (100, 50), (0.1, 0.1))
    p["nid"][[3, 10, 120]] = -1
    with ctx_for(n, (1.0, 1.0), (1.0, 1.0), (1.0, 1.0), (100, 50)) as ctx:
        ctx.upload(p, (100, 50))
        g, cnt = ctx.download()
        assert list(cnt) == [98, 49]
        keep = p[p["nid"] >= 0]
        assert np.array_equal(g["pid"], keep["pid"])
        with pytest.raises(emg.EmgError) as e:
            ctx.upload(mt_particles(n, (101, 50), (0.1, 0.1)), (101, 50))
        assert e.value.code == -7
        bad = mt_particles(n, (100, 50), (0.1, 0.1))
        bad["x"][5] = 8.5
        with pytest.raises(emg.EmgError) as e:
            ctx.upload(bad, (100, 50))
        assert e.value.code == -5
        with pytest.raises(emg.EmgError) as e:
            ctx.push_es(2)            # no fields yet
        assert e.value.code in (-4,)


def test_device_init_statistics():
    n = (32, 32, 32)
    counts = (400000, 200000)
    t = units.rotation(90.0, 0.0)
    spec = [dict(count=counts[0], vth_perp=0.05, vth_para=0.02, drift=(0.01, 0, 0), t=t),
            dict(count=counts[1], vth_perp=0.005, vth_para=0.005, t=t, ramp_ratio=100.0)]
    with ctx_for(n, (1e-3, -1e-3), (1.0, -0.01), (1e-3, 0.1), counts) as ctx:
        ctx.init_device(spec, 1234)
        g, cnt = ctx.download()
        ctx2 = ctx_for(n, (1e-3, -1e-3), (1.0, -0.01), (1e-3, 0.1), counts)
        ctx2.init_device(spec, 1234)
        g2, _ = ctx2.download()
        ctx2.close()
    assert list(cnt) == list(counts)
    for f in FIELDS:
        assert np.array_equal(g[f], g2[f])   # deterministic stream
    e = g[:counts[0]]
    # B along x (phiz = 90): v_para is the x component after rotation
    assert abs(e["vx"].std() - 0.02) < 2e-4 and abs(e["vy"].std() - 0.05) < 5e-4
    assert abs(e["vx"].mean() - 0.01) < 2e-4
    for f, L in zip(("x", "y", "z"), n):
        assert e[f].min() >= 0 and e[f].max() < L
    z = g[counts[0]:]["z"]
    h, _ = np.histogram(z, bins=8, range=(0, n[2]))
    ratio = h[-1] / h[0]
    assert 40 < ratio < 90   # exp ramp of 100 across the box: bin-average ratio ~ 100^(7/8)


def test_fp32_spill_path_keeps_every_particle():
    """Move 600 particles of one tile into its neighbour in one step: 512 fit
    the outbox, the rest are placed directly (the destination's 16 end slots,
    then the spill region, which the next step processes); nothing is lost and
    the result matches the FP64 oracle."""
    n = (16, 16, 16)
    k = 600
    p = np.zeros(k + 20, dtype=emg.PARTICLE)
    rng = np.random.default_rng(7)
    p["x"][:k] = 4.0 + rng.uniform(0.2, 0.5, k)      # tile x in [4, 8)
    p["y"][:k] = rng.uniform(0.5, 3.5, k)
    p["z"][:k] = rng.uniform(0.5, 3.5, k)
    p["vx"][:k] = -0.4                                  # moves -0.8: into tile x in [0, 4)
    p["x"][k:] = rng.uniform(0.5, 3.5, 20)
    p["y"][k:] = rng.uniform(0.5, 3.5, 20)
    p["z"][k:] = rng.uniform(0.5, 3.5, 20)
    p["pid"] = np.arange(k + 20)
    counts = (k + 20,)
    eb = np.zeros((n[2] + 3, n[1] + 3, n[0] + 3, 6))
    with ctx_for(n, (1e-3,), (1.0,), (1e-3,), counts, precision=emg.FP32_FIXED, track_pid=True) as ctx:
        ctx.upload(p, counts)
        start, _ = ctx.download()
        ctx.set_fields(eb)
        ctx.push_es(2)
        ctx.push_es(2)
        r = ctx.rho()
        g, cnt = ctx.download()
    assert cnt[0] == k + 20
    o = by_pid(start)
    wrk = oracle.grdcrt(eb, n, (0, 0, 0), (0, 0, 0))
    for _ in range(2):
        oracle.psolve2(o, counts, (1.0,), wrk, n, 2, np.array(n, dtype=float))
    g = by_pid(g)
    assert np.array_equal(g["pid"], o["pid"])
    for f, L in zip(("x", "y", "z"), n):
        d = np.abs(g[f] - o[f])
        assert np.minimum(d, L - d).max() <= 1e-5
    ro = oracle.charge(o, counts, (1e-3,), n)
    oracle.fold_rho(ro, n)
    assert np.abs(r - oracle.rho_nodes(ro)).max() <= 1e-6 * np.abs(oracle.rho_nodes(ro)).max()
