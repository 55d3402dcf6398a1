"""Benchmark of the particle hot path on one MI355X.

Workload (default): the magnetised scale-up -- 128^3 periodic cells, two
species of 64 particles per cell (268,435,456 particles), B along z with
wce/wpe = 0.5, vth_e 1.0, vth_i 0.05, mass ratio 100, dt 0.05.  Particles
are generated on the device from a seeded counter-based stream; the fields
are E = -grad(phi) of the initial charge density (host Poisson solve) plus
the uniform B0.

A step is one push + deposit pass over every particle:
  * FP32 form (the headline `value`): emg_push_es in EMG_FP32_FIXED -- one
    fused kernel: LDS-staged gather, FP32 Boris, cell+offset move, int64
    fixed-point CIC deposit, re-bucketing;
  * FP64 form (beside it): emg_push_es + emg_deposit_rho in EMG_PLAIN_FP64,
    the reference arithmetic (psolve2 then charge).
The host field solve is not part of a step (it stays on the host in the
reference); inputs are resident in HBM before the timed region.  The FP64
form is sorted once before its steps (spsort at initialisation; the
reference re-sorts every isort = 500 steps), and the sort's kernel time is
reported beside it, also amortised over 500 steps.

Further legs, reported inside `config` (each can be switched off):
  * adjoint: the 64^3 ES window of 50 steps with checkpoints every 10 --
    forward time, reverse-sweep time and their ratio, gradients against
    central differences;
  * em_twostream: the EM two-stream workload (256x256x128, 295,279,002
    particles), psolve1 push + zigzag current, FP32 form with int64
    fixed-point J beside the FP64 form, and the whole EM step with the Yee
    update on the device (bfield, grdcrt, push, bfield, efield);
  * ramp_densest_slab: the densest 512x512x32 z-slab of the 1G exp(z) ramp
    (474,679,216 particles), the worst-loaded share of 8 GPUs;
  * ramp_full_1G: the whole 512x512x256 exp(z) ramp, 1,073,741,824 particles
    on one GPU, both forms (one context after the other);
  * energy_drift: the plain ES configuration (32^3, 1,048,576 particles) run
    self-consistently on the device (deposit, field solve, push) in both
    forms, with the drift of kinetic + field energy;
  * em_energy_drift: the EM two-stream workload at full size with the Yee
    update on the device, 100 steps in both forms, the same for kinetic +
    electric + magnetic energy.

Prints one JSON line.  `roofline` is the fused FP32 kernel: algorithmic bytes
per launch (56 B per particle + 24 B per field node + 8 B per charge node and
species) over its average launch time from HIP events on the context stream.
`roofline.copy_gbs` is a plain 1 GiB device copy timed on the same GPU (read +
write bytes over its time), the practical ceiling of a kernel that streams
its state once, and `frac_of_copy` the kernel's share of it.
`roofline.traffic` is the HBM traffic per launch measured with rocprofv3
FETCH_SIZE / WRITE_SIZE counters on the same workload (profiles/, read here
when it was measured for this workload and tile; else null).
`cpu_baseline` times the oracle (FP64 C restatement of psolve2 + charge +
fold, one thread) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s push+deposit, FP32/fixed-pt vs FP64 GPU; adjoint/forward time"
HBM_PEAK_GBS = 8000.0
# The FP64 form re-sorts itself by cell before every 10th push (emg_config
# sort_every): its LDS windows need cell-ordered runs, and thermal drift
# spreads a run beyond its window within ~20 steps (at the reference's CPU
# cadence, isort = 500, the deposit slows from 6 to 35 ms in 60 steps on the
# 128^3 workload).  With the fused push + charge kernel, 48 steps on the
# 128^3 workload cost 10.2, 9.9, 9.2, 9.4, 9.6 and 11.7 ms per step, sorts
# included, at a cadence of 6, 8, 10, 12, 16 and 24.  A sort only permutes
# the particles.
FP64_SORT_EVERY = 10
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_fx_step_mag128_pmc.txt")


def measured_traffic(workload, tile):
    """HBM bytes per launch of the step kernel from the committed PMC summary
    (lines `key = value`), if it was measured for this workload and tile."""
    try:
        kv = {}
        for line in open(TRAFFIC_FILE):
            if "=" in line and not line.lstrip().startswith("#"):
                k, v = line.split("=", 1)
                kv[k.strip()] = v.strip()
        if kv.get("workload") != workload or int(kv.get("tile", -1)) != tile:
            return None
        return float(kv["traffic_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """Read + write bandwidth of a plain device-to-device copy of `nbytes` on
    this GPU (torch copy_, timed with events on its stream): the practical
    ceiling of a kernel that streams its state in and out once."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def build_fields(ctx, su, n):
    """E of the initial density (host Poisson) as the eb halo array."""
    from ctca_amd import fieldsolve
    ctx.deposit_rho(0)
    rho = ctx.rho() + fieldsolve.background(su.q, su.npin, n)
    phi = fieldsolve.poisson(rho)
    return fieldsolve.eb_halo(fieldsolve.efield(phi))


def time_steps(ctx, form, steps, warmup):
    from ctca_amd import emg

    def one():
        if form == "em":
            ctx.push_em(2)
            return
        ctx.push_es(2)
        if form == "fp64":
            ctx.deposit_rho(0)

    for _ in range(warmup):
        one()
    ctx.synchronize()
    ctx.enable_timing(True)
    ctx.reset_stats()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
    ctx.synchronize()
    t1 = time.perf_counter()
    stats = {k: ctx.kernel_stats(k) for k in (emg.KERNEL_PUSH, emg.KERNEL_DEPOSIT)}
    ctx.enable_timing(False)
    return t1 - t0, stats


def auto_tile(n, tile_arg=0):
    return tile_arg if tile_arg > 0 else (8 if all(v % 8 == 0 for v in n) else 4)


def kernel_ms(stats, which):
    ms, count = stats[which]
    return ms / max(count, 1)


def timed_sort(ctx):
    """emg_sort under the kernel timers; returns its kernel time in ms."""
    from ctca_amd import emg
    ctx.synchronize()
    ctx.enable_timing(True)
    ctx.reset_stats()
    ctx.sort()
    ctx.synchronize()
    ms = kernel_ms({emg.KERNEL_SORT: ctx.kernel_stats(emg.KERNEL_SORT)}, emg.KERNEL_SORT)
    ctx.enable_timing(False)
    return ms


def em_leg(steps32, steps64, warmup, tile_arg=0):
    """The EM two-stream workload (256x256x128 periodic cells; 268,435,456
    background electrons at 32 per cell, vth 1.0, plus 26,843,546 beam
    electrons drifting +5.0 along x; immobile neutralising ions): a step is
    the psolve1 push + Umeda zigzag current.  FP32 cell+offset form with
    int64 fixed-point J beside the plain FP64 form (sorted at initialisation,
    as esses/emses sort at step 0); E of the initial density, no B0."""
    from ctca_amd import emg, units
    w = units.em_twostream()
    su = units.to_sunits(w)
    n = w.n
    nsp = len(w.species)
    counts = su.npin
    npart = int(counts.sum())
    ncell = n[0] * n[1] * n[2]
    spec = units.device_species(w, su)
    tile = auto_tile(n, tile_arg)
    ctx = emg.Context(n, su.q, su.qm, su.rm, counts, precision=emg.FP32_FIXED, e0=su.e0,
                      b0=su.b0, tile=tile)
    try:
        ctx.init_device(spec, w.seed)
        eb = build_fields(ctx, su, n)
        ctx.set_fields(eb)
        ctx.push_em(2)
        j32 = ctx.current()
        el32, st32 = time_steps(ctx, "em", steps32, warmup)
        # the whole EM step with the fields on the device, in the reference's
        # order (emses.F90:365-434): bfield, grdcrt, psolve1 + current fold,
        # bfield, efield; tcs = 2 cv^2 in S-units (inital.F90:78)
        tcs = 2.0 * su.cs

        def full():
            ctx.bfield()
            ctx.grdcrt()
            ctx.push_em(2)
            ctx.bfield()
            ctx.efield(tcs, 2.0)

        for _ in range(warmup):
            full()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps32):
            full()
        ctx.synchronize()
        el_full = time.perf_counter() - t0
    finally:
        ctx.close()
    k32 = kernel_ms(st32, emg.KERNEL_PUSH)
    bytes32 = 56.0 * npart + 24.0 * ncell + 24.0 * ncell * nsp
    out = {"workload": w.name, "grid": list(n), "particles": npart, "tile": tile,
           "fp32": {"particle_steps_per_s": npart * steps32 / el32,
                    "ms_per_step": el32 / steps32 * 1e3, "steps": steps32, "kernel_ms": k32,
                    "bytes_per_launch": bytes32,
                    "gbs_algorithmic": bytes32 / (k32 * 1e-3) / 1e9,
                    "hbm_frac": bytes32 / (k32 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "device_resident_step_ms": el_full / steps32 * 1e3,
                    "device_resident_particle_steps_per_s": npart * steps32 / el_full}}
    if steps64:
        ctx = emg.Context(n, su.q, su.qm, su.rm, counts, precision=emg.PLAIN_FP64, e0=su.e0,
                          b0=su.b0, sort_every=FP64_SORT_EVERY)
        try:
            ctx.init_device(spec, w.seed)
            sort_ms = timed_sort(ctx)
            ctx.set_fields(eb)
            ctx.push_em(2)
            j64 = ctx.current()
            el64, st64 = time_steps(ctx, "em", steps64, warmup)
        finally:
            ctx.close()
        k64 = kernel_ms(st64, emg.KERNEL_PUSH)
        bytes64 = 96.0 * npart + 48.0 * (n[0] + 1) * (n[1] + 1) * (n[2] + 1) + 24.0 * ncell * nsp
        v64 = npart * steps64 / el64
        out["fp64"] = {"particle_steps_per_s": v64, "ms_per_step": el64 / steps64 * 1e3,
                       "steps": steps64, "kernel_ms": k64,
                       "gbs_algorithmic": bytes64 / (k64 * 1e-3) / 1e9,
                       "sort_ms": sort_ms, "sort_every": FP64_SORT_EVERY}
        out["fp32_over_fp64"] = out["fp32"]["particle_steps_per_s"] / v64
        # the zigzag current is discontinuous in the end point where the relay
        # point switches from a midpoint to a cell face, so the few particles
        # whose FP32 and FP64 end points straddle a face set the maximum; the
        # RMS shows the arithmetic
        d = j32 - j64
        out["current_err_fp32_vs_fp64_one_step"] = {
            "max_over_max": float(np.abs(d).max() / np.abs(j64).max()),
            "rms_over_rms": float(np.sqrt((d * d).mean() / (j64 * j64).mean())),
            "nodes_above_1e-5_of_max": int((np.abs(d) > 1e-5 * np.abs(j64).max()).sum()),
            "nodes": int(d.size)}
    return out


def slab_leg(steps32, steps64, warmup, tile_arg=0, workload=None):
    """An ES workload timed like the headline: by default the densest z-slab
    of the 1G ramp (512x512x32 cells, 474,679,216 particles: the share of one
    of 8 GPUs that carries the most particles); with workload=units.ramp512
    the whole ramp (512x512x256, 1,073,741,824 particles) on one GPU.  ES
    push + fixed-point charge in the FP32 form, the FP64 form beside it (the
    two contexts one after the other)."""
    from ctca_amd import emg, units
    w = (workload or units.ramp512_slab)()
    su = units.to_sunits(w)
    n = w.n
    nsp = len(w.species)
    counts = su.npin
    npart = int(counts.sum())
    ncell = n[0] * n[1] * n[2]
    spec = units.device_species(w, su)
    tile = auto_tile(n, tile_arg)
    ctx = emg.Context(n, su.q, su.qm, su.rm, counts, precision=emg.FP32_FIXED, e0=su.e0,
                      b0=su.b0, tile=tile)
    try:
        ctx.init_device(spec, w.seed)
        eb = build_fields(ctx, su, n)
        ctx.set_fields(eb)
        ctx.push_es(2)
        el32, st32 = time_steps(ctx, "fp32", steps32, warmup)
    finally:
        ctx.close()
    k32 = kernel_ms(st32, emg.KERNEL_PUSH)
    bytes32 = 56.0 * npart + 24.0 * ncell + 8.0 * ncell * nsp
    out = {"workload": w.name, "grid": list(n), "particles": npart, "tile": tile,
           "fp32": {"particle_steps_per_s": npart * steps32 / el32,
                    "ms_per_step": el32 / steps32 * 1e3, "steps": steps32, "kernel_ms": k32,
                    "gbs_algorithmic": bytes32 / (k32 * 1e-3) / 1e9,
                    "hbm_frac": bytes32 / (k32 * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    if steps64:
        ctx = emg.Context(n, su.q, su.qm, su.rm, counts, precision=emg.PLAIN_FP64, e0=su.e0,
                          b0=su.b0, sort_every=FP64_SORT_EVERY)
        try:
            ctx.init_device(spec, w.seed)
            sort_ms = timed_sort(ctx)
            ctx.set_fields(eb)
            el64, st64 = time_steps(ctx, "fp64", steps64, warmup)
        finally:
            ctx.close()
        v64 = npart * steps64 / el64
        out["fp64"] = {"particle_steps_per_s": v64, "ms_per_step": el64 / steps64 * 1e3,
                       "steps": steps64,
                       "push_deposit_kernel_ms": kernel_ms(st64, emg.KERNEL_PUSH),
                       "deposit_launches": st64[emg.KERNEL_DEPOSIT][1],
                       "sort_ms": sort_ms, "sort_every": FP64_SORT_EVERY}
        out["fp32_over_fp64"] = out["fp32"]["particle_steps_per_s"] / v64
    return out


def energy_drift(steps):
    """The plain ES configuration (32^3, 2 x 524,288 particles, dt 0.1) run
    self-consistently on the device for `steps` steps in both forms (deposit,
    emg_solve_es, push): total energy W = kinetic + field (0.5 sum E^2 / V)
    per step, and its largest relative change from the first step."""
    from ctca_amd import emg, fieldsolve, units
    w = units.es32()
    su = units.to_sunits(w)
    n = w.n
    counts = su.npin
    spec = units.device_species(w, su)
    bk = fieldsolve.background(su.q, counts, n)
    out = {"workload": w.name, "steps": steps, "particles": int(counts.sum())}
    for name, prec in (("fp64", emg.PLAIN_FP64), ("fp32", emg.FP32_FIXED)):
        ctx = emg.Context(n, su.q, su.qm, su.rm, counts, precision=prec, e0=su.e0, b0=su.b0)
        try:
            ctx.init_device(spec, w.seed)
            if prec == emg.PLAIN_FP64:
                ctx.sort()
            ctx.deposit_rho(0)
            ws = []
            for _ in range(steps):
                if prec == emg.PLAIN_FP64:
                    ctx.deposit_rho(0)
                ctx.solve_es(bk)
                fe = ctx.field_energy(su.cs)
                ctx.push_es(2)
                ws.append(float(ctx.energy().sum() + fe[0]))
        finally:
            ctx.close()
        ws = np.array(ws)
        out[name] = {"energy_first": ws[0], "energy_last": ws[-1],
                     "max_rel_drift": float(np.abs(ws - ws[0]).max() / abs(ws[0]))}
    return out


def em_energy_drift(steps):
    """The EM two-stream workload at full size run with the fields on the
    device (bfield, grdcrt, psolve1 + current, bfield, efield; emses.F90:
    365-434) for `steps` steps in both forms, from the same seeded state and
    the E of its initial density: total energy W = kinetic + 0.5 sum E^2 / V
    + 0.5 cs sum B^2 / V (energy.F90:30-70) after each step, its largest
    relative change from the first step, and the largest difference of the
    two forms' histories."""
    from ctca_amd import emg, units
    w = units.em_twostream()
    su = units.to_sunits(w)
    n = w.n
    spec = units.device_species(w, su)
    tcs = 2.0 * su.cs
    out = {"workload": w.name, "steps": steps, "particles": int(su.npin.sum())}
    hist = {}
    eb = None
    for name, prec in (("fp32", emg.FP32_FIXED), ("fp64", emg.PLAIN_FP64)):
        kw = dict(tile=auto_tile(n)) if prec == emg.FP32_FIXED else dict(sort_every=FP64_SORT_EVERY)
        ctx = emg.Context(n, su.q, su.qm, su.rm, su.npin, precision=prec, e0=su.e0, b0=su.b0, **kw)
        try:
            ctx.init_device(spec, w.seed)
            if eb is None:
                eb = build_fields(ctx, su, n)
            ctx.set_fields(eb)
            ws = []
            for _ in range(steps):
                ctx.bfield()
                ctx.grdcrt()
                ctx.push_em(2)
                ctx.bfield()
                ctx.efield(tcs, 2.0)
                fe = ctx.field_energy(su.cs)
                ws.append(float(ctx.energy().sum() + fe[0] + fe[1]))
        finally:
            ctx.close()
        ws = np.array(ws)
        hist[name] = ws
        out[name] = {"energy_first": ws[0], "energy_last": ws[-1],
                     "max_rel_drift": float(np.abs(ws - ws[0]).max() / abs(ws[0]))}
    out["max_rel_diff_fp32_fp64"] = float(np.abs(hist["fp32"] - hist["fp64"]).max()
                                          / abs(hist["fp64"][0]))
    return out


def cpu_sample(w, su, nsample):
    """A bounded sample of the workload for the oracle: nsample particles split
    over the species, each species ordered by the reference's own
    particle_sort_ (z slabs) as spsort leaves it at initialisation."""
    import ctypes as C

    import oracle
    n = w.n
    g = oracle.rng_start(w.seed)
    parts = []
    for s in w.species:
        sp = oracle.SpeciesInit(npin=nsample // len(w.species), peth=s.vth, path=s.vth)
        p = oracle.inipcl_species(g, sp, len(parts) + 1, n, su.t)
        p["vx"] *= su.renv
        p["vy"] *= su.renv
        p["vz"] *= su.renv
        ref = oracle.ref_lib()
        if ref is not None:
            ref.particle_sort_.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
            ref.particle_sort_(p.ctypes.data_as(C.c_void_p), C.byref(C.c_int(len(p))))
        parts.append(p)
    return parts


def cpu_baseline(w, su, eb, nsample, steps):
    """The oracle on host cores: psolve2 + charge + periodic fold, one thread
    (one reference MPI rank), on the z-sorted sample."""
    import oracle
    n = w.n
    parts = cpu_sample(w, su, nsample)
    counts = [len(p) for p in parts]
    p = np.concatenate(parts)
    wrk = oracle.grdcrt(eb, n, su.e0, su.b0)
    sl = np.array(n, dtype=float)
    t0 = time.perf_counter()
    for _ in range(steps):
        oracle.psolve2(p, counts, su.qm, wrk, n, 2, sl)
        r = oracle.charge(p, counts, su.q, n)
        oracle.fold_rho(r, n)
    t1 = time.perf_counter()
    total = sum(counts)
    return {"value": total * steps / (t1 - t0), "unit": "particle-steps/s", "cores": 1,
            "kind": "port",
            "sample": f"{n[0]}x{n[1]}x{n[2]} grid, {total:,} particles ({len(counts)} species, "
                      f"z-sorted by the reference's particle_sort_), {steps} steps of psolve2 + "
                      f"charge + periodic fold, FP64, 1 thread, {t1 - t0:.1f} s"}


def cpu_baseline_threads(w, su, eb, nsample, steps, threads):
    """The same work spread over host threads: each thread pushes and deposits
    a contiguous share of every species (a z range, after the sort) into its
    own charge array; the arrays are summed pairwise in the pool and folded
    once per step.  The oracle's C calls release the GIL."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle
    n = w.n
    parts = cpu_sample(w, su, nsample)
    shares = []
    for t in range(threads):
        sub = [p[len(p) * t // threads: len(p) * (t + 1) // threads] for p in parts]
        shares.append((np.concatenate(sub), [len(x) for x in sub]))
    wrk = oracle.grdcrt(eb, n, su.e0, su.b0)
    sl = np.array(n, dtype=float)

    def work(t):
        p, cnt = shares[t]
        oracle.psolve2(p, cnt, su.qm, wrk, n, 2, sl)
        return oracle.charge(p, cnt, su.q, n)

    def add(ab):
        np.add(ab[0], ab[1], out=ab[0])
        return ab[0]

    with ThreadPoolExecutor(threads) as ex:
        t0 = time.perf_counter()
        for _ in range(steps):
            rs = list(ex.map(work, range(threads)))
            while len(rs) > 1:
                rs = list(ex.map(add, zip(rs[0::2], rs[1::2]))) + (rs[-1:] if len(rs) % 2 else [])
            oracle.fold_rho(rs[0], n)
        t1 = time.perf_counter()
    total = sum(len(p) for p in parts)
    return {"value": total * steps / (t1 - t0), "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{n[0]}x{n[1]}x{n[2]} grid, {total:,} particles, {steps} steps of psolve2 "
                      f"+ charge + periodic fold, FP64, {threads} threads (per-thread charge "
                      f"arrays summed each step), {t1 - t0:.1f} s"}


def adjoint_window(steps, checkpoint):
    """The adjoint-window configuration (64^3 ES, electrons 32 ppc + a 5% beam
    at v_d = 3.0, E0 = 0.01, K steps, checkpoints every `checkpoint`): time
    of the forward window and of the reverse sweep (which recomputes each
    window's states), and the gradients against central differences."""
    from ctca_amd import adjoint, emg, units
    w = units.adjoint64()
    win = adjoint.ESWindow(w, steps=steps, checkpoint=checkpoint)
    try:
        vd, e0 = w.species[1].drift[0], w.e0[0]
        win.forward(vd, e0)                      # warm-up (allocations, first launches)
        win.ctx.synchronize()
        t0 = time.perf_counter()
        loss = win.forward(vd, e0)
        win.ctx.synchronize()
        t1 = time.perf_counter()
        g_vd, g_e0 = win.backward()
        win.ctx.synchronize()
        t2 = time.perf_counter()
        # Finite differences, declared before the run and applied to both
        # parameters alike: central differences at the configuration's
        # h = 1e-3 (the headline FD figure, pass/fail against 1e-4), and the
        # same quotient at each step of a fixed ladder relative to the
        # parameter, every step reported, none selected.  The loss is only
        # piecewise smooth (a particle that crosses a cell face kinks it) and
        # h = 1e-3 moves E0 by 10%, so large steps carry kink and curvature
        # terms and tiny ones meet the round-off of the FP64 atomics.
        def cd(hv, he):
            return ((win.forward(vd + hv, e0) - win.forward(vd - hv, e0)) / (2 * hv),
                    (win.forward(vd, e0 + he) - win.forward(vd, e0 - he)) / (2 * he))
        h = 1e-3
        fd_h = cd(h, h)
        ladder = {f"rel {r:g}": cd(r * abs(vd), r * abs(e0)) for r in (1e-4, 1e-5, 1e-6, 1e-7, 1e-8)}
    finally:
        win.close()

    def errs(v):
        return {"dL_dvd": v[0], "dL_dE0": v[1], "rel_err_vd": abs(g_vd - v[0]) / abs(v[0]),
                "rel_err_E0": abs(g_e0 - v[1]) / abs(v[1])}
    head = errs(fd_h)
    return {"workload": w.name, "particles": int(win.counts.sum()), "steps": steps,
            "checkpoint_every": checkpoint, "loss": loss,
            "forward_s": t1 - t0, "adjoint_s": t2 - t1, "adjoint_over_forward": (t2 - t1) / (t1 - t0),
            "dL_dvd": g_vd, "dL_dE0": g_e0,
            "fd_protocol": "central differences at h = 1e-3 on both parameters (the "
                           "configuration's); bar 1e-4 relative",
            "fd_h_1e-3": head,
            "rel_err_vd": head["rel_err_vd"], "rel_err_E0": head["rel_err_E0"],
            "fd_bar": 1e-4,
            "fd_pass": bool(head["rel_err_vd"] <= 1e-4 and head["rel_err_E0"] <= 1e-4),
            "fd_ladder_relative_steps": {k: errs(v) for k, v in ladder.items()},
            "fp64_reference_gradient": "tests/test_gpu_adjoint.py::"
                                       "test_adjoint_window_64cubed_k50_full_size: the reverse "
                                       "sweep equals torch.autograd through an independent FP64 "
                                       "restatement to 1e-10",
            "precision": "FP64 particle path, FP64 torch grid operators",
            "note": "adjoint time includes recomputing every window's states"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mag128_2sp_64ppc")
    ap.add_argument("--tile", type=int, default=0, help="4 or 8 (0: 8 where the grid allows)")
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--cpu-sample", type=int, default=16_777_216)
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=50)
    ap.add_argument("--no-em", action="store_true")
    ap.add_argument("--em-steps", type=int, default=40)
    ap.add_argument("--no-slab", action="store_true")
    ap.add_argument("--no-drift", action="store_true")
    ap.add_argument("--drift-steps", type=int, default=100)
    ap.add_argument("--em-drift-steps", type=int, default=100,
                    help="EM two-stream energy history (0: skip)")
    ap.add_argument("--slab-steps", type=int, default=40)
    ap.add_argument("--no-full-ramp", action="store_true",
                    help="skip the whole 512x512x256 ramp (1,073,741,824 particles)")
    ap.add_argument("--full-ramp-steps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the particle path runs on one MI355X that owns the "
              "whole domain; sharding over several GPUs is designed (z slabs with particle "
              "migration between neighbours) but not a measured mode.", file=sys.stderr)
        sys.exit(2)

    # torch must take the device before libemgpu does: in the other order
    # torch.cuda.is_available() is False in this process (torch runs the
    # adjoint leg's grid operators and the copy-bandwidth probe)
    import torch
    torch.cuda.is_available()
    from ctca_amd import emg, units
    w = units.WORKLOADS[args.workload]()
    su = units.to_sunits(w)
    n = w.n
    nsp = len(w.species)
    counts = su.npin
    npart = int(counts.sum())
    ncell = n[0] * n[1] * n[2]
    spec = units.device_species(w, su)
    tile = auto_tile(n, args.tile)
    log(f"[bench] {w.name}: {npart:,} particles, grid {n}")

    # ---- FP32 cell + offset, int64 fixed point (headline)
    ctx = emg.Context(n, su.q, su.qm, su.rm, counts, precision=emg.FP32_FIXED, e0=su.e0,
                      b0=su.b0, tile=tile)
    ctx.init_device(spec, w.seed)
    eb = build_fields(ctx, su, n)
    ctx.set_fields(eb)
    ctx.push_es(2)
    rho32 = [ctx.rho(species=k) for k in range(nsp)]
    el32, st32 = time_steps(ctx, "fp32", args.steps, args.warmup)
    ek32 = ctx.energy()
    # the whole ES step on the device: field solve from the deposited charge
    # (spectral Poisson, E = -grad phi, grdcrt) and the fused push + charge
    from ctca_amd import fieldsolve
    bk = fieldsolve.background(su.q, counts, n)
    nfull = min(args.steps, 100)
    for _ in range(args.warmup):
        ctx.solve_es(bk)
        ctx.push_es(2)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(nfull):
        ctx.solve_es(bk)
        ctx.push_es(2)
    ctx.synchronize()
    el_full = time.perf_counter() - t0
    ctx.close()
    v32 = npart * args.steps / el32
    push_ms, push_n = st32[emg.KERNEL_PUSH]
    kern_ms = push_ms / max(push_n, 1)
    bytes_launch = 56.0 * npart + 24.0 * ncell + 8.0 * ncell * nsp
    achieved = bytes_launch / (kern_ms * 1e-3) / 1e9
    copy_gbs = copy_bandwidth()
    log(f"[bench] fp32: {el32 / args.steps * 1e3:.3f} ms/step, kernel {kern_ms:.3f} ms, "
        f"{achieved:.0f} GB/s algorithmic")

    # ---- plain FP64 (reference arithmetic) beside it
    fp64 = None
    if not args.no_fp64:
        steps64 = args.fp64_steps or max(args.steps // 4, 10)
        ctx = emg.Context(n, su.q, su.qm, su.rm, counts, precision=emg.PLAIN_FP64, e0=su.e0,
                          b0=su.b0, sort_every=FP64_SORT_EVERY)
        ctx.init_device(spec, w.seed)
        sort_ms = timed_sort(ctx)   # spsort at initialisation (esses.F90:277-284)
        ctx.set_fields(eb)
        ctx.push_es(2)
        ctx.deposit_rho(2)
        rho64 = [ctx.rho(species=k) for k in range(nsp)]
        el64, st64 = time_steps(ctx, "fp64", steps64, args.warmup)
        ctx.close()
        v64 = npart * steps64 / el64
        pm, pn = st64[emg.KERNEL_PUSH]
        dm, dn = st64[emg.KERNEL_DEPOSIT]
        derr = [float(np.abs(rho32[k] - rho64[k]).max() / np.abs(rho64[k]).max()) for k in range(nsp)]
        # one fused kernel (k_push_dep64: psolve2 + charge at the new
        # positions); emg_deposit_rho after it launches nothing
        pk = pm / max(pn, 1)
        bytes64 = 96.0 * npart + 48.0 * (n[0] + 1) * (n[1] + 1) * (n[2] + 1) + 8.0 * ncell * nsp
        fp64 = {"particle_steps_per_s": v64, "ms_per_step": el64 / steps64 * 1e3, "steps": steps64,
                "push_deposit_kernel_ms": pk, "deposit_launches": dn,
                "gbs_algorithmic": bytes64 / (pk * 1e-3) / 1e9,
                "hbm_frac": bytes64 / (pk * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "fp32_over_fp64": v32 / v64,
                "density_err_fp32_vs_fp64_one_step": derr,
                "sort_ms": sort_ms, "sort_every": FP64_SORT_EVERY,
                "note": "ms_per_step includes the sort before every 10th push"}
        log(f"[bench] fp64: {el64 / steps64 * 1e3:.3f} ms/step; fp32/fp64 = {v32 / v64:.2f}x; "
            f"density err {derr}")

    adj = None
    if not args.no_adjoint:
        adj = adjoint_window(args.adjoint_steps, 10)
        log(f"[bench] adjoint window: forward {adj['forward_s']:.2f} s, adjoint {adj['adjoint_s']:.2f} s "
            f"(x{adj['adjoint_over_forward']:.2f}); central FD at h = 1e-3: rel err "
            f"{adj['rel_err_vd']:.1e} (v_d), {adj['rel_err_E0']:.1e} (E0), within 1e-4: "
            f"{adj['fd_pass']}")

    drift = None
    if not args.no_drift:
        drift = energy_drift(args.drift_steps)
        log(f"[bench] energy drift over {args.drift_steps} self-consistent steps: fp64 "
            f"{drift['fp64']['max_rel_drift']:.2e}, fp32 {drift['fp32']['max_rel_drift']:.2e}")

    em = None
    if not args.no_em:
        em = em_leg(args.em_steps, max(args.em_steps // 4, 5) if not args.no_fp64 else 0,
                    args.warmup, args.tile)
        log(f"[bench] em: fp32 {em['fp32']['ms_per_step']:.3f} ms/step, kernel "
            f"{em['fp32']['kernel_ms']:.3f} ms ({em['fp32']['gbs_algorithmic']:.0f} GB/s)")

    em_drift = None
    if not args.no_em and not args.no_drift and args.em_drift_steps > 0:
        em_drift = em_energy_drift(args.em_drift_steps)
        log(f"[bench] em energy drift over {args.em_drift_steps} steps: fp32 "
            f"{em_drift['fp32']['max_rel_drift']:.2e}, fp64 {em_drift['fp64']['max_rel_drift']:.2e}, "
            f"forms differ by {em_drift['max_rel_diff_fp32_fp64']:.1e}")

    slab = None
    if not args.no_slab:
        slab = slab_leg(args.slab_steps, max(args.slab_steps // 8, 5) if not args.no_fp64 else 0,
                        args.warmup, args.tile)
        log(f"[bench] ramp slab: fp32 {slab['fp32']['ms_per_step']:.3f} ms/step, kernel "
            f"{slab['fp32']['kernel_ms']:.3f} ms ({slab['fp32']['gbs_algorithmic']:.0f} GB/s)")

    full_ramp = None
    if not args.no_full_ramp:
        full_ramp = slab_leg(args.full_ramp_steps,
                             max(args.full_ramp_steps // 4, 3) if not args.no_fp64 else 0,
                             args.warmup, args.tile, workload=units.ramp512)
        log(f"[bench] full ramp: fp32 {full_ramp['fp32']['ms_per_step']:.3f} ms/step, kernel "
            f"{full_ramp['fp32']['kernel_ms']:.3f} ms ({full_ramp['fp32']['gbs_algorithmic']:.0f} GB/s)")

    cpu = cpu_all = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, su, eb, args.cpu_sample, args.cpu_steps)
        log(f"[bench] cpu baseline {cpu['value']:.3e} particle-steps/s")
        threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        # four times the particles per step, so each thread's share outweighs the
        # per-step work on the whole grid (its charge array, the sum, the fold)
        cpu_all = cpu_baseline_threads(w, su, eb, 4 * args.cpu_sample, args.cpu_steps, threads)
        log(f"[bench] cpu baseline, {threads} threads: {cpu_all['value']:.3e} particle-steps/s")

    out = {
        "metric": METRIC,
        "value": v32,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el32 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded Maxwellian particles generated on the device, E from the "
                "initial density (host Poisson), uniform B0",
        "config": {"workload": w.name, "grid": list(n), "particles": npart, "species": nsp,
                   "ppc_per_species": npart // nsp // ncell, "dt": w.dt, "wc": w.wc,
                   "form": "FP32 cell index + in-cell offset push, int64 fixed-point CIC charge "
                           f"(2^24 per unit weight), tile {tile}^3 cells",
                   "kinetic_energy_fp32": ek32.tolist(),
                   "device_resident_step": {
                       "ms_per_step": el_full / nfull * 1e3, "steps": nfull,
                       "particle_steps_per_s": npart * nfull / el_full,
                       "what": "emg_solve_es (charge + background -> hipFFT Poisson -> E -> "
                               "grdcrt) + emg_push_es (fused push + charge), fields on the "
                               "device"},
                   "fp64": fp64,
                   "adjoint": adj,
                   "cpu_all_cores": cpu_all,
                   "energy_drift": drift,
                   "em_energy_drift": em_drift,
                   "em_twostream": em,
                   "ramp_densest_slab": slab,
                   "ramp_full_1G": full_ramp},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "copy_gbs": copy_gbs, "frac_of_copy": achieved / copy_gbs,
                     "traffic": measured_traffic(w.name, tile),
                     "kernel": f"k_fx_step<{tile},true,1> (push + deposit + re-bucket)",
                     "kernel_ms": kern_ms, "bytes_per_launch": bytes_launch},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
