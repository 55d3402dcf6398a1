"""Full-size runs of the benchmarked workloads, checked through properties
that hold at any size (the oracle is too slow for these particle counts):

  * the FP32 cell + offset form against the plain FP64 form, one step from
    the same seeded device state and the same fields: density within 1e-6 of
    the species maximum, kinetic energy within 1e-5, the zigzag current's RMS
    error within 1e-5;
  * charge conservation: every particle deposits exactly unit weight, so the
    folded density of species s sums to q_s N_s (closed form);
  * the total current sum_nodes J equals in both forms (each particle adds
    q times its displacement, whatever cells it crosses);
  * the FP64 cell sort is a permutation: the density before and after it
    agrees to round-off;
  * no particle is lost or gained.

Workloads: magnetised 128^3 with 2 species at 64 per cell (268,435,456
particles), the EM two-stream 256x256x128 (295,279,002), the densest
z-slab 512x512x32 of the exponential ramp (474,679,216), and the whole ramp
512x512x256 with 8 per cell per species on average (1,073,741,824
particles on one GPU, the two forms in turn)."""
import numpy as np
import pytest

from ctca_amd import emg, units
from ctca_amd import fieldsolve

pytestmark = pytest.mark.gpu


def fields_of(ctx, su, n):
    ctx.deposit_rho(0)
    rho = ctx.rho() + fieldsolve.background(su.q, su.npin, n)
    return fieldsolve.eb_halo(fieldsolve.efield(fieldsolve.poisson(rho)))


def setup(workload):
    w = workload()
    su = units.to_sunits(w)
    return w, su, units.device_species(w, su)


def context(w, su, precision):
    kw = dict(tile=8) if precision == emg.FP32_FIXED else {}
    return emg.Context(w.n, su.q, su.qm, su.rm, su.npin, precision=precision, e0=su.e0,
                       b0=su.b0, **kw)


def check_charge(rho, q, count):
    """sum over the periodic nodes of the folded density = q N"""
    total = float(np.sum(rho, dtype=np.float64))
    assert abs(total - q * count) <= 1e-12 * abs(q * count), (total, q * count)


@pytest.mark.parametrize("workload", [units.mag128, units.ramp512_slab, units.ramp512],
                         ids=["mag128_268M", "ramp512_slab_475M", "ramp512_full_1G"])
def test_es_step_fp32_against_fp64_full_size(workload):
    w, su, spec = setup(workload)
    nsp = len(w.species)
    with context(w, su, emg.FP32_FIXED) as ctx:
        ctx.init_device(spec, w.seed)
        eb = fields_of(ctx, su, w.n)
        ctx.set_fields(eb)
        ctx.push_es(2)
        rho32 = [ctx.rho(species=k) for k in range(nsp)]
        ke32 = ctx.energy()
        assert ctx.counts().tolist() == su.npin.tolist()
    with context(w, su, emg.PLAIN_FP64) as ctx:
        ctx.init_device(spec, w.seed)
        ctx.deposit_rho(0)
        before = [ctx.rho(species=k) for k in range(nsp)]
        ctx.sort()
        ctx.deposit_rho(0)
        for k in range(nsp):
            after = ctx.rho(species=k)
            assert np.abs(after - before[k]).max() <= 1e-12 * np.abs(before[k]).max(), k
        ctx.set_fields(eb)
        ctx.push_es(2)
        ctx.deposit_rho(0)
        rho64 = [ctx.rho(species=k) for k in range(nsp)]
        ke64 = ctx.energy()
        assert ctx.counts().tolist() == su.npin.tolist()
    for k in range(nsp):
        err = np.abs(rho32[k] - rho64[k]).max() / np.abs(rho64[k]).max()
        assert err <= 1e-6, (k, err)
        check_charge(rho32[k], su.q[k], su.npin[k])
        check_charge(rho64[k], su.q[k], su.npin[k])
    assert np.abs(ke32 - ke64).max() <= 1e-5 * np.abs(ke64).max()


def test_em_step_fp32_against_fp64_full_size():
    w, su, spec = setup(units.em_twostream)
    with context(w, su, emg.FP32_FIXED) as ctx:
        ctx.init_device(spec, w.seed)
        eb = fields_of(ctx, su, w.n)
        ctx.set_fields(eb)
        ctx.push_em(2)
        j32 = ctx.current()
        assert ctx.counts().tolist() == su.npin.tolist()
    with context(w, su, emg.PLAIN_FP64) as ctx:
        ctx.init_device(spec, w.seed)
        ctx.sort()
        ctx.set_fields(eb)
        ctx.push_em(2)
        j64 = ctx.current()
        assert ctx.counts().tolist() == su.npin.tolist()
    d = j32 - j64
    # the relay point of the zigzag jumps where an end point crosses a cell
    # face, so the few particles whose FP32 and FP64 end points straddle a
    # face set the maximum; the RMS measures the arithmetic
    assert np.sqrt((d * d).mean() / (j64 * j64).mean()) <= 1e-5
    assert np.abs(d).max() <= 2e-3 * np.abs(j64).max()
    t32 = j32.reshape(-1, 3).sum(axis=0, dtype=np.float64)
    t64 = j64.reshape(-1, 3).sum(axis=0, dtype=np.float64)
    scale = np.abs(j64).reshape(-1, 3).sum(axis=0, dtype=np.float64)
    assert np.all(np.abs(t32 - t64) <= 1e-6 * scale), (t32, t64)
