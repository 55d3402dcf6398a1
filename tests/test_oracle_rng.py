"""The oracle's MT19937 / RANU0 / RANN0 / inipcl restatement against the
reference's own mt19937arm.c (golden vectors from oracle/_ref) and against
the published MT19937 known answer; the product's host initialiser
(emg_host_inipcl) against the oracle, bit for bit."""
import math
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_mt19937_known_answer():
    # published mt19937ar: init_genrand(5489) first outputs
    g = oracle.MT(5489)
    assert [g.u32() for _ in range(3)] == [3499211612, 581869302, 3890346734]


@pytest.mark.parametrize("seed", [5489, 10, 1234])
def test_mt19937_matches_reference_stream(seed):
    gold = np.load(os.path.join(GOLD, "mt19937_ref.npz"))[f"u32_seed{seed}"]
    g = oracle.MT(seed)
    got = np.array([g.u32() for _ in range(len(gold))], dtype=np.uint64)
    assert np.array_equal(got, gold)


def test_res53_matches_mtrand_real8():
    # RANU0 = mtrand_real8_ = genrand_res53 (wrapper.F90:116-118, mt19937arm.c:208-210)
    gold = np.load(os.path.join(GOLD, "mt19937_ref.npz"))["real8_seed10"]
    g = oracle.MT(10)
    got = g.ranu0(len(gold))
    assert np.array_equal(got, gold)


def test_rann0_pairs_and_odd_tail():
    gold = np.load(os.path.join(GOLD, "mt19937_ref.npz"))["real8_seed10"]
    am, sd = 0.3, 1.7
    # even n: Box-Muller over the pairs (u1, u2) of the uniform stream
    g = oracle.MT(10)
    out = g.rann0(am, sd, 6)
    u = gold[:6]
    for i in range(0, 6, 2):
        r = sd * math.sqrt(-2.0 * math.log(u[i]))
        assert out[i] == am + r * math.cos(6.28318530717958646 * u[i + 1])
        assert out[i + 1] == am + r * math.sin(6.28318530717958646 * u[i + 1])
    # odd n: 4 from the array draw, the tail from two fresh uniforms
    g = oracle.MT(10)
    out = g.rann0(am, sd, 5)
    assert out[4] == am + sd * math.sqrt(-2.0 * math.log(gold[4])) * math.cos(6.28318530717958646 * gold[5])
    # n == 1 consumes two uniforms and uses the cosine branch
    g = oracle.MT(10)
    one = g.rann0(am, sd, 1)[0]
    assert one == am + sd * math.sqrt(-2.0 * math.log(gold[0])) * math.cos(6.28318530717958646 * gold[1])
    assert g.res53() == gold[2]


def test_rng_start_consumes_three_res53():
    # inital.F90:715-736: one RANU0(1) and one RANN0(1) before inipcl
    gold = np.load(os.path.join(GOLD, "mt19937_ref.npz"))["real8_seed10"]
    g = oracle.rng_start(10)
    assert g.res53() == gold[3]


def test_inipcl_draw_order():
    gold = np.load(os.path.join(GOLD, "mt19937_ref.npz"))["real8_seed10"]
    g = oracle.rng_start(10)
    sp = oracle.SpeciesInit(npin=300, peth=1.0, path=1.0)
    n = (8, 8, 8)
    p0 = oracle.inipcl_species(g, sp, 1, n, np.eye(3))
    # positions: x = n * u over the first three RANU0 arrays (inipcl.F90:83-104)
    u = gold[3:3 + 3 * 300]
    assert np.array_equal(p0["x"], 8.0 * u[0:300])
    assert np.array_equal(p0["y"], 8.0 * u[300:600])
    assert np.array_equal(p0["z"], 8.0 * u[600:900])
    assert (p0["spec"] == 1).all() and (p0["nid"] == 0).all()


def test_inipcl_moments():
    g = oracle.rng_start(10)
    sp = oracle.SpeciesInit(npin=200000, peth=1.3, path=0.7)
    sp.vdz = 0.25
    p = oracle.inipcl_species(g, sp, 1, (16, 16, 16), np.eye(3))
    assert abs(p["vx"].mean()) < 0.02 and abs(p["vz"].mean() - 0.25) < 0.01
    assert abs(p["vx"].std() - 1.3) < 0.01 and abs(p["vz"].std() - 0.7) < 0.01
    assert p["x"].min() >= 0 and p["x"].max() < 16


def test_product_host_init_matches_oracle():
    from ctca_amd import emg
    n = (8, 8, 8)
    t = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])  # phixy = 90
    spec = [dict(npin=1001, peth=1.0, path=0.5, vdx=0.1), dict(npin=500, peth=0.1, path=0.1, vdri=0.2,
                                                                vdthz=30.0, vdthxy=45.0)]
    prod = emg.host_inipcl(spec, n, t, 10)
    g = oracle.rng_start(10)
    o = []
    for k, s in enumerate(spec):
        si = oracle.SpeciesInit(npin=s["npin"], peth=s["peth"], path=s["path"])
        for key in ("vdx", "vdri", "vdthz", "vdthxy"):
            if key in s:
                setattr(si, key, s[key])
        o.append(oracle.inipcl_species(g, si, k + 1, n, t))
    o = np.concatenate(o)
    for f in ("x", "y", "z", "vx", "vy", "vz", "spec"):
        assert np.array_equal(prod[f], o[f]), f
