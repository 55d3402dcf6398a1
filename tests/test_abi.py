"""The C ABI library: loads, exports every entry point include/emgpu.h
declares, and rejects bad arguments with an error code (no device calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from ctca_amd import emg, units

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "emgpu.h")).read()
    return sorted(set(re.findall(r"^int\s+(emg_\w+)\s*\(", src, re.M)))


def test_header_and_binding_agree():
    assert declared() == sorted(emg.EXPORTS)


def test_library_exports_every_symbol():
    L = emg.lib()
    for name in declared():
        assert hasattr(L, name), name


def test_bad_arguments_return_codes():
    L = emg.lib()
    h = C.c_void_p()
    cfg = emg.Config()
    cfg.nspec = 0
    assert L.emg_create(C.byref(cfg), C.byref(h)) == -1
    assert "nspec" in emg.last_error()
    cfg.nspec = 1
    cfg.n = (C.c_int32 * 3)(1, 4, 4)
    assert L.emg_create(C.byref(cfg), C.byref(h)) == -1
    assert L.emg_create(None, C.byref(h)) == -1
    assert L.emg_push_es(None, 2) == -1
    assert L.emg_destroy(None) == 0


def test_host_inipcl_rejects_bad_input():
    L = emg.lib()
    sp = (emg.HostSpecies * 1)()
    sp[0].npin = -1
    n = (C.c_int32 * 3)(4, 4, 4)
    t = (C.c_double * 9)(1, 0, 0, 0, 1, 0, 0, 0, 1)
    assert L.emg_host_inipcl(sp, 1, n, t, 10, None) == -1
    sp[0].npin = 5
    assert L.emg_host_inipcl(sp, 1, n, t, 10, None) == -1  # null output


def test_particle_record_layout():
    # oh_part.h:23-28: six doubles then pid, preside, nid, spec (int32)
    assert emg.PARTICLE.itemsize == 64
    assert [emg.PARTICLE.fields[f][1] for f in ("pid", "preside", "nid", "spec")] == [48, 52, 56, 60]


def test_units_es32_matches_reference_formulas():
    w = units.es32()
    su = units.to_sunits(w)
    # EMSES-U q = wp^2 V / npin / qm: electrons -0.0625, ions +0.0625 (16 ppc)
    # rescal renq = qm(1) (dt/2)^2 = -0.0025
    assert np.allclose(su.q, [0.0625 * 0.0025, -0.0625 * 0.0025])
    assert np.allclose(su.qm, [1.0, -0.01])
    assert su.renv == pytest.approx(0.05)
    assert np.allclose(su.q * su.npin, [81.92, -81.92])   # neutral


def test_units_mag128_b0():
    su = units.to_sunits(units.mag128())
    # b0 = wc / qm(1) = -0.5, renb = qm(1) dt/2 = -0.025 -> +0.0125 along z
    assert np.allclose(su.b0, [0, 0, 0.0125])
    # kick rotation parameter of the electrons qm*ustep/2*b0 = wc dt / 2
    assert su.qm[0] * su.b0[2] == pytest.approx(0.5 * 0.05 / 2)


def test_units_ramp512_b_along_x():
    su = units.to_sunits(units.ramp512())
    assert np.allclose(su.b0, [0.2 * 0.05 / 2, 0, 0])
