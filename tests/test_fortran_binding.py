"""The Fortran iso_c_binding module (ctca_amd/fortran/emg_gpu.F90) compiles
with amdflang, links against libemgpu.so and passes the reference's own
type(oh_particle) records through the boundary: a host-only program gets the
same seeded particles as the Python binding (no device calls)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from ctca_amd import emg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("amdflang") or "/opt/rocm/llvm/bin/amdflang"


@pytest.mark.skipif(not os.path.exists(FLANG), reason="amdflang not available")
def test_fortran_module_and_host_call(tmp_path):
    lib_dir = os.path.join(ROOT, "ctca_amd")
    assert os.path.exists(os.path.join(lib_dir, "libemgpu.so")), "build the library first"
    mod = os.path.join(ROOT, "ctca_amd", "fortran", "emg_gpu.F90")
    prog = os.path.join(ROOT, "tests", "fortran", "boundary_check.F90")
    exe = str(tmp_path / "boundary_check")
    subprocess.run([FLANG, "-O1", "-J", str(tmp_path), mod, prog, "-o", exe,
                    "-L", lib_dir, "-lemgpu", "-Wl,-rpath," + lib_dir],
                   check=True, capture_output=True, text=True, cwd=tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, cwd=tmp_path, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "nspec" in r.stdout
    got = np.fromfile(tmp_path / "particles.bin", dtype=emg.PARTICLE)
    want = emg.host_inipcl([dict(npin=300, peth=1.0, path=1.0),
                            dict(npin=201, peth=0.1, path=0.2, vdx=0.5)],
                           (8, 6, 4), np.eye(3), 10)
    assert got.shape == want.shape
    for f in emg.PARTICLE.names:
        assert np.array_equal(got[f], want[f]), f
