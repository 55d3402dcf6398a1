"""The host-side seeded initialisation (emg_host_inipcl, host_init.cpp) built
with AddressSanitizer and UndefinedBehaviorSanitizer by g++ (host code only:
no GPU code is built here), run on a small case and on its error paths; its
records must equal the ones the library itself produces."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from ctca_amd import emg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GXX = shutil.which("g++")

DRIVER = r"""
#include <cstdio>
#include <vector>
#include "emgpu.h"
int main(int argc, char **argv) {
  emg_host_species sp[2] = {};
  sp[0].npin = 1000; sp[0].peth = 1.0; sp[0].path = 1.0; sp[0].vdri = 0.3; sp[0].vdthz = 90.0;
  sp[1].npin = 333;  sp[1].peth = 0.1; sp[1].path = 0.05; sp[1].spe = 0.2; sp[1].speth = 30.0;
  const int32_t n[3] = {16, 12, 8};
  const double t[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::vector<emg_particle> out(1333);
  if (emg_host_inipcl(sp, 2, n, t, 10u, out.data()) != EMG_OK) return 2;
  // error paths: no species, a negative count, a null output
  if (emg_host_inipcl(sp, 0, n, t, 10u, out.data()) != EMG_ERR_ARG) return 3;
  emg_host_species bad = sp[0];
  bad.npin = -1;
  if (emg_host_inipcl(&bad, 1, n, t, 10u, out.data()) != EMG_ERR_ARG) return 4;
  if (emg_host_inipcl(sp, 2, n, t, 10u, nullptr) != EMG_ERR_ARG) return 5;
  FILE *f = std::fopen(argv[1], "wb");
  if (!f) return 6;
  std::fwrite(out.data(), sizeof(emg_particle), out.size(), f);
  std::fclose(f);
  return 0;
}
"""


@pytest.mark.skipif(GXX is None, reason="g++ not found")
def test_host_init_under_asan_and_ubsan(tmp_path):
    drv = tmp_path / "drv.cpp"
    drv.write_text(DRIVER)
    exe = tmp_path / "drv"
    subprocess.run([GXX, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "ctca_amd", "csrc", "host_init.cpp"), str(drv),
                    "-o", str(exe)], check=True, capture_output=True, text=True)
    out = tmp_path / "p.bin"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    got = np.fromfile(out, dtype=emg.PARTICLE)
    spec = [dict(npin=1000, peth=1.0, path=1.0, vdri=0.3, vdthz=90.0),
            dict(npin=333, peth=0.1, path=0.05, spe=0.2, speth=30.0)]
    ref = emg.host_inipcl(spec, (16, 12, 8), np.eye(3), 10)
    assert got.tobytes() == ref.tobytes()
