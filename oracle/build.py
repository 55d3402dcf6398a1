"""Build recipe of the oracle.  TEST INFRASTRUCTURE ONLY.

* oracle/liboracle.so : the FP64 C restatement (oracle.c), gcc, no FP contraction.
* oracle/_ref/libref.so : the reference's own C sources on this path, compiled
  directly from where they lie under /root/reference (nothing is copied):
    lib/mtarm/mt19937arm.c     (MT19937 + the Fortran shims mt_init_/mtrand_real8_)
    lib/psort/psort-struct.c   (particle_sort_ / particle_spec_ qsort comparators)
  Built only where /root/reference exists; a machine without those sources
  uses the prebuilt file.  The Fortran path (psolve2/charge/psolve1) needs HDF5, MPI,
  FFTW and unpinned fpm git dependencies and is unbuildable here.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")

CFLAGS = ["-O3", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))


def build_oracle():
    tmp = LIB + ".tmp"
    _run(["gcc"] + CFLAGS + [os.path.join(HERE, "oracle.c"), "-o", tmp, "-lm"])
    os.replace(tmp, LIB)
    return LIB


def build_ref():
    if not os.path.isdir(REF):
        return None
    os.makedirs(os.path.dirname(REF_LIB), exist_ok=True)
    srcs = [os.path.join(REF, "lib/mtarm/mt19937arm.c"), os.path.join(REF, "lib/psort/psort-struct.c")]
    inc = os.path.join(REF, "lib/ohhelp/ohhelp-1.1.1")
    tmp = REF_LIB + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-I", inc] + srcs + ["-o", tmp])
    os.replace(tmp, REF_LIB)
    return REF_LIB


def build():
    build_oracle()
    build_ref()


if __name__ == "__main__":
    build()
    print(LIB, REF_LIB if os.path.exists(REF_LIB) else "(no reference: _ref not built)")
