"""Build recipe of libemgpu.so (hipcc for gfx950, in-tree).

Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
selection disabled.  The FP64 mirror kernels are compiled without FP
contraction so that each multiply and add rounds separately, in source order.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libemgpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", INC, "-I", CSRC, "-Wall", "-Wno-unused-function",
          "-Werror=misleading-indentation"]

SOURCES = [
    ("k_fp64.hip", ["-ffp-contract=off"]),
    ("k_common.hip", []),
    ("k_fx.hip", []),
    ("k_adj.hip", []),
    ("k_yee.hip", ["-ffp-contract=off"]),
    ("k_poisson.hip", []),
    ("emg_api.hip", []),
]
HOST_SOURCES = [("host_init.cpp", ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-I", INC])]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build failed: " + os.path.basename(cmd[-3] if len(cmd) > 3 else cmd[0]))
    return r


def build(verbose=False):
    """Compile every source (always) and link ctca_amd/libemgpu.so."""
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    procs = []
    for src, extra in SOURCES:
        o = os.path.join(OBJ, src + ".o")
        cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", o]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(o)
    for src, flags in HOST_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        cmd = ["g++"] + flags + ["-c", os.path.join(CSRC, src), "-o", o]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(o)
    failed = []
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            failed.append(" ".join(cmd) + "\n" + out + err)
        elif verbose and (out or err):
            sys.stderr.write(out + err)
    if failed:
        sys.stderr.write("\n".join(failed))
        raise RuntimeError("libemgpu build failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}"] + objs + ["-lhipfft", "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
