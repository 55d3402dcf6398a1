"""Write the golden vectors of tests/golden/ from the reference's own C code.

oracle/_ref/libref.so is lib/mtarm/mt19937arm.c and lib/psort/psort-struct.c
compiled from /root/reference (oracle/build.py).  This script records:

  * mt19937_ref.npz: genrand_int32 streams for three seeds (init_genrand),
    and the Fortran-facing mt_init_/mtrand_real8_ stream for seed 10 (the
    rank-0 seed of src/main/inital.F90:715) -- what RANU0 returns;
  * psort_ref.npz: particle_sort_ applied to a seeded set of 64-byte particle
    records, i.e. the order the reference's z comparator (psort-struct.c:15-20)
    produces with this glibc qsort.

Run:  python tests/golden/make_golden.py   (needs oracle/_ref/libref.so)
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main():
    import oracle
    from oracle import build
    build.build_ref()
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libref.so is not built (needs /root/reference)")
    ref.genrand_int32.restype = C.c_ulong
    ref.init_genrand.argtypes = [C.c_ulong]
    ref.mtrand_real8_.restype = C.c_double
    ref.mt_init_.argtypes = [C.POINTER(C.c_int)]
    streams = {}
    for seed in (5489, 10, 1234):
        ref.init_genrand(seed)
        streams[f"u32_seed{seed}"] = np.array([ref.genrand_int32() for _ in range(2000)], dtype=np.uint64)
    s = C.c_int(10)
    ref.mt_init_(C.byref(s))
    streams["real8_seed10"] = np.array([ref.mtrand_real8_() for _ in range(1000)])
    np.savez_compressed(os.path.join(HERE, "mt19937_ref.npz"), **streams)

    rng = np.random.default_rng(7)
    n = 4096
    p = np.zeros(n, dtype=oracle.PARTICLE)
    p["x"] = rng.uniform(0, 16, n)
    p["y"] = rng.uniform(0, 16, n)
    p["z"] = rng.uniform(0, 16, n)
    p["pid"] = np.arange(n)
    p["spec"] = 1
    before = p.copy()
    ref.particle_sort_.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    cnt = C.c_int(n)
    ref.particle_sort_(p.ctypes.data_as(C.c_void_p), C.byref(cnt))
    np.savez_compressed(os.path.join(HERE, "psort_ref.npz"), z_in=before["z"], pid_out=p["pid"],
                        z_out=p["z"])
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
