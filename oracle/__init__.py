"""ctypes front-end of the oracle: the CPU FP64 restatement of the reference's
particle path (oracle.c) and the reference's own C sources built in _ref/.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker / CPU baseline.
"""
import ctypes as C
import os

import numpy as np

from . import fields  # noqa: F401  (re-export)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

# oh_part.h:23-28 / oh_type.F90:15-28
PARTICLE = np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8"), ("vx", "f8"), ("vy", "f8"),
                     ("vz", "f8"), ("pid", "i4"), ("preside", "i4"), ("nid", "i4"), ("spec", "i4")])
assert PARTICLE.itemsize == 64


class _MT(C.Structure):
    _fields_ = [("s", C.c_uint32 * 624), ("idx", C.c_int)]


class SpeciesInit(C.Structure):
    """inipcl inputs of one species (EMSES-U units), oracle.h orc_species_init."""
    _fields_ = [("npin", C.c_int64), ("peth", C.c_double), ("path", C.c_double),
                ("spe", C.c_double), ("speth", C.c_double), ("spa", C.c_double),
                ("vdri", C.c_double), ("vdthz", C.c_double), ("vdthxy", C.c_double),
                ("vdx", C.c_double), ("vdy", C.c_double), ("vdz", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build
            build.build_oracle()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def ref_lib():
    """The reference's mt19937arm.c + psort-struct.c, or None if not built."""
    if not os.path.exists(REF_PATH):
        return None
    return C.CDLL(REF_PATH)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _i3(v):
    return (C.c_int * 3)(*[int(x) for x in v])


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


class MT:
    def __init__(self, seed=None):
        self.g = _MT()
        if seed is not None:
            lib().orc_mt_seed(C.byref(self.g), C.c_uint32(seed))

    def u32(self):
        lib().orc_mt_u32.restype = C.c_uint32
        return lib().orc_mt_u32(C.byref(self.g))

    def res53(self):
        lib().orc_mt_res53.restype = C.c_double
        return lib().orc_mt_res53(C.byref(self.g))

    def ranu0(self, n):
        out = np.empty(n)
        lib().orc_ranu0(C.byref(self.g), _p(out), C.c_int64(n))
        return out

    def rann0(self, am, sd, n):
        out = np.empty(n)
        lib().orc_rann0(C.byref(self.g), C.c_double(am), C.c_double(sd), _p(out), C.c_int64(n))
        return out


def rng_start(seed):
    m = MT()
    lib().orc_rng_start(C.byref(m.g), C.c_uint32(seed))
    return m


def inipcl_species(mt, sp: SpeciesInit, spec, n, t):
    p = np.zeros(sp.npin, dtype=PARTICLE)
    tt = (C.c_double * 9)(*[float(x) for x in np.asarray(t).ravel()])
    lib().orc_inipcl_species(C.byref(mt.g), _p(p), C.byref(sp), C.c_int(spec), _i3(n), tt)
    return p


def wrk_shape(n):
    return (n[2] + 9, n[1] + 9, n[0] + 9, 6)


def grdcrt(eb, n, e0, b0, elo=(-1, -1, -1)):
    """eb: halo array [k, j, i, 6] whose first index is elo."""
    eb = np.ascontiguousarray(eb, dtype=np.float64)
    ehi = [elo[d] + eb.shape[2 - d] - 1 for d in range(3)]
    wrk = np.zeros(wrk_shape(n))
    lib().orc_grdcrt(_p(eb), _i3(elo), _i3(ehi), _i3(n), _d3(e0), _d3(b0), _p(wrk))
    return wrk


def positn(p, ustep, sl):
    lib().orc_positn(_p(p), C.c_int64(len(p)), C.c_int(ustep), _d3(sl))


def psolve2(p, counts, qm, wrk, n, ustep, sl):
    cnt = np.asarray(counts, dtype=np.int64)
    qmv = np.asarray(qm, dtype=np.float64)
    lib().orc_psolve2(_p(p), _p(cnt), C.c_int(len(cnt)), _p(qmv), _p(wrk), _i3(n), C.c_int(ustep), _d3(sl))


def rho_shape(n, ncomp=1):
    s = (n[2] + 3, n[1] + 3, n[0] + 3)
    return s if ncomp == 1 else s + (ncomp,)


def charge(p, counts, q, n, with_species=False):
    cnt = np.asarray(counts, dtype=np.int64)
    qv = np.asarray(q, dtype=np.float64)
    rho = np.zeros(rho_shape(n))
    rhodg = np.zeros(rho_shape(n, len(cnt))) if with_species else None
    lib().orc_charge(_p(p), _p(cnt), C.c_int(len(cnt)), _p(qv), _i3(n), _p(rho),
                     _p(rhodg) if with_species else None)
    return (rho, rhodg) if with_species else rho


def fold_rho(rho, n, ncomp=1):
    lib().orc_fold_rho(_p(rho), C.c_int(ncomp), _i3(n))


def chgntr(q, npin, n):
    qv = np.asarray(q, dtype=np.float64)
    npv = np.asarray(npin, dtype=np.int64)
    rhobk = np.zeros(rho_shape(n))
    lib().orc_chgntr(_p(qv), _p(npv), C.c_int(len(qv)), _i3(n), _p(rhobk))
    return rhobk


def add_background(rho, rhobk, n):
    lib().orc_add_background(_p(rho), _p(rhobk), _i3(n))


def energy(p, counts, rm, sl):
    cnt = np.asarray(counts, dtype=np.int64)
    rmv = np.asarray(rm, dtype=np.float64)
    out = np.zeros(3 * len(cnt))
    lib().orc_energy(_p(p), _p(cnt), C.c_int(len(cnt)), _p(rmv), _d3(sl), _p(out))
    return out.reshape(len(cnt), 3)


def field_energy(eb, n, cs, sl, elo=(-1, -1, -1)):
    eb = np.ascontiguousarray(eb, dtype=np.float64)
    ehi = [elo[d] + eb.shape[2 - d] - 1 for d in range(3)]
    out = np.zeros(2)
    lib().orc_field_energy(_p(eb), _i3(elo), _i3(ehi), _i3(n), C.c_double(cs), _d3(sl), _p(out))
    return out


def aj_shape(n):
    return (n[2] + 9, n[1] + 9, n[0] + 9, 3)


def psolve1(p, counts, qm, q, wrk, n, ustep, sl, aj=None):
    cnt = np.asarray(counts, dtype=np.int64)
    qmv = np.asarray(qm, dtype=np.float64)
    qv = np.asarray(q, dtype=np.float64)
    if aj is None:
        aj = np.zeros(aj_shape(n))
    lib().orc_psolve1(_p(p), _p(cnt), C.c_int(len(cnt)), _p(qmv), _p(qv), _p(wrk), _i3(n),
                      C.c_int(ustep), _d3(sl), _p(aj))
    return aj


def fold_current(aj, n):
    lib().orc_fold_current(_p(aj), _i3(n))


def rhoupd(rho, aj, n):
    lib().orc_rhoupd(_p(rho), _p(aj), _i3(n))


def rho_nodes(rho):
    """Interior nodes 0..n-1 of rho(1,-1:n+1)^3 as a compact [k, j, i] array."""
    return rho[1:-2, 1:-2, 1:-2]


def aj_cells(aj):
    """Interior 0..n-1 of aj(3,-4:n+4)^3 as compact [k, j, i, 3]."""
    return aj[4:-5, 4:-5, 4:-5]


def wrk_nodes(wrk):
    """Nodal fields 0..n of wrk(6,-4:n+4)^3 as [k, j, i, 6]."""
    return wrk[4:-4, 4:-4, 4:-4]
