"""Checkpoint / resume in the reference's snapshot format.

The reference's ES job writes its state at the end of a run into
./SNAPSHOT1/esdat%04d.h5, one HDF5 file per rank (save_esdat,
src/io/snap.F90:203-356), and a continued job (jobnum(1) = 1 for esdat,
2 for the emdat files of the same layout) reads ./SNAPSHOT0/ in four places:

  * ohinit.F90:225-244   'sdoms'  must equal the job's subdomain, else stop;
  * inipcl.F90:359-426   'np', 'rens', the particles 'px' .. 'pvz', 'pres',
                         'pis', then specsort (records grouped by species);
  * rescal.F90:102-119, 186-192   positions and velocities scaled by
                         renr / rens(1) and renv / rens(2);
  * inifld.F90:59-74     'rhobk', the background charge on the nodes.

The datasets (snap.F90:272-344, written through hdf5io.F90 wrt1i/wrt2i/wrt1d/
wrt3d as H5T_NATIVE_INTEGER (32-bit) or H5T_NATIVE_DOUBLE):

  nnode  int[1]      (rank 0 only)      nspec  int[1]
  np     int[nspec]  particles per species (primary + secondary)
  npc    int[1]      bodies             sdoms  int[3][2]   (Fortran sdoms(2,3))
  rens   double[4]   renr, renv, renrho, renq
  px py pz pvx pvy pvz   double[N]      pres pis   int[N]
  rhobk  double[nz+1][ny+1][nx+1]   (Fortran rhobk(1,0:nx,0:ny,0:nz,3))
  chgacm double[npc]

The HDF5 Fortran interface reverses dimensions, so Fortran's sdoms(2,3) is a
(3, 2) dataset and rhobk(0:nx,0:ny,0:nz) a (nz+1, ny+1, nx+1) one: numpy's
C order reads them as [axis][lower, upper] and [k][j][i].

The device state comes down through emg_download_particles and goes back up
through emg_upload_particles; this module only formats it.  HDF5 is reached
through ctypes (the system's libhdf5, loaded on first use); nothing here
touches the GPU.
"""
import ctypes as C
import ctypes.util
import glob
import os

import numpy as np

from . import emg

SAVE_DIR = "SNAPSHOT1"   # where a job writes (snap.F90:268)
LOAD_DIR = "SNAPSHOT0"   # where a continued job reads (inipcl.F90:362)
ESDAT, EMDAT = "esdat", "emdat"

_CHUNK = 1 << 22          # particles per transfer (bounds the host temporaries)


class SnapshotError(RuntimeError):
    pass


def snapshot_file(directory, myid=0, kind=ESDAT):
    """<directory>/<kind>%04d.h5 (snap.F90:268, inipcl.F90:361-365)."""
    if kind not in (ESDAT, EMDAT):
        raise ValueError(f"kind must be {ESDAT!r} or {EMDAT!r}")
    return os.path.join(directory, f"{kind}{int(myid):04d}.h5")


# ---------------------------------------------------------------- HDF5 access
_h5 = None
_HID = C.c_int64
_HSIZE = C.c_ulonglong
_P_DEFAULT = 0
_S_ALL = 0
_ACC_RDONLY = 0x0000
_ACC_TRUNC = 0x0002
_T_INTEGER, _T_FLOAT = 0, 1


def _find_hdf5():
    env = os.environ.get("EMG_HDF5_LIB")
    if env:
        return env
    name = ctypes.util.find_library("hdf5")
    if name:
        return name
    for pat in ("/opt/conda/lib/libhdf5.so*", "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so*",
                "/usr/lib/x86_64-linux-gnu/libhdf5.so*", "/usr/lib64/libhdf5.so*"):
        hits = sorted(glob.glob(pat))
        if hits:
            return hits[0]
    return None


def hdf5_available():
    try:
        _lib()
        return True
    except SnapshotError:
        return False


def _lib():
    global _h5
    if _h5 is not None:
        return _h5
    path = _find_hdf5()
    if not path:
        raise SnapshotError("libhdf5 not found (set EMG_HDF5_LIB to its path)")
    try:
        h = C.CDLL(path)
    except OSError as e:
        raise SnapshotError(f"cannot load {path}: {e}") from e
    sig = {
        "H5open": (C.c_int, []),
        "H5Eset_auto2": (C.c_int, [_HID, C.c_void_p, C.c_void_p]),
        "H5Fcreate": (_HID, [C.c_char_p, C.c_uint, _HID, _HID]),
        "H5Fopen": (_HID, [C.c_char_p, C.c_uint, _HID]),
        "H5Fclose": (C.c_int, [_HID]),
        "H5Screate_simple": (_HID, [C.c_int, C.POINTER(_HSIZE), C.POINTER(_HSIZE)]),
        "H5Sselect_hyperslab": (C.c_int, [_HID, C.c_int, C.POINTER(_HSIZE), C.POINTER(_HSIZE),
                                          C.POINTER(_HSIZE), C.POINTER(_HSIZE)]),
        "H5Sget_simple_extent_ndims": (C.c_int, [_HID]),
        "H5Sget_simple_extent_dims": (C.c_int, [_HID, C.POINTER(_HSIZE), C.POINTER(_HSIZE)]),
        "H5Sclose": (C.c_int, [_HID]),
        "H5Dcreate2": (_HID, [_HID, C.c_char_p, _HID, _HID, _HID, _HID, _HID]),
        "H5Dopen2": (_HID, [_HID, C.c_char_p, _HID]),
        "H5Dget_space": (_HID, [_HID]),
        "H5Dget_type": (_HID, [_HID]),
        "H5Dwrite": (C.c_int, [_HID, _HID, _HID, _HID, _HID, C.c_void_p]),
        "H5Dread": (C.c_int, [_HID, _HID, _HID, _HID, _HID, C.c_void_p]),
        "H5Dclose": (C.c_int, [_HID]),
        "H5Tget_class": (C.c_int, [_HID]),
        "H5Tget_size": (C.c_size_t, [_HID]),
        "H5Tclose": (C.c_int, [_HID]),
        "H5Lexists": (C.c_int, [_HID, C.c_char_p, _HID]),
    }
    for name, (res, args) in sig.items():
        f = getattr(h, name)
        f.restype = res
        f.argtypes = args
    if h.H5open() < 0:
        raise SnapshotError("H5open failed")
    h.H5Eset_auto2(0, None, None)   # errors come back as return codes, not prints
    h.T_INT = _HID.in_dll(h, "H5T_NATIVE_INT_g").value
    h.T_DOUBLE = _HID.in_dll(h, "H5T_NATIVE_DOUBLE_g").value
    _h5 = h
    return h


def _ok(rc, what):
    if rc < 0:
        raise SnapshotError(f"HDF5: {what} failed")
    return rc


def _dims(shape):
    return (_HSIZE * len(shape))(*[int(v) for v in shape])


class _File:
    def __init__(self, path, write):
        h = _lib()
        self.h = h
        b = os.fsencode(path)
        self.id = (h.H5Fcreate(b, _ACC_TRUNC, _P_DEFAULT, _P_DEFAULT) if write
                   else h.H5Fopen(b, _ACC_RDONLY, _P_DEFAULT))
        if self.id < 0:
            raise SnapshotError(f"cannot {'create' if write else 'open'} {path}")
        self.path = path

    def close(self):
        if self.id >= 0:
            _ok(self.h.H5Fclose(self.id), "close")
            self.id = -1

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def has(self, name):
        return self.h.H5Lexists(self.id, name.encode(), _P_DEFAULT) > 0

    # a dataset of `shape` and memory type `mtype`, filled by `fill(dset, filespace)`
    def _create(self, name, shape, mtype):
        h = self.h
        sp = _ok(h.H5Screate_simple(len(shape), _dims(shape), None), f"dataspace {name}")
        try:
            d = _ok(h.H5Dcreate2(self.id, name.encode(), mtype, sp, _P_DEFAULT, _P_DEFAULT,
                                 _P_DEFAULT), f"create {name}")
        except SnapshotError:
            h.H5Sclose(sp)
            raise
        return d, sp

    def write(self, name, arr, mtype):
        """A whole array in one transfer (small datasets)."""
        a = np.ascontiguousarray(arr)
        d, sp = self._create(name, a.shape, mtype)
        try:
            if a.size:
                _ok(self.h.H5Dwrite(d, mtype, _S_ALL, _S_ALL, _P_DEFAULT, a.ctypes.data),
                    f"write {name}")
        finally:
            self.h.H5Dclose(d)
            self.h.H5Sclose(sp)

    def write_column(self, name, col, mtype, ctype):
        """A 1-D dataset from a (possibly strided) column, in chunks."""
        h = self.h
        n = len(col)
        d, sp = self._create(name, (n,), mtype)
        try:
            for a in range(0, n, _CHUNK):
                b = min(n, a + _CHUNK)
                buf = np.ascontiguousarray(col[a:b], dtype=ctype)
                ms = _ok(h.H5Screate_simple(1, _dims((b - a,)), None), "memory space")
                try:
                    _ok(h.H5Sselect_hyperslab(sp, 0, _dims((a,)), None, _dims((b - a,)), None),
                        "select")
                    _ok(h.H5Dwrite(d, mtype, ms, sp, _P_DEFAULT, buf.ctypes.data), f"write {name}")
                finally:
                    h.H5Sclose(ms)
        finally:
            h.H5Dclose(d)
            h.H5Sclose(sp)

    def info(self, name):
        """(shape, 'i' | 'f', element size) of a dataset."""
        h = self.h
        d = h.H5Dopen2(self.id, name.encode(), _P_DEFAULT)
        if d < 0:
            raise SnapshotError(f"{self.path}: no dataset {name!r}")
        try:
            sp = _ok(h.H5Dget_space(d), "space")
            t = _ok(h.H5Dget_type(d), "type")
            try:
                nd = _ok(h.H5Sget_simple_extent_ndims(sp), "rank")
                dims = (_HSIZE * max(nd, 1))()
                _ok(h.H5Sget_simple_extent_dims(sp, dims, None), "dims")
                cls = h.H5Tget_class(t)
                kind = {_T_INTEGER: "i", _T_FLOAT: "f"}.get(cls, "?")
                return tuple(int(dims[k]) for k in range(nd)), kind, int(h.H5Tget_size(t))
            finally:
                h.H5Tclose(t)
                h.H5Sclose(sp)
        finally:
            h.H5Dclose(d)

    def read(self, name, ctype, shape=None):
        """A whole dataset converted to ctype ('i4' or 'f8'); its rank and
        extent are checked against `shape` when given (read1i/read1d check
        the rank and that the file's extent is not larger, hdf5io.F90:165-183)."""
        h = self.h
        fshape, _, _ = self.info(name)
        if shape is not None and tuple(shape) != fshape:
            raise SnapshotError(f"{self.path}: dataset {name!r} has shape {fshape}, "
                                f"expected {tuple(shape)}")
        out = np.empty(fshape, dtype=ctype)
        mtype = h.T_INT if np.dtype(ctype) == np.int32 else h.T_DOUBLE
        d = _ok(h.H5Dopen2(self.id, name.encode(), _P_DEFAULT), f"open {name}")
        try:
            if out.size:
                _ok(h.H5Dread(d, mtype, _S_ALL, _S_ALL, _P_DEFAULT, out.ctypes.data),
                    f"read {name}")
        finally:
            h.H5Dclose(d)
        return out

    def read_column(self, name, n, ctype, dst):
        """A 1-D dataset of n elements into dst (a possibly strided column)."""
        h = self.h
        fshape, _, _ = self.info(name)
        if fshape != (n,):
            raise SnapshotError(f"{self.path}: dataset {name!r} has shape {fshape}, expected ({n},)")
        mtype = h.T_INT if np.dtype(ctype) == np.int32 else h.T_DOUBLE
        d = _ok(h.H5Dopen2(self.id, name.encode(), _P_DEFAULT), f"open {name}")
        sp = _ok(h.H5Dget_space(d), "space")
        try:
            for a in range(0, n, _CHUNK):
                b = min(n, a + _CHUNK)
                buf = np.empty(b - a, dtype=ctype)
                ms = _ok(h.H5Screate_simple(1, _dims((b - a,)), None), "memory space")
                try:
                    _ok(h.H5Sselect_hyperslab(sp, 0, _dims((a,)), None, _dims((b - a,)), None),
                        "select")
                    _ok(h.H5Dread(d, mtype, ms, sp, _P_DEFAULT, buf.ctypes.data), f"read {name}")
                finally:
                    h.H5Sclose(ms)
                dst[a:b] = buf
        finally:
            h.H5Sclose(sp)
            h.H5Dclose(d)


# ------------------------------------------------------------ the esdat format
def periodic_sdoms(n):
    """sdoms(1:2, 1:3) of the one periodic rank: [0, n_d] per axis."""
    return np.array([[0, int(n[0])], [0, int(n[1])], [0, int(n[2])]], dtype=np.int32)


def _rhobk_nodes(rhobk, n):
    """The file's rhobk(0:nx, 0:ny, 0:nz) as [k][j][i]: zeros if None; a
    periodic-node array (nz, ny, nx) is extended by its images (node n_d is
    node 0 on one periodic rank)."""
    nx, ny, nz = (int(v) for v in n)
    if rhobk is None:
        return np.zeros((nz + 1, ny + 1, nx + 1))
    a = np.asarray(rhobk, dtype=np.float64)
    if a.shape == (nz + 1, ny + 1, nx + 1):
        return a
    if a.shape == (nz, ny, nx):
        return np.pad(a, [(0, 1), (0, 1), (0, 1)], mode="wrap")
    raise ValueError(f"rhobk shape {a.shape}: expected {(nz, ny, nx)} or {(nz + 1, ny + 1, nx + 1)}")


def write_esdat(path, pbuf, counts, n, rhobk=None, chgacm=None, rens=(1.0, 1.0, 1.0, 1.0),
                nnode=1, myid=0, sdoms=None):
    """save_esdat (snap.F90:203-356) for one rank: the records of `pbuf`
    (PARTICLE, the species segments of `counts`, as emg_download_particles
    returns them) in their order, positions and velocities as they are (the
    reference writes them unscaled: its rescale lines are commented out,
    snap.F90:250-260).  'nnode' goes only into rank 0's file."""
    pbuf = np.asarray(pbuf)
    if pbuf.dtype != emg.PARTICLE:
        raise ValueError("pbuf must be PARTICLE records")
    counts = np.asarray(counts, dtype=np.int64)
    if int(counts.sum()) != len(pbuf):
        raise ValueError(f"counts sum to {int(counts.sum())}, pbuf holds {len(pbuf)}")
    if len(pbuf) > np.iinfo(np.int32).max:
        raise ValueError("more particles than a 32-bit 'np' can count")
    chg = np.zeros(0) if chgacm is None else np.asarray(chgacm, dtype=np.float64).ravel()
    sd = periodic_sdoms(n) if sdoms is None else np.asarray(sdoms, dtype=np.int32)
    if sd.shape != (3, 2):
        raise ValueError("sdoms must be [axis][lower, upper]")
    rb = _rhobk_nodes(rhobk, (sd[0, 1] - sd[0, 0], sd[1, 1] - sd[1, 0], sd[2, 1] - sd[2, 0]))
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with _File(path, write=True) as f:
        ti, td = f.h.T_INT, f.h.T_DOUBLE
        if int(myid) == 0:
            f.write("nnode", np.array([nnode], dtype=np.int32), ti)
        f.write("nspec", np.array([len(counts)], dtype=np.int32), ti)
        f.write("np", counts.astype(np.int32), ti)
        f.write("npc", np.array([len(chg)], dtype=np.int32), ti)
        f.write("sdoms", sd, ti)
        f.write("rens", np.asarray(rens, dtype=np.float64).reshape(4), td)
        for ds, fld in (("px", "x"), ("py", "y"), ("pz", "z"), ("pvx", "vx"), ("pvy", "vy"),
                        ("pvz", "vz")):
            f.write_column(ds, pbuf[fld], td, np.float64)
        f.write_column("pres", pbuf["preside"], ti, np.int32)
        f.write_column("pis", pbuf["spec"], ti, np.int32)
        f.write("rhobk", rb, td)
        f.write("chgacm", chg, td)


class Snapshot:
    """What a continued job takes from one esdat/emdat file."""

    def __init__(self, nnode, nspec, np_, npc, sdoms, rens, pbuf, rhobk, chgacm):
        self.nnode = nnode        # None in the files of ranks > 0
        self.nspec = nspec
        self.np = np_             # particles per species
        self.npc = npc
        self.sdoms = sdoms        # [axis][lower, upper]
        self.rens = rens          # renr, renv, renrho, renq of the writing job
        self.pbuf = pbuf          # PARTICLE records in the file's order
        self.rhobk = rhobk        # [k][j][i], nodes 0..n_d
        self.chgacm = chgacm

    @property
    def n(self):
        return tuple(int(self.sdoms[d, 1] - self.sdoms[d, 0]) for d in range(3))


def read_esdat(path):
    """Every dataset of one file, the particles as PARTICLE records in the
    file's order (pid = nid = 0, preside from 'pres', spec from 'pis')."""
    with _File(path, write=False) as f:
        nnode = int(f.read("nnode", np.int32, (1,))[0]) if f.has("nnode") else None
        nspec = int(f.read("nspec", np.int32, (1,))[0])
        np_ = f.read("np", np.int32, (nspec,)).astype(np.int64)
        npc = int(f.read("npc", np.int32, (1,))[0])
        sdoms = f.read("sdoms", np.int32, (3, 2))
        rens = f.read("rens", np.float64, (4,))
        total = int(np_.sum())
        pbuf = np.zeros(total, dtype=emg.PARTICLE)
        for ds, fld in (("px", "x"), ("py", "y"), ("pz", "z"), ("pvx", "vx"), ("pvy", "vy"),
                        ("pvz", "vz")):
            f.read_column(ds, total, np.float64, pbuf[fld])
        f.read_column("pres", total, np.int32, pbuf["preside"])
        f.read_column("pis", total, np.int32, pbuf["spec"])
        ext = tuple(int(sdoms[d, 1] - sdoms[d, 0]) + 1 for d in (2, 1, 0))
        rhobk = f.read("rhobk", np.float64, ext)
        chgacm = f.read("chgacm", np.float64, (npc,))
    return Snapshot(nnode, nspec, np_, npc, sdoms, rens, pbuf, rhobk, chgacm)


def check_sdoms(snap, n, sdoms=None):
    """ohinit.F90:233-241: a continued job stops when the file's subdomain
    differs from its own."""
    want = periodic_sdoms(n) if sdoms is None else np.asarray(sdoms, dtype=np.int32)
    if not np.array_equal(np.asarray(snap.sdoms), want):
        raise SnapshotError("sdoms is inconsistent with continued job data: "
                            f"file {snap.sdoms.tolist()}, job {want.tolist()}")


def specsort(pbuf, nspec):
    """The records grouped by species 1..nspec, stable within a species (the
    reference sorts the loaded buffer by 'spec' before use, inipcl.F90:408),
    with the counts per species."""
    spec = pbuf["spec"]
    if len(spec) and (spec.min() < 1 or spec.max() > nspec):
        raise SnapshotError(f"species index outside 1..{nspec}")
    order = np.argsort(spec, kind="stable")
    counts = np.bincount(spec, minlength=nspec + 1)[1:nspec + 1].astype(np.int64)
    return pbuf[order], counts


def rescale(pbuf, rens_file, renr, renv):
    """rescal.F90:186-192: x * (1 / rens(1)) * renr and v * (1 / rens(2)) *
    renv, in that order of operations (in place)."""
    ir, iv = 1.0 / float(rens_file[0]), 1.0 / float(rens_file[1])
    for f in ("x", "y", "z"):
        pbuf[f] = pbuf[f] * ir * float(renr)
    for f in ("vx", "vy", "vz"):
        pbuf[f] = pbuf[f] * iv * float(renv)
    return pbuf


def rescale_rhobk(rhobk, rens_file, renrho):
    """rescal.F90:234: rhobk * (1 / rens(3)) * renrho."""
    return rhobk * (1.0 / float(rens_file[2])) * float(renrho)


# ------------------------------------------------------- device state <-> file
def save(ctx, path, rhobk=None, chgacm=None, rens=(1.0, 1.0, 1.0, 1.0), nnode=1, myid=0):
    """Download a context's particles and write them with the job's rhobk
    and chgacm (esses_finalize -> save_esdat, esses.F90:481-494)."""
    pbuf, counts = ctx.download()
    write_esdat(path, pbuf, counts, ctx.n, rhobk=rhobk, chgacm=chgacm, rens=rens, nnode=nnode,
                myid=myid)


def load(ctx, path, renr=None, renv=None, renrho=None):
    """A continued job's start from one file: the sdoms check, the species
    sort, the rescale to this job's units when renr/renv/renrho are given
    (with the file's own values the scale is exactly 1), the upload, and the
    background charge of the file's rhobk at the periodic nodes (removed if
    it is zero everywhere).  Returns the Snapshot."""
    snap = read_esdat(path)
    check_sdoms(snap, ctx.n)
    if snap.nspec != ctx.nspec:
        raise SnapshotError(f"file holds {snap.nspec} species, the context {ctx.nspec}")
    pbuf, counts = specsort(snap.pbuf, snap.nspec)
    if not np.array_equal(counts, snap.np):
        raise SnapshotError(f"'pis' counts {counts.tolist()} differ from 'np' {snap.np.tolist()}")
    if renr is not None or renv is not None:
        rescale(pbuf, snap.rens, snap.rens[0] if renr is None else renr,
                snap.rens[1] if renv is None else renv)
    pbuf["pid"] = 0
    pbuf["nid"] = 0
    ctx.upload(pbuf, counts)
    rb = snap.rhobk if renrho is None else rescale_rhobk(snap.rhobk, snap.rens, renrho)
    per = np.ascontiguousarray(rb[:-1, :-1, :-1])
    ctx.set_background(per if np.any(per) else None)
    return snap
