"""ctypes binding of libemgpu.so (the C ABI declared in include/emgpu.h).

The product path runs only through this library: if libemgpu.so is missing or
no GPU is visible, every call raises.  There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libemgpu.so")

MAX_SPECIES = 8
PLAIN_FP64 = 0
FP32_FIXED = 1

KERNEL_PUSH, KERNEL_DEPOSIT, KERNEL_SORT, KERNEL_RELOCATE, KERNEL_ENERGY = range(5)

ERRORS = {0: "EMG_OK", -1: "EMG_ERR_ARG", -2: "EMG_ERR_HIP", -3: "EMG_ERR_NOMEM",
          -4: "EMG_ERR_STATE", -5: "EMG_ERR_DOMAIN", -6: "EMG_ERR_UNSUPPORTED",
          -7: "EMG_ERR_CAPACITY"}

# struct S_particle, lib/ohhelp/ohhelp-1.1.1/oh_part.h:23-28
PARTICLE = np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8"), ("vx", "f8"), ("vy", "f8"),
                     ("vz", "f8"), ("pid", "i4"), ("preside", "i4"), ("nid", "i4"), ("spec", "i4")])
assert PARTICLE.itemsize == 64

EMG_STATE_SLOTS = 64  # saved particle states per context (emg_state_save)

# every entry point declared in include/emgpu.h
EXPORTS = (
    "emg_create", "emg_destroy", "emg_last_error", "emg_set_stream", "emg_synchronize",
    "emg_upload_particles", "emg_download_particles", "emg_particle_counts", "emg_set_fields",
    "emg_positn", "emg_push_es", "emg_deposit_rho", "emg_get_rho", "emg_get_rho_noalloc",
    "emg_get_rhodg", "emg_energy",
    "emg_push_em", "emg_get_current", "emg_sort", "emg_init_device", "emg_host_inipcl",
    "emg_enable_timing", "emg_kernel_stats", "emg_reset_kernel_stats",
    "emg_state_save", "emg_state_load", "emg_adj_reset", "emg_adj_step", "emg_adj_get_field",
    "emg_adj_add_density", "emg_adj_velocity_sums",
    "emg_bfield", "emg_efield", "emg_grdcrt", "emg_get_fields", "emg_field_energy",
    "emg_solve_es", "emg_get_potential", "emg_set_background", "emg_get_ajdg",
)


class EmgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("n", C.c_int32 * 3), ("nspec", C.c_int32), ("precision", C.c_int32),
                ("tile", C.c_int32), ("track_pid", C.c_int32), ("sort_every", C.c_int32),
                ("capacity", C.c_int64 * MAX_SPECIES),
                ("q", C.c_double * MAX_SPECIES), ("qm", C.c_double * MAX_SPECIES),
                ("rm", C.c_double * MAX_SPECIES), ("e0", C.c_double * 3), ("b0", C.c_double * 3)]


class DeviceSpecies(C.Structure):
    _fields_ = [("count", C.c_int64), ("vth_perp", C.c_double), ("vth_para", C.c_double),
                ("drift", C.c_double * 3), ("t", C.c_double * 9), ("ramp_ratio", C.c_double)]


class HostSpecies(C.Structure):
    _fields_ = [("npin", C.c_int64), ("peth", C.c_double), ("path", C.c_double),
                ("spe", C.c_double), ("speth", C.c_double), ("spa", C.c_double),
                ("vdri", C.c_double), ("vdthz", C.c_double), ("vdthxy", C.c_double),
                ("vdx", C.c_double), ("vdy", C.c_double), ("vdz", C.c_double)]


_lib = None


def lib():
    """Load libemgpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "'python -m ctca_amd.build'")
        L = C.CDLL(LIB_PATH)
        i32p = C.POINTER(C.c_int32)
        vp = C.c_void_p
        sig = {
            "emg_create": [C.POINTER(Config), C.POINTER(vp)],
            "emg_destroy": [vp],
            "emg_last_error": [C.c_char_p, C.c_size_t],
            "emg_set_stream": [vp, vp],
            "emg_synchronize": [vp],
            "emg_upload_particles": [vp, vp, vp, C.c_int32],
            "emg_download_particles": [vp, vp, vp],
            "emg_particle_counts": [vp, vp],
            "emg_set_fields": [vp, vp, i32p, i32p],
            "emg_positn": [vp, C.c_int32],
            "emg_push_es": [vp, C.c_int32],
            "emg_deposit_rho": [vp, C.c_int32],
            "emg_get_rho": [vp, vp, i32p, i32p],
            "emg_get_rho_noalloc": [vp, vp, i32p, i32p],
            "emg_get_rhodg": [vp, C.c_int32, vp, i32p, i32p],
            "emg_energy": [vp, vp],
            "emg_push_em": [vp, C.c_int32],
            "emg_get_current": [vp, vp, i32p, i32p],
            "emg_get_ajdg": [vp, C.c_int32, vp, i32p, i32p],
            "emg_sort": [vp],
            "emg_init_device": [vp, C.POINTER(DeviceSpecies), C.c_int32, C.c_uint64],
            "emg_host_inipcl": [C.POINTER(HostSpecies), C.c_int32, i32p, C.POINTER(C.c_double),
                                C.c_uint32, vp],
            "emg_enable_timing": [vp, C.c_int32],
            "emg_kernel_stats": [vp, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int64)],
            "emg_reset_kernel_stats": [vp],
            "emg_state_save": [vp, C.c_int32],
            "emg_state_load": [vp, C.c_int32],
            "emg_adj_reset": [vp],
            "emg_adj_step": [vp, C.c_int32],
            "emg_adj_get_field": [vp, vp],
            "emg_adj_add_density": [vp, vp, vp],
            "emg_adj_velocity_sums": [vp, vp],
            "emg_bfield": [vp],
            "emg_efield": [vp, C.c_double, C.c_double],
            "emg_grdcrt": [vp],
            "emg_get_fields": [vp, vp, i32p, i32p],
            "emg_field_energy": [vp, C.c_double, vp],
            "emg_solve_es": [vp, C.c_double],
            "emg_get_potential": [vp, vp],
            "emg_set_background": [vp, vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib = L
    return _lib


def last_error():
    buf = C.create_string_buffer(1024)
    lib().emg_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def _chk(rc):
    if rc != 0:
        raise EmgError(rc, last_error())
    return rc


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


def _i3(v):
    return (C.c_int32 * 3)(*[int(x) for x in v])


def host_inipcl(species, n, t, seed):
    """emg_host_inipcl: the reference's MT19937 seeded initial records (EMSES-U)."""
    ns = len(species)
    arr = (HostSpecies * ns)()
    total = 0
    for k, s in enumerate(species):
        arr[k] = HostSpecies(int(s["npin"]), s.get("peth", 1.0), s.get("path", 1.0),
                             s.get("spe", 0.0), s.get("speth", 0.0), s.get("spa", 0.0),
                             s.get("vdri", 0.0), s.get("vdthz", 0.0), s.get("vdthxy", 0.0),
                             s.get("vdx", 0.0), s.get("vdy", 0.0), s.get("vdz", 0.0))
        total += int(s["npin"])
    out = np.zeros(total, dtype=PARTICLE)
    tt = (C.c_double * 9)(*[float(v) for v in np.asarray(t, dtype=np.float64).ravel()])
    _chk(lib().emg_host_inipcl(arr, ns, _i3(n), tt, C.c_uint32(seed), _ptr(out)))
    return out


class Context:
    """One GPU-resident particle population (include/emgpu.h emg_ctx)."""

    def __init__(self, n, q, qm, rm, capacity, precision=PLAIN_FP64, e0=(0, 0, 0),
                 b0=(0, 0, 0), tile=0, track_pid=False, sort_every=0):
        ns = len(q)
        cfg = Config()
        cfg.n = (C.c_int32 * 3)(*[int(v) for v in n])
        cfg.nspec = ns
        cfg.precision = int(precision)
        cfg.tile = int(tile)
        cfg.track_pid = 1 if track_pid else 0
        cfg.sort_every = int(sort_every)
        for k in range(ns):
            cfg.capacity[k] = int(capacity[k])
            cfg.q[k] = float(q[k])
            cfg.qm[k] = float(qm[k])
            cfg.rm[k] = float(rm[k])
        for d in range(3):
            cfg.e0[d] = float(e0[d])
            cfg.b0[d] = float(b0[d])
        self.n = tuple(int(v) for v in n)
        self.nspec = ns
        self.precision = precision
        self.h = C.c_void_p()
        _chk(lib().emg_create(C.byref(cfg), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().emg_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- particles
    def upload(self, pbuf, counts):
        pbuf = np.ascontiguousarray(pbuf, dtype=PARTICLE)
        cnt = np.asarray(counts, dtype=np.int64)
        _chk(lib().emg_upload_particles(self.h, _ptr(pbuf), _ptr(cnt), self.nspec))

    def counts(self):
        cnt = np.zeros(self.nspec, dtype=np.int64)
        _chk(lib().emg_particle_counts(self.h, _ptr(cnt)))
        return cnt

    def download(self):
        cnt = self.counts()
        out = np.zeros(int(cnt.sum()), dtype=PARTICLE)
        cnt2 = np.zeros(self.nspec, dtype=np.int64)
        _chk(lib().emg_download_particles(self.h, _ptr(out), _ptr(cnt2)))
        return out, cnt2

    def init_device(self, species, seed):
        arr = (DeviceSpecies * self.nspec)()
        for k, s in enumerate(species):
            d = DeviceSpecies()
            d.count = int(s["count"])
            d.vth_perp = float(s["vth_perp"])
            d.vth_para = float(s["vth_para"])
            for a in range(3):
                d.drift[a] = float(s.get("drift", (0, 0, 0))[a])
            t = np.asarray(s.get("t", np.eye(3)), dtype=np.float64).ravel()
            for a in range(9):
                d.t[a] = float(t[a])
            d.ramp_ratio = float(s.get("ramp_ratio", 0.0))
            arr[k] = d
        _chk(lib().emg_init_device(self.h, arr, self.nspec, C.c_uint64(seed)))

    # ---- fields and steps
    def set_fields(self, eb, lo=(-1, -1, -1)):
        """eb: numpy [k, j, i, 6] array = Fortran eb(6, lo:hi, ...)."""
        eb = np.ascontiguousarray(eb, dtype=np.float64)
        hi = [lo[d] + eb.shape[2 - d] - 1 for d in range(3)]
        _chk(lib().emg_set_fields(self.h, _ptr(eb), _i3(lo), _i3(hi)))

    def positn(self, ustep):
        _chk(lib().emg_positn(self.h, int(ustep)))

    def push_es(self, ustep):
        _chk(lib().emg_push_es(self.h, int(ustep)))

    def deposit_rho(self, func=0):
        _chk(lib().emg_deposit_rho(self.h, int(func)))

    def charge(self, func, rho=None, rhodg=None):
        """charge(ps, func) with the periodic fold into the caller's arrays,
        as the reference leaves them (charge.F90:44-50, 59-143; the fold of
        esses.F90:555-562 and 594-601): func 0 fills rho, func 2 rho and
        rhodg, func 1 only rhodg and leaves rho as it was (the EM loop keeps
        the continuity-updated rho of rhoupd, emses.F90:474-476).  rho is
        [k, j, i] over -1..n+1, rhodg [component, k, j, i] over -1..n+1 with
        component is (0-based) holding species is; the other components
        (the reference's 2 nspec + 1) are zeroed like the reference zeroes
        them."""
        func = int(func)
        if func not in (0, 1, 2):
            raise ValueError("func must be 0, 1 or 2")
        shape = (self.n[2] + 3, self.n[1] + 3, self.n[0] + 3)
        if func in (0, 2) and (rho is None or rho.shape != shape):
            raise ValueError(f"rho must have the shape {shape} (-1..n+1)")
        if func in (1, 2) and (rhodg is None or rhodg.shape[1:] != shape
                               or rhodg.shape[0] < self.nspec):
            raise ValueError(f"rhodg must have the shape (>= nspec, *{shape})")
        self.deposit_rho(func)
        lo, hi = (-1, -1, -1), tuple(v + 1 for v in self.n)
        if func in (0, 2):
            rho[...] = self.rho(lo, hi)
        if func in (1, 2):
            rhodg[...] = 0.0
            for k in range(self.nspec):
                rhodg[k] = self.rho(lo, hi, species=k)

    def push_em(self, ustep):
        _chk(lib().emg_push_em(self.h, int(ustep)))

    # ---- device-resident EM fields (the Yee update on the uploaded eb)
    def bfield(self):
        _chk(lib().emg_bfield(self.h))

    def efield(self, tcs, dlt=2.0):
        _chk(lib().emg_efield(self.h, float(tcs), float(dlt)))

    def grdcrt(self):
        _chk(lib().emg_grdcrt(self.h))

    def fields(self, lo=(-1, -1, -1), hi=None):
        """eb(6, lo:hi) as numpy [k, j, i, 6]; default: the box emg_set_fields took
        for an eb of the usual extent -1..n+1."""
        if hi is None:
            hi = tuple(v + 1 for v in self.n)
        out = np.zeros((hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1, 6))
        _chk(lib().emg_get_fields(self.h, _ptr(out), _i3(lo), _i3(hi)))
        return out

    def solve_es(self, rhobk=0.0):
        _chk(lib().emg_solve_es(self.h, float(rhobk)))

    def set_background(self, rhobk):
        """Per-node background charge [k, j, i] over the periodic nodes, added
        by every following solve_es (None removes it)."""
        if rhobk is None:
            _chk(lib().emg_set_background(self.h, None))
            return
        a = np.ascontiguousarray(rhobk, dtype=np.float64)
        if a.shape != (self.n[2], self.n[1], self.n[0]):
            raise ValueError(f"background shape {a.shape} is not (nz, ny, nx) = "
                             f"{(self.n[2], self.n[1], self.n[0])}")
        _chk(lib().emg_set_background(self.h, _ptr(a)))

    def potential(self):
        """phi of the last device field solve, numpy [k, j, i]."""
        out = np.zeros((self.n[2], self.n[1], self.n[0]))
        _chk(lib().emg_get_potential(self.h, _ptr(out)))
        return out

    def field_energy(self, cs):
        out = np.zeros(2)
        _chk(lib().emg_field_energy(self.h, float(cs), _ptr(out)))
        return out

    def sort(self):
        _chk(lib().emg_sort(self.h))

    def synchronize(self):
        _chk(lib().emg_synchronize(self.h))

    def _bounds(self, lo, hi):
        if lo is None:
            lo = (0, 0, 0)
        if hi is None:
            hi = tuple(v - 1 for v in self.n)
        return lo, hi

    def rho(self, lo=None, hi=None, species=None):
        """rho(1, lo:hi)^3 as numpy [k, j, i]; default: the periodic nodes 0..n-1."""
        lo, hi = self._bounds(lo, hi)
        out = np.zeros((hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1))
        if species is None:
            _chk(lib().emg_get_rho(self.h, _ptr(out), _i3(lo), _i3(hi)))
        else:
            _chk(lib().emg_get_rhodg(self.h, int(species), _ptr(out), _i3(lo), _i3(hi)))
        return out

    def rho_noalloc(self, lo=None, hi=None):
        """rho(1, lo:hi)^3 like rho(), read through emg_get_rho_noalloc (no
        device allocation: the library's static staging array)."""
        lo, hi = self._bounds(lo, hi)
        out = np.zeros((hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1))
        _chk(lib().emg_get_rho_noalloc(self.h, _ptr(out), _i3(lo), _i3(hi)))
        return out

    def current(self, lo=None, hi=None, species=None):
        """aj(3, lo:hi)^3 as numpy [k, j, i, 3]; with species, that species'
        ajdg alone."""
        lo, hi = self._bounds(lo, hi)
        out = np.zeros((hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1, 3))
        if species is None:
            _chk(lib().emg_get_current(self.h, _ptr(out), _i3(lo), _i3(hi)))
        else:
            _chk(lib().emg_get_ajdg(self.h, int(species), _ptr(out), _i3(lo), _i3(hi)))
        return out

    def energy(self):
        out = np.zeros(3 * self.nspec)
        _chk(lib().emg_energy(self.h, _ptr(out)))
        return out.reshape(self.nspec, 3)

    # ---- adjoint window (PLAIN_FP64)
    def state_save(self, slot):
        _chk(lib().emg_state_save(self.h, int(slot)))

    def state_load(self, slot):
        _chk(lib().emg_state_load(self.h, int(slot)))

    def adj_reset(self):
        _chk(lib().emg_adj_reset(self.h))

    def adj_step(self, ustep):
        _chk(lib().emg_adj_step(self.h, int(ustep)))

    def adj_field(self):
        """Adjoint of the nodal E: numpy [k, j, i, 3], periodic nodes 0..n-1."""
        out = np.zeros((self.n[2], self.n[1], self.n[0], 3))
        _chk(lib().emg_adj_get_field(self.h, _ptr(out)))
        return out

    def adj_add_density(self, g, qfac):
        g = np.ascontiguousarray(g, dtype=np.float64)
        assert g.shape == (self.n[2], self.n[1], self.n[0])
        q = np.ascontiguousarray(qfac, dtype=np.float64)
        assert q.shape == (self.nspec,)
        _chk(lib().emg_adj_add_density(self.h, _ptr(g), _ptr(q)))

    def adj_velocity_sums(self):
        out = np.zeros(3 * self.nspec)
        _chk(lib().emg_adj_velocity_sums(self.h, _ptr(out)))
        return out.reshape(self.nspec, 3)

    # ---- timing
    def enable_timing(self, on=True):
        _chk(lib().emg_enable_timing(self.h, 1 if on else 0))

    def reset_stats(self):
        _chk(lib().emg_reset_kernel_stats(self.h))

    def kernel_stats(self, which):
        ms = C.c_double()
        n = C.c_int64()
        _chk(lib().emg_kernel_stats(self.h, int(which), C.byref(ms), C.byref(n)))
        return ms.value, n.value
