"""Workloads of the particle path and the reference's unit conversions.

A workload is given in the reference's input units (EMSES-U: dr = 1, the
electron plasma frequency wp(1) = 1, velocities in dr per unit time).
inital / rescal turn those into the S-units the particle kernels run in:

  * q(is)  = wp(is)^2 slx sly slz / npin(is) / qm(is),  rm = q / qm
    (src/main/inital.F90:231, 271);
  * b0 = wc / qm(1) rotated by T(phiz, phixy) (inital.F90:116-140);
  * rescal (src/main/rescal.F90:38-193): dth = dt/2,
    renv = dth/dr, renq = qm1 dth^2, rene = qm1 dth^2, renb = qm1 dth,
    renm = (qm1 dth)^2, qm -> qm/qm1, positions unchanged (dr = 1).

In S-units a full step is ustep = 2 and the kick factor is qm * ustep / 2
(src/pcl2field/psolve2.F90:147-150).
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


@dataclass
class Species:
    name: str
    qm: float                 # charge-to-mass ratio (EMSES-U); species 0 must be the electrons
    wp: float                 # plasma frequency (EMSES-U)
    npin: int                 # macro particles
    vth: float                # isotropic thermal speed (peth = path)
    drift: tuple = (0.0, 0.0, 0.0)   # cartesian drift (vdx, vdy, vdz)
    vth_para: Optional[float] = None  # parallel thermal speed if anisotropic
    mobile: bool = True       # immobile species contribute only the neutralising background


@dataclass
class Workload:
    name: str
    n: tuple                  # cells (nx, ny, nz)
    dt: float
    steps: int
    species: List[Species]
    wc: float = 0.0           # electron cyclotron frequency (EMSES-U)
    phiz: float = 0.0         # B direction angles (deg), inital.F90:116-140
    phixy: float = 0.0
    cv: float = 0.0           # speed of light (EMSES-U), EM only
    e0: tuple = (0.0, 0.0, 0.0)   # uniform applied E (EMSES-U)
    em: bool = False
    ramp_ratio: float = 0.0   # >1: electron/ion density ~ exp(z ln(ratio) / Lz)
    seed: int = 10            # inital: (myid + 1) * 10 for rank 0
    notes: str = ""


def rotation(phiz, phixy):
    """T(phiz, phixy), inital.F90:126-135 (row-major)."""
    pi = np.arccos(-1.0)
    csz, snz = np.cos(pi * phiz / 180.0), np.sin(pi * phiz / 180.0)
    csxy, snxy = np.cos(pi * phixy / 180.0), np.sin(pi * phixy / 180.0)
    return np.array([[csz * csxy, -snxy, snz * csxy],
                     [csz * snxy, csxy, snz * snxy],
                     [-snz, 0.0, csz]])


@dataclass
class SUnits:
    """Everything the kernels need, in S-units."""
    n: tuple
    q: np.ndarray
    qm: np.ndarray
    rm: np.ndarray
    b0: np.ndarray
    e0: np.ndarray
    cv: float
    cs: float
    renv: float
    renq: float
    rene: float
    renrho: float
    rened: float
    t: np.ndarray
    npin: np.ndarray
    mobile: np.ndarray
    dt: float
    extra: dict = field(default_factory=dict)


def to_sunits(w: Workload) -> SUnits:
    sp = w.species
    qm1 = sp[0].qm
    dr = 1.0
    sl = np.array(w.n, dtype=np.float64) * dr
    vol = sl[0] * sl[1] * sl[2]
    q_u = np.array([s.wp ** 2 * vol / s.npin / s.qm for s in sp])
    rm_u = q_u / np.array([s.qm for s in sp])
    dth = w.dt / 2.0
    renv = dth / dr
    renq = qm1 * (dth / dr) ** 2 / dr
    rene = qm1 * dth ** 2 / dr
    renb = qm1 * dth
    renrho = qm1 * dth ** 2
    renm = (qm1 * dth / dr) ** 2 / dr
    rened = qm1 ** 2 * dth ** 4 / dr ** 2
    t = rotation(w.phiz, w.phixy)
    b0 = w.wc / qm1
    b0v = b0 * t[:, 2]
    if b0 != 0.0:
        b0v = np.where(np.abs(b0v / b0) < 1e-10, 0.0, b0v)
    cv = w.cv * renv
    return SUnits(n=tuple(w.n), q=q_u * renq, qm=np.array([s.qm / qm1 for s in sp]),
                  rm=rm_u * renm, b0=b0v * renb, e0=np.array(w.e0) * rene, cv=cv, cs=cv * cv,
                  renv=renv, renq=renq, rene=rene, renrho=renrho, rened=rened, t=t,
                  npin=np.array([s.npin for s in sp], dtype=np.int64),
                  mobile=np.array([s.mobile for s in sp]), dt=w.dt)


def host_species(w: Workload):
    """inipcl inputs (EMSES-U) of the mobile species, for the MT19937 host init."""
    out = []
    for s in w.species:
        if not s.mobile:
            continue
        vpa = s.vth if s.vth_para is None else s.vth_para
        out.append(dict(npin=s.npin, peth=s.vth, path=vpa, spe=0.0, speth=0.0, spa=0.0,
                        vdri=0.0, vdthz=0.0, vdthxy=0.0, vdx=s.drift[0], vdy=s.drift[1],
                        vdz=s.drift[2]))
    return out


def device_species(w: Workload, su: SUnits, count_scale=1.0):
    """emg_init_device inputs (S-units) of the mobile species."""
    out = []
    for k, s in enumerate(w.species):
        if not s.mobile:
            continue
        vpa = s.vth if s.vth_para is None else s.vth_para
        out.append(dict(count=int(round(s.npin * count_scale)), vth_perp=s.vth * su.renv,
                        vth_para=vpa * su.renv, drift=tuple(d * su.renv for d in s.drift),
                        t=su.t, ramp_ratio=w.ramp_ratio))
    return out


def scale_to_s(p, su: SUnits):
    """rescal.F90:178-185 on particle records (positions unchanged for dr = 1)."""
    p["vx"] *= su.renv
    p["vy"] *= su.renv
    p["vz"] *= su.renv


# ---------------------------------------------------------------------------
# The five workloads of the particle path
# ---------------------------------------------------------------------------

def es32():
    """Plain ES, CPU-runnable: 32^3 periodic, e- (qm=-1, vth 1) and ions
    (qm=+0.01, vth 0.1), 16 particles/cell each, dt 0.1, 100 steps, FP64."""
    nc = 32 ** 3
    return Workload("es32_2sp_16ppc", (32, 32, 32), 0.1, 100, [
        Species("electron", -1.0, 1.0, 16 * nc, 1.0),
        Species("ion", 0.01, 0.1, 16 * nc, 0.1),
    ], seed=10)


def mag128():
    """Magnetised scale-up: 128^3, two species of 64 particles/cell
    (268,435,456 particles), B along z with wce/wpe = 0.5, vth_e 1.0,
    vth_i 0.05, mass ratio 100, dt 0.05, 200 steps."""
    nc = 128 ** 3
    return Workload("mag128_2sp_64ppc", (128, 128, 128), 0.05, 200, [
        Species("electron", -1.0, 1.0, 64 * nc, 1.0),
        Species("ion", 0.01, 0.1, 64 * nc, 0.05),
    ], wc=0.5, seed=20)


def em_twostream():
    """EM two-stream: 256x256x128, background e- 32 ppc (vth 1.0) plus a beam
    of 10% of the background density drifting +5.0 along x at equal macro
    weight (26,843,546 particles), immobile neutralising ions, c = 20,
    dt = 0.02, 500 steps."""
    nc = 256 * 256 * 128
    nb = 32 * nc
    nbeam = int(round(0.1 * nb))
    return Workload("em_twostream_256x256x128", (256, 256, 128), 0.02, 500, [
        Species("electron", -1.0, 1.0, nb, 1.0),
        Species("beam", -1.0, float(np.sqrt(0.1)), nbeam, 1.0, drift=(5.0, 0.0, 0.0)),
    ], cv=20.0, em=True, seed=30,
        notes="beam density 10% of the background; ions as a uniform background")


def adjoint64():
    """Adjoint window: 64^3 ES periodic, e- 32 ppc (vth 1.0) plus a 5% density
    beam drifting v_d = 3.0 along x, immobile ions, uniform applied E0 = 0.01
    along x, K = 50 steps (dt 0.1), checkpoints every 10 steps."""
    nc = 64 ** 3
    ne = 32 * nc
    nbeam = int(round(0.05 * ne))
    return Workload("adjoint64_beam", (64, 64, 64), 0.1, 50, [
        Species("electron", -1.0, 1.0, ne, 1.0),
        Species("beam", -1.0, float(np.sqrt(0.05)), nbeam, 1.0, drift=(3.0, 0.0, 0.0)),
    ], e0=(0.01, 0.0, 0.0), seed=40)


def ramp512():
    """Large nonuniform plasma: 512x512x256, 8 particles/cell on average for
    electrons and for ions (536,870,912 each, 1,073,741,824 total), density
    ~ exp(z ln 100 / Lz), vth_e 1.0, vth_i 0.1, B along x with wce/wpe = 0.2,
    dt 0.05, 50 steps."""
    nc = 512 * 512 * 256
    return Workload("ramp512_2sp_8ppc", (512, 512, 256), 0.05, 50, [
        Species("electron", -1.0, 1.0, 8 * nc, 1.0),
        Species("ion", 0.01, 0.1, 8 * nc, 0.1),
    ], wc=0.2, phiz=90.0, phixy=0.0, ramp_ratio=100.0, seed=50)


def ramp512_slab():
    """One GPU's share of ramp512 split into 8 z-slabs: the densest slab
    (z in [224, 256), 512x512x32 cells) holds a fraction (100 - 100^(7/8)) / 99
    = 44.2% of each species (237,339,608 particles each), with the density
    rising by 100^(1/8) across it.  Macro-particle charge as in the full box
    (wp scaled by the slab's density), so the kernels see the real shape and
    occupancy of the worst-loaded GPU."""
    full = 8 * 512 * 512 * 256
    frac = (100.0 - 100.0 ** (7.0 / 8.0)) / 99.0
    npin = int(round(frac * full))
    wp = float(np.sqrt(frac * 8.0))   # same q per particle as the full box
    return Workload("ramp512_slab_densest", (512, 512, 32), 0.05, 50, [
        Species("electron", -1.0, wp, npin, 1.0),
        Species("ion", 0.01, 0.1 * wp, npin, 0.1),
    ], wc=0.2, phiz=90.0, phixy=0.0, ramp_ratio=100.0 ** (1.0 / 8.0), seed=51)


WORKLOADS = {w().name: w for w in (es32, mag128, em_twostream, adjoint64, ramp512,
                                   ramp512_slab)}
