"""Host drivers of the GPU particle path, in the order of the reference's main
loops, for one periodic rank.

ES (src/main/esses.F90): esses_initialize 151-297 then, per step,
esses_mainstep 300-478: grdcrt -> psolve2 -> (spsort) -> charge + fold ->
add_background_charge -> Poisson -> esfld2 -> energy.  ustep = 2, except the
last step of a run (esses.F90:115-119).

EM (src/main/emses.F90:345-573): bfield(1/2) -> grdcrt -> psolve1 (+ current
fold) -> bfield(1/2) -> efield(-2 aj) -> charge (diagnostic).

The particle work runs in libemgpu through ctca_amd.emg; the field solve is
the host side (ctca_amd.fieldsolve).
"""
import numpy as np

from . import emg, fieldsolve, units


class ESRun:
    """An electrostatic run of a workload on the GPU path."""

    def __init__(self, w: units.Workload, precision=emg.PLAIN_FP64, init="host",
                 count_scale=1.0, tile=0, particles=None):
        self.w = w
        self.su = su = units.to_sunits(w)
        self.n = tuple(w.n)
        mobile = [k for k, s in enumerate(w.species) if s.mobile]
        self.q = su.q[mobile]
        npin = np.array([int(round(w.species[k].npin * count_scale)) for k in mobile])
        self.ctx = emg.Context(self.n, su.q[mobile], su.qm[mobile], su.rm[mobile], npin,
                               precision=precision, e0=su.e0, b0=su.b0, tile=tile)
        if particles is not None:
            p, counts = particles
            self.ctx.upload(p, counts)
        elif init == "host":
            p = emg.host_inipcl(units.host_species(w), self.n, su.t, w.seed)
            units.scale_to_s(p, su)
            self.ctx.upload(p, npin)
        else:
            self.ctx.init_device(units.device_species(w, su, count_scale), w.seed)
        self.rhobk = fieldsolve.background(su.q, su.npin, self.n)
        self.istep = 0
        self.history = []

    def solve_fields(self, rho):
        phi = fieldsolve.poisson(rho)
        e = fieldsolve.efield(phi)
        self.ebc = np.concatenate([e, np.zeros_like(e)], axis=-1)
        return fieldsolve.eb_halo(e)

    def charge(self):
        """esses_charge + add_background_charge: folded rho + background."""
        self.ctx.deposit_rho(0)
        rho = self.ctx.rho()
        return rho + self.rhobk

    def initialize(self):
        """esses_initialize from chgntr on: positn(1,1) half push, charge,
        Poisson, E, energy (esses.F90:190-263)."""
        self.ctx.positn(1)
        rho = self.charge()
        self.eb = self.solve_fields(rho)
        self.rho = rho
        return self.energy()

    def energy(self):
        ek = self.ctx.energy()
        ee, eb = fieldsolve.field_energy(self.ebc, self.su.cs)
        self.history.append((self.istep, ee, eb, ek.copy()))
        return ek

    def step(self, ustep=2):
        """esses_mainstep(ustep)."""
        self.istep += 1
        self.ctx.set_fields(self.eb)
        self.ctx.push_es(ustep)
        rho = self.charge()
        self.eb = self.solve_fields(rho)
        self.rho = rho
        return self.energy()

    def run(self, nstep):
        for k in range(1, nstep + 1):
            self.step(2 if k != nstep else 1)

    def close(self):
        self.ctx.close()


class EMRun:
    """An electromagnetic run (zigzag current + host Yee solve)."""

    def __init__(self, w: units.Workload, precision=emg.PLAIN_FP64, init="host",
                 count_scale=1.0, particles=None):
        self.w = w
        self.su = su = units.to_sunits(w)
        self.n = tuple(w.n)
        npin = np.array([int(round(s.npin * count_scale)) for s in w.species])
        self.ctx = emg.Context(self.n, su.q, su.qm, su.rm, npin, precision=precision,
                               e0=su.e0, b0=su.b0)
        if particles is not None:
            p, counts = particles
            self.ctx.upload(p, counts)
        elif init == "host":
            p = emg.host_inipcl(units.host_species(w), self.n, su.t, w.seed)
            units.scale_to_s(p, su)
            self.ctx.upload(p, npin)
        else:
            self.ctx.init_device(units.device_species(w, su, count_scale), w.seed)
        nz, ny, nx = self.n[2], self.n[1], self.n[0]
        self.ebc = np.zeros((nz, ny, nx, 6))
        self.tcs = 2.0 * su.cs

    def step(self, ustep=2):
        """emses main step for the fields that the particles see: B half step,
        relocation, push + zigzag current, B half step, E full step."""
        self.ebc = fieldsolve.b_half(self.ebc)
        self.ctx.set_fields(fieldsolve.eb_halo(self.ebc[..., :3], self.ebc[..., 3:]))
        self.ctx.push_em(ustep)
        aj = self.ctx.current()
        self.ebc = fieldsolve.b_half(self.ebc)
        self.ebc = fieldsolve.e_full(self.ebc, aj, self.tcs)
        return aj

    def close(self):
        self.ctx.close()
