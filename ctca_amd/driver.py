"""Host drivers of the GPU particle path, in the order of the reference's main
loops, for one periodic rank.

ES (src/main/esses.F90): esses_initialize 151-297 then, per step,
esses_mainstep 300-478: grdcrt -> psolve2 -> (spsort) -> charge + fold ->
add_background_charge -> Poisson -> esfld2 -> energy.  ustep = 2, except the
last step of a run (esses.F90:115-119).

EM (src/main/emses.F90:345-573): bfield(1/2) -> grdcrt -> psolve1 (+ current
fold) -> bfield(1/2) -> efield(-2 aj) -> rhoupd (rho by continuity) ->
charge(1, 1) (the per-species diagnostic rhodg; rho keeps its continuity
update).

The particle work runs in libemgpu through ctca_amd.emg; the field solve is
the host side (ctca_amd.fieldsolve).

Job continuation follows &jobcon jobnum(1:2): jobnum(1) = 1 starts from
./SNAPSHOT0/esdat0000.h5 (inipcl.F90:359-426) and takes the half push
positn(1, 1) (esses.F90:198-201, a continued job only); jobnum(2) = 1 writes
./SNAPSHOT1/esdat0000.h5 at the end (esses_finalize, esses.F90:481-494).
"""
import os

import numpy as np

from . import emg, fieldsolve, units


class ESRun:
    """An electrostatic run of a workload on the GPU path."""

    def __init__(self, w: units.Workload, precision=emg.PLAIN_FP64, init="host",
                 count_scale=1.0, tile=0, particles=None, jobnum=(0, 0), root="."):
        self.w = w
        self.su = su = units.to_sunits(w)
        self.n = tuple(w.n)
        self.jobnum = (int(jobnum[0]), int(jobnum[1]))
        if self.jobnum[0] not in (0, 1) or self.jobnum[1] not in (0, 1):
            raise ValueError("jobnum(1) must be 0 or 1 (ES), jobnum(2) 0 or 1")
        self.root = root
        mobile = [k for k, s in enumerate(w.species) if s.mobile]
        self.q = su.q[mobile]
        npin = np.array([int(round(w.species[k].npin * count_scale)) for k in mobile])
        self.ctx = emg.Context(self.n, su.q[mobile], su.qm[mobile], su.rm[mobile], npin,
                               precision=precision, e0=su.e0, b0=su.b0, tile=tile)
        self.rhobk = fieldsolve.background(su.q, su.npin, self.n)
        if self.jobnum[0] == 1:
            from . import snapshot
            snap = snapshot.load(self.ctx, snapshot.snapshot_file(
                os.path.join(root, snapshot.LOAD_DIR), 0))
            self.ctx.set_background(None)   # the host solve adds the background itself
            self.rhobk = np.ascontiguousarray(snap.rhobk[:-1, :-1, :-1])
        elif particles is not None:
            p, counts = particles
            self.ctx.upload(p, counts)
        elif init == "host":
            p = emg.host_inipcl(units.host_species(w), self.n, su.t, w.seed)
            units.scale_to_s(p, su)
            self.ctx.upload(p, npin)
        else:
            self.ctx.init_device(units.device_species(w, su, count_scale), w.seed)
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
        """esses_initialize from chgntr on: the half push positn(1,1) of a
        continued job, charge, Poisson, E, energy (esses.F90:190-263)."""
        if self.jobnum[0] == 1:
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

    def finalize(self):
        """esses_finalize: with jobnum(2) = 1 the particles and the background
        go to ./SNAPSHOT1/esdat0000.h5 (save_esdat, snap.F90:203-356)."""
        if self.jobnum[1] == 1:
            from . import snapshot
            rb = np.broadcast_to(np.asarray(self.rhobk, dtype=np.float64),
                                 (self.n[2], self.n[1], self.n[0]))
            snapshot.save(self.ctx, snapshot.snapshot_file(
                os.path.join(self.root, snapshot.SAVE_DIR), 0), rhobk=rb)

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
        # rho(1, -1:n+1) and rhodg(2 nspec + 1, -1:n+1) of the reference
        # (ohinit.F90:155-159, 189); rho from the initial charge (emses.F90
        # initialisation), then by continuity
        self.rho = np.zeros((nz + 3, ny + 3, nx + 3))
        self.rhodg = np.zeros((2 * len(w.species) + 1, nz + 3, ny + 3, nx + 3))
        self.ctx.charge(0, rho=self.rho)

    def step(self, ustep=2):
        """emses main step: B half step, relocation, push + zigzag current,
        B half step, E full step, rhoupd, charge(1, 1)."""
        self.ebc = fieldsolve.b_half(self.ebc)
        self.ctx.set_fields(fieldsolve.eb_halo(self.ebc[..., :3], self.ebc[..., 3:]))
        self.ctx.push_em(ustep)
        aj = self.ctx.current()
        self.ebc = fieldsolve.b_half(self.ebc)
        self.ebc = fieldsolve.e_full(self.ebc, aj, self.tcs)
        self.rho = fieldsolve.rhoupd(self.rho, aj)        # emses.F90:474
        self.ctx.charge(1, rhodg=self.rhodg)               # emses.F90:475
        return aj

    def close(self):
        self.ctx.close()
