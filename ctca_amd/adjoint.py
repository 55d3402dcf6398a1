"""Differentiable electrostatic window on the GPU particle path.

The reference has no adjoint; this harness differentiates a window of its ES
loop (src/main/esses.F90:300-478) with a hand-written reverse step:

  forward, per step k (S-units, ustep = 2):
    rho_k  = deposit(x_k) + background             (charge.F90, chgntr.F90)
    phi_k  = P rho_k                               (periodic spectral Poisson,
                                                    inifft.F90:390-480, zero mode dropped)
    e_k    = G phi_k + E0 x                        (staggered E = -grad phi,
                                                    esfild.F90:351-367; E0 as grdcrt adds e0x)
    (x, v) <- push_es(x, v; e_k)                   (psolve2.F90, libemgpu)
  loss:  L = |F|^2,  F = sum_nodes rho_e(x_K) exp(-2 pi i n_x / N_x)
         (the k = (1, 0, 0) mode of the electron charge density at step K)
  parameters: the beam drift v_d (added to the beam's initial vx) and E0.

Reverse sweep: libemgpu's emg_adj_step maps the particle adjoints across one
push and returns the adjoint of the nodal E (the transposed gather); here it
goes back through the grdcrt relocation R, the staggered gradient G and the
symmetric operator P to the adjoint of rho, which emg_adj_add_density turns
into the particles' position adjoints (the transposed CIC).  The particle
state is saved every `checkpoint` steps and the states of a window are
recomputed before its reverse sweep.  Grid operators run in torch (FP64) on
the GPU when there is one; the particle work is the HIP library.
"""
import math

import numpy as np
import torch

from . import emg, fieldsolve, units


# ------------------------------------------------------------------ grid ops
# arrays are [nz, ny, nx] (scalars) or [nz, ny, nx, 3] (vectors); axis of
# component d (x, y, z) is 2 - d

def _eig(n, dev):
    k = torch.arange(n, dtype=torch.float64, device=dev)
    return (2.0 * torch.sin(math.pi * k / n)) ** 2


def poisson(rho):
    """P: -lap7 phi = rho on the periodic grid, zero mode dropped; P is symmetric."""
    nz, ny, nx = rho.shape
    dev = rho.device
    div = (_eig(nz, dev)[:, None, None] + _eig(ny, dev)[None, :, None]
           + _eig(nx, dev)[: nx // 2 + 1][None, None, :])
    div[0, 0, 0] = 1.0
    spec = torch.fft.rfftn(rho) / div
    spec[0, 0, 0] = 0.0
    return torch.fft.irfftn(spec, s=rho.shape)


def grad(phi):
    """G: staggered E_d(i + 1/2) = phi(i) - phi(i + 1) along each axis."""
    return torch.stack([phi - torch.roll(phi, -1, 2 - d) for d in range(3)], dim=-1)


def grad_t(e):
    """G^T."""
    return sum(e[..., d] - torch.roll(e[..., d], 1, 2 - d) for d in range(3))


def relocate(e):
    """R (grdcrt.F90:158-169 without e0): nodal E_d(i) = (E_d(i - 1/2) + E_d(i + 1/2)) / 2."""
    return torch.stack([0.5 * (e[..., d] + torch.roll(e[..., d], 1, 2 - d)) for d in range(3)],
                       dim=-1)


def relocate_t(w):
    """R^T."""
    return torch.stack([0.5 * (w[..., d] + torch.roll(w[..., d], -1, 2 - d)) for d in range(3)],
                       dim=-1)


def mode_weights(n, dev):
    """exp(-2 pi i n_x / N_x) as (cos, sin) broadcast over [nz, ny, nx]."""
    th = 2.0 * math.pi * torch.arange(n[0], dtype=torch.float64, device=dev) / n[0]
    return torch.cos(th)[None, None, :], torch.sin(th)[None, None, :]


# ------------------------------------------------------------------ window
class ESWindow:
    """A K-step electrostatic window with beam drift v_d and applied E0 as
    the differentiable parameters (EMSES-U values)."""

    def __init__(self, w: units.Workload, steps=None, checkpoint=10, electrons=(0, 1),
                 beam=1, count_scale=1.0, device=None):
        self.w = w
        self.su = su = units.to_sunits(w)
        self.n = tuple(w.n)
        self.K = int(steps or w.steps)
        self.C = int(checkpoint)
        if self.K % self.C:
            raise ValueError("steps must be a multiple of the checkpoint interval")
        if self.K // self.C + self.C > emg.EMG_STATE_SLOTS:
            raise ValueError("too many saved states for one context")
        # (torch has to see the device before libemgpu initialises it in this
        # process; this call comes before the context is created)
        self.dev = torch.device(device or ("cuda" if torch.cuda.is_available() else "cpu"))
        self.beam = beam
        self.electrons = electrons
        species = units.host_species(w)
        for s in species:
            s["npin"] = int(round(s["npin"] * count_scale))
            s["vdx"] = 0.0           # the drift is a parameter, added below
        self.counts = np.array([s["npin"] for s in species], dtype=np.int64)
        p = emg.host_inipcl(species, self.n, su.t, w.seed)
        units.scale_to_s(p, su)
        p["pid"] = np.arange(len(p))
        self.p0 = p
        self.beam_slice = slice(int(self.counts[:beam].sum()), int(self.counts[:beam + 1].sum()))
        mob = [k for k, s in enumerate(w.species) if s.mobile]
        self.q = su.q[mob]
        self.rhobk = fieldsolve.background(self.q, self.counts, self.n)
        self.ctx = emg.Context(self.n, self.q, su.qm[mob], su.rm[mob], self.counts,
                               precision=emg.PLAIN_FP64, e0=(0.0, 0.0, 0.0), b0=su.b0)
        cos, sin = mode_weights(self.n, self.dev)
        self.cos, self.sin = cos, sin
        self.ustep = 2
        self.last = None

    def close(self):
        self.ctx.close()

    # ---- forward pieces
    def _t(self, a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(self.dev)

    def _fields(self, e0_s):
        """Deposit, solve and upload the fields of the current state; returns
        the staggered e (torch) for reuse in the reverse sweep."""
        self.ctx.deposit_rho(0)
        rho = self._t(self.ctx.rho() + self.rhobk)
        e = grad(poisson(rho))
        e[..., 0] += e0_s
        self.ctx.set_fields(fieldsolve.eb_halo(e.cpu().numpy()))
        return e

    def _loss_and_density_adjoint(self):
        """L and dL/d rho_e at the current state."""
        self.ctx.deposit_rho(2)
        rho_e = sum(self._t(self.ctx.rho(species=s)) for s in self.electrons)
        fr = (rho_e * self.cos).sum()
        fi = -(rho_e * self.sin).sum()
        loss = float(fr * fr + fi * fi)
        lam_rho = 2.0 * (fr * self.cos - fi * self.sin) * torch.ones_like(rho_e)
        return loss, lam_rho

    def _start(self, vd):
        p = self.p0.copy()
        p["vx"][self.beam_slice] += vd * self.su.renv
        self.ctx.upload(p, self.counts)

    def forward(self, vd, e0):
        """L(v_d, E0) over K steps; saves the checkpoints."""
        e0_s = e0 * self.su.rene
        self._start(vd)
        for k in range(self.K):
            if k % self.C == 0:
                self.ctx.state_save(k // self.C)
            self._fields(e0_s)
            self.ctx.push_es(self.ustep)
        loss, lam_rho = self._loss_and_density_adjoint()
        self.last = (vd, e0, lam_rho)
        return loss

    def backward(self):
        """(dL/dv_d, dL/dE0) at the last forward point, by the reverse sweep."""
        vd, e0, lam_rho = self.last
        e0_s = e0 * self.su.rene
        nwin = self.K // self.C
        self.ctx.adj_reset()
        qfac = np.zeros(len(self.q))
        for s in self.electrons:
            qfac[s] = self.q[s]
        self.ctx.adj_add_density(lam_rho.cpu().numpy(), qfac)   # state is x_K here
        lam_e0 = 0.0
        for wdx in range(nwin - 1, -1, -1):
            # recompute the window's states (slots nwin .. nwin + C - 1) and fields
            self.ctx.state_load(wdx)
            fields = []
            for j in range(self.C):
                self.ctx.state_save(nwin + j)
                fields.append(self._fields(e0_s))
                self.ctx.push_es(self.ustep)
            for j in range(self.C - 1, -1, -1):
                self.ctx.state_load(nwin + j)
                self.ctx.set_fields(fieldsolve.eb_halo(fields[j].cpu().numpy()))
                self.ctx.adj_step(self.ustep)
                lam_w = self._t(self.ctx.adj_field())
                lam_e0 += float(lam_w[..., 0].sum())
                lam_phi = grad_t(relocate_t(lam_w))
                lam_rho = poisson(lam_phi)
                self.ctx.adj_add_density(lam_rho.cpu().numpy(), self.q)
        sums = self.ctx.adj_velocity_sums()
        g_vd = float(sums[self.beam, 0]) * self.su.renv
        g_e0 = lam_e0 * self.su.rene
        return g_vd, g_e0


class WindowLoss(torch.autograd.Function):
    """L = window(v_d, E0) with the hand-written reverse sweep as its backward."""

    @staticmethod
    def forward(ctx, vd, e0, window):
        loss = window.forward(float(vd), float(e0))
        ctx.window = window
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        g_vd, g_e0 = ctx.window.backward()
        g = float(grad_out)
        return (torch.tensor(g * g_vd, dtype=torch.float64),
                torch.tensor(g * g_e0, dtype=torch.float64), None)


def window_loss(vd, e0, window):
    return WindowLoss.apply(vd, e0, window)
