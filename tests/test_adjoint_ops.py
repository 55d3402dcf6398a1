"""Grid operators of the adjoint window (ctca_amd/adjoint.py): each forward
operator equals the field solve it restates, and each transpose passes the
dot-product test <A u, w> = <u, A^T w>; the Poisson operator is symmetric."""
import numpy as np
import pytest
import torch

import oracle
from ctca_amd import adjoint, fieldsolve

N = (8, 6, 10)


def rnd(shape, seed):
    return torch.from_numpy(np.random.default_rng(seed).normal(size=shape))


def dot(a, b):
    return float((a * b).sum())


def test_forward_operators_match_the_field_solve():
    phi = rnd((N[2], N[1], N[0]), 1)
    e = adjoint.grad(phi)
    assert np.allclose(e.numpy(), fieldsolve.efield(phi.numpy()), atol=0, rtol=0)
    rho = rnd((N[2], N[1], N[0]), 2)
    rho -= rho.mean()
    assert np.allclose(adjoint.poisson(rho).numpy(), fieldsolve.poisson(rho.numpy()), atol=1e-12)
    # relocation: the E part of the oracle's grdcrt (nodes 0..n-1)
    eb = fieldsolve.eb_halo(e.numpy())
    wrk = oracle.wrk_nodes(oracle.grdcrt(eb, N, (0, 0, 0), (0, 0, 0)))[:-1, :-1, :-1, :3]
    assert np.allclose(adjoint.relocate(e).numpy(), wrk, atol=1e-14)


@pytest.mark.parametrize("op,op_t,shape_in,shape_out", [
    (adjoint.grad, adjoint.grad_t, (N[2], N[1], N[0]), (N[2], N[1], N[0], 3)),
    (adjoint.relocate, adjoint.relocate_t, (N[2], N[1], N[0], 3), (N[2], N[1], N[0], 3)),
    (adjoint.poisson, adjoint.poisson, (N[2], N[1], N[0]), (N[2], N[1], N[0])),
])
def test_transposes(op, op_t, shape_in, shape_out):
    u = rnd(shape_in, 3)
    w = rnd(shape_out, 4)
    assert dot(op(u), w) == pytest.approx(dot(u, op_t(w)), rel=1e-12, abs=1e-12)


def test_mode_weights_pick_the_first_x_mode():
    cos, sin = adjoint.mode_weights(N, torch.device("cpu"))
    x = torch.arange(N[0], dtype=torch.float64)
    rho = torch.cos(2 * np.pi * x / N[0])[None, None, :].expand(N[2], N[1], N[0])
    fr = float((rho * cos).sum())
    assert fr == pytest.approx(0.5 * N[0] * N[1] * N[2])
    assert float((rho * sin).sum()) == pytest.approx(0.0, abs=1e-9)
