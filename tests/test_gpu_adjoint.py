"""The adjoint window: gradients of the loss with respect to the beam drift
and the applied field from the hand-written reverse sweep (libemgpu
emg_adj_*) against central finite differences of the FP64 forward, through
the torch.autograd.Function as well."""
import numpy as np
import pytest
import torch

from ctca_amd import adjoint, units

pytestmark = pytest.mark.gpu


def small_workload():
    nc = 16 ** 3
    return units.Workload("adjoint16_beam", (16, 16, 16), 0.1, 6, [
        units.Species("electron", -1.0, 1.0, 16 * nc, 1.0),
        units.Species("beam", -1.0, float(np.sqrt(0.05)), int(0.05 * 16 * nc), 1.0,
                      drift=(3.0, 0.0, 0.0)),
    ], e0=(0.01, 0.0, 0.0), seed=41)


def test_adjoint_gradients_match_central_differences():
    win = adjoint.ESWindow(small_workload(), steps=6, checkpoint=3)
    try:
        vd, e0 = 3.0, 0.01
        loss = win.forward(vd, e0)
        g_vd, g_e0 = win.backward()
        hv, he = 3e-6, 1e-6
        fd_vd = (win.forward(vd + hv, e0) - win.forward(vd - hv, e0)) / (2 * hv)
        fd_e0 = (win.forward(vd, e0 + he) - win.forward(vd, e0 - he)) / (2 * he)
        # the same gradients through torch.autograd
        tv = torch.tensor(vd, dtype=torch.float64, requires_grad=True)
        te = torch.tensor(e0, dtype=torch.float64, requires_grad=True)
        L = adjoint.window_loss(tv, te, win)
        L.backward()
    finally:
        win.close()
    assert loss > 0
    print("L", loss, "dL/dvd", g_vd, fd_vd, "dL/dE0", g_e0, fd_e0)
    assert abs(g_vd - fd_vd) <= 1e-4 * abs(fd_vd)
    assert abs(g_e0 - fd_e0) <= 1e-4 * abs(fd_e0)
    assert float(L) == pytest.approx(loss, rel=1e-12)
    assert float(tv.grad) == pytest.approx(g_vd, rel=1e-12)
    assert float(te.grad) == pytest.approx(g_e0, rel=1e-12)
