"""emg_get_rho_noalloc: the rho read-out staged through the library's static
device array instead of a per-call allocation.  Called through the C ABI,
its output equals emg_get_rho's bit for bit on two calls in a row (the second
a smaller box, which reads what the first left in the shared array only
where it writes again), in both forms, ghost planes included; a box beyond
the array's 2^24 values is refused with EMG_ERR_CAPACITY."""
import ctypes as C

import numpy as np
import pytest

from ctca_amd import emg

pytestmark = pytest.mark.gpu


def read(ctx, name, lo, hi):
    out = np.full((hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1), np.nan)
    rc = getattr(emg.lib(), name)(ctx.h, out.ctypes.data_as(C.c_void_p),
                                  (C.c_int32 * 3)(*lo), (C.c_int32 * 3)(*hi))
    return rc, out


@pytest.mark.parametrize("precision", [emg.PLAIN_FP64, emg.FP32_FIXED], ids=["fp64", "fp32"])
def test_noalloc_readout_equals_get_rho(precision):
    n = (16, 12, 8)
    counts = (20000, 12000)
    spec = [dict(npin=counts[0], peth=0.08, path=0.08), dict(npin=counts[1], peth=0.01, path=0.01)]
    p = emg.host_inipcl(spec, n, np.eye(3), 10)
    with emg.Context(n, (0.002, -0.002), (1.0, -0.01), (0.002, 0.2), counts,
                     precision=precision) as ctx:
        ctx.upload(p, counts)
        ctx.deposit_rho(0)
        big = ((-1, -1, -1), tuple(v + 1 for v in n))
        small = ((2, 1, 0), (9, 5, 3))
        for lo, hi in (big, small):
            rc, a = read(ctx, "emg_get_rho_noalloc", lo, hi)
            assert rc == 0, emg.last_error()
            rc, b = read(ctx, "emg_get_rho", lo, hi)
            assert rc == 0, emg.last_error()
            assert np.array_equal(a, b), (lo, hi)
            assert np.abs(a).max() > 0.0
        # beyond 2^24 values: refused, nothing written
        lo, hi = (0, 0, 0), (256, 256, 256)
        out = np.zeros(1)
        rc = emg.lib().emg_get_rho_noalloc(ctx.h, out.ctypes.data_as(C.c_void_p),
                                           (C.c_int32 * 3)(*lo), (C.c_int32 * 3)(*hi))
        assert rc == -7 and "2^24" in emg.last_error()
