"""reinitializeMedium3d (smoothed-sign reinitialization, fused ENO3
right-hand side + TVD "RK2" stages, all-face BC) on the GPU against the
oracle's restatement with the ENO3 coefficient 1/6, over the whole ghost box:
bit-exact in FP64 and FP32 (both evaluate the reference's expression order
with IEEE division and square root; reference:
src/reinitialization/reinitialize_top.c:205-295,
src/reinitialization/qss_reinitialization3d.f:39-470)."""
import ctypes as C

import numpy as np
import pytest

from lsmpqs_amd import hiplib
from lsmpqs_amd.abi import make_grid, default_options
import oracle_py as op
from lsm_cases import case


def fields(dt):
    g, mask, phi0 = case((22, 19, 17), dt)
    rough = np.random.default_rng(5).random(phi0.shape) * 4 - 2
    return g, [np.ascontiguousarray(phi0.astype(dt)), np.ascontiguousarray(rough.astype(dt))]


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("tmax_r", [0.4, 1.6])
def test_reinit_medium_matches_oracle(dt, tmax_r):
    print(f"case: reinitializeMedium3d {np.dtype(dt).name} tmax_r={tmax_r}", flush=True)
    lib = hiplib.load(dt)
    g, fs = fields(dt)
    o = default_options(dt)
    o.tmax_r = tmax_r
    R = C.c_double if dt == np.float64 else C.c_float
    for f in fs:
        a, b = f.copy(), f.copy()
        lib.reinitializeMedium3d(a.ctypes.data_as(C.c_void_p), C.byref(g), C.byref(o))
        hiplib.check(lib)
        op.oracle_lib(dt).orc_reinitializeMedium3d(b.ctypes.data_as(C.c_void_p), C.byref(g),
                                                   C.byref(o), R(1.0 / 6.0))
        assert not np.array_equal(a, f)             # it did something
        assert np.array_equal(a, b), float(np.abs(a - b).max())
