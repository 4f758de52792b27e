"""Input errors of the 3D entry points are reported through
lsmpqs_hip_last_error (and the return code where the entry point has one)
before any device call, so these run without a GPU: a box below the drivers'
minimum (3 ghost layers and 4 fill voxels per axis), a batch without boxes or
with one record given to two boxes, an adjoint run without steps or with an
unknown LSMPQS_ADJOINT_STORE mode."""
import ctypes as C

import numpy as np
import pytest

from lsm_cases import options, DEVNULL
from lsmpqs_amd import hiplib
from lsmpqs_amd.abi import make_grid, Arrays
from lsmpqs_amd.adjoint import AdjointArgs, _lib as adjoint_lib


def small_box(dt, fill):
    g = make_grid((fill, fill, fill), dtype=dt)
    A = Arrays(g, dt)
    shape = tuple(g.grid_dims_ghostbox[i] for i in (2, 1, 0))
    A["mask"] = -np.ones(shape, dtype=dt)
    A["phi"] = np.ones(shape, dtype=dt)
    return g, A


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_box_below_minimum_is_rejected(dt):
    lib = hiplib.load(dt)
    g, A = small_box(dt, 3)
    o = options(dt)
    lib.drain_model3d(C.byref(o), A.ptr(), C.byref(g), DEVNULL)
    assert "4 fill voxels per axis" in hiplib.last_error(lib)
    lib.lsmpqs_release(A.ptr())


def test_batch_without_boxes_or_with_shared_records_is_rejected():
    dt = np.float32
    lib = hiplib.load(dt)
    st = hiplib.BatchStats()
    rc = lib.lsmpqs_batch_model3d(0, None, None, None, None, 1, C.byref(st))
    assert rc == -2 and "n >= 1" in hiplib.last_error(lib)
    g, A = small_box(dt, 8)
    o = options(dt)
    opts = (C.c_void_p * 2)(C.addressof(o), C.addressof(o))
    recs = (C.c_void_p * 2)(C.addressof(A.rec), C.addressof(A.rec))
    grids = (C.c_void_p * 2)(C.addressof(g), C.addressof(g))
    rc = lib.lsmpqs_batch_model3d(2, opts, recs, grids, None, 1, C.byref(st))
    assert rc == -2 and "its own QSS_DataArrays and Options" in hiplib.last_error(lib)


def test_adjoint_without_steps_is_rejected():
    dt = np.float64
    lib = adjoint_lib(dt)
    g, A = small_box(dt, 8)
    o = options(dt)
    theta = np.zeros(g.num_gridpts, dtype=dt)
    a = AdjointArgs()
    a.nsteps, a.checkpoint_every = 0, 1
    rc = lib.lsmpqs_adjoint_run(C.addressof(o), C.addressof(A.rec), C.addressof(g),
                                theta.ctypes.data, None, C.byref(a))
    assert rc == -2 and "nsteps and checkpoint_every" in hiplib.last_error(lib)


def test_adjoint_with_unknown_store_mode_is_rejected(monkeypatch):
    dt = np.float64
    lib = adjoint_lib(dt)
    g, A = small_box(dt, 8)
    o = options(dt)
    theta = np.zeros(g.num_gridpts, dtype=dt)
    a = AdjointArgs()
    a.nsteps, a.checkpoint_every = 4, 2
    monkeypatch.setenv("LSMPQS_ADJOINT_STORE", "everything")
    rc = lib.lsmpqs_adjoint_run(C.addressof(o), C.addressof(A.rec), C.addressof(g),
                                theta.ctypes.data, None, C.byref(a))
    assert rc == -2 and "LSMPQS_ADJOINT_STORE must be" in hiplib.last_error(lib)
