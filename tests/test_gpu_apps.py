"""The pre- and post-processing computations on the GPU (lsmpqs_mask_from_image,
lsmpqs_init_from_mask, lsmpqs_diff_steps) against the oracle's restatement of
the reference programs (createMaskUBC.c, createInitFromMask.c,
makeDiffSteps.c; the oracle is pinned against those programs in
tests/test_oracle_apps.py): bit-exact in FP64 and FP32 for the sphere and the
lens, bit-exact in FP64 for the reinitialized mask."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as op
from lsm_cases import void_api
from lsmpqs_amd import hiplib
from lsmpqs_amd.abi import types_for
from test_oracle_apps import segmented, step_pair

pytestmark = pytest.mark.gpu


def lib_for(dtype):
    lib = void_api(hiplib.load(dtype))
    lib.lsmpqs_mask_from_image.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                           C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
    return lib


def gpu_mask(lib, img, dtype):
    nz, ny, nx = img.shape
    o = C.c_void_p(lib.createOptionsDefault())
    g, m = C.c_void_p(), C.c_void_p()
    rc = lib.lsmpqs_mask_from_image(img.ctypes.data, nx, ny, nz, o, C.byref(g), C.byref(m))
    assert rc == 0, hiplib.last_error(lib)
    n = (nx + 6) * (ny + 6) * (nz + 6)
    R = C.c_double if dtype == np.float64 else C.c_float
    mask = np.ctypeslib.as_array(C.cast(m, C.POINTER(R)), (n,)).copy()
    C.CDLL("libc.so.6").free(m)
    return o, g, mask


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(18, 20, 22), (70, 40, 96)])
def test_mask_from_image(dtype, shape):
    print(f"case: mask_from_image {np.dtype(dtype).name} {shape}", flush=True)
    lib = lib_for(dtype)
    img = segmented(shape, 4)
    o, g, mask = gpu_mask(lib, img, dtype)
    want = np.zeros_like(mask)
    nz, ny, nx = shape
    op.oracle_lib(dtype).orc_mask_from_image(img.ctypes.data_as(C.c_void_p), nx, ny, nz, g, o,
                                              want.ctypes.data_as(C.c_void_p))
    lib.destroyOptions(o)
    lib.destroyGrid(g)
    if dtype == np.float64:
        assert np.array_equal(mask, want)
    else:
        assert np.abs(mask - want).max() <= 1e-6
    # the reinitialization keeps every fill voxel on its side of the grain surface
    m3 = mask.reshape(nz + 6, ny + 6, nx + 6)[3:-3, 3:-3, 3:-3]
    assert np.array_equal(m3 > 0, img == 255)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("center", [0, 1])
def test_init_from_mask(dtype, center):
    print(f"case: init_from_mask {np.dtype(dtype).name} center {center}", flush=True)
    lib = lib_for(dtype)
    o, g, mask = gpu_mask(lib, segmented((30, 26, 34), 5), dtype)
    _, Options, _ = types_for(dtype)
    C.cast(o, C.POINTER(Options)).contents.center_inlet = center
    got, want = np.zeros_like(mask), np.zeros_like(mask)
    assert lib.lsmpqs_init_from_mask(mask.ctypes.data_as(C.c_void_p), g, o,
                                     got.ctypes.data_as(C.c_void_p)) == 0, hiplib.last_error(lib)
    op.oracle_lib(dtype).orc_init_from_mask(mask.ctypes.data_as(C.c_void_p), g, o,
                                             want.ctypes.data_as(C.c_void_p))
    lib.destroyOptions(o)
    lib.destroyGrid(g)
    assert np.array_equal(got, want)
    assert (got < 0).any() and (got > 0).any()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("percolating", [True, False])
@pytest.mark.parametrize("shape", [(18, 20, 22), (80, 72, 150)])
def test_diff_steps(dtype, percolating, shape):
    print(f"case: diff_steps {np.dtype(dtype).name} {shape} percolating={percolating}", flush=True)
    lib = lib_for(dtype)
    o, g, _ = gpu_mask(lib, segmented(shape, 6), dtype)
    gshape = tuple(s + 6 for s in shape)
    a, b = step_pair(gshape, dtype, 7, percolating)
    got, got_c = np.zeros(a.size, dtype), np.full(a.size, 7, dtype)
    want, want_c = np.zeros(a.size, dtype), np.full(a.size, 7, dtype)
    rc = lib.lsmpqs_diff_steps(a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p), g,
                               got.ctypes.data_as(C.c_void_p), got_c.ctypes.data_as(C.c_void_p))
    perc = op.oracle_lib(dtype).orc_diff_steps(a.ctypes.data_as(C.c_void_p),
                                               b.ctypes.data_as(C.c_void_p), g,
                                               want.ctypes.data_as(C.c_void_p),
                                               want_c.ctypes.data_as(C.c_void_p))
    lib.destroyOptions(o)
    lib.destroyGrid(g)
    assert rc == perc == (1 if percolating else 0)
    assert np.array_equal(got, want)
    assert np.array_equal(got_c, want_c)
    if percolating:
        assert (got_c != got).any()
