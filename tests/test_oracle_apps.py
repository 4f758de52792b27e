"""Pins of the oracle's restatement of the pre- and post-processing programs
(orc_mask_from_image, orc_init_from_mask, orc_diff_steps) against the
reference's own programs, compiled from src/applications/createMaskUBC.c,
createInitFromMask.c and makeDiffSteps.c by oracle/Makefile.ref and run here
on seeded inputs; outputs are compared file against array."""
import ctypes as C
import gzip
import os
import subprocess

import numpy as np
import pytest

import oracle_py as op
from lsm_cases import void_api
from lsmpqs_amd import hiplib

APP = os.path.join(os.path.dirname(op.__file__), "_ref", "app_{}", "{}")


def app(dtype, name):
    return APP.format("f64" if dtype == np.float64 else "f32", name)


needs_apps = pytest.mark.skipif(not os.path.exists(app(np.float64, "createMaskUBC")),
                                reason="reference programs (oracle/_ref/app_*) not built")


def segmented(shape, seed, frac=0.45):
    """(nz, ny, nx) 0/255 image of smooth random grains, x fastest in memory."""
    from scipy.ndimage import gaussian_filter
    f = gaussian_filter(np.random.default_rng(seed).standard_normal(shape), 1.5)
    return np.where(f > np.quantile(f, 1 - frac), 255, 0).astype(np.uint8)


def read_array(path, dtype):
    raw = gzip.open(path).read() if path.endswith(".gz") else open(path, "rb").read()
    dims = np.frombuffer(raw[:12], dtype=np.int32)
    return dims, np.frombuffer(raw[12:], dtype=dtype).copy()


def grid_from(lib, path):
    return C.c_void_p(void_api(lib).readGridFromBinaryFile(path.encode()))


def run(cmd, cwd):
    # createInitFromMask frees an uninitialised pointer after writing its file;
    # the output is complete whatever its exit status
    return subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=600)


def mask_case(dtype, tmp_path, shape=(18, 20, 22), seed=4):
    img = segmented(shape, seed)
    (tmp_path / "img.raw").write_bytes(img.tobytes())
    nz, ny, nx = shape
    r = run([app(dtype, "createMaskUBC"), "img.raw", str(nx), str(ny), str(nz)], str(tmp_path))
    assert r.returncode == 0, r.stderr
    return img


@needs_apps
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_mask_from_image_matches_createMaskUBC(dtype, tmp_path):
    img = mask_case(dtype, tmp_path)
    nz, ny, nx = img.shape
    lib = hiplib.load(dtype)
    g = grid_from(lib, str(tmp_path / "grid.gz"))
    dims, want = read_array(str(tmp_path / "mask.gz"), dtype)
    o = C.c_void_p(void_api(lib).createOptionsDefault())
    got = np.zeros_like(want)
    op.oracle_lib(dtype).orc_mask_from_image(img.ctypes.data_as(C.c_void_p), nx, ny, nz, g, o,
                                              got.ctypes.data_as(C.c_void_p))
    lib.destroyOptions(o)
    lib.destroyGrid(g)
    assert np.abs(want).max() > 0
    if dtype == np.float64:
        assert np.array_equal(got, want)
    else:  # the reference's single-precision build mixes double literals into the reinit
        assert np.abs(got - want).max() <= 1e-6


@needs_apps
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("center", [0, 1])
def test_init_from_mask_matches_createInitFromMask(dtype, center, tmp_path):
    mask_case(dtype, tmp_path, seed=5)
    (tmp_path / "opts").write_text(f"center_inlet {center}\n")
    r = run([app(dtype, "createInitFromMask"), "mask.gz", "grid.gz", "opts"], str(tmp_path))
    assert (tmp_path / "data_init.gz").exists(), r.stdout + r.stderr
    lib = hiplib.load(dtype)
    g = grid_from(lib, str(tmp_path / "grid.gz"))
    _, mask = read_array(str(tmp_path / "mask.gz"), dtype)
    _, want = read_array(str(tmp_path / "data_init.gz"), dtype)
    o = C.c_void_p(void_api(lib).createOptionsFromInputFile(str(tmp_path / "opts").encode()))
    got = np.zeros_like(want)
    op.oracle_lib(dtype).orc_init_from_mask(mask.ctypes.data_as(C.c_void_p), g, o,
                                             got.ctypes.data_as(C.c_void_p))
    lib.destroyOptions(o)
    lib.destroyGrid(g)
    assert np.array_equal(got, want)
    assert (want < 0).any() and (want > 0).any()


def step_pair(shape, dtype, seed, percolating):
    """Two level sets over the ghost box: a drained state a and a later b."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.standard_normal(shape), 2.0)
    f /= f.std()
    a = (1.0 + 0.1 * f).astype(dtype)                       # nothing drained yet
    # the lens is b < dx/2: isolated blobs (about 4 % of the voxels), and with
    # `percolating` a tube through the box along x at mid y, z
    b = f + 1.75
    if percolating:
        nz, ny, nx = shape
        b[nz // 2 - 1:nz // 2 + 2, ny // 2 - 1:ny // 2 + 2, :] = -1.0
    b = b.astype(dtype)
    return np.ascontiguousarray(a), np.ascontiguousarray(b)


@needs_apps
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("percolating", [True, False])
def test_diff_steps_matches_makeDiffSteps(dtype, percolating, tmp_path):
    mask_case(dtype, tmp_path, seed=6)   # grid.gz of a 22 x 20 x 18 box
    lib = hiplib.load(dtype)
    g = grid_from(lib, str(tmp_path / "grid.gz"))
    dims, _ = read_array(str(tmp_path / "mask.gz"), dtype)
    shape = (int(dims[2]), int(dims[1]), int(dims[0]))
    a, b = step_pair(shape, dtype, 7, percolating)
    for k, arr in ((1, a), (2, b)):
        lib.writeDataArrayQSS(arr.ctypes.data_as(C.c_void_p), g,
                              str(tmp_path / f"data_step_{k}").encode(), 1)
    r = run([app(dtype, "makeDiffSteps"), "1", "2"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    want_diff = np.fromfile(str(tmp_path / "diff_steps_1_2.raw"), dtype=dtype)
    conn_file = tmp_path / "diff_steps_1_2_connected_only.raw"
    diff = np.zeros(a.size, dtype)
    conn = np.zeros(a.size, dtype)
    perc = op.oracle_lib(dtype).orc_diff_steps(a.ctypes.data_as(C.c_void_p),
                                               b.ctypes.data_as(C.c_void_p), g,
                                               diff.ctypes.data_as(C.c_void_p),
                                               conn.ctypes.data_as(C.c_void_p))
    lib.destroyGrid(g)
    assert np.array_equal(diff, want_diff)
    assert bool(perc) == percolating == conn_file.exists()
    assert ("No percolation" in r.stdout) == (not percolating)
    if percolating:
        assert np.array_equal(conn, np.fromfile(str(conn_file), dtype=dtype))
        assert (conn != diff).any()
