"""File-level entry points without a GPU, checked against the reference build
(oracle/_ref, compiled from the reference's own sources):

* createGridSetDx / createGridSetDxDyDz / createGridSetGridDims give the
  reference's Grid record, field for field, for every accuracy level
  (reference: src/geometry/qss_grid.c:51-250,789-1036);
* the grid binary file is byte-identical both ways, plain and .gz
  (qss_grid.c:614-787);
* printGrid / printOptions write the reference's text (qss_grid.c:323-440,
  src/utilities/qss_options.c:497-541);
* createOptionsFromInputFile reads an option file -- every key, invalid flag
  values, a key with a capital letter, a line longer than the 79-character
  read -- into the same Options record (qss_options.c:118-494).
"""
import ctypes as C
import gzip
import os
import shutil
import sys

import numpy as np
import pytest

from lsmpqs_amd import hiplib
from lsmpqs_amd.abi import types_for

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle_py as op  # noqa: E402

DTYPES = [np.float32, np.float64]
LIBC = C.CDLL("libc.so.6")
LIBC.fopen.restype = C.c_void_p
LIBC.fopen.argtypes = [C.c_char_p, C.c_char_p]
LIBC.fclose.argtypes = [C.c_void_p]
LIBC.fflush.argtypes = [C.c_void_p]


def libs(dtype):
    if not op.have_ref(dtype):
        pytest.skip("reference build (oracle/_ref) not available")
    ours, ref = hiplib.load(dtype), op.ref_lib(dtype)
    Grid, Options, _ = types_for(dtype)
    R = C.c_double if np.dtype(dtype) == np.float64 else C.c_float
    for lib in (ours, ref):
        lib.createGridSetDx.restype = C.POINTER(Grid)
        lib.createGridSetDx.argtypes = [C.c_int, R, C.POINTER(R), C.POINTER(R), C.c_int]
        lib.createGridSetDxDyDz.restype = C.POINTER(Grid)
        lib.createGridSetDxDyDz.argtypes = [C.c_int, R, R, R, C.POINTER(R), C.POINTER(R), C.c_int]
        lib.createGridSetGridDims.restype = C.POINTER(Grid)
        lib.createGridSetGridDims.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(R),
                                              C.POINTER(R), C.c_int]
        lib.readGridFromBinaryFile.restype = C.POINTER(Grid)
        lib.readGridFromBinaryFile.argtypes = [C.c_char_p]
        lib.writeGridToBinaryFile.argtypes = [C.POINTER(Grid), C.c_char_p, C.c_int]
        lib.destroyGrid.argtypes = [C.POINTER(Grid)]
        lib.printGrid.argtypes = [C.POINTER(Grid), C.c_void_p]
        lib.createOptionsFromInputFile.restype = C.POINTER(Options)
        lib.createOptionsFromInputFile.argtypes = [C.c_char_p]
        lib.printOptions.argtypes = [C.POINTER(Options), C.c_void_p]
        lib.destroyOptions.argtypes = [C.POINTER(Options)]
    return ours, ref, Grid, Options, R


def record_bytes(ptr, cls):
    return bytes(C.string_at(C.addressof(ptr.contents), C.sizeof(cls)))


def grid_cases(R):
    lo = (R * 3)(0.0, 0.5, -1.0)
    hi = (R * 3)(2.0, 1.7, 0.6)
    return lo, hi


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("accuracy", [0, 1, 2, 3])
@pytest.mark.parametrize("num_dims", [3, 2])
def test_grid_constructors_match_reference(dtype, accuracy, num_dims):
    ours, ref, Grid, _, R = libs(dtype)
    lo, hi = grid_cases(R)
    dims = (C.c_int * 3)(40, 24, 17)
    for call in (lambda L: L.createGridSetDx(num_dims, 0.04, lo, hi, accuracy),
                 lambda L: L.createGridSetDxDyDz(num_dims, 0.04, 0.05, 0.1, lo, hi, accuracy),
                 lambda L: L.createGridSetGridDims(num_dims, dims, lo, hi, accuracy)):
        a, b = call(ours), call(ref)
        assert record_bytes(a, Grid) == record_bytes(b, Grid)
        ours.destroyGrid(a)
        ref.destroyGrid(b)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("gz", [0, 1])
def test_grid_file_roundtrip_with_reference(dtype, gz, tmp_path):
    ours, ref, Grid, _, R = libs(dtype)
    lo, hi = grid_cases(R)
    g = ref.createGridSetDxDyDz(3, 0.04, 0.05, 0.1, lo, hi, 3)
    names = {}
    for tag, L in (("ours", ours), ("ref", ref)):
        name = str(tmp_path / f"grid_{tag}")
        L.writeGridToBinaryFile(g, name.encode(), gz)
        names[tag] = name + (".gz" if gz else "")
        assert os.path.exists(names[tag])
    raw = {t: (gzip.open(n).read() if gz else open(n, "rb").read()) for t, n in names.items()}
    assert raw["ours"] == raw["ref"]
    for writer in ("ours", "ref"):
        for L in (ours, ref):
            back = L.readGridFromBinaryFile(names[writer].encode())
            assert record_bytes(back, Grid) == record_bytes(g, Grid)
            L.destroyGrid(back)
            assert os.path.exists(names[writer])        # compressed again after reading
    ref.destroyGrid(g)


def printed(lib, fn, rec, path):
    fp = LIBC.fopen(str(path).encode(), b"w")
    getattr(lib, fn)(rec, fp)
    LIBC.fclose(fp)
    return open(path).read()


OPTION_FILE = """\
outfile run_log.txt
dx 0.02
tmax 42.5
tplot 0.75
accuracy very_high
a 0.031
amax 0.9
amin 0.011
b 0.045
cmax 12.5
checkpoint 1
check_connectivity 0
center_inlet 1
do_reinit 0
do_mask 1
dc 0.2
eps_coefficient 1.25
err_check_zone 0.6
eps_stop 0.0001
extrapol_id 2
init_step 3
mask_reinit 0
narrow_band 1
overlap 0.3
reservoir_inlet 7
save_data 0
stop_touch 0
subcell_fix 0
space_accur 5
time_accur 2
theta 30
vol_frac_max 0.95
vol_frac_min 0.02
use_var_theta 0
use_satn_stop 0
print_details 0
   b    0.05
Accuracy high
do_mask 2
""" + "tmax " + "7" * 90 + "\n"


@pytest.mark.parametrize("dtype", DTYPES)
def test_option_file_and_printouts_match_reference(dtype, tmp_path):
    ours, ref, Grid, Options, R = libs(dtype)
    path = tmp_path / "opts.in"
    path.write_text(OPTION_FILE)
    a = ours.createOptionsFromInputFile(str(path).encode())
    b = ref.createOptionsFromInputFile(str(path).encode())
    skip = {"extrapol_func", "space_deriv_func", "reinit_func"}   # code addresses
    for name, _ in Options._fields_:
        if name in skip:
            continue
        va, vb = getattr(a.contents, name), getattr(b.contents, name)
        if isinstance(va, C.Array):
            va, vb = bytes(va), bytes(vb)
        assert va == vb, name
    assert bool(a.contents.extrapol_func) and bool(b.contents.extrapol_func)
    assert a.contents.tmax == b.contents.tmax            # the 90-digit line is split
    assert printed(ours, "printOptions", a, tmp_path / "o1") == \
        printed(ref, "printOptions", b, tmp_path / "o2")
    lo, hi = grid_cases(R)
    for nd in (3, 2):
        g = ref.createGridSetDx(nd, 0.04, lo, hi, 1)
        assert printed(ours, "printGrid", g, tmp_path / "g1") == \
            printed(ref, "printGrid", g, tmp_path / "g2")
        ref.destroyGrid(g)
    ours.destroyOptions(a)
    ref.destroyOptions(b)


@pytest.mark.parametrize("dtype", DTYPES)
def test_outfile_sets_path(dtype, tmp_path):
    ours, _, _, _, _ = libs(dtype)
    path = tmp_path / "opts.in"
    path.write_text("outfile results/deep/log.txt\n")
    o = ours.createOptionsFromInputFile(str(path).encode())
    assert o.contents.outfile == b"results/deep/log.txt"
    assert o.contents.path == b"results/deep/"
    ours.destroyOptions(o)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dims", [(7, 5, 3), (262, 262, 40)])
def test_gzip_array_files_read_back(dtype, dims, tmp_path):
    """writeDataArrayQSS(..., 1) deflates on several threads into a gzip file of
    independent members (more than one above 8 MB); gunzip, Python's gzip and
    readDataArrayQSS read back the reference layout (int32 dims, values) and
    no uncompressed file is left, as after the reference's write + gzip -f
    (reference: src/utilities/qss_data_arrays.c:214-228)."""
    lib = hiplib.load(dtype)
    Grid, _, _ = types_for(dtype)
    g = Grid()
    for a in range(3):
        g.grid_dims_ghostbox[a] = dims[a]
    g.num_gridpts = int(np.prod(dims))
    data = np.random.default_rng(5).standard_normal(g.num_gridpts).astype(dtype)
    name = str(tmp_path / "arr")
    lib.writeDataArrayQSS(data.ctypes.data_as(C.c_void_p), C.byref(g), name.encode(), 1)
    assert os.path.exists(name + ".gz") and not os.path.exists(name)
    raw = gzip.open(name + ".gz").read()
    assert np.array_equal(np.frombuffer(raw[:12], np.int32), np.array(dims, np.int32))
    assert np.array_equal(np.frombuffer(raw[12:], dtype), data)
    back_dims = (C.c_int * 3)()
    lib.readDataArrayQSS.restype = C.c_void_p
    ptr = lib.readDataArrayQSS(back_dims, (name + ".gz").encode())
    assert ptr and list(back_dims) == list(dims)
    got = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_float if dtype == np.float32
                                                      else C.c_double)), (g.num_gridpts,))
    assert np.array_equal(got, data)
    C.CDLL("libc.so.6").free(C.c_void_p(ptr))
