"""ctypes access to the oracle (liboracle_*.so) and to the reference build
(_ref/libqssref_*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py.  The product library never uses it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_LIBS = {}


def _real(dtype):
    return C.c_double if np.dtype(dtype) == np.float64 else C.c_float


def _suffix(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def oracle_lib(dtype):
    key = ("oracle", _suffix(dtype))
    if key not in _LIBS:
        path = os.path.join(HERE, f"liboracle_{_suffix(dtype)}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run make in oracle/)")
        _LIBS[key] = C.CDLL(path)
    return _LIBS[key]


def ref_path(dtype):
    return os.path.join(HERE, "_ref", f"libqssref_{_suffix(dtype)}.so")


def have_ref(dtype=np.float64):
    return os.path.exists(ref_path(dtype))


def ref_lib(dtype):
    key = ("ref", _suffix(dtype))
    if key not in _LIBS:
        path = ref_path(dtype)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -f Makefile.ref in oracle/)")
        _LIBS[key] = C.CDLL(path)
    return _LIBS[key]


def pad_phi_bin(A, g):
    """Give A's phi_bin array readable room in front of it before the
    reference build labels it: the reference's unionCoords3d reads phi_bin
    at a neighbour's index before testing that the neighbour lies in the box
    (src/connectivity/connectivity.c:138-146), up to nx*ny + nx + 1 bytes
    before the array.  It discards those bytes; they only have to be
    mapped.  A is a lsmpqs_amd.abi.Arrays record; returns A."""
    nx, ny = g.grid_dims_ghostbox[0], g.grid_dims_ghostbox[1]
    pad = nx * ny + nx + 64
    n = A["phi_bin"].size
    buf = np.zeros(n + 2 * pad, dtype=np.uint8)
    A.a["phi_bin"] = buf[pad:pad + n].reshape(A["phi_bin"].shape)
    A.rec.phi_bin = A.a["phi_bin"].ctypes.data
    return A


def set_omp_threads(n):
    """Thread count the reference build's OpenMP team will use."""
    gomp = C.CDLL("libgomp.so.1")
    gomp.omp_set_num_threads(C.c_int(int(n)))


class Scalar:
    """A by-reference scalar argument whose value can be read back."""

    def __init__(self, value, ctype):
        self.c = ctype(value)

    @property
    def value(self):
        return self.c.value


def fcall(lib, name, dtype, *args):
    """Call a Fortran-ABI routine: every argument is passed by reference.
    numpy arrays pass their data pointer, Python ints become int*, Python
    floats become QSSLIB_REAL*, Scalar objects pass their storage."""
    R = _real(dtype)
    keep = []
    cargs = []
    for a in args:
        if isinstance(a, np.ndarray):
            cargs.append(C.c_void_p(a.ctypes.data))
        elif isinstance(a, Scalar):
            cargs.append(C.byref(a.c))
        elif isinstance(a, bool):
            raise TypeError("bool argument")
        elif isinstance(a, (int, np.integer)):
            v = C.c_int(int(a))
            keep.append(v)
            cargs.append(C.byref(v))
        elif isinstance(a, (float, np.floating)):
            v = R(float(a))
            keep.append(v)
            cargs.append(C.byref(v))
        else:
            cargs.append(a)
    fn = getattr(lib, name)
    fn.restype = None
    fn(*cargs)


def gb6(g):
    return (g.ilo_gb, g.ihi_gb, g.jlo_gb, g.jhi_gb, g.klo_gb, g.khi_gb)


def fb6(g):
    return (g.ilo_fb, g.ihi_fb, g.jlo_fb, g.jhi_fb, g.klo_fb, g.khi_fb)


def driver(lib, name, restype=None):
    fn = getattr(lib, name)
    fn.restype = restype
    return fn


class OrcStats(C.Structure):
    _fields_ = [("outer_steps", C.c_int), ("inner_steps", C.c_int), ("last_dt", C.c_double),
                ("satn", C.c_double), ("max_abs_err", C.c_double),
                ("stale_norm_voxels", C.c_int)]


def oracle_stats(dtype):
    s = OrcStats()
    lib = oracle_lib(dtype)
    lib.orc_get_stats(C.byref(s))
    return s
