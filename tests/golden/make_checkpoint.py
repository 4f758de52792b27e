"""Writes tests/golden/checkpoint_f64.npz by running the REFERENCE build
(oracle/_ref/libqssref_f64.so, compiled from /root/reference by
oracle/Makefile.ref) with options->checkpoint = 1.

    python tests/golden/make_checkpoint.py

With the option set, the reference's outer loop writes checkpoint_phi,
checkpoint_phi_prev, mask_w and mask_nw (gzip, working directory) after
every outer step (src/level_set_evolution/constCurvModel3d.c:301-310 and
:576-585) and logs "connectivity time" (:314, :589).  Two seeded 16^3
drainage runs of one curvature step, 4 outer steps each (tmax = 4 tplot):
  novar_*  ENO2 + "RK2", theta 0 (constCurvModel3dNoVar)
  var_*    WENO5 + TVD-RK3, theta 30 deg (constCurvModel3d)
The fixture holds the four files' arrays after the run (ghost box, FP64,
header dims checked), the number of "connectivity time" lines and of outer
steps in the log (data only).
"""
import ctypes as C
import gzip
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "lsmpqs-1.0_amd"), os.path.join(ROOT, "oracle"),
                os.path.join(ROOT, "tests")]

import oracle_py as op  # noqa: E402
from lsm_cases import arrays, case, options  # noqa: E402

DT = np.float64
N = 16
THREADS = 2
FILES = ("checkpoint_phi", "checkpoint_phi_prev", "mask_w", "mask_nw")
CASES = {"novar": dict(space=2, time=2, theta_deg=0), "var": dict(space=5, time=3, theta_deg=30)}
OUT = os.path.join(HERE, "checkpoint_f64.npz")


def case_options(name, dtype):
    o = options(dtype, steps=1, **CASES[name])
    o.checkpoint = 1
    o.tmax = 4 * o.tplot
    return o


def read_array(path, dtype, shape):
    raw = gzip.open(path, "rb").read()
    dims = np.frombuffer(raw[:12], dtype=np.int32)
    assert tuple(dims[::-1]) == tuple(shape), (dims, shape)
    return np.frombuffer(raw[12:], dtype=dtype).reshape(shape).copy()


def run(lib, name, dtype, workdir, extra=()):
    """Run drain_model3d of case `name` in workdir; returns (arrays of the
    four checkpoint files, connectivity-time lines, outer steps)."""
    g, mask, phi0 = case(N, dtype)
    A = op.pad_phi_bin(arrays(g, mask, phi0, dtype), g)
    o = case_options(name, dtype)
    libc = C.CDLL("libc.so.6")
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    log = os.path.join(workdir, "run.log")
    cwd = os.getcwd()
    os.chdir(workdir)
    try:
        fp = C.c_void_p(libc.fopen(log.encode(), b"w"))
        lib.drain_model3d(C.byref(o), A.ptr(), C.byref(g), fp, *extra)
        libc.fclose(fp)
    finally:
        os.chdir(cwd)
    files = {f: read_array(os.path.join(workdir, f + ".gz"), dtype, A.shape) for f in FILES}
    text = open(log).read()
    return files, text.count("connectivity time = "), text.count("nw phase satn"), A


def main():
    lib = op.ref_lib(DT)
    op.set_omp_threads(THREADS)
    out = {"threads": THREADS}
    for name in CASES:
        work = tempfile.mkdtemp(prefix="lsmckpt_")
        try:
            files, nconn, nouter, A = run(lib, name, DT, work)
        finally:
            shutil.rmtree(work, ignore_errors=True)
        for f, arr in files.items():
            out[f"{name}_{f}"] = arr
        out[f"{name}_conn_lines"] = nconn
        out[f"{name}_outer"] = nouter
        print(name, "outer", nouter, "connectivity lines", nconn, flush=True)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
