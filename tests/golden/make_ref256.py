"""Writes tests/golden/ref256_drain_imbibe.npz by running the REFERENCE build
(oracle/_ref/libqssref_f64.so, compiled from /root/reference by
oracle/Makefile.ref) on the 256^3 drainage + imbibition case.

    python tests/golden/make_ref256.py [threads]

The case: 256^3 seeded Boolean sphere pack (radius U[6,12], porosity 0.36,
seed 1002), HJ-WENO5 + TVD-RK3, contact angle 30 deg, trapping on,
eps_stop 1e-4 dx; drain_model3d over 2 curvature steps from a = 0.02
(b = 0.04, dc = 0.25, tmax 10), then imbibe_model3d over 2 steps from the
last drainage pressure.  The full fields are 140 MB each, so the fixture
keeps what a comparison needs at 1e-10 (data only):
  sat_*        FP64 smoothed non-wetting saturation of the state
  plane_sum_*  per-z-plane sums of phi over the fill box
  idx, val_*   4096 seeded voxels of the |phi| < 3 band and their phi
  nneg_*       number of fill voxels with phi < 0
  nw_trap_*, w_trap_*  number of voxels with mask_nw > 0, mask_w > 0
  outer_*      number of outer steps in the log
"""
import ctypes as C
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "lsmpqs-1.0_amd"), os.path.join(ROOT, "oracle"),
                os.path.join(ROOT, "tests")]

import oracle_py as op  # noqa: E402
from lsm_cases import arrays, fill, saturation  # noqa: E402
from lsmpqs_amd.abi import default_options  # noqa: E402
from lsmpqs_amd.inputs import pack_case  # noqa: E402

N = 256
SEED = 1002
NSAMPLE = 4096
OUT = os.path.join(HERE, "ref256_drain_imbibe.npz")


def case_options(dtype):
    """Options of the case (shared with the GPU test)."""
    o = default_options(dtype)
    o.order_space_accur = 5
    o.order_time_accur = 3
    o.theta = 30 * 3.14159 / 180
    o.a = 0.02
    o.b = 0.04
    o.dc = 0.25
    o.amax = o.a + o.b * o.dc * 1.5
    o.amin = o.a - o.b * o.dc * 0.5
    o.tmax = 10.0
    o.eps_stop = 1e-4
    o.check_connectivity = 1
    return o


def summarise(A, g, tag, out, rng_idx=None):
    phi = A["phi"]
    f = fill(g)
    out[f"sat_{tag}"] = saturation(phi, A["mask"], g)
    out[f"plane_sum_{tag}"] = phi[f].sum(axis=(1, 2))
    out[f"nneg_{tag}"] = int((phi[f] < 0).sum())
    out[f"nw_trap_{tag}"] = int((A["mask_nw"][f] > 0).sum())
    out[f"w_trap_{tag}"] = int((A["mask_w"][f] > 0).sum())
    if rng_idx is not None:
        out[f"val_{tag}"] = phi.ravel()[rng_idx]


def band_sample(phi, g, seed):
    f = np.zeros(phi.shape, dtype=bool)
    f[fill(g)] = True
    cand = np.flatnonzero((np.abs(phi) < 3.0) & f)
    return np.sort(np.random.default_rng(seed).choice(cand, NSAMPLE, replace=False))


def count_outer(path):
    return sum(1 for line in open(path) if "nw phase satn" in line)


def main():
    T = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    dt = np.float64
    g, mask, phi0, _ = pack_case((N, N, N), (6.0, 12.0), 0.36, SEED, dt)
    A = arrays(g, mask, phi0, dt)
    o = case_options(dt)
    lib = op.ref_lib(dt)
    op.set_omp_threads(T)
    libc = C.CDLL("libc.so.6")
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    work = tempfile.mkdtemp(prefix="lsmref256_")
    cwd = os.getcwd()
    out = {"threads": T, "n": N, "seed": SEED}
    os.chdir(work)
    try:
        logs = {}
        for tag, fn in (("drain", lib.drain_model3d), ("imbibe", lib.imbibe_model3d)):
            path = os.path.join(work, f"{tag}.log")
            fp = C.c_void_p(libc.fopen(path.encode(), b"w"))
            if tag == "imbibe":
                o.a = o.a + o.b * o.dc
            fn(C.byref(o), A.ptr(), C.byref(g), fp)
            libc.fclose(fp)
            logs[tag] = path
            if tag == "drain":
                idx = band_sample(A["phi"], g, SEED)
                out["idx"] = idx
            summarise(A, g, tag, out, idx)
            out[f"outer_{tag}"] = count_outer(path)
            print(tag, "done: S_nw", out[f"sat_{tag}"], "outer", out[f"outer_{tag}"], flush=True)
    finally:
        os.chdir(cwd)
    np.savez(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
