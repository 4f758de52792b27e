"""The library with the reference's quirks switched off ("fixed mode":
lsmpqs_set_quirks(0): the ENO3 x line reads phi(i+3,j,k), the normal term
is -vel_n |grad phi|_G for every scheme, no stale normal terms) against the
oracle with the same switches off (orc_set_quirks(0), pinned by closed forms
in tests/test_oracle_closed_forms.py).  This is the mode the batched
subvolumes (non-cubic boxes, where the reference's phi(i+3,j,j) read leaves
the array) and the adjoint run in.  FP64, to 1e-10 (bit for bit on the
constant-coefficient path); the batched boxes box by box."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_py as op
from lsm_cases import run_model, maxdiff, case, options, DEVNULL, saturation

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


class fixed_mode:
    """Quirk switches of the oracle and of the library set to `bits`."""

    def __init__(self, dtype, bits=0):
        self.dtype, self.bits = dtype, bits

    def __enter__(self):
        from lsmpqs_amd import hiplib
        orc = op.oracle_lib(self.dtype)
        orc.orc_set_quirks.restype = C.c_uint
        orc.orc_set_quirks.argtypes = [C.c_uint]
        self.old_orc = orc.orc_set_quirks(self.bits)
        self.old_lib = hiplib.load(self.dtype).lsmpqs_set_quirks(self.bits)
        return self

    def __exit__(self, *exc):
        from lsmpqs_amd import hiplib
        op.oracle_lib(self.dtype).orc_set_quirks(self.old_orc)
        hiplib.load(self.dtype).lsmpqs_set_quirks(self.old_lib)


CASES = [
    ("eno3c_rk3_theta0", (24, 40, 16), dict(space=3, time=3, theta_deg=0)),
    ("eno3_rk3_theta45", (40, 28, 20), dict(space=3, time=3, theta_deg=45)),
    ("weno5_rk3_theta30", (28, 36, 20), dict(space=5, time=3, theta_deg=30)),
    ("eno2_rk2_theta60", (32, 32, 32), dict(space=2, time=2, theta_deg=60)),
]


@pytest.mark.parametrize("name,dims,opts", CASES, ids=[c[0] for c in CASES])
def test_fixed_mode_drain_matches_fixed_oracle(name, dims, opts):
    print(f"case fixed-mode drain {name} {dims} fp64", flush=True)
    with fixed_mode(np.float64):
        Ao, _, _ = run_model("oracle", np.float64, dims, opts)
        Ah, _, _ = run_model("hip", np.float64, dims, opts)
    for f in ("phi", "mask_w", "mask_nw"):
        d = maxdiff(Ao[f], Ah[f])
        print(f"  {f}: max|diff| = {d:.3e}", flush=True)
        assert d <= TOL64, (name, f, d)
    # the fixed mode is a different trajectory from the reference mode
    Ar, _, _ = run_model("oracle", np.float64, dims, opts)
    assert maxdiff(Ar["phi"], Ao["phi"]) > 0 or opts["space"] == 2


def test_fixed_mode_fp32_against_fixed_fp64_oracle():
    """The FP32 library (fast arithmetic) in fixed mode against the FP64
    oracle in fixed mode: saturation within 1e-4."""
    dims, opts = (28, 36, 20), dict(space=5, time=3, theta_deg=30)
    print("case fixed-mode fp32 weno5 (28,36,20)", flush=True)
    with fixed_mode(np.float64), fixed_mode(np.float32):
        A64, _, g = run_model("oracle", np.float64, dims, opts)
        A32, _, g32 = run_model("hip", np.float32, dims, opts)
    s64 = saturation(A64["phi"], A64["mask"], g)
    s32 = saturation(A32["phi"], A32["mask"], g32)
    print(f"  saturation fp64 oracle {s64:.8f} fp32 gpu {s32:.8f}", flush=True)
    assert abs(s64 - s32) < 1e-4


SHAPES = [(20, 24, 18), (28, 20, 22), (16, 30, 24), (24, 24, 24), (18, 22, 30)]
STARTS = [0.02, 0.025, 0.018, 0.03, 0.022]


def _boxes(dt):
    from lsmpqs_amd.abi import Arrays
    out = []
    for i, (shape, a) in enumerate(zip(SHAPES, STARTS)):
        g, mask, phi0 = case(shape, dt, seed=2000 + i)
        o = options(dt, space=3, time=3, a=a, b=0.04, dc=0.25, steps=3 + (i % 2), tmax=15.0)
        A = Arrays(g, dt)
        A["mask"] = mask
        A["phi"] = phi0
        out.append((g, o, A))
    return out


def test_fixed_mode_batched_boxes_match_fixed_oracle():
    """Five non-cubic boxes, each with its own curvature schedule, drained
    by lsmpqs_batch_model3d (one persistent brick-queue stage kernel per RK
    stage, replayed graph) in fixed mode: every box against the oracle's
    fixed-mode drain_model3d of that box alone.  ENO3 + TVD-RK3, theta 0,
    trapping on, FP64: bit for bit."""
    from lsmpqs_amd import hiplib
    dt = np.float64
    print("case fixed-mode batch of 5 boxes fp64", flush=True)
    lib = hiplib.load(dt)
    orc = op.oracle_lib(dt)
    os.environ["LSMPQS_NO_STEP_FILES"] = "1"
    T = 4
    lib.lsmpqs_set_dt_slabs(T)
    with fixed_mode(dt):
        ref, batched = _boxes(dt), _boxes(dt)
        for g, o, A in ref:
            orc.orc_drain_model3d(C.byref(o), A.ptr(), C.byref(g), DEVNULL, C.c_int(T))
        n = len(batched)
        opts = (C.c_void_p * n)(*[C.addressof(o) for _, o, _ in batched])
        recs = (C.c_void_p * n)(*[C.addressof(A.rec) for _, _, A in batched])
        grids = (C.c_void_p * n)(*[C.addressof(g) for g, _, _ in batched])
        st = hiplib.BatchStats()
        rc = lib.lsmpqs_batch_model3d(n, opts, recs, grids, None, 1, C.byref(st))
        assert rc == 0, hiplib.last_error(lib)
        for _, _, A in batched:
            lib.lsmpqs_release(A.ptr())
    for i, ((_, _, Ar), (_, _, Ab)) in enumerate(zip(ref, batched)):
        for f in ("phi", "mask_w", "mask_nw"):
            d = maxdiff(Ar[f], Ab[f])
            print(f"  box {i} {SHAPES[i]} {f}: max|diff| {d:.2e}", flush=True)
            assert d == 0.0, (i, f, d)
