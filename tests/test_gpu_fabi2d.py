"""The 2D kernel-level entry points (qss2dgetrhs2nd_, qss2dgetrhs3rd_, the
constant-coefficient forms, the 2D TVD-RK stages and the one-edge signed
linear extrapolation, in lsm_model2d.hip) against the reference build's
routines of the same names (oracle/_ref), run on the same seeded 2D fields.

FP64, IEEE add/multiply/divide/sqrt in the reference's order: bit for bit.
The reference's constant-coefficient 2D right-hand sides add an unassigned
temp_vel (qss_spatial_derivatives2d.f:74,215), so those are pinned instead by
equality with the variable-coefficient routines on uniform fields.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_py as op

DT = np.float64
NX, NY = 37, 29            # ghost box; fill box [3, N-4]
GB = (0, NX - 1, 0, NY - 1)
FB = (3, NX - 4, 3, NY - 4)


def fields(seed=5):
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.arange(NX, dtype=DT), np.arange(NY, dtype=DT))
    phi = np.hypot(X - 17.2, Y - 13.9) - 8.0 + 0.05 * rng.normal(size=X.shape)
    vn = rng.normal(size=X.shape)
    vn[::5] = 3e-6          # rows with |vel_n| < 1e-5 reuse the previous normal term
    b = rng.uniform(0, 0.1, size=X.shape)
    vx, vy = rng.normal(size=X.shape), rng.normal(size=X.shape)
    return dict(phi=phi, vn=vn, b=b, vx=vx, vy=vy, u1=phi + 0.1 * vx, un=phi - 0.2 * vy)


def run_all(lib, dx=0.5, dy=0.75):
    f = fields()
    out = {}
    for nm in ("qss2dgetrhs2nd_", "qss2dgetrhs3rd_"):
        rhs = np.zeros_like(f["phi"])
        h = op.Scalar(0.0, C.c_double)
        op.fcall(lib, nm, DT, f["phi"], f["vn"], f["b"], f["vx"], f["vy"], h, rhs, *GB, *FB, dx, dy)
        out[nm] = rhs
        out[nm + "H"] = h.value
    u = np.zeros_like(f["phi"])
    op.fcall(lib, "qss2drk1step_", DT, u, *GB, f["u1"], *GB, f["b"], *GB, *FB, 0.3)
    out["rk1"] = u
    u = np.zeros_like(f["phi"])
    op.fcall(lib, "qss2dtvdrk2stage2_", DT, u, *GB, f["u1"], *GB, f["un"], *GB, f["vx"], *GB, *FB, 0.3)
    out["rk2"] = u
    u = np.zeros_like(f["phi"])
    op.fcall(lib, "qss2dtvdrk3stage3_", DT, u, *GB, f["u1"], *GB, f["un"], *GB, f["vy"], *GB, *FB, 0.3)
    out["rk3"] = u
    bc = f["vx"].copy()
    for edge in range(4):
        op.fcall(lib, "qss2dsignedlinearextrapolation_", DT, bc, *GB, *FB, edge)
    out["bc"] = bc
    return out


@pytest.mark.gpu
@pytest.mark.skipif(not op.have_ref(DT), reason="reference build not present")
def test_kernel_entry_points_2d_match_reference_build():
    print("\n[fabi2d] 2D kernel-level entry points vs reference build", flush=True)
    from lsmpqs_amd import hiplib
    ours = hiplib.load(DT)
    a = run_all(ours)
    hiplib.check(ours)
    r = run_all(op.ref_lib(DT))
    for k in r:
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(r[k]), err_msg=k)


@pytest.mark.gpu
def test_const_rhs_2d_equals_var_rhs_on_uniform_fields():
    print("\n[fabi2d] const vs uniform var right-hand sides", flush=True)
    from lsmpqs_amd import hiplib
    lib = hiplib.load(DT)
    f = fields()
    one, zero = np.ones_like(f["phi"]), np.zeros_like(f["phi"])
    for s in ("2nd", "3rd"):
        rv, rc = np.zeros_like(one), np.zeros_like(one)
        hv, hc = op.Scalar(0.0, C.c_double), op.Scalar(0.0, C.c_double)
        op.fcall(lib, f"qss2dgetrhs{s}_", DT, f["phi"], 0.07 * one, 0.04 * one, zero, zero, hv, rv,
                 *GB, *FB, 1.0, 1.0)
        op.fcall(lib, f"qss2dgetrhs{s}const_", DT, f["phi"], 0.07, 0.04, hc, rc, *GB, *FB, 1.0, 1.0)
        hiplib.check(lib)
        np.testing.assert_array_equal(rv, rc, err_msg=s)
        assert hv.value == hc.value
