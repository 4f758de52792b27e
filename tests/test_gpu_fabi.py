"""The kernel-level entry points (Fortran calling convention, lsm_fabi.hip)
against the reference build's routines of the same names: the golden vectors
of tests/golden/kernels_f64.npz (written by tests/golden/make_golden.py from
oracle/_ref on a 12^3 box, |vel_n| < 1e-5 on every fourth plane), and, for
the routines those vectors do not cover, the reference build run beside it.

FP64.  The right-hand sides, RK stages, extrapolation, distance and
reinitialization right-hand side use only IEEE add, multiply, divide and
sqrt in the reference's order and must agree bit for bit; the volume sums
in a different order (rounding tolerance); the coefficient routines call
sin (smoothed Heaviside) and cos (device vs glibc, last bits).
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle_py as op

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

DT = np.float64
EXACT = ("getrhs3d2nd", "getrhs3d3rd", "getrhs3d5th", "getrhs3d2ndconst", "getrhs3d3rdconst",
         "rk1", "rk2", "rk3", "bc", "dist0", "reinit_rhs", "maxnorm")


def _ours():
    from lsmpqs_amd import hiplib
    return hiplib, hiplib.load(DT)


@pytest.mark.gpu
def test_kernel_entry_points_match_golden_vectors():
    print("\n[fabi] kernel-level entry points vs golden vectors (12^3)", flush=True)
    hiplib, lib = _ours()
    gold = np.load(os.path.join(HERE, "golden", "kernels_f64.npz"))
    g, inp = mg.kernel_inputs()
    out = mg.kernel_outputs(lib, "", "_", g, inp)
    hiplib.check(lib)
    for k in EXACT:
        np.testing.assert_array_equal(out[k], gold[k], err_msg=k)
        if k.startswith("getrhs"):
            assert float(out[k + "_H"]) == float(gold[k + "_H"]), k
    assert abs(float(out["volume"]) - float(gold["volume"])) <= 1e-12 * abs(float(gold["volume"]))


@pytest.mark.gpu
@pytest.mark.skipif(not op.have_ref(DT), reason="reference build not present")
def test_coefficient_entry_points_match_reference_build():
    print("\n[fabi] imposeMaskParallel, central gradient, coefficients vs reference", flush=True)
    hiplib, ours = _ours()
    ref = op.ref_lib(DT)
    g, inp = mg.kernel_inputs()
    gb, fb = op.gb6(g), op.fb6(g)
    mask = inp["phi"] - 1.0
    res = {}
    for tag, lib in (("ours", ours), ("ref", ref)):
        r = {}
        pm = np.zeros_like(mask) - 7.0
        op.fcall(lib, "imposemaskparallel_", DT, pm, mask, inp["vx"], 0.3, *gb, 4, 9)
        r["impose"] = pm
        gx, gy, gz = (np.zeros_like(mask) for _ in range(3))
        op.fcall(lib, "getcentralgrad3dorder2_", DT, gx, gy, gz, *gb, mask, *gb, *fb, 1.0, 1.0, 1.0)
        r.update(gx=gx, gy=gy, gz=gz)
        va = np.zeros_like(mask)
        op.fcall(lib, "setvarnorm3d_", DT, mask, gx, gy, gz, va, 0.02, 0.7, *gb, *fb, 1.0)
        r["var_a"] = va
        vb, vx, vy, vz = (np.zeros_like(mask) for _ in range(4))
        bm, um = op.Scalar(0.0, C.c_double), op.Scalar(0.0, C.c_double)
        op.fcall(lib, "setvarcurvadv3d_", DT, mask, gx, gy, gz, vb, vx, vy, vz, bm, um, 0.04,
                 *gb, *fb, 1.0)
        r.update(var_b=vb, vx=vx, vy=vy, vz=vz, bmax=bm.value, umax=um.value)
        res[tag] = r
    hiplib.check(ours)
    for k in ("impose", "gx", "gy", "gz"):
        np.testing.assert_array_equal(res["ours"][k], res["ref"][k], err_msg=k)
    for k in ("var_a", "var_b", "vx", "vy", "vz"):
        np.testing.assert_allclose(res["ours"][k], res["ref"][k], rtol=0, atol=1e-14, err_msg=k)
    for k in ("bmax", "umax"):
        assert abs(res["ours"][k] - res["ref"][k]) <= 1e-14 * max(1.0, abs(res["ref"][k])), k


def test_kernel_entry_points_exported_and_declared():
    """Every kernel-level routine is exported by both precisions' libraries and
    declared in include/lsmpqs_hip.h (no GPU call)."""
    from lsmpqs_amd import hiplib
    hdr = open(os.path.join(os.path.dirname(HERE), "include", "lsmpqs_hip.h")).read()
    names = [n for n in hiplib.EXPORTS if n.endswith("_")]
    assert len(names) == 17 + 8
    for dt in (np.float32, np.float64):
        path = hiplib.lib_path(dt)
        if not os.path.exists(path):
            pytest.skip("library not built")
        lib = C.CDLL(path)
        for n in names:
            assert hasattr(lib, n), n
            assert f"void {n}(" in hdr, n
