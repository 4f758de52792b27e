"""Direct calls of constCurvModel3d and constCurvModel3dNoVar (the outer /
inner pseudo-time loop at fixed pressure; reference:
src/level_set_evolution/constCurvModel3d.c:50-334, :336-610) against the
oracle's orc_constCurvModel3d / orc_constCurvModel3dNoVar on the same
records.  The inputs are the state drain_model3d hands its first call: the
oracle's set-up (trapped masks, reservoir, reinitialized mask, mask
gradient, overlap, inlet indices), the coefficient fields of the first
curvature step (setVarCurvAdv3d, setVarNorm3d) and one trapping pass.  The
library's entry points stage every host array themselves; phi, phi_prev,
mask_w, mask_nw, the returned pseudo-time and, on the contact-angle path
with conserve_imbibe, the coefficient fields must come back equal to 1e-10
(FP64), bit for bit on the constant path."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as op
from lsm_cases import arrays, case, options, maxdiff, DEVNULL

pytestmark = pytest.mark.gpu

T = 4
COEFFS = ("normal_velocity", "curvature_coeff", "external_velocity_x", "external_velocity_y",
          "external_velocity_z")


def prepared(dt, n, opts, a0, imbibe=False):
    """(grid, options, arrays) after the oracle's set-up and the coefficient
    and trapping work before the first constCurvModel3d call."""
    g, mask, phi0 = case(n, dt)
    A = arrays(g, mask, phi0, dt)
    o = options(dt, **opts)
    orc = op.oracle_lib(dt)
    orc.orc_model_setup3d(C.byref(o), A.ptr(), C.byref(g), C.c_int(0 if imbibe else 1))
    gb, fb = op.gb6(g), op.fb6(g)
    R = C.c_double if dt == np.float64 else C.c_float
    if o.theta > 0:
        bm, um = op.Scalar(0.0, R), op.Scalar(0.0, R)
        op.fcall(orc, "orc_setvarcurvadv3d", dt, A["mask"], A["mask_x"], A["mask_y"], A["mask_z"],
                 A["curvature_coeff"], A["external_velocity_x"], A["external_velocity_y"],
                 A["external_velocity_z"], bm, um, float(o.b), *gb, *fb, float(g.dx[0]))
        o.b_max_over_dx, o.max_U_over_dx = bm.value, um.value
        op.fcall(orc, "orc_setvarnorm3d", dt, A["mask"], A["mask_x"], A["mask_y"], A["mask_z"],
                 A["normal_velocity"], a0, float(o.theta), *gb, *fb, float(g.dx[0]))
    else:
        o.b_max_over_dx = 2 * o.b * 3.0
        o.max_U_over_dx = 0.0
    if imbibe:
        o.conserve_imbibe = 1
    orc.orc_trapComponents_mask(A.ptr(), C.byref(g), C.byref(o))
    return g, o, A


def copy_of(g, A, dt):
    B = arrays(g, A["mask"], A["phi"], dt)
    for f in A.a:
        B[f] = A[f]
    return B


CASES = [
    ("var_weno5_rk3_theta30", dict(space=5, time=3, theta_deg=30), True, False),
    ("var_eno2_rk2_theta45_imbibe", dict(space=2, time=2, theta_deg=45), True, True),
    ("novar_eno2_rk2", dict(space=2, time=2, theta_deg=0), False, False),
    ("novar_eno3_rk3", dict(space=3, time=3, theta_deg=0), False, False),
]


@pytest.mark.parametrize("name,opts,var,imbibe", CASES, ids=[c[0] for c in CASES])
def test_constcurvmodel3d_entry_points_match_oracle(name, opts, var, imbibe):
    from lsmpqs_amd import hiplib
    dt = np.float64
    a0 = 0.03
    print(f"\ncase: {'constCurvModel3d' if var else 'constCurvModel3dNoVar'} {name} 28^3 fp64",
          flush=True)
    g, o, A = prepared(dt, 28, opts, a0, imbibe)
    Ao, Ah = copy_of(g, A, dt), copy_of(g, A, dt)
    oo, oh = options(dt, **opts), options(dt, **opts)
    C.memmove(C.addressof(oo), C.addressof(o), C.sizeof(o))
    C.memmove(C.addressof(oh), C.addressof(o), C.sizeof(o))
    orc = op.oracle_lib(dt)
    lib = hiplib.load(dt)
    lib.lsmpqs_set_dt_slabs(T)
    if var:
        orc.orc_constCurvModel3d.restype = C.c_double
        t_o = orc.orc_constCurvModel3d(C.byref(oo), Ao.ptr(), C.byref(g), DEVNULL, C.c_int(T))
        t_h = lib.constCurvModel3d(C.byref(oh), Ah.ptr(), C.byref(g), DEVNULL)
    else:
        orc.orc_constCurvModel3dNoVar.restype = C.c_double
        orc.orc_constCurvModel3dNoVar.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                  C.c_double, C.c_int]
        t_o = orc.orc_constCurvModel3dNoVar(C.byref(oo), Ao.ptr(), C.byref(g), DEVNULL, a0, T)
        t_h = lib.constCurvModel3dNoVar(C.byref(oh), Ah.ptr(), C.byref(g), DEVNULL, a0)
    hiplib.check(lib)
    lib.lsmpqs_release(Ah.ptr())
    print(f"  returned t: oracle {t_o} library {t_h}", flush=True)
    assert t_h == t_o and t_h > 0
    tol = 1e-10 if var else 0.0
    fields = ["phi", "phi_prev", "mask_w", "mask_nw"] + (list(COEFFS) if imbibe else [])
    for f in fields:
        d = maxdiff(Ao[f], Ah[f])
        print(f"  {f}: max|diff| {d:.2e}", flush=True)
        assert d <= tol, (f, d)
