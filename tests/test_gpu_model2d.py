"""The 2D models (drain_model2d, imbibe_model2d, constCurvModel2d[NoVar],
reinitializeSubcellFix2d) against vectors the reference build wrote
(tests/golden/model2d_f64.npz, made by tests/golden/make_model2d.py from
src/level_set_evolution/drain_model2d.c, imbibe_model2d.c,
constCurvModel2d.c and src/reinitialization/reinitialize_top.c:123-203).

FP64 throughout.  The GPU evaluates cos and sin (wetting coefficients,
smoothed Heaviside) with the device math library and the reference with
glibc, so the fields agree to rounding, not bit for bit; the tolerances
below are written for that.  The constant-coefficient 2D right-hand side
(theta = 0) has no reference vector: the reference adds an unassigned
temp_vel there (qss_spatial_derivatives2d.f:74,215).  It is pinned instead
by being equal, bit for bit, to the variable-coefficient path run with
uniform coefficient fields, which the reference vectors pin.
"""
import ctypes as C
import os
import sys
import tempfile

import numpy as np
import pytest

from lsmpqs_amd.abi import Arrays, default_options
from lsm_cases import DEVNULL

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_model2d as mk  # noqa: E402

FIX = os.path.join(HERE, "golden", "model2d_f64.npz")
REF_SO = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libqssref_f64.so")
PHI_TOL = 1e-8


def fixture():
    return np.load(FIX)


def fill2d(g):
    return (slice(g.jlo_fb, g.jhi_fb + 1), slice(g.ilo_fb, g.ihi_fb + 1))


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference build not present")
def test_fixture_is_the_reference_output():
    """Re-running the reference build reproduces the committed vectors exactly."""
    d = fixture()
    out, _ = mk.run_ref("eno3_rk3_t45")
    for k, v in out.items():
        np.testing.assert_array_equal(v, d[k], err_msg=k)
    rr = mk.run_ref_reinit()
    np.testing.assert_array_equal(rr["reinit/out"], d["reinit/out"])


def test_fixture_sanity():
    """Properties the reference output must have: the solid stays imposed,
    the saturation moved, the trapped masks are -20 dx where nothing is trapped."""
    d = fixture()
    for name in mk.CASES:
        mask, phi = d[f"{name}/in_mask"], d[f"{name}/phi"]
        assert np.all(phi[mask > 0] >= mask[mask > 0] - 3.0 - 1e-12), name
        assert not np.array_equal(phi, d[f"{name}/in_phi0"]), name
        assert int(d[f"{name}/n_drain_steps"]) == 2, name
        for m in ("mask_w", "mask_nw"):
            f = d[f"{name}/{m}"]
            assert np.all(np.isfinite(f)), name


def _hip():
    from lsmpqs_amd import hiplib
    return hiplib, hiplib.load(np.float64)


def _run_hip(name, d):
    space, time, theta_deg, center, imbibe = mk.CASES[name]
    hiplib, lib = _hip()
    lib.lsmpqs_set_dt_slabs(mk.SLABS)
    g = mk.grid_for(name, lib)
    A = Arrays(g, np.float64)
    A["mask"] = d[f"{name}/in_mask"].reshape(A.shape)
    A["phi"] = d[f"{name}/in_phi0"].reshape(A.shape)
    o = mk.case_options(space, time, theta_deg, center)
    cwd = os.getcwd()
    tmp = tempfile.mkdtemp(prefix="lsm2d_hip_")
    os.chdir(tmp)
    saved = os.environ.pop("LSMPQS_NO_STEP_FILES", None)
    try:
        lib.drain_model2d(C.byref(o), A.ptr(), C.byref(g), DEVNULL)
        hiplib.check(lib)
        n_drain = len([f for f in os.listdir(tmp) if f.startswith("data_step_")])
        if imbibe:
            o.amin = o.a + o.b * o.dc * 1.001 - 2 * o.b * o.dc * 1.001
            o.a = o.a + o.b * o.dc
            lib.imbibe_model2d(C.byref(o), A.ptr(), C.byref(g), DEVNULL)
            hiplib.check(lib)
        n_files = len([f for f in os.listdir(tmp) if f.startswith("data_step_")])
    finally:
        os.chdir(cwd)
        if saved is not None:
            os.environ["LSMPQS_NO_STEP_FILES"] = saved
    return A, o, g, n_drain, n_files


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(mk.CASES))
def test_model2d_matches_reference(name):
    print(f"\n[model2d] case {name}", flush=True)
    d = fixture()
    A, o, g, n_drain, n_files = _run_hip(name, d)
    assert n_drain == int(d[f"{name}/n_drain_steps"])
    assert n_files == int(d[f"{name}/n_step_files"])
    shape = d[f"{name}/in_mask"].shape
    for f in ("overlap", "phi_nw_ind", "phi_w_ind", "conserve_imbibe"):
        assert getattr(o, f) == d[f"{name}/opt_{f}"], f
    for f in ("b_max_over_dx", "max_U_over_dx"):
        assert abs(getattr(o, f) - float(d[f"{name}/opt_{f}"])) <= 1e-12 * max(1.0, abs(getattr(o, f))), f
    fb = fill2d(g)
    # phi over the fill box and the whole ghost box
    phi, ref = A["phi"].reshape(shape), d[f"{name}/phi"]
    err = np.max(np.abs(phi - ref))
    assert err < PHI_TOL, f"phi max diff {err:.3e}"
    # the interface (sign of phi) and the trapped-phase masks
    assert np.array_equal(phi[fb] < 0, ref[fb] < 0)
    for f in ("mask", "mask_w", "mask_nw", "mask_disconn_init", "mask_x", "mask_y",
              "normal_velocity", "curvature_coeff", "external_velocity_x", "external_velocity_y"):
        got, exp = A[f].reshape(shape), d[f"{name}/{f}"]
        e = np.max(np.abs(got[fb] - exp[fb]))
        assert e < PHI_TOL, f"{f} max diff {e:.3e}"


@pytest.mark.gpu
def test_reinit2d_matches_reference():
    print("\n[model2d] reinitializeSubcellFix2d", flush=True)
    d = fixture()
    hiplib, lib = _hip()
    g = mk.grid2d(lib)
    o = default_options(np.float64)
    o.tmax_r = float(d["reinit/tmax_r"])
    data = np.ascontiguousarray(d["reinit/in"].copy())
    lib.reinitializeSubcellFix2d(data.ctypes.data_as(C.c_void_p), C.byref(g), C.byref(o))
    hiplib.check(lib)
    # no transcendental functions on this path: every operation is an IEEE
    # double add, multiply, divide or sqrt in the reference's order
    np.testing.assert_array_equal(data, d["reinit/out"])


@pytest.mark.gpu
def test_novar_equals_var_with_uniform_fields():
    """constCurvModel2dNoVar(a) == constCurvModel2d with vel_n = a, b = b,
    V = 0 everywhere and no solid (the var model's extra mask imposition is
    then a no-op), one outer step, no trapping: the same kernels apart from
    the scalar/field coefficient loads and a -0 advective term."""
    print("\n[model2d] NoVar vs uniform var", flush=True)
    d = fixture()
    hiplib, lib = _hip()
    lib.lsmpqs_set_dt_slabs(mk.SLABS)
    g = mk.grid2d(lib)
    name = "eno3_rk3_t45"
    res = []
    for var in (False, True):
        A = Arrays(g, np.float64)
        A["mask"] = -50.0
        A["phi"] = d[f"{name}/in_phi0"].reshape(A.shape) - 0.3
        o = default_options(np.float64)
        o.order_space_accur, o.order_time_accur = 3, 3
        o.check_connectivity = 0
        o.tmax = o.tplot
        o.b = 0.04
        inv = 1.0 / g.dx[0]
        o.b_max_over_dx = 2 * o.b * (inv * inv + inv * inv)
        o.max_U_over_dx = 0.0
        a = 0.07
        if var:
            A["normal_velocity"] = a
            A["curvature_coeff"] = o.b
            t = lib.constCurvModel2d(C.byref(o), A.ptr(), C.byref(g), DEVNULL)
        else:
            t = lib.constCurvModel2dNoVar(C.byref(o), A.ptr(), C.byref(g), DEVNULL, a)
        hiplib.check(lib)
        assert t > 0
        res.append((t, A["phi"].copy()))
    assert res[0][0] == res[1][0]
    np.testing.assert_array_equal(res[0][1], res[1][1])


OPTIONS_2D = """\
outfile run.log
accuracy medium
space_accur 3
time_accur 3
theta 30
a 0.05
b 0.04
dc 0.5
amax 0.0901
amin 0.0
tmax 4
tplot 0.5
check_connectivity 1
save_data 1
print_details 1
"""


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference build not present")
def test_drain_top_dispatches_2d_grids(tmp_path):
    """drainTop on 2D files runs drain_model2d (drain_top.c:116-119): the step
    files and data_final.gz match the reference build's drainTop on the same files."""
    import gzip
    import oracle_py as op
    print("\n[model2d] drainTop on a 40x32 grid", flush=True)
    d = fixture()
    hiplib, ours = _hip()
    ref = op.ref_lib(np.float64)
    g = mk.grid2d(ours)
    name = "eno3_rk3_t45"
    files = {k: str(tmp_path / k) for k in ("data", "mask", "grid", "opts")}
    for key, f in (("in_phi0", "data"), ("in_mask", "mask")):
        arr = np.ascontiguousarray(d[f"{name}/{key}"])
        ours.writeDataArrayQSS(arr.ctypes.data_as(C.c_void_p), C.byref(g), files[f].encode(), 0)
    ours.writeGridToBinaryFile(C.byref(g), files["grid"].encode(), 0)
    open(files["opts"], "w").write(OPTIONS_2D)
    op.set_omp_threads(mk.SLABS)
    ours.lsmpqs_set_dt_slabs(mk.SLABS)
    saved = os.environ.pop("LSMPQS_NO_STEP_FILES", None)
    rcs = {}
    try:
        for tag, lib in (("ours", ours), ("ref", ref)):
            wd = tmp_path / tag
            os.makedirs(wd)
            cwd = os.getcwd()
            os.chdir(wd)
            try:
                lib.createOptionsFromInputFile.restype = C.c_void_p
                lib.createOptionsFromInputFile.argtypes = [C.c_char_p]
                o = lib.createOptionsFromInputFile(files["opts"].encode())
                rcs[tag] = lib.drainTop(C.c_void_p(o), files["data"].encode(),
                                        files["grid"].encode(), files["mask"].encode())
                lib.destroyOptions.argtypes = [C.c_void_p]
                lib.destroyOptions(C.c_void_p(o))
            finally:
                os.chdir(cwd)
    finally:
        if saved is not None:
            os.environ["LSMPQS_NO_STEP_FILES"] = saved
    assert rcs["ours"] == 0, hiplib.last_error(ours)
    assert rcs["ref"] == 0
    outs = sorted(f for f in os.listdir(tmp_path / "ref") if f.startswith("data_"))
    assert "data_final.gz" in outs and any(f.startswith("data_step") for f in outs)
    assert outs == sorted(f for f in os.listdir(tmp_path / "ours") if f.startswith("data_"))
    for f in outs:
        a, b = (gzip.open(str(tmp_path / t / f)).read() for t in ("ours", "ref"))
        assert a[:12] == b[:12], f
        x, y = np.frombuffer(a[12:], np.float64), np.frombuffer(b[12:], np.float64)
        assert np.max(np.abs(x - y)) < PHI_TOL, f
    log_o = open(tmp_path / "ours" / "run.log").read()
    log_r = open(tmp_path / "ref" / "run.log").read()
    assert log_o.count("nw phase satn") == log_r.count("nw phase satn") > 0


@pytest.mark.gpu
def test_helpers_on_2d_grids_match_reference():
    """getMainInd_new + trapComponents_mask (side and centre inlet),
    findConnectivity2d, signedLinearExtrapolationBCqss, initializeDisconnectedMasks
    and setThetaOverlap on a 2D grid, bit for bit against the reference build's
    outputs (integer labels and IEEE add/multiply/sqrt arithmetic only)."""
    print("\n[model2d] helper entry points on a 2D grid", flush=True)
    d = fixture()
    hiplib, lib = _hip()
    g = mk.grid2d(lib)
    mask, blobs = d["helpers/mask"], d["helpers/phi"]
    for center in (0, 1):
        print(f"[model2d] trapComponents_mask center_inlet={center}", flush=True)
        A = Arrays(g, np.float64)
        A["mask"] = mask.reshape(A.shape)
        A["phi"] = np.maximum(0.35 - (blobs + 0.1), mask).reshape(A.shape)
        A["mask_nw"] = -20.0
        o = default_options(np.float64)
        o.center_inlet = center
        lib.getMainInd_new(C.byref(o), A.ptr(), C.byref(g))
        key = f"helpers/trap{center}"
        assert o.phi_nw_ind == int(d[f"{key}/phi_nw_ind"])
        assert o.phi_w_ind == int(d[f"{key}/phi_w_ind"])
        lib.trapComponents_mask(A.ptr(), C.byref(g), C.byref(o))
        hiplib.check(lib)
        for f in ("phi", "mask_nw", "mask_w", "connectivity", "phi_bin"):
            np.testing.assert_array_equal(A[f].reshape(mask.shape), d[f"{key}/{f}"], err_msg=f)
    A = Arrays(g, np.float64)
    A["phi_bin"] = (blobs + 0.1 > 0.35).astype(np.uint8).reshape(A.shape)
    lib.findConnectivity2d(A.ptr(), C.byref(g))
    hiplib.check(lib)
    np.testing.assert_array_equal(A["connectivity"].reshape(mask.shape), d["helpers/conn"])
    f = np.ascontiguousarray(d["helpers/bc_in"].copy())
    lib.signedLinearExtrapolationBCqss(f.ctypes.data_as(C.c_void_p), C.byref(g), 9)
    hiplib.check(lib)
    np.testing.assert_array_equal(f, d["helpers/bc_out"])
    f = np.zeros_like(blobs)
    lib.initializeDisconnectedMasks(f.ctypes.data_as(C.c_void_p), C.byref(g))
    hiplib.check(lib)
    np.testing.assert_array_equal(f, d["helpers/disconn"])
    for deg in (10, 30):
        o = default_options(np.float64)
        o.theta = deg * 3.14159 / 180
        lib.setThetaOverlap(None, C.byref(g), C.byref(o))
        assert o.overlap == float(d[f"helpers/overlap_{deg}"])


@pytest.mark.gpu
def test_model2d_fp32_tracks_fp64_reference():
    """The FP32 library on the ENO3/RK3 45-degree case against the FP64
    reference vectors: same curvature steps, saturation within 2e-3, the
    interface (sign of phi) equal on at least 99.5 % of the fill box."""
    print("\n[model2d] FP32 eno3_rk3_t45", flush=True)
    from lsmpqs_amd import hiplib
    d = fixture()
    name = "eno3_rk3_t45"
    space, time, theta_deg, center, _ = mk.CASES[name]
    lib = hiplib.load(np.float32)
    lib.lsmpqs_set_dt_slabs(mk.SLABS)
    g = mk.grid2d(lib, np.float32)
    A = Arrays(g, np.float32)
    A["mask"] = d[f"{name}/in_mask"].reshape(A.shape)
    A["phi"] = d[f"{name}/in_phi0"].reshape(A.shape)
    import make_model2d
    o = default_options(np.float32)
    for f in ("order_space_accur", "order_time_accur", "theta", "a", "b", "dc", "amax", "amin",
              "tmax", "center_inlet"):
        setattr(o, f, getattr(make_model2d.case_options(space, time, theta_deg, center), f))
    saved = os.environ.get("LSMPQS_NO_STEP_FILES")
    os.environ["LSMPQS_NO_STEP_FILES"] = "1"
    try:
        lib.drain_model2d(C.byref(o), A.ptr(), C.byref(g), DEVNULL)
    finally:
        if saved is None:
            del os.environ["LSMPQS_NO_STEP_FILES"]
        else:
            os.environ["LSMPQS_NO_STEP_FILES"] = saved
    hiplib.check(lib)
    shape = d[f"{name}/in_mask"].shape
    fb = fill2d(g)
    phi, ref, mask = A["phi"].reshape(shape).astype(np.float64), d[f"{name}/phi"], d[f"{name}/in_mask"]

    def satn(f):
        q = f[fb] / 1.5
        h = np.where(f[fb] < -1.5, 1.0, np.where(f[fb] < 1.5, 0.5 * (1 - q - np.sin(np.pi * q) / np.pi), 0.0))
        return h.sum()
    s32, s64 = satn(phi) / satn(mask), satn(ref) / satn(mask)
    assert abs(s32 - s64) < 2e-3, (s32, s64)
    agree = np.mean((phi[fb] < 0) == (ref[fb] < 0))
    assert agree >= 0.995, agree


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference build not present")
def test_const_curv_model2d_direct_call_matches_reference_build():
    """constCurvModel2d called directly (trapping on, indices from
    getMainInd_new, coefficient fields supplied by the caller) against the
    reference build's constCurvModel2d on the same records."""
    import oracle_py as op
    print("\n[model2d] constCurvModel2d direct call", flush=True)
    d = fixture()
    name = "eno3_rk3_t45"
    hiplib, ours = _hip()
    ref = op.ref_lib(np.float64)
    op.set_omp_threads(mk.SLABS)
    ours.lsmpqs_set_dt_slabs(mk.SLABS)
    ref.constCurvModel2d.restype = C.c_double
    out = {}
    for tag, lib in (("ours", ours), ("ref", ref)):
        g = mk.grid2d(lib)
        A = Arrays(g, np.float64)
        for f, key in (("mask", "mask"), ("phi", "in_phi0"), ("normal_velocity", "normal_velocity"),
                       ("curvature_coeff", "curvature_coeff"),
                       ("external_velocity_x", "external_velocity_x"),
                       ("external_velocity_y", "external_velocity_y")):
            A[f] = d[f"{name}/{key}"].reshape(A.shape)
        A["mask_w"] = -20.0
        A["mask_nw"] = -20.0
        o = mk.case_options(3, 3, 45.0, 0)
        o.tmax = 2.0
        o.overlap = float(d[f"{name}/opt_overlap"])
        o.b_max_over_dx = float(d[f"{name}/opt_b_max_over_dx"])
        o.max_U_over_dx = float(d[f"{name}/opt_max_U_over_dx"])
        lib.getMainInd_new(C.byref(o), A.ptr(), C.byref(g))
        # the reference calls options->space_deriv_func, which this sets
        # (a NULL slot crashes it); this library chooses its kernels itself
        lib.setSpaceDerivFunc(C.byref(g), C.byref(o))
        t = lib.constCurvModel2d(C.byref(o), A.ptr(), C.byref(g), DEVNULL)
        out[tag] = (float(t), A["phi"].copy(), A["mask_w"].copy(), A["mask_nw"].copy())
    hiplib.check(ours)
    assert out["ours"][0] == pytest.approx(out["ref"][0], rel=1e-12, abs=1e-12)
    for i, f in ((1, "phi"), (2, "mask_w"), (3, "mask_nw")):
        e = float(np.max(np.abs(out["ours"][i] - out["ref"][i])))
        assert e < PHI_TOL, f"{f} max diff {e:.3e}"


def test_model2d_input_errors_need_no_gpu():
    """Input errors of the 2D entry points are reported before any device
    call: a 3D grid, use_var_theta, a single-edge boundary index."""
    from lsmpqs_amd import hiplib
    from lsmpqs_amd.abi import make_grid
    path = hiplib.lib_path(np.float64)
    if not os.path.exists(path):
        pytest.skip("library not built")
    lib = hiplib.load(np.float64)
    g3 = make_grid((8, 8, 8), np.float64)
    A3 = Arrays(g3, np.float64)
    o = default_options(np.float64)
    o.theta = 0.5
    lib.drain_model2d(C.byref(o), A3.ptr(), C.byref(g3), DEVNULL)
    assert "2D grid" in hiplib.last_error(lib)
    g2 = mk.grid2d(lib)
    A2 = Arrays(g2, np.float64)
    o.use_var_theta = 1
    lib.imbibe_model2d(C.byref(o), A2.ptr(), C.byref(g2), DEVNULL)
    assert "use_var_theta" in hiplib.last_error(lib)
    f = np.zeros(A2.shape)
    lib.signedLinearExtrapolationBCqss(f.ctypes.data_as(C.c_void_p), C.byref(g2), 0)
    assert "bdry_location_idx 9" in hiplib.last_error(lib)
