"""The drainage workloads at their full stated sizes (lsmpqs_amd/workloads.py):

* 64^3 pack (radius 8, porosity 0.40), HJ-ENO2 + "TVD-RK2", contact angle 0,
  10 curvature steps from kappa 0.05 to 0.30, reinitialization every 20
  time steps, FP64: the GPU against the oracle over the whole curve, with
  the reference's 26-connected trapping and with 6-connected trapping (the
  6-connected labels also against scipy);
* 512^3 pack (radius U[6,12], porosity 0.36), WENO5 + TVD-RK3, 30 deg, FP64:
  one drainage curvature step (17 outer steps) against the reference build's
  run (tests/golden/ref512_c2.npz, tests/golden/make_ref512.py; 27 minutes
  on 8 threads here);
* 512^3 pack (radius U[10,20], porosity 0.35), WENO5 + TVD-RK3, 39.999 deg:
  the FP64 GPU against the reference build's FP64 run
  (tests/golden/ref512_c3.npz) and the FP32 GPU against it (saturation
  within 1e-4; band phi L-inf and interface mismatch reported);
* the same pack at 1024^3: the FP32 GPU against the FP64 GPU over two
  curvature steps.
(reference: src/level_set_evolution/drain_model3d.c:49-165,
constCurvModel3d.c:50-610, src/connectivity/connectivity.c:148-217)."""
import ctypes as C
import importlib.util
import os

import numpy as np
import pytest

from lsm_cases import arrays, fill, saturation, maxdiff, DEVNULL
import oracle_py as op

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL64 = 1e-10


def _mk512():
    spec = importlib.util.spec_from_file_location("make_ref512",
                                                  os.path.join(HERE, "golden", "make_ref512.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _canonical_labels(binv, structure):
    """Largest linear index of each component (0 on background), from scipy."""
    from scipy import ndimage
    lab, n = ndimage.label(binv, structure=structure)
    flat = lab.ravel()
    top = np.full(n + 1, -1, dtype=np.int64)
    np.maximum.at(top, flat, np.arange(flat.size))
    return np.where(flat > 0, top[flat], 0).reshape(binv.shape).astype(np.int32)


class connectivity:
    """6- or 26-connected trapping in the oracle and the library."""

    def __init__(self, dtype, conn):
        self.dtype, self.conn = dtype, conn

    def __enter__(self):
        self.old_env = os.environ.get("LSMPQS_CONNECTIVITY")
        os.environ["LSMPQS_CONNECTIVITY"] = str(self.conn)
        self.old = op.oracle_lib(self.dtype).orc_set_connectivity(self.conn)

    def __exit__(self, *exc):
        if self.old_env is None:
            os.environ.pop("LSMPQS_CONNECTIVITY", None)
        else:
            os.environ["LSMPQS_CONNECTIVITY"] = self.old_env
        op.oracle_lib(self.dtype).orc_set_connectivity(self.old)


@pytest.mark.parametrize("conn", [26, 6])
def test_drain64_eno2_rk2_curve_matches_oracle(conn):
    from lsmpqs_amd import hiplib
    from lsmpqs_amd import workloads as W
    dt = np.float64
    T = 8
    print(f"\ncase: 64^3 ENO2/RK2 theta 0, 10 curvature steps, {conn}-connected, fp64", flush=True)
    g, mask, phi0, info = W.drain64_inputs(dt)
    res = {}
    with connectivity(dt, conn):
        for kind in ("oracle", "hip"):
            A = arrays(g, mask, phi0, dt)
            o = W.drain64_options(dt)
            if kind == "oracle":
                op.oracle_lib(dt).orc_drain_model3d(C.byref(o), A.ptr(), C.byref(g), DEVNULL,
                                                    C.c_int(T))
            else:
                lib = hiplib.load(dt)
                lib.lsmpqs_set_dt_slabs(T)
                os.environ["LSMPQS_NO_STEP_FILES"] = "1"
                lib.drain_model3d(C.byref(o), A.ptr(), C.byref(g), DEVNULL)
                hiplib.check(lib)
            res[kind] = A
            print(f"  {kind} done: S_nw {saturation(A['phi'], A['mask'], g):.6f}", flush=True)
        for f in ("phi", "mask_w", "mask_nw", "mask"):
            d = maxdiff(res["oracle"][f], res["hip"][f])
            print(f"  {f}: max|diff| {d:.2e}", flush=True)
            assert d <= TOL64, (f, d)
        if conn == 6:
            # labels of the final non-wetting phase, 6-connected, against scipy
            A = res["hip"]
            A["phi_bin"] = (A["phi"] < 0).astype(np.uint8)
            A["connectivity"] = -1
            lib.findConnectivity3d(A.ptr(), C.byref(g))
            hiplib.check(lib)
            from scipy import ndimage
            want = _canonical_labels(A["phi_bin"], ndimage.generate_binary_structure(3, 1))
            assert np.array_equal(A["connectivity"], want)
            want26 = _canonical_labels(A["phi_bin"], np.ones((3, 3, 3)))
            print(f"  6-connected components {len(np.unique(want)) - 1}, 26-connected "
                  f"{len(np.unique(want26)) - 1}", flush=True)
    lib.lsmpqs_release(res["hip"].ptr())


def _lean(dtype, mask, phi0):
    from lsmpqs_amd.abi import types_for
    _, _, DataArrays = types_for(dtype)
    keep = {"phi": np.array(phi0, dtype=dtype), "mask": np.array(mask, dtype=dtype),
            "mask_w": np.zeros(mask.shape, dtype), "mask_nw": np.zeros(mask.shape, dtype)}
    rec = DataArrays()
    for f, arr in keep.items():
        setattr(rec, f, arr.ctypes.data)
    return rec, keep


def _drain_gpu(dtype, g, mask, phi0, o, slabs):
    from lsmpqs_amd import hiplib
    lib = hiplib.load(dtype)
    lib.lsmpqs_set_dt_slabs(slabs)
    os.environ["LSMPQS_NO_STEP_FILES"] = "1"
    rec, keep = _lean(dtype, mask, phi0)
    lib.drain_model3d(C.byref(o), C.byref(rec), C.byref(g), DEVNULL)
    hiplib.check(lib)
    st = hiplib.Timing()
    lib.lsmpqs_session_stats(C.byref(rec), C.byref(g), C.byref(st))
    lib.lsmpqs_release(C.byref(rec))
    return keep, int(st.outer_steps)


def _compare_with_fixture(keep, g, ref, tag, outer, exact):
    phi = keep["phi"].astype(np.float64)
    f = fill(g)
    sat = saturation(phi, keep["mask"], g)
    dS = abs(sat - float(ref["sat_drain"]))
    band = np.abs(phi.ravel()[ref["idx"]] - ref["val_drain"])
    neg = phi < 0
    dplane = np.abs(phi[f].sum(axis=(1, 2)) - ref["plane_sum_drain"]).max()
    print(f"  {tag}: S_nw {sat:.8f} (reference {float(ref['sat_drain']):.8f}, |dS| {dS:.2e}), "
          f"band phi max|diff| {band.max():.2e} (p99.9 {np.quantile(band, 0.999):.2e}), "
          f"plane-sum max diff {dplane:.2e}, phi<0 voxels {int(neg[f].sum())} "
          f"(reference {int(ref['nneg_drain'])}), outer {outer} "
          f"(reference {int(ref['outer_drain'])})", flush=True)
    if exact:
        assert outer == int(ref["outer_drain"])
        assert dS <= 1e-8
        assert band.max() <= TOL64
        assert dplane <= 1e-10 * np.abs(ref["plane_sum_drain"]).max()
        assert int(neg[f].sum()) == int(ref["nneg_drain"])
        assert int((keep["mask_nw"][f] > 0).sum()) == int(ref["nw_trap_drain"])
        assert int((keep["mask_w"][f] > 0).sum()) == int(ref["w_trap_drain"])
    return sat, band, neg


def test_drain512_weno5_theta30_fp64_matches_reference():
    mk = _mk512()
    path = mk.out_path("c2")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    ref = np.load(path)
    dt = np.float64
    print("\ncase: 512^3 WENO5/RK3 theta 30, 1 drainage step, fp64, vs reference", flush=True)
    g, mask, phi0, _ = mk.case_inputs("c2", dt)
    keep, outer = _drain_gpu(dt, g, mask, phi0, mk.case_options("c2", dt), int(ref["threads"]))
    _compare_with_fixture(keep, g, ref, "fp64 gpu", outer, exact=True)


def test_drain512_precision_pack_fp64_and_fp32_against_reference():
    mk = _mk512()
    path = mk.out_path("c3")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    ref = np.load(path)
    print("\ncase: 512^3 r U[10,20] WENO5/RK3 39.999 deg, 1 drainage step, fp64 + fp32 "
          "vs reference fp64", flush=True)
    g64, mask, phi0, _ = mk.case_inputs("c3", np.float64)
    keep64, outer64 = _drain_gpu(np.float64, g64, mask, phi0, mk.case_options("c3", np.float64),
                                 int(ref["threads"]))
    _, _, neg64 = _compare_with_fixture(keep64, g64, ref, "fp64 gpu", outer64, exact=True)
    del keep64
    from lsmpqs_amd.abi import make_grid
    g32 = make_grid((mk.N,) * 3, np.float32)
    o32 = mk.case_options("c3", np.float32)
    keep32, outer32 = _drain_gpu(np.float32, g32, mask.astype(np.float32),
                                 phi0.astype(np.float32), o32, int(ref["threads"]))
    assert abs(float(o32.overlap) - float(ref["overlap"])) < 1e-7
    sat32, band32, neg32 = _compare_with_fixture(keep32, g32, ref, "fp32 gpu", outer32,
                                                 exact=False)
    # interface voxels of the FP64 state (a face neighbour in the other
    # phase), which the fixture pins through its counts and band samples
    f = fill(g64)
    n64 = neg64[f]
    iface = np.zeros_like(n64)
    for ax in range(3):
        d = n64 != np.roll(n64, 1, axis=ax)
        iface |= d | np.roll(d, -1, axis=ax)
    mism = int((neg32[f] != n64).sum())
    print(f"  fp32: voxels in the other phase {mism} of {int(iface.sum())} interface voxels "
          f"({mism / max(1, int(iface.sum())):.2e})", flush=True)
    assert abs(sat32 - float(ref["sat_drain"])) <= 1e-4
    assert mism <= 0.01 * int(iface.sum())


def test_fp32_against_fp64_at_1024():
    """FP32 and FP64 GPU runs of two drainage curvature steps on the 1024^3
    pack (radius U[10,20], porosity 0.35, seed 1003, 39.999 deg, pseudo-time
    5 per step): saturation within 1e-4; band phi L-inf and interface
    mismatch reported."""
    from lsmpqs_amd import workloads as W
    from lsmpqs_amd.abi import make_grid
    from lsmpqs_amd.inputs import sphere_pack_mask, initial_phi
    n = 1024
    spec = W.WENO5_PRECISION
    print("\ncase: 1024^3 r U[10,20] WENO5/RK3 39.999 deg, 2 drainage steps, fp32 vs fp64",
          flush=True)
    g64 = make_grid((n, n, n), np.float64)
    # the pack is built on the GPU (same seeded spheres; 1.1e9 voxels)
    mask, _, radii, por = sphere_pack_mask(g64, spec["radius"], spec["porosity"], spec["seed"],
                                           device="cuda")
    phi0 = initial_phi(g64, mask)
    print(f"  pack: {len(radii)} spheres, porosity {por:.4f}", flush=True)
    states = {}
    for dt in (np.float32, np.float64):
        g = make_grid((n, n, n), dt)
        o = W.weno5_options(spec, dt, steps=2)
        o.tmax = 5.0
        keep, outer = _drain_gpu(dt, g, mask, phi0, o, 16)
        states[np.dtype(dt).name] = (keep["phi"], keep["mask"], outer)
        print(f"  {np.dtype(dt).name}: outer steps {outer}", flush=True)
    del phi0
    p32, m32, o32 = states["float32"]
    p64, m64, o64 = states["float64"]
    f = fill(g64)
    s32 = saturation(p32, m32, g64)
    s64 = saturation(p64, m64, g64)
    a32 = p32[f].astype(np.float64)
    a64 = p64[f]
    band = np.abs(a64) < 3.0
    linf = float(np.abs(a32[band] - a64[band]).max())
    neg64 = a64 < 0
    iface = np.zeros_like(neg64)
    for ax in range(3):
        d = neg64 != np.roll(neg64, 1, axis=ax)
        iface |= d | np.roll(d, -1, axis=ax)
    mism = int(((a32 < 0) != neg64).sum())
    print(f"  S_nw fp32 {s32:.8f} fp64 {s64:.8f} |dS| {abs(s32 - s64):.2e}; band phi L-inf "
          f"{linf:.3e}; voxels in the other phase {mism} of {int(iface.sum())} interface voxels",
          flush=True)
    assert o32 == o64
    assert abs(s32 - s64) <= 1e-4
