"""GPU parity: the HIP library (through its C ABI) against the oracle on the
same seeded inputs.  FP64 paths must agree to 1e-10 max-abs on phi and the
trapped masks (bit-exact wherever no transcendental function is involved);
labels and trapped partitions bit-exact."""
import ctypes as C

import numpy as np
import pytest

from lsm_cases import case, arrays, options, run_model, maxdiff, saturation, DEVNULL, fill
from lsmpqs_amd import hiplib
import oracle_py as op

pytestmark = pytest.mark.gpu

TOL64 = 1e-10

DRAIN = [
    ("eno2_rk2_theta0", dict(space=2, time=2, theta_deg=0)),
    ("eno2_rk1_theta0", dict(space=2, time=1, theta_deg=0)),
    ("eno3c_rk3_theta0", dict(space=5, time=3, theta_deg=0)),
    ("weno5_rk3_theta30", dict(space=5, time=3, theta_deg=30)),
    ("eno3_rk3_theta45", dict(space=3, time=3, theta_deg=45)),
    ("eno2_rk2_theta60", dict(space=2, time=2, theta_deg=60)),
]


@pytest.mark.parametrize("name,opts", DRAIN, ids=[d[0] for d in DRAIN])
def test_drain_fp64_matches_oracle(name, opts):
    print(f"case drain {name} 32^3 fp64", flush=True)
    Ao, _, g = run_model("oracle", np.float64, 32, opts)
    Ah, _, _ = run_model("hip", np.float64, 32, opts)
    for f in ("phi", "mask_w", "mask_nw", "mask"):
        d = maxdiff(Ao[f], Ah[f])
        print(f"  {f}: max|diff| = {d:.3e} exact={np.array_equal(Ao[f], Ah[f])}")
        assert d <= TOL64, f"{name} {f} differs by {d}"
    assert (Ao["phi"] < 0).sum() == (Ah["phi"] < 0).sum()


def test_imbibe_after_drain_fp64():
    opts = dict(space=5, time=3, theta_deg=30)
    print("case drain+imbibe weno5 theta30 32^3 fp64", flush=True)
    Ao, _, _ = run_model("oracle", np.float64, 32, opts, imbibe_after=True)
    Ah, _, _ = run_model("hip", np.float64, 32, opts, imbibe_after=True)
    for f in ("phi", "mask_w", "mask_nw", "normal_velocity"):
        d = maxdiff(Ao[f], Ah[f])
        print(f"  {f}: {d:.3e}")
        assert d <= TOL64


@pytest.mark.parametrize("T", [1, 3, 7])
def test_dt_slabs_follow_thread_count(T):
    opts = dict(space=2, time=2, theta_deg=0)
    print(f"case slabs T={T}", flush=True)
    Ao, _, _ = run_model("oracle", np.float64, 24, opts, T=T)
    Ah, _, _ = run_model("hip", np.float64, 24, opts, T=T)
    assert maxdiff(Ao["phi"], Ah["phi"]) <= TOL64


def test_noncubic_grid():
    opts = dict(space=2, time=3, theta_deg=30)
    print("case noncubic (40,28,20)", flush=True)
    Ao, _, _ = run_model("oracle", np.float64, (40, 28, 20), opts)
    Ah, _, _ = run_model("hip", np.float64, (40, 28, 20), opts)
    for f in ("phi", "mask_w", "mask_nw"):
        assert maxdiff(Ao[f], Ah[f]) <= TOL64


def _random_bin(shape, p, seed):
    rng = np.random.default_rng(seed)
    return (rng.random(shape) < p).astype(np.uint8)


@pytest.mark.parametrize("p", [0.2, 0.35, 0.6])
def test_labels_bit_exact(p):
    print(f"case labels p={p}", flush=True)
    g, mask, phi0 = case(20)
    lib = hiplib.load(np.float64)
    res = []
    for kind in ("oracle", "hip"):
        A = arrays(g, mask, phi0, np.float64)
        A["phi_bin"] = _random_bin(A.shape, p, 7)
        if kind == "oracle":
            op.oracle_lib(np.float64).orc_findConnectivity3d(A.ptr(), C.byref(g))
        else:
            lib.findConnectivity3d(A.ptr(), C.byref(g))
            hiplib.check(lib)
            lib.lsmpqs_release(A.ptr())
        res.append(A["connectivity"].copy())
    assert np.array_equal(res[0], res[1])


@pytest.mark.parametrize("tmax_r", [None, 1.6])
def test_reinit_and_bc_bit_exact(tmax_r):
    """tmax_r None: the reference default (0.4, one iteration at dx = 1, which
    the library runs as one fused pass); 1.6: four iterations from a copy of
    phi0 and a stored D0 field."""
    print(f"case reinit+bc tmax_r={tmax_r}", flush=True)
    g, mask, phi0 = case(32)
    lib = hiplib.load(np.float64)
    orc = op.oracle_lib(np.float64)
    o = options(np.float64)
    if tmax_r is not None:
        o.tmax_r = tmax_r
    a = mask.copy()
    b = mask.copy()
    orc.orc_reinitializeSubcellFix3d(a.ctypes.data_as(C.c_void_p), C.byref(g), C.byref(o))
    lib.reinitializeSubcellFix3d(b.ctypes.data_as(C.c_void_p), C.byref(g), C.byref(o))
    hiplib.check(lib)
    assert np.array_equal(a, b)
    rng = np.random.default_rng(3)
    a = rng.normal(size=mask.shape)
    b = a.copy()
    orc.orc_signedLinearExtrapolationBCqss(a.ctypes.data_as(C.c_void_p), C.byref(g), 9)
    lib.signedLinearExtrapolationBCqss(b.ctypes.data_as(C.c_void_p), C.byref(g), 9)
    hiplib.check(lib)
    assert np.array_equal(a, b)


def test_trap_components_bit_exact():
    print("case trap", flush=True)
    # a drained state with trapped blobs: run the oracle drain, then trap once on both
    Ao, o, g = run_model("oracle", np.float64, 32, dict(space=2, time=2, theta_deg=0))
    lib = hiplib.load(np.float64)
    outs = []
    for kind in ("oracle", "hip"):
        A = arrays(g, Ao["mask"], Ao["phi"], np.float64)
        A["mask_nw"] = Ao["mask_nw"]
        A["mask_w"] = Ao["mask_w"]
        if kind == "oracle":
            op.oracle_lib(np.float64).orc_trapComponents_mask(A.ptr(), C.byref(g), C.byref(o))
        else:
            lib.trapComponents_mask(A.ptr(), C.byref(g), C.byref(o))
            hiplib.check(lib)
            lib.lsmpqs_release(A.ptr())
        outs.append(A)
    for f in ("phi", "mask_nw", "mask_w", "connectivity", "phi_bin"):
        assert np.array_equal(outs[0][f], outs[1][f]), f


FP32_CASES = [
    ("weno5_rk3_theta30", dict(space=5, time=3, theta_deg=30)),
    ("eno2_rk2_theta0", dict(space=2, time=2, theta_deg=0)),
    ("eno3_rk3_theta45", dict(space=3, time=3, theta_deg=45)),
]


@pytest.mark.parametrize("name,opts", FP32_CASES, ids=[c[0] for c in FP32_CASES])
def test_fp32_fast_path_against_fp64_oracle(name, opts):
    """The FP32 library (fast WENO5 arithmetic, FP64 reductions) against the
    FP64 oracle: saturation within 1e-4, errors reported for saturation, phi
    L-infinity in the |phi| < 3 band and interface voxels."""
    print(f"case fp32 fast {name}", flush=True)
    A64, _, g = run_model("oracle", np.float64, 32, opts)
    Ah, _, g32 = run_model("hip", np.float32, 32, opts)
    s64 = saturation(A64["phi"], A64["mask"], g)
    sh = saturation(Ah["phi"], Ah["mask"], g32)
    f = fill(g)
    band = np.abs(A64["phi"][f]) < 3.0
    linf = float(np.max(np.abs(A64["phi"][f][band] - Ah["phi"][f][band].astype(np.float64))))
    mism = int(np.sum((A64["phi"][f] < 0) != (Ah["phi"][f] < 0)))
    print(f"  saturation fp64 {s64:.8f} fp32 {sh:.8f} |dS|={abs(s64 - sh):.2e}; "
          f"band Linf {linf:.2e}; interface mismatch {mism} voxels")
    assert abs(s64 - sh) < 1e-4
    assert mism <= max(10, band.sum() // 1000)


def test_fp32_exact_order_matches_fp32_oracle():
    """With the exact-arithmetic switch the FP32 library evaluates the
    reference's expressions in order and follows the FP32 oracle."""
    opts = dict(space=5, time=3, theta_deg=30)
    print("case fp32 exact-order weno5", flush=True)
    lib = hiplib.load(np.float32)
    old = lib.lsmpqs_set_quirks(1 | 2 | 4 | 8)
    try:
        Ah, _, g32 = run_model("hip", np.float32, 32, opts)
    finally:
        lib.lsmpqs_set_quirks(old)
    A32, _, _ = run_model("oracle", np.float32, 32, opts)
    s32 = saturation(A32["phi"], A32["mask"], g32)
    sh = saturation(Ah["phi"], Ah["mask"], g32)
    print(f"  saturation fp32-oracle {s32:.8f} fp32-gpu-exact {sh:.8f}")
    assert abs(sh - s32) < 1e-5


def _canonical_labels(binv):
    """Largest linear index of each 26-connected component (0 on background),
    from scipy's labelling."""
    from scipy import ndimage
    lab, n = ndimage.label(binv, structure=np.ones((3, 3, 3)))
    flat = lab.ravel()
    top = np.full(n + 1, -1, dtype=np.int64)
    np.maximum.at(top, flat, np.arange(flat.size))
    return np.where(flat > 0, top[flat], 0).reshape(binv.shape).astype(np.int32)


@pytest.mark.parametrize("shape", [(96, 187, 90), (178, 107, 122)])
def test_labels_of_blob_fields_are_stable(shape):
    """Labels of thresholded smooth random fields (long, branching components
    across many tiles), five times each, against scipy's partition with the
    reference's root rule.  The final compression pass once let a finished
    entry be re-pointed at a non-root ancestor by another thread's path
    halving; every run must now give the exact labels."""
    from scipy.ndimage import gaussian_filter
    from lsmpqs_amd.abi import make_grid, Arrays
    print(f"case labels blobs {shape}", flush=True)
    lib = hiplib.load(np.float32)
    g = make_grid(shape, np.float32)
    rng = np.random.default_rng(sum(shape))
    for thr in (0.0, 0.4):
        A = Arrays(g, np.float32)
        f = gaussian_filter(rng.standard_normal(A.shape), 2.0)
        A["phi_bin"] = (f / f.std() > thr).astype(np.uint8)
        want = _canonical_labels(A["phi_bin"])
        for rep in range(5):
            A["connectivity"] = -7
            lib.findConnectivity3d(A.ptr(), C.byref(g))
            hiplib.check(lib)
            assert np.array_equal(A["connectivity"], want), (thr, rep)
        lib.lsmpqs_release(A.ptr())


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_graph_replayed_steps_equal_direct_launches(dt, monkeypatch):
    """Inner steps replayed from captured HIP graphs (LSMPQS_GRAPHS=1)
    against the same steps launched directly (the default): drainage then
    imbibition, RK3 and "RK2" (phi and phi_prev swap buffers every outer
    step, so the cached graphs rotate), bit for bit in phi and the trapped
    masks; every replay is followed by the drivers' own eager kernels and
    synchronizations."""
    print(f"case graphs vs direct launches {np.dtype(dt).name}", flush=True)
    for opts in (dict(space=5, time=3, theta_deg=30), dict(space=2, time=2, theta_deg=0)):
        res = []
        for graphs in ("0", "1"):
            monkeypatch.setenv("LSMPQS_GRAPHS", graphs)
            A, _, _ = run_model("hip", dt, 24, opts, imbibe_after=True)
            res.append(A)
        for f in ("phi", "mask_w", "mask_nw"):
            assert np.array_equal(res[0][f], res[1][f]), (opts, f)
