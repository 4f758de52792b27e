"""Batched drainage/imbibition of uneven boxes (lsmpqs_batch_model3d): boxes of
different sizes, each with its own curvature schedule, drained through one
persistent brick-queue stage kernel per RK stage inside a replayed HIP graph,
must end bit-identical to the same boxes drained one at a time with
drain_model3d / imbibe_model3d (phi and both trapped-phase masks), in FP32 and
FP64.  HJ-ENO3 + TVD-RK3, theta 0, trapping on (the batched case of the
reference's drain / imbibe programs)."""
import ctypes as C
import os

import numpy as np
import pytest

from lsmpqs_amd import hiplib
from lsmpqs_amd.abi import Arrays
from lsm_cases import case, options, DEVNULL

SHAPES = [(20, 24, 18), (28, 20, 22), (16, 30, 24), (24, 24, 24), (18, 22, 30)]
STARTS = [0.02, 0.025, 0.018, 0.03, 0.022]


def boxes(dt, theta_deg=0.0):
    out = []
    for i, (shape, a) in enumerate(zip(SHAPES, STARTS)):
        g, mask, phi0 = case(shape, dt, seed=2000 + i)
        o = options(dt, space=3, time=3, theta_deg=theta_deg, a=a, b=0.04, dc=0.25,
                    steps=3 + (i % 2), tmax=15.0)
        A = Arrays(g, dt)
        A["mask"] = mask
        A["phi"] = phi0
        out.append((g, o, A))
    return out


def run_single(lib, items, drain):
    for g, o, A in items:
        (lib.drain_model3d if drain else lib.imbibe_model3d)(C.byref(o), A.ptr(), C.byref(g), DEVNULL)
        hiplib.check(lib)
        lib.lsmpqs_release(A.ptr())


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("drain", [True, False])
@pytest.mark.parametrize("theta_deg", [0.0, 45.0])
def test_batched_boxes_match_single_runs(dt, drain, theta_deg):
    """theta 0: constCurvModel3dNoVar's path; 45 deg: the contact-angle path,
    whose imbibition also zeroes the coefficients of trapped wetting voxels
    in the outer steps."""
    print(f"case: batch {np.dtype(dt).name} drain={drain} theta={theta_deg} "
          f"boxes={len(SHAPES)}", flush=True)
    lib = hiplib.load(dt)
    lib.lsmpqs_set_dt_slabs(4)
    os.environ["LSMPQS_NO_STEP_FILES"] = "1"
    single, batched = boxes(dt, theta_deg), boxes(dt, theta_deg)
    if not drain:  # imbibe from a partly drained state
        run_single(lib, single, True)
        run_single(lib, batched, True)
        for items in (single, batched):
            for g, o, A in items:
                o.a = o.a + o.b * o.dc * 2
    run_single(lib, single, drain)
    n = len(batched)
    opts = (C.c_void_p * n)(*[C.addressof(o) for _, o, _ in batched])
    recs = (C.c_void_p * n)(*[C.addressof(A.rec) for _, _, A in batched])
    grids = (C.c_void_p * n)(*[C.addressof(g) for g, _, _ in batched])
    st = hiplib.BatchStats()
    rc = lib.lsmpqs_batch_model3d(n, opts, recs, grids, None, 1 if drain else 0, C.byref(st))
    assert rc == 0, hiplib.last_error(lib)
    for _, _, A in batched:
        lib.lsmpqs_release(A.ptr())
    assert st.boxes == n and st.batched_steps > 0 and st.box_steps >= st.batched_steps
    assert st.brick_depth in (8, 16, 32, 64) and st.workgroups > 0
    assert st.batched_outer > 0 and st.box_outer >= st.batched_outer
    print(f"batched steps {st.batched_steps}, box steps {st.box_steps}, brick depth "
          f"{st.brick_depth}, workgroups {st.workgroups}; batched outer steps "
          f"{st.batched_outer} serving {st.box_outer}", flush=True)
    for (g1, o1, A1), (g2, o2, A2) in zip(single, batched):
        for f in ("phi", "mask_w", "mask_nw"):
            assert np.array_equal(A1[f], A2[f]), f


@pytest.mark.gpu
@pytest.mark.parametrize("groups", [1, 2, 3])
def test_box_groups_do_not_change_results(groups, monkeypatch):
    """The batch runs its boxes in groups of consecutive indices, each with its
    own stream and rounds (LSMPQS_BATCH_GROUPS, default 4); any grouping gives
    the single runs' results bit for bit (FP32 drainage on the contact-angle
    path)."""
    dt = np.float32
    print(f"case: batch fp32 drain theta=45 groups={groups} boxes={len(SHAPES)}", flush=True)
    monkeypatch.setenv("LSMPQS_BATCH_GROUPS", str(groups))
    lib = hiplib.load(dt)
    lib.lsmpqs_set_dt_slabs(4)
    os.environ["LSMPQS_NO_STEP_FILES"] = "1"
    single, batched = boxes(dt, 45.0), boxes(dt, 45.0)
    run_single(lib, single, True)
    n = len(batched)
    opts = (C.c_void_p * n)(*[C.addressof(o) for _, o, _ in batched])
    recs = (C.c_void_p * n)(*[C.addressof(A.rec) for _, _, A in batched])
    grids = (C.c_void_p * n)(*[C.addressof(g) for g, _, _ in batched])
    st = hiplib.BatchStats()
    rc = lib.lsmpqs_batch_model3d(n, opts, recs, grids, None, 1, C.byref(st))
    assert rc == 0, hiplib.last_error(lib)
    for _, _, A in batched:
        lib.lsmpqs_release(A.ptr())
    assert st.boxes == n and st.box_steps >= st.batched_steps > 0
    for (g1, o1, A1), (g2, o2, A2) in zip(single, batched):
        for f in ("phi", "mask_w", "mask_nw"):
            assert np.array_equal(A1[f], A2[f]), f


@pytest.mark.gpu
def test_box_with_invalid_input_fails_the_batch_without_stalling_it():
    """A box below the drivers' minimum size fails on the host before any
    device work; the other boxes still run to their end (their groups' rounds
    continue without it) and the batch returns -2 with that box's message."""
    from lsmpqs_amd.abi import make_grid
    dt = np.float32
    print("case: batch fp32 with one undersized box", flush=True)
    lib = hiplib.load(dt)
    lib.lsmpqs_set_dt_slabs(4)
    os.environ["LSMPQS_NO_STEP_FILES"] = "1"
    items = boxes(dt)
    g_bad = make_grid((3, 3, 3), dtype=dt)
    A_bad = Arrays(g_bad, dt)
    shape = tuple(g_bad.grid_dims_ghostbox[i] for i in (2, 1, 0))
    A_bad["mask"] = -np.ones(shape, dtype=dt)
    A_bad["phi"] = np.ones(shape, dtype=dt)
    o_bad = options(dt, space=3, time=3)
    items.insert(2, (g_bad, o_bad, A_bad))
    n = len(items)
    opts = (C.c_void_p * n)(*[C.addressof(o) for _, o, _ in items])
    recs = (C.c_void_p * n)(*[C.addressof(A.rec) for _, _, A in items])
    grids = (C.c_void_p * n)(*[C.addressof(g) for g, _, _ in items])
    st = hiplib.BatchStats()
    rc = lib.lsmpqs_batch_model3d(n, opts, recs, grids, None, 1, C.byref(st))
    msg = hiplib.last_error(lib)
    for _, _, A in items:
        lib.lsmpqs_release(A.ptr())
    assert rc == -2 and "box 2" in msg and "4 fill voxels" in msg, msg
    assert st.box_steps > 0  # the valid boxes ran


@pytest.mark.gpu
def test_sixty_four_subvolumes_of_a_768_pack():
    """The full batched case: 64 subvolumes (dims uniform in {48..192} per
    axis) of one seeded 768^3 pack (radius U[6,12], porosity 0.36, seed 1004),
    each with its own 15-step curvature schedule (kappa from U[0.03,0.08] to
    U[0.2,0.4], b = 0.04), FP32 HJ-ENO3 + TVD-RK3, trapping on, ENO3 read
    quirk off.  The batched drainage equals the 64 single-box runs bit for bit
    (phi and both trapped-phase masks)."""
    from lsmpqs_amd.inputs import subvolume_batch
    from lsmpqs_amd.abi import default_options, make_grid, types_for
    dt = np.float32
    print("case: batch 64 subvolumes of a 768^3 pack, ENO3 RK3 fp32", flush=True)
    lib = hiplib.load(dt)
    boxes = subvolume_batch(seed=1004, dtype=dt)
    _, _, DataArrays = types_for(dt)

    def records():
        out = []
        for bx in boxes:
            # copies: drain_model3d writes phi and the reinitialized mask back
            keep = {"phi": bx["phi0"].copy(), "mask": bx["mask"].copy(),
                    "mask_w": np.zeros_like(bx["mask"]), "mask_nw": np.zeros_like(bx["mask"])}
            rec = DataArrays()
            for f, arr in keep.items():
                setattr(rec, f, arr.ctypes.data)
            o = default_options(dt)
            o.order_space_accur, o.order_time_accur = 3, 3
            ks, ke = bx["kappa"]
            o.b = bx["b"]
            o.a = ks * o.b
            o.dc = (ke - ks) / 14
            o.amax = ke * o.b + 1e-3 * o.b * o.dc
            o.tmax = 20.0
            out.append((make_grid(tuple(bx["dims"]), dt), o, rec, keep))
        return out

    old = lib.lsmpqs_set_quirks(0)
    lib.lsmpqs_set_quirks(old & ~1)
    lib.lsmpqs_set_dt_slabs(16)
    os.environ["LSMPQS_NO_STEP_FILES"] = "1"
    try:
        single, batched = records(), records()
        for g, o, rec, _ in single:
            lib.drain_model3d(C.byref(o), C.byref(rec), C.byref(g), DEVNULL)
            hiplib.check(lib)
            lib.lsmpqs_release(C.byref(rec))
        n = len(batched)
        opts = (C.c_void_p * n)(*[C.addressof(o) for _, o, _, _ in batched])
        recs = (C.c_void_p * n)(*[C.addressof(r) for _, _, r, _ in batched])
        grids = (C.c_void_p * n)(*[C.addressof(g) for g, _, _, _ in batched])
        st = hiplib.BatchStats()
        rc = lib.lsmpqs_batch_model3d(n, opts, recs, grids, None, 1, C.byref(st))
        assert rc == 0, hiplib.last_error(lib)
        for _, _, rec, _ in batched:
            lib.lsmpqs_release(C.byref(rec))
    finally:
        lib.lsmpqs_set_quirks(old)
    print(f"batched steps {st.batched_steps}, box steps {st.box_steps}, brick depth "
          f"{st.brick_depth}, workgroups {st.workgroups}", flush=True)
    assert st.boxes == 64 and st.box_steps > st.batched_steps
    moved = 0
    for (_, _, _, k1), (_, _, _, k2), bx in zip(single, batched, boxes):
        for f in ("phi", "mask_w", "mask_nw"):
            assert np.array_equal(k1[f], k2[f]), f
        moved += int(np.any(k1["phi"] != bx["phi0"]))
    assert moved == 64
