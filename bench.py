#!/usr/bin/env python3
"""Throughput of the LSMPQS level-set hot path on one MI355X.

A step is one inner time step of constCurvModel3d on a 512^3 seeded sphere
pack (radius U[6,12], porosity 0.36, seed 1002): HJ-WENO5 + TVD-RK3 with the
contact-angle (theta = 30 deg) variable-coefficient path and the trapped-phase
masks imposed, i.e. time-step probe, three fused RHS+RK stage kernels, ghost
extrapolation and mask imposition.  The value is voxel-updates/s of the FP32
library; the FP64 library is measured on the same input and reported beside
it.  Inputs are device-resident before the timed region (set up by the
library's own drain_model3d set-up on the GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--no-cpu]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "lsmpqs-1.0_amd"), os.path.join(ROOT, "oracle")]

from lsmpqs_amd import hiplib  # noqa: E402
from lsmpqs_amd.abi import types_for, make_grid, default_options  # noqa: E402
from lsmpqs_amd.inputs import sphere_pack_mask, initial_phi  # noqa: E402

HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12   # unpacked vector lane-ops, T/s
THETA_DEG = 30.0


def lean_arrays(dtype, mask, phi0):
    """QSS_DataArrays record with only phi and mask (the rest NULL), on copies:
    the model drivers write both back (mask reinitialized), as the reference's do."""
    _, _, DataArrays = types_for(dtype)
    rec = DataArrays()
    keep = [np.array(phi0, dtype=dtype, order="C"), np.array(mask, dtype=dtype, order="C")]
    rec.phi = keep[0].ctypes.data
    rec.mask = keep[1].ctypes.data
    return rec, keep


def options(dtype, a, b):
    o = default_options(dtype)
    o.order_space_accur = 5
    o.order_time_accur = 3
    o.theta = THETA_DEG * 3.14159 / 180
    o.a = a
    o.b = b
    o.check_connectivity = 1
    return o


def gpu_leg(dtype, g64, mask, phi0, steps, warmup, a, b):
    lib = hiplib.load(dtype)
    g = make_grid(tuple(g64.grid_dims), dtype)
    rec, keep = lean_arrays(dtype, mask, phi0)
    o = options(dtype, a, b)
    lib.lsmpqs_set_dt_slabs(16)
    if lib.lsmpqs_model_prepare(C.byref(o), C.byref(rec), C.byref(g), 1) != 0:
        raise hiplib.LsmpqsError(hiplib.last_error(lib))
    t = hiplib.Timing()
    if warmup:
        if lib.lsmpqs_bench_inner_steps(C.byref(o), C.byref(rec), C.byref(g), warmup, 1, a, 0,
                                        C.byref(t)) != 0:
            raise hiplib.LsmpqsError(hiplib.last_error(lib))
    t = hiplib.Timing()
    w0 = time.perf_counter()
    rc = lib.lsmpqs_bench_inner_steps(C.byref(o), C.byref(rec), C.byref(g), steps, 1, a, 0,
                                      C.byref(t))
    wall = time.perf_counter() - w0
    if rc != 0:
        raise hiplib.LsmpqsError(hiplib.last_error(lib))
    # per-kernel durations in a separate pass (events around each stage launch)
    tp = hiplib.Timing()
    tp.profile = 1
    if lib.lsmpqs_bench_inner_steps(C.byref(o), C.byref(rec), C.byref(g), max(3, steps // 2), 1, a,
                                    0, C.byref(tp)) != 0:
        raise hiplib.LsmpqsError(hiplib.last_error(lib))
    lib.lsmpqs_release(C.byref(rec))
    nvox = g64.grid_dims[0] * g64.grid_dims[1] * g64.grid_dims[2]
    return dict(elapsed_ms=t.elapsed_ms, wall_s=wall, nvox=nvox, last_dt=t.last_dt,
                stage_avg_ms=[tp.stage_ms[i] / max(1, tp.stage_launches[i]) for i in range(3)],
                nstale=t.nstale)


def counter_record(n, key):
    """Per-launch counters of the last-stage kernel at n^3 from the committed
    rocprofv3 summary (profiles/weno5_rk3_512_counters.md, its json block):
    HBM bytes = FETCH_SIZE x fetch_scale + WRITE_SIZE x write_scale (KiB,
    scales from a plain copy kernel), VALU wave instructions."""
    path = os.path.join(ROOT, "profiles", "weno5_rk3_512_counters.md")
    try:
        text = open(path).read()
        blob = json.loads(text.rsplit("```json", 1)[1].split("```", 1)[0])
    except (OSError, IndexError, ValueError):
        return None
    rec = blob.get(key)
    if not rec or rec.get("grid") != n:
        return None
    return rec


def stage_roofline(res, itemsize, key, n):
    """Roofline of the last-stage kernel (stage 3 with imposition): its
    algorithmic traffic per fill voxel is 6 fields (phi^(2), phi^n, mask,
    mask_w, mask_nw read; phi^(n+1) written), the compulsory traffic with the
    contact-angle coefficients derived from the mask.  `traffic` is the HBM
    bytes per launch the counters measured (the library also streams the five
    stored coefficient fields and the imposition floor); `valu` is the vector
    ALU issue rate against 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz."""
    per_voxel = 6 * itemsize
    t = res["stage_avg_ms"][2] * 1e-3
    ach = per_voxel * res["nvox"] / t / 1e9
    out = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
           "kernel": "stage_kernel<WENO5,var,stage3,final>",
           "alg_bytes_per_launch": per_voxel * res["nvox"]}
    rec = counter_record(n, key)
    if rec:
        traffic = (rec["fetch_kib"] * rec["fetch_scale"] + rec["write_kib"] * rec["write_scale"]) * 1024
        out["traffic"] = int(traffic)
        out["traffic_rate_gbs"] = round(traffic / t / 1e9, 1)
        lane_ops = rec["valu_insts"] * 64
        out["valu"] = {"achieved_tops": round(lane_ops / t / 1e12, 2), "peak_tops": VALU_PEAK_TOPS,
                       "frac": round(lane_ops / t / 1e12 / VALU_PEAK_TOPS, 4),
                       "unit": "T lane-ops/s", "insts_per_voxel": round(lane_ops / res["nvox"], 1)}
    return out


def curve_leg(dtype, g64, mask, phi0, curv_steps, tmax, a, b, step_files=False):
    """Drainage-curve time: drain_model3d end to end on the GPU (set-up,
    curv_steps curvature steps, each running outer steps until the reference's
    stopping tests fire or pseudo-time tmax; step files off), host wall time
    from host arrays in to host arrays out."""
    lib = hiplib.load(dtype)
    g = make_grid(tuple(g64.grid_dims), dtype)
    rec, keep = lean_arrays(dtype, mask, phi0)
    o = options(dtype, a, b)
    o.dc = 0.25
    o.amax = a + b * o.dc * (curv_steps - 0.5)
    o.tmax = tmax
    lib.lsmpqs_set_dt_slabs(16)
    os.environ["LSMPQS_NO_STEP_FILES"] = "0" if step_files else "1"
    cwd = os.getcwd()
    if step_files:  # data_step_N.gz of every curvature step, into a scratch directory
        import tempfile
        os.chdir(tempfile.mkdtemp(prefix="lsm_steps_"))
    try:
        w0 = time.perf_counter()
        lib.drain_model3d(C.byref(o), C.byref(rec), C.byref(g), None)
        wall = time.perf_counter() - w0
    finally:
        os.chdir(cwd)
    hiplib.check(lib)
    st = hiplib.Timing()
    lib.lsmpqs_session_stats(C.byref(rec), C.byref(g), C.byref(st))
    lib.lsmpqs_release(C.byref(rec))
    return {"wall_s": round(wall, 3), "outer_steps": int(st.outer_steps),
            "inner_steps": int(st.inner_steps), "inner_loop_s": round(st.inner_loop_ms * 1e-3, 3),
            "nw_voxels": int((keep[0] < 0).sum())}


def adjoint_leg(dtype, n=128, steps=200, checkpoint_every=20):
    """Gradient of (S_w - 0.45)^2 with respect to a per-voxel contact-angle
    field (40 deg + Gaussian random field, sigma 5 deg, correlation length
    10; n^3 pack of radius 8, seed 1005) and b, through `steps` WENO5 TVD-RK3
    steps at a0/b = 0.15, b = 0.04: host wall time of the forward map alone
    and of forward + checkpointed reverse pass (lsmpqs_adjoint_run)."""
    from lsmpqs_amd.adjoint import saturation_loss
    from lsmpqs_amd.inputs import contact_angle_case
    g, mask, phi0, theta = contact_angle_case(n, dtype=dtype)
    o = default_options(dtype)
    o.order_space_accur, o.order_time_accur = 5, 3
    b = 0.04
    a0, target = 0.15 * b, 0.45
    saturation_loss(phi0, mask, theta, g, o, 2, a0, b, target, gradient=False)  # warm-up
    w0 = time.perf_counter()
    rf = saturation_loss(phi0, mask, theta, g, o, steps, a0, b, target, gradient=False)
    tf = time.perf_counter() - w0
    w0 = time.perf_counter()
    rg = saturation_loss(phi0, mask, theta, g, o, steps, a0, b, target,
                         checkpoint_every=checkpoint_every)
    tg = time.perf_counter() - w0
    nvox = n ** 3
    return {"grid": n, "steps": steps, "checkpoint_every": checkpoint_every,
            "forward_s": round(tf, 4), "gradient_s": round(tg, 4),
            "gradient_over_forward": round(tg / tf, 2),
            "forward_voxel_updates_per_s": round(nvox * steps / tf, 1),
            "sw": round(rg.sw, 8), "loss": rg.loss, "dloss_db": rg.dloss_db,
            "same_loss_as_forward": rf.loss == rg.loss}


def batch_leg(tmax=20.0, seed=1004, sequential=True):
    """64 subvolumes (dims U{48..192} per axis) of one seeded 768^3 pack,
    each with its own 15-step curvature schedule (kappa from U[0.03,0.08] to
    U[0.2,0.4], b = 0.04), FP32 HJ-ENO3 + TVD-RK3, theta 0, trapping on,
    ENO3 read quirk off (undefined for non-cubic boxes): one batched
    drainage run (lsmpqs_batch_model3d), and for comparison the same boxes
    one after another through drain_model3d."""
    from lsmpqs_amd.inputs import subvolume_batch
    dt_ = np.float32
    lib = hiplib.load(dt_)
    boxes = subvolume_batch(seed=seed, dtype=dt_)
    old = lib.lsmpqs_set_quirks(0)
    lib.lsmpqs_set_quirks(old & ~1)
    lib.lsmpqs_set_dt_slabs(16)
    os.environ["LSMPQS_NO_STEP_FILES"] = "1"

    def records():
        out = []
        for bx in boxes:
            rec, keep = lean_arrays(dt_, bx["mask"], bx["phi0"])
            o = default_options(dt_)
            o.order_space_accur, o.order_time_accur = 3, 3
            o.theta = 0
            ks, ke = bx["kappa"]
            o.b = bx["b"]
            o.a = ks * o.b
            o.dc = (ke - ks) / 14
            o.amax = ke * o.b + 1e-3 * o.b * o.dc
            o.tmax = tmax
            o.check_connectivity = 1
            g = make_grid(tuple(bx["dims"]), dt_)
            out.append((g, o, rec, keep))
        return out

    def voxel_updates(items):
        tot = 0
        for g, o, rec, _ in items:
            st = hiplib.Timing()
            lib.lsmpqs_session_stats(C.byref(rec), C.byref(g), C.byref(st))
            tot += int(st.inner_steps) * g.grid_dims[0] * g.grid_dims[1] * g.grid_dims[2]
        return tot

    libc = C.CDLL("libc.so.6")
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    devnull = C.c_void_p(libc.fopen(b"/dev/null", b"w"))  # the drivers' per-step log lines
    try:
        items = records()
        n = len(items)
        opts = (C.c_void_p * n)(*[C.addressof(o) for _, o, _, _ in items])
        recs = (C.c_void_p * n)(*[C.addressof(r) for _, _, r, _ in items])
        grids = (C.c_void_p * n)(*[C.addressof(g) for g, _, _, _ in items])
        st = hiplib.BatchStats()
        w0 = time.perf_counter()
        rc = lib.lsmpqs_batch_model3d(n, opts, recs, grids, None, 1, C.byref(st))
        wall = time.perf_counter() - w0
        if rc != 0:
            raise hiplib.LsmpqsError(hiplib.last_error(lib))
        vu = voxel_updates(items)
        for _, _, rec, _ in items:
            lib.lsmpqs_release(C.byref(rec))
        out = {"boxes": n, "voxels": int(sum(np.prod(b["dims"]) for b in boxes)),
               "batched_wall_s": round(wall, 3), "batched_gpu_s": round(st.batched_ms * 1e-3, 3),
               "batched_steps": int(st.batched_steps), "box_steps": int(st.box_steps),
               "brick_depth": int(st.brick_depth), "workgroups": int(st.workgroups),
               "voxel_updates": vu,
               "voxel_updates_per_s_batched_gpu": round(vu / (st.batched_ms * 1e-3), 1)}
        if sequential:
            items2 = records()
            w0 = time.perf_counter()
            for g, o, rec, _ in items2:
                lib.drain_model3d(C.byref(o), C.byref(rec), C.byref(g), devnull)
                hiplib.check(lib)
            out["sequential_wall_s"] = round(time.perf_counter() - w0, 3)
            out["same_result"] = all(np.array_equal(a[3][0], b[3][0])
                                     for a, b in zip(items, items2))
            for _, _, rec, _ in items2:
                lib.lsmpqs_release(C.byref(rec))
            out["speedup_over_sequential"] = round(out["sequential_wall_s"] / wall, 2)
        return out
    finally:
        lib.lsmpqs_set_quirks(old)


def precision_leg(n=512, curv_steps=40, tmax=20.0, seed=1003):
    """FP32 against FP64 over one drainage curve on the GPU: n^3 pack (radius
    U[10,20], porosity 0.35), WENO5 + TVD-RK3, theta 39.999 deg (both
    precisions then take overlap 0.3 dx; at exactly 40 deg the reference's
    threshold test rounds differently in the two precisions), trapping on,
    curv_steps curvature steps of 0.25 b from a = 0.02 (b = 0.04), pseudo-time
    tmax per step.  Reports wall time per precision and, on the final states,
    the saturation difference, phi L-infinity in the |phi| < 3 dx band of the
    FP64 state, the voxels whose phase (phi < 0) differs, and the trapped-mask
    differences."""
    g64 = make_grid((n, n, n), np.float64)
    mask, _, radii, por = sphere_pack_mask(g64, (10, 20), 0.35, seed)
    phi0 = initial_phi(g64, mask)
    res, states = {}, {}
    for key, dt_ in (("f32", np.float32), ("f64", np.float64)):
        lib = hiplib.load(dt_)
        g = make_grid((n, n, n), dt_)
        _, _, DataArrays = types_for(dt_)
        keep = {"phi": np.array(phi0, dtype=dt_), "mask": np.array(mask, dtype=dt_),
                "mask_w": np.zeros(mask.shape, dt_), "mask_nw": np.zeros(mask.shape, dt_)}
        rec = DataArrays()
        for f, arr in keep.items():
            setattr(rec, f, arr.ctypes.data)
        o = options(dt_, 0.02, 0.04)
        o.theta = 39.999 * 3.14159 / 180
        o.dc = 0.25
        o.amax = o.a + o.b * o.dc * (curv_steps - 0.5)
        o.tmax = tmax
        lib.lsmpqs_set_dt_slabs(16)
        os.environ["LSMPQS_NO_STEP_FILES"] = "1"
        w0 = time.perf_counter()
        lib.drain_model3d(C.byref(o), C.byref(rec), C.byref(g), None)
        wall = time.perf_counter() - w0
        hiplib.check(lib)
        st = hiplib.Timing()
        lib.lsmpqs_session_stats(C.byref(rec), C.byref(g), C.byref(st))
        lib.lsmpqs_release(C.byref(rec))
        res[key] = {"wall_s": round(wall, 3), "outer_steps": int(st.outer_steps),
                    "inner_steps": int(st.inner_steps),
                    "inner_loop_s": round(st.inner_loop_ms * 1e-3, 3), "overlap": float(o.overlap)}
        states[key] = keep
        print(f"# precision {key}: {res[key]}", file=sys.stderr, flush=True)
    fb = (slice(3, n + 3),) * 3
    p32 = states["f32"]["phi"][fb].astype(np.float64)
    p64 = states["f64"]["phi"][fb]
    m = mask[fb]

    def sat(p):
        q = p / 1.5
        v = np.where(p < -1.5, 1.0, np.where(p < 1.5, 0.5 * (1 - q - np.sin(np.pi * q) / np.pi), 0.0))
        qm = m / 1.5
        vm = np.where(m < -1.5, 1.0, np.where(m < 1.5, 0.5 * (1 - qm - np.sin(np.pi * qm) / np.pi), 0.0))
        return float(v.sum() / vm.sum())

    band = np.abs(p64) < 3.0
    neg64 = p64 < 0
    iface = np.zeros_like(neg64)
    for ax in range(3):
        d = neg64 != np.roll(neg64, 1, axis=ax)
        iface |= d | np.roll(d, -1, axis=ax)
    flips = int(((p32 < 0) != neg64).sum())
    trap_diff = ((states["f32"]["mask_nw"][fb] > 0) != (states["f64"]["mask_nw"][fb] > 0)) | \
                ((states["f32"]["mask_w"][fb] > 0) != (states["f64"]["mask_w"][fb] > 0))
    dev = np.abs(p32 - p64)
    same_trap = band & ~trap_diff
    out = {"grid": n, "spheres": int(len(radii)), "porosity": round(float(por), 4),
           "curvature_steps": curv_steps, "tmax": tmax, "f32": res["f32"], "f64": res["f64"],
           "fp64_over_fp32_wall": round(res["f64"]["wall_s"] / res["f32"]["wall_s"], 3),
           "sat_nw_f32": sat(p32), "sat_nw_f64": sat(p64),
           "sat_abs_diff": abs(sat(p32) - sat(p64)),
           "phi_linf_band": float(dev[band].max()),
           "phi_linf_band_same_trapping": float(dev[same_trap].max()) if same_trap.any() else 0.0,
           "phi_p999_band": float(np.quantile(dev[band], 0.999)),
           "band_voxels": int(band.sum()),
           "phase_mismatch_voxels": flips, "interface_voxels": int(iface.sum()),
           "trapped_nw_mismatch": int(((states["f32"]["mask_nw"][fb] > 0)
                                       != (states["f64"]["mask_nw"][fb] > 0)).sum()),
           "trapped_w_mismatch": int(((states["f32"]["mask_w"][fb] > 0)
                                      != (states["f64"]["mask_w"][fb] > 0)).sum())}
    return out


def disk_pack_2d(g, seed, porosity=0.55, rmin=6.0, rmax=14.0):
    """(mask, phi0) on a 2D grid: signed distance to a Boolean disk pack
    (> 0 in solid, clipped at -6), non-wetting phase in a slab of 12 dx at the
    x-low face with the solid imposed.  Arrays (ny, nx), i fastest."""
    rng = np.random.default_rng(seed)
    nx, ny = g.grid_dims_ghostbox[0], g.grid_dims_ghostbox[1]
    x = g.x_lo_ghostbox[0] + g.dx[0] * np.arange(nx)
    y = g.x_lo_ghostbox[1] + g.dx[1] * np.arange(ny)
    mask = np.full((ny, nx), -6.0)
    area = (g.x_hi[0] - g.x_lo[0]) * (g.x_hi[1] - g.x_lo[1])
    ndisk = int(np.ceil(-np.log(porosity) * area / (np.pi * (rmin * rmin + rmin * rmax + rmax * rmax) / 3)))
    for _ in range(ndisk):
        cx = rng.uniform(g.x_lo[0] + 16, g.x_hi[0])
        cy = rng.uniform(g.x_lo[1], g.x_hi[1])
        r = rng.uniform(rmin, rmax)
        i0, i1 = np.searchsorted(x, cx - r - 7), np.searchsorted(x, cx + r + 7)
        j0, j1 = np.searchsorted(y, cy - r - 7), np.searchsorted(y, cy + r + 7)
        d = r - np.hypot(x[None, i0:i1] - cx, y[j0:j1, None] - cy)
        np.maximum(mask[j0:j1, i0:i1], d, out=mask[j0:j1, i0:i1])
    phi0 = np.broadcast_to(x[None, :] - (g.x_lo[0] + 12.0), mask.shape)
    return mask, np.maximum(phi0, mask)


def model2d_leg(n=1024, curv_steps=2, tmax=4.0, seed=1006, cpu=True):
    """drain_model2d end to end on an n x n disk pack (HJ-ENO3 + TVD-RK3,
    theta 45 deg, trapping on, curv_steps curvature steps from a = 0.05 by
    0.02, pseudo-time tmax each, step files off): host wall time per
    precision on the GPU, and the reference build's drain_model2d (FP64,
    OpenMP) on the same input, with the largest difference of the final phi."""
    import tempfile
    from lsmpqs_amd.abi import Arrays
    out = {"grid": [n, n], "curvature_steps": curv_steps, "tmax": tmax}
    finals = {}
    libc = C.CDLL("libc.so.6")
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    devnull = C.c_void_p(libc.fopen(b"/dev/null", b"w"))  # both models write their log here

    def grid_of(lib, dt):
        Grid, _, _ = types_for(dt)
        R = C.c_double if dt == np.float64 else C.c_float
        lib.createGridSetDx.restype = C.POINTER(Grid)
        lib.createGridSetDx.argtypes = [C.c_int, R, C.POINTER(R), C.POINTER(R), C.c_int]
        return lib.createGridSetDx(2, 1.0, (R * 3)(0, 0, 0), (R * 3)(n, n, 0), 1).contents

    def opts(dt):
        o = default_options(dt)
        o.order_space_accur, o.order_time_accur = 3, 3
        o.theta = 45 * 3.14159 / 180
        o.a, o.b, o.dc = 0.05, 0.04, 0.5
        o.amax = o.a + o.b * o.dc * (curv_steps - 1) + 1e-3
        o.tmax = tmax
        return o

    def run(lib, dt, mask, phi0, threads_env):
        g = grid_of(lib, dt)
        A = Arrays(g, dt)
        A["mask"] = mask.reshape(A.shape)
        A["phi"] = phi0.reshape(A.shape)
        o = opts(dt)
        cwd = os.getcwd()
        os.chdir(tempfile.mkdtemp(prefix="lsm2d_bench_"))
        try:
            t0 = time.perf_counter()
            lib.drain_model2d(C.byref(o), A.ptr(), C.byref(g), devnull)
            el = time.perf_counter() - t0
        finally:
            os.chdir(cwd)
        return el, A["phi"].reshape(mask.shape).astype(np.float64)

    g64 = None
    for key in ("f32", "f64"):
        dt = np.float32 if key == "f32" else np.float64
        lib = hiplib.load(dt)
        if g64 is None:
            g64 = grid_of(hiplib.load(np.float64), np.float64)
            mask, phi0 = disk_pack_2d(g64, seed)
        lib.lsmpqs_set_dt_slabs(16)
        os.environ["LSMPQS_NO_STEP_FILES"] = "1"
        el, fin = run(lib, dt, mask.astype(dt), phi0.astype(dt), None)
        hiplib.check(lib)
        out[key + "_s"] = round(el, 3)
        finals[key] = fin
    if cpu:
        import oracle_py as op
        if op.have_ref(np.float64):
            threads = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
            op.set_omp_threads(threads)
            el, fin = run(op.ref_lib(np.float64), np.float64, mask, phi0, threads)
            out["reference_f64_s"] = round(el, 3)
            out["reference_threads"] = threads
            out["max_abs_diff_f64_vs_reference"] = float(np.max(np.abs(finals["f64"] - fin)))
    out["max_abs_diff_f32_vs_f64"] = float(np.max(np.abs(finals["f32"] - finals["f64"])))
    return out


def cpu_leg(n, steps=8):
    """The reference build (oracle/_ref, OpenMP, FP64) on a bounded sample:
    one outer step of constCurvModel3d at n^3 whose pseudo-time window holds
    `steps` inner steps; the time is the reference's own inner-loop timer
    (the "Level set time" line it writes to its log)."""
    import tempfile
    import oracle_py as op
    from lsmpqs_amd.abi import Arrays
    dt = np.float64
    if not op.have_ref(dt):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    op.set_omp_threads(threads)
    lib = op.ref_lib(dt)
    g = make_grid((n, n, n), dt)
    mask, _, _, _ = sphere_pack_mask(g, (6, 12), 0.36, 1002)
    A = Arrays(g, dt)
    A["mask"] = mask
    A["phi"] = initial_phi(g, mask)
    o = options(dt, 0.02, 0.04)
    gb, fb = op.gb6(g), op.fb6(g)
    op.fcall(lib, "getcentralgrad3dorder2_", dt, A["mask_x"], A["mask_y"], A["mask_z"], *gb,
             A["mask"], *gb, *fb, 1.0, 1.0, 1.0)
    bm, um = op.Scalar(0.0, C.c_double), op.Scalar(0.0, C.c_double)
    op.fcall(lib, "setvarcurvadv3d_", dt, A["mask"], A["mask_x"], A["mask_y"], A["mask_z"],
             A["curvature_coeff"], A["external_velocity_x"], A["external_velocity_y"],
             A["external_velocity_z"], bm, um, 0.04, *gb, *fb, 1.0)
    op.fcall(lib, "setvarnorm3d_", dt, A["mask"], A["mask_x"], A["mask_y"], A["mask_z"],
             A["normal_velocity"], 0.02, float(o.theta), *gb, *fb, 1.0)
    o.b_max_over_dx, o.max_U_over_dx = bm.value, um.value
    # first time step exactly as the reference forms it: H/dx at each slab's last voxel
    nsl = g.khi_fb - g.klo_fb + 1
    H = -1.0
    for t in range(threads):
        klo = g.klo_fb + nsl * t // threads
        khi = min(g.klo_fb + nsl * (t + 1) // threads - 1, g.khi_fb)
        if klo > khi:
            continue
        mh = op.Scalar(0.0, C.c_double)
        rhs = np.zeros_like(A["phi"])
        op.fcall(lib, "getrhs3d5th_", dt, A["phi"], A["normal_velocity"], A["curvature_coeff"],
                 A["external_velocity_x"], A["external_velocity_y"], A["external_velocity_z"], mh,
                 rhs, *gb, g.ilo_fb, g.ihi_fb, g.jlo_fb, g.jhi_fb, klo, khi, 1.0, 1.0, 1.0)
        H = max(H, mh.value)
    dt1 = 0.5 / (H + o.b_max_over_dx + o.max_U_over_dx)
    o.check_connectivity = 0
    lib.setSpaceDerivFunc(C.byref(g), C.byref(o))
    o.tmax = 1e-30
    o.tplot = dt1 * (steps - 0.5)
    log = tempfile.NamedTemporaryFile(prefix="lsm_ref_", suffix=".log", delete=False).name
    libc = C.CDLL("libc.so.6")
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    fp = libc.fopen(log.encode(), b"w")
    lib.constCurvModel3d.restype = C.c_double
    lib.constCurvModel3d(C.byref(o), A.ptr(), C.byref(g), C.c_void_p(fp))
    libc.fclose(C.c_void_p(fp))
    secs = None
    for line in open(log):
        if "Level set time" in line:
            secs = float(line.split("=")[1])
    os.unlink(log)
    if not secs:
        return None
    per_step = secs / steps
    return {"value": round(n ** 3 / per_step, 1), "unit": "voxel-updates/s", "cores": threads,
            "kind": "reference",
            "sample": f"reference constCurvModel3d (OpenMP {threads} threads, FP64) at {n}^3, "
                      f"WENO5+TVD-RK3 theta 30: {steps} inner steps of one outer step, timed by "
                      f"the reference's own inner-loop timer ({per_step * 1e3:.1f} ms/step)"}


def calibration_copy(mbytes):
    """One device-to-device copy of `mbytes` MiB (hipMemcpy's blit kernel), so a
    counter run records FETCH_SIZE / WRITE_SIZE of a plain copy of known size."""
    hip = C.CDLL("libamdhip64.so")
    n = int(mbytes) << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"# calibration copy: {mbytes} MiB device to device", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--curve-n", type=int, default=256)
    ap.add_argument("--curve-steps", type=int, default=3)
    ap.add_argument("--curve-tmax", type=float, default=20.0)
    ap.add_argument("--precisions", default="f32,f64")
    ap.add_argument("--calib-copy", type=int, default=0, help="MiB copied once before the run")
    ap.add_argument("--curve-only", action="store_true", help="only the drainage-curve leg")
    ap.add_argument("--curve-step-files", action="store_true",
                    help="drainage-curve leg with data_step_N.gz files written")
    ap.add_argument("--adjoint-only", action="store_true", help="only the adjoint leg")
    ap.add_argument("--batch-only", action="store_true", help="only the batched-subvolume leg")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--model2d-only", action="store_true",
                    help="only the 2D drainage leg (--model2d-n, --curve-steps, --curve-tmax)")
    ap.add_argument("--model2d-n", type=int, default=2048)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-subvolume leg")
    ap.add_argument("--precision-only", action="store_true",
                    help="only the FP32-vs-FP64 drainage-curve comparison (--n, --curve-steps, "
                         "--curve-tmax)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(z-slab sharding is described in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)
    if args.calib_copy:
        calibration_copy(args.calib_copy)
    if args.adjoint_only:
        out = {k: adjoint_leg(np.float32 if k == "f32" else np.float64)
               for k in args.precisions.split(",")}
        print(json.dumps({"adjoint": out}), flush=True)
        return
    if args.precision_only:
        print(json.dumps({"precision": precision_leg(args.n, args.curve_steps, args.curve_tmax)}),
              flush=True)
        return
    if args.model2d_only:
        print(json.dumps({"model2d": model2d_leg(args.model2d_n, args.curve_steps, args.curve_tmax,
                                                 cpu=not args.no_cpu)}), flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batch": batch_leg()}), flush=True)
        return
    if args.curve_only:
        gc = make_grid((args.curve_n,) * 3, np.float64)
        mc, _, _, _ = sphere_pack_mask(gc, (6, 12), 0.36, 1002)
        pc = initial_phi(gc, mc)
        out = {}
        for key in args.precisions.split(","):
            dt_ = np.float32 if key == "f32" else np.float64
            out[key] = curve_leg(dt_, gc, mc, pc, args.curve_steps, args.curve_tmax, 0.02, 0.04,
                                 args.curve_step_files)
        print(json.dumps({"drainage_curve": out, "grid": args.curve_n}), flush=True)
        return
    precs = args.precisions.split(",")
    n = args.n
    g64 = make_grid((n, n, n), np.float64)
    t0 = time.perf_counter()
    mask, _, radii, por = sphere_pack_mask(g64, (6, 12), 0.36, 1002)
    phi0 = initial_phi(g64, mask)
    print(f"# input: {n}^3 pack, {len(radii)} spheres, porosity {por:.4f}, "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    a, b = 0.02, 0.04
    r32 = gpu_leg(np.float32, g64, mask, phi0, args.steps, args.warmup, a, b)
    print(f"# fp32 done: {r32['elapsed_ms'] / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
    if "f64" not in precs:
        print(json.dumps({"f32_ms_per_step": r32["elapsed_ms"] / args.steps,
                          "stage_avg_ms": r32["stage_avg_ms"]}), flush=True)
        return
    r64 = gpu_leg(np.float64, g64, mask, phi0, args.steps, args.warmup, a, b)
    print(f"# fp64 done: {r64['elapsed_ms'] / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
    del mask, phi0
    curve = None
    if args.curve_n > 0:
        gc = make_grid((args.curve_n,) * 3, np.float64)
        mc, _, _, _ = sphere_pack_mask(gc, (6, 12), 0.36, 1002)
        pc = initial_phi(gc, mc)
        curve = {"grid": args.curve_n, "curvature_steps": args.curve_steps, "tmax": args.curve_tmax}
        for dt_, key in ((np.float32, "f32"), (np.float64, "f64")):
            curve[key] = curve_leg(dt_, gc, mc, pc, args.curve_steps, args.curve_tmax, a, b)
            print(f"# curve {key}: {curve[key]}", file=sys.stderr, flush=True)
        del mc, pc
    v32 = r32["nvox"] * args.steps / (r32["elapsed_ms"] * 1e-3)
    v64 = r64["nvox"] * args.steps / (r64["elapsed_ms"] * 1e-3)
    adjoint = None
    if not args.no_adjoint:
        adjoint = {"f32": adjoint_leg(np.float32), "f64": adjoint_leg(np.float64)}
        print(f"# adjoint: {adjoint}", file=sys.stderr, flush=True)
    batch = None
    if not args.no_batch:
        batch = batch_leg()
        print(f"# batch: {batch}", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_leg(args.cpu_n)
    line = {
        "metric": "voxel-updates/s per RK3 step (WENO5, theta 30, trapped masks), FP32 vs FP64, 1 MI355X",
        "value": round(v32, 1),
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["elapsed_ms"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Boolean sphere pack, radius U[6,12], porosity 0.36, seed 1002)",
        "config": {"workload": f"{n}^3 drainage step, WENO5 + TVD-RK3, theta 30 deg, dx 1, "
                               f"16 dt slabs, trapping masks on", "grid": n, "scheme": "WENO5+TVD-RK3"},
        "roofline": stage_roofline(r32, 4, "f32", n),
        "cpu_baseline": cpu,
        "fp64": {"value": round(v64, 1), "ms_per_step": round(r64["elapsed_ms"] / args.steps, 4),
                 "roofline": stage_roofline(r64, 8, "f64", n)},
        "fp32_over_fp64": round(v32 / v64, 3),
        "drainage_curve": curve,
        "adjoint": adjoint,
        "batch": batch,
        "stage_avg_ms": {"f32": [round(x, 4) for x in r32["stage_avg_ms"]],
                         "f64": [round(x, 4) for x in r64["stage_avg_ms"]]},
        "host_wall_s": {"f32": round(r32["wall_s"], 4), "f64": round(r64["wall_s"], 4)},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
