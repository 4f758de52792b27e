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

Beside it, in the same JSON line: the ENO2 "RK2" and ENO3 RK3 step rates
(contact angle 0) on the same pack; the drainage-curve time on the
workloads of lsmpqs_amd/workloads.py (256^3 drain + imbibe 25 + 25 steps,
512^3 FP32 vs FP64 40 steps in reference mode and FP32 with the quirks off,
64^3 ENO2 10 steps against the reference build); the adjoint; the batched
subvolumes; and the reference build's CPU rate.

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


def options(dtype, a, b, space=5, time_order=3, theta_deg=THETA_DEG):
    o = default_options(dtype)
    o.order_space_accur = space
    o.order_time_accur = time_order
    o.theta = theta_deg * 3.14159 / 180
    o.a = a
    o.b = b
    o.check_connectivity = 1
    return o


def gpu_leg(dtype, g64, mask, phi0, steps, warmup, a, b, space=5, time_order=3,
            theta_deg=THETA_DEG):
    """K inner steps of constCurvModel3d's loop on the device state the
    library's own drainage set-up left (time-step probe, fused RHS + RK
    stages, ghost extrapolation, imposition), timed with HIP events on the
    library's stream; then a shorter pass with events around every stage
    kernel for the per-kernel durations."""
    lib = hiplib.load(dtype)
    g = make_grid(tuple(g64.grid_dims), dtype)
    rec, keep = lean_arrays(dtype, mask, phi0)
    o = options(dtype, a, b, space, time_order, theta_deg)
    var = 1 if theta_deg > 0 else 0
    lib.lsmpqs_set_dt_slabs(16)
    if lib.lsmpqs_model_prepare(C.byref(o), C.byref(rec), C.byref(g), 1) != 0:
        raise hiplib.LsmpqsError(hiplib.last_error(lib))
    t = hiplib.Timing()
    if warmup:
        if lib.lsmpqs_bench_inner_steps(C.byref(o), C.byref(rec), C.byref(g), warmup, var, a, 0,
                                        C.byref(t)) != 0:
            raise hiplib.LsmpqsError(hiplib.last_error(lib))
    t = hiplib.Timing()
    w0 = time.perf_counter()
    rc = lib.lsmpqs_bench_inner_steps(C.byref(o), C.byref(rec), C.byref(g), steps, var, a, 0,
                                      C.byref(t))
    wall = time.perf_counter() - w0
    if rc != 0:
        raise hiplib.LsmpqsError(hiplib.last_error(lib))
    # per-kernel durations in a separate pass (events around each stage launch)
    tp = hiplib.Timing()
    tp.profile = 1
    if lib.lsmpqs_bench_inner_steps(C.byref(o), C.byref(rec), C.byref(g), max(3, steps // 2), var,
                                    a, 0, C.byref(tp)) != 0:
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


# Algorithmic bytes per voxel-update of a whole inner step (fields read and
# written once per stage, coefficients derived from the mask, trapped-phase
# masks imposed): contact-angle path RK3 13 fields, theta 0 RK3 11, theta 0
# "RK2" 8.
STEP_FIELDS = {("var", 3): 13, ("const", 3): 11, ("const", 2): 8}


def scheme_leg(dtype, g64, mask, phi0, steps, warmup, space, time_order):
    """Inner-step rate of a contact-angle-0 scheme (the constant-coefficient
    right-hand sides: ENO2 "RK2", ENO3 RK3) on the same pack: voxel-updates/s
    and the whole step's algorithmic HBM rate against 8 TB/s."""
    r = gpu_leg(dtype, g64, mask, phi0, steps, warmup, 0.02, 0.04, space, time_order, 0.0)
    itemsize = np.dtype(dtype).itemsize
    ms = r["elapsed_ms"] / steps
    vu = r["nvox"] / (ms * 1e-3)
    fields = STEP_FIELDS[("const", time_order)]
    ach = fields * itemsize * vu / 1e9
    return {"scheme": {2: "ENO2", 3: "ENO3"}[space] + ("+RK2" if time_order == 2 else "+TVD-RK3"),
            "ms_per_step": round(ms, 4), "voxel_updates_per_s": round(vu, 1),
            "stage_avg_ms": [round(x, 4) for x in r["stage_avg_ms"][:time_order]],
            "step_roofline": {"bound": "hbm", "alg_bytes_per_voxel_update": fields * itemsize,
                              "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(ach / HBM_PEAK_GBS, 4)}}


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


def model_leg(dtype, g64, mask, phi0, o, drain_steps, imbibe_steps, slabs=16, quirks=None):
    """drain_model3d over drain_steps curvature steps, then (imbibe_steps > 0)
    imbibe_model3d from the last drainage pressure, on the GPU from host
    arrays (step files off): host wall time of each curve, outer and inner
    steps, GPU time in inner loops, final non-wetting saturation."""
    from lsmpqs_amd.workloads import saturation_of
    lib = hiplib.load(dtype)
    g = make_grid(tuple(g64.grid_dims), dtype)
    rec, keep = lean_arrays(dtype, mask, phi0)
    _, _, DataArrays = types_for(dtype)
    extra = {f: np.zeros(keep[0].shape, dtype) for f in ("mask_w", "mask_nw")}
    rec.mask_w, rec.mask_nw = extra["mask_w"].ctypes.data, extra["mask_nw"].ctypes.data
    lib.lsmpqs_set_dt_slabs(slabs)
    os.environ["LSMPQS_NO_STEP_FILES"] = "1"
    old = lib.lsmpqs_set_quirks(quirks) if quirks is not None else None
    out = {}
    try:
        for name, steps, fn in (("drain", drain_steps, lib.drain_model3d),
                                ("imbibe", imbibe_steps, lib.imbibe_model3d)):
            if steps <= 0:
                continue
            if name == "drain":
                o.amax = o.a + o.b * o.dc * (steps - 0.5)
            else:
                o.a = o.a + o.b * o.dc * (drain_steps - 1)
                o.amin = o.a - o.b * o.dc * (steps - 0.5)
            w0 = time.perf_counter()
            fn(C.byref(o), C.byref(rec), C.byref(g), None)
            wall = time.perf_counter() - w0
            hiplib.check(lib)
            st = hiplib.Timing()
            lib.lsmpqs_session_stats(C.byref(rec), C.byref(g), C.byref(st))
            out[name] = {"curvature_steps": steps, "wall_s": round(wall, 3),
                         "outer_steps": int(st.outer_steps), "inner_steps": int(st.inner_steps),
                         "inner_loop_s": round(st.inner_loop_ms * 1e-3, 3),
                         "sat_nw": round(saturation_of(keep[0], keep[1], g), 8)}
    finally:
        if old is not None:
            lib.lsmpqs_set_quirks(old)
        lib.lsmpqs_release(C.byref(rec))
    return out, keep[0]


def curves_leg(n_pair=256, n_prec=512):
    """The drainage-curve half of the metric on three workloads:
    * the 256^3 pack of radius U[6,12] (porosity 0.36, seed 1002), WENO5 +
      TVD-RK3, 30 deg: 25 drainage then 25 imbibition curvature steps,
      eps_stop 1e-4 dx, pseudo-time 20 per step, FP32 and FP64;
    * the 512^3 pack of radius U[10,20] (porosity 0.35, seed 1003), WENO5 +
      TVD-RK3, 39.999 deg, 40 drainage steps, pseudo-time 20 per step: FP32
      and FP64 in reference mode (precision_leg, with the FP32-vs-FP64
      errors), and FP32 with the reference quirks off, whose WENO5 normal term
      has the physical sign, so the non-wetting phase invades and trapping
      works on a moving front."""
    from lsmpqs_amd import workloads as W
    spec = W.WENO5_DRAIN_IMBIBE
    g64 = make_grid((n_pair,) * 3, np.float64)
    mask, _, _, por = sphere_pack_mask(g64, spec["radius"], spec["porosity"], spec["seed"])
    phi0 = initial_phi(g64, mask)
    pair = {"grid": n_pair, "porosity": round(float(por), 4), "tmax_per_step": 20.0}
    for key, dt_ in (("f32", np.float32), ("f64", np.float64)):
        o = W.weno5_options(spec, dt_)
        o.tmax = 20.0
        pair[key], _ = model_leg(dt_, g64, mask, phi0, o, spec["steps"], spec["steps"])
        print(f"# curve pair {key}: {pair[key]}", file=sys.stderr, flush=True)
    del mask, phi0
    prec = precision_leg(n_prec, W.WENO5_PRECISION["steps"], 20.0)
    print(f"# precision: {prec}", file=sys.stderr, flush=True)
    spec3 = W.WENO5_PRECISION
    g3 = make_grid((n_prec,) * 3, np.float64)
    mask, _, _, _ = sphere_pack_mask(g3, spec3["radius"], spec3["porosity"], spec3["seed"])
    phi0 = initial_phi(g3, mask)
    o = W.weno5_options(spec3, np.float32)
    o.tmax = 20.0
    fixed, _ = model_leg(np.float32, g3, mask, phi0, o, spec3["steps"], 0, quirks=0)
    fixed["grid"] = n_prec
    print(f"# fixed-mode curve: {fixed}", file=sys.stderr, flush=True)
    return {"weno5_drain25_imbibe25": pair, "weno5_precision_drain40": prec,
            "weno5_fixed_mode_drain40_f32": fixed}


def drain64_leg(threads):
    """The 64^3 ENO2 drainage curve (radius 8, porosity 0.40, 10 curvature
    steps from kappa 0.05 to 0.30, reinitialization every 20 time steps,
    trapping on): wall time of drain_model3d on the GPU (FP64 and FP32) and
    of the reference build's drain_model3d (OpenMP, `threads` threads, FP64
    and FP32 as shipped), the same `threads` time-step slabs on both."""
    import tempfile
    import oracle_py as op
    from lsmpqs_amd import workloads as W
    from lsmpqs_amd.abi import Arrays
    from lsmpqs_amd.workloads import saturation_of
    libc = C.CDLL("libc.so.6")
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    devnull = C.c_void_p(libc.fopen(b"/dev/null", b"w"))
    out = {"grid": 64, "curvature_steps": W.DRAIN64["steps"], "threads": threads}
    finals = {}
    for kind in ("gpu", "reference"):
        if kind == "reference" and not op.have_ref(np.float64):
            continue
        for key, dt_ in (("f64", np.float64), ("f32", np.float32)):
            g, mask, phi0, _ = W.drain64_inputs(dt_)
            A = Arrays(g, dt_)
            A["mask"], A["phi"] = mask, phi0
            o = W.drain64_options(dt_)
            cwd = os.getcwd()
            if kind == "gpu":
                lib = hiplib.load(dt_)
                lib.lsmpqs_set_dt_slabs(threads)
                os.environ["LSMPQS_NO_STEP_FILES"] = "1"
            else:
                op.pad_phi_bin(A, g)
                lib = op.ref_lib(dt_)
                op.set_omp_threads(threads)
                os.chdir(tempfile.mkdtemp(prefix="lsm_ref64_"))
            try:
                w0 = time.perf_counter()
                lib.drain_model3d(C.byref(o), A.ptr(), C.byref(g), devnull)
                wall = time.perf_counter() - w0
            finally:
                os.chdir(cwd)
            if kind == "gpu":
                hiplib.check(lib)
                lib.lsmpqs_release(A.ptr())
            finals[(kind, key)] = A["phi"].astype(np.float64)
            out[f"{kind}_{key}_s"] = round(wall, 3)
            out[f"{kind}_{key}_sat_nw"] = round(saturation_of(A["phi"], A["mask"], g), 8)
    if ("reference", "f64") in finals:
        out["max_abs_diff_gpu_f64_vs_reference_f64"] = float(
            np.max(np.abs(finals[("gpu", "f64")] - finals[("reference", "f64")])))
        out["reference_f64_over_gpu_f64"] = round(out["reference_f64_s"] / out["gpu_f64_s"], 2)
    return out


def adjoint_leg(dtype, n=128, steps=200, checkpoint_every=20):
    """Gradient of (S_w - 0.45)^2 with respect to a per-voxel contact-angle
    field (40 deg + Gaussian random field, sigma 5 deg, correlation length
    10; n^3 pack of radius 8, seed 1005) and b, through `steps` WENO5 TVD-RK3
    steps at a0/b = 0.15, b = 0.04: host wall time of the forward map alone
    and of forward + reverse pass (lsmpqs_adjoint_run; the forward keeps every
    step when it fits in device memory, checkpoint_every applies otherwise)."""
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
    out = {"grid": n, "steps": steps, "checkpoint_every": checkpoint_every,
           "forward_s": round(tf, 4), "gradient_s": round(tg, 4),
           "gradient_over_forward": round(tg / tf, 2),
           "forward_voxel_updates_per_s": round(nvox * steps / tf, 1),
           "sw": round(rg.sw, 8), "loss": rg.loss, "dloss_db": rg.dloss_db,
           "same_loss_as_forward": rf.loss == rg.loss}
    if np.dtype(dtype) == np.float32:
        # fit of the per-voxel angle field to the saturation of the same field 8 deg wider
        from lsmpqs_amd.adjoint import fit_angle_field
        shifted = (theta + np.float32(np.radians(8.0))).astype(dtype)
        tgt = saturation_loss(phi0, mask, shifted, g, o, steps, a0, b, target, dt=rg.dt,
                              gradient=False).sw
        w0 = time.perf_counter()
        _, losses = fit_angle_field(phi0, mask, theta, g, o, steps, a0, b, tgt, rg.dt,
                                    iterations=6, checkpoint_every=checkpoint_every)
        out["field_fit"] = {"target_sw": round(tgt, 8), "losses": [float("%.4e" % x) for x in losses],
                            "gauss_newton_steps": len(losses) - 1,
                            "wall_s": round(time.perf_counter() - w0, 3)}
    return out


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
               "batched_wall_s": round(wall, 3),
               "voxel_updates": vu,
               "voxel_updates_per_s_wall": round(vu / wall, 1),
               "batched_steps": int(st.batched_steps), "box_steps": int(st.box_steps),
               "brick_depth": int(st.brick_depth), "workgroups": int(st.workgroups),
               "batched_outer_steps": int(st.batched_outer), "box_outer_steps": int(st.box_outer),
               # summed event intervals of the rounds; the two box groups'
               # rounds run on their own streams and overlap on the GPU
               "inner_round_gpu_s": round(st.batched_ms * 1e-3, 3),
               "outer_round_gpu_s": round(st.batched_outer_ms * 1e-3, 3)}
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
    # band voxels at least 4 voxels (26-neighbour steps) from any voxel whose
    # trapped state differs between the precisions
    from scipy import ndimage
    near_trap = ndimage.binary_dilation(trap_diff, structure=np.ones((3, 3, 3), bool),
                                        iterations=3)
    away = band & ~near_trap
    out = {"grid": n, "spheres": int(len(radii)), "porosity": round(float(por), 4),
           "curvature_steps": curv_steps, "tmax": tmax, "f32": res["f32"], "f64": res["f64"],
           "fp64_over_fp32_wall": round(res["f64"]["wall_s"] / res["f32"]["wall_s"], 3),
           "sat_nw_f32": sat(p32), "sat_nw_f64": sat(p64),
           "sat_abs_diff": abs(sat(p32) - sat(p64)),
           "phi_linf_band": float(dev[band].max()),
           "phi_linf_band_same_trapping": float(dev[same_trap].max()) if same_trap.any() else 0.0,
           "phi_linf_band_away_from_trapping_diff": float(dev[away].max()) if away.any() else 0.0,
           "band_voxels_near_trapping_diff": int((band & near_trap).sum()),
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


def cpu_threads():
    """Host threads for the CPU baseline: the CPUs this process may run on,
    capped by OMP_NUM_THREADS when it is set (the box's CPU share for one
    GPU), and the CPU model name."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return n, model, len(os.sched_getaffinity(0))


def cpu_leg(n, steps=8):
    """The reference build (oracle/_ref, OpenMP) on a bounded sample of the
    headline workload: one outer step of constCurvModel3d at n^3 (WENO5 +
    TVD-RK3, 30 deg, trapping off) whose pseudo-time window holds `steps`
    inner steps, in FP32 (as the reference ships) and FP64; the time is the
    reference's own inner-loop timer (its "Level set time" log line).  dt is
    formed as the reference forms it (H/dx at each slab's last voxel), from
    the reference's own right-hand side at those voxels."""
    import tempfile
    import oracle_py as op
    from lsmpqs_amd.abi import Arrays
    if not op.have_ref(np.float64):
        return None
    threads, model, visible = cpu_threads()
    op.set_omp_threads(threads)
    g64 = make_grid((n, n, n), np.float64)
    mask64, _, _, _ = sphere_pack_mask(g64, (6, 12), 0.36, 1002)
    phi64 = initial_phi(g64, mask64)
    res = {}
    for key, dt in (("f32", np.float32), ("f64", np.float64)):
        lib = op.ref_lib(dt)
        R = C.c_double if dt == np.float64 else C.c_float
        g = make_grid((n, n, n), dt)
        A = Arrays(g, dt)
        A["mask"] = mask64
        A["phi"] = phi64
        o = options(dt, 0.02, 0.04)
        gb, fb = op.gb6(g), op.fb6(g)
        op.fcall(lib, "getcentralgrad3dorder2_", dt, A["mask_x"], A["mask_y"], A["mask_z"], *gb,
                 A["mask"], *gb, *fb, 1.0, 1.0, 1.0)
        bm, um = op.Scalar(0.0, R), op.Scalar(0.0, R)
        op.fcall(lib, "setvarcurvadv3d_", dt, A["mask"], A["mask_x"], A["mask_y"], A["mask_z"],
                 A["curvature_coeff"], A["external_velocity_x"], A["external_velocity_y"],
                 A["external_velocity_z"], bm, um, 0.04, *gb, *fb, 1.0)
        op.fcall(lib, "setvarnorm3d_", dt, A["mask"], A["mask_x"], A["mask_y"], A["mask_z"],
                 A["normal_velocity"], 0.02, float(o.theta), *gb, *fb, 1.0)
        o.b_max_over_dx, o.max_U_over_dx = bm.value, um.value
        # the first dt as the reference forms it: its right-hand side routine
        # returns H/dx of the last voxel of the range, so each slab's value
        # comes from that one voxel
        nsl = g.khi_fb - g.klo_fb + 1
        H = -1.0
        for t in range(threads):
            klo = g.klo_fb + nsl * t // threads
            khi = min(g.klo_fb + nsl * (t + 1) // threads - 1, g.khi_fb)
            if klo > khi:
                continue
            mh = op.Scalar(0.0, R)
            op.fcall(lib, "getrhs3d5th_", dt, A["phi"], A["normal_velocity"], A["curvature_coeff"],
                     A["external_velocity_x"], A["external_velocity_y"],
                     A["external_velocity_z"], mh, A["lse_rhs"], *gb, g.ihi_fb, g.ihi_fb,
                     g.jhi_fb, g.jhi_fb, khi, khi, 1.0, 1.0, 1.0)
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
        lib.constCurvModel3d.restype = R
        lib.constCurvModel3d(C.byref(o), A.ptr(), C.byref(g), C.c_void_p(fp))
        libc.fclose(C.c_void_p(fp))
        secs = None
        for line in open(log):
            if "Level set time" in line:
                secs = float(line.split("=")[1])
        os.unlink(log)
        del A
        if not secs:
            return None
        res[key] = {"value": round(n ** 3 * steps / secs, 1), "ms_per_step": round(secs / steps * 1e3, 2)}
        print(f"# cpu reference {key}: {res[key]}", file=sys.stderr, flush=True)
    return {"value": res["f32"]["value"], "unit": "voxel-updates/s", "cores": threads,
            "kind": "reference",
            "sample": f"reference constCurvModel3d (OpenMP, {threads} threads, FP32 as shipped) at "
                      f"{n}^3 on the headline pack, WENO5 + TVD-RK3, 30 deg: {steps} inner steps "
                      f"of one outer step, timed by the reference's own inner-loop timer; the "
                      f"512^3 box would need 24 GB of host arrays and a minute of the "
                      f"reference's serial set-up per precision",
            "f64_value": res["f64"]["value"], "ms_per_step": {k: v["ms_per_step"] for k, v in res.items()},
            "cpu_model": model, "cpus_visible": visible}


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
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--curve-n", type=int, default=256)
    ap.add_argument("--curve-steps", type=int, default=3)
    ap.add_argument("--curve-tmax", type=float, default=20.0)
    ap.add_argument("--precisions", default="f32,f64")
    ap.add_argument("--calib-copy", type=int, default=0, help="MiB copied once before the run")
    ap.add_argument("--curve-only", action="store_true",
                    help="only a drainage curve on the headline pack (--curve-n, --curve-steps, "
                         "--curve-tmax)")
    ap.add_argument("--curve-step-files", action="store_true",
                    help="drainage-curve leg with data_step_N.gz files written")
    ap.add_argument("--adjoint-only", action="store_true", help="only the adjoint leg")
    ap.add_argument("--batch-only", action="store_true", help="only the batched-subvolume leg")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--model2d-only", action="store_true",
                    help="only the 2D drainage leg (--model2d-n, --curve-steps, --curve-tmax)")
    ap.add_argument("--model2d-n", type=int, default=2048)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-subvolume leg")
    ap.add_argument("--no-curves", action="store_true",
                    help="skip the drainage-curve workloads (256^3 drain+imbibe, 512^3 precision "
                         "and fixed-mode curves, 64^3 ENO2 curve against the reference)")
    ap.add_argument("--no-schemes", action="store_true", help="skip the ENO2 / ENO3 step rates")
    ap.add_argument("--precision-only", action="store_true",
                    help="only the FP32-vs-FP64 drainage-curve comparison (--n, --curve-steps, "
                         "--curve-tmax)")
    ap.add_argument("--stage-only", action="store_true",
                    help="only the headline inner steps in FP32 (for kernel profiles)")
    ap.add_argument("--schemes-only", action="store_true",
                    help="only the ENO2 / ENO3 step rates (--n, --steps, --warmup; for kernel "
                         "profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path "
              "(a z-slab split with 3-plane halos over xGMI is the design, not built)",
              file=sys.stderr)
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
    n = args.n
    g64 = make_grid((n, n, n), np.float64)
    t0 = time.perf_counter()
    mask, _, radii, por = sphere_pack_mask(g64, (6, 12), 0.36, 1002)
    phi0 = initial_phi(g64, mask)
    print(f"# input: {n}^3 pack, {len(radii)} spheres, porosity {por:.4f}, "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    a, b = 0.02, 0.04
    if args.schemes_only:
        out = {}
        for (sp, to) in ((2, 2), (3, 3)):
            for key in args.precisions.split(","):
                dt_ = np.float32 if key == "f32" else np.float64
                r = scheme_leg(dt_, g64, mask, phi0, args.steps, args.warmup, sp, to)
                out[f"{r['scheme']}_{key}"] = r
        print(json.dumps({"schemes": out}), flush=True)
        return
    r32 = gpu_leg(np.float32, g64, mask, phi0, args.steps, args.warmup, a, b)
    print(f"# fp32 done: {r32['elapsed_ms'] / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
    if args.stage_only or "f64" not in args.precisions.split(","):
        print(json.dumps({"f32_ms_per_step": r32["elapsed_ms"] / args.steps,
                          "stage_avg_ms": r32["stage_avg_ms"]}), flush=True)
        return
    r64 = gpu_leg(np.float64, g64, mask, phi0, args.steps, args.warmup, a, b)
    print(f"# fp64 done: {r64['elapsed_ms'] / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
    schemes = None
    if not args.no_schemes:
        schemes = {}
        for (sp, to) in ((2, 2), (3, 3)):
            for key, dt_ in (("f32", np.float32), ("f64", np.float64)):
                r = scheme_leg(dt_, g64, mask, phi0, args.steps, args.warmup, sp, to)
                schemes[f"{r['scheme']}_{key}"] = r
                print(f"# scheme {r['scheme']} {key}: {r}", file=sys.stderr, flush=True)
    del mask, phi0
    v32 = r32["nvox"] * args.steps / (r32["elapsed_ms"] * 1e-3)
    v64 = r64["nvox"] * args.steps / (r64["elapsed_ms"] * 1e-3)
    curves = None
    if not args.no_curves:
        curves = curves_leg(256, n)
        curves["eno2_drain64_vs_reference"] = drain64_leg(cpu_threads()[0])
        print(f"# drain64: {curves['eno2_drain64_vs_reference']}", file=sys.stderr, flush=True)
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
        "metric": "voxel-updates/s per RK3 step (WENO5, theta 30, trapped masks) & drainage-curve "
                  "time, FP32 vs FP64, 1 MI355X",
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
        "config": {"workload": f"{n}^3 drainage inner step, WENO5 + TVD-RK3, theta 30 deg, dx 1, "
                               f"16 dt slabs, trapping masks on", "grid": n, "scheme": "WENO5+TVD-RK3"},
        "roofline": stage_roofline(r32, 4, "f32", n),
        "cpu_baseline": cpu,
        "fp64": {"value": round(v64, 1), "ms_per_step": round(r64["elapsed_ms"] / args.steps, 4),
                 "roofline": stage_roofline(r64, 8, "f64", n)},
        "fp32_over_fp64": round(v32 / v64, 3),
        "schemes": schemes,
        "drainage_curves": curves,
        "adjoint": adjoint,
        "batch": batch,
        "stage_avg_ms": {"f32": [round(x, 4) for x in r32["stage_avg_ms"]],
                         "f64": [round(x, 4) for x in r64["stage_avg_ms"]]},
        "host_wall_s": {"f32": round(r32["wall_s"], 4), "f64": round(r64["wall_s"], 4)},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
