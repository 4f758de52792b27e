"""Gradient of the saturation loss through the level-set time steps
(lsmpqs_adjoint_run) on the GPU.

The reverse pass is checked against central finite differences of the same
forward map in FP64 along 16 seeded random directions of the contact-angle
field (and along b), for each spatial scheme and Runge-Kutta order, to
1e-3 of the largest directional derivative (directions whose probes straddle
a switch of the scheme are detected and left out, at most 4 of 16); the
FP32 library's gradient against the FP64 one; the result against the
checkpoint interval; and the torch.autograd wrapper against the raw entry
point, with a fit of a uniform contact angle to a target saturation and a
fit of the full-size per-voxel angle field.  The
forward map is the variable-coefficient inner step of the reference's
constCurvModel3d (src/level_set_evolution/constCurvModel3d.c:127-230) with
setVarNorm3dTheta's per-voxel angle
(src/spatial_derivatives/qss_spatial_derivatives3d.f:94-170)."""
import numpy as np
import pytest

from lsmpqs_amd.abi import default_options
from lsmpqs_amd.adjoint import saturation_loss, ThetaLoss
from lsm_cases import case

N = 20
STEPS = 12
A0, B = 0.05, 0.04
TARGET = 0.5
FD_H = 3e-7
FD_TOL = 1e-3


def problem(dt, space, time, seed=7):
    g, mask, phi0 = case(N, dt, seed=3100)
    o = default_options(dt)
    o.order_space_accur = space
    o.order_time_accur = time
    rng = np.random.default_rng(seed)
    theta = (0.4 + 0.6 * rng.random(phi0.shape)).astype(dt)
    return g, o, np.ascontiguousarray(phi0), np.ascontiguousarray(mask), theta


def fill_of(g):
    return (slice(g.klo_fb, g.khi_fb + 1), slice(g.jlo_fb, g.jhi_fb + 1),
            slice(g.ilo_fb, g.ihi_fb + 1))


def smooth_between(lm, l0, lp, h):
    """The forward map is piecewise smooth: the reference's scheme switches
    discontinuously (ENO stencils at |D2| ties of opposite sign, the
    curvature cut at |grad phi|^2 = 1e-5, the normal-term cut at
    |vel_n| = 1e-5, the sign of the ghost extrapolation at phi = 0).  A
    central difference whose two probes straddle such a switch measures the
    jump, not the derivative; its one-sided differences then disagree by
    jump / h.  True where they agree to 5 %."""
    fwd, bwd = (lp - l0) / h, (l0 - lm) / h
    return abs(fwd - bwd) <= 0.05 * max(abs(fwd), abs(bwd), 1e-30)


def loss(g, o, phi0, mask, theta, b=B, dt=0.0):
    return saturation_loss(phi0, mask, theta, g, o, STEPS, A0, b, TARGET, dt=dt,
                           gradient=False).loss


@pytest.mark.gpu
@pytest.mark.parametrize("space,time", [(2, 1), (2, 2), (3, 3), (5, 3), (5, 2)])
def test_gradient_matches_finite_differences_fp64(space, time):
    print(f"case: adjoint FD fp64 scheme {space} RK{time} n={N} steps={STEPS}", flush=True)
    g, o, phi0, mask, theta = problem(np.float64, space, time)
    r = saturation_loss(phi0, mask, theta, g, o, STEPS, A0, B, TARGET, checkpoint_every=4)
    assert np.isfinite(r.loss) and r.dt > 0
    assert r.loss == loss(g, o, phi0, mask, theta, dt=r.dt)  # loss-only run, same map
    f = fill_of(g)
    ghost = np.ones(theta.shape, dtype=bool)
    ghost[f] = False
    assert not r.dtheta[ghost].any()
    assert np.abs(r.dtheta).max() > 0
    rng = np.random.default_rng(100 + 10 * space + time)
    l0 = r.loss
    got, want, skipped = [], [], 0
    for _ in range(16):
        v = np.zeros_like(theta)
        v[f] = rng.standard_normal(v[f].shape)
        lp = loss(g, o, phi0, mask, theta + FD_H * v, dt=r.dt)
        lm = loss(g, o, phi0, mask, theta - FD_H * v, dt=r.dt)
        if not smooth_between(lm, l0, lp, FD_H):
            skipped += 1
            continue
        want.append((lp - lm) / (2 * FD_H))
        got.append(float(np.sum(r.dtheta * v)))
    print(f"directions crossing a switch of the scheme: {skipped} of 16", flush=True)
    assert skipped <= 4
    got, want = np.array(got), np.array(want)
    scale = np.abs(want).max()
    assert scale > 0
    print(f"adjoint vs central differences: max error {np.abs(got - want).max() / scale:.2e} "
          f"of the largest directional derivative", flush=True)
    assert np.abs(got - want).max() <= FD_TOL * scale, (got, want)
    # b
    for hb in (1e-6, 3e-7, 1e-7):
        lp = loss(g, o, phi0, mask, theta, B + hb, dt=r.dt)
        lm = loss(g, o, phi0, mask, theta, B - hb, dt=r.dt)
        if smooth_between(lm, l0, lp, hb):
            break
    else:
        pytest.fail("every b step crosses a switch of the scheme")
    fd_b = (lp - lm) / (2 * hb)
    assert abs(r.dloss_db - fd_b) <= FD_TOL * max(abs(fd_b), 1e-12), (r.dloss_db, fd_b)


@pytest.mark.gpu
@pytest.mark.parametrize("dt,space,time", [(np.float64, 5, 3), (np.float64, 2, 2),
                                           (np.float64, 3, 1), (np.float32, 5, 3)])
@pytest.mark.parametrize("store", ["all", "segment", "none"])
def test_checkpoint_interval_does_not_change_the_gradient(store, dt, space, time, monkeypatch):
    """Whole trajectory kept, checkpointed segments with their stage values
    kept, and segments recomputed twice: the same loss and final state bit
    for bit, the same gradient up to the atomics' summation order."""
    print(f"case: adjoint checkpoints {np.dtype(dt).name} space {space} time {time}, "
          f"store {store}", flush=True)
    g, o, phi0, mask, theta = problem(dt, space, time)
    monkeypatch.delenv("LSMPQS_ADJOINT_STORE", raising=False)
    ref = saturation_loss(phi0, mask, theta, g, o, STEPS, A0, B, TARGET, checkpoint_every=STEPS)
    monkeypatch.setenv("LSMPQS_ADJOINT_STORE", store)
    rs = [ref] + [saturation_loss(phi0, mask, theta, g, o, STEPS, A0, B, TARGET, checkpoint_every=k)
                  for k in (1, 5, STEPS)]
    for r in rs[1:]:
        assert r.loss == rs[0].loss
        assert np.array_equal(r.phi, rs[0].phi)
        # the transposed stencils accumulate with atomics: order-dependent rounding only
        tol = 1e-10 if dt == np.float64 else 1e-4
        assert np.abs(r.dtheta - rs[0].dtheta).max() <= tol * np.abs(rs[0].dtheta).max()
        assert abs(r.dloss_db - rs[0].dloss_db) <= tol * abs(rs[0].dloss_db)


@pytest.mark.gpu
@pytest.mark.parametrize("space,time", [(3, 3), (5, 3)])
def test_fp32_gradient_matches_fp64(space, time):
    print(f"case: adjoint fp32 vs fp64 scheme {space} RK{time}", flush=True)
    g64, o64, phi64, mask64, th64 = problem(np.float64, space, time)
    g32, o32, phi32, mask32, th32 = problem(np.float32, space, time)
    r64 = saturation_loss(phi64, mask64, th64, g64, o64, STEPS, A0, B, TARGET)
    r32 = saturation_loss(phi32, mask32, th32, g32, o32, STEPS, A0, B, TARGET, dt=r64.dt)
    assert abs(r32.sw - r64.sw) <= 1e-5
    d64, d32 = r64.dtheta, r32.dtheta.astype(np.float64)
    rel = np.linalg.norm(d32 - d64) / np.linalg.norm(d64)
    print(f"fp32 vs fp64 gradient: relative L2 {rel:.3e}", flush=True)
    assert rel <= 1e-4
    assert abs(r32.dloss_db - r64.dloss_db) <= 1e-4 * abs(r64.dloss_db)


@pytest.mark.gpu
def test_autograd_wrapper_and_uniform_angle_fit():
    import torch
    print("case: adjoint autograd + fit fp64 ENO3 RK3", flush=True)
    g, o, phi0, mask, theta = problem(np.float64, 3, 3)
    prob = {"phi0": phi0, "mask": mask, "grid": g, "options": o, "nsteps": STEPS, "a0": A0,
            "target_sw": TARGET}
    th = torch.tensor(theta, requires_grad=True)
    b = torch.tensor(B, dtype=torch.float64, requires_grad=True)
    L = ThetaLoss.apply(th, b, prob)
    L.backward()
    r = saturation_loss(phi0, mask, theta, g, o, STEPS, A0, B, TARGET)
    assert float(L.detach()) == r.loss
    # two runs: the atomics of the transposed stencils round in different orders
    assert np.abs(th.grad.numpy() - r.dtheta).max() <= 1e-10 * np.abs(r.dtheta).max()
    assert abs(float(b.grad) - r.dloss_db) <= 1e-10 * abs(r.dloss_db)
    # Gauss-Newton on one uniform angle: loss = res^2, d res / d t = (dL/dt) / (2 res)
    prob["dt"] = r.dt
    t = torch.tensor(0.6, dtype=torch.float64, requires_grad=True)
    field = torch.ones(phi0.shape, dtype=torch.float64)
    bb = torch.tensor(B, dtype=torch.float64)
    # a target the map reaches: the saturation at angle 0.9
    prob["target_sw"] = saturation_loss(phi0, mask, np.full_like(theta, 0.9), g, o, STEPS, A0, B,
                                        TARGET, dt=r.dt, gradient=False).sw
    L0 = ThetaLoss.apply(t * field, bb, prob)
    losses = []
    for _ in range(4):
        t.grad = None
        L = ThetaLoss.apply(t * field, bb, prob)
        L.backward()
        res = prob["last"].sw - prob["target_sw"]
        losses.append(float(L))
        if float(t.grad) == 0.0:
            break
        with torch.no_grad():
            t -= res / (float(t.grad) / (2 * res))
    print(f"fit: loss {losses}, angle {float(t):.4f}", flush=True)
    assert float(L0) >= 0
    assert losses[-1] <= 1e-2 * losses[0]


@pytest.mark.gpu
def test_contact_angle_field_128():
    """The full-size adjoint case: 128^3 pack (radius 8, seed 1005), angle
    field 40 deg + Gaussian random field (sigma 5 deg, correlation length 10),
    200 WENO5 TVD-RK3 steps at equilibrium curvature a0/b = 0.15 (b = 0.04),
    loss (S_w - 0.45)^2, checkpoints every 20 steps.  FP64 gradient against
    central differences along 16 seeded directions (8 white-noise, 8 smooth
    random fields); the FP32 gradient against the FP64 one.  Over 200 steps
    on 2.1M voxels the scheme's switches lie dense enough that steps of
    1e-7..1e-6 rad straddle some in most directions; at 1e-9 rad none is
    met and FP64 rounding still leaves the difference quotient good to about
    1e-5."""
    H128 = 1e-9
    from lsmpqs_amd.inputs import contact_angle_case, gaussian_random_field
    steps, b, a0, target, ck = 200, 0.04, 0.15 * 0.04, 0.45, 20
    print("case: adjoint 128^3 WENO5 RK3 200 steps, fp64 FD + fp32", flush=True)
    g, mask, phi0, theta = contact_angle_case(128, dtype=np.float64)
    o = default_options(np.float64)
    o.order_space_accur, o.order_time_accur = 5, 3
    r64 = saturation_loss(phi0, mask, theta, g, o, steps, a0, b, target, checkpoint_every=ck)
    print(f"fp64: S_w {r64.sw:.6f} loss {r64.loss:.6e} dt {r64.dt:.4f} dL/db {r64.dloss_db:.6e}",
          flush=True)

    def L(th, bb=b):
        return saturation_loss(phi0, mask, th, g, o, steps, a0, bb, target, dt=r64.dt,
                               gradient=False).loss

    f = fill_of(g)
    rng = np.random.default_rng(1005)
    got, want, skipped, dirs = [], [], 0, []
    for q in range(16):
        v = np.zeros_like(theta)
        if q < 8:
            v[f] = rng.standard_normal(v[f].shape)
        else:
            v[f] = gaussian_random_field(theta.shape, 1.0, 10.0, 2000 + q)[f]
        dirs.append(v)
        lp, lm = L(theta + H128 * v), L(theta - H128 * v)
        if not smooth_between(lm, r64.loss, lp, H128):
            skipped += 1
            continue
        want.append((lp - lm) / (2 * H128))
        got.append(float(np.sum(r64.dtheta * v)))
    got, want = np.array(got), np.array(want)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"directions crossing a switch: {skipped} of 16; max error {err:.2e} of the largest "
          f"directional derivative", flush=True)
    assert skipped <= 4
    assert err <= FD_TOL
    # the run the case names is FP32; its gradient against the FP64 one
    g32, mask32, phi32, th32 = contact_angle_case(128, dtype=np.float32)
    o32 = default_options(np.float32)
    o32.order_space_accur, o32.order_time_accur = 5, 3
    r32 = saturation_loss(phi32, mask32, th32, g32, o32, steps, a0, b, target, dt=r64.dt,
                          checkpoint_every=ck)
    rel = np.linalg.norm(r32.dtheta.astype(np.float64) - r64.dtheta) / np.linalg.norm(r64.dtheta)
    print(f"fp32: S_w {r32.sw:.6f} (fp64 {r64.sw:.6f}); gradient relative L2 vs fp64 {rel:.3e}; "
          f"dL/db {r32.dloss_db:.6e}", flush=True)
    d32 = np.array([float(np.sum(r32.dtheta.astype(np.float64) * v)) for v in dirs])
    d64 = np.array([float(np.sum(r64.dtheta * v)) for v in dirs])
    derr = np.abs(d32 - d64).max() / np.abs(d64).max()
    print(f"fp32 directional derivatives vs fp64: max error {derr:.2e}", flush=True)
    assert abs(r32.sw - r64.sw) <= 1e-4
    assert rel <= 2e-3
    assert derr <= FD_TOL
    assert abs(r32.dloss_db - r64.dloss_db) <= 1e-3 * abs(r64.dloss_db)


@pytest.mark.gpu
def test_fit_of_a_contact_angle_field_128():
    """The full-size case fitted: the per-voxel angle field (128^3 pack, 40
    deg + Gaussian random field, FP32, 200 WENO5 TVD-RK3 steps, checkpoints
    every 20) is moved by Gauss-Newton steps on the whole field
    (fit_angle_field) to the saturation another field gives (the same field
    8 deg wider, a target the map reaches); the loss falls by 10^3 or more
    within six reverse passes and the fitted field stays in [0, 90] deg."""
    from lsmpqs_amd.adjoint import fit_angle_field
    from lsmpqs_amd.inputs import contact_angle_case
    steps, b, a0, ck = 200, 0.04, 0.15 * 0.04, 20
    print("case: fit of a 128^3 angle field, fp32 WENO5 RK3 200 steps", flush=True)
    g, mask, phi0, theta = contact_angle_case(128, dtype=np.float32)
    o = default_options(np.float32)
    o.order_space_accur, o.order_time_accur = 5, 3
    r0 = saturation_loss(phi0, mask, theta, g, o, steps, a0, b, 0.45, checkpoint_every=ck,
                         gradient=False)
    shifted = (theta + np.float32(np.radians(8.0))).astype(np.float32)
    target = saturation_loss(phi0, mask, shifted, g, o, steps, a0, b, 0.45, dt=r0.dt,
                             gradient=False).sw
    fitted, losses = fit_angle_field(phi0, mask, theta, g, o, steps, a0, b, target, r0.dt,
                                     iterations=6, checkpoint_every=ck)
    f = fill_of(g)
    moved = np.degrees(np.abs(fitted[f].astype(np.float64) - theta[f]))
    print(f"S_w start {r0.sw:.6f}, target {target:.6f}; losses {['%.3e' % x for x in losses]}; "
          f"angle change max {moved.max():.2f} deg, rms {np.sqrt(np.mean(moved ** 2)):.3f} deg, "
          f"{int((moved > 0.01).sum())} voxels moved by more than 0.01 deg", flush=True)
    assert losses[-1] <= 1e-3 * losses[0]
    assert fitted.min() >= 0.0 and fitted.max() <= np.float32(0.5 * np.pi)


def _theta_fixture():
    import os
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_theta_coeff as mt
    return mt, np.load(os.path.join(here, "golden", "theta_coeff_f64.npz"))


def test_theta_fixture_is_the_reference_output():
    """The committed setVarNorm3dTheta vectors are the reference build's (CPU)."""
    import oracle_py as op
    if not op.have_ref(np.float64):
        pytest.skip("reference build not present")
    mt, d = _theta_fixture()
    g, mask, _, theta = mt.inputs()
    np.testing.assert_array_equal(mt.reference_var_a(op.ref_lib(np.float64), g, mask, theta),
                                  d["var_a"])


@pytest.mark.gpu
def test_angle_field_speed_matches_reference_setvarnorm3dtheta():
    """The per-voxel contact-angle speed the adjoint's forward map uses
    (vel_n = h1 a0 - s h2 cos(theta) |grad mask|) against the reference
    build's setVarNorm3dTheta on the same mask and angle field
    (tests/golden/theta_coeff_f64.npz, tests/golden/make_theta_coeff.py).
    FP64; cos and the Heaviside's sin are the device's and glibc's, so the
    tolerance is a few ulps of the speed."""
    import ctypes as C
    from lsmpqs_amd import hiplib
    from lsmpqs_amd.abi import Arrays
    from lsmpqs_amd.adjoint import AdjointArgs, _lib
    print("\n[adjoint] setVarNorm3dTheta speed vs reference, 20^3 fp64", flush=True)
    mt, d = _theta_fixture()
    g, mask, phi0, theta = mt.inputs()
    lib = _lib(np.float64)
    A = Arrays(g, np.float64)
    A["mask"] = mask
    A["phi"] = phi0
    A["normal_velocity"] = -7.0
    th = np.ascontiguousarray(theta)
    o = default_options(np.float64)
    o.order_space_accur, o.order_time_accur = 5, 3
    a = AdjointArgs()
    a.nsteps, a.checkpoint_every = 1, 1
    a.dt, a.a0, a.b, a.target_sw = 0.0, float(d["a0"]), B, TARGET
    rc = lib.lsmpqs_adjoint_run(C.addressof(o), C.addressof(A.rec), C.addressof(g),
                                th.ctypes.data, None, C.byref(a))
    lib.lsmpqs_release(A.ptr())
    assert rc == 0, hiplib.last_error(lib)
    f = fill_of(g)
    diff = float(np.max(np.abs(A["normal_velocity"][f] - d["var_a"][f])))
    print(f"  max|vel_n - reference| = {diff:.2e}", flush=True)
    assert diff <= 1e-13
