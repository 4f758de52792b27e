"""The seeded drainage workloads this library is measured and tested on:
their inputs (Boolean sphere packs, inputs.pack_case) and their Options.

A curvature schedule "kappa from k0 to k1 in m steps" is the reference's
drain_model3d loop a0 = a + s b dc, s = 0 .. m-1 (reference:
src/level_set_evolution/drain_model3d.c:102-163) with a = k0 b,
dc = (k1 - k0) / (m - 1) and amax half a step beyond k1 b, so the
equilibrium curvature a0 / b runs from k0 to k1.
"""
import numpy as np

from .abi import default_options
from .inputs import pack_case

# 64^3 pack (radius 8, porosity 0.40, seed 1001), HJ-ENO2 + the reference's
# "TVD-RK2", contact angle 0, 10 curvature steps from kappa 0.05 to 0.30 at
# b = 0.04, reinitialization every 20 time steps, trapping on, FP64.
DRAIN64 = dict(dims=(64, 64, 64), radius=8.0, porosity=0.40, seed=1001, space=2, time=2,
               theta_deg=0.0, b=0.04, kappa=(0.05, 0.30), steps=10, steps_per_reinit=20)

# 256^3 / 512^3 packs (radius U[6,12], porosity 0.36, seed 1002), HJ-WENO5 +
# TVD-RK3, contact angle 30 deg, a drainage curve then an imbibition curve
# of 25 curvature steps each (0.25 b apart from a = 0.02), convergence
# max |dphi| < 1e-4 dx.
WENO5_DRAIN_IMBIBE = dict(radius=(6.0, 12.0), porosity=0.36, seed=1002, space=5, time=3,
                          theta_deg=30.0, a=0.02, b=0.04, dc=0.25, steps=25, eps_stop=1e-4)

# 512^3 / 1024^3 packs (radius U[10,20], porosity 0.35, seed 1003), HJ-WENO5 +
# TVD-RK3, contact angle 39.999 deg (overlap 0.3 dx in both precisions; at
# exactly 40 deg the reference's threshold test rounds differently in its
# FP32 and FP64 builds), 40 curvature steps 0.25 b apart from a = 0.02.
WENO5_PRECISION = dict(radius=(10.0, 20.0), porosity=0.35, seed=1003, space=5, time=3,
                       theta_deg=39.999, a=0.02, b=0.04, dc=0.25, steps=40)


def drain64_inputs(dtype=np.float64):
    w = DRAIN64
    return pack_case(w["dims"], w["radius"], w["porosity"], w["seed"], dtype)


def drain64_options(dtype=np.float64, steps=None):
    """Options of the 64^3 ENO2 drainage.  tplot is 20 time steps: the first
    step's dt at theta 0 is 0.5 / (H + b_max_over_dx) with b_max_over_dx =
    2 b (3 / dx^2) = 0.24 and H = a0 |grad phi|_G / dx, small next to it (a0
    <= 0.012), so dt ~ 0.5 / 0.24 and 20 steps ~ 41.7 pseudo-time units."""
    w = DRAIN64
    o = default_options(dtype)
    o.order_space_accur, o.order_time_accur = w["space"], w["time"]
    o.theta = w["theta_deg"] * 3.14159 / 180
    m = w["steps"] if steps is None else steps
    k0, k1 = w["kappa"]
    o.b = w["b"]
    o.a = k0 * o.b
    o.dc = (k1 - k0) / (w["steps"] - 1)
    o.amax = o.a + o.b * o.dc * (m - 0.5)
    dt0 = 0.5 / (2 * o.b * 3.0)
    o.tplot = w["steps_per_reinit"] * dt0
    o.check_connectivity = 1
    return o


def weno5_options(spec, dtype=np.float64, steps=None, drain=True):
    """Options of a WENO5 workload (WENO5_DRAIN_IMBIBE or WENO5_PRECISION):
    `steps` curvature steps (default the workload's) of a drainage curve, or
    of an imbibition curve from the pressure the drainage ended at."""
    o = default_options(dtype)
    o.order_space_accur, o.order_time_accur = spec["space"], spec["time"]
    o.theta = spec["theta_deg"] * 3.14159 / 180
    o.a, o.b, o.dc = spec["a"], spec["b"], spec["dc"]
    m = spec["steps"] if steps is None else steps
    o.amax = o.a + o.b * o.dc * (m - 0.5)
    if not drain:
        o.a = o.a + o.b * o.dc * (m - 1)
        o.amin = o.a - o.b * o.dc * (m - 0.5)
    if "eps_stop" in spec:
        o.eps_stop = spec["eps_stop"]
    o.check_connectivity = 1
    return o


def weno5_inputs(spec, n, dtype=np.float64):
    return pack_case((n, n, n), spec["radius"], spec["porosity"], spec["seed"], dtype)


def saturation_of(phi, mask, grid, eps=1.5):
    """Non-wetting saturation V(phi < 0) / V(mask < 0) over the fill box with
    the reference's smoothed Heaviside of width eps dx, summed in FP64
    (reference: src/utilities/qss_util3d.f:288-358,
    src/level_set_evolution/constCurvModel3d.c:66-76,292-295)."""
    f = (slice(grid.klo_fb, grid.khi_fb + 1), slice(grid.jlo_fb, grid.jhi_fb + 1),
         slice(grid.ilo_fb, grid.ihi_fb + 1))
    w = eps * float(grid.dx[0])

    def vol(a):
        a = np.asarray(a[f], dtype=np.float64)
        q = a / w
        v = np.where(a < -w, 1.0, np.where(a < w, 0.5 * (1 - q - np.sin(np.pi * q) / np.pi), 0.0))
        return float(v.sum())

    return vol(phi) / vol(mask)
