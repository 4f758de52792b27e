"""Gradient of a wetting-saturation loss through the level-set time steps.

`saturation_loss` runs the library's lsmpqs_adjoint_run: nsteps TVD-RK steps
of the variable-coefficient level-set equation (the reference's
constCurvModel3d inner step, src/level_set_evolution/constCurvModel3d.c:127-230,
with a per-voxel contact angle as in setVarNorm3dTheta,
src/spatial_derivatives/qss_spatial_derivatives3d.f:94-170) at a fixed time
step, then loss = (S_w - target)^2, and returns the loss with its gradient
with respect to the contact-angle field and the curvature coefficient b,
computed by the library's hand-written reverse pass on the GPU.

`ThetaLoss` wraps it as a torch.autograd.Function so a contact-angle field
(and b) can be fitted with torch optimisers; `fit_angle_field` fits a
per-voxel angle field to a target saturation by Gauss-Newton steps.  Tensors travel to the library
as host arrays over the ghost box; there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import hiplib
from .abi import types_for


class AdjointArgs(C.Structure):
    _fields_ = [("nsteps", C.c_int), ("checkpoint_every", C.c_int), ("dt", C.c_double),
                ("a0", C.c_double), ("b", C.c_double), ("target_sw", C.c_double),
                ("phi_cap", C.c_double), ("loss", C.c_double), ("sw", C.c_double),
                ("dt_used", C.c_double), ("dloss_db", C.c_double), ("cap_used", C.c_double)]


def _lib(dtype):
    lib = hiplib.load(dtype)
    f = lib.lsmpqs_adjoint_run
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.POINTER(AdjointArgs)]
    return lib


class AdjointResult:
    """loss, sw, dt, dloss_db (floats); dtheta and phi (ghost-box arrays)."""

    def __init__(self, args, dtheta, phi):
        self.loss = args.loss
        self.sw = args.sw
        self.dt = args.dt_used
        self.dloss_db = args.dloss_db
        self.cap = args.cap_used
        self.dtheta = dtheta
        self.phi = phi


def saturation_loss(phi0, mask, theta, grid, options, nsteps, a0, b, target_sw, dt=0.0,
                    checkpoint_every=None, gradient=True, phi_cap=0.0):
    """Run the forward map and its reverse pass.

    phi0, mask, theta: arrays over the ghost box (shape nz+2g, ny+2g, nx+2g)
    of the library's QSSLIB_REAL; options: the Options record (scheme from
    order_space_accur, RK order from order_time_accur, overlap,
    eps_coefficient); dt <= 0 takes the CFL rule on phi0; phi is capped at
    phi_cap after every step (<= 0: 6 dx).  Raises
    hiplib.LsmpqsError on a library error."""
    dtype = np.dtype(phi0.dtype)
    lib = _lib(dtype)
    _, _, DataArrays = types_for(dtype)
    phi = np.ascontiguousarray(phi0, dtype=dtype).copy()
    mask = np.ascontiguousarray(mask, dtype=dtype)
    theta = np.ascontiguousarray(theta, dtype=dtype)
    if not (phi.shape == mask.shape == theta.shape) or phi.size != grid.num_gridpts:
        raise ValueError("phi0, mask and theta must cover the grid's ghost box")
    rec = DataArrays()
    rec.phi = phi.ctypes.data
    rec.mask = mask.ctypes.data
    dtheta = np.zeros_like(phi) if gradient else None
    a = AdjointArgs()
    a.nsteps = int(nsteps)
    a.checkpoint_every = int(checkpoint_every or max(1, int(round(np.sqrt(nsteps)))))
    a.dt, a.a0, a.b, a.target_sw = float(dt), float(a0), float(b), float(target_sw)
    a.phi_cap = float(phi_cap)
    try:
        rc = lib.lsmpqs_adjoint_run(C.addressof(options), C.addressof(rec), C.addressof(grid),
                                    theta.ctypes.data,
                                    dtheta.ctypes.data if gradient else None, C.byref(a))
        if rc != 0:
            raise hiplib.LsmpqsError(hiplib.last_error(lib))
    finally:
        # the device session is keyed on this temporary record: free it on
        # every path, so a later record at the same address starts clean
        lib.lsmpqs_release(C.c_void_p(C.addressof(rec)))
    return AdjointResult(a, dtheta, phi)


def fit_angle_field(phi0, mask, theta0, grid, options, nsteps, a0, b, target_sw, dt,
                    iterations=6, checkpoint_every=None, lo=0.0, hi=0.5 * np.pi, region=None):
    """Fit a per-voxel contact-angle field to a target wetting saturation by
    Gauss-Newton steps on the whole field: with the residual r = S_w - target
    and g = dS_w/dtheta = (dloss/dtheta) / (2 r) from the reverse pass, the
    minimum-norm step theta -= r g / (g . g), clipped to [lo, hi], on the
    voxels of `region` (a boolean mask over the ghost box; default the fill
    box, where the gradient lives).  Stops early when the loss is 0 or the
    gradient vanishes.  Returns (theta, [loss of each evaluated field])."""
    theta = np.array(theta0, copy=True)
    if region is None:
        region = np.zeros(theta.shape, dtype=bool)
        region[grid.klo_fb:grid.khi_fb + 1, grid.jlo_fb:grid.jhi_fb + 1,
               grid.ilo_fb:grid.ihi_fb + 1] = True
    losses = []
    for _ in range(iterations):
        r = saturation_loss(phi0, mask, theta, grid, options, nsteps, a0, b, target_sw, dt=dt,
                            checkpoint_every=checkpoint_every)
        losses.append(r.loss)
        res = r.sw - target_sw
        if res == 0.0:
            break
        g = np.where(region, r.dtheta.astype(np.float64) / (2.0 * res), 0.0)
        gg = float(np.sum(g * g))
        if gg == 0.0:
            break
        step = theta.astype(np.float64) - (res / gg) * g
        theta = np.clip(step, lo, hi).astype(theta.dtype)
    else:
        losses.append(saturation_loss(phi0, mask, theta, grid, options, nsteps, a0, b, target_sw,
                                      dt=dt, gradient=False).loss)
    return theta, losses


def _torch():
    import torch
    return torch


class ThetaLoss:
    """loss = ThetaLoss.apply(theta, b, problem): theta a tensor over the ghost
    box, b a 0-d tensor; problem a dict with phi0, mask, grid, options, nsteps,
    a0, target_sw and optionally dt, checkpoint_every.  Gradients flow to
    theta and b."""

    @staticmethod
    def apply(theta, b, problem):
        return _function().apply(theta, b, problem)


_FN = None


def _function():
    global _FN
    if _FN is not None:
        return _FN
    torch = _torch()

    class _ThetaLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, b, problem):
            dtype = np.dtype(problem["phi0"].dtype)
            th = theta.detach().to("cpu", dtype=torch.float64).numpy().astype(dtype)
            r = saturation_loss(problem["phi0"], problem["mask"], th, problem["grid"],
                                problem["options"], problem["nsteps"], problem["a0"],
                                float(b.detach().cpu()), problem["target_sw"],
                                dt=problem.get("dt", 0.0),
                                checkpoint_every=problem.get("checkpoint_every"))
            problem["last"] = r
            ctx.save_for_backward(torch.from_numpy(r.dtheta.astype(np.float64)).to(theta.device),
                                  torch.tensor(r.dloss_db, dtype=torch.float64, device=b.device))
            return torch.tensor(r.loss, dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, grad):
            dth, db = ctx.saved_tensors
            return (grad * dth).to(grad.dtype), (grad * db).to(grad.dtype), None

    _FN = _ThetaLoss
    return _FN
