"""Writes tests/golden/theta_coeff_f64.npz by running the REFERENCE build
(oracle/_ref/libqssref_f64.so, compiled from /root/reference by
oracle/Makefile.ref): the per-voxel contact-angle speed of
setVarNorm3dTheta (src/spatial_derivatives/qss_spatial_derivatives3d.f:94-170),
vel_n = h1 a0 - s h2 cos(theta) |grad mask|, on the mask of a seeded 20^3
pack with a seeded angle field, after getCentralGrad3dOrder2 (:2987-3048)
of the mask.

    python tests/golden/make_theta_coeff.py

The inputs are regenerated from the seeds by inputs(); the fixture holds
var_a over the ghost box (0 off the fill box) and a0 (data only).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "lsmpqs-1.0_amd"), os.path.join(ROOT, "oracle"),
                os.path.join(ROOT, "tests")]

import oracle_py as op  # noqa: E402
from lsm_cases import case  # noqa: E402

DT = np.float64
A0 = 0.05
OUT = os.path.join(HERE, "theta_coeff_f64.npz")


def inputs():
    """(grid, mask, phi0, theta): the 20^3 pack of seed 3100 and an angle
    field uniform in [0.4, 1.0) rad (seed 7)."""
    g, mask, phi0 = case(20, DT, seed=3100)
    theta = 0.4 + 0.6 * np.random.default_rng(7).random(mask.shape)
    return g, mask, phi0, theta


def reference_var_a(lib, g, mask, theta):
    gb, fb = op.gb6(g), op.fb6(g)
    gx, gy, gz = (np.zeros_like(mask) for _ in range(3))
    op.fcall(lib, "getcentralgrad3dorder2_", DT, gx, gy, gz, *gb, mask, *gb, *fb, 1.0, 1.0, 1.0)
    va = np.zeros_like(mask)
    op.fcall(lib, "setvarnorm3dtheta_", DT, mask, gx, gy, gz, va, A0, theta, *gb, *fb, 1.0)
    return va


def main():
    g, mask, _, theta = inputs()
    va = reference_var_a(op.ref_lib(DT), g, mask, theta)
    np.savez_compressed(OUT, var_a=va, a0=A0)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
