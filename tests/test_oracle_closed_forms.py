"""Pins of the oracle that do not depend on the reference: closed forms,
invariants and brute force.  Each is chosen so a plausible slip (a dropped
term, a swapped sign or operand, a wrong index) fails it."""
import ctypes as C

import numpy as np
import pytest
from scipy import ndimage

import oracle_py as op
from lsmpqs_amd.abi import make_grid, default_options, Arrays

DT = np.float64


def grid_and_coords(n=12):
    g = make_grid((n, n, n), DT)
    shape = (g.grid_dims_ghostbox[2], g.grid_dims_ghostbox[1], g.grid_dims_ghostbox[0])
    z, y, x = np.meshgrid(*[np.arange(s, dtype=DT) for s in shape], indexing="ij")
    return g, shape, x, y, z


def fill(g):
    return (slice(g.klo_fb, g.khi_fb + 1), slice(g.jlo_fb, g.jhi_fb + 1),
            slice(g.ilo_fb, g.ihi_fb + 1))


def rhs_var(name, phi, vn, b, g, vx=None):
    z = np.zeros_like(phi)
    vx = z if vx is None else vx
    rhs = np.zeros_like(phi)
    mh = op.Scalar(0.0, C.c_double)
    op.fcall(op.oracle_lib(DT), "orc_" + name, DT, phi, vn, b, vx, z, z, mh, rhs,
             *op.gb6(g), *op.fb6(g), 1.0, 1.0, 1.0)
    return rhs, mh.value


@pytest.mark.parametrize("name,sign", [("getrhs3d2nd", -1), ("getrhs3d3rd", -1), ("getrhs3d5th", +1)])
def test_plane_gives_minus_a(name, sign):
    # phi = -(x - x0): |grad phi| = 1, kappa = 0, so rhs = -a |grad phi| for
    # ENO2/ENO3 and +a for the reference's WENO5 (normal term enters with +)
    g, shape, x, y, z = grid_and_coords()
    phi = -(x - 7.3)
    a = 0.37
    rhs, H = rhs_var(name, phi, np.full(shape, a), np.full(shape, 0.05), g)
    assert np.allclose(rhs[fill(g)], sign * a, atol=1e-13)
    # H/dx = |a| / sqrt(1 + dx^2) * (1/dx) at the last voxel
    assert abs(H - a / np.sqrt(2.0)) < 1e-14


def test_advection_term_is_upwinded():
    # plane phi = 2x; V = (vx, 0, 0): rhs = -vx * 2 with either sign of vx
    g, shape, x, y, z = grid_and_coords()
    phi = 2.0 * x
    for vxv in (0.7, -0.4):
        rhs, _ = rhs_var("getrhs3d2nd", phi, np.full(shape, 1e-7), np.zeros(shape), g,
                         vx=np.full(shape, vxv))
        # |vel_n| < 1e-5 everywhere: the normal term is never assigned (0 in the oracle)
        assert np.allclose(rhs[fill(g)], -vxv * 2.0, atol=1e-12)


def test_eno2_exact_on_quadratics_and_eno3_on_cubics():
    g, shape, x, y, z = grid_and_coords()
    # ENO2 one-sided derivatives are exact for quadratics: curvature-free
    # check through the normal term with vn > 0 on phi = (x-c)^2 + (y-d)^2
    # where all upwind pieces are nonnegative (x > c, y > d everywhere)
    phi = (x + 5.0) ** 2 + 0.5 * (y + 4.0) ** 2
    rhs, _ = rhs_var("getrhs3d2nd", phi, np.full(shape, 1.0), np.zeros(shape), g)
    exact = -np.sqrt((2 * (x + 5.0)) ** 2 + (y + 4.0) ** 2)
    assert np.allclose(rhs[fill(g)], exact[fill(g)], rtol=1e-13)
    # ENO3 on a cubic along y (the y line has no index quirk)
    phi = (y + 6.0) ** 3 / 30.0
    rhs, _ = rhs_var("getrhs3d3rd", phi, np.full(shape, -1.0), np.zeros(shape), g)
    exact = np.abs(3 * (y + 6.0) ** 2 / 30.0)  # vn < 0: rhs = +|grad phi|
    assert np.allclose(rhs[fill(g)], exact[fill(g)], rtol=1e-12)


def test_curvature_of_sphere():
    # phi = |x - c| - R: b * kappa |grad phi| with kappa = 2/r for the wide stencil
    g, shape, x, y, z = grid_and_coords(24)
    c = np.array([14.2, 13.7, 15.1])
    r = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
    phi = r - 6.0
    rhs, _ = rhs_var("getrhs3d2nd", phi, np.full(shape, 1e-7), np.full(shape, 1.0), g)
    f = np.zeros(shape, bool)
    f[fill(g)] = True
    band = (np.abs(phi) < 1.0) & f
    # second-order wide differences: relative error O((2 dx / r)^2) ~ 4% at r = 5
    assert np.allclose(rhs[band], 2.0 / r[band], rtol=0.04)


def rk(u, rhs, dt, g, stage):
    lib = op.oracle_lib(DT)
    out = np.zeros_like(u)
    gb, fb = op.gb6(g), op.fb6(g)
    if stage == 1:
        op.fcall(lib, "orc_qss3drk1step", DT, out, *gb, u, *gb, rhs, *gb, *fb, dt)
    return out


def test_rk_amplification_factors():
    # linear test u' = lam u: RK1 -> 1+z; the reference's "TVD-RK2" stage 2
    # (0.75 u + 0.25 (u1 + dt L(u1))) -> 1 + z/2 + z^2/4; RK3 -> 1+z+z^2/2+z^3/6
    g, shape, *_ = grid_and_coords(6)
    lib = op.oracle_lib(DT)
    gb, fb = op.gb6(g), op.fb6(g)
    lam, dt = -0.8, 0.3
    zz = lam * dt
    u = np.full(shape, 1.0)
    u1 = np.zeros(shape); u2 = np.zeros(shape); u3 = np.zeros(shape)
    op.fcall(lib, "orc_qss3drk1step", DT, u1, *gb, u, *gb, lam * u, *gb, *fb, dt)
    op.fcall(lib, "orc_qss3dtvdrk2stage2", DT, u2, *gb, u1, *gb, u, *gb, lam * u1, *gb, *fb, dt)
    f = fill(g)
    assert np.allclose(u1[f], 1 + zz, rtol=1e-15)
    assert np.allclose(u2[f], 1 + zz / 2 + zz ** 2 / 4, rtol=1e-15)
    # genuine TVD-RK3: stage 2 with 0.75/0.25 applied to the RK1 result, then stage 3
    op.fcall(lib, "orc_qss3dtvdrk3stage3", DT, u3, *gb, u2, *gb, u, *gb, lam * u2, *gb, *fb, dt)
    # the reference uses its stage-2 routine for RK3's second stage too, so
    # u2 = u(0.75 + 0.25(1+z)^2) and u3 = u/3 + 2/3 u2 (1+z)
    u2c = 0.75 + 0.25 * (1 + zz) ** 2
    assert np.allclose(u3[f], 1 / 3 + 2 / 3 * u2c * (1 + zz), rtol=1e-15)
    assert abs((1 / 3 + 2 / 3 * u2c * (1 + zz)) - (1 + zz + zz ** 2 / 2 + zz ** 3 / 6)) < 1e-15


def test_signed_extrapolation_exact_on_distance_like_field():
    # phi = +-(5 + |x-xm| + |y-ym| + |z-zm|) grows in magnitude away from the
    # box centre with unit slope, which is what signed linear extrapolation
    # continues exactly, including edges and corners
    g, shape, x, y, z = grid_and_coords(10)
    m = [(g.ilo_fb + g.ihi_fb) / 2 + 0.25, (g.jlo_fb + g.jhi_fb) / 2 + 0.25,
         (g.klo_fb + g.khi_fb) / 2 + 0.25]
    for s in (1.0, -1.0):
        exact = s * (5 + np.abs(x - m[0]) + np.abs(y - m[1]) + np.abs(z - m[2]))
        phi = np.zeros(shape)
        phi[fill(g)] = exact[fill(g)]
        op.oracle_lib(DT).orc_signedLinearExtrapolationBCqss(phi.ctypes.data_as(C.c_void_p),
                                                             C.byref(g), 9)
        assert np.allclose(phi, exact, atol=1e-12)


def test_reinit_fixed_point_on_plane_and_iteration_count():
    g, shape, x, y, z = grid_and_coords(10)
    o = default_options(DT)
    phi = x - 7.37
    d0 = np.zeros(shape)
    op.fcall(op.oracle_lib(DT), "orc_qss3dcomputedistanceforsubcellfix", DT, d0, phi,
             *op.gb6(g), *op.fb6(g), 1.0, 1.0, 1.0)
    near = np.abs(phi) < 1
    f = np.zeros(shape, bool)
    f[fill(g)] = True
    assert np.allclose(d0[near & f], phi[near & f], atol=1e-15)
    assert np.all(d0[~near & f] == -1000.0)
    out = phi.copy()
    op.oracle_lib(DT).orc_reinitializeSubcellFix3d(out.ctypes.data_as(C.c_void_p), C.byref(g),
                                                   C.byref(o))
    assert np.allclose(out[fill(g)], phi[fill(g)], atol=1e-13)
    # tmax_r = 0.4 at dx = 1 and dt_r = 0.5: exactly one iteration; a field
    # that is not a distance moves by exactly dt_r * rhs
    sq = 3.0 * (x - 7.37)
    out = sq.copy()
    op.oracle_lib(DT).orc_reinitializeSubcellFix3d(out.ctypes.data_as(C.c_void_p), C.byref(g),
                                                   C.byref(o))
    away = (np.abs(sq) > 4) & f
    # away from the interface rhs = sgn (1 - |grad|) = sgn * (1 - 3)
    assert np.allclose(out[away], sq[away] + 0.5 * np.sign(sq[away]) * (1 - 3), atol=1e-12)


def test_volume_closed_forms():
    g, shape, x, y, z = grid_and_coords(20)
    lib = op.oracle_lib(DT)
    vol = op.Scalar(0.0, C.c_double)
    op.fcall(lib, "orc_qss3dvolumeregionphilessthanzero", DT, vol, np.full(shape, -5.0),
             *op.gb6(g), *op.fb6(g), 1.0, 1.0, 1.0, 1.5)
    assert vol.value == 20 ** 3
    c = 13.1
    phi = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2) - 7.0
    op.fcall(lib, "orc_qss3dvolumeregionphilessthanzero", DT, vol, phi, *op.gb6(g), *op.fb6(g),
             1.0, 1.0, 1.0, 1.5)
    # integral of 1 - H_eps(r - R) over space: 4/3 pi R^3 + 4 pi R eps^2 (1/3 - 2/pi^2)
    R, eps = 7.0, 1.5
    want = 4 / 3 * np.pi * R ** 3 + 4 * np.pi * R * eps ** 2 * (1 / 3 - 2 / np.pi ** 2)
    assert abs(vol.value - want) / want < 2e-3
    h = op.Scalar(0.0, C.c_double)
    for xv, want in ((0.0, 0.5), (-1.5, 0.0), (1.5, 1.0), (2.0, 1.0), (-3.0, 0.0)):
        op.fcall(lib, "orc_qssheaviside", DT, h, xv, 1.5)
        assert abs(h.value - want) < 1e-15


def test_maxnorm_band_and_empty():
    g, shape, *_ = grid_and_coords(8)
    lib = op.oracle_lib(DT)
    m = op.Scalar(0.0, C.c_double)
    op.fcall(lib, "orc_qss3dmaxnormdifflocal", DT, m, np.full(shape, 5.0), *op.gb6(g),
             np.zeros(shape), *op.gb6(g), *op.fb6(g), 0.8)
    assert m.value == -1.0
    a = np.full(shape, 0.5)
    b = a.copy()
    b[7, 6, 5] = -0.25
    op.fcall(lib, "orc_qss3dmaxnormdifflocal", DT, m, a, *op.gb6(g), b, *op.gb6(g), *op.fb6(g), 0.8)
    assert m.value == 0.75


@pytest.mark.parametrize("p,seed", [(0.15, 1), (0.3, 2), (0.5, 3)])
def test_labels_equal_bruteforce_26(p, seed):
    g, shape, *_ = grid_and_coords(9)
    A = Arrays(g, DT)
    rng = np.random.default_rng(seed)
    A["phi_bin"] = (rng.random(shape) < p).astype(np.uint8)
    op.oracle_lib(DT).orc_findConnectivity3d(A.ptr(), C.byref(g))
    lab, n = ndimage.label(A["phi_bin"], structure=np.ones((3, 3, 3)))
    flat = np.arange(lab.size).reshape(shape)
    want = np.zeros(shape, dtype=np.int64)
    for c in range(1, n + 1):
        sel = lab == c
        want[sel] = flat[sel].max()
    assert np.array_equal(A["connectivity"], want)


@pytest.mark.parametrize("dtype,table", [
    (np.float64, {10: 0.0, 20: 0.3, 30: 0.3, 40: 0.5, 50: 1.0, 60: 1.0}),
    (np.float32, {10: 0.0, 20: 0.0, 30: 0.3, 40: 0.3, 50: 1.0, 60: 1.0}),
])
def test_overlap_thresholds(dtype, table):
    g = make_grid((8, 8, 8), dtype)
    for deg, want in table.items():
        o = default_options(dtype)
        o.theta = deg * 3.14159 / 180
        op.oracle_lib(dtype).orc_setThetaOverlap(C.byref(o), C.byref(g))
        assert abs(o.overlap - want) < 1e-6, (deg, o.overlap)


# ---------------------------------------------------------------------------
# The textbook ("fixed") scheme behind the oracle's quirk switches
# (orc_set_quirks): the batched subvolumes and the adjoint run with the
# reference's quirks off, so each switch is pinned here on its own.

class quirks:
    """Context manager: the oracle's quirk bits set to `bits` inside."""

    def __init__(self, bits):
        self.bits = bits

    def __enter__(self):
        lib = op.oracle_lib(DT)
        lib.orc_set_quirks.restype = C.c_uint
        lib.orc_set_quirks.argtypes = [C.c_uint]
        self.old = lib.orc_set_quirks(self.bits)

    def __exit__(self, *exc):
        op.oracle_lib(DT).orc_set_quirks(self.old)


def noncubic(nx=14, ny=11, nz=7):
    g = make_grid((nx, ny, nz), DT)
    shape = (g.grid_dims_ghostbox[2], g.grid_dims_ghostbox[1], g.grid_dims_ghostbox[0])
    z, y, x = np.meshgrid(*[np.arange(s, dtype=DT) for s in shape], indexing="ij")
    return g, shape, x, y, z


@pytest.mark.parametrize("name", ["getrhs3d2nd", "getrhs3d3rd", "getrhs3d5th"])
def test_fixed_mode_plane_gives_minus_a_for_every_scheme(name):
    """Quirks off: a tilted plane (|grad phi| = 1, kappa = 0) gives rhs = -a
    for WENO5 as for ENO2/ENO3; with the sign switch back on WENO5 gives +a
    (so the switch is what flips it)."""
    g, shape, x, y, z = noncubic()
    n = np.array([0.48, -0.6, 0.64])
    n /= np.linalg.norm(n)
    phi = n[0] * x + n[1] * y + n[2] * z - 3.1
    a = 0.37
    with quirks(0):
        rhs, _ = rhs_var(name, phi, np.full(shape, a), np.full(shape, 0.05), g)
    assert np.allclose(rhs[fill(g)], -a, atol=1e-12)
    if name == "getrhs3d5th":
        with quirks(2):
            rhs, _ = rhs_var(name, phi, np.full(shape, a), np.full(shape, 0.05), g)
        assert np.allclose(rhs[fill(g)], +a, atol=1e-12)


def test_fixed_mode_eno3_exact_on_cubic_along_x_noncubic_box():
    """Quirks off: ENO3 is exact for a cubic along x on a box whose y range
    exceeds its z range; with the read switch on, the x line takes
    phi(i+3,j,j), which is another voxel's value wherever j != k, and the
    result is wrong there (and only there)."""
    g, shape, x, y, z = noncubic()
    # increasing in x with |phi_xx| falling to the right, so the upwind side
    # (vn < 0 takes D+) widens its stencil to the right, where phi(i+3, .)
    # enters through the rightmost third difference (0.02 for this cubic); a
    # small slope in z makes the misread value lower that difference, and
    # ENO then picks it, for j - k in 1..3
    phi = -(40.0 - x) ** 3 / 300.0 - 0.01 * z
    vn = np.full(shape, -1.0)                   # vn < 0: rhs = +|grad phi|
    exact = np.sqrt((3 * (40.0 - x) ** 2 / 300.0) ** 2 + 0.01 ** 2)
    with quirks(0):
        rhs, _ = rhs_var("getrhs3d3rd", phi, vn, np.zeros(shape), g)
    f = fill(g)
    assert np.allclose(rhs[f], exact[f], rtol=1e-12)
    with quirks(1):
        rhs_q, _ = rhs_var("getrhs3d3rd", phi, vn, np.zeros(shape), g)
    jj, kk = y[f], z[f]
    off = ~np.isclose(rhs_q[f], exact[f], rtol=1e-9)
    assert off.any()
    assert not (off & (jj == kk)).any()


def test_fixed_mode_small_speed_has_no_normal_term():
    """Quirks off: where |vel_n| < 1e-5 the normal term is 0 (flat phi in
    x, no curvature, no advection: rhs = 0 there); with the stale switch on
    those voxels repeat the previous voxel's -vel_n |grad phi|."""
    g, shape, x, y, z = noncubic()
    phi = x - 5.2
    vn = np.full(shape, 0.8)
    vn[:, :, 8] = 3e-6                          # one x plane of tiny speeds
    with quirks(0):
        rhs, _ = rhs_var("getrhs3d2nd", phi, vn, np.zeros(shape), g)
    inside = _fillmask(g, shape)
    small = np.zeros(shape, bool)
    small[:, :, 8] = True
    assert np.all(rhs[small & inside] == 0.0)
    assert np.allclose(rhs[~small & inside], -0.8, atol=1e-12)
    with quirks(4):
        rhs_q, _ = rhs_var("getrhs3d2nd", phi, vn, np.zeros(shape), g)
    assert np.allclose(rhs_q[small & inside], -0.8, atol=1e-12)


def _fillmask(g, shape):
    m = np.zeros(shape, bool)
    m[fill(g)] = True
    return m


def test_six_connected_labels_match_scipy():
    """orc_set_connectivity(6): face-neighbour components, each labelled with
    its largest linear index, against scipy's 6-connected labelling (and
    different from the 26-connected partition on the same field)."""
    g = make_grid((13, 11, 9), DT)
    A = Arrays(g, DT)
    rng = np.random.default_rng(5)
    A["phi_bin"] = (rng.random(A.shape) < 0.35).astype(np.uint8)
    lib = op.oracle_lib(DT)
    res = {}
    for conn, struct in ((6, None), (26, np.ones((3, 3, 3)))):
        old = lib.orc_set_connectivity(conn)
        try:
            lib.orc_findConnectivity3d(A.ptr(), C.byref(g))
        finally:
            lib.orc_set_connectivity(old)
        lab, nlab = ndimage.label(A["phi_bin"], structure=struct)
        flat = lab.ravel()
        top = np.full(nlab + 1, -1, dtype=np.int64)
        np.maximum.at(top, flat, np.arange(flat.size))
        want = np.where(flat > 0, top[flat], 0).reshape(lab.shape)
        np.testing.assert_array_equal(A["connectivity"], want)
        res[conn] = A["connectivity"].copy()
    assert not np.array_equal(res[6], res[26])
