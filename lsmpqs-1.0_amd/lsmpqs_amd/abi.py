"""ctypes mirrors of the LSMPQS C-ABI records (Grid, Options, QSS_DataArrays).

The layouts follow the reference headers field for field
(reference: src/geometry/qss_grid.h:29-81, src/utilities/qss_options.h:34-168,
src/utilities/qss_data_arrays.h:24-73) for both values of QSSLIB_REAL, so the
same Python objects can be handed to the HIP library, the oracle and the
reference build.  Arrays stay in numpy; the records hold raw pointers.
"""
import ctypes as C

import numpy as np

_CACHE = {}


def real_ctype(dtype):
    return C.c_double if np.dtype(dtype) == np.float64 else C.c_float


def types_for(dtype):
    """Return (Grid, Options, DataArrays) ctypes classes for QSSLIB_REAL = dtype."""
    dtype = np.dtype(dtype)
    if dtype in _CACHE:
        return _CACHE[dtype]
    R = real_ctype(dtype)
    P = C.c_void_p

    class Grid(C.Structure):
        _fields_ = [
            ("num_dims", C.c_int),
            ("x_lo", R * 3), ("x_hi", R * 3),
            ("x_lo_ghostbox", R * 3), ("x_hi_ghostbox", R * 3),
            ("grid_dims", C.c_int * 3), ("grid_dims_ghostbox", C.c_int * 3),
            ("dx", R * 3),
            ("num_gridpts", C.c_int),
        ] + [(n, C.c_int) for n in (
            "ilo_gb", "ihi_gb", "jlo_gb", "jhi_gb", "klo_gb", "khi_gb",
            "ilo_fb", "ihi_fb", "jlo_fb", "jhi_fb", "klo_fb", "khi_fb",
            "ilo_D1_fb", "ihi_D1_fb", "jlo_D1_fb", "jhi_D1_fb", "klo_D1_fb", "khi_D1_fb",
            "ilo_D2_fb", "ihi_D2_fb", "jlo_D2_fb", "jhi_D2_fb", "klo_D2_fb", "khi_D2_fb",
            "ilo_D3_fb", "ihi_D3_fb", "jlo_D3_fb", "jhi_D3_fb", "klo_D3_fb", "khi_D3_fb",
            "num_nb_levels")] + [
            ("mark_gb", C.c_ubyte), ("mark_D1", C.c_ubyte), ("mark_D2", C.c_ubyte),
            ("mark_D3", C.c_ubyte), ("mark_fb", C.c_ubyte),
            ("beta", R), ("gamma", R),
        ]

    class Options(C.Structure):
        _fields_ = [
            ("outfile", C.c_char * 256), ("path", C.c_char * 256),
            ("dx", R), ("tmax", R), ("tmax_r", R), ("tplot", R),
            ("accuracy", C.c_char * 20), ("accuracy_id", C.c_int),
            ("a", R), ("b", R),
            ("save_data", C.c_int), ("do_reinit", C.c_int), ("do_mask", C.c_int),
            ("eps_stop", R), ("subcell_fix", C.c_int), ("reinit_func", P),
            ("dc", R), ("cmax", R), ("amax", R), ("amin", R),
            ("vol_frac_max", R), ("vol_frac_min", R),
            ("mask_reinit", C.c_int), ("narrow_band", C.c_int), ("init_step", C.c_int),
            ("overlap", R), ("stop_touch", C.c_int), ("eps_coefficient", R),
            ("extrapol_id", C.c_int), ("extrapol_func", P), ("space_deriv_func", P),
            ("order_space_accur", C.c_int), ("order_time_accur", C.c_int),
            ("checkpoint", C.c_int), ("reservoir_inlet", C.c_int),
            ("theta", R), ("C", R), ("b_max_over_dx", R), ("max_U_over_dx", R),
            ("check_connectivity", C.c_int), ("phi_w_ind", C.c_int), ("phi_nw_ind", C.c_int),
            ("err_check_zone", R), ("use_var_theta", C.c_int),
            ("theta_1", R), ("theta_2", R),
            ("print_details", C.c_int), ("use_satn_stop", C.c_int),
            ("conserve_imbibe", C.c_int), ("center_inlet", C.c_int),
        ]

    class DataArrays(C.Structure):
        _fields_ = [(n, P) for n in (
            "phi", "phi_next", "phi_extra", "phi_prev", "mask", "lse_rhs",
            "normal_velocity", "curvature_coeff",
            "external_velocity_x", "external_velocity_y", "external_velocity_z",
            "mask_x", "mask_y", "mask_z", "scratch1", "scratch2",
            "connectivity", "phi_bin", "mask_w", "mask_nw", "mask_disconn_init",
            "theta", "overlap")]

    _CACHE[dtype] = (Grid, Options, DataArrays)
    return _CACHE[dtype]


REAL_FIELDS = ("phi", "phi_next", "phi_extra", "phi_prev", "mask", "lse_rhs",
               "normal_velocity", "curvature_coeff",
               "external_velocity_x", "external_velocity_y", "external_velocity_z",
               "mask_x", "mask_y", "mask_z", "scratch1", "scratch2",
               "mask_w", "mask_nw", "mask_disconn_init", "theta", "overlap")


def make_grid(dims, dtype=np.float64, dx=1.0, ghost=3):
    """Grid as createMaskUBC builds it: MEDIUM accuracy (3 ghost layers, ENO2
    fill box [3, n+2]) and x_lo = 3 dx (reference:
    src/applications/createMaskUBC.c:63,90-94; src/geometry/qss_grid.c:52-112,
    874-895)."""
    Grid, _, _ = types_for(dtype)
    g = Grid()
    g.num_dims = 3
    for a in range(3):
        n = int(dims[a])
        g.dx[a] = dx
        g.grid_dims[a] = n
        g.x_lo[a] = ghost * dx
        g.x_hi[a] = ghost * dx + n * dx
        g.grid_dims_ghostbox[a] = n + 2 * ghost
        g.x_lo_ghostbox[a] = g.x_lo[a] - ghost * dx
        g.x_hi_ghostbox[a] = g.x_hi[a] + ghost * dx
    N = [g.grid_dims_ghostbox[a] for a in range(3)]
    g.num_gridpts = N[0] * N[1] * N[2]
    g.ilo_gb, g.ihi_gb = 0, N[0] - 1
    g.jlo_gb, g.jhi_gb = 0, N[1] - 1
    g.klo_gb, g.khi_gb = 0, N[2] - 1
    for lvl in (1, 2, 3):
        for ax, nm in enumerate("ijk"):
            setattr(g, f"{nm}lo_D{lvl}_fb", lvl)
            setattr(g, f"{nm}hi_D{lvl}_fb", N[ax] - 1 - lvl)
    for ax, nm in enumerate("ijk"):
        setattr(g, f"{nm}lo_fb", 3)
        setattr(g, f"{nm}hi_fb", N[ax] - 4)
    return g


def default_options(dtype=np.float64):
    """Options with the reference defaults (reference: src/utilities/qss_options.c:42-117)."""
    _, Options, _ = types_for(dtype)
    o = Options()
    o.outfile = b"out_file"
    o.path = b"./"
    o.dx = 0.04
    o.tmax = 150.0
    o.tplot = 0.5
    o.accuracy = b"medium"
    o.accuracy_id = 1
    o.a = 0.15
    o.b = 0.04
    o.save_data = 1
    o.do_reinit = 1
    o.do_mask = 1
    o.eps_stop = 0.014
    o.narrow_band = 0
    o.overlap = 0
    o.init_step = 0
    o.print_details = 1
    o.tmax_r = 10 * o.dx
    o.subcell_fix = 1
    o.dc = 0.1
    o.cmax = 1.0 / (o.dx / 2.0)
    o.amax = 10
    o.amin = 0.05
    o.vol_frac_max = 0.97
    o.vol_frac_min = 0
    o.mask_reinit = 1
    o.stop_touch = 1
    o.eps_coefficient = 1.5
    o.extrapol_id = 0
    o.checkpoint = 0
    o.reservoir_inlet = 1
    o.theta = 0
    o.C = 1
    o.check_connectivity = 1
    o.err_check_zone = 20 * o.dx
    o.phi_w_ind = 0
    o.phi_nw_ind = 0
    o.use_var_theta = 0
    o.order_space_accur = 3
    o.order_time_accur = 3
    o.use_satn_stop = 1
    o.conserve_imbibe = 0
    o.center_inlet = 0
    return o


class Arrays:
    """numpy storage for every QSS_DataArrays field over one ghost box, plus
    the ctypes record pointing at it."""

    def __init__(self, grid, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        _, _, DataArrays = types_for(dtype)
        n = grid.num_gridpts
        shape = (grid.grid_dims_ghostbox[2], grid.grid_dims_ghostbox[1],
                 grid.grid_dims_ghostbox[0])
        self.shape = shape
        self.a = {}
        for f in REAL_FIELDS:
            self.a[f] = np.zeros(shape, dtype=self.dtype)
        self.a["connectivity"] = np.zeros(shape, dtype=np.int32)
        self.a["phi_bin"] = np.zeros(shape, dtype=np.uint8)
        self.rec = DataArrays()
        for f, arr in self.a.items():
            assert arr.size == n
            setattr(self.rec, f, arr.ctypes.data)

    def __getitem__(self, name):
        return self.a[name]

    def __setitem__(self, name, value):
        self.a[name][...] = value

    def ptr(self):
        return C.byref(self.rec)
