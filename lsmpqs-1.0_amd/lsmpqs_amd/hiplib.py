"""Loader of the gfx950 C-ABI library (liblsmpqs_hip_f32.so / _f64.so).

There is no CPU fallback: loading fails loudly when the shared object is
missing, and every entry point that reports an error raises.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(PKG, "lib")

_LIBS = {}


class LsmpqsError(RuntimeError):
    pass


class Timing(C.Structure):
    _fields_ = [("profile", C.c_int), ("elapsed_ms", C.c_double),
                ("stage_ms", C.c_double * 3), ("stage_launches", C.c_longlong * 3),
                ("inner_steps", C.c_longlong), ("last_dt", C.c_double), ("nstale", C.c_int),
                ("outer_steps", C.c_longlong), ("inner_loop_ms", C.c_double)]


class BatchStats(C.Structure):
    _fields_ = [("boxes", C.c_int), ("batched_steps", C.c_longlong), ("box_steps", C.c_longlong),
                ("brick_depth", C.c_int), ("workgroups", C.c_int), ("batched_ms", C.c_double),
                ("autotune_best_ms", C.c_double), ("batched_outer", C.c_longlong),
                ("box_outer", C.c_longlong), ("batched_outer_ms", C.c_double)]


# every exported entry point declared in include/lsmpqs_hip.h and qss_data_arrays.h
EXPORTS = (
    "constCurvModel3d", "constCurvModel3dNoVar", "drain_model3d", "imbibe_model3d",
    "trapComponents_mask", "findConnectivity3d", "getMainInd_new", "reinitializeSubcellFix3d",
    "drain_model2d", "imbibe_model2d", "constCurvModel2d", "constCurvModel2dNoVar",
    "reinitializeSubcellFix2d", "findConnectivity2d",
    "getrhs3d2ndconst_", "getrhs3d3rdconst_", "getrhs3d2nd_", "getrhs3d3rd_", "getrhs3d5th_",
    "qss3drk1step_", "qss3dtvdrk2stage2_", "qss3dtvdrk3stage3_", "signedlinearextrapolation3d_",
    "imposemaskparallel_", "qss3dvolumeregionphilessthanzero_", "qss3dmaxnormdifflocal_",
    "getcentralgrad3dorder2_", "setvarnorm3d_", "setvarcurvadv3d_",
    "qss3dcomputedistanceforsubcellfix_", "qss3dcomputereinitializationeqnrhssubcellfixorder1_",
    "qss2dgetrhs2nd_", "qss2dgetrhs3rd_", "qss2dgetrhs2ndconst_", "qss2dgetrhs3rdconst_",
    "qss2drk1step_", "qss2dtvdrk2stage2_", "qss2dtvdrk3stage3_", "qss2dsignedlinearextrapolation_",
    "signedLinearExtrapolationBCqss", "initializeDisconnectedMasks", "createReservoirInlet3d",
    "setSpaceDerivFunc", "setThetaOverlap", "createOptionsDefault", "destroyOptions",
    "writeDataArrayQSS", "readDataArrayQSS", "allocateQSSDataArrays",
    "allocateMemoryForQSSDataArrays", "destroyQSSDataArrays", "lsmpqs_hip_last_error",
    "lsmpqs_set_dt_slabs", "lsmpqs_set_quirks", "lsmpqs_release", "lsmpqs_bench_inner_steps",
    "lsmpqs_download", "lsmpqs_model_prepare", "lsmpqs_session_stats",
    "drainTop", "imbibeTop", "createOptionsFromInputFile", "printOptions", "createGridSetDx",
    "createGridSetDxDyDz", "createGridSetGridDims", "destroyGrid", "printGrid",
    "writeGridToBinaryFile", "readGridFromBinaryFile", "reinitializeMedium3d",
    "lsmpqs_batch_model3d", "lsmpqs_adjoint_run", "writeDataArrayRaw", "lsmpqs_mask_from_image",
    "lsmpqs_init_from_mask", "lsmpqs_diff_steps", "lsmpqs_volume_noalloc",
)


def lib_path(dtype):
    suf = "f64" if np.dtype(dtype) == np.float64 else "f32"
    return os.path.join(LIBDIR, f"liblsmpqs_hip_{suf}.so")


def load(dtype):
    """ctypes handle of the library for QSSLIB_REAL = dtype (raises if not built)."""
    dtype = np.dtype(dtype)
    if dtype in _LIBS:
        return _LIBS[dtype]
    path = lib_path(dtype)
    if not os.path.exists(path):
        raise LsmpqsError(f"{path} is not built; run make in lsmpqs-1.0_amd/")
    lib = C.CDLL(path)
    R = C.c_double if dtype == np.float64 else C.c_float
    lib.constCurvModel3d.restype = R
    lib.constCurvModel3dNoVar.restype = R
    lib.constCurvModel3dNoVar.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, R]
    lib.constCurvModel2d.restype = R
    lib.constCurvModel2dNoVar.restype = R
    lib.constCurvModel2dNoVar.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, R]
    lib.lsmpqs_hip_last_error.restype = C.c_char_p
    lib.lsmpqs_bench_inner_steps.restype = C.c_int
    lib.lsmpqs_bench_inner_steps.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                            R, C.c_int, C.POINTER(Timing)]
    lib.lsmpqs_set_quirks.restype = C.c_uint
    lib.lsmpqs_set_quirks.argtypes = [C.c_uint]
    lib.lsmpqs_download.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p]
    _LIBS[dtype] = lib
    return lib


def last_error(lib):
    return lib.lsmpqs_hip_last_error().decode()


def check(lib):
    msg = last_error(lib)
    if msg:
        raise LsmpqsError(msg)
