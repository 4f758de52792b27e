"""lsmpqs_volume_noalloc (the allocation-free volume over device arrays)
against qss3dvolumeregionphilessthanzero_ (host arrays, the library's
staging path) on the same fields: two calls in a row, the second on a smaller
box, in both precisions.  The two entry points share the reduction kernels,
so the results must be equal bit for bit."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as op
from lsm_cases import case


def _volume_host(lib, dt, phi, g, eps):
    out = op.Scalar(0.0, C.c_double if dt == np.float64 else C.c_float)
    op.fcall(lib, "qss3dvolumeregionphilessthanzero_", dt, out, phi, *op.gb6(g), *op.fb6(g),
             1.0, 1.0, 1.0, eps)
    return out.value


def _volume_dev(lib, torch, dt, phi, g, eps, stream):
    R = C.c_double if dt == np.float64 else C.c_float
    tdt = torch.float64 if dt == np.float64 else torch.float32
    d_phi = torch.from_numpy(np.ascontiguousarray(phi)).to("cuda")
    d_vol = torch.full((1,), -7.0, dtype=tdt, device="cuda")
    gb = (C.c_int * 6)(*op.gb6(g))
    fb = (C.c_int * 6)(*op.fb6(g))
    lib.lsmpqs_volume_noalloc.restype = C.c_int
    lib.lsmpqs_volume_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, R, R, R,
                                          R, C.c_void_p]
    rc = lib.lsmpqs_volume_noalloc(d_vol.data_ptr(), d_phi.data_ptr(), gb, fb, 1.0, 1.0, 1.0, eps,
                                   stream.cuda_stream)
    assert rc == 0, lib.lsmpqs_hip_last_error()
    stream.synchronize()
    return float(d_vol.cpu().numpy()[0])


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_volume_noalloc_matches_staging_entry_point(dt):
    import torch
    from lsmpqs_amd import hiplib
    lib = hiplib.load(dt)
    stream = torch.cuda.Stream()
    for n in (40, 20):
        print(f"[no-alloc] volume, {np.dtype(dt).name}, {n}^3", flush=True)
        g, mask, phi0 = case(n, dt)
        for field in (phi0, mask):
            ref = _volume_host(lib, dt, field, g, 1.5)
            hiplib.check(lib)
            got = _volume_dev(lib, torch, dt, field, g, 1.5, stream)
            assert got == ref, (n, got, ref)
            assert got > 0


@pytest.mark.gpu
def test_volume_noalloc_rejects_bad_input():
    """Input errors come back as -2 with a message, before any device work."""
    import torch
    from lsmpqs_amd import hiplib
    dt = np.float64
    lib = hiplib.load(dt)
    d = torch.zeros(8, dtype=torch.float64, device="cuda")
    gb = (C.c_int * 6)(0, -1, 0, 1, 0, 1)     # empty x range
    fb = (C.c_int * 6)(0, 0, 0, 0, 0, 0)
    R = C.c_double
    lib.lsmpqs_volume_noalloc.restype = C.c_int
    lib.lsmpqs_volume_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, R, R, R,
                                          R, C.c_void_p]
    assert lib.lsmpqs_volume_noalloc(d.data_ptr(), d.data_ptr(), gb, fb, 1.0, 1.0, 1.0, 1.5,
                                     None) == -2
    assert "empty ghost box" in hiplib.last_error(lib)
    assert lib.lsmpqs_volume_noalloc(None, d.data_ptr(), gb, fb, 1.0, 1.0, 1.0, 1.5, None) == -2
