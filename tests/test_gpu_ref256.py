"""256^3 drainage + imbibition on the GPU (FP64) against the reference build's
run of the same case, recorded in tests/golden/ref256_drain_imbibe.npz by
tests/golden/make_ref256.py (8 OpenMP threads, so 8 time-step slabs here):
seeded pack of radius U[6,12], porosity 0.36, seed 1002, WENO5 + TVD-RK3,
theta 30 deg, trapping on, eps_stop 1e-4; 2 drainage then 2 imbibition
curvature steps (reference: src/level_set_evolution/drain_model3d.c:49-165,
imbibe_model3d.c:49-161).  Checked after each curve: the smoothed saturation
to 1e-8, phi at 4096 seeded band voxels to 1e-10, the per-z-plane sums of
phi, the phase and trapped-mask voxel counts exactly, and the number of outer
steps."""
import ctypes as C
import importlib.util
import os

import numpy as np
import pytest

from lsmpqs_amd import hiplib
from lsm_cases import arrays, fill, saturation

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "ref256_drain_imbibe.npz")


def case_module():
    spec = importlib.util.spec_from_file_location("make_ref256",
                                                  os.path.join(HERE, "golden", "make_ref256.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.gpu
def test_drain_imbibe_256_matches_reference(tmp_path):
    if not os.path.exists(FIXTURE):
        pytest.skip("fixture not generated")
    ref = np.load(FIXTURE)
    mk = case_module()
    from lsmpqs_amd.inputs import pack_case
    dt = np.float64
    print("case: 256^3 WENO5 RK3 theta 30, drain 2 + imbibe 2, FP64, 8 slabs", flush=True)
    g, mask, phi0, _ = pack_case((mk.N,) * 3, (6.0, 12.0), 0.36, int(ref["seed"]), dt)
    A = arrays(g, mask, phi0, dt)
    o = mk.case_options(dt)
    lib = hiplib.load(dt)
    lib.lsmpqs_set_dt_slabs(int(ref["threads"]))
    os.environ["LSMPQS_NO_STEP_FILES"] = "1"
    libc = C.CDLL("libc.so.6")
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    f = fill(g)
    idx = ref["idx"]
    for tag, fn in (("drain", lib.drain_model3d), ("imbibe", lib.imbibe_model3d)):
        path = str(tmp_path / f"{tag}.log")
        fp = C.c_void_p(libc.fopen(path.encode(), b"w"))
        if tag == "imbibe":
            o.a = o.a + o.b * o.dc
        fn(C.byref(o), A.ptr(), C.byref(g), fp)
        libc.fclose(fp)
        hiplib.check(lib)
        phi = A["phi"]
        sat = saturation(phi, A["mask"], g)
        dphi = np.abs(phi.ravel()[idx] - ref[f"val_{tag}"]).max()
        dplane = np.abs(phi[f].sum(axis=(1, 2)) - ref[f"plane_sum_{tag}"]).max()
        outer = sum(1 for line in open(path) if "nw phase satn" in line)
        print(f"{tag}: |dS| {abs(sat - float(ref[f'sat_{tag}'])):.2e}, band phi max diff "
              f"{dphi:.2e}, plane-sum max diff {dplane:.2e}, outer {outer} "
              f"(reference {int(ref[f'outer_{tag}'])})", flush=True)
        assert outer == int(ref[f"outer_{tag}"])
        assert abs(sat - float(ref[f"sat_{tag}"])) <= 1e-8
        assert dphi <= 1e-10
        assert dplane <= 1e-10 * np.abs(ref[f"plane_sum_{tag}"]).max()
        assert int((phi[f] < 0).sum()) == int(ref[f"nneg_{tag}"])
        assert int((A["mask_nw"][f] > 0).sum()) == int(ref[f"nw_trap_{tag}"])
        assert int((A["mask_w"][f] > 0).sum()) == int(ref[f"w_trap_{tag}"])
    lib.lsmpqs_release(A.ptr())
