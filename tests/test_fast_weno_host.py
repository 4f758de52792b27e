"""The single-precision fast WENO5 of the HIP library, compiled for the host:
finite everywhere (no denormal weight sums reaching the hardware reciprocal)
and within float accuracy of the double-precision scheme."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_fast_weno_face_is_finite_and_accurate():
    exe = os.path.join(tempfile.mkdtemp(prefix="lsm_fw_"), "fast_weno_check")
    src = os.path.join(ROOT, "tests", "native", "fast_weno_check.cpp")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                    "-I", os.path.join(ROOT, "lsmpqs-1.0_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                    "-DQSSLIB_REAL=float", src, "-o", exe], check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    nonfinite, worst = int(out[0]), float(out[1])
    assert nonfinite == 0
    assert worst < 2e-6


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_weno_face_adjoint_matches_differences():
    """weno5_face_adj, the reverse-mode derivative the FP32 adjoint uses for the
    fast WENO5 faces, against central differences of weno5_face (double), and
    its float instantiation against its double one; plus the flat-window
    branch, whose derivative is the middle candidate's stencil; and
    weno5_line_adj (the FP64 adjoint's reference-order WENO5) against central
    differences of weno5."""
    exe = os.path.join(tempfile.mkdtemp(prefix="lsm_wa_"), "weno_adj_check")
    src = os.path.join(ROOT, "tests", "native", "weno_adj_check.cpp")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                    "-I", os.path.join(ROOT, "lsmpqs-1.0_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                    "-DQSSLIB_REAL=float", src, "-o", exe], check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    fd, f32, flat, line = (float(x) for x in out)
    assert fd < 1e-6
    assert f32 < 1e-4
    assert flat < 1e-15
    assert line < 1e-6   # weno5_line_adj, the FP64 adjoint's reference-order WENO5


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_fast_eno_matches_osher_shu_construction():
    """eno2_fast_raw / eno3_fast_raw, the FP32 library's ENO2 and ENO3
    one-sided derivatives, against the Osher-Shu ENO construction written
    from its definition (tests/native/fast_eno_check.cpp): exact to rounding
    in double, to float accuracy in float away from near-ties of a stencil
    choice."""
    exe = os.path.join(tempfile.mkdtemp(prefix="lsm_fe_"), "fast_eno_check")
    src = os.path.join(ROOT, "tests", "native", "fast_eno_check.cpp")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                    "-I", os.path.join(ROOT, "lsmpqs-1.0_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                    "-DQSSLIB_REAL=float", src, "-o", exe], check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    w64, w32, n32 = float(out[0]), float(out[1]), int(out[2])
    assert w64 < 1e-12
    assert w32 < 1e-5
    assert n32 > 1000000
