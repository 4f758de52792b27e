"""Checkpoint files of the outer loop (options->checkpoint = 1) against the
ones the reference build wrote for the same runs
(tests/golden/checkpoint_f64.npz, made by tests/golden/make_checkpoint.py;
reference: src/level_set_evolution/constCurvModel3d.c:301-314, :576-589):
checkpoint_phi, checkpoint_phi_prev, mask_w and mask_nw (gzip, the
reference's array-file layout) after the last outer step, and one
"connectivity time" log line per outer step.  FP64; the constant path
(theta 0) uses no transcendental function and must agree bit for bit, the
contact-angle path to 1e-10."""
import os
import sys

import numpy as np
import pytest

import oracle_py as op

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_checkpoint as mc  # noqa: E402

FIX = os.path.join(HERE, "golden", "checkpoint_f64.npz")


@pytest.mark.skipif(not op.have_ref(np.float64), reason="reference build not present")
def test_fixture_is_the_reference_output(tmp_path):
    """The committed vectors are what the reference build writes (CPU)."""
    d = np.load(FIX)
    op.set_omp_threads(int(d["threads"]))
    files, nconn, nouter, _ = mc.run(op.ref_lib(np.float64), "novar", np.float64, str(tmp_path))
    assert nconn == int(d["novar_conn_lines"]) and nouter == int(d["novar_outer"])
    for f in mc.FILES:
        np.testing.assert_array_equal(files[f], d[f"novar_{f}"], err_msg=f)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["novar", "var"])
def test_checkpoint_files_match_reference(name, tmp_path):
    from lsmpqs_amd import hiplib
    print(f"\n[checkpoint] {name} 16^3 fp64", flush=True)
    d = np.load(FIX)
    lib = hiplib.load(np.float64)
    lib.lsmpqs_set_dt_slabs(int(d["threads"]))
    os.environ["LSMPQS_NO_STEP_FILES"] = "1"
    files, nconn, nouter, A = mc.run(lib, name, np.float64, str(tmp_path))
    hiplib.check(lib)
    lib.lsmpqs_release(A.ptr())
    assert nouter == int(d[f"{name}_outer"])
    assert nconn == nouter == int(d[f"{name}_conn_lines"])
    tol = 0.0 if name == "novar" else 1e-10
    for f in mc.FILES:
        diff = float(np.max(np.abs(files[f] - d[f"{name}_{f}"])))
        print(f"  {f}: max|diff| {diff:.2e}", flush=True)
        assert diff <= tol, f


@pytest.mark.gpu
def test_checkpoint_files_fp32_layout(tmp_path):
    """The FP32 library writes the same files in its precision (int32 dims,
    then float32 values); checkpoint_phi is the state after the last outer
    step, close to the FP64 reference's."""
    from lsmpqs_amd import hiplib
    print("\n[checkpoint] var 16^3 fp32", flush=True)
    d = np.load(FIX)
    lib = hiplib.load(np.float32)
    lib.lsmpqs_set_dt_slabs(int(d["threads"]))
    os.environ["LSMPQS_NO_STEP_FILES"] = "1"
    files, nconn, nouter, A = mc.run(lib, "var", np.float32, str(tmp_path))
    hiplib.check(lib)
    lib.lsmpqs_release(A.ptr())
    assert nconn == nouter == int(d["var_outer"])
    for f in mc.FILES:
        assert files[f].dtype == np.float32
    band = np.abs(d["var_checkpoint_phi"]) < 3
    assert np.max(np.abs(files["checkpoint_phi"][band] - d["var_checkpoint_phi"][band])) < 0.05
