"""The oracle against the reference: golden vectors written by the reference
build (tests/golden/, always available) and, where oracle/_ref is built, the
reference itself on fresh inputs.  FP64 must be bit-exact."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import oracle_py as op  # noqa: E402
from lsm_cases import run_model, maxdiff  # noqa: E402
import make_golden as mg  # noqa: E402

DT = np.float64


def test_kernels_match_golden_vectors():
    gold = np.load(os.path.join(HERE, "golden", "kernels_f64.npz"))
    g, inp = mg.kernel_inputs()
    out = mg.kernel_outputs(op.oracle_lib(DT), "orc_", "", g, inp)
    assert set(out) == set(gold.files)
    for k in gold.files:
        assert np.array_equal(out[k], gold[k]), k


@pytest.mark.parametrize("name", sorted(mg.DRAIN_CASES))
def test_drain_matches_golden(name):
    gold = np.load(os.path.join(HERE, "golden", f"drain_{name}.npz"))
    A, _, _ = run_model("oracle", DT, mg.DRAIN_N, mg.DRAIN_CASES[name], T=mg.DRAIN_T)
    for f in ("phi", "mask_w", "mask_nw", "mask"):
        assert np.array_equal(A[f], gold[f]), f


needs_ref = pytest.mark.skipif(not op.have_ref(DT), reason="oracle/_ref not built here")


@needs_ref
@pytest.mark.parametrize("opts", [
    dict(space=2, time=1, theta_deg=0),
    dict(space=5, time=3, theta_deg=0),
    dict(space=3, time=3, theta_deg=45),
    dict(space=2, time=2, theta_deg=60),
], ids=["eno2rk1", "eno3c", "eno3t45", "eno2t60"])
def test_drain_matches_reference_build(opts):
    A, _, _ = run_model("oracle", DT, 24, opts, T=3)
    B, _, _ = run_model("ref", DT, 24, opts, T=3)
    for f in ("phi", "mask_w", "mask_nw"):
        assert maxdiff(A[f], B[f]) == 0.0, f


@needs_ref
def test_imbibition_matches_reference_build():
    opts = dict(space=3, time=3, theta_deg=30)
    A, _, _ = run_model("oracle", DT, 24, opts, T=2, imbibe_after=True)
    B, _, _ = run_model("ref", DT, 24, opts, T=2, imbibe_after=True)
    for f in ("phi", "mask_w", "mask_nw", "normal_velocity", "curvature_coeff"):
        assert np.array_equal(A[f], B[f]), f


@needs_ref
def test_kernels_match_reference_on_fresh_inputs():
    g, inp = mg.kernel_inputs(seed=99)
    a = mg.kernel_outputs(op.oracle_lib(DT), "orc_", "", g, inp)
    b = mg.kernel_outputs(op.ref_lib(DT), "", "_", g, inp)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@needs_ref
def test_fp32_oracle_close_to_fp32_reference():
    # the FP32 reference mixes real*8 literals into real*4 expressions, so
    # the FP32 oracle is checked to a tolerance, not bit for bit
    opts = dict(space=2, time=2, theta_deg=0)
    A, _, _ = run_model("oracle", np.float32, 20, opts, T=2)
    B, _, _ = run_model("ref", np.float32, 20, opts, T=2)
    band = np.abs(B["phi"]) < 3
    assert maxdiff(A["phi"][band], B["phi"][band]) < 1e-3
