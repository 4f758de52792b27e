"""Edge geometries through drain_model3d, the HIP library (through its C ABI)
against the oracle on the same seeded inputs, FP64: a box without grains (one
pore space, nothing to trap), a tight pack (58 % solid, the wetting phase
trapped in many small pores), the smallest box the library accepts (4^3 fill
voxels, 3 ghost layers) and a 6^3 one, whose non-wetting phase vanishes (the
drivers' early return on a vanishing volume), and a thin non-cubic slab
(40 x 8 x 6) on the contact-angle path.  phi and both trapped-phase masks to 1e-10, the phases
voxel for voxel."""
import numpy as np
import pytest

from lsm_cases import run_model, maxdiff

pytestmark = pytest.mark.gpu

TOL64 = 1e-10

CASES = [
    ("no_grains", 24, dict(radius=4.0, porosity=1.0, seed=11), dict(space=2, time=2, theta_deg=0)),
    ("tight_pack", 24, dict(radius=4.0, porosity=0.18, seed=12), dict(space=3, time=3, theta_deg=0)),
    ("smallest_box", 4, dict(radius=1.5, porosity=0.5, seed=13),
     dict(space=5, time=3, theta_deg=30)),
    ("box_6", 6, dict(radius=2.0, porosity=0.5, seed=13), dict(space=5, time=3, theta_deg=30)),
    ("thin_slab", (40, 8, 6), dict(radius=3.0, porosity=0.4, seed=14),
     dict(space=3, time=3, theta_deg=45)),
]


@pytest.mark.parametrize("name,n,kw,opts", CASES, ids=[c[0] for c in CASES])
def test_edge_geometry_matches_oracle(name, n, kw, opts):
    print(f"case edge {name} {n} fp64", flush=True)
    Ao, _, _ = run_model("oracle", np.float64, n, opts, **kw)
    Ah, _, _ = run_model("hip", np.float64, n, opts, **kw)
    for f in ("phi", "mask_w", "mask_nw"):
        d = maxdiff(Ao[f], Ah[f])
        print(f"  {f}: max|diff| = {d:.3e}")
        assert d <= TOL64, f"{name} {f} differs by {d}"
    assert np.array_equal(Ao["phi"] < 0, Ah["phi"] < 0)
    assert np.array_equal(Ao["mask_w"] > 0, Ah["mask_w"] > 0)
    assert np.array_equal(Ao["mask_nw"] > 0, Ah["mask_nw"] > 0)
