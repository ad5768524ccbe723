"""C3 in FP32 against C3 in FP64 from the same device-drawn plasma (256^3
cells, 2 x 1.07e9 particles, 32^3 floating cube, B0 = 0.2 z): the FP32 push,
deposit and V-cycle smoothing keep rho, phi, the particle phase-space moments
and the energies within the bars below of the all-FP64 run, right after
initialisation and after three steps (RootGrid.cpp:82-143; the conductor
algebra, the multigrid defect correction and the reductions are FP64 in both).

Bars (relative unless stated; each is a few times the error measured on an
MI355X, so a change that loses FP32 accuracy fails it):
  after initialisation: rho L2 5e-5, phi L2 5e-6;
  after three steps:    rho L2 5e-4, phi L2 1e-4, mean velocity 1e-5 sigma,
                        velocity variance 1e-5, particle count 1e-6,
                        particle energy 1e-5, field energy 1e-4.
"""
import os

import pytest

pytestmark = pytest.mark.gpu

pytest.importorskip("fopis_amd.lib")
from fopis_amd import precision_check  # noqa: E402
from tests.cases import CONFIGS  # noqa: E402


def test_c3_fp32_tracks_fp64():
    res = precision_check.main(["--config", os.path.join(CONFIGS, "c3.json"), "--steps", "3"])
    print(res, flush=True)
    a0, a3 = res["after_initialize"], res["after_3_steps"]
    assert a0["rho"]["rel_l2"] <= 5e-5, a0["rho"]
    assert a0["phi"]["rel_l2"] <= 5e-6, a0["phi"]
    assert a3["rho"]["rel_l2"] <= 5e-4, a3["rho"]
    assert a3["phi"]["rel_l2"] <= 1e-4, a3["phi"]
    for sp in a3["species"]:
        assert abs(sp["count_diff"]) <= 1e-6 * sp["count_64"], sp
        assert max(abs(x) for x in sp["mean_v_diff_over_sigma"]) <= 1e-5, sp
        assert max(abs(x) for x in sp["var_v_rel_diff"]) <= 1e-5, sp
    assert abs(a3["particle_energy_rel_diff"]) <= 1e-5, a3
    assert abs(a3["field_energy_rel_diff"]) <= 1e-4, a3
