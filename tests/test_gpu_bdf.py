"""The pyTACS-style front end on the GPU with the reference's own fixture
meshes: two_hexs.bdf (PLOAD4 face pressures on CHEXA, six subcases;
tests/integration_tests/test_elast_linhexa_pressure_3d.py) and plate.bdf
(CQUAD4 shells in a CORD2R frame; tests/integration_tests/
test_shell_plate_quad.py), against the function values those tests hold."""
import os

import numpy as np
import pytest

from tacs_amd import constitutive, elements, functions
from tacs_amd.pytacs import pyTACS

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "bdf")

# test_elast_linhexa_pressure_3d.py:24-75
TWO_HEX_COMPLIANCE = {
    "Clamp_+X": 25788.7451098405, "Clamp_+Y": 11409.61367995888, "Clamp_-X": 0.0,
    "Clamp_-Y": 11409.61367995888, "Clamp_+Z": 36425.924866297384,
    "Clamp_-Z": 36425.924866297384,
}
# test_shell_plate_quad.py:27-74 (ksWeight 10)
PLATE = {"point_load_compliance": 683.8571611645635, "point_load_ks_vmfailure": 0.5757488025912149,
         "pressure_compliance": 4679.345460332448, "pressure_ks_vmfailure": 1.2938623156883402}


def test_two_hexs_bdf_subcases():
    fea = pyTACS(os.path.join(GOLD, "two_hexs.bdf"))
    fea.initialize()
    probs = fea.createTACSProbsFromBDF()
    assert sorted(probs) == sorted(TWO_HEX_COMPLIANCE)
    funcs = {}
    for name, p in probs.items():
        # the reference test's settings (test_elast_linhexa_pressure_3d.py:103-105)
        p.setOption("L2Convergence", 1e-20)
        p.setOption("L2ConvergenceRel", 1e-20)
        p.addFunction("compliance", functions.Compliance)
        p.solve()
        p.evalFunctions(funcs)
    for name, ref in TWO_HEX_COMPLIANCE.items():
        c = funcs[f"{name}_compliance"]
        assert abs(c - ref) <= 1e-9 * max(ref, 1.0), (name, c, ref)


def test_plate_bdf_point_load_and_pressure():
    fea = pyTACS(os.path.join(GOLD, "plate.bdf"))

    # the reference test's element callback (test_shell_plate_quad.py:98-118)
    def elem_call_back(dv_num, comp_id, comp_descript, elem_descripts, global_dvs, **kwargs):
        prop = constitutive.MaterialProperties(rho=2500.0, E=70e9, nu=0.3, ys=464.0e6)
        con = constitutive.IsoShellConstitutive(prop, t=0.005, tNum=dv_num)
        return elements.Quad4Shell(None, con), [100.0]

    fea.initialize(elem_call_back)
    assert fea.dvNum == 2 and fea.selectCompIDs(include="PLATE") == [0]
    sp = fea.createStaticProblem(name="point_load")
    sp.addLoadToNodes(81, np.array([0.0, 0.0, 1e4, 0.0, 0.0, 0.0]), nastranOrdering=True)
    pp = fea.createStaticProblem(name="pressure")
    pp.addPressureToComponents(fea.selectCompIDs(include="PLATE"), 100e3)
    funcs = {}
    for p in (sp, pp):
        p.setOption("L2ConvergenceRel", 1e-14)
        p.setOption("subSpaceSize", 200)
        p.setOption("nRestarts", 200)
        p.addFunction("ks_vmfailure", functions.KSFailure, ksWeight=10.0)
        p.addFunction("compliance", functions.Compliance)
        assert p.solve() == 1
        p.evalFunctions(funcs)
    for k, ref in PLATE.items():
        assert abs(funcs[k] - ref) <= 1e-9 * ref, (k, funcs[k], ref)
