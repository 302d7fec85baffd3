"""Nastran BDF reader and the pyTACS-style model conversion (tacs/pytacs.py,
tacs/pymeshloader.py, tacs/problems/base.py) on the CPU: field formats,
continuations and coordinate frames; the reference's two-hex fixture
(tests/integration_tests/input_files/two_hexs.bdf) converted and solved by the
oracle against the compliance values its test holds
(tests/integration_tests/test_elast_linhexa_pressure_3d.py:24-75)."""
import os

import numpy as np
import pytest

import oracle as O
from tacs_amd.bdf import BDF, nastran_float
from tacs_amd.pytacs import bc_groups, load_arrays, mesh_arrays

GOLD = os.path.join(os.path.dirname(__file__), "golden", "bdf")

# test_elast_linhexa_pressure_3d.py:24-75
TWO_HEX_COMPLIANCE = {
    "Clamp_+X": 25788.7451098405, "Clamp_+Y": 11409.61367995888, "Clamp_-X": 0.0,
    "Clamp_-Y": 11409.61367995888, "Clamp_+Z": 36425.924866297384,
    "Clamp_-Z": 36425.924866297384,
}


def test_nastran_reals():
    cases = {"7.+10": 7e10, "1.-5": 1e-5, "-100000.": -1e5, ".3": 0.3, "2.7+3": 2700.0,
             "1.0D+2": 100.0, "1e3": 1000.0, "-.5": -0.5, "  ": None, "3": 3.0}
    for s, v in cases.items():
        assert nastran_float(s) == v
    with pytest.raises(ValueError):
        nastran_float("abc")


def test_field_formats_and_continuations(tmp_path):
    small = """BEGIN BULK
GRID           1       0      0.      0.      0.
GRID           2       0      1.      0.      0.
GRID*                  3               0              1.              1.
*                     0.
GRID,4,,0.,1.,0.
CQUAD4         1       1       1       2       3       4
MAT1,7,7.+10,,.3,2700.,,,,+
+,270.+6
PSHELL         1       7    .005
SPC1           1     123       1    THRU       2
FORCE          3       4       0      2.      0.      0.      1.
PLOAD2         4    -50.       1
LOAD           5      2.     1.5       3     -1.       4
ENDDATA
"""
    p = tmp_path / "m.bdf"
    p.write_text(small)
    b = BDF(str(p))
    assert [list(b.nodes[n]) for n in (1, 2, 3, 4)] == [[0, 0, 0], [1, 0, 0], [1, 1, 0],
                                                          [0, 1, 0]]
    assert b.elements[1] == ("CQUAD4", 1, [1, 2, 3, 4])
    assert b.materials[7] == {"E": 7e10, "nu": 0.3, "rho": 2700.0, "alpha": 0.0, "St": 270e6}
    assert b.properties[1]["t"] == 0.005 and b.spcs[1] == [([1, 2], "123", 0.0)]
    etype, vpn, conn, X, comps = mesh_arrays(b)
    assert etype == "CQUAD4" and vpn == 6 and conn.tolist() == [[0, 1, 3, 2]]
    F, pres = load_arrays(b, 5, vpn)
    f = np.zeros(24)
    f[6 * 3 + 2] = 2.0 * 1.5 * 2.0  # LOAD 5 = 2 * (1.5 FORCE 3 - 1 PLOAD2 4)
    assert np.array_equal(F, f) and pres == [(0, 0, 2.0 * -1.0 * -50.0)]
    (nodes, dofs, vals), = bc_groups(b, vpn)
    assert nodes.tolist() == [0, 1] and dofs == [0, 1, 2] and vals is None


def test_coordinate_frames(tmp_path):
    txt = """BEGIN BULK
CORD2R         3       0      .5      .5      0.      .5      .5      1.+
+            1.5      .5      0.
CORD2C         4       0      1.      0.      0.      1.      0.      1.+
+             2.      0.      0.
GRID           1       3     -.5     -.5      0.
GRID           2       4      2.     90.      3.
ENDDATA
"""
    p = tmp_path / "c.bdf"
    p.write_text(txt)
    b = BDF(str(p))
    assert np.allclose(b.nodes[1], [0.0, 0.0, 0.0])
    # r = 2 at 90 degrees about the frame's z axis through (1, 0, 0)
    assert np.allclose(b.nodes[2], [1.0, 2.0, 3.0])


def test_plate_fixture_mesh():
    b = BDF(os.path.join(GOLD, "plate.bdf"))
    etype, vpn, conn, X, comps = mesh_arrays(b)
    assert etype == "CQUAD4" and conn.shape == (100, 4) and X.shape == (121, 3)
    assert np.allclose(X.min(0), [0, 0, 0]) and np.allclose(X.max(0), [1, 1, 0])
    assert [b.comments[p] for p in b.properties] == ["PLATE", "DUMMY"]
    # every element is a positively oriented unit-square cell of 0.1
    d1 = X[conn[:, 1]] - X[conn[:, 0]]
    d2 = X[conn[:, 2]] - X[conn[:, 0]]
    assert np.allclose(np.cross(d1, d2)[:, 2], 0.01)
    (nodes, dofs, vals), = bc_groups(b, vpn)
    assert len(nodes) == 40 and dofs == list(range(6))


@pytest.mark.parametrize("name", list(TWO_HEX_COMPLIANCE))
def test_two_hex_pressures_compliance(name):
    """Model, BCs and PLOAD4 face pressures of two_hexs.bdf, solved densely
    with the oracle's element restatements; the compliance matches the value
    the reference's test holds."""
    b = BDF(os.path.join(GOLD, "two_hexs.bdf"))
    sc = [s for s in b.subcases if s["name"] == name][0]
    etype, vpn, conn, X, comps = mesh_arrays(b)
    assert etype == "CHEXA" and conn.shape == (2, 8)
    mat = b.materials[b.properties[1]["mid"]]
    C = O.iso_C21(mat["E"], mat["nu"])
    nn = X.shape[0]
    rowp, cols = O.node_csr(nn, conn, 8)
    A, res = O.assemble_solid(2, conn, X.reshape(-1), np.zeros(3 * nn), np.ones(2), C, rowp,
                              cols)
    F, pres = load_arrays(b, sc["load"], vpn)
    assert len(pres) in (1, 2) and all(p == 1e5 for _, _, p in pres)
    e, f, p = zip(*pres)
    O.add_pressures(2, conn, X.reshape(-1), np.array(e), np.array(f), np.array(p), res)
    res -= F
    K = np.zeros((3 * nn, 3 * nn))
    for i in range(nn):
        for j in range(rowp[i], rowp[i + 1]):
            K[3 * i:3 * i + 3, 3 * cols[j]:3 * cols[j] + 3] = A[9 * j:9 * j + 9].reshape(3, 3)
    free = np.ones(3 * nn, bool)
    for nodes, dofs, vals in bc_groups(b, vpn):
        for n in nodes:
            free[3 * n + np.array(dofs)] = False
    u = np.zeros(3 * nn)
    u[free] = -np.linalg.solve(K[np.ix_(free, free)], res[free])
    c = O.compliance(2, conn, X.reshape(-1), u, np.ones(2), C)
    ref = TWO_HEX_COMPLIANCE[name]
    assert abs(c - ref) <= 1e-10 * max(abs(ref), 1.0)
