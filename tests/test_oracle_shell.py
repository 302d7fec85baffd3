"""Pins of the oracle's Quad4 shell restatement (oracle/shell_oracle.c).

- element residual, Jacobian, stiffness, consistent mass and geometric
  stiffness matrices, the nonlinear model's residual and Jacobian,
  thickness adjoint product, failure index, detXd and pressure load against
  values the reference's own element code wrote (tests/golden/elements.npz,
  made by tests/golden/make_fixtures.py);
- the mass of a flat rectangle in closed form: total mass rho t a b for a
  rigid translation and the rotary inertia rho t^3 a b / 12 for a uniform
  in-plane rotation, zero for the drilling rotation;
- the 10x10 clamped plate of tests/integration_tests/test_shell_plate_quad.py
  (pressure 100 kPa and a 10 kN centre load): compliance and KS failure
  (ksWeight 10) against the golden values held in that test (:32-46);
- rigid-body null space and symmetry of the element stiffness;
- central finite differences of the KS failure sensitivities and of the
  thickness adjoint product.
"""
import os

import numpy as np
import pytest

import oracle as O
from tacs_amd.meshes import PlateMesh

HERE = os.path.dirname(os.path.abspath(__file__))
E, NU, YS = 70e9, 0.3, 270e6


@pytest.fixture(scope="module")
def fixtures():
    return np.load(os.path.join(HERE, "golden", "elements.npz"))


@pytest.mark.parametrize("case", range(4))
def test_quad4_against_reference_element(fixtures, case):
    k = f"quad4_{case}"
    X, v, psi = fixtures[k + "_X"], fixtures[k + "_v"], fixtures[k + "_psi"]
    t = float(fixtures[k + "_t"].reshape(-1)[0])
    res, mat = O.quad4_element(X, v, E, NU, t)
    R, M = fixtures[k + "_res"], fixtures[k + "_mat"].reshape(24, 24)
    assert np.abs(res - R).max() <= 1e-14 * np.abs(R).max()
    assert np.abs(mat - M).max() <= 1e-14 * np.abs(M).max()
    adj = O.quad4_adj_product(X, v, psi, E, NU, t)
    ref_adj = float(fixtures[k + "_adj"].reshape(-1)[0])
    assert abs(adj - ref_adj) <= 1e-13 * abs(ref_adj)
    for n in range(4):
        f, det = O.quad4_failure(X, v, E, NU, YS, t, n)
        assert abs(f - fixtures[k + "_fail"][n]) <= 1e-13 * abs(fixtures[k + "_fail"][n])
        assert abs(det - fixtures[k + "_det"][n]) <= 1e-14 * abs(fixtures[k + "_det"][n])
    pres = O.quad4_pressure(X, 1e4)
    assert np.abs(pres - fixtures[k + "_pres"]).max() <= 1e-14 * np.abs(fixtures[k + "_pres"]).max()


@pytest.mark.parametrize("case", range(4))
def test_quad4_mat_types_against_reference_element(fixtures, case):
    """getMatType(TACS_STIFFNESS_MATRIX) and getMatType(TACS_MASS_MATRIX) of
    the reference element (density 2700) against the oracle."""
    k = f"quad4_{case}"
    X = fixtures[k + "_X"]
    t = float(fixtures[k + "_t"].reshape(-1)[0])
    ref = fixtures[k + "_mattype"]
    _, K = O.quad4_element(X, np.zeros(24), E, NU, t)
    assert np.abs(K.reshape(-1) - ref[0]).max() <= 1e-14 * np.abs(ref[0]).max()
    M = O.quad4_mass(X, 2700.0, t)
    assert np.abs(M.reshape(-1) - ref[1]).max() <= 1e-14 * np.abs(ref[1]).max()
    # alpha scales the whole matrix
    assert np.array_equal(O.quad4_mass(X, 2700.0, t, 2.0), 2.0 * M)
    # geometric stiffness along the state v (TACS_GEOMETRIC_STIFFNESS_MATRIX)
    G = O.quad4_geo(X, fixtures[k + "_v"], E, NU, t)
    assert np.abs(G.reshape(-1) - ref[2]).max() <= 1e-12 * np.abs(ref[2]).max()
    # the nonlinear model (TACSQuad4NonlinearShell) at the state 30 v
    r, K = O.quad4_nl_element(X, 30.0 * fixtures[k + "_v"], E, NU, t)
    Rr, Kr = fixtures[k + "_nl_res"], fixtures[k + "_nl_mat"]
    assert np.abs(r - Rr).max() <= 1e-13 * np.abs(Rr).max()
    assert np.abs(K.reshape(-1) - Kr).max() <= 1e-13 * np.abs(Kr).max()


def test_quad4_geometric_stiffness_properties():
    """G is symmetric, linear in the path, zero for a zero path, and the
    tangent of the nonlinear model's stiffness: (K_T(h p) - K_T(-h p)) / 2h."""
    rng = np.random.default_rng(21)
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float) * [0.3, 0.2, 1.0]
    X = (X + 0.02 * rng.standard_normal((4, 3))).reshape(-1)
    p = 1e-3 * rng.standard_normal(24)
    t = 0.01
    G = O.quad4_geo(X, p, E, NU, t)
    assert np.abs(G - G.T).max() <= 1e-12 * np.abs(G).max()
    assert np.abs(O.quad4_geo(X, 2.5 * p, E, NU, t) - 2.5 * G).max() <= 1e-12 * np.abs(G).max()
    assert np.abs(O.quad4_geo(X, np.zeros(24), E, NU, t)).max() == 0.0
    h = 1e-2
    fd = (O.quad4_nl_element(X, h * p, E, NU, t)[1] - O.quad4_nl_element(X, -h * p, E, NU, t)[1])
    assert np.abs(fd / (2 * h) - G).max() <= 1e-9 * np.abs(G).max()


def test_quad4_mass_closed_form():
    """Flat a x b rectangle in the x-y plane: v^T M v equals rho t a b for a
    rigid translation, rho t^3 a b / 12 for a unit rotation about x or y and
    0 about z (the director q x n vanishes); M is symmetric and positive
    semi-definite."""
    a, b, rho, t = 0.7, 0.4, 2700.0, 0.02
    X = np.array([[0, 0, 0], [a, 0, 0], [0, b, 0], [a, b, 0]], float).reshape(-1)
    M = O.quad4_mass(X, rho, t)
    assert np.abs(M - M.T).max() <= 1e-15 * np.abs(M).max()
    for k in range(3):
        v = np.zeros(24)
        v[k::6] = 1.0
        assert abs(v @ M @ v - rho * t * a * b) <= 1e-13 * rho * t * a * b
    for k, expect in ((3, rho * t ** 3 * a * b / 12), (4, rho * t ** 3 * a * b / 12), (5, 0.0)):
        v = np.zeros(24)
        v[k::6] = 1.0
        assert abs(v @ M @ v - expect) <= 1e-13 * rho * t ** 3 * a * b / 12
    assert np.linalg.eigvalsh(M).min() >= -1e-12 * np.abs(M).max()


def _warped_quad(rng):
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float) * 0.3
    X += rng.uniform(-0.03, 0.03, X.shape)
    return X.reshape(-1)


def test_quad4_rigid_body_and_symmetry():
    rng = np.random.default_rng(3)
    for _ in range(3):
        X = _warped_quad(rng)
        _, K = O.quad4_element(X, np.zeros(24), E, NU, 0.01)
        assert np.abs(K - K.T).max() <= 1e-14 * np.abs(K).max()
        Xn = X.reshape(4, 3)
        for ax in range(3):
            tr = np.zeros((4, 6))
            tr[:, ax] = 1.0
            th = np.zeros(3)
            th[ax] = 1.0
            rot = np.zeros((4, 6))
            rot[:, :3] = np.cross(th, Xn - Xn.mean(0))
            rot[:, 3:] = th
            for mode in (tr, rot):
                r = K @ mode.reshape(-1)
                assert np.abs(r).max() <= 1e-10 * np.abs(K).max() * np.abs(mode).max()


def _plate_dense(m, t):
    rowp, cols = O.node_csr(m.nnodes, m.conn, 4)
    A, _ = O.assemble_shell(m.conn, m.X, None, np.full(m.nelem, t), E, NU, rowp, cols)
    n = m.nnodes
    K = np.zeros((6 * n, 6 * n))
    for i in range(n):
        for j in range(rowp[i], rowp[i + 1]):
            K[6 * i:6 * i + 6, 6 * cols[j]:6 * cols[j] + 6] = A[36 * j:36 * j + 36].reshape(6, 6)
    return K


@pytest.mark.parametrize("case,comp,ks", [
    ("pressure", 4679.345460332448, 1.2938623156883402),
    ("point_load", 683.8571611645635, 0.5757488025912149),
])
def test_plate_golden(case, comp, ks):
    """tests/integration_tests/test_shell_plate_quad.py: 1 m x 1 m plate of
    10 x 10 Quad4 shells, t = 5 mm, E = 70 GPa, nu = 0.3, ys = 464 MPa,
    perimeter clamped in all six dofs."""
    m = PlateMesh(10, 10)
    t, ys = 0.005, 464e6
    K = _plate_dense(m, t)
    n = m.nnodes
    f = np.zeros(6 * n)
    if case == "pressure":
        O.shell_add_pressure(m.conn, m.X, np.arange(m.nelem), np.full(m.nelem, 100e3), f)
        f = -f
    else:
        f[6 * m.center_node() + 2] = 1e4
    free = np.ones(6 * n, bool)
    for b in m.edge_nodes():
        free[6 * b:6 * b + 6] = False
    u = np.zeros(6 * n)
    u[free] = np.linalg.solve(K[np.ix_(free, free)], f[free])
    te = np.full(m.nelem, t)
    c = O.shell_compliance(m.conn, m.X, u, te, E, NU)
    k = O.shell_ks_failure(m.conn, m.X, u, te, E, NU, ys, 10.0)[0]
    # the golden values come from an iterative solve; they agree to ~3e-12
    assert abs(c - comp) <= 1e-10 * comp
    assert abs(k - ks) <= 1e-10 * ks


def _small_shell(rng, n=3):
    m = PlateMesh(n, n, 0.4, 0.4)
    X = m.X.copy()
    X[:, 2] += rng.uniform(-0.01, 0.01, m.nnodes)
    t = rng.uniform(0.005, 0.02, m.nelem)
    u = rng.normal(0.0, 1e-4, 6 * m.nnodes)
    return m, X.reshape(-1), t, u


def test_shell_ks_sensitivities_fd():
    rng = np.random.default_rng(5)
    m, X, t, u = _small_shell(rng)
    w = 30.0
    val, mx, s = O.shell_ks_failure(m.conn, X, u, t, E, NU, YS, w)
    dfdu, dfdt = O.shell_ks_sens(m.conn, X, u, t, E, NU, YS, w, mx, s)
    h = 1e-9
    for i in rng.choice(u.size, 12, replace=False):
        up, um = u.copy(), u.copy()
        up[i] += h
        um[i] -= h
        fd = (O.shell_ks_failure(m.conn, X, up, t, E, NU, YS, w)[0]
              - O.shell_ks_failure(m.conn, X, um, t, E, NU, YS, w)[0]) / (2 * h)
        assert abs(fd - dfdu[i]) <= 1e-5 * np.abs(dfdu).max()
    h = 1e-7
    for e in range(m.nelem):
        tp, tm = t.copy(), t.copy()
        tp[e] += h
        tm[e] -= h
        fd = (O.shell_ks_failure(m.conn, X, u, tp, E, NU, YS, w)[0]
              - O.shell_ks_failure(m.conn, X, u, tm, E, NU, YS, w)[0]) / (2 * h)
        assert abs(fd - dfdt[e]) <= 1e-5 * np.abs(dfdt).max()


def test_shell_adjoint_product_fd():
    rng = np.random.default_rng(7)
    m, X, t, u = _small_shell(rng)
    psi = rng.normal(size=u.size)
    rowp, cols = O.node_csr(m.nnodes, m.conn, 4)
    prod = O.shell_adj_res_product(m.conn, X, u, psi, t, E, NU)
    h = 1e-7
    for e in range(m.nelem):
        tp, tm = t.copy(), t.copy()
        tp[e] += h
        tm[e] -= h
        rp = O.assemble_shell(m.conn, X, u, tp, E, NU, rowp, cols)[1]
        rm = O.assemble_shell(m.conn, X, u, tm, E, NU, rowp, cols)[1]
        fd = psi @ (rp - rm) / (2 * h)
        assert abs(fd - prod[e]) <= 1e-6 * np.abs(prod).max()


def test_shell_compliance_sensitivities_fd():
    rng = np.random.default_rng(9)
    m, X, t, u = _small_shell(rng)
    dfdu, dfdt = O.shell_compliance_sens(m.conn, X, u, t, E, NU)
    c0 = O.shell_compliance(m.conn, X, u, t, E, NU)
    # compliance is quadratic in u: the central difference is exact up to rounding
    h = 1e-7
    for i in rng.choice(u.size, 12, replace=False):
        up, um = u.copy(), u.copy()
        up[i] += h
        um[i] -= h
        fd = (O.shell_compliance(m.conn, X, up, t, E, NU)
              - O.shell_compliance(m.conn, X, um, t, E, NU)) / (2 * h)
        assert abs(fd - dfdu[i]) <= 1e-6 * np.abs(dfdu).max()
    h = 1e-7
    for e in range(m.nelem):
        tp, tm = t.copy(), t.copy()
        tp[e] += h
        tm[e] -= h
        fd = (O.shell_compliance(m.conn, X, u, tp, E, NU)
              - O.shell_compliance(m.conn, X, u, tm, E, NU)) / (2 * h)
        assert abs(fd - dfdt[e]) <= 1e-6 * abs(c0)


KS_WEIGHT = 30.0


@pytest.mark.parametrize("case", range(4))
def test_quad4_xpt_against_complex_step_reference(fixtures, case):
    """Element node sensitivities (psi^T dR/dX, pressure psi^T dF/dX, the
    compliance and one-element continuous KS terms) against the complex-step
    values of the reference's complex build."""
    k = f"quad4_{case}"
    X, v, psi = fixtures[k + "_X"], fixtures[k + "_v"], fixtures[k + "_psi"]
    t = float(fixtures[k + "_t"].reshape(-1)[0])

    def close(a, b):
        return np.abs(a - b).max() <= 1e-11 * np.abs(b).max()

    assert close(O.quad4_adj_xpt(X, v, psi, E, NU, t), fixtures[k + "_adjxpt"])
    assert close(O.quad4_pressure_xpt(X, 1e4, psi), fixtures[k + "_presxpt"])
    U = fixtures[k + "_sed"]
    det = fixtures[k + "_det"]
    assert close(O.quad4_point_xpt(2, X, v, E, NU, YS, t, U, det), fixtures[k + "_compxpt"])
    F = fixtures[k + "_fail"]
    ksum = np.sum(det * np.exp(KS_WEIGHT * (F - F.max())))
    ef = np.exp(KS_WEIGHT * (F - F.max())) / ksum
    assert close(O.quad4_point_xpt(1, X, v, E, NU, YS, t, ef / KS_WEIGHT, det * ef),
                 fixtures[k + "_ksxpt"])
    # the strain energy density the compliance term uses
    e, detw = O.quad4_strains(X, v, drill=False)
    sed = np.array([e[q] @ O.shell_stress(E, NU, t, e[q]) for q in range(4)])
    assert np.abs(sed - U).max() <= 1e-13 * np.abs(U).max()


def test_shell_xpt_mesh_translation_invariance_and_total_derivative():
    """Mesh sums of the node sensitivities: a rigid translation of every node
    changes neither the residual nor the functions (at fixed u), so each
    coordinate direction sums to zero over the nodes; the compliance and KS
    gradients match central differences of the oracle's mesh totals."""
    m = PlateMesh(3, 3, 0.3, 0.3)
    rng = np.random.default_rng(8)
    X = m.X.copy()
    X[:, 2] += 0.01 * rng.standard_normal(m.nnodes)
    X = X.reshape(-1)
    t = rng.uniform(0.005, 0.02, m.nelem)
    u = 1e-3 * rng.standard_normal(6 * m.nnodes)
    psi = rng.standard_normal(6 * m.nnodes)
    gA = O.shell_xpt_sens(0, m.conn, X, u, t, E, NU, psi=psi)
    gC = O.shell_xpt_sens(1, m.conn, X, u, t, E, NU)
    ks, fmax, ksum = O.shell_ks_failure(m.conn, X, u, t, E, NU, YS, KS_WEIGHT)
    gK = O.shell_xpt_sens(2, m.conn, X, u, t, E, NU, ys=YS, ks_weight=KS_WEIGHT,
                          max_fail=fmax, ks_sum=ksum)
    gP = O.shell_pressure_xpt(m.conn, X, np.arange(m.nelem), np.full(m.nelem, 1e4), psi)
    for g in (gA, gC, gK, gP):
        s = g.reshape(-1, 3).sum(axis=0)
        assert np.abs(s).max() <= 1e-9 * np.abs(g).max()
    for i in rng.choice(X.size, 6, replace=False):
        h = 1e-6
        Xp, Xm = X.copy(), X.copy()
        Xp[i] += h
        Xm[i] -= h
        dC = (O.shell_compliance(m.conn, Xp, u, t, E, NU)
              - O.shell_compliance(m.conn, Xm, u, t, E, NU)) / (2 * h)
        assert abs(dC - gC[i]) <= 1e-6 * np.abs(gC).max()
        dK = (O.shell_ks_failure(m.conn, Xp, u, t, E, NU, YS, KS_WEIGHT)[0]
              - O.shell_ks_failure(m.conn, Xm, u, t, E, NU, YS, KS_WEIGHT)[0]) / (2 * h)
        assert abs(dK - gK[i]) <= 1e-6 * np.abs(gK).max()


def test_quad4_ref_axis_transform_against_reference(fixtures):
    """TACSShellRefAxisTransform: residual, Jacobian and failure index of the
    reference build against the oracle's frame; the results do not depend
    on the axis (isotropic material, orthonormal frame), while the natural
    transform's frame, whose t1 keeps part of the normal on a warped
    element, gives different matrices."""
    axis = fixtures["quad4_axis"]
    try:
        for case in range(4):
            k = f"quad4_{case}"
            X, v = fixtures[k + "_X"], fixtures[k + "_v"]
            t = float(fixtures[k + "_t"].reshape(-1)[0])
            O.shell_set_ref_axis(axis)
            res, mat = O.quad4_element(X, v, E, NU, t)
            R, M = fixtures[k + "_axis_res"], fixtures[k + "_axis_mat"].reshape(24, 24)
            assert np.abs(res - R).max() <= 1e-14 * np.abs(R).max()
            assert np.abs(mat - M).max() <= 1e-14 * np.abs(M).max()
            for n in range(4):
                f, _ = O.quad4_failure(X, v, E, NU, YS, t, n)
                assert abs(f - fixtures[k + "_axis_fail"][n]) <= 1e-13 * abs(f)
            O.shell_set_ref_axis([0.0, 0.2, -1.0] if case % 2 else [1.0, 0.0, 0.0])
            _, mat2 = O.quad4_element(X, v, E, NU, t)
            assert np.abs(mat2 - M).max() <= 1e-12 * np.abs(M).max()
            O.shell_set_ref_axis(None)
            _, mat3 = O.quad4_element(X, v, E, NU, t)
            assert np.abs(mat3 - M).max() >= 1e-6 * np.abs(M).max()
    finally:
        O.shell_set_ref_axis(None)


@pytest.mark.parametrize("case", range(4))
def test_quad4_thickness_offset_against_reference(fixtures, case):
    """TACSIsoShellConstitutive with tOffset != 0 (membrane-bending coupling,
    shifted top and bottom faces, mass moment m1): residual, Jacobian, mass
    matrix, thickness adjoint product, failure index and its thickness
    derivative against the reference build."""
    k = f"quad4_{case}"
    X, v, psi = fixtures[k + "_X"], fixtures[k + "_v"], fixtures[k + "_psi"]
    t = float(fixtures[k + "_t"].reshape(-1)[0])
    toff = float(fixtures[k + "_toff"][0])
    try:
        O.shell_set_offset(toff)
        res, mat = O.quad4_element(X, v, E, NU, t)
        R, K = fixtures[k + "_off_res"], fixtures[k + "_off_mat"].reshape(24, 24)
        assert np.abs(res - R).max() <= 1e-14 * np.abs(R).max()
        assert np.abs(mat - K).max() <= 1e-14 * np.abs(K).max()
        M = O.quad4_mass(X, 2700.0, t)
        Mr = fixtures[k + "_off_mass"].reshape(24, 24)
        assert np.abs(M - Mr).max() <= 1e-14 * np.abs(Mr).max()
        adj = O.quad4_adj_product(X, v, psi, E, NU, t)
        ref_adj = float(fixtures[k + "_off_adj"][0])
        assert abs(adj - ref_adj) <= 1e-13 * abs(ref_adj)
        e, _ = O.quad4_strains(X, v, drill=False)
        for n in range(4):
            f, _, dfdt = O.shell_failure(E, NU, YS, t, e[n])
            assert abs(f - fixtures[k + "_off_fail"][n]) <= 1e-13 * abs(f)
            ref_dt = fixtures[k + "_off_dfail"][n]
            assert abs(dfdt - ref_dt) <= 1e-12 * abs(ref_dt)
        # the offset couples membrane and bending: K differs from tOffset = 0
        O.shell_set_offset(0.0)
        _, K0 = O.quad4_element(X, v, E, NU, t)
        assert np.abs(K0 - K).max() >= 1e-4 * np.abs(K).max()
    finally:
        O.shell_set_offset(0.0)
