"""The oracle against element-level values computed by the reference's own
sources (tests/golden/elements.npz, written by tests/golden/make_fixtures.py
from oracle/_ref), and, where oracle/_ref is built, against fresh inputs."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "elements.npz"), allow_pickle=False)
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libtacs_ref.so")


@pytest.mark.parametrize("key", ["hex8_0", "hex8_1", "hex27_0", "hex27_1"])
def test_solid_element_matches_reference(key):
    order = 2 if key.startswith("hex8") else 3
    X, u, t = FIX[key + "_X"], FIX[key + "_u"], FIX[key + "_t"][0]
    res, mat = O.solid_element(order, X, u, O.iso_C21(70e9, 0.3), t=t, rho=2700.0, alpha=1.0,
                               gamma=0.25)
    rr, rm = FIX[key + "_res"], FIX[key + "_mat"]
    assert np.abs(res - rr).max() <= 1e-13 * np.abs(rr).max()
    assert np.abs(mat.reshape(-1) - rm).max() <= 1e-13 * np.abs(rm).max()
    trac = FIX[key + "_trac"]
    for face in range(6):
        ro = np.zeros(res.size)
        O.lib().orc_traction_face(order, face, O._p(np.ascontiguousarray(X)),
                                  O._p(np.ascontiguousarray(trac)), O._p(ro))
        ref = FIX[key + "_trac_res"][face]
        assert np.abs(ro - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("key", ["hex8_0", "hex8_1", "hex27_0", "hex27_1"])
def test_solid_pressure_matches_reference(key):
    """TACSPressure3D face residuals of the reference build on distorted
    hexes (every face), and the closed form on a box: the face residual sums
    to -p n A, i.e. a force p n A along the outward normal n (positive
    pressure pulls; pyTACS flips the sign of Nastran PLOAD4 on solids)."""
    order = 2 if key.startswith("hex8") else 3
    X, pv = FIX[key + "_X"], float(FIX[key + "_p"][0])
    for face in range(6):
        ref = FIX[key + "_p_res"][face]
        ro = O.pressure_face(order, face, X, pv)
        assert np.abs(ro - ref).max() <= 1e-13 * np.abs(ref).max()
    k = [-1.0, 1.0] if order == 2 else [-1.0, 0.0, 1.0]
    box = np.array([[x, y, z] for z in k for y in k for x in k]) * [0.5, 0.75, 1.25] + 2.0
    dims = np.array([1.0, 1.5, 2.5])
    for face in range(6):
        ax, sgn = face // 2, (1.0 if face % 2 else -1.0)
        n = np.zeros(3)
        n[ax] = sgn
        area = np.prod(np.delete(dims, ax))
        r = O.pressure_face(order, face, box.reshape(-1), 3.0).reshape(-1, 3)
        assert np.allclose(r.sum(0), -3.0 * n * area, rtol=1e-13, atol=1e-13)
        # only the nodes of the face carry load
        on = np.isclose(box[:, ax], 2.0 + sgn * dims[ax] / 2)
        assert np.abs(r[~on]).max() == 0.0


@pytest.mark.parametrize("key", ["hex8_0", "hex8_1", "hex27_0", "hex27_1"])
def test_solid_matvec_matches_reference(key):
    """Matrix-free element products (TACSElement3D::getMatVecProductData +
    addMatVecProduct) of the reference build for the Jacobian (alpha 0.8,
    gamma 0.3), stiffness and mass types, and the same products from the
    oracle's element matrices (K x with K = alpha Ks + gamma M)."""
    order = 2 if key.startswith("hex8") else 3
    X, t, x = FIX[key + "_X"], float(FIX[key + "_t"][0]), FIX[key + "_mvx"]
    C = O.iso_C21(70e9, 0.3)
    coefs = [(0.8, 0.3), (0.8, 0.0), (0.0, 0.3)]
    for mt, (kc, mc) in enumerate(coefs):
        ref = FIX[key + "_mv"][mt]
        y = O.solid_matvec_element(order, X, x, C, t=t, rho=2700.0, kc=kc, mc=mc)
        assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()
        _, K = O.solid_element(order, X, np.zeros_like(x), C, t=t, rho=2700.0, alpha=kc, gamma=mc,
                               want_res=False)
        assert np.abs(K @ x - y).max() <= 1e-13 * np.abs(y).max()
    # the mass product of a constant field integrates rho t over the element
    ones = np.tile([1.0, 0.0, 0.0], order ** 3)
    m = O.solid_matvec_element(order, X, ones, C, t=t, rho=2700.0, kc=0.0, mc=1.0)
    vol = sum(O.hex_quadrature(order, q)[0] * abs(np.linalg.det(
        O.hex_basis(order, O.hex_quadrature(order, q)[1])[1].reshape(-1, 3).T
        @ X.reshape(-1, 3))) for q in range(order ** 3))
    assert np.isclose(m[0::3].sum(), 2700.0 * t * vol, rtol=1e-12)
    assert np.abs(m[1::3]).max() == 0.0 and np.abs(m[2::3]).max() == 0.0


@pytest.mark.parametrize("key", ["hex8_0", "hex8_1", "hex27_0", "hex27_1"])
def test_solid_mat_types_match_reference(key):
    """Stiffness, mass and geometric-stiffness element matrices
    (TACSElement3D::getMatType) of the reference build against the oracle;
    the geometric stiffness is symmetric, acts per displacement component
    and annihilates rigid translations (sum_b grad N_b = 0)."""
    order = 2 if key.startswith("hex8") else 3
    X, u, t = FIX[key + "_X"], FIX[key + "_u"], float(FIX[key + "_t"][0])
    ref = FIX[key + "_mattype"]
    C = O.iso_C21(70e9, 0.3)
    z = np.zeros_like(u)
    _, K = O.solid_element(order, X, z, C, t=t, rho=2700.0, alpha=1.0, gamma=0.0, want_res=False)
    _, M = O.solid_element(order, X, z, C, t=t, rho=2700.0, alpha=0.0, gamma=1.0, want_res=False)
    G = O.solid_geo_element(order, X, u, C, t=t)
    for k, mine in enumerate((K, M, G)):
        assert np.abs(mine.reshape(-1) - ref[k]).max() <= 1e-13 * np.abs(ref[k]).max()
    assert np.abs(G - G.T).max() <= 1e-14 * np.abs(G).max()
    for j in range(3):
        e = np.zeros(3)
        e[j] = 1.0
        assert np.abs(G @ np.tile(e, order ** 3)).max() <= 1e-12 * np.abs(G).max()
    # no coupling between displacement components
    Gb = G.reshape(order ** 3, 3, order ** 3, 3)
    assert np.abs(Gb[:, 0, :, 1]).max() == 0.0 and np.abs(Gb[:, 1, :, 2]).max() == 0.0
    # linear in the state
    assert np.allclose(O.solid_geo_element(order, X, 2.5 * u, C, t=t), 2.5 * G, rtol=1e-13,
                       atol=1e-13 * np.abs(G).max())


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (no reference tree)")
def test_solid_element_fresh_inputs_against_reference_build():
    R = ctypes.CDLL(REF, mode=ctypes.RTLD_GLOBAL)
    rng = np.random.default_rng(99)
    c = ctypes.c_double
    for order in (2, 3):
        npe = order ** 3
        k = [-1.0, 1.0] if order == 2 else [-1.0, 0.0, 1.0]
        base = np.array([[x, y, z] for z in k for y in k for x in k])
        for _ in range(3):
            X = (base * rng.uniform(0.5, 2.0, 3) + 0.05 * rng.standard_normal((npe, 3))).reshape(-1)
            u = rng.standard_normal(3 * npe)
            res = np.zeros(3 * npe)
            mat = np.zeros(9 * npe * npe)
            R.ref_solid_element(order, X.ctypes.data_as(ctypes.c_void_p),
                                u.ctypes.data_as(ctypes.c_void_p), c(200e9), c(0.25), c(7800.0),
                                c(0.6), c(0.5), c(2.0), res.ctypes.data_as(ctypes.c_void_p),
                                mat.ctypes.data_as(ctypes.c_void_p))
            r2, m2 = O.solid_element(order, X, u, O.iso_C21(200e9, 0.25), t=0.6, rho=7800.0,
                                     alpha=0.5, gamma=2.0)
            assert np.abs(r2 - res).max() <= 1e-13 * np.abs(res).max()
            assert np.abs(m2.reshape(-1) - mat).max() <= 1e-13 * np.abs(mat).max()
