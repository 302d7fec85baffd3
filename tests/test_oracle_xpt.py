"""Pins of the oracle's node-coordinate sensitivities (oracle/tacs_oracle.c:
orc_compliance_xptsens, orc_adj_res_xpt_product, orc_traction_xpt_product):

- element level against the reference's own TACSElement3D::addAdjResXptProduct
  and the compliance point-quantity chain (tests/golden/elements.npz);
- mesh level against central finite differences of the oracle's compliance,
  KS failure, element residual and traction residual on a distorted 2x2x2
  mesh.
"""
import os

import numpy as np
import pytest

import oracle as O
from tacs_amd.meshes import HexMesh

HERE = os.path.dirname(os.path.abspath(__file__))
C = O.iso_C21(70e9, 0.3)


@pytest.fixture(scope="module")
def fixtures():
    return np.load(os.path.join(HERE, "golden", "elements.npz"))


@pytest.mark.parametrize("key", ["hex8_0", "hex8_1", "hex27_0", "hex27_1"])
def test_xpt_sens_against_reference_element(fixtures, key):
    order = 2 if key.startswith("hex8") else 3
    npe = order ** 3
    X, u, psi = fixtures[key + "_X"], fixtures[key + "_u"], fixtures[key + "_psi"]
    t = np.array([fixtures[key + "_t"][0]])
    conn = np.arange(npe, dtype=np.int32)
    adj = O.adj_res_xpt_product(order, conn, X, u, psi, t, C)
    ref = fixtures[key + "_adjxpt"]
    assert np.abs(adj - ref).max() <= 1e-12 * np.abs(ref).max()
    cx = O.compliance_xptsens(order, conn, X, u, t, C)
    ref = fixtures[key + "_compxpt"]
    assert np.abs(cx - ref).max() <= 1e-12 * np.abs(ref).max()


def _distorted(order=2):
    m = HexMesh(2, 2, 2, order=order, L=(1.0, 0.8, 1.2))
    rng = np.random.default_rng(4)
    X = m.X + 0.03 * rng.standard_normal(m.X.shape)
    return m, X.reshape(-1), rng


@pytest.mark.parametrize("order", [2, 3])
def test_xpt_sens_central_differences(order):
    m, X, rng = _distorted(order)
    u = 1e-3 * rng.standard_normal(3 * m.nnodes)
    psi = rng.standard_normal(3 * m.nnodes)
    t = rng.uniform(0.3, 1.0, m.nelem)
    rowp, cols = O.node_csr(m.nnodes, m.conn, m.npe)
    ae, af = m.elems_on_x(1)
    trac = np.tile([0.3, -0.5, 1.0], (ae.size, 1))
    dc = O.compliance_xptsens(order, m.conn, X, u, t, C)
    dr = O.adj_res_xpt_product(order, m.conn, X, u, psi, t, C)
    dt = O.traction_xpt_product(order, m.conn, X, ae, af, trac, psi)

    def comp(Xp):
        return O.compliance(order, m.conn, Xp, u, t, C)

    ks = O.ks_failure(order, m.conn, X, u, C, 270e6, 30.0)
    dk = O.ks_xptsens(order, m.conn, X, u, C, 270e6, 30.0, ks[1], ks[2])

    def ksf(Xp):
        return O.ks_failure(order, m.conn, Xp, u, C, 270e6, 30.0)[0]

    def res_el(Xp):
        return psi @ O.assemble_solid(order, m.conn, Xp, u, t, C, rowp, cols, want_mat=False)[1]

    def res_tr(Xp):
        r = np.zeros(3 * m.nnodes)
        O.add_tractions(order, m.conn, Xp, ae, af, trac, r)
        return psi @ r

    h = 1e-6
    for i in rng.choice(X.size, 15, replace=False):
        Xp, Xm = X.copy(), X.copy()
        Xp[i] += h
        Xm[i] -= h
        for f, d in ((comp, dc), (res_el, dr), (res_tr, dt), (ksf, dk)):
            fd = (f(Xp) - f(Xm)) / (2 * h)
            assert abs(fd - d[i]) <= 1e-6 * np.abs(d).max(), (f.__name__, i, fd, d[i])


def _solve_dense(m, X, t, ae, af, trac):
    rowp, cols = O.node_csr(m.nnodes, m.conn, m.npe)
    A, res = O.assemble_solid(m.order, m.conn, X, np.zeros(3 * m.nnodes), t, C, rowp, cols)
    O.add_tractions(m.order, m.conn, X, ae, af, trac, res)
    n = m.nnodes
    K = np.zeros((3 * n, 3 * n))
    Ab = A.reshape(-1, 3, 3)
    for i in range(n):
        for j in range(rowp[i], rowp[i + 1]):
            K[3 * i:3 * i + 3, 3 * cols[j]:3 * cols[j] + 3] = Ab[j]
    free = np.ones(3 * n, bool)
    for b in m.nodes_on_x(0):
        free[3 * b:3 * b + 3] = False
    u = np.zeros(3 * n)
    u[free] = -np.linalg.solve(K[np.ix_(free, free)], res[free])
    return u, K, free


def test_total_compliance_node_gradient_fd():
    """dC/dX = dC/dX|_u - psi^T dR/dX with K^T psi = dC/du (the adjoint chain
    the assembler applies) against central differences of the re-solved
    compliance."""
    m, X, rng = _distorted(2)
    t = rng.uniform(0.5, 1.0, m.nelem)
    ae, af = m.elems_on_x(1)
    trac = np.tile([0.0, 0.0, 1.0], (ae.size, 1))
    u, K, free = _solve_dense(m, X, t, ae, af, trac)
    dcdu = O.compliance_svsens(2, m.conn, X, u, t, C)
    dcdu[~free] = 0.0
    psi = np.zeros_like(u)
    psi[free] = np.linalg.solve(K[np.ix_(free, free)].T, dcdu[free])
    g = (O.compliance_xptsens(2, m.conn, X, u, t, C)
         - O.adj_res_xpt_product(2, m.conn, X, u, psi, t, C)
         - O.traction_xpt_product(2, m.conn, X, ae, af, trac, psi))

    def comp(Xp):
        up, _, _ = _solve_dense(m, Xp, t, ae, af, trac)
        return O.compliance(2, m.conn, Xp, up, t, C)

    h = 1e-6
    for i in rng.choice(X.size, 10, replace=False):
        Xp, Xm = X.copy(), X.copy()
        Xp[i] += h
        Xm[i] -= h
        fd = (comp(Xp) - comp(Xm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * np.abs(g).max()
