"""Per-element materials of the solid family (one TACSSolidConstitutive per
element with different isotropic TACSMaterialProperties, as
TACSAssembler::setElements allows, src/TACSAssembler.cpp:623): every element
routine against the oracle summed over the material groups (element
assembly and the functions' element sums are additive; the KS aggregate
combines the groups' maxima and sums)."""
import numpy as np
import pytest

import oracle as O
from tacs_amd import TACS, functions
from tacs_amd.problems import SolidProblem

pytestmark = pytest.mark.gpu

RTOL = 1e-10
MATS = [dict(E=70e9, nu=0.3, rho=2700.0, ys=270e6),
        dict(E=200e9, nu=0.27, rho=7800.0, ys=400e6),
        dict(E=110e9, nu=0.34, rho=4500.0, ys=800e6)]


def _bc(p, v):
    out = v.copy().reshape(-1, 3)
    out[p.bc_nodes] = 0.0
    return out.reshape(-1)


@pytest.mark.parametrize("order,n", [(2, 3), (3, 2)])
def test_solid_per_element_materials_match_oracle(order, n):
    p = SolidProblem(n, order=order, dv="element", materials=MATS)
    a, m = p.assembler, p.mesh
    rng = np.random.default_rng(4 + order)
    X = m.X + 0.03 / n * rng.standard_normal(m.X.shape)
    p.X.getArray()[:] = X.reshape(-1)
    a.setNodes(p.X)
    X = X.reshape(-1)
    t = rng.uniform(0.5, 1.5, m.nelem)
    p.set_thickness(t)
    u = 1e-4 * rng.standard_normal(3 * m.nnodes)
    uv = a.createVec()
    uv.getArray()[:] = u
    a.setVariables(uv)
    conn = m.conn
    rowp, cols = O.node_csr(m.nnodes, conn, order ** 3)
    groups = [np.arange(g, m.nelem, len(MATS)) for g in range(len(MATS))]
    Cs = [O.iso_C21(d["E"], d["nu"]) for d in MATS]

    def osum(fn):
        return sum(fn(idx, Cs[g], MATS[g]) for g, idx in enumerate(groups))

    # Jacobian (alpha K + gamma M) with both Jacobian kernels, and the residual
    oA = osum(lambda idx, C, d: O.assemble_solid(order, conn[idx], X, u, t[idx], C, rowp, cols,
                                                 rho=d["rho"], alpha=1.3, gamma=0.4,
                                                 want_res=False)[0])
    O.apply_bc_mat(3, p.bc_nodes, np.full(p.bc_nodes.size, 7, np.int32), rowp, cols, oA)
    ores = osum(lambda idx, C, d: O.assemble_solid(order, conn[idx], X, u, t[idx], C, rowp, cols,
                                                   want_mat=False)[1])
    trac = np.tile(p.traction, (p.aux_elems.size, 1))
    O.add_tractions(order, conn, X, p.aux_elems, p.aux_faces, trac, ores)
    ores = _bc(p, ores)
    for kern in (1, 2):
        a.setJacobianKernel(kern)
        a.assembleJacobian(1.3, 0.0, 0.4, p.res, p.mat)
        A = p.mat.getArrays()[2].reshape(-1)
        assert np.abs(A - oA).max() <= RTOL * np.abs(oA).max(), kern
        assert np.abs(p.res.getArray() - ores).max() <= RTOL * np.abs(ores).max(), kern
    a.setJacobianKernel(0)
    # mass matrix type
    a.assembleMatType(TACS.MASS_MATRIX, p.mat)
    oM = osum(lambda idx, C, d: O.assemble_solid(order, conn[idx], X, u, t[idx], C, rowp, cols,
                                                 rho=d["rho"], alpha=0.0, gamma=1.0,
                                                 want_res=False)[0])
    O.apply_bc_mat(3, p.bc_nodes, np.full(p.bc_nodes.size, 7, np.int32), rowp, cols, oM)
    assert np.abs(p.mat.getArrays()[2].reshape(-1) - oM).max() <= RTOL * np.abs(oM).max()
    # matrix-free product (stiffness + mass)
    xv = rng.standard_normal(3 * m.nnodes)
    x, y = a.createVec(), a.createVec()
    x.getArray()[:] = xv
    mf = TACS.MatrixFreeMat(a)
    mf.assembleMatrixFreeData(TACS.JACOBIAN_MATRIX, 0.8, 0.0, 0.3)
    mf.mult(x, y)
    oy = _bc(p, osum(lambda idx, C, d: O.matfree_product(order, conn[idx], X, t[idx], C, xv,
                                                         rho=d["rho"], kc=0.8, mc=0.3)))
    assert np.abs(y.getArray() - oy).max() <= 1e-12 * np.abs(oy).max()
    # functions and their sensitivities
    comp = functions.Compliance(a)
    ksf = functions.KSFailure(a, ksWeight=20.0)
    vals = a.evalFunctions([comp, ksf])
    oc = osum(lambda idx, C, d: O.compliance(order, conn[idx], X, u, t[idx], C))
    parts = [O.ks_failure(order, conn[idx], X, u, Cs[g], MATS[g]["ys"], 20.0)
             for g, idx in enumerate(groups)]
    fmax = max(q[1] for q in parts)
    ksum = sum(q[2] * np.exp(20.0 * (q[1] - fmax)) for q in parts)
    oks = fmax + np.log(ksum) / 20.0
    assert abs(vals[0] - oc) <= RTOL * abs(oc)
    assert abs(vals[1] - oks) <= RTOL * abs(oks)
    dfdu = [a.createVec(), a.createVec()]
    a.addSVSens([comp, ksf], dfdu)
    odc = osum(lambda idx, C, d: O.compliance_svsens(order, conn[idx], X, u, t[idx], C))
    odk = osum(lambda idx, C, d: O.ks_svsens(order, conn[idx], X, u, C, d["ys"], 20.0, fmax, ksum))
    assert np.abs(dfdu[0].getArray() - odc).max() <= RTOL * np.abs(odc).max()
    assert np.abs(dfdu[1].getArray() - odk).max() <= RTOL * np.abs(odk).max()
    dfdx = a.createDesignVec()
    a.addDVSens([comp], [dfdx], 1.0)
    odt = np.zeros(m.nelem)
    for g, idx in enumerate(groups):
        odt[idx] = O.compliance_dvsens(order, conn[idx], X, u, Cs[g])
    assert np.abs(dfdx.getArray() - odt).max() <= RTOL * np.abs(odt).max()
    dX = a.createNodeVec()
    a.addXptSens([comp], [dX])
    oX = osum(lambda idx, C, d: O.compliance_xptsens(order, conn[idx], X, u, t[idx], C))
    assert np.abs(dX.getArray() - oX).max() <= RTOL * np.abs(oX).max()
