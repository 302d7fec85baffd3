"""The configurations' defining workloads on the GPU at their real sizes,
checked against the oracle:

* 20^3 8-node cube (E = 70e9, nu = 0.3, x = 0 clamped, 1e6 Pa traction on
  x = 1): the whole matrix and residual against the oracle's assembly, FP64
  GMRES(30) against the oracle's GMRES, compliance and its node gradient
  against the oracle's at the oracle's solution.
* 128^3 8-node cube, mixed-precision recycling solve + adjoint: the FP64
  true residual of the GPU solution through the oracle's block-CSR product,
  sampled Jacobian rows against the oracle's assembly of their elements, the
  gradient against the FP64 recycling path (the 1e-6 bar).
* 96^3 SIMP cube through the autograd function: the same residual, row and
  gradient checks.
* 27-node cube with 16 load cases as one block (FP32-valued operator,
  recycling per load case): every solution against the oracle's dense solve
  (4^3), and the true residual of every load case at 40^3.
* 32 cubes of edge 8..24 (seed 11) in one union solve: every cube's FP64
  true residual, compliance = f^T u, sampled cubes' matrices against the
  oracle.
"""
import numpy as np
import pytest

import oracle as O
from tacs_amd import TACS, functions
from tacs_amd.problems import MultiCubeProblem, SolidProblem

pytestmark = pytest.mark.gpu


def _check_rows(mesh, order, X, t, C, rows, rowp, cols, A, bc_nodes=()):
    """GPU block rows `rows` against the oracle's assembly of every element
    incident to them (a local system over those elements' nodes)."""
    npe = order ** 3
    el = np.nonzero(np.isin(mesh.conn, rows).any(axis=1))[0]
    sub = mesh.conn[el]
    nodes = np.unique(sub)
    lconn = np.searchsorted(nodes, sub).astype(np.int32)
    lrowp, lcols = O.node_csr(nodes.size, lconn, npe)
    Xl = X.reshape(-1, 3)[nodes].reshape(-1)
    lA, _ = O.assemble_solid(order, lconn, Xl, None, t[el], C, lrowp, lcols, want_res=False)
    lbc = np.nonzero(np.isin(nodes, bc_nodes))[0].astype(np.int32)
    if lbc.size:
        O.apply_bc_mat(3, lbc, np.full(lbc.size, 7, np.int32), lrowp, lcols, lA)
    scale = np.abs(lA).max()
    lA = lA.reshape(-1, 9)
    for r in rows:
        li = int(np.searchsorted(nodes, r))
        gc = cols[rowp[r]:rowp[r + 1]]
        oc = nodes[lcols[lrowp[li]:lrowp[li + 1]]]
        assert np.array_equal(gc, oc), r
        gA = A[rowp[r]:rowp[r + 1]].reshape(-1, 9)
        oA = lA[lrowp[li]:lrowp[li + 1]]
        assert np.abs(gA - oA).max() <= 1e-10 * scale, r


def _true_residual(rowp, cols, A, u, res):
    """|K u + res| / |res| in FP64 through the oracle's block-CSR product."""
    r = O.bsr_mult(3, rowp, cols, A.reshape(-1), u) + res
    return np.linalg.norm(r) / np.linalg.norm(res)


def test_cfg1_cube20_full_against_oracle():
    p = SolidProblem(20)
    a, m = p.assembler, p.mesh
    X = m.X.reshape(-1)
    t = np.ones(m.nelem)
    p.assemble()
    rowp, cols, A = p.mat.getArrays()
    res = p.res.getArray().copy()
    C = O.iso_C21(70e9, 0.3)
    orp, ocols = O.node_csr(m.nnodes, m.conn, 8)
    oA, ores = O.assemble_solid(2, m.conn, X, np.zeros(3 * m.nnodes), t, C, orp, ocols)
    trac = np.tile(p.traction, (p.aux_elems.size, 1))
    O.add_tractions(2, m.conn, X, p.aux_elems, p.aux_faces, trac, ores)
    bcv = np.full(p.bc_nodes.size, 7, np.int32)
    O.apply_bc_mat(3, p.bc_nodes, bcv, orp, ocols, oA)
    O.apply_bc_vec(3, p.bc_nodes, bcv, ores, u=np.zeros(3 * m.nnodes))
    assert np.array_equal(rowp, orp) and np.array_equal(cols, ocols)
    assert np.abs(A.reshape(-1) - oA).max() <= 1e-10 * np.abs(oA).max()
    assert np.abs(res - ores).max() <= 1e-10 * np.abs(ores).max()
    # the reference's solver settings: GMRES(30), MGS, block-Jacobi, 1e-8
    flag, its, pc, ksm = p.solve_fp64(m=30, restarts=10000, rtol=1e-8)
    assert flag == 1
    u = p.u.getArray().copy()
    D = O.bsr_diag_inv(3, orp, ocols, oA)
    uo, oflag, oits, orn = O.gmres(3, orp, ocols, oA, -ores, D, m=30, nrestart=10000, rtol=1e-8)
    assert oflag == 1
    # same algorithm in the same arithmetic: the step counts agree closely
    assert abs(its - oits) <= 0.05 * oits + 2, (its, oits)
    assert _true_residual(orp, ocols, oA, u, ores) <= 1.01e-8
    assert np.abs(u - uo).max() <= 1e-5 * np.abs(uo).max()
    # compliance and its node gradient (adjoint psi = 2 u for the compliance)
    comp = functions.Compliance(a)
    c = a.evalFunctions([comp])[0]
    co = O.compliance(2, m.conn, X, uo, t, C)
    assert abs(c - co) <= 1e-7 * abs(co)
    dfdu = a.createVec()
    a.addSVSens([comp], [dfdu])
    a.applyBCs(dfdu)
    psi = a.createVec()
    assert ksm.solve(dfdu, psi) == 1
    g = a.createNodeVec()
    a.addXptSens([comp], [g])
    a.addAdjointResXptSensProducts([psi], [g], -1.0)
    po = 2.0 * uo
    go = (O.compliance_xptsens(2, m.conn, X, uo, t, C)
          - O.adj_res_xpt_product(2, m.conn, X, uo, po, t, C)
          - O.traction_xpt_product(2, m.conn, X, p.aux_elems, p.aux_faces, trac, po))
    assert np.abs(g.getArray() - go).max() <= 1e-5 * np.abs(go).max()


def _solve_and_gradient(p, ksm, pc):
    """Forward solve, compliance, adjoint solve (same solver) and dC/dt."""
    a = p.assembler
    pc.factor()
    assert ksm.solve(p.res, p.u) == 1
    p.u.scale(-1.0)
    a.setVariables(p.u)
    comp = functions.Compliance(a)
    c = a.evalFunctions([comp])[0]
    dfdu = a.createVec()
    a.addSVSens([comp], [dfdu])
    a.applyBCs(dfdu)
    psi = a.createVec()
    assert ksm.solve(dfdu, psi) == 1
    dfdx = a.createDesignVec()
    a.addDVSens([comp], [dfdx], 1.0)
    a.addAdjointResProducts([psi], [dfdx], -1.0)
    return c, p.u.getArray().copy(), dfdx.getArray().copy()


def test_cfg2_cube128_mixed_solve_full_size():
    p = SolidProblem(128, dv="element")
    m = p.mesh
    p.assemble()
    res = p.res.getArray().copy()
    pc = TACS.Pc(p.mat)
    ksm = TACS.MixedKSM(p.mat, pc, 60, 100000, deflation=20)
    ksm.setTolerances(1e-8, 1e-30)
    c, u, g = _solve_and_gradient(p, ksm, pc)
    rowp, cols, A = p.mat.getArrays()
    # the FP64 true residual of the mixed solution
    assert _true_residual(rowp, cols, A, u, res) <= 1.01e-8
    # sampled rows: corners, edges, faces, the clamped face and random interior nodes
    n1 = m.grid[0]
    rng = np.random.default_rng(128)
    pick = [0, n1 - 1, n1 * n1 - 1, m.nnodes - 1, m.nnodes // 2, int(p.bc_nodes[37])]
    pick += [int(v) for v in rng.choice(m.nnodes, 24, replace=False)]
    C = O.iso_C21(70e9, 0.3)
    _check_rows(m, 2, m.X, np.ones(m.nelem), C, np.unique(pick), rowp, cols, A, p.bc_nodes)
    del A
    # gradient against the FP64 operator with the same recycling algorithm
    p.assemble()
    rk = TACS.RecyclingKSM(p.mat, pc, m=60, deflation=20, prec=64)
    rk.setTolerances(1e-8, 1e-30)
    c64, u64, g64 = _solve_and_gradient(p, rk, pc)
    assert np.abs(g - g64).max() <= 1e-6 * np.abs(g64).max()
    assert np.abs(u - u64).max() <= 1e-6 * np.abs(u64).max()
    assert abs(c - c64) <= 1e-7 * abs(c64)


def test_cfg3_simp96_full_size():
    import torch
    from tacs_amd.autograd import SolidComplianceProblem, simp_compliance
    from tacs_amd.meshes import simp_thickness

    prob = SolidProblem(96, E=1.0, nu=0.3, traction=(0.0, 0.0, 1.0), dv="element",
                        traction_patch=8)
    m = prob.mesh
    rng = np.random.default_rng(0)
    x0 = rng.uniform(0.1, 1.0, m.nelem)
    sp = SolidComplianceProblem(prob, p=3.0, E0=1.0, Emin=1e-9, rtol=1e-8, m=60, deflation=20,
                                restarts=2000)
    x = torch.tensor(x0, dtype=torch.float64, device="cuda", requires_grad=True)
    c = simp_compliance(x, sp)
    c.backward()
    g = x.grad.detach().cpu().numpy()
    # residual at the forward's state: re-assemble the residual at u = 0
    a = prob.assembler
    u = prob.u.getArray().copy()
    a.zeroVariables()
    a.assembleRes(prob.res)
    res = prob.res.getArray().copy()
    rowp, cols, A = prob.mat.getArrays()
    assert _true_residual(rowp, cols, A, u, res) <= 1.01e-8
    t = simp_thickness(x0)
    rows = np.unique(rng.choice(m.nnodes, 24, replace=False))
    _check_rows(m, 2, m.X, t, O.iso_C21(1.0, 0.3), rows, rowp, cols, A, prob.bc_nodes)
    del A
    # the FP64 recycling path gives the same gradient
    sp64 = SolidComplianceProblem(prob, p=3.0, E0=1.0, Emin=1e-9, rtol=1e-8, m=60, mixed=False,
                                  deflation=20, restarts=2000)
    x64 = torch.tensor(x0, dtype=torch.float64, device="cuda", requires_grad=True)
    c64 = simp_compliance(x64, sp64)
    c64.backward()
    g64 = x64.grad.detach().cpu().numpy()
    assert abs(c.item() - c64.item()) <= 1e-7 * abs(c64.item())
    assert np.abs(g - g64).max() <= 1e-6 * np.abs(g64).max()


def _face_loads(p, R, seed=5):
    """Per-node tractions on the x = 1 face, N(0, 1e6) Pa per component, as
    nodal forces (the batched workload's loads)."""
    a = p.assembler
    rng = np.random.default_rng(seed)
    face = p.mesh.nodes_on_x(1)
    vecs = []
    for _ in range(R):
        v = a.createVec()
        arr = v.getArray()
        arr[:] = 0.0
        arr.reshape(-1, 3)[face] = 1e6 * rng.standard_normal((face.size, 3)) / face.size
        a.applyBCs(v)
        vecs.append(v)
    return vecs


def test_cfg5a_hex27_16_load_cases_against_dense_oracle():
    p = SolidProblem(4, order=3)
    m = p.mesh
    p.assemble()
    pc = TACS.Pc(p.mat)
    pc.factor()
    bs = _face_loads(p, 16)
    xs = [p.assembler.createVec() for _ in bs]
    ksm = TACS.BatchedKSM(p.mat, pc, m=60, restarts=100000, nrhs=16, prec=32, deflation=20)
    ksm.setTolerances(1e-10, 1e-30)
    assert ksm.solve(bs, xs) == 1
    it, fl, r0, rn, cyc = ksm.getStats()
    assert np.all(fl == 1) and np.all(it > 0)
    C = O.iso_C21(70e9, 0.3)
    orp, ocols = O.node_csr(m.nnodes, m.conn, 27)
    oA, _ = O.assemble_solid(3, m.conn, m.X.reshape(-1), None, np.ones(m.nelem), C, orp, ocols,
                             want_res=False)
    O.apply_bc_mat(3, p.bc_nodes, np.full(p.bc_nodes.size, 7, np.int32), orp, ocols, oA)
    n = m.nnodes
    K = np.zeros((3 * n, 3 * n))
    Ab = oA.reshape(-1, 3, 3)
    for i in range(n):
        for j in range(orp[i], orp[i + 1]):
            K[3 * i:3 * i + 3, 3 * ocols[j]:3 * ocols[j] + 3] = Ab[j]
    # dense solve on the free dofs (unit-diagonal BC rows next to 1e11 entries
    # would spoil its accuracy)
    free = np.ones(3 * n, bool)
    for bn in p.bc_nodes:
        free[3 * bn:3 * bn + 3] = False
    Kf = K[np.ix_(free, free)]
    for b, x in zip(bs, xs):
        uo = np.zeros(3 * n)
        uo[free] = np.linalg.solve(Kf, b.getArray()[free])
        xv = x.getArray()
        assert not xv[~free].any()
        assert np.abs(xv - uo).max() <= 1e-8 * np.abs(uo).max()
    # a second solve with the same matrix reuses the recycled spaces: fewer steps
    for x in xs:
        x.zeroEntries()
    assert ksm.solve(bs, xs) == 1
    it2 = ksm.getStats()[0]
    assert it2.sum() < it.sum()


def test_cfg5a_hex27_40cube_16_load_cases_full_size():
    p = SolidProblem(40, order=3, dv="element")
    p.assemble()
    pc = TACS.Pc(p.mat)
    pc.factor()
    bs = _face_loads(p, 16)
    xs = [p.assembler.createVec() for _ in bs]
    ksm = TACS.BatchedKSM(p.mat, pc, m=60, restarts=100000, nrhs=16, prec=32, deflation=20)
    ksm.setTolerances(1e-8, 1e-30)
    assert ksm.solve(bs, xs) == 1
    assert np.all(ksm.getStats()[1] == 1)
    rowp, cols, A = p.mat.getArrays()
    Af = A.reshape(-1)
    for b, x in zip(bs, xs):
        bv = b.getArray()
        r = O.bsr_mult(3, rowp, cols, Af, x.getArray()) - bv
        assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(bv)


def test_cfg5b_32_uneven_cubes_union_solve():
    rng = np.random.default_rng(11)
    sizes = [int(v) for v in rng.integers(8, 25, 32)]
    p = MultiCubeProblem(sizes, dv="element")
    a = p.assembler
    p.assemble()
    res = p.res.getArray().copy()
    pc = TACS.Pc(p.mat)
    pc.factor()
    rhs = a.createVec()
    p.scale_cubes(p.res, rhs)
    rtol = 1e-8
    ksm = TACS.MixedKSM(p.mat, pc, 60, 100000, deflation=20)
    ksm.setTolerances(1e-30, rtol)
    assert ksm.solve(rhs, p.u) == 1
    p.unscale_cubes(p.u)
    p.u.scale(-1.0)
    a.setVariables(p.u)
    u = p.u.getArray().copy()
    rowp, cols, A = p.mat.getArrays()
    Ku = O.bsr_mult(3, rowp, cols, A.reshape(-1), u)
    for c in range(len(sizes)):
        sl = p.dof(c)
        assert np.linalg.norm((Ku + res)[sl]) <= 1.01 * rtol * np.linalg.norm(res[sl]), c
    # compliance (e.s integrated) equals f^T u at equilibrium
    comp = functions.Compliance(a)
    cval = a.evalFunctions([comp])[0]
    assert abs(cval - (-res) @ u) <= 1e-7 * abs(cval)
    # the smallest and the largest cube's matrices against the oracle
    C = O.iso_C21(70e9, 0.3)
    for c in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        mc = p.meshes[c]
        orp, ocols = O.node_csr(mc.nnodes, mc.conn, 8)
        oA, _ = O.assemble_solid(2, mc.conn, mc.X.reshape(-1), None, np.ones(mc.nelem), C, orp,
                                 ocols, want_res=False)
        bn = mc.nodes_on_x(0)
        O.apply_bc_mat(3, bn, np.full(bn.size, 7, np.int32), orp, ocols, oA)
        n0, n1 = int(p.node_off[c]), int(p.node_off[c + 1])
        assert np.array_equal(rowp[n0:n1 + 1] - rowp[n0], orp)
        assert np.array_equal(cols[rowp[n0]:rowp[n1]] - n0, ocols)
        blk = A[rowp[n0]:rowp[n1]].reshape(-1)
        assert np.abs(blk - oA).max() <= 1e-10 * np.abs(oA).max()
