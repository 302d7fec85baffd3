"""Many independent hex8 cubes of uneven sizes in one assembler (the
block-diagonal union): one assembly launch for all cubes, one solve, every
cube to its own relative residual.  Checked against the oracle per cube."""
import numpy as np
import pytest

import oracle as O
from tacs_amd import TACS
from tacs_amd.problems import MultiCubeProblem

pytestmark = pytest.mark.gpu

SIZES = [3, 2, 4, 3]


def _oracle_cube(n, traction):
    from tacs_amd.meshes import HexMesh
    m = HexMesh(n, n, n)
    C = O.iso_C21(70e9, 0.3)
    rowp, cols = O.node_csr(m.nnodes, m.conn, 8)
    A, res = O.assemble_solid(2, m.conn, m.X.reshape(-1), np.zeros(3 * m.nnodes),
                              np.ones(m.nelem), C, rowp, cols)
    ae, af = m.elems_on_x(1)
    O.add_tractions(2, m.conn, m.X.reshape(-1), ae, af, np.tile(traction, (ae.size, 1)), res)
    bcn = m.nodes_on_x(0)
    bcv = np.full(bcn.size, 7, np.int32)
    O.apply_bc_mat(3, bcn, bcv, rowp, cols, A)
    O.apply_bc_vec(3, bcn, bcv, res, u=np.zeros(3 * m.nnodes))
    return rowp, cols, A, res


def test_union_assembly_matches_oracle_per_cube():
    p = MultiCubeProblem(SIZES)
    p.assemble()
    rowp, cols, A = p.mat.getArrays()
    r = p.res.getArray()
    for c, n in enumerate(SIZES):
        orp, ocols, oA, ores = _oracle_cube(n, p.traction)
        n0, n1 = int(p.node_off[c]), int(p.node_off[c + 1])
        rp = rowp[n0:n1 + 1] - rowp[n0]
        assert np.array_equal(rp, orp)
        cc = cols[rowp[n0]:rowp[n1]] - n0
        assert np.array_equal(cc, ocols)
        blk = A[rowp[n0]:rowp[n1]].reshape(-1)
        assert np.abs(blk - oA).max() <= 1e-10 * np.abs(oA).max()
        assert np.abs(r[p.dof(c)] - ores).max() <= 1e-10 * np.abs(ores).max()


@pytest.mark.parametrize("deflation", [0, 8])
def test_union_solve_reaches_each_cube_tolerance(deflation):
    p = MultiCubeProblem(SIZES)
    p.assemble()
    pc = TACS.Pc(p.mat)
    pc.factor()
    rtol = 1e-9
    ksm = TACS.MixedKSM(p.mat, pc, 30, 1000, deflation=deflation)
    ksm.setTolerances(1e-30, p.cube_atol(p.res, rtol))
    assert ksm.solve(p.res, p.u) == 1
    u = -p.u.getArray()
    for c, n in enumerate(SIZES):
        orp, ocols, oA, ores = _oracle_cube(n, p.traction)
        K = np.zeros((ores.size, ores.size))
        Ab = oA.reshape(-1, 3, 3)
        for i in range(orp.size - 1):
            for j in range(orp[i], orp[i + 1]):
                K[3 * i:3 * i + 3, 3 * ocols[j]:3 * ocols[j] + 3] = Ab[j]
        uc = u[p.dof(c)]
        assert np.linalg.norm(K @ uc + ores) <= 1.5 * rtol * np.linalg.norm(ores), c
        # dense solve on the free dofs (the identity BC rows would spoil its scaling)
        free = np.ones(ores.size, bool)
        for b in p.meshes[c].nodes_on_x(0):
            free[3 * b:3 * b + 3] = False
        uo = np.zeros_like(uc)
        uo[free] = -np.linalg.solve(K[np.ix_(free, free)], ores[free])
        assert not uc[~free].any()
        assert np.abs(uc - uo).max() <= 1e-6 * np.abs(uo).max(), c


@pytest.mark.parametrize("prec,deflation,cheb", [(64, 0, 1), (32, 0, 1), (32, 8, 1), (64, 8, 1),
                                                (32, 8, 6)])
def test_segmented_solve_per_cube_against_oracle(prec, deflation, cheb):
    """Per-cube Krylov spaces (SegmentedKSM): every cube converges on its own
    test and matches the oracle's dense solve of that cube (with block-Jacobi
    or its Chebyshev polynomial)."""
    p = MultiCubeProblem(SIZES)
    p.assemble()
    pc = TACS.Pc(p.mat)
    if cheb > 1:
        pc.setChebyshev(cheb, 60.0)
    pc.factor()
    rtol = 1e-10
    ksm = TACS.SegmentedKSM(p.mat, pc, p.node_off, m=20, restarts=1000, prec=prec,
                            deflation=deflation)
    ksm.setTolerances(rtol, 1e-30)
    assert ksm.solve(p.res, p.u) == 1
    its, flags, r0, rn, cycles = ksm.getStats()
    assert list(flags) == [1] * len(SIZES)
    assert np.all(rn <= 1.0001 * rtol * r0)
    # the cubes stop on their own: the 2^3 cube needs fewer steps than the 4^3
    assert its[SIZES.index(2)] < its[SIZES.index(4)]
    u = -p.u.getArray()
    for c, n in enumerate(SIZES):
        orp, ocols, oA, ores = _oracle_cube(n, p.traction)
        nn = orp.size - 1
        K = np.zeros((3 * nn, 3 * nn))
        Ab = oA.reshape(-1, 3, 3)
        for i in range(nn):
            for s in range(orp[i], orp[i + 1]):
                K[3 * i:3 * i + 3, 3 * ocols[s]:3 * ocols[s] + 3] = Ab[s]
        uc = u[p.dof(c)]
        assert np.linalg.norm(K @ uc + ores) <= 1.5 * rtol * np.linalg.norm(ores), c
        # dense solve on the free dofs (the identity BC rows would spoil its scaling)
        free = np.ones(ores.size, bool)
        for b in p.meshes[c].nodes_on_x(0):
            free[3 * b:3 * b + 3] = False
        uo = np.zeros_like(uc)
        uo[free] = -np.linalg.solve(K[np.ix_(free, free)], ores[free])
        assert not uc[~free].any()
        assert np.abs(uc - uo).max() <= 1e-6 * np.abs(uo).max(), c


def test_segmented_matches_union_and_refuses_coupled_rows():
    p = MultiCubeProblem(SIZES)
    p.assemble()
    pc = TACS.Pc(p.mat)
    pc.factor()
    seg = TACS.SegmentedKSM(p.mat, pc, p.node_off, m=30, prec=64, deflation=0)
    seg.setTolerances(1e-11, 1e-30)
    x1 = p.assembler.createVec()
    assert seg.solve(p.res, x1) == 1
    ksm = TACS.KSM(p.mat, pc, 60, 1000)
    ksm.setTolerances(1e-30, p.cube_atol(p.res, 1e-11))
    x2 = p.assembler.createVec()
    assert ksm.solve(p.res, x2) == 1
    a, b = x1.getArray(), x2.getArray()
    assert np.abs(a - b).max() <= 1e-8 * np.abs(b).max()
    # offsets that split a cube: its rows couple the two ranges
    bad = np.array(p.node_off).copy()
    bad[1] -= 1
    with pytest.raises(TACS.TACSError):
        TACS.SegmentedKSM(p.mat, pc, bad, m=20, prec=64, deflation=0)
    # offsets that do not cover the matrix
    with pytest.raises(TACS.TACSError):
        TACS.SegmentedKSM(p.mat, pc, np.array(p.node_off[:-1]), m=20, prec=64, deflation=0)
