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
