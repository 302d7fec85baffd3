"""GMRES with each restart cycle captured as one HIP graph against the
step-by-step solver: the same step counts and the same solutions (the same
kernels on the same data), over several replays of the graph with eager work
and synchronisations between them, for modified and classical Gram-Schmidt,
FP64 and flexible; solutions against the oracle's dense solve."""
import numpy as np
import pytest

from tacs_amd import TACS, lib
from tacs_amd.problems import SolidProblem

from test_gpu_solid import _dense, _oracle_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ortho", [TACS.MODIFIED_GRAM_SCHMIDT, TACS.CLASSICAL_GRAM_SCHMIDT])
@pytest.mark.parametrize("flexible", [0, 1])
def test_cycle_graph_matches_step_by_step(ortho, flexible):
    p = SolidProblem(5)
    p.assemble()
    a = p.assembler
    pc = TACS.Pc(p.mat)
    pc.factor()
    rng = np.random.default_rng(3)
    rhs = []
    for k in range(3):
        v = a.createVec()
        v.getArray()[:] = rng.standard_normal(v.getSize()) if k else p.res.getArray()
        a.applyBCs(v)
        rhs.append(v)
    res = {}
    for graph in (False, True):
        ksm = TACS.KSM(p.mat, pc, 20, 200, isFlexible=flexible)
        ksm.setTolerances(1e-11, 1e-30)
        if ortho != TACS.MODIFIED_GRAM_SCHMIDT:
            ksm.setOrthoType(ortho)
        ksm.setCycleGraph(graph)
        out = []
        for b in rhs:  # several solves: the graph is replayed many times
            x = a.createVec()
            assert ksm.solve(b, x) == 1
            out.append((ksm.getIterCount(), x.getArray().copy()))
            # eager work and a synchronisation between replays
            y = a.createVec()
            p.mat.mult(x, y)
            lib().tacs_device_synchronize()
        res[graph] = out
    for (it0, x0), (it1, x1) in zip(res[False], res[True]):
        assert it0 == it1
        assert np.array_equal(x0, x1)
    # and the solution is the oracle's
    # (dense solve on the free dofs: unit BC rows beside 1e11 entries would
    # spoil its accuracy)
    _, _, orp, ocols, oA, ores = _oracle_system(p)
    K = _dense(orp, ocols, oA)
    free = np.ones(K.shape[0], bool)
    for bn in p.bc_nodes:
        free[3 * bn:3 * bn + 3] = False
    u = np.zeros(K.shape[0])
    u[free] = np.linalg.solve(K[np.ix_(free, free)], rhs[0].getArray()[free])
    x = res[True][0][1]
    assert not x[~free].any()
    assert np.abs(x - u).max() <= 1e-8 * np.abs(u).max()


def test_cycle_graph_follows_reassembly():
    """The graph reads the matrix's product layout: after a reassembly with
    other values the captured cycle solves the new system."""
    p = SolidProblem(4, dv="element")
    a = p.assembler
    p.assemble()
    pc = TACS.Pc(p.mat)
    pc.factor()
    ksm = TACS.KSM(p.mat, pc, 20, 200)
    ksm.setTolerances(1e-11, 1e-30)
    ksm.setCycleGraph(True)
    x = a.createVec()
    assert ksm.solve(p.res, x) == 1
    p.set_thickness(np.random.default_rng(1).uniform(0.5, 2.0, p.mesh.nelem))
    p.assemble()
    pc.factor()
    assert ksm.solve(p.res, x) == 1
    r = a.createVec()
    p.mat.mult(x, r)
    rv = r.getArray() - p.res.getArray()
    assert np.linalg.norm(rv) <= 1e-10 * np.linalg.norm(p.res.getArray())
