"""SIMP compliance gradient through the torch.autograd.Function built on the
adjoint kernels: checked against the oracle's closed form and against central
finite differences (h = 1e-6) of the GPU forward solve."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_simp_compliance_gradient():
    import torch
    from tacs_amd.autograd import SolidComplianceProblem, simp_compliance
    from tacs_amd.meshes import simp_thickness, simp_dthickness
    from tacs_amd.problems import SolidProblem

    n = 6
    prob = SolidProblem(n, E=1.0, nu=0.3, traction=(0.0, 0.0, 1.0), dv="element",
                        traction_patch=2)
    rng = np.random.default_rng(11)
    x0 = rng.uniform(0.1, 1.0, prob.mesh.nelem)
    sp = SolidComplianceProblem(prob, p=3.0, E0=1.0, Emin=1e-9, rtol=1e-12, m=60)
    x = torch.tensor(x0, dtype=torch.float64, device="cuda", requires_grad=True)
    c = simp_compliance(x, sp)
    c.backward()
    g = x.grad.detach().cpu().numpy()

    # oracle: dense solve at t(x), closed form dC/dt_e = -u_e^T K0_e u_e
    m = prob.mesh
    C = O.iso_C21(1.0, 0.3)
    t = simp_thickness(x0)
    rowp, cols = O.node_csr(m.nnodes, m.conn, 8)
    A, res = O.assemble_solid(2, m.conn, m.X, np.zeros(3 * m.nnodes), t, C, rowp, cols)
    O.add_tractions(2, m.conn, m.X, prob.aux_elems, prob.aux_faces,
                    np.tile([0.0, 0.0, 1.0], (prob.aux_elems.size, 1)), res)
    bcv = np.full(prob.bc_nodes.size, 7, np.int32)
    O.apply_bc_mat(3, prob.bc_nodes, bcv, rowp, cols, A)
    O.apply_bc_vec(3, prob.bc_nodes, bcv, res)
    nn = rowp.size - 1
    K = np.zeros((3 * nn, 3 * nn))
    Ab = A.reshape(-1, 3, 3)
    for i in range(nn):
        for j in range(rowp[i], rowp[i + 1]):
            K[3 * i:3 * i + 3, 3 * cols[j]:3 * cols[j] + 3] = Ab[j]
    free = np.ones(3 * nn, bool)
    for nd in prob.bc_nodes:
        free[3 * nd:3 * nd + 3] = False
    u = np.zeros(3 * nn)
    u[free] = -np.linalg.solve(K[np.ix_(free, free)], res[free])
    c_ref = O.compliance(2, m.conn, m.X, u, t, C)
    assert abs(c.item() - c_ref) <= 1e-9 * abs(c_ref)
    g_ref = -O.compliance_dvsens(2, m.conn, m.X, u, C) * simp_dthickness(x0)
    assert np.abs(g - g_ref).max() <= 1e-7 * np.abs(g_ref).max()

    # central finite differences of the GPU forward, h = 1e-6
    idx = rng.choice(m.nelem, 8, replace=False)
    h = 1e-6
    for e in idx:
        vals = []
        for s in (1.0, -1.0):
            xp = x0.copy()
            xp[e] += s * h
            vals.append(sp.forward(torch.tensor(xp, dtype=torch.float64, device="cuda")))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g[e]) <= 1e-4 * abs(g[e]) + 1e-8 * abs(c_ref)


def test_backward_uses_its_own_forward_state():
    """forward(x1), forward(x2), then the backward of the first node: the
    gradient is the one at x1 (the problem re-solves at x1), equal to a
    fresh forward/backward at x1 and to central differences there."""
    import torch
    from tacs_amd.autograd import SolidComplianceProblem, simp_compliance
    from tacs_amd.problems import SolidProblem

    prob = SolidProblem(4, E=1.0, nu=0.3, traction=(0.0, 0.0, 1.0), dv="element",
                        traction_patch=2)
    rng = np.random.default_rng(5)
    x1 = rng.uniform(0.1, 1.0, prob.mesh.nelem)
    x2 = rng.uniform(0.1, 1.0, prob.mesh.nelem)
    sp = SolidComplianceProblem(prob, p=3.0, E0=1.0, Emin=1e-9, rtol=1e-12, m=60, mixed=False,
                                restarts=500)
    xa = torch.tensor(x1, dtype=torch.float64, device="cuda", requires_grad=True)
    xb = torch.tensor(x2, dtype=torch.float64, device="cuda", requires_grad=True)
    ca = simp_compliance(xa, sp)
    cb = simp_compliance(xb, sp)
    ca.backward()
    ga = xa.grad.detach().cpu().numpy()
    cb.backward()
    gb = xb.grad.detach().cpu().numpy()
    # fresh passes at each point
    for x0, g in ((x1, ga), (x2, gb)):
        xf = torch.tensor(x0, dtype=torch.float64, device="cuda", requires_grad=True)
        simp_compliance(xf, sp).backward()
        gf = xf.grad.detach().cpu().numpy()
        assert np.abs(g - gf).max() <= 1e-9 * np.abs(gf).max()
    assert np.abs(ga - gb).max() > 1e-3 * np.abs(ga).max()  # the points differ
    e = int(np.argmax(np.abs(ga)))
    h = 1e-6
    vals = []
    for s in (1.0, -1.0):
        xp = x1.copy()
        xp[e] += s * h
        vals.append(sp.forward(torch.tensor(xp, dtype=torch.float64, device="cuda")))
    fd = (vals[0] - vals[1]) / (2 * h)
    assert abs(fd - ga[e]) <= 1e-4 * abs(ga[e])
