"""GPU parity of the solid assembly / solve / adjoint path against the oracle
(CPU restatement of the reference) on the same inputs."""
import numpy as np
import pytest

import oracle as O
from tacs_amd import TACS, functions
from tacs_amd.problems import SolidProblem

pytestmark = pytest.mark.gpu

# FP64 parity bar: values within 1e-10 relative (normwise where entries
# cancel), sparsity pattern and BC rows bit-exact.
RTOL = 1e-10


def _oracle_system(p, t_elem=None, apply_bcs=True):
    m = p.mesh
    C = O.iso_C21(p.props.props["E"], p.props.props["nu"])
    rowp, cols = O.node_csr(m.nnodes, m.conn, m.npe)
    t = np.ones(m.nelem) if t_elem is None else t_elem
    A, res = O.assemble_solid(p.order, m.conn, m.X, np.zeros(3 * m.nnodes), t, C, rowp, cols)
    trac = np.tile(p.traction, (p.aux_elems.size, 1))
    O.add_tractions(p.order, m.conn, m.X, p.aux_elems, p.aux_faces, trac, res)
    bcv = np.full(p.bc_nodes.size, 7, np.int32)
    if apply_bcs:
        O.apply_bc_mat(3, p.bc_nodes, bcv, rowp, cols, A)
        O.apply_bc_vec(3, p.bc_nodes, bcv, res, u=np.zeros(3 * m.nnodes))
    return C, t, rowp, cols, A, res


def _dense(rowp, cols, A, b=3):
    n = rowp.size - 1
    K = np.zeros((b * n, b * n))
    Ab = A.reshape(-1, b, b)
    for i in range(n):
        for j in range(rowp[i], rowp[i + 1]):
            K[b * i:b * i + b, b * cols[j]:b * cols[j] + b] = Ab[j]
    return K


@pytest.mark.parametrize("n,order", [(4, 2), (2, 3)])
def test_jacobian_and_residual_match_oracle(n, order):
    p = SolidProblem(n, order=order)
    p.assemble()
    rowp, cols, A = p.mat.getArrays()
    C, t, orp, ocols, oA, ores = _oracle_system(p)
    assert np.array_equal(rowp, orp) and np.array_equal(cols, ocols)
    scale = np.abs(oA).max()
    assert np.abs(A.reshape(-1) - oA).max() <= RTOL * scale
    # BC rows are exact: zero rows with a unit diagonal
    bc = np.zeros(rowp.size - 1, bool)
    bc[p.bc_nodes] = True
    Ab = A.reshape(-1, 3, 3)
    for i in np.nonzero(bc)[0][:10]:
        for j in range(rowp[i], rowp[i + 1]):
            expect = np.eye(3) if cols[j] == i else np.zeros((3, 3))
            assert np.array_equal(Ab[j], expect)
    r = p.res.getArray()
    assert np.abs(r - ores).max() <= RTOL * np.abs(ores).max()


def test_spmv_matches_oracle():
    p = SolidProblem(6)
    p.assemble()
    rowp, cols, A = p.mat.getArrays()
    x = p.assembler.createVec()
    rng = np.random.default_rng(0)
    xv = rng.standard_normal(x.getSize())
    x.getArray()[:] = xv
    y = p.assembler.createVec()
    p.mat.mult(x, y)
    yo = O.bsr_mult(3, rowp, cols, A.reshape(-1), xv)
    assert np.abs(y.getArray() - yo).max() <= 1e-13 * np.abs(A).max() * np.abs(xv).sum() / xv.size * 30


def _ragged_pattern(nrows, seed, b=3):
    """Random sorted-unique block-CSR pattern with empty rows, rows up to 40
    blocks long and a row count that is not a multiple of 64."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 41, nrows)
    lens[rng.random(nrows) < 0.1] = 0
    rows = [np.sort(rng.choice(nrows, size=min(k, nrows), replace=False)) for k in lens]
    rowp = np.zeros(nrows + 1, np.int32)
    rowp[1:] = np.cumsum([r.size for r in rows])
    cols = np.concatenate(rows).astype(np.int32)
    A = rng.standard_normal(b * b * cols.size)
    return rowp, cols, A


@pytest.mark.parametrize("nrows", [1, 63, 65, 1000, 4161])
def test_sliced_ell_fill_variants_identical(nrows):
    """Every launch variant of the sliced ELL fill (one lane per row; one wave
    per slice staging 4 or 7 slots in LDS) writes bit-identical FP64,
    FP32 and bfloat16 copies and columns on a ragged 3x3 pattern, and leaves
    the entries beyond the last slot untouched."""
    import ctypes

    import torch

    from tacs_amd import lib

    b, b2 = 3, 9
    rowp, cols, A = _ragged_pattern(nrows, nrows + b, b)
    ns = (nrows + 63) // 64
    width = np.zeros(ns, np.int32)
    coff = np.zeros(ns + 1, np.int64)
    slots = lib().tacs_hip_sell_slots(nrows, rowp.ctypes.data, width.ctypes.data, coff.ctypes.data)
    dev = lambda a: torch.from_numpy(a).to("cuda")
    rp, cp, Ad, wd, od = dev(rowp), dev(cols), dev(A), dev(width), dev(coff)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    L = lib()
    nv = L.tacs_hip_sell_fill_nvariants(b)
    assert nv == 4 and L.tacs_hip_sell_fill_nvariants(6) == 1
    outs = []
    for v in range(nv):
        cs = torch.full((slots + 64,), -7, dtype=torch.int32, device="cuda")
        v64 = torch.full((b2 * slots + 64,), np.nan, dtype=torch.float64, device="cuda")
        v32 = torch.full((b2 * slots + 64,), np.nan, dtype=torch.float32, device="cuda")
        v16 = torch.full((b2 * slots + 64,), 12345, dtype=torch.int16, device="cuda")
        torch.cuda.synchronize()
        assert L.tacs_hip_sell_fill_variant(v, b, nrows, P(rp), P(cp), P(wd), P(od), P(Ad),
                                            P(v64), P(v32), P(v16), P(cs), None) == 0
        torch.cuda.synchronize()
        outs.append([t.cpu().numpy() for t in (v64, v32, v16, cs)])
    for o in outs:
        for a0, a1 in zip(outs[0], o):
            assert np.array_equal(a0.view(np.uint8), a1.view(np.uint8))
    v64h, v32h, v16h, csh = outs[0]
    assert not np.isnan(v64h[:b2 * slots]).any() and csh[:slots].min() >= 0
    assert np.isnan(v64h[b2 * slots:]).all() and (csh[slots:] == -7).all()
    assert (v16h[b2 * slots:] == 12345).all()
    # the copies agree with the FP64 values: FP32 rounding, bfloat16 of that
    assert np.array_equal(v32h[:b2 * slots], v64h[:b2 * slots].astype(np.float32))
    u = v32h[:b2 * slots].view(np.uint32).astype(np.uint64)
    bf = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    assert np.array_equal(v16h[:b2 * slots].view(np.uint16), bf)
    assert L.tacs_hip_sell_fill_variant(nv, b, nrows, P(rp), P(cp), P(wd), P(od), P(Ad), P(v64),
                                        None, None, None, None) != 0
    assert L.tacs_hip_sell_fill_variant(1, 6, nrows, P(rp), P(cp), P(wd), P(od), P(Ad), P(v64),
                                        None, None, None, None) != 0


@pytest.mark.parametrize("b,nrows", [(3, 1), (3, 63), (3, 1000), (3, 4161), (6, 1), (6, 1000),
                                     (6, 2113)])
def test_sliced_ell_spmv_matches_oracle(b, nrows):
    """The sliced ELL copy of a 3x3 or 6x6 BCSR matrix (slices of 64 rows,
    zero padding blocks) and its products with FP64 values and vectors, FP32
    values and FP64 vectors, FP32 values and vectors, against the oracle's
    block-CSR product (BCSRMat::mult) on a ragged pattern."""
    import ctypes

    import torch

    from tacs_amd import lib

    b2 = b * b
    rowp, cols, A = _ragged_pattern(nrows, nrows + b, b)
    ns = (nrows + 63) // 64
    width = np.zeros(ns, np.int32)
    coff = np.zeros(ns + 1, np.int64)
    slots = lib().tacs_hip_sell_slots(nrows, rowp.ctypes.data, width.ctypes.data, coff.ctypes.data)
    lens = np.diff(rowp)
    for s in range(ns):
        assert width[s] == lens[64 * s:64 * s + 64].max()
    assert coff[-1] == slots == 64 * width.sum()
    assert lib().tacs_hip_sell_slots(nrows, rowp.ctypes.data, None, None) == slots
    dev = lambda a: torch.from_numpy(a).to("cuda")
    rp, cp, Ad, wd, od = dev(rowp), dev(cols), dev(A), dev(width), dev(coff)
    cs = torch.full((max(slots, 1),), -1, dtype=torch.int32, device="cuda")
    v64 = torch.full((b2 * max(slots, 1),), np.nan, dtype=torch.float64, device="cuda")
    v32 = torch.full((b2 * max(slots, 1),), np.nan, dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(7)
    xh = rng.standard_normal(b * nrows)
    x64, y64 = dev(xh), torch.zeros(b * nrows, dtype=torch.float64, device="cuda")
    x32, y32 = x64.float(), torch.zeros(b * nrows, dtype=torch.float32, device="cuda")
    ym = torch.zeros_like(y64)
    torch.cuda.synchronize()
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    L = lib()
    assert L.tacs_hip_sell_fill(b, 64, nrows, P(rp), P(cp), P(wd), P(od), P(Ad), P(v64), P(cs),
                                None) == 0
    assert L.tacs_hip_sell_fill(b, 32, nrows, P(rp), P(cp), P(wd), P(od), P(Ad), P(v32), None,
                                None) == 0
    assert L.tacs_hip_sell_spmv(b, 64, 64, nrows, P(wd), P(od), P(cs), P(v64), P(x64), P(y64),
                                None) == 0
    assert L.tacs_hip_sell_spmv(b, 32, 64, nrows, P(wd), P(od), P(cs), P(v32), P(x64), P(ym),
                                None) == 0
    assert L.tacs_hip_sell_spmv(b, 32, 32, nrows, P(wd), P(od), P(cs), P(v32), P(x32), P(y32),
                                None) == 0
    # rejected before any launch: FP64 values with FP32 vectors, other block sizes
    assert L.tacs_hip_sell_spmv(b, 64, 32, nrows, P(wd), P(od), P(cs), P(v64), P(x64), P(y64),
                                None) != 0
    assert L.tacs_hip_sell_spmv(4, 64, 64, nrows, P(wd), P(od), P(cs), P(v64), P(x64), P(y64),
                                None) != 0
    assert L.tacs_hip_sell_fill(5, 64, nrows, P(rp), P(cp), P(wd), P(od), P(Ad), P(v64), None,
                                None) != 0
    torch.cuda.synchronize()
    # every slot written: columns in range, padding blocks zero in the row's own column
    csh = cs.cpu().numpy()[:slots]
    vh = v64.cpu().numpy()[:b2 * slots]
    assert csh.min() >= 0 and csh.max() < nrows and not np.isnan(vh).any()
    for s in range(ns):
        for t in range(width[s]):
            for l in range(64):
                r = 64 * s + l
                if r >= nrows or t >= lens[r]:
                    assert vh[b2 * coff[s] + 64 * b2 * t + l::64][:b2].tolist() == [0.0] * b2
                    assert csh[coff[s] + 64 * t + l] == (r if r < nrows else 0)
    yo = O.bsr_mult(b, rowp, cols, A, xh)
    scale = np.abs(A).max() * np.abs(xh).max() * 40 * b
    assert np.abs(y64.cpu().numpy() - yo).max() <= 1e-14 * scale
    A32 = A.astype(np.float32).astype(np.float64)
    yo32 = O.bsr_mult(b, rowp, cols, A32, xh)
    assert np.abs(ym.cpu().numpy() - yo32).max() <= 1e-14 * scale
    yo3232 = O.bsr_mult(b, rowp, cols, A32, xh.astype(np.float32).astype(np.float64))
    assert np.abs(y32.cpu().numpy() - yo3232).max() <= 1e-6 * scale


def test_products_follow_reassembly():
    """The matrix's product copies (FP32 values, sliced ELL) are refreshed
    when assembly changes the values: y = A x after a second assembly with
    other thicknesses matches the oracle's product with the new matrix, in
    the FP64, FP32-valued and FP32 products."""
    import torch

    from tacs_amd import lib

    p = SolidProblem(5, dv="element")
    rng = np.random.default_rng(3)
    x = p.assembler.createVec()
    xv = rng.standard_normal(x.getSize())
    x.getArray()[:] = xv
    y = p.assembler.createVec()
    for k in range(2):
        p.set_thickness(rng.uniform(0.2, 1.0, p.mesh.nelem))
        p.assemble()
        p.mat.mult(x, y)
        rowp, cols, A = p.mat.getArrays()
        yo = O.bsr_mult(3, rowp, cols, A.reshape(-1), xv)
        assert np.abs(y.getArray() - yo).max() <= 1e-13 * np.abs(yo).max()
    xd = torch.tensor(xv, device="cuda")
    ym = torch.zeros_like(xd)
    x32 = xd.float()
    y32 = torch.zeros(xv.size, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert lib().tacs_mat_mult_prec(p.mat.ptr, 32, 64, xd.data_ptr(), ym.data_ptr()) == 0
    assert lib().tacs_mat_mult_prec(p.mat.ptr, 32, 32, x32.data_ptr(), y32.data_ptr()) == 0
    assert lib().tacs_mat_mult_prec(p.mat.ptr, 64, 32, x32.data_ptr(), y32.data_ptr()) != 0
    torch.cuda.synchronize()
    A32 = A.reshape(-1).astype(np.float32).astype(np.float64)
    yo32 = O.bsr_mult(3, rowp, cols, A32, xv)
    assert np.abs(ym.cpu().numpy() - yo32).max() <= 1e-13 * np.abs(yo32).max()
    yo3232 = O.bsr_mult(3, rowp, cols, A32, xv.astype(np.float32).astype(np.float64))
    assert np.abs(y32.cpu().numpy() - yo3232).max() <= 1e-6 * np.abs(yo3232).max()


@pytest.mark.parametrize("order", [2, 3])
def test_pressure_aux_elements_match_oracle(order):
    """TACSPressure3D aux elements (several per element, every face, with a
    traction on the x = 1 face as well) in the assembled residual, against the
    oracle's restatement of TACSPressure3D::addResidual."""
    from tacs_amd import elements

    p = SolidProblem(3, order=order)
    m = p.mesh
    rng = np.random.default_rng(order)
    q = 12
    elems = rng.choice(m.nelem, q)
    faces = rng.integers(0, 6, q)
    pv = rng.standard_normal(q) * 3e7
    aux = TACS.AuxElements()
    for e in p.aux_elems:
        aux.addElement(int(e), elements.Traction3D(3, 1, p.basis, list(p.traction)))
    for e, f, v in zip(elems, faces, pv):
        aux.addElement(int(e), elements.Pressure3D(3, int(f), p.basis, float(v)))
    p.assembler.setAuxElements(aux)
    p.assemble()
    r = p.res.getArray().copy()
    C, t, rowp, cols, A, res = _oracle_system(p)
    pres = np.zeros(3 * m.nnodes)
    O.add_pressures(order, m.conn, m.X, elems, faces, pv, pres)
    bcv = np.full(p.bc_nodes.size, 7, np.int32)
    O.apply_bc_vec(3, p.bc_nodes, bcv, pres)
    expect = res + pres
    assert np.abs(r - expect).max() <= 1e-12 * np.abs(expect).max()
    assert np.abs(pres).max() > 0.1 * np.abs(expect).max()


def test_golden_cube_compliance_and_ks():
    """tests/integration_tests/test_elast_linhexa_element_3d.py:11-27: 3x3x3
    cube of side 10, rho=2570, E=70e9, nu=0.3, ys=350e6, x=0 clamped,
    nodal forces 1e8 on every dof; compliance and KS (ksWeight 10)."""
    p = SolidProblem(3, L=10.0, E=70e9, nu=0.3, rho=2570.0, ys=350e6, traction=(0, 0, 0))
    a = p.assembler
    a.zeroVariables()
    a.assembleJacobian(1.0, 0.0, 0.0, p.res, p.mat)
    f = a.createVec()
    f.getArray()[:] = 1e8
    a.applyBCs(f)
    p.res.axpy(-1.0, f)
    pc = TACS.Pc(p.mat)
    pc.factor()
    ksm = TACS.KSM(p.mat, pc, 100, 2)
    ksm.setTolerances(1e-12, 1e-30)
    assert ksm.solve(p.res, p.u) == 1
    p.u.scale(-1.0)
    a.setVariables(p.u)
    funcs = [functions.KSFailure(a, ksWeight=10.0), functions.Compliance(a)]
    vals = a.evalFunctions(funcs)
    assert abs(vals[1] - 178995399.11928475) <= 1e-9 * 178995399.11928475
    assert abs(vals[0] - 1.1074017881777185) <= 1e-9


@pytest.mark.parametrize("solver", ["fp64", "fp64_cgs", "mixed", "mixed_dr"])
def test_solve_functions_and_adjoint(solver):
    n = 6
    p = SolidProblem(n, dv="element")
    rng = np.random.default_rng(1)
    t = rng.uniform(0.1, 1.0, p.mesh.nelem)
    p.set_thickness(t)
    p.assemble()
    if solver == "mixed":
        flag, its, pc, ksm = p.solve_mixed(rtol=1e-10)
    elif solver == "mixed_dr":
        flag, its, pc, ksm = p.solve_mixed(m=40, rtol=1e-10, deflation=12)
    else:
        ortho = TACS.CLASSICAL_GRAM_SCHMIDT if solver == "fp64_cgs" else TACS.MODIFIED_GRAM_SCHMIDT
        flag, its, pc, ksm = p.solve_fp64(m=30, rtol=1e-10, ortho=ortho)
    assert flag == 1
    C, tt, rowp, cols, A, res = _oracle_system(p, t)
    K = _dense(rowp, cols, A)
    u = p.u.getArray().copy()
    # the solver's own stopping test, evaluated with the oracle matrix in FP64
    assert np.linalg.norm(K @ u + res) <= 1.05e-10 * np.linalg.norm(res)
    # reference solution: dense solve of the system with the Dirichlet dofs
    # removed (the unit BC rows make the full matrix badly scaled)
    free = np.ones(u.size, bool)
    for nd in p.bc_nodes:
        free[3 * nd:3 * nd + 3] = False
    u_ref = np.zeros_like(u)
    u_ref[free] = -np.linalg.solve(K[np.ix_(free, free)], res[free])
    assert np.all(u[~free] == 0.0)
    assert np.abs(u - u_ref).max() <= 1e-6 * np.abs(u_ref).max()
    a = p.assembler
    comp = functions.Compliance(a)
    ks = functions.KSFailure(a, ksWeight=100.0)
    vals = a.evalFunctions([comp, ks])
    # function kernels at the same u: FP64 parity
    c_gpu_u = O.compliance(2, p.mesh.conn, p.mesh.X, u, t, C)
    ks_gpu_u = O.ks_failure(2, p.mesh.conn, p.mesh.X, u, C, 270e6, 100.0)
    assert abs(vals[0] - c_gpu_u) <= RTOL * abs(c_gpu_u)
    assert abs(vals[1] - ks_gpu_u[0]) <= RTOL * abs(ks_gpu_u[0])
    # and at the reference solution, to the solve's accuracy
    c_ref = O.compliance(2, p.mesh.conn, p.mesh.X, u_ref, t, C)
    assert abs(vals[0] - c_ref) <= 1e-8 * abs(c_ref)
    # state sensitivities at the GPU solution vs the oracle at the same u
    dfdu = [a.createVec(), a.createVec()]
    a.addSVSens([comp, ks], dfdu)
    g0 = O.compliance_svsens(2, p.mesh.conn, p.mesh.X, u, t, C)
    kso = O.ks_failure(2, p.mesh.conn, p.mesh.X, u, C, 270e6, 100.0)
    g1 = O.ks_svsens(2, p.mesh.conn, p.mesh.X, u, C, 270e6, 100.0, kso[1], kso[2])
    assert np.abs(dfdu[0].getArray() - g0).max() <= RTOL * np.abs(g0).max()
    assert np.abs(dfdu[1].getArray() - g1).max() <= 1e-9 * np.abs(g1).max()
    # adjoint of the compliance: K psi = dC/du, total dC/dt = dC/dt - psi^T dR/dt
    a.applyBCs(dfdu[0])
    psi = a.createVec()
    ksm.solve(dfdu[0], psi)
    dfdx = a.createDesignVec()
    a.addDVSens([comp], [dfdx], 1.0)
    a.addAdjointResProducts([psi], [dfdx], -1.0)
    grad = dfdx.getArray().copy()
    psi_v = psi.getArray().copy()
    g_ref = O.compliance_dvsens(2, p.mesh.conn, p.mesh.X, u, C) - \
        O.adj_res_product(2, p.mesh.conn, p.mesh.X, u, psi_v, C)
    assert np.abs(grad - g_ref).max() <= 1e-9 * np.abs(g_ref).max()
    # closed form: dC/dt_e = -u_e^T K0_e u_e  (< 0)
    closed = -O.compliance_dvsens(2, p.mesh.conn, p.mesh.X, u_ref, C)
    assert np.abs(grad - closed).max() <= 1e-6 * np.abs(closed).max()


def test_solver_residual_monitors(tmp_path):
    """KSM.setMonitor (KSMPrintFile): GMRES prints |r| at every restart and
    every freq-th step (GMRES::solve src/bpmat/KSM.cpp:815, :893); the
    recycling solver prints the true residual at each cycle start."""
    import re

    p = SolidProblem(5)
    p.assemble()
    pc = TACS.Pc(p.mat)
    pc.factor()
    ksm = TACS.KSM(p.mat, pc, 20, 50)
    ksm.setTolerances(1e-10, 1e-30)
    f1 = tmp_path / "gmres.txt"
    ksm.setMonitor(descript="GMRES", freq=5, filename=f1)
    x = p.assembler.createVec()
    assert ksm.solve(p.res, x) == 1
    its = ksm.getIterCount()
    del ksm  # closes the file
    lines = f1.read_text().splitlines()
    pat = re.compile(r"GMRES\[\s*(\d+)\]:\s+(\S+)$")
    recs = [pat.match(l) for l in lines]
    assert all(recs)
    steps = [int(r.group(1)) for r in recs]
    vals = [float(r.group(2)) for r in recs]
    assert steps[0] == 0 and abs(vals[0] - np.linalg.norm(p.res.getArray())) <= 1e-12 * vals[0]
    assert all(k == 0 or k % 5 == 0 for k in steps)
    nrestart = (its + 19) // 20
    assert steps.count(0) == nrestart and len(steps) == nrestart + sum(
        min(20, its - 20 * c) // 5 for c in range(nrestart))
    mk = TACS.MixedKSM(p.mat, pc, 30, 100, deflation=8)
    mk.setTolerances(1e-10, 1e-30)
    f2 = tmp_path / "rec.txt"
    mk.setMonitor(descript="rec", freq=1, filename=f2)
    assert mk.solve(p.res, p.assembler.createVec()) == 1
    cycles = mk.getOuterCount()
    del mk
    lines = f2.read_text().splitlines()
    assert len(lines) == cycles + 1 and all(l.startswith("rec[  0]: ") for l in lines)
    r = [float(l.split()[-1]) for l in lines]
    assert r[-1] <= 1e-10 * r[0] < r[-2]
