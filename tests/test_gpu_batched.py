"""GPU parity of the batched multi right-hand-side GMRES (one SpMM per step for
all load cases): every solution against the oracle's dense solve of the
oracle-assembled stiffness matrix, and iteration counts against one
single right-hand-side GMRES per load case (classical Gram-Schmidt twice,
the recurrence of GMRES::solve src/bpmat/KSM.cpp:785)."""
import numpy as np
import pytest

from tacs_amd import TACS
from tacs_amd.problems import SolidProblem

from test_gpu_solid import _dense, _oracle_system

pytestmark = pytest.mark.gpu


def _load_cases(p, rng):
    """Five load cases: the traction load, a random load, point loads, a zero
    load and a scaled copy of the random load; zero on the BC dofs."""
    a = p.assembler
    n = 3 * p.mesh.nnodes
    cases = [-p.res.getArray().copy(), rng.standard_normal(n), np.zeros(n), np.zeros(n)]
    pts = rng.choice(p.mesh.nnodes, 4, replace=False)
    for k in pts:
        cases[2][3 * k:3 * k + 3] = 1e5 * rng.standard_normal(3)
    cases.append(1e3 * cases[1])
    vecs = []
    for c in cases:
        v = a.createVec()
        v.getArray()[:] = c
        a.applyBCs(v)
        vecs.append(v)
    return vecs


@pytest.fixture(scope="module")
def system():
    p = SolidProblem(3, order=3)
    p.assemble()
    pc = TACS.Pc(p.mat)
    pc.factor()
    _, _, rowp, cols, A, _ = _oracle_system(p)
    K = _dense(rowp, cols, A)
    return p, pc, K


@pytest.mark.parametrize("prec", [64, 32])
def test_batched_solutions_match_dense_solve(system, prec):
    p, pc, K = system
    a = p.assembler
    rng = np.random.default_rng(11)
    bs = _load_cases(p, rng)
    xs = [a.createVec() for _ in bs]
    ksm = TACS.BatchedKSM(p.mat, pc, m=25, restarts=200, nrhs=8, prec=prec)
    ksm.setTolerances(1e-12, 1e-30)
    assert ksm.solve(bs, xs) == 1
    it, fl, r0, rn, cyc = ksm.getStats()
    assert np.all(fl == 1)
    for j, (b, x) in enumerate(zip(bs, xs)):
        bv, xv = b.getArray(), x.getArray()
        if not bv.any():
            assert it[j] == 0 and not xv.any()
            continue
        xo = np.linalg.solve(K, bv)
        assert np.abs(xv - xo).max() <= 1e-8 * np.abs(xo).max(), j
        assert np.linalg.norm(bv - K @ xv) <= 2e-12 * np.linalg.norm(bv), j
    # the scaled copy converges in the same number of steps, to the scaled solution
    assert it[4] == it[1]
    x1, x4 = xs[1].getArray(), xs[4].getArray()
    assert np.abs(x4 - 1e3 * x1).max() <= 1e-9 * np.abs(x4).max()


def test_batched_iteration_counts_match_single_gmres(system):
    """Without restarts (m above the step count) the batched recurrences take
    the steps of single GMRES solves; restarted runs on this small mesh
    stagnate and their counts follow the rounding of the products."""
    p, pc, K = system
    a = p.assembler
    rng = np.random.default_rng(12)
    bs = _load_cases(p, rng)[:3]
    xs = [a.createVec() for _ in bs]
    m = 250
    ksm = TACS.BatchedKSM(p.mat, pc, m=m, restarts=200, nrhs=3, prec=64)
    ksm.setTolerances(1e-10, 1e-30)
    assert ksm.solve(bs, xs) == 1
    it = ksm.getStats()[0]
    single = TACS.KSM(p.mat, pc, m, 200)
    single.setTolerances(1e-10, 1e-30)
    single.setOrthoType(TACS.CLASSICAL_GRAM_SCHMIDT)
    for j, b in enumerate(bs):
        x = a.createVec()
        assert single.solve(b, x) == 1
        assert abs(single.getIterCount() - it[j]) <= 1, (j, single.getIterCount(), it[j])
        xv, sv = xs[j].getArray(), x.getArray()
        assert np.abs(xv - sv).max() <= 1e-8 * np.abs(sv).max()


def test_batched_initial_guess_and_fewer_rhs(system):
    """zero_guess=0 from half the solution, with fewer right-hand sides than
    the solver was created for, then a single right-hand side."""
    p, pc, K = system
    a = p.assembler
    rng = np.random.default_rng(13)
    bs = _load_cases(p, rng)[:2]
    ksm = TACS.BatchedKSM(p.mat, pc, m=30, restarts=100, nrhs=6, prec=64)
    ksm.setTolerances(1e-11, 1e-30)
    xs = []
    for b in bs:
        x = a.createVec()
        x.getArray()[:] = 0.5 * np.linalg.solve(K, b.getArray())
        xs.append(x)
    assert ksm.solve(bs, xs, zero_guess=0) == 1
    it, fl, r0, rn, cyc = ksm.getStats()
    for j, b in enumerate(bs):
        xo = np.linalg.solve(K, b.getArray())
        assert np.abs(xs[j].getArray() - xo).max() <= 1e-8 * np.abs(xo).max()
        assert rn[j] <= 1e-11 * r0[j]
    x1 = a.createVec()
    assert ksm.solve(bs[1:], [x1]) == 1
    xo = np.linalg.solve(K, bs[1].getArray())
    assert np.abs(x1.getArray() - xo).max() <= 1e-8 * np.abs(xo).max()


def test_batched_rejects_too_many_rhs(system):
    p, pc, K = system
    a = p.assembler
    ksm = TACS.BatchedKSM(p.mat, pc, m=10, restarts=10, nrhs=2)
    bs = [a.createVec() for _ in range(3)]
    with pytest.raises(RuntimeError):
        ksm.solve(bs, [a.createVec() for _ in range(3)])


@pytest.mark.parametrize("prec", [32, 64])
def test_batched_with_recycled_space(system, prec):
    """The recycled space of a deflated mixed GMRES solve of the first load
    case deflates the batch: the same solutions as the dense solve."""
    p, pc, K = system
    a = p.assembler
    rng = np.random.default_rng(14)
    bs = _load_cases(p, rng)
    mk = TACS.MixedKSM(p.mat, pc, 30, 200, deflation=8)
    mk.setTolerances(1e-10, 1e-30)
    x0 = a.createVec()
    assert mk.solve(bs[0], x0) == 1
    plain = TACS.BatchedKSM(p.mat, pc, m=20, restarts=200, nrhs=8, prec=prec)
    plain.setTolerances(1e-12, 1e-30)
    xs = [a.createVec() for _ in bs]
    assert plain.solve(bs, xs) == 1
    it_plain = plain.getStats()[0]
    ksm = TACS.BatchedKSM(p.mat, pc, m=20, restarts=200, nrhs=8, prec=prec)
    ksm.setTolerances(1e-12, 1e-30)
    assert ksm.setRecycleSpace(mk) == 8
    xs = [a.createVec() for _ in bs]
    assert ksm.solve(bs, xs) == 1
    it, fl, r0, rn, cyc = ksm.getStats()
    assert np.all(fl == 1)
    for j, (b, x) in enumerate(zip(bs, xs)):
        bv, xv = b.getArray(), x.getArray()
        if not bv.any():
            assert not xv.any()
            continue
        xo = np.linalg.solve(K, bv)
        assert np.abs(xv - xo).max() <= 1e-8 * np.abs(xo).max(), j
        assert np.linalg.norm(bv - K @ xv) <= 2e-12 * np.linalg.norm(bv), j
    # the load case the space came from converges much faster; for the others
    # the gain depends on the spectrum (none is asserted on this small mesh)
    assert it[0] < it_plain[0] // 2, (it, it_plain)
    assert it[4] == it[1]
    # a change of the matrix values drops the space
    p.assemble()
    pc.factor()
    x1 = a.createVec()
    assert ksm.solve(bs[1:2], [x1]) == 1
    it_after = ksm.getStats()[0][0]
    assert plain.solve(bs[1:2], [a.createVec()]) == 1
    assert it_after == plain.getStats()[0][0]


@pytest.mark.parametrize("prec,delayed", [(32, True), (64, True), (32, False)])
def test_batched_gcrodr_matches_dense_and_single_recycling(system, prec, delayed):
    """Per right-hand side recycling (batched GCRO-DR): solutions against the
    dense solve; with the FP32-valued operator each system takes the steps of
    one MixedKSM with the same deflation solving it alone (the same cycle,
    batched)."""
    p, pc, K = system
    a = p.assembler
    rng = np.random.default_rng(15)
    bs = _load_cases(p, rng)
    ksm = TACS.BatchedKSM(p.mat, pc, m=20, restarts=300, nrhs=8, prec=prec, deflation=6)
    ksm.setDelayedReorthogonalization(delayed)
    ksm.setTolerances(1e-10, 1e-30)
    xs = [a.createVec() for _ in bs]
    assert ksm.solve(bs, xs) == 1
    it, fl, r0, rn, cyc = ksm.getStats()
    assert np.all(fl == 1)
    for j, (b, x) in enumerate(zip(bs, xs)):
        bv, xv = b.getArray(), x.getArray()
        if not bv.any():
            assert it[j] == 0 and not xv.any()
            continue
        xo = np.linalg.solve(K, bv)
        assert np.abs(xv - xo).max() <= 1e-7 * np.abs(xo).max(), j
        assert np.linalg.norm(bv - K @ xv) <= 2e-10 * np.linalg.norm(bv), j
    assert it[4] == it[1]
    if prec == 32:
        for j in (0, 1, 2):
            mk = TACS.MixedKSM(p.mat, pc, 20, 300, deflation=6)
            mk.setTolerances(1e-10, 1e-30)
            assert mk.solve(bs[j], a.createVec()) == 1
            single = mk.getIterCount()
            assert abs(single - it[j]) <= max(3, 0.05 * single), (j, single, it[j])
    # the spaces carry over to the next solve with the same matrix
    xs2 = [a.createVec() for _ in bs]
    assert ksm.solve(bs, xs2) == 1
    it2 = ksm.getStats()[0]
    assert it2[1] < it[1], (it, it2)


def test_batched_gcrodr_seeded_from_shared_space(system):
    p, pc, K = system
    a = p.assembler
    rng = np.random.default_rng(16)
    bs = _load_cases(p, rng)[:3]
    mk = TACS.MixedKSM(p.mat, pc, 30, 200, deflation=8)
    mk.setTolerances(1e-10, 1e-30)
    assert mk.solve(bs[0], a.createVec()) == 1
    ksm = TACS.BatchedKSM(p.mat, pc, m=20, restarts=300, nrhs=3, prec=32, deflation=8)
    ksm.setTolerances(1e-10, 1e-30)
    assert ksm.setRecycleSpace(mk) == 8
    xs = [a.createVec() for _ in bs]
    assert ksm.solve(bs, xs) == 1
    it = ksm.getStats()[0]
    fresh = TACS.BatchedKSM(p.mat, pc, m=20, restarts=300, nrhs=3, prec=32, deflation=8)
    fresh.setTolerances(1e-10, 1e-30)
    assert fresh.solve(bs, [a.createVec() for _ in bs]) == 1
    assert it[0] < fresh.getStats()[0][0]
    for b, x in zip(bs, xs):
        xo = np.linalg.solve(K, b.getArray())
        assert np.abs(x.getArray() - xo).max() <= 1e-7 * np.abs(xo).max()


@pytest.mark.parametrize("R", [1, 3, 4, 12, 16, 33, 64])
def test_spmm_matches_oracle(system, R):
    """The multi right-hand-side SpMM (FP64 values and FP32-rounded values)
    against the oracle's block-CSR product, one right-hand side at a time."""
    import torch

    import oracle as O
    from tacs_amd import lib

    p, pc, K = system
    rowp, cols, A = p.mat.getArrays()
    rp, cp, ap = p.mat.getDeviceArrays()
    b, nr, nnz = p.mat.getInfo()
    rng = np.random.default_rng(R)
    Xh = rng.standard_normal((3 * nr, R))
    X = torch.tensor(Xh.reshape(-1), dtype=torch.float64, device="cuda")
    Y = torch.zeros_like(X)
    A32 = torch.tensor(A.reshape(-1), dtype=torch.float64, device="cuda").float()
    torch.cuda.synchronize()
    assert lib().tacs_hip_bsr3_spmm(nr, R, rp, cp, ap, X.data_ptr(), Y.data_ptr(), None) == 0
    torch.cuda.synchronize()
    Yh = Y.cpu().numpy().reshape(3 * nr, R)
    assert lib().tacs_hip_bsr3_spmm_mixed(nr, R, rp, cp, A32.data_ptr(), X.data_ptr(),
                                          Y.data_ptr(), None) == 0
    torch.cuda.synchronize()
    Y32 = Y.cpu().numpy().reshape(3 * nr, R)
    A32h = A.reshape(-1).astype(np.float32).astype(np.float64)
    for j in range(R):
        yo = O.bsr_mult(3, rowp, cols, A.reshape(-1), Xh[:, j])
        assert np.abs(Yh[:, j] - yo).max() <= 1e-13 * np.abs(yo).max()
        yo32 = O.bsr_mult(3, rowp, cols, A32h, Xh[:, j])
        assert np.abs(Y32[:, j] - yo32).max() <= 1e-13 * np.abs(yo32).max()


@pytest.mark.parametrize("R,p,q", [(1, 9, 5), (3, 5, 7), (16, 61, 20), (64, 40, 32)])
def test_batched_products_and_basis_change(R, p, q):
    """Per right-hand side products X_j^T Y_j and basis changes W_j P_j of
    interleaved block vectors against numpy, incl. the chunked launch
    shapes (large R x p x q)."""
    import torch

    from tacs_amd import lib

    n = 3 * 517
    rng = np.random.default_rng(R + p + q)
    X = rng.standard_normal((p, n, R))
    Y = rng.standard_normal((q, n, R))
    Xd = torch.tensor(X.reshape(-1), device="cuda")
    Yd = torch.tensor(Y.reshape(-1), device="cuda")
    ldo, sR = p + 3, (p + 3) * q + 5
    out = torch.zeros(R * sR, dtype=torch.float64, device="cuda")
    scratch = torch.empty(lib().tacs_hip_btsgemm_scratch_bytes(R) // 8 + 1, dtype=torch.float64,
                          device="cuda")
    torch.cuda.synchronize()
    assert lib().tacs_hip_btsgemm(n, R, p, Xd.data_ptr(), q, Yd.data_ptr(), n * R,
                                  scratch.data_ptr(), out.data_ptr(), sR, ldo, None) == 0
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    for j in range(R):
        ref = X[:, :, j] @ Y[:, :, j].T  # p x q
        got = np.array([[o[j * sR + b * ldo + v] for b in range(q)] for v in range(p)])
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # basis change: out_c(:, j) = sum_v X_v(:, j) P_j[v + c*p], c < min(q, 32)
    nout = min(q, 32)
    P = rng.standard_normal((R, nout, p))  # P[j][c][v]
    sP = p * nout + 7
    Pd = torch.zeros(R * sP, dtype=torch.float64, device="cuda")
    Ph = np.zeros(R * sP)
    for j in range(R):
        Ph[j * sP:j * sP + p * nout] = P[j].reshape(-1)  # v fastest within column c
    Pd.copy_(torch.tensor(Ph))
    Od = torch.full((nout * n * R,), 2.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert lib().tacs_hip_bbasis_combine(n, R, p, Xd.data_ptr(), n * R, Pd.data_ptr(), sP, p, nout,
                                         Od.data_ptr(), n * R, 1, None) == 0
    torch.cuda.synchronize()
    Oh = Od.cpu().numpy().reshape(nout, n, R)
    for j in range(R):
        ref = 2.0 + np.einsum("vn,cv->cn", X[:, :, j], P[j])
        assert np.abs(Oh[:, :, j] - ref).max() <= 1e-12 * np.abs(ref).max()
