"""Chebyshev polynomial preconditioner of the recycling and batched solvers:
the Lanczos estimate of lmax(D^{-1} A) against the dense eigenvalues
of the oracle's matrix, and solves with it against the oracle's dense
solution (solid cube, several load cases, a shell plate)."""
import numpy as np
import pytest

from tacs_amd import TACS
from tacs_amd.problems import ShellProblem, SolidProblem

pytestmark = pytest.mark.gpu


def _dense(p):
    rowp, cols, A = p.mat.getArrays()
    b = A.shape[1]
    n = rowp.size - 1
    K = np.zeros((b * n, b * n))
    for i in range(n):
        for s in range(rowp[i], rowp[i + 1]):
            K[b * i:b * i + b, b * cols[s]:b * cols[s] + b] = A[s]
    return K, rowp, cols, A


def _dinv_a_lmax(K, b):
    n = K.shape[0] // b
    D = np.zeros_like(K)
    for i in range(n):
        sl = slice(b * i, b * i + b)
        D[sl, sl] = np.linalg.inv(K[sl, sl])
    return np.max(np.abs(np.linalg.eigvals(D @ K)))


@pytest.mark.parametrize("order,n", [(2, 5), (3, 3)])
def test_lanczos_estimate(order, n):
    p = SolidProblem(n, order=order, dv="element")
    p.set_thickness(np.random.default_rng(2).uniform(0.5, 1.5, p.mesh.nelem))
    p.assemble()
    pc = TACS.Pc(p.mat)
    assert pc.getChebyshev() == (1, 20.0, 0.0)
    pc.setChebyshev(4, 30.0)
    pc.factor()
    deg, ratio, lmax = pc.getChebyshev()
    assert (deg, ratio) == (4, 30.0)
    K, *_ = _dense(p)
    true = _dinv_a_lmax(K, 3)
    # 1.05 x the largest Ritz value (from below): an upper bound once the
    # Ritz value is within 5 %, and at most 5 % above the true value
    assert true <= lmax <= 1.05 * true * (1 + 1e-6), (lmax, true)
    with pytest.raises(TACS.TACSError):
        pc.setChebyshev(0)
    with pytest.raises(TACS.TACSError):
        pc.setChebyshev(4, 1.0)


@pytest.mark.parametrize("deg,ratio,values", [(2, 20.0, 32), (6, 60.0, 32), (10, 300.0, 32),
                                              (10, 300.0, 16)])
def test_recycling_solve_with_chebyshev_matches_dense(deg, ratio, values):
    p = SolidProblem(5, dv="element")
    p.set_thickness(np.random.default_rng(3).uniform(0.5, 1.5, p.mesh.nelem))
    p.assemble()
    pc = TACS.Pc(p.mat)
    pc.setChebyshev(deg, ratio, values)
    pc.factor()
    ksm = TACS.MixedKSM(p.mat, pc, 30, 1000, deflation=8)
    ksm.setTolerances(1e-11, 1e-30)
    x = p.assembler.createVec()
    assert ksm.solve(p.res, x) == 1
    its1 = ksm.getIterCount()
    assert ksm.solve(p.res, x) == 1  # again, with the recycled space
    K, *_ = _dense(p)
    r = p.res.getArray()
    free = np.ones(r.size, bool)
    for bn in p.bc_nodes:
        free[3 * bn:3 * bn + 3] = False
    uo = np.zeros_like(r)
    uo[free] = np.linalg.solve(K[np.ix_(free, free)], r[free])
    u = x.getArray()
    assert np.linalg.norm(K @ u - r) <= 2e-11 * np.linalg.norm(r)
    assert np.abs(u - uo).max() <= 1e-8 * np.abs(uo).max()
    # fewer outer steps than block-Jacobi alone
    pc1 = TACS.Pc(p.mat)
    pc1.factor()
    k1 = TACS.MixedKSM(p.mat, pc1, 30, 1000, deflation=8)
    k1.setTolerances(1e-11, 1e-30)
    assert k1.solve(p.res, x) == 1
    assert its1 < k1.getIterCount()


@pytest.mark.parametrize("R,values,deg", [(4, 32, 5), (16, 16, 5), (16, 32, 5), (32, 16, 5),
                                          (16, 16, 3), (16, 16, 4), (16, 16, 6), (16, 16, 7)])
def test_batched_solve_with_chebyshev_matches_dense(R, values, deg):
    """values 16 with R = 16, 32: the polynomial's products read an FP32 copy
    of the iterate, each step fused into its product (bsr3_spmm_quad_cheb);
    degrees 3-7 take every order of its in-place, transition and three-term
    steps (the last iterate must land in the output for even and odd
    degrees)."""
    p = SolidProblem(3, order=3)
    p.assemble()
    pc = TACS.Pc(p.mat)
    pc.setChebyshev(deg, 50.0, values=values)
    pc.factor()
    rng = np.random.default_rng(5)
    a = p.assembler
    bs = [a.createVec() for _ in range(R)]
    xs = [a.createVec() for _ in range(R)]
    K, *_ = _dense(p)
    free = np.ones(K.shape[0], bool)
    for bn in p.bc_nodes:
        free[3 * bn:3 * bn + 3] = False
    for b in bs:
        v = rng.standard_normal(K.shape[0])
        v[~free] = 0.0
        b.getArray()[:] = v
    ksm = TACS.BatchedKSM(p.mat, pc, m=30, restarts=1000, nrhs=R, prec=32, deflation=8)
    ksm.setTolerances(1e-10, 1e-30)
    assert ksm.solve(bs, xs) == 1
    for b, x in zip(bs, xs):
        r = b.getArray()
        uo = np.zeros_like(r)
        uo[free] = np.linalg.solve(K[np.ix_(free, free)], r[free])
        assert np.abs(x.getArray() - uo).max() <= 1e-7 * np.abs(uo).max()


def test_shell_recycling_solve_with_chebyshev():
    p = ShellProblem(8, dv="element", pressure=1e4, bc="ss", t0=0.01)
    p.set_thickness(np.random.default_rng(4).uniform(0.005, 0.02, p.mesh.nelem))
    p.assemble()
    pc = TACS.Pc(p.mat)
    pc.setChebyshev(4, 50.0)
    pc.factor()
    ksm = TACS.MixedKSM(p.mat, pc, 60, 1000, deflation=10)
    ksm.setTolerances(1e-9, 1e-30)
    x = p.assembler.createVec()
    assert ksm.solve(p.res, x) == 1
    K, *_ = _dense(p)
    r = p.res.getArray()
    assert np.linalg.norm(K @ x.getArray() - r) <= 2e-9 * np.linalg.norm(r)


def _bf16(a):
    """float64 -> float32 -> bfloat16 (round to nearest even) -> float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("kind", ["solid", "shell"])
@pytest.mark.parametrize("values", [64, 32, 16])
def test_fused_chebyshev_step_matches_numpy(kind, values):
    """One fused Chebyshev step (tacs_mat_cheb_step: a = A z, w = D^-1 (rhs - a),
    d = c1 d + c2 w, zout = z + d) against numpy with the matrix values and
    inverse diagonal blocks rounded as the kernel holds them (FP64; FP32;
    bfloat16 values with FP32 inverse blocks).  The FP64 arithmetic after
    the rounding is the same, so the bar is 1e-10 normwise."""
    from tacs_amd._lib import lib
    if kind == "solid":
        p = SolidProblem(6, dv="element")
        p.set_thickness(np.random.default_rng(5).uniform(0.5, 1.5, p.mesh.nelem))
    else:
        p = ShellProblem(9, warp=0.02, seed=4)
    p.assemble()
    a = p.assembler
    K, rowp, cols, A = _dense(p)
    b = A.shape[1]
    n = rowp.size - 1
    pc = TACS.Pc(p.mat)
    pc.factor()
    Dinv = np.zeros((n, b, b))
    xv, yv = a.createVec(), a.createVec()
    for c in range(b):
        e = np.zeros((n, b))
        e[:, c] = 1.0
        xv.getArray()[:] = e.reshape(-1)
        pc.applyFactor(xv, yv)
        Dinv[:, :, c] = yv.getArray().reshape(n, b)
    if values == 64:
        Kr, Dr = K, Dinv
    elif values == 32:
        Kr = K.astype(np.float32).astype(np.float64)
        Dr = Dinv.astype(np.float32).astype(np.float64)
    else:
        Kr = _bf16(K)
        Dr = Dinv.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(values + b)
    z = rng.standard_normal(b * n)
    rhs = Kr @ rng.standard_normal(b * n)
    d0 = rng.standard_normal(b * n)
    c1, c2 = 0.37, 1.9
    vecs = [a.createVec() for _ in range(4)]
    for v, arr in zip(vecs, (z, np.zeros_like(z), rhs, d0)):
        v.getArray()[:] = arr
    ptrs = [v.getDevicePointer() for v in vecs]
    rc = lib().tacs_mat_cheb_step(p.mat.ptr, pc.ptr, values, ptrs[0], ptrs[1], ptrs[2], ptrs[3],
                                  c1, c2)
    assert rc == 0
    w = np.einsum("nij,nj->ni", Dr, (rhs - Kr @ z).reshape(n, b)).reshape(-1)
    d_exp = c1 * d0 + c2 * w
    z_exp = z + d_exp
    d_got, z_got = vecs[3].getArray().copy(), vecs[1].getArray().copy()
    assert np.abs(d_got - d_exp).max() <= 1e-10 * np.abs(d_exp).max()
    assert np.abs(z_got - z_exp).max() <= 1e-10 * np.abs(z_exp).max()
    # z and rhs unchanged
    assert np.array_equal(vecs[0].getArray(), z) and np.array_equal(vecs[2].getArray(), rhs)


@pytest.mark.parametrize("kind", ["solid", "shell"])
@pytest.mark.parametrize("values", [64, 32, 16])
def test_fused_chebyshev_step3_matches_numpy(kind, values):
    """The three-term fused step (tacs_mat_cheb_step3: zprev <- (1 + c1) z
    - c1 zprev + c2 D^-1 (rhs - A z), the two-term step with d = z - zprev)
    against numpy with the values and inverse blocks rounded as the kernel
    holds them; 1e-10 normwise.  Two steps in a row also match two-term
    steps from the same start."""
    from tacs_amd._lib import lib
    if kind == "solid":
        p = SolidProblem(6, dv="element")
        p.set_thickness(np.random.default_rng(5).uniform(0.5, 1.5, p.mesh.nelem))
    else:
        p = ShellProblem(9, warp=0.02, seed=4)
    p.assemble()
    a = p.assembler
    K, rowp, cols, A = _dense(p)
    b = A.shape[1]
    n = rowp.size - 1
    pc = TACS.Pc(p.mat)
    pc.factor()
    Dinv = np.zeros((n, b, b))
    xv, yv = a.createVec(), a.createVec()
    for c in range(b):
        e = np.zeros((n, b))
        e[:, c] = 1.0
        xv.getArray()[:] = e.reshape(-1)
        pc.applyFactor(xv, yv)
        Dinv[:, :, c] = yv.getArray().reshape(n, b)
    if values == 64:
        Kr, Dr = K, Dinv
    elif values == 32:
        Kr = K.astype(np.float32).astype(np.float64)
        Dr = Dinv.astype(np.float32).astype(np.float64)
    else:
        Kr = _bf16(K)
        Dr = Dinv.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(3 * values + b)
    z = rng.standard_normal(b * n)
    zp = rng.standard_normal(b * n)
    rhs = Kr @ rng.standard_normal(b * n)
    c1, c2 = 0.37, 1.9
    vz, vp, vr = a.createVec(), a.createVec(), a.createVec()
    for v, arr in zip((vz, vp, vr), (z, zp, rhs)):
        v.getArray()[:] = arr
    pz, pp, pr = vz.getDevicePointer(), vp.getDevicePointer(), vr.getDevicePointer()
    assert lib().tacs_mat_cheb_step3(p.mat.ptr, pc.ptr, values, pz, pp, pr, c1, c2) == 0
    w = np.einsum("nij,nj->ni", Dr, (rhs - Kr @ z).reshape(n, b)).reshape(-1)
    z_exp = (1.0 + c1) * z - c1 * zp + c2 * w
    z_got = vp.getArray().copy()
    assert np.abs(z_got - z_exp).max() <= 1e-10 * np.abs(z_exp).max()
    assert np.array_equal(vz.getArray(), z) and np.array_equal(vr.getArray(), rhs)
    # z == zprev is refused
    assert lib().tacs_mat_cheb_step3(p.mat.ptr, pc.ptr, values, pz, pz, pr, c1, c2) != 0
