"""Launch variants and the tuner: every variant of the sliced ELL product,
the multi right-hand-side product and the solid Jacobian rows kernel gives
the same values (bit for bit), the element-block form agrees with the rows
to rounding, and the tuners (Mat.autotune, Mat.autotuneSpMM,
Assembler.autotune) keep a valid variant after which the results are
unchanged."""
import ctypes

import numpy as np
import pytest

from tacs_amd import TACS
from tacs_amd.problems import ShellProblem, SolidProblem

pytestmark = pytest.mark.gpu


def _products(p, L):
    """y = A x for the three kinds through the library's product entry."""
    b, nr, nnz = p.mat.getInfo()
    n = b * nr
    import torch
    x = torch.linspace(-1.0, 1.0, n, dtype=torch.float64, device="cuda")
    x32 = x.float()
    out = []
    for kind, pv, pw in ((0, 64, 64), (1, 32, 64), (2, 32, 32)):
        xa = x32 if kind == 2 else x
        y = torch.zeros_like(xa)
        assert L.tacs_mat_mult_prec(p.mat.ptr, pv, pw, xa.data_ptr(), y.data_ptr()) == 0
        torch.cuda.synchronize()
        out.append(y.cpu().numpy())
    return out


@pytest.mark.parametrize("kind_of_mesh", ["hex8", "hex27", "shell"])
def test_sell_variants_identical(kind_of_mesh):
    L = TACS.lib()
    if kind_of_mesh == "shell":
        p = ShellProblem(9)
    else:
        p = SolidProblem(7 if kind_of_mesh == "hex8" else 3,
                         order=2 if kind_of_mesh == "hex8" else 3)
    p.assemble()
    b = p.mat.getInfo()[0]
    nv = L.tacs_hip_sell_spmv_nvariants(b)
    assert nv == (8 if b == 3 else 4)
    ref = _products(p, L)
    for v in range(nv):
        for kind in range(3):
            p.mat.setSellVariant(kind, v)
            assert p.mat.getSellVariant(kind) == v
        got = _products(p, L)
        for kind in range(3):
            assert np.array_equal(got[kind], ref[kind]), (v, kind)
    for kind in range(3):
        p.mat.setSellVariant(kind, -1)
    r = p.mat.autotune(reps=3)
    assert len(r["ms"]) == 3 and all(len(t) == nv for t in r["ms"])
    assert all(0 <= c < nv for c in r["chosen"])
    assert all(t > 0 for row in r["ms"] for t in row)
    got = _products(p, L)
    for kind in range(3):
        assert np.array_equal(got[kind], ref[kind])
    with pytest.raises(TACS.TACSError):
        p.mat.setSellVariant(0, nv)


@pytest.mark.parametrize("nrhs", [4, 8, 12, 6, 16, 5])
@pytest.mark.parametrize("prec", [64, 32])
def test_spmm_variants_identical(nrhs, prec):
    import torch
    L = TACS.lib()
    p = SolidProblem(3, order=3)
    p.assemble()
    b, nr, nnz = p.mat.getInfo()
    rp, cp, ap = p.mat.getDeviceArrays()
    rowp, cols, Ah = p.mat.getArrays()
    A = torch.from_numpy(Ah.reshape(-1).copy()).cuda()
    if prec == 32:
        A = A.float()
    g = torch.Generator(device="cpu").manual_seed(nrhs)
    X = torch.randn(3 * nr * nrhs, generator=g, dtype=torch.float64).cuda()
    nv = L.tacs_hip_bsr3_spmm_nvariants(nrhs)
    # two right-hand sides per lane (even R <= 12), four per lane (R = 16, 20, ...)
    assert nv == {4: 2, 8: 2, 12: 2, 6: 2, 16: 5, 5: 1}[nrhs]
    outs = []
    for v in range(nv):
        Y = torch.zeros_like(X)
        assert L.tacs_hip_bsr3_spmm_variant(v, prec, nr, nrhs, rp, cp, A.data_ptr(),
                                            X.data_ptr(), Y.data_ptr(), None) == 0
        torch.cuda.synchronize()
        outs.append(Y.cpu().numpy())
    for v in range(1, nv):
        assert np.array_equal(outs[v], outs[0]), v
    assert L.tacs_hip_bsr3_spmm_variant(nv, prec, nr, nrhs, rp, cp, A.data_ptr(), X.data_ptr(),
                                        X.data_ptr(), None) != 0
    # the dense product of the same values
    Ad = np.zeros((3 * nr, 3 * nr))
    for i in range(nr):
        for s in range(rowp[i], rowp[i + 1]):
            Ad[3 * i:3 * i + 3, 3 * cols[s]:3 * cols[s] + 3] = Ah[s]
    if prec == 32:
        Ad = Ad.astype(np.float32).astype(np.float64)
    Xh = X.cpu().numpy().reshape(3 * nr, nrhs)
    ref = Ad @ Xh
    assert np.abs(outs[0].reshape(3 * nr, nrhs) - ref).max() <= 1e-12 * np.abs(ref).max()
    ms = p.mat.autotuneSpMM(nrhs, prec=prec, reps=2)
    assert len(ms) == nv
    sc = L.tacs_tune_size_class(nr)
    assert 0 <= TACS.getTuning(TACS.TUNE_SPMM, (100 if prec == 32 else 0) + nrhs, sc) < nv
    TACS.setTuning(TACS.TUNE_SPMM, (100 if prec == 32 else 0) + nrhs, sc, -1)


@pytest.mark.parametrize("order,n", [(2, 5), (3, 3)])
def test_jacobian_variants_identical(order, n):
    L = TACS.lib()
    p = SolidProblem(n, order=order, dv="element")
    rng = np.random.default_rng(order)
    p.set_thickness(rng.uniform(0.5, 1.5, p.mesh.nelem))
    a = p.assembler
    nv = L.tacs_hip_solid_jacobian_nvariants(order)
    assert nv == (5 if order == 2 else 4)
    vals = []
    for v in range(nv + 1):
        a.setJacobianVariant(v)
        assert a.getJacobianVariant() == v
        assert a.getJacobianKernel() == (2 if v == nv else 1)
        a.assembleJacobian(1.0, 0.0, 0.25, None, p.mat)
        vals.append(p.mat.getArrays()[2].copy())
    for v in range(1, nv):
        assert np.array_equal(vals[v], vals[0]), v
    scale = np.abs(vals[0]).max()
    assert np.abs(vals[nv] - vals[0]).max() <= 1e-13 * scale
    a.setJacobianVariant(-1)
    ms = a.autotune(p.mat, reps=1)
    assert len(ms) == nv + 1 and all(t > 0 for t in ms)
    chosen = a.getJacobianVariant()
    assert 0 <= chosen <= nv
    a.assembleJacobian(1.0, 0.0, 0.25, None, p.mat)
    got = p.mat.getArrays()[2]
    assert np.array_equal(got, vals[chosen])
    # the rows kernel only: the element-block form is not a choice
    a.setJacobianKernel(1)
    assert len(a.autotune(p.mat, reps=1)) == nv
    assert a.getJacobianVariant() < nv
    a.setJacobianKernel(0)
    a.setJacobianVariant(-1)
    with pytest.raises(TACS.TACSError):
        a.setJacobianVariant(nv + 1)
