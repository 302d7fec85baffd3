"""torch.autograd over the thickness DVs of shells and solids
(ThicknessFunction: compliance or KS failure, the adjoint in backward)
against central finite differences of the forward function and against the
mixed-precision path."""
import numpy as np
import pytest
import torch

from tacs_amd.autograd import ThicknessFunctionProblem, thickness_function
from tacs_amd.problems import ShellProblem, SolidProblem

pytestmark = pytest.mark.gpu


def _fd_check(prob, t0, idx, h_rel, tol):
    t = torch.tensor(t0, dtype=torch.float64, device="cuda", requires_grad=True)
    f = thickness_function(t, prob)
    f.backward()
    g = t.grad.detach().cpu().numpy()
    assert np.all(np.isfinite(g))
    for e in idx:
        h = h_rel * t0[e]
        tp, tm = t0.copy(), t0.copy()
        tp[e] += h
        tm[e] -= h
        fp = prob.forward(torch.tensor(tp, device="cuda"))
        fm = prob.forward(torch.tensor(tm, device="cuda"))
        fd = (fp - fm) / (2 * h)
        assert abs(g[e] - fd) <= tol * max(abs(fd), 1e-3 * np.abs(g).max()), (e, g[e], fd)
    return float(f.detach()), g


@pytest.mark.parametrize("func", ["ks", "compliance"])
def test_shell_thickness_gradient_central_differences(func):
    p = ShellProblem(8, warp=0.01, seed=4, dv="element", ys=270e6)
    rng = np.random.default_rng(4)
    t0 = rng.uniform(0.005, 0.02, p.mesh.nelem)
    prob = ThicknessFunctionProblem(p, func=func, ks_weight=100.0, rtol=1e-12, m=120, mixed=False, deflation=0,
                                    restarts=500)
    idx = rng.choice(p.mesh.nelem, 6, replace=False)
    f64, g64 = _fd_check(prob, t0, idx, 1e-5, 2e-5)
    # the recycling mixed-precision path gives the same gradient to its tolerance
    pm = ThicknessFunctionProblem(p, func=func, ks_weight=100.0, rtol=1e-10, m=40,
                                  mixed=True, deflation=12, restarts=2000)
    t = torch.tensor(t0, dtype=torch.float64, device="cuda", requires_grad=True)
    fm = thickness_function(t, pm)
    fm.backward()
    gm = t.grad.detach().cpu().numpy()
    assert abs(float(fm.detach()) - f64) <= 1e-8 * abs(f64)
    assert np.abs(gm - g64).max() <= 1e-6 * np.abs(g64).max()


def test_solid_ks_thickness_gradient_central_differences():
    p = SolidProblem(4, dv="element")
    rng = np.random.default_rng(9)
    t0 = rng.uniform(0.5, 1.5, p.mesh.nelem)
    prob = ThicknessFunctionProblem(p, func="ks", ks_weight=50.0, rtol=1e-12, m=60, mixed=False, deflation=0,
                                    restarts=500)
    _fd_check(prob, t0, rng.choice(p.mesh.nelem, 6, replace=False), 1e-5, 2e-5)
