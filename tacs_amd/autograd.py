"""Structural functions of the GPU path as differentiable torch functions.

SIMPCompliance: the compliance of a solid SIMP problem in the element
densities.  ThicknessFunction: the compliance or KS failure of a solid or shell
problem in its per-element thickness variables (the shells' thickness DVs of
TACSIsoShellConstitutive, the solids' t of TACSSolidConstitutive).

Forward: t = (Emin + x^p (E0 - Emin)) / E0 (the reference's solid
constitutive model is linear in its thickness variable t,
src/constitutive/TACSSolidConstitutive.cpp:142, so SIMP is applied on top of
it), assembleJacobian, the solve, and the compliance.  Backward: the adjoint
solve K psi = dC/du (addSVSens + applyBCs), then
dC/dt = addDVSens - addAdjointResProducts (TACSAssembler.cpp:4767, :5033) in
the GPU kernels, which recompute the element strains from u and psi instead
of storing them, and the chain rule dt/dx = p x^(p-1) (E0 - Emin)/E0.

The problem objects hold one state (design variables, assembled matrix,
displacements).  Every forward stamps it with a new version; a backward whose
forward is no longer the latest (f(x1) + f(x2), a line search, a retained
graph) first runs its own forward again at the saved point, so the gradient
is always taken at the point its value was.
"""
import torch

from . import TACS, functions, lib
from .meshes import simp_thickness, simp_dthickness


def _sync():
    torch.cuda.synchronize()
    lib().tacs_device_synchronize()


class SolidComplianceProblem:
    """Holds a SolidProblem (one DV per element, tNum = element index) and its
    solver for repeated forward / backward passes."""

    def __init__(self, problem, p=3.0, E0=1.0, Emin=1e-9, rtol=1e-8, m=120, mixed=True,
                 deflation=0, restarts=100000, chebyshev=None):
        self.prob = problem
        self.p, self.E0, self.Emin, self.rtol = p, E0, Emin, rtol
        a = problem.assembler
        self.dv = a.createDesignVec()
        self.dfdt = a.createDesignVec()
        self.pc = TACS.Pc(problem.mat)
        if chebyshev and deflation:
            # (degree, ratio): the recycling solvers' Chebyshev preconditioner
            self.pc.setChebyshev(*chebyshev)
        if mixed:
            # deflation > 0: the recycling solver (its space carries over to
            # the adjoint solve with the same matrix)
            self.ksm = TACS.MixedKSM(problem.mat, self.pc, m, restarts, deflation=deflation)
        elif deflation:
            # the same recycling on the FP64 operator
            self.ksm = TACS.RecyclingKSM(problem.mat, self.pc, m, restarts, deflation, prec=64)
        else:
            self.ksm = TACS.KSM(problem.mat, self.pc, min(m, 512), restarts)
        self.ksm.setTolerances(rtol, 1e-30)
        self.comp = functions.Compliance(a)
        self.iters = []
        self.version = 0

    def forward(self, x):
        a = self.prob.assembler
        self.version += 1
        t = simp_thickness(x, self.p, self.E0, self.Emin).to(torch.float64).contiguous()
        _sync()
        self.dv.copyFromDevice(t.data_ptr(), t.numel())
        a.setDesignVars(self.dv)
        a.zeroVariables()
        a.assembleJacobian(1.0, 0.0, 0.0, self.prob.res, self.prob.mat)
        self.pc.factor()
        if self.ksm.solve(self.prob.res, self.prob.u) != 1:
            raise RuntimeError("forward solve did not converge")
        self.iters.append(self.ksm.getIterCount())
        self.prob.u.scale(-1.0)
        a.setVariables(self.prob.u)
        return a.evalFunctions([self.comp])[0]

    def gradient_t(self):
        """dC/dt on the device (torch tensor) at the current state."""
        a = self.prob.assembler
        dfdu = a.createVec()
        a.addSVSens([self.comp], [dfdu])
        a.applyBCs(dfdu)
        psi = a.createVec()
        if self.ksm.solve(dfdu, psi) != 1:
            raise RuntimeError("adjoint solve did not converge")
        self.iters.append(self.ksm.getIterCount())
        self.dfdt.zeroEntries()
        a.addDVSens([self.comp], [self.dfdt], 1.0)
        a.addAdjointResProducts([psi], [self.dfdt], -1.0)
        n = self.dfdt.getSize()
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        self.dfdt.copyToDevice(out.data_ptr(), n)
        _sync()
        return out


class SIMPCompliance(torch.autograd.Function):
    """C(x) for element densities x (1-D float64 CUDA tensor)."""

    @staticmethod
    def forward(ctx, x, prob):
        c = prob.forward(x)
        ctx.prob = prob
        ctx.version = prob.version
        ctx.save_for_backward(x)
        return torch.tensor(c, dtype=torch.float64, device=x.device)

    @staticmethod
    def backward(ctx, grad):
        (x,) = ctx.saved_tensors
        prob = ctx.prob
        if prob.version != ctx.version:  # another forward moved the state
            prob.forward(x.detach())
            ctx.version = prob.version
        dfdt = prob.gradient_t()
        dtdx = simp_dthickness(x, prob.p, prob.E0, prob.Emin)
        return grad * dfdt * dtdx, None


def simp_compliance(x, prob):
    return SIMPCompliance.apply(x, prob)


class ThicknessFunctionProblem:
    """f(t) and df/dt for one DV per element (tNum = element index) of a
    SolidProblem or ShellProblem: f is the compliance or the KS failure
    (KS_CONTINUOUS, ksWeight).  Forward: setDesignVars, assembleJacobian, the
    solve, evalFunctions.  Backward: addSVSens, the adjoint solve with the
    same matrix, addDVSens - addAdjointResProducts (TACSAssembler.cpp:4767,
    :5033), the element strains recomputed in the kernels from u and psi."""

    def __init__(self, problem, func="ks", ks_weight=100.0, rtol=1e-8, m=60, mixed=True,
                 deflation=20, restarts=100000, chebyshev=None):
        self.prob = problem
        a = problem.assembler
        self.dv = a.createDesignVec()
        self.dfdt = a.createDesignVec()
        self.pc = TACS.Pc(problem.mat)
        if chebyshev and deflation:
            self.pc.setChebyshev(*chebyshev)
        if mixed:
            self.ksm = TACS.MixedKSM(problem.mat, self.pc, m, restarts, deflation=deflation)
        elif deflation:
            self.ksm = TACS.RecyclingKSM(problem.mat, self.pc, m, restarts, deflation, prec=64)
        else:
            self.ksm = TACS.KSM(problem.mat, self.pc, min(m, 512), restarts)
        self.ksm.setTolerances(rtol, 1e-30)
        if func == "ks":
            self.func = functions.KSFailure(a, ksWeight=ks_weight)
        elif func == "compliance":
            self.func = functions.Compliance(a)
        else:
            raise ValueError(f"unknown function {func!r}")
        self.iters = []
        self.version = 0

    def forward(self, t):
        a = self.prob.assembler
        self.version += 1
        t = t.to(torch.float64).contiguous()
        _sync()
        self.dv.copyFromDevice(t.data_ptr(), t.numel())
        a.setDesignVars(self.dv)
        a.zeroVariables()
        a.assembleJacobian(1.0, 0.0, 0.0, self.prob.res, self.prob.mat)
        self.pc.factor()
        if self.ksm.solve(self.prob.res, self.prob.u) != 1:
            raise RuntimeError("forward solve did not converge")
        self.iters.append(self.ksm.getIterCount())
        self.prob.u.scale(-1.0)
        a.setVariables(self.prob.u)
        return a.evalFunctions([self.func])[0]

    def gradient_t(self):
        """df/dt on the device (torch tensor) at the current state."""
        a = self.prob.assembler
        dfdu = a.createVec()
        a.addSVSens([self.func], [dfdu])
        a.applyBCs(dfdu)
        psi = a.createVec()
        if self.ksm.solve(dfdu, psi) != 1:
            raise RuntimeError("adjoint solve did not converge")
        self.iters.append(self.ksm.getIterCount())
        self.dfdt.zeroEntries()
        a.addDVSens([self.func], [self.dfdt], 1.0)
        a.addAdjointResProducts([psi], [self.dfdt], -1.0)
        n = self.dfdt.getSize()
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        self.dfdt.copyToDevice(out.data_ptr(), n)
        _sync()
        return out


class ThicknessFunction(torch.autograd.Function):
    """f(t) for per-element thicknesses t (1-D float64 CUDA tensor)."""

    @staticmethod
    def forward(ctx, t, prob):
        f = prob.forward(t)
        ctx.prob = prob
        ctx.version = prob.version
        ctx.save_for_backward(t)
        return torch.tensor(f, dtype=torch.float64, device=t.device)

    @staticmethod
    def backward(ctx, grad):
        (t,) = ctx.saved_tensors
        prob = ctx.prob
        if prob.version != ctx.version:  # another forward moved the state
            prob.forward(t.detach())
            ctx.version = prob.version
        return grad * prob.gradient_t(), None


def thickness_function(t, prob):
    return ThicknessFunction.apply(t, prob)
