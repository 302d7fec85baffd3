"""Writes tests/golden/elements.npz: element-level reference values computed by
the reference's own sources (oracle/_ref/libtacs_ref.so, built by
oracle/build_ref.sh from /root/reference) on seeded inputs.  Run in a
container that has the reference tree:  python tests/golden/make_fixtures.py

The Quad4 shell node-coordinate sensitivities come from the complex build of
the same sources (oracle/_ref/cplx/libtacs_ref_cplx.so: the reference's
default element derivatives take complex steps there), run in a child
process because the two builds define the same C++ symbols.
"""
import ctypes
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libtacs_ref.so")


def ref_lib():
    L = ctypes.CDLL(REF, mode=ctypes.RTLD_GLOBAL)
    L.ref_quad4_adj_product.restype = ctypes.c_double
    L.ref_quad4_failure.restype = ctypes.c_double
    return L


def _d(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def hex_nodes(order):
    k = [-1.0, 1.0] if order == 2 else [-1.0, 0.0, 1.0]
    return np.array([[x, y, z] for z in k for y in k for x in k])


def make(seed=2024):
    L = ref_lib()
    rng = np.random.default_rng(seed)
    c = ctypes.c_double
    out = {}
    for order in (2, 3):
        npe = order ** 3
        for s in range(2):
            X = (0.5 * hex_nodes(order) @ np.diag([1.0, 0.8, 1.3])
                 + 0.06 * rng.standard_normal((npe, 3))).reshape(-1)
            u = 1e-3 * rng.standard_normal(3 * npe)
            t = rng.uniform(0.2, 1.0)
            res = np.zeros(3 * npe)
            mat = np.zeros(9 * npe * npe)
            L.ref_solid_element(order, _d(X), _d(u), c(70e9), c(0.3), c(2700.0), c(t), c(1.0),
                                c(0.25), _d(res), _d(mat))
            key = f"hex{npe}_{s}"
            out[key + "_X"], out[key + "_u"], out[key + "_t"] = X, u, np.array([t])
            out[key + "_res"], out[key + "_mat"] = res, mat
            trac = rng.standard_normal(3)
            tr = np.zeros((6, 3 * npe))
            for face in range(6):
                L.ref_solid_traction(order, face, _d(X), _d(trac), _d(tr[face]))
            out[key + "_trac"], out[key + "_trac_res"] = trac, tr
    # node-coordinate sensitivities of the solid elements (separate stream so
    # the values above do not move)
    rng2 = np.random.default_rng(seed + 1)
    for order in (2, 3):
        npe = order ** 3
        for s in range(2):
            key = f"hex{npe}_{s}"
            X, u, t = out[key + "_X"], out[key + "_u"], float(out[key + "_t"][0])
            psi = rng2.standard_normal(3 * npe)
            adj = np.zeros(3 * npe)
            L.ref_solid_adj_xpt(order, _d(X), _d(u), _d(psi), c(70e9), c(0.3), c(t), _d(adj))
            cx = np.zeros(3 * npe)
            L.ref_solid_compliance_xpt(order, _d(X), _d(u), c(70e9), c(0.3), c(t), _d(cx))
            out[key + "_psi"], out[key + "_adjxpt"], out[key + "_compxpt"] = psi, adj, cx
    for s in range(4):
        X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float) * [0.3, 0.2, 1.0]
        X = (X + 0.02 * rng.standard_normal((4, 3))).reshape(-1)
        v = 1e-3 * rng.standard_normal(24)
        psi = rng.standard_normal(24)
        t = rng.uniform(0.005, 0.02)
        res = np.zeros(24)
        mat = np.zeros(576)
        L.ref_quad4_shell(_d(X), _d(v), c(70e9), c(0.3), c(270e6), c(t), c(1.0), _d(res), _d(mat))
        adj = L.ref_quad4_adj_product(_d(X), _d(v), _d(psi), c(70e9), c(0.3), c(270e6), c(t))
        fail = np.zeros(4)
        det = np.zeros(4)
        for n in range(4):
            dd = ctypes.c_double()
            fail[n] = L.ref_quad4_failure(_d(X), _d(v), c(70e9), c(0.3), c(270e6), c(t), n,
                                          ctypes.byref(dd))
            det[n] = dd.value
        pres = np.zeros(24)
        L.ref_quad4_pressure(_d(X), c(1e4), _d(pres))
        key = f"quad4_{s}"
        out[key + "_X"], out[key + "_v"], out[key + "_psi"] = X, v, psi
        out[key + "_t"], out[key + "_res"], out[key + "_mat"] = np.array([t]), res, mat
        out[key + "_adj"], out[key + "_fail"], out[key + "_det"] = np.array([adj]), fail, det
        out[key + "_pres"] = pres
    # solid face pressures (TACSPressure3D), separate stream
    rng3 = np.random.default_rng(seed + 2)
    for order in (2, 3):
        npe = order ** 3
        for s in range(2):
            key = f"hex{npe}_{s}"
            X = out[key + "_X"]
            pv = float(rng3.standard_normal())
            pr = np.zeros((6, 3 * npe))
            for face in range(6):
                L.ref_solid_pressure(order, face, _d(X), c(pv), _d(pr[face]))
            out[key + "_p"], out[key + "_p_res"] = np.array([pv]), pr
    # matrix-free element products (Jacobian, stiffness, mass), separate stream
    rng4 = np.random.default_rng(seed + 3)
    for order in (2, 3):
        npe = order ** 3
        for s in range(2):
            key = f"hex{npe}_{s}"
            X, t = out[key + "_X"], float(out[key + "_t"][0])
            x = rng4.standard_normal(3 * npe)
            mv = np.zeros((3, 3 * npe))
            for mt in range(3):
                L.ref_solid_matvec(order, mt, _d(X), _d(x), c(70e9), c(0.3), c(2700.0), c(t),
                                   c(0.8), c(0.3), _d(mv[mt]))
            out[key + "_mvx"], out[key + "_mv"] = x, mv
    # element matrices by type (stiffness, mass, geometric stiffness at u)
    for order in (2, 3):
        npe = order ** 3
        for s in range(2):
            key = f"hex{npe}_{s}"
            X, u, t = out[key + "_X"], out[key + "_u"], float(out[key + "_t"][0])
            mt = np.zeros((3, 9 * npe * npe))
            for k in range(3):
                L.ref_solid_mat_type(order, k + 1, _d(X), _d(u), c(70e9), c(0.3), c(2700.0), c(t),
                                     _d(mt[k]))
            out[key + "_mattype"] = mt
    # Quad4 shell element matrices by type (stiffness, mass, geometric
    # stiffness along the state v)
    for s in range(4):
        key = f"quad4_{s}"
        X, v, t = out[key + "_X"], out[key + "_v"], float(out[key + "_t"][0])
        mt = np.zeros((3, 576))
        for k in range(3):
            L.ref_quad4_mat_type(k + 1, _d(X), _d(v), c(70e9), c(0.3), c(t), _d(mt[k]))
        out[key + "_mattype"] = mt
        # TACSQuad4NonlinearShell at the state 30 v (residual, Jacobian)
        vn = 30.0 * v
        res, mat = np.zeros(24), np.zeros(576)
        L.ref_quad4_nonlinear_shell(_d(X), _d(vn), c(70e9), c(0.3), c(t), _d(res), _d(mat))
        out[key + "_nl_res"], out[key + "_nl_mat"] = res, mat
    # Quad4 shells with TACSShellRefAxisTransform (residual, Jacobian, failure)
    axis = np.array([0.3, 1.0, 0.2])
    out["quad4_axis"] = axis
    for s in range(4):
        key = f"quad4_{s}"
        X, v, t = out[key + "_X"], out[key + "_v"], float(out[key + "_t"][0])
        res, mat, fail = np.zeros(24), np.zeros(576), np.zeros(4)
        L.ref_quad4_shell_axis(_d(X), _d(v), c(70e9), c(0.3), c(270e6), c(t), _d(axis), _d(res),
                               _d(mat), _d(fail))
        out[key + "_axis_res"], out[key + "_axis_mat"], out[key + "_axis_fail"] = res, mat, fail
    # Quad4 shells with a thickness offset (TACSIsoShellConstitutive tOffset)
    for s in range(4):
        key = f"quad4_{s}"
        X, v, psi = out[key + "_X"], out[key + "_v"], out[key + "_psi"]
        t = float(out[key + "_t"][0])
        toff = (-0.3, 0.5, 0.15, -0.5)[s]
        res, mat, mass = np.zeros(24), np.zeros(576), np.zeros(576)
        adj, fail, dfail = np.zeros(1), np.zeros(4), np.zeros(4)
        L.ref_quad4_shell_offset(_d(X), _d(v), _d(psi), c(70e9), c(0.3), c(270e6), c(t),
                                 c(toff), _d(res), _d(mat), _d(mass), _d(adj), _d(fail),
                                 _d(dfail))
        out[key + "_toff"] = np.array([toff])
        out[key + "_off_res"], out[key + "_off_mat"], out[key + "_off_mass"] = res, mat, mass
        out[key + "_off_adj"], out[key + "_off_fail"], out[key + "_off_dfail"] = adj, fail, dfail
    # Quad4 shell node-coordinate sensitivities (complex build, child process)
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "in.npz")
        dst = os.path.join(tmp, "out.npz")
        np.savez(src, **{k: v for k, v in out.items() if k.startswith("quad4_")})
        subprocess.run([sys.executable, os.path.abspath(__file__), "--cplx", src, dst], check=True)
        with np.load(dst) as f:
            out.update({k: f[k] for k in f.files})
    np.savez_compressed(os.path.join(HERE, "elements.npz"), **out)
    return out


KS_WEIGHT = 30.0


def make_cplx(src, dst):
    """Element node sensitivities of the four Quad4 cases: psi^T dR/dX, the
    pressure load's psi^T dF/dX (p = 1e4), the compliance element term
    (TACSCompliance::getElementXptSens: dfdq = w detXd, dfddetXd = w U) and
    the KS failure element term of a one-element KS_CONTINUOUS aggregate
    (ksWeight 30; TACSKSFailure::getElementXptSens)."""
    C = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "cplx", "libtacs_ref_cplx.so"))
    C.refc_quad4_point.restype = ctypes.c_double
    c = ctypes.c_double
    res = {}
    with np.load(src) as f:
        for s in range(4):
            key = f"quad4_{s}"
            X, v, psi = f[key + "_X"].copy(), f[key + "_v"].copy(), f[key + "_psi"].copy()
            t = float(f[key + "_t"][0])
            adj = np.zeros(12)
            C.refc_quad4_adj_xpt(_d(X), _d(v), _d(psi), c(70e9), c(0.3), c(t), c(1.0), _d(adj))
            pres = np.zeros(12)
            C.refc_quad4_pressure_xpt(_d(X), c(1e4), _d(psi), c(1.0), _d(pres))
            U, F, det = np.zeros(4), np.zeros(4), np.zeros(4)
            for n in range(4):
                dd = ctypes.c_double()
                U[n] = C.refc_quad4_point(2, _d(X), _d(v), c(70e9), c(0.3), c(270e6), c(t), n,
                                          ctypes.byref(dd))
                det[n] = dd.value
                F[n] = C.refc_quad4_point(1, _d(X), _d(v), c(70e9), c(0.3), c(270e6), c(t), n,
                                          ctypes.byref(dd))
            comp = np.zeros(12)
            C.refc_quad4_point_xpt(2, _d(X), _d(v), c(70e9), c(0.3), c(270e6), c(t), _d(U),
                                   _d(det.copy()), _d(comp))
            fmax = F.max()
            ksum = np.sum(det * np.exp(KS_WEIGHT * (F - fmax)))
            expf = np.exp(KS_WEIGHT * (F - fmax)) / ksum
            ks = np.zeros(12)
            C.refc_quad4_point_xpt(1, _d(X), _d(v), c(70e9), c(0.3), c(270e6), c(t),
                                   _d(expf / KS_WEIGHT), _d(det * expf), _d(ks))
            res[key + "_adjxpt"], res[key + "_presxpt"] = adj, pres
            res[key + "_compxpt"], res[key + "_ksxpt"] = comp, ks
            res[key + "_sed"] = U
    np.savez(dst, **res)


if __name__ == "__main__":
    if len(sys.argv) == 4 and sys.argv[1] == "--cplx":
        make_cplx(sys.argv[2], sys.argv[3])
    else:
        make()
        print("wrote", os.path.join(HERE, "elements.npz"))
