"""ctypes front end of the oracle library (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  It is the CPU restatement of the reference's
algorithms that the HIP path is checked against; the product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.orc_hex_quadrature.restype = ctypes.c_double
        L.orc_hex_face_quadrature.restype = ctypes.c_double
        L.orc_node_csr.restype = ctypes.c_long
        L.orc_assemble_solid.restype = ctypes.c_long
        L.orc_gmres.restype = ctypes.c_int
        L.orc_gmres_capped.restype = ctypes.c_int
        L.orc_bsr_diag_inv.restype = ctypes.c_int
        L.orc_compliance.restype = ctypes.c_double
        L.orc_quad4_adj_product.restype = ctypes.c_double
        L.orc_quad4_failure.restype = ctypes.c_double
        L.orc_shell_failure.restype = ctypes.c_double
        L.orc_shell_compliance.restype = ctypes.c_double
        L.orc_assemble_shell.restype = ctypes.c_long
        L.orc_assemble_shell_mat.restype = ctypes.c_long
        L.orc_par_assemble_solid.restype = ctypes.c_long
        L.orc_par_assemble_shell.restype = ctypes.c_long
        L.orc_par_gmres.restype = ctypes.c_int
        L.orc_color_elements.restype = ctypes.c_int
        L.orc_par_threads.restype = ctypes.c_int
        L.orc_hex_quadrature.argtypes = None
        _lib = L
    return _lib


def _p(a, dt=np.float64):
    if a is None:
        return None
    assert a.dtype == dt and a.flags.c_contiguous, (a.dtype, dt)
    return a.ctypes.data_as(_d if dt == np.float64 else _i)


def _dd(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ii(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def iso_C21(E, nu):
    C = np.zeros(21)
    lib().orc_iso_C21(ctypes.c_double(E), ctypes.c_double(nu), _p(C))
    return C


def hex_quadrature(order, n):
    pt = np.zeros(3)
    w = lib().orc_hex_quadrature(order, n, _p(pt))
    return w, pt


def hex_basis(order, pt):
    npe = order ** 3
    N = np.zeros(npe)
    Nxi = np.zeros(3 * npe)
    lib().orc_hex_basis(order, _p(_dd(pt)), _p(N), _p(Nxi))
    return N, Nxi.reshape(npe, 3)


def solid_element(order, X, u, C21, t=1.0, rho=0.0, alpha=1.0, gamma=0.0,
                  want_res=True, want_mat=True):
    nv = 3 * order ** 3
    res = np.zeros(nv) if want_res else None
    mat = np.zeros((nv, nv)) if want_mat else None
    lib().orc_solid_element(order, _p(_dd(X)), _p(_dd(u)), _p(_dd(C21)),
                            ctypes.c_double(t), ctypes.c_double(rho),
                            ctypes.c_double(alpha), ctypes.c_double(gamma),
                            _p(res), _p(mat))
    return res, mat


def solid_geo_element(order, Xe, ue, C21, t=1.0):
    """Geometric stiffness element matrix at the state ue
    (TACSElement3D::getMatType, TACS_GEOMETRIC_STIFFNESS_MATRIX)."""
    nv = 3 * order ** 3
    mat = np.zeros((nv, nv))
    lib().orc_solid_geo_element(order, _p(_dd(Xe)), _p(_dd(ue)), _p(_dd(C21)),
                                ctypes.c_double(t), _p(mat))
    return mat


def assemble_geo(order, conn, X, u, t_elem, C21, rowp, cols, A, gc=1.0):
    """A += gc * geometric stiffness at u (in place); returns missing slots."""
    conn = _ii(conn)
    return lib().orc_assemble_geo(order, conn.shape[0], _p(conn, np.int32), _p(_dd(X)),
                                  _p(_dd(u)), _p(_dd(t_elem)), _p(_dd(C21)), ctypes.c_double(gc),
                                  _p(_ii(rowp), np.int32), _p(_ii(cols), np.int32), _p(A))


def solid_matvec_element(order, Xe, xe, C21, t=1.0, rho=0.0, kc=1.0, mc=0.0):
    """Element matrix-free product (TACSElement3D::addMatVecProduct)."""
    c = ctypes.c_double
    y = np.zeros(3 * order ** 3)
    lib().orc_solid_matvec_element(order, _p(_dd(Xe)), _p(_dd(xe)), _p(_dd(C21)), c(t), c(rho),
                                   c(kc), c(mc), _p(y))
    return y


def matfree_product(order, conn, X, t_elem, C21, x, rho=0.0, kc=1.0, mc=0.0):
    """sum_e P_e^T K_e P_e x (TACSAssembler::addMatrixFreeVecProduct without BCs)."""
    c = ctypes.c_double
    conn = _ii(conn)
    x = _dd(x)
    y = np.zeros_like(x)
    lib().orc_matfree_product(order, conn.shape[0], _p(conn, np.int32), _p(_dd(X)),
                              _p(_dd(t_elem)), _p(_dd(C21)), c(rho), c(kc), c(mc), _p(x), _p(y))
    return y


def node_csr(nnodes, conn, npe):
    conn = _ii(conn).reshape(-1)
    nelem = conn.size // npe
    rowp = np.zeros(nnodes + 1, np.int32)
    nnz = lib().orc_node_csr(nnodes, nelem, npe, _p(conn, np.int32),
                             _p(rowp, np.int32), None, ctypes.c_long(0))
    cols = np.zeros(nnz, np.int32)
    got = lib().orc_node_csr(nnodes, nelem, npe, _p(conn, np.int32),
                             _p(rowp, np.int32), _p(cols, np.int32),
                             ctypes.c_long(nnz))
    assert got == nnz
    return rowp, cols


def assemble_solid(order, conn, X, u, t_elem, C21, rowp, cols, rho=0.0,
                   alpha=1.0, gamma=0.0, want_mat=True, want_res=True):
    npe = order ** 3
    conn = _ii(conn).reshape(-1)
    nelem = conn.size // npe
    nnodes = rowp.size - 1
    A = np.zeros(cols.size * 9) if want_mat else None
    res = np.zeros(3 * nnodes) if want_res else None
    missing = lib().orc_assemble_solid(
        order, nelem, _p(conn, np.int32), _p(_dd(X)),
        _p(_dd(u)) if u is not None else None, _p(_dd(t_elem)), _p(_dd(C21)),
        ctypes.c_double(rho), ctypes.c_double(alpha), ctypes.c_double(gamma),
        _p(_ii(rowp), np.int32), _p(_ii(cols), np.int32), _p(A), _p(res))
    assert missing == 0
    return A, res


def add_tractions(order, conn, X, aux_elem, aux_face, aux_trac, res):
    aux_elem = _ii(aux_elem)
    lib().orc_add_tractions(order, aux_elem.size, _p(aux_elem, np.int32),
                            _p(_ii(aux_face), np.int32), _p(_dd(aux_trac)),
                            _p(_ii(conn).reshape(-1), np.int32), _p(_dd(X)),
                            _p(res))
    return res


def pressure_face(order, face, Xe, p):
    """Residual of one TACSPressure3D face (element vector, 3 * npe)."""
    res = np.zeros(3 * order ** 3)
    lib().orc_pressure_face(order, face, _p(_dd(Xe)), ctypes.c_double(p), _p(res))
    return res


def add_pressures(order, conn, X, aux_elem, aux_face, aux_p, res):
    aux_elem = _ii(aux_elem)
    lib().orc_add_pressures(order, aux_elem.size, _p(aux_elem, np.int32),
                            _p(_ii(aux_face), np.int32), _p(_dd(aux_p)),
                            _p(_ii(conn).reshape(-1), np.int32), _p(_dd(X)), _p(res))
    return res


def apply_bc_mat(b, bc_nodes, bc_vars, rowp, cols, A):
    bc_nodes = _ii(bc_nodes)
    lib().orc_apply_bc_mat(b, bc_nodes.size, _p(bc_nodes, np.int32),
                           _p(_ii(bc_vars), np.int32), _p(_ii(rowp), np.int32),
                           _p(_ii(cols), np.int32), _p(A))


def apply_bc_vec(b, bc_nodes, bc_vars, x, bc_vals=None, u=None, lam=1.0):
    bc_nodes = _ii(bc_nodes)
    lib().orc_apply_bc_vec(b, bc_nodes.size, _p(bc_nodes, np.int32),
                           _p(_ii(bc_vars), np.int32),
                           _p(_dd(bc_vals)) if bc_vals is not None else None,
                           _p(_dd(u)) if u is not None else None,
                           ctypes.c_double(lam), _p(x))


def bsr_mult(b, rowp, cols, A, x):
    n = rowp.size - 1
    y = np.zeros(b * n)
    lib().orc_bsr_mult(b, n, _p(_ii(rowp), np.int32), _p(_ii(cols), np.int32),
                       _p(_dd(A)), _p(_dd(x)), _p(y))
    return y


def bsr_diag_inv(b, rowp, cols, A):
    n = rowp.size - 1
    D = np.zeros(n * b * b)
    bad = lib().orc_bsr_diag_inv(b, n, _p(_ii(rowp), np.int32),
                                 _p(_ii(cols), np.int32), _p(_dd(A)), _p(D))
    assert bad == 0
    return D


def gmres(b, rowp, cols, A, rhs, Dinv=None, m=30, nrestart=10, rtol=1e-8,
          atol=1e-30, x0=None, max_its=0):
    n = rowp.size - 1
    x = np.zeros(b * n) if x0 is None else _dd(x0).copy()
    iters = ctypes.c_int(0)
    rn = ctypes.c_double(0.0)
    flag = lib().orc_gmres_capped(b, n, _p(_ii(rowp), np.int32), _p(_ii(cols), np.int32),
                                  _p(_dd(A)), _p(_dd(Dinv)) if Dinv is not None else None,
                                  _p(_dd(rhs)), _p(x), m, nrestart,
                                  ctypes.c_double(rtol), ctypes.c_double(atol),
                                  1 if x0 is None else 0, max_its, ctypes.byref(iters),
                                  ctypes.byref(rn))
    return x, flag, iters.value, rn.value


def compliance(order, conn, X, u, t_elem, C21):
    conn = _ii(conn).reshape(-1)
    return lib().orc_compliance(order, conn.size // order ** 3, _p(conn, np.int32),
                                _p(_dd(X)), _p(_dd(u)), _p(_dd(t_elem)),
                                _p(_dd(C21)))


def compliance_svsens(order, conn, X, u, t_elem, C21, alpha=1.0):
    conn = _ii(conn).reshape(-1)
    dfdu = np.zeros(np.asarray(u).size)
    lib().orc_compliance_svsens(order, conn.size // order ** 3,
                                _p(conn, np.int32), _p(_dd(X)), _p(_dd(u)),
                                _p(_dd(t_elem)), _p(_dd(C21)),
                                ctypes.c_double(alpha), _p(dfdu))
    return dfdu


def compliance_dvsens(order, conn, X, u, C21, scale=1.0):
    conn = _ii(conn).reshape(-1)
    nelem = conn.size // order ** 3
    out = np.zeros(nelem)
    lib().orc_compliance_dvsens(order, nelem, _p(conn, np.int32), _p(_dd(X)),
                                _p(_dd(u)), _p(_dd(C21)), ctypes.c_double(scale),
                                _p(out))
    return out


def adj_res_product(order, conn, X, u, psi, C21, scale=1.0):
    conn = _ii(conn).reshape(-1)
    nelem = conn.size // order ** 3
    out = np.zeros(nelem)
    lib().orc_adj_res_product(order, nelem, _p(conn, np.int32), _p(_dd(X)),
                              _p(_dd(u)), _p(_dd(psi)), _p(_dd(C21)),
                              ctypes.c_double(scale), _p(out))
    return out


def ks_failure(order, conn, X, u, C21, ys, ks_weight, alpha=1.0, safety=1.0):
    conn = _ii(conn).reshape(-1)
    out = np.zeros(3)
    lib().orc_ks_failure(order, conn.size // order ** 3, _p(conn, np.int32),
                         _p(_dd(X)), _p(_dd(u)), _p(_dd(C21)),
                         ctypes.c_double(ys), ctypes.c_double(ks_weight),
                         ctypes.c_double(alpha), ctypes.c_double(safety), _p(out))
    return out


def ks_svsens(order, conn, X, u, C21, ys, ks_weight, max_fail, ks_sum,
              alpha=1.0, safety=1.0):
    conn = _ii(conn).reshape(-1)
    dfdu = np.zeros(np.asarray(u).size)
    lib().orc_ks_svsens(order, conn.size // order ** 3, _p(conn, np.int32),
                        _p(_dd(X)), _p(_dd(u)), _p(_dd(C21)), ctypes.c_double(ys),
                        ctypes.c_double(ks_weight), ctypes.c_double(safety),
                        ctypes.c_double(max_fail), ctypes.c_double(ks_sum),
                        ctypes.c_double(alpha), _p(dfdu))
    return dfdu


# ---------------------------------------------------------------- shells
_c = ctypes.c_double


def quad4_strains(X, v, drill=True):
    e = np.zeros((4, 9))
    detw = np.zeros(4)
    lib().orc_quad4_strains(_p(_dd(X)), _p(_dd(v)), 1 if drill else 0, _p(e), _p(detw))
    return e, detw


def quad4_element(X, v, E, nu, t, alpha=1.0):
    res = np.zeros(24)
    mat = np.zeros(576)
    lib().orc_quad4_element(_p(_dd(X)), _p(_dd(v)), _c(E), _c(nu), _c(t), _c(alpha),
                            _p(res), _p(mat))
    return res, mat.reshape(24, 24)


def shell_set_ref_axis(axis):
    """Frame of the oracle's shell functions: TACSShellRefAxisTransform(axis),
    or the natural transform for None."""
    if axis is None:
        lib().orc_shell_set_ref_axis(None)
    else:
        a = np.ascontiguousarray(axis, dtype=np.float64)
        lib().orc_shell_set_ref_axis(_p(a))


def shell_set_offset(toff):
    """tOffset of the oracle's shell constitutive model (0: mid-surface)."""
    lib().orc_shell_set_offset(_c(toff))


def quad4_mass(X, rho, t, alpha=1.0):
    mat = np.zeros(576)
    lib().orc_quad4_mass(_p(_dd(X)), _c(rho), _c(t), _c(alpha), _p(mat))
    return mat.reshape(24, 24)


def quad4_adj_product(X, v, psi, E, nu, t):
    return lib().orc_quad4_adj_product(_p(_dd(X)), _p(_dd(v)), _p(_dd(psi)), _c(E),
                                       _c(nu), _c(t))


def quad4_failure(X, v, E, nu, ys, t, n):
    det = ctypes.c_double(0.0)
    f = lib().orc_quad4_failure(_p(_dd(X)), _p(_dd(v)), _c(E), _c(nu), _c(ys), _c(t), n,
                                ctypes.byref(det))
    return f, det.value


def quad4_pressure(X, p):
    res = np.zeros(24)
    lib().orc_quad4_pressure(_p(_dd(X)), _c(p), _p(res))
    return res


def shell_stress(E, nu, t, e):
    s = np.zeros(9)
    lib().orc_shell_stress(_c(E), _c(nu), _c(t), _p(_dd(e)), _p(s))
    return s


def shell_failure(E, nu, ys, t, e):
    de = np.zeros(9)
    dt = ctypes.c_double(0.0)
    f = lib().orc_shell_failure(_c(E), _c(nu), _c(ys), _c(t), _p(_dd(e)), _p(de),
                                ctypes.byref(dt))
    return f, de, dt.value


def assemble_shell(conn, X, u, t_elem, E, nu, rowp, cols):
    conn = _ii(conn).reshape(-1)
    nnodes = np.asarray(X).size // 3
    A = np.zeros(36 * cols.size)
    res = np.zeros(6 * nnodes)
    miss = lib().orc_assemble_shell(conn.size // 4, _p(conn, np.int32), _p(_dd(X)),
                                    _p(_dd(u)) if u is not None else None,
                                    _p(_dd(t_elem)), _c(E), _c(nu),
                                    _p(_ii(rowp), np.int32), _p(_ii(cols), np.int32),
                                    _p(A), _p(res))
    assert miss == 0
    return A, res


def assemble_shell_mat(conn, X, t_elem, E, nu, rho, kc, mc, rowp, cols, gc=0.0, u=None):
    """kc K + mc M + gc G(u) of Quad4 shells assembled into 6x6 block rows."""
    conn = _ii(conn).reshape(-1)
    A = np.zeros(36 * cols.size)
    miss = lib().orc_assemble_shell_mat(conn.size // 4, _p(conn, np.int32), _p(_dd(X)),
                                        _p(_dd(u)) if u is not None else None,
                                        _p(_dd(t_elem)), _c(E), _c(nu), _c(rho), _c(kc),
                                        _c(mc), _c(gc), _p(_ii(rowp), np.int32),
                                        _p(_ii(cols), np.int32), _p(A))
    assert miss == 0
    return A


def quad4_nl_element(X, u, E, nu, t):
    """Residual and tangent stiffness of the nonlinear shell model at u."""
    res = np.zeros(24)
    mat = np.zeros(576)
    lib().orc_quad4_nl_element(_p(_dd(X)), _p(_dd(u)), _c(E), _c(nu), _c(t), _c(1.0),
                               _p(res), _p(mat))
    return res, mat.reshape(24, 24)


def quad4_geo(X, p, E, nu, t):
    """Geometric stiffness of the linear shell element along the path p."""
    mat = np.zeros(576)
    lib().orc_quad4_geo(_p(_dd(X)), _p(_dd(p)), _c(E), _c(nu), _c(t), _c(1.0), _p(mat))
    return mat.reshape(24, 24)


def shell_add_pressure(conn, X, aux_elems, pressures, res):
    conn = _ii(conn).reshape(-1)
    aux = _ii(aux_elems)
    lib().orc_shell_add_pressure(aux.size, _p(aux, np.int32), _p(_dd(pressures)),
                                 _p(conn, np.int32), _p(_dd(X)), _p(res))
    return res


def shell_compliance(conn, X, u, t_elem, E, nu):
    conn = _ii(conn).reshape(-1)
    return lib().orc_shell_compliance(conn.size // 4, _p(conn, np.int32), _p(_dd(X)),
                                      _p(_dd(u)), _p(_dd(t_elem)), _c(E), _c(nu))


def shell_ks_failure(conn, X, u, t_elem, E, nu, ys, ks_weight, alpha=1.0, safety=1.0):
    conn = _ii(conn).reshape(-1)
    out = np.zeros(3)
    lib().orc_shell_ks_failure(conn.size // 4, _p(conn, np.int32), _p(_dd(X)), _p(_dd(u)),
                               _p(_dd(t_elem)), _c(E), _c(nu), _c(ys), _c(ks_weight),
                               _c(alpha), _c(safety), _p(out))
    return out


def shell_ks_sens(conn, X, u, t_elem, E, nu, ys, ks_weight, max_fail, ks_sum, safety=1.0):
    """(dKS/du, dKS/dt_e) of the continuous KS failure function."""
    conn = _ii(conn).reshape(-1)
    nelem = conn.size // 4
    dfdu = np.zeros(np.asarray(u).size)
    dfdt = np.zeros(nelem)
    lib().orc_shell_ks_sens(nelem, _p(conn, np.int32), _p(_dd(X)), _p(_dd(u)),
                            _p(_dd(t_elem)), _c(E), _c(nu), _c(ys), _c(ks_weight),
                            _c(safety), _c(max_fail), _c(ks_sum), _p(dfdu), _p(dfdt))
    return dfdu, dfdt


def shell_adj_res_product(conn, X, u, psi, t_elem, E, nu, scale=1.0):
    conn = _ii(conn).reshape(-1)
    nelem = conn.size // 4
    out = np.zeros(nelem)
    lib().orc_shell_adj_res_product(nelem, _p(conn, np.int32), _p(_dd(X)), _p(_dd(u)),
                                    _p(_dd(psi)), _p(_dd(t_elem)), _c(E), _c(nu),
                                    _c(scale), _p(out))
    return out


def shell_compliance_sens(conn, X, u, t_elem, E, nu):
    """(dC/du, dC/dt_e at fixed u) of the shell compliance."""
    conn = _ii(conn).reshape(-1)
    nelem = conn.size // 4
    dfdu = np.zeros(np.asarray(u).size)
    dfdt = np.zeros(nelem)
    lib().orc_shell_compliance_sens(nelem, _p(conn, np.int32), _p(_dd(X)), _p(_dd(u)),
                                    _p(_dd(t_elem)), _c(E), _c(nu), _p(dfdu), _p(dfdt))
    return dfdu, dfdt


# ---------------------------------------------------------------- node coordinates
def quad4_adj_xpt(X, v, psi, E, nu, t, scale=1.0):
    out = np.zeros(12)
    lib().orc_quad4_adj_xpt(_p(_dd(X)), _p(_dd(v)), _p(_dd(psi)), _c(E), _c(nu), _c(t),
                            _c(scale), _p(out))
    return out


def quad4_pressure_xpt(X, p, psi, scale=1.0):
    out = np.zeros(12)
    lib().orc_quad4_pressure_xpt(_p(_dd(X)), _c(p), _p(_dd(psi)), _c(scale), _p(out))
    return out


def quad4_point_xpt(quantity, X, v, E, nu, ys, t, dfddet, dfdq):
    """d/dX sum_q (dfddet_q detXd_q + dfdq_q q_q); quantity 2: strain energy
    density, 1: failure index."""
    out = np.zeros(12)
    lib().orc_quad4_point_xpt(int(quantity), _p(_dd(X)), _p(_dd(v)), _c(E), _c(nu), _c(ys),
                              _c(t), _p(_dd(dfddet)), _p(_dd(dfdq)), _p(out))
    return out


def shell_xpt_sens(mode, conn, X, u, t_elem, E, nu, psi=None, ys=270e6, ks_weight=0.0,
                   max_fail=0.0, ks_sum=1.0, safety=1.0, scale=1.0):
    """Node sensitivities of the shell mesh: mode 0 scale psi^T dR/dX, 1 the
    compliance, 2 the continuous KS failure (aggregate max and sum given)."""
    conn = _ii(conn).reshape(-1)
    out = np.zeros(np.asarray(X).size)
    lib().orc_shell_xpt_sens(int(mode), conn.size // 4, _p(conn, np.int32), _p(_dd(X)),
                             _p(_dd(u)), _p(_dd(psi)) if psi is not None else None,
                             _p(_dd(t_elem)), _c(E), _c(nu), _c(ys), _c(ks_weight),
                             _c(max_fail), _c(ks_sum), _c(safety), _c(scale), _p(out))
    return out


def shell_pressure_xpt(conn, X, aux_elems, pressures, psi, scale=1.0):
    conn = _ii(conn).reshape(-1)
    aux = _ii(aux_elems)
    out = np.zeros(np.asarray(X).size)
    lib().orc_shell_pressure_xpt(aux.size, _p(aux, np.int32), _p(_dd(pressures)),
                                 _p(conn, np.int32), _p(_dd(X)), _p(_dd(psi)), _c(scale),
                                 _p(out))
    return out


def compliance_xptsens(order, conn, X, u, t_elem, C21, scale=1.0):
    conn = _ii(conn).reshape(-1)
    X = _dd(X).reshape(-1)
    out = np.zeros(X.size)
    lib().orc_compliance_xptsens(order, conn.size // order ** 3, _p(conn, np.int32), _p(X),
                                 _p(_dd(u)), _p(_dd(t_elem)), _p(_dd(C21)), _c(scale), _p(out))
    return out


def adj_res_xpt_product(order, conn, X, u, psi, t_elem, C21, scale=1.0):
    conn = _ii(conn).reshape(-1)
    X = _dd(X).reshape(-1)
    out = np.zeros(X.size)
    lib().orc_adj_res_xpt_product(order, conn.size // order ** 3, _p(conn, np.int32), _p(X),
                                  _p(_dd(u)), _p(_dd(psi)), _p(_dd(t_elem)), _p(_dd(C21)),
                                  _c(scale), _p(out))
    return out


def traction_xpt_product(order, conn, X, aux_elems, aux_faces, trac, psi, scale=1.0):
    conn = _ii(conn).reshape(-1)
    X = _dd(X).reshape(-1)
    out = np.zeros(X.size)
    ae = _ii(aux_elems)
    lib().orc_traction_xpt_product(order, ae.size, _p(ae, np.int32), _p(_ii(aux_faces), np.int32),
                                   _p(_dd(trac).reshape(-1)), _p(conn, np.int32), _p(X),
                                   _p(_dd(psi)), _c(scale), _p(out))
    return out


def ks_xptsens(order, conn, X, u, C21, ys, ks_weight, max_fail, ks_sum, safety=1.0, scale=1.0):
    conn = _ii(conn).reshape(-1)
    X = _dd(X).reshape(-1)
    out = np.zeros(X.size)
    lib().orc_ks_xptsens(order, conn.size // order ** 3, _p(conn, np.int32), _p(X), _p(_dd(u)),
                         _p(_dd(C21)), _c(ys), _c(ks_weight), _c(safety), _c(max_fail),
                         _c(ks_sum), _c(scale), _p(out))
    return out


# ---------------------------------------------------------------- multi-core baseline
# cpu_baseline.c: the same algorithms over all host cores (OpenMP), for the
# cpu_baseline leg of bench.py; equal to the serial functions up to summation
# order.
def par_threads():
    return lib().orc_par_threads()


def color_elements(nnodes, conn, npe):
    """Greedy element colouring (no two elements of a colour share a node):
    (color_ptr, color_elems)."""
    conn = _ii(conn).reshape(-1)
    nelem = conn.size // npe
    cptr = np.zeros(65, np.int32)
    celems = np.zeros(max(nelem, 1), np.int32)
    nc = lib().orc_color_elements(nelem, npe, nnodes, _p(conn, np.int32), _p(cptr, np.int32),
                                  _p(celems, np.int32))
    assert nc > 0
    return cptr[:nc + 1].copy(), celems[:nelem]


def par_assemble_solid(order, colors, conn, X, u, t_elem, C21, rowp, cols, rho=0.0,
                       alpha=1.0, gamma=0.0, want_mat=True, want_res=True):
    cptr, celems = colors
    nnodes = rowp.size - 1
    A = np.zeros(cols.size * 9) if want_mat else None
    res = np.zeros(3 * nnodes) if want_res else None
    missing = lib().orc_par_assemble_solid(
        order, cptr.size - 1, _p(_ii(cptr), np.int32), _p(_ii(celems), np.int32),
        _p(_ii(conn).reshape(-1), np.int32), _p(_dd(X)),
        _p(_dd(u)) if u is not None else None, _p(_dd(t_elem)), _p(_dd(C21)),
        ctypes.c_double(rho), ctypes.c_double(alpha), ctypes.c_double(gamma),
        _p(_ii(rowp), np.int32), _p(_ii(cols), np.int32), _p(A), _p(res))
    assert missing == 0
    return A, res


def par_bsr_mult(b, rowp, cols, A, x):
    n = rowp.size - 1
    y = np.zeros(b * n)
    lib().orc_par_bsr_mult(b, n, _p(_ii(rowp), np.int32), _p(_ii(cols), np.int32),
                           _p(_dd(A)), _p(_dd(x)), _p(y))
    return y


def par_gmres(b, rowp, cols, A, rhs, Dinv=None, m=30, nrestart=10, rtol=1e-8,
              atol=1e-30, x0=None, max_its=0):
    n = rowp.size - 1
    x = np.zeros(b * n) if x0 is None else _dd(x0).copy()
    iters = ctypes.c_int(0)
    rn = ctypes.c_double(0.0)
    flag = lib().orc_par_gmres(b, n, _p(_ii(rowp), np.int32), _p(_ii(cols), np.int32),
                               _p(_dd(A)), _p(_dd(Dinv)) if Dinv is not None else None,
                               _p(_dd(rhs)), _p(x), m, nrestart, ctypes.c_double(rtol),
                               ctypes.c_double(atol), 1 if x0 is None else 0, max_its,
                               ctypes.byref(iters), ctypes.byref(rn))
    return x, flag, iters.value, rn.value


def par_assemble_shell(colors, conn, X, u, t_elem, E, nu, rowp, cols, want_mat=True,
                       want_res=True):
    cptr, celems = colors
    nnodes = rowp.size - 1
    A = np.zeros(cols.size * 36) if want_mat else None
    res = np.zeros(6 * nnodes) if want_res else None
    missing = lib().orc_par_assemble_shell(
        cptr.size - 1, _p(_ii(cptr), np.int32), _p(_ii(celems), np.int32),
        _p(_ii(conn).reshape(-1), np.int32), _p(_dd(X)), _p(_dd(u)) if u is not None else None,
        _p(_dd(t_elem)), ctypes.c_double(E), ctypes.c_double(nu), _p(_ii(rowp), np.int32),
        _p(_ii(cols), np.int32), _p(A), _p(res))
    assert missing == 0
    return A, res
