"""Mirror of tacs.TACS for the assembly / solve / adjoint path (tacs/TACS.pyx:
Vec :594, Mat :960, Pc :1106, KSM :1256, Assembler :1595, AuxElements :912).

Same class and method names, argument orders and defaults as the reference's
Cython wrappers; the objects live on one GPU.  Differences, by design:
``Pc(mat)`` builds the block-Jacobi preconditioner (the reference builds a
Schur-complement direct factorisation), and ``Assembler.create`` ignores its
communicator argument (single rank).
"""
import ctypes
import json
import os

import numpy as np

from ._lib import lib, check, TACSError
from ._obj import TACSHandle

dtype = np.float64
NORMAL = 0
TRANSPOSE = 1
CLASSICAL_GRAM_SCHMIDT = 0
MODIFIED_GRAM_SCHMIDT = 1
# ElementMatrixType (src/elements/TACSElementTypes.h:113, tacs/TACS.pyx:55-57)
JACOBIAN_MATRIX = 0
STIFFNESS_MATRIX = 1
MASS_MATRIX = 2
GEOMETRIC_STIFFNESS_MATRIX = 3


def _ptr_array(objs):
    arr = (ctypes.c_void_p * len(objs))(*[o.ptr for o in objs])
    return arr


class Vec(TACSHandle):
    def getSize(self):
        n, b = ctypes.c_int(), ctypes.c_int()
        lib().tacs_vec_get_size(self.ptr, ctypes.byref(n), ctypes.byref(b))
        return n.value

    def getBlockSize(self):
        n, b = ctypes.c_int(), ctypes.c_int()
        lib().tacs_vec_get_size(self.ptr, ctypes.byref(n), ctypes.byref(b))
        return b.value

    def getArray(self):
        """Host view of the vector (TACSBVec::getArray).  Writes through the
        view reach the device before every later device read, until a device
        operation writes the vector (solve output, axpy, scale, zeroEntries,
        assembly into it); after such a write the view is stale and getArray
        must be called again to see the device values."""
        n = self.getSize()
        p = lib().tacs_vec_get_array(self.ptr)
        if n == 0:
            return np.zeros(0)
        if not p:
            check(1, "Vec.getArray")
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_double)), shape=(n,))

    def getDevicePointer(self):
        """Raw device address of the values (device copy becomes current)."""
        return lib().tacs_vec_device_array(self.ptr)

    def copyFromDevice(self, ptr, n):
        """Copy n doubles from device address ptr into the vector."""
        check(lib().tacs_vec_copy_from_device(self.ptr, ptr, n), "Vec.copyFromDevice")

    def copyToDevice(self, ptr, n):
        """Copy the vector into n doubles at device address ptr."""
        check(lib().tacs_vec_copy_to_device(self.ptr, ptr, n), "Vec.copyToDevice")

    def norm(self):
        return lib().tacs_vec_norm(self.ptr)

    def dot(self, vec):
        return lib().tacs_vec_dot(self.ptr, vec.ptr)

    def copyValues(self, vec):
        check(lib().tacs_vec_copy_values(self.ptr, vec.ptr), "Vec.copyValues")

    def scale(self, alpha):
        lib().tacs_vec_scale(self.ptr, alpha)

    def axpy(self, alpha, vec):
        lib().tacs_vec_axpy(self.ptr, alpha, vec.ptr)

    def axpby(self, alpha, beta, vec):
        lib().tacs_vec_axpby(self.ptr, alpha, beta, vec.ptr)

    def zeroEntries(self):
        lib().tacs_vec_zero_entries(self.ptr)

    def set(self, val):
        lib().tacs_vec_set(self.ptr, val)

    def setRand(self, lower=-1.0, upper=1.0):
        lib().tacs_vec_set_rand(self.ptr, lower, upper)


class Mat(TACSHandle):
    def mult(self, x, y):
        check(lib().tacs_mat_mult(self.ptr, x.ptr, y.ptr), "Mat.mult")

    def zeroEntries(self):
        lib().tacs_mat_zero_entries(self.ptr)

    def prepareProducts(self, prec=64):
        """Build the products' sliced ELL copy of the current values (FP64 or
        FP32) now instead of at the first product."""
        check(lib().tacs_mat_prepare_products(self.ptr, int(prec)), "Mat.prepareProducts")

    def autotune(self, reps=20):
        """Times every launch variant of the sliced ELL product of each kind
        (0: FP64, 1: FP32 values with FP64 vectors, 2: FP32) on this matrix
        and keeps the fastest per kind (all variants give identical
        products).  Returns {"ms": [[per variant] per kind], "chosen": [...]}.
        No counterpart in the reference."""
        maxv = 8
        ms = np.zeros(3 * maxv)
        nv = ctypes.c_int()
        check(lib().tacs_mat_autotune(self.ptr, int(reps), ms.ctypes.data, maxv, ctypes.byref(nv)),
              "Mat.autotune")
        return {"ms": ms.reshape(3, maxv)[:, :nv.value].tolist(),
                "chosen": [self.getSellVariant(k) for k in range(3)]}

    def getSellVariant(self, kind):
        v = ctypes.c_int()
        check(lib().tacs_mat_get_sell_variant(self.ptr, int(kind), ctypes.byref(v)),
              "Mat.getSellVariant")
        return v.value

    def setSellVariant(self, kind, variant):
        """Force the product variant of a kind (-1: the tuning table's)."""
        check(lib().tacs_mat_set_sell_variant(self.ptr, int(kind), int(variant)),
              "Mat.setSellVariant")

    def autotuneSpMM(self, nrhs, prec=32, reps=10):
        """Times the variants of the nrhs right-hand-side block product and
        enters the fastest in the tuning table (size class of this matrix);
        returns the times per variant."""
        ms = np.zeros(8)
        nv = ctypes.c_int()
        check(lib().tacs_mat_autotune_spmm(self.ptr, int(nrhs), int(prec), int(reps),
                                           ms.ctypes.data, 8, ctypes.byref(nv)),
              "Mat.autotuneSpMM")
        return ms[:nv.value].tolist()

    def getInfo(self):
        b, n, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_long()
        lib().tacs_mat_get_info(self.ptr, ctypes.byref(b), ctypes.byref(n), ctypes.byref(nnz))
        return b.value, n.value, nnz.value

    def getArrays(self):
        """Host copies (rowp, cols, A[nnzb, b, b]) of the block-CSR matrix
        (BCSRMat::getArrays)."""
        b, n, nnz = self.getInfo()
        rp, cp, ap = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        lib().tacs_mat_get_arrays(self.ptr, ctypes.byref(rp), ctypes.byref(cp), ctypes.byref(ap))
        I = ctypes.POINTER(ctypes.c_int)
        D = ctypes.POINTER(ctypes.c_double)
        rowp = np.ctypeslib.as_array(ctypes.cast(rp, I), shape=(n + 1,)).copy()
        cols = np.ctypeslib.as_array(ctypes.cast(cp, I), shape=(nnz,)).copy()
        A = np.ctypeslib.as_array(ctypes.cast(ap, D), shape=(nnz * b * b,)).copy()
        return rowp, cols, A.reshape(nnz, b, b)

    def getDeviceArrays(self):
        rp, cp, ap = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        lib().tacs_mat_device_arrays(self.ptr, ctypes.byref(rp), ctypes.byref(cp),
                                     ctypes.byref(ap))
        return rp.value, cp.value, ap.value


class MatrixFreeMat(TACSHandle):
    """TACSMatrixFreeMat (src/TACSMatrixFreeMat.h): the solid elements'
    Jacobian (or stiffness / mass matrix) applied element by element from
    quadrature-point data, with no assembled matrix.  ``mult`` returns K x
    with the Dirichlet rows zeroed; it is also a KSM operator."""

    def __init__(self, assembler):
        self.assembler = assembler
        super().__init__(lib().tacs_matfree_create(assembler.ptr), keep=(assembler,))

    def assembleMatrixFreeData(self, matType=JACOBIAN_MATRIX, alpha=1.0, beta=0.0, gamma=0.0):
        check(lib().tacs_matfree_assemble_data(self.ptr, int(matType), alpha, beta, gamma),
              "MatrixFreeMat.assembleMatrixFreeData")

    def mult(self, x, y):
        check(lib().tacs_mat_mult(self.ptr, x.ptr, y.ptr), "MatrixFreeMat.mult")

    def multTranspose(self, x, y):
        check(lib().tacs_matfree_mult_transpose(self.ptr, x.ptr, y.ptr),
              "MatrixFreeMat.multTranspose")

    def multDevice(self, x_ptr, y_ptr):
        """y = K x on device arrays (raw pointers) of the assembler's length."""
        check(lib().tacs_matfree_mult_device(self.ptr, x_ptr, y_ptr), "MatrixFreeMat.multDevice")


class Pc(TACSHandle):
    """Preconditioner of an assembled matrix: block-Jacobi (the inverse
    diagonal blocks, BCSRMat::factorDiag src/bpmat/BCSRMat.cpp:1527).

    The reference's TACS.Pc (tacs/TACS.pyx:1106) builds a Schur-complement
    (TACSSchurPc) or additive-Schwarz ILU(k) factorisation from lev_fill,
    fill and reorder_schur; those factorisations are not in this build, so
    its arguments are accepted and a warning says that block-Jacobi runs
    (stronger preconditioners change the iteration counts, not the
    solutions, at a given tolerance)."""

    _warned = False

    def __init__(self, mat, lev_fill=None, fill=None, reorder_schur=None, **kw):
        if not Pc._warned and (lev_fill is not None or fill is not None or kw):
            import warnings

            warnings.warn("TACS.Pc: block-Jacobi preconditioner (the reference's Schur / "
                          "ILU(k) factorisations are not part of this build)", stacklevel=2)
            Pc._warned = True
        self.mat = mat
        super().__init__(lib().tacs_block_jacobi_pc_create(mat.ptr), keep=(mat,))

    def factor(self):
        check(lib().tacs_pc_factor(self.ptr), "Pc.factor")

    def applyFactor(self, x, y):
        lib().tacs_pc_apply(self.ptr, x.ptr, y.ptr)

    def setChebyshev(self, degree, ratio=20.0, values=32):
        """Chebyshev polynomial (degree - 1 operator products per
        application) of the block-Jacobi iteration as the preconditioner of
        the recycling and batched solvers, on [lmax / ratio, lmax] of the
        eigenvalues of D^{-1} A (lmax estimated at the next factor(): 1.05 x
        the largest Ritz value of 30 Lanczos steps).
        degree 1: plain block-Jacobi.  values=16: reduced precision inside
        the polynomial of FP32-operator solvers: single systems read a
        bfloat16 copy of the matrix values (the polynomial stays linear);
        the batched solver with 16 or more right-hand sides (3x3 blocks)
        reads an FP32 copy of the iterate (not exactly linear; the solver's
        FP64 restarts absorb it).  No counterpart in the reference."""
        check(lib().tacs_pc_set_chebyshev(self.ptr, int(degree), float(ratio), int(values)),
              "Pc.setChebyshev")

    def getChebyshev(self):
        """(degree, ratio, lmax): lmax as the last factor() estimated it."""
        d, r, l = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        check(lib().tacs_pc_get_chebyshev(self.ptr, ctypes.byref(d), ctypes.byref(r),
                                          ctypes.byref(l)), "Pc.getChebyshev")
        return d.value, r.value, l.value


class KSM(TACSHandle):
    """GMRES(m) with right preconditioning (GMRES::solve src/bpmat/KSM.cpp:785)."""

    def __init__(self, mat, pc, m, restarts=0, isFlexible=0):
        super().__init__(lib().tacs_gmres_create(mat.ptr, pc.ptr if pc else None, m, restarts,
                                                 isFlexible), keep=(mat, pc))

    def solve(self, b, x, zero_guess=1):
        return lib().tacs_ksm_solve(self.ptr, b.ptr, x.ptr, zero_guess)

    def setTolerances(self, rtol, atol):
        lib().tacs_ksm_set_tolerances(self.ptr, rtol, atol)

    def setCycleGraph(self, flag=True):
        """Run each restart cycle's Arnoldi steps as one captured HIP graph
        (assembled operator): no host round trip per step, the same iterates
        (the first step that meets the tolerance is taken after the cycle)."""
        check(lib().tacs_gmres_set_cycle_graph(self.ptr, int(bool(flag))), "KSM.setCycleGraph")

    def setOrthoType(self, ortho):
        check(lib().tacs_gmres_set_ortho_type(self.ptr, ortho), "KSM.setOrthoType")

    def setMonitor(self, comm=None, descript="GMRES", freq=10, filename=None):
        """Residual monitor (KSM.setMonitor tacs/TACS.pyx:1326, KSMPrintStdout
        src/bpmat/KSM.cpp:261): prints "descript[iter]: residual" at every
        restart and every freq-th step, on rank 0 of comm (None: rank 0).
        filename: write the lines to that file instead (KSMPrintFile)."""
        rank = comm.rank if comm is not None else 0
        d = descript.encode()
        if filename is None:
            mon = TACSHandle(lib().tacs_ksm_print_stdout_create(d, rank, int(freq)))
        else:
            mon = TACSHandle(lib().tacs_ksm_print_file_create(str(filename).encode(), d, rank,
                                                              int(freq)))
        check(lib().tacs_ksm_set_monitor(self.ptr, mon.ptr), "setMonitor")

    def getIterCount(self):
        return lib().tacs_ksm_get_iter_count(self.ptr)

    def getResidualNorm(self):
        return lib().tacs_ksm_get_res_norm(self.ptr)


class MixedKSM(KSM):
    """Mixed-precision restarted GMRES(m) with block-Jacobi: every cycle
    restarts from the FP64 residual b - A x; its Arnoldi steps use the FP32
    matrix, preconditioner and basis with FP64 Gram-Schmidt dot products and
    Hessenberg; the correction is accumulated in FP64.  No counterpart in the
    reference.

    deflation=k > 0 turns on Krylov subspace recycling (GCRO-DR): each cycle
    keeps the k harmonic Ritz vectors with the smallest harmonic Ritz values,
    and the space carries over to later solves with the same matrix."""

    def __init__(self, mat, pc, m=30, restarts=1000, deflation=0):
        TACSHandle.__init__(self, lib().tacs_mixed_gmres_create(mat.ptr, pc.ptr, m, restarts),
                            keep=(mat, pc))
        if deflation:
            self.setDeflation(deflation)

    def setDeflation(self, k):
        check(lib().tacs_mixed_gmres_set_deflation(self.ptr, int(k)), "MixedKSM.setDeflation")

    def getOuterCount(self):
        return lib().tacs_mixed_gmres_get_outer_count(self.ptr)


class BatchedKSM(TACSHandle):
    """Batched restarted GMRES(m) with block-Jacobi for up to nrhs (<= 64)
    right-hand sides of one 3x3-block matrix: every Arnoldi step of all the
    systems shares one multi-RHS SpMM; each system keeps its own Hessenberg,
    stopping test and iteration count.  prec=64 runs the Arnoldi steps on the
    FP64 matrix, prec=32 on its FP32-rounded values (FP64 vectors); every
    cycle restarts from the FP64 residual.  No counterpart in the reference,
    which runs one KSM.solve per load case."""

    def __init__(self, mat, pc, m=30, restarts=100, nrhs=16, prec=64, deflation=0):
        super().__init__(lib().tacs_batched_gmres_create(mat.ptr, pc.ptr, m, restarts, nrhs,
                                                         prec), keep=(mat, pc))
        self.nrhs = nrhs
        if deflation:
            self.setDeflation(deflation)

    def setDeflation(self, k):
        """Krylov subspace recycling per right-hand side (GCRO-DR): each system
        keeps a k-dimensional space, updated every cycle from its harmonic
        Ritz vectors and kept across solves with the same matrix values."""
        check(lib().tacs_batched_gmres_set_deflation(self.ptr, int(k)), "BatchedKSM.setDeflation")

    def setTolerances(self, rtol, atol):
        lib().tacs_batched_gmres_set_tolerances(self.ptr, rtol, atol)

    def setMonitor(self, comm=None, descript="GMRES", freq=10, filename=None):
        """Residual monitor (KSM.setMonitor tacs/TACS.pyx:1326; here the true residual of each
        system at the start of every cycle, KSMPrintStdout
        src/bpmat/KSM.cpp:261): prints "descript[iter]: residual" at every
        restart and every freq-th step, on rank 0 of comm (None: rank 0).
        filename: write the lines to that file instead (KSMPrintFile)."""
        rank = comm.rank if comm is not None else 0
        d = descript.encode()
        if filename is None:
            mon = TACSHandle(lib().tacs_ksm_print_stdout_create(d, rank, int(freq)))
        else:
            mon = TACSHandle(lib().tacs_ksm_print_file_create(str(filename).encode(), d, rank,
                                                              int(freq)))
        check(lib().tacs_ksm_set_monitor(self.ptr, mon.ptr), "setMonitor")

    def setDelayedReorthogonalization(self, flag=True):
        """Gram-Schmidt of the recycling Arnoldi steps: delayed classical
        Gram-Schmidt (DCGS2, the default: two sweeps over the basis per step)
        or, flag False, classical Gram-Schmidt twice (four sweeps)."""
        check(lib().tacs_batched_gmres_set_delayed_reorth(self.ptr, int(bool(flag))),
              "BatchedKSM.setDelayedReorthogonalization")

    def setRecycleSpace(self, ksm):
        """Deflate every right-hand side by the recycled space (U, C) that the
        MixedKSM ksm (with deflation, same matrix) holds after a solve: each
        cycle starts with x += M^{-1} U C^T r, r -= C C^T r and its Arnoldi
        steps run on (I - C C^T) A M^{-1}.  Returns the space's dimension."""
        k = lib().tacs_batched_gmres_set_recycle_space(self.ptr, ksm.ptr)
        self._raise_on_error(k, "BatchedKSM.setRecycleSpace")
        return k

    def solve(self, bs, xs, zero_guess=1):
        """xs[j] = A^{-1} bs[j]; returns 1 when every system converged."""
        if len(bs) != len(xs):
            raise ValueError("BatchedKSM.solve: as many solutions as right-hand sides")
        arr = ctypes.c_void_p * len(bs)
        flag = lib().tacs_batched_gmres_solve(self.ptr, len(bs), arr(*[b.ptr for b in bs]),
                                              arr(*[x.ptr for x in xs]), zero_guess)
        self._raise_on_error(flag, "BatchedKSM.solve")
        self._last = len(bs)
        return flag

    @staticmethod
    def _raise_on_error(flag, what):
        # negative: rejected arguments; a message without a negative flag: a
        # kernel launch or HIP call failed during the solve
        msg = lib().tacs_last_error()
        if flag < 0 or msg:
            raise TACSError(f"{what} failed (code {flag}): {msg.decode() if msg else ''}")

    def solveBlock(self, nrhs, B_ptr, X_ptr, zero_guess=1):
        """Solve on device block vectors (B[i*nrhs + j], X[i*nrhs + j])."""
        flag = lib().tacs_batched_gmres_solve_block(self.ptr, nrhs, B_ptr, X_ptr, zero_guess)
        self._raise_on_error(flag, "BatchedKSM.solveBlock")
        self._last = nrhs
        return flag

    def getStats(self):
        """(iterations, flags, initial residual norms, final residual norms,
        cycles) of the last solve, per right-hand side."""
        n = getattr(self, "_last", self.nrhs)
        it = np.zeros(n, np.int32)
        fl = np.zeros(n, np.int32)
        r0 = np.zeros(n)
        rn = np.zeros(n)
        cyc = lib().tacs_batched_gmres_get_stats(self.ptr, n, it.ctypes.data, fl.ctypes.data,
                                                 r0.ctypes.data, rn.ctypes.data)
        return it, fl, r0, rn, cyc


class SegmentedKSM(BatchedKSM):
    """Independent systems on the diagonal blocks of one block-diagonal 3x3
    matrix (the meshes of a multi-mesh assembler): block j is the node range
    [node_off[j], node_off[j+1]) and is solved as its own system, with its
    own Krylov space, recycled space (deflation > 0, GCRO-DR), Hessenberg
    matrix and stopping test, all advancing together (one batched step
    serves every block; a converged block stops counting steps).  At most 64
    blocks; rows coupling two blocks are refused.  No counterpart in the
    reference, which builds one assembler and one GMRES per mesh."""

    def __init__(self, mat, pc, node_off, m=60, restarts=100000, prec=32, deflation=20):
        off = np.ascontiguousarray(node_off, dtype=np.int32)
        TACSHandle.__init__(self, lib().tacs_batched_gmres_create_segments(
            mat.ptr, pc.ptr, m, restarts, off.size - 1, off.ctypes.data, prec), keep=(mat, pc))
        self.nrhs = off.size - 1
        if deflation:
            self.setDeflation(deflation)

    def solve(self, b, x, zero_guess=1):
        """x = A^{-1} b block by block; returns 1 when every block converged
        (getStats: per block)."""
        flag = lib().tacs_batched_gmres_solve_segments(self.ptr, b.ptr, x.ptr, zero_guess)
        self._raise_on_error(flag, "SegmentedKSM.solve")
        self._last = self.nrhs
        return flag


class RecyclingKSM:
    """Krylov subspace recycling (GCRO-DR with delayed classical Gram-Schmidt)
    for one right-hand side through the KSM interface (solve, setTolerances,
    getIterCount): the batched solver with one right-hand side, on the FP64
    operator (prec=64) or its FP32-rounded values with FP64 vectors
    (prec=32).  The recycled space carries over to later solves with the same
    matrix values (the adjoint solve).  No counterpart in the reference."""

    def __init__(self, mat, pc, m=60, restarts=100000, deflation=20, prec=64):
        self.batched = BatchedKSM(mat, pc, m=m, restarts=restarts, nrhs=1, prec=prec,
                                  deflation=deflation)

    def setTolerances(self, rtol, atol):
        self.batched.setTolerances(rtol, atol)

    def setMonitor(self, *args, **kw):
        self.batched.setMonitor(*args, **kw)

    def solve(self, b, x, zero_guess=1):
        return self.batched.solve([b], [x], zero_guess)

    def getIterCount(self):
        return int(self.batched.getStats()[0][0])


class AuxElements(TACSHandle):
    def __init__(self, size=100):
        super().__init__(lib().tacs_aux_elements_create(size))
        self._elems = []

    def addElement(self, num, elem):
        self._elems.append(elem)
        check(lib().tacs_aux_elements_add(self.ptr, int(num), elem.ptr), "AuxElements.addElement")


class Assembler(TACSHandle):
    @staticmethod
    def create(comm, varsPerNode, numOwnedNodes, numElements, numDependentNodes=0):
        return Assembler(varsPerNode, numOwnedNodes, numElements, numDependentNodes)

    def __init__(self, varsPerNode, numOwnedNodes, numElements, numDependentNodes=0):
        super().__init__(lib().tacs_assembler_create(varsPerNode, numOwnedNodes, numElements,
                                                     numDependentNodes))
        self._elements = None
        self._aux = None

    # -- mesh definition -------------------------------------------------
    def setElementConnectivity(self, ptr, conn):
        ptr = np.ascontiguousarray(ptr, dtype=np.intc)
        conn = np.ascontiguousarray(conn, dtype=np.intc)
        if ptr.size - 1 != self.getNumElements():
            raise ValueError("Connectivity must match number of elements")
        check(lib().tacs_assembler_set_element_connectivity(self.ptr, ptr.ctypes.data,
                                                            conn.ctypes.data),
              "setElementConnectivity")

    def setElements(self, elements):
        if len(elements) != self.getNumElements():
            raise ValueError("Element list must match number of elements")
        self._elements = list(elements)
        check(lib().tacs_assembler_set_elements(self.ptr, _ptr_array(self._elements)),
              "setElements")

    def addBCs(self, nodes, _vars=None, values=None):
        nodes = np.ascontiguousarray(nodes, dtype=np.intc)
        nb = -1
        vp = vv = None
        if _vars is not None:
            _vars = np.ascontiguousarray(_vars, dtype=np.intc)
            nb = _vars.size
            vp = _vars.ctypes.data
        if values is not None:
            values = np.ascontiguousarray(values, dtype=np.float64)
            vv = values.ctypes.data
        check(lib().tacs_assembler_add_bcs(self.ptr, nodes.size, nodes.ctypes.data, nb, vp, vv),
              "addBCs")

    def initialize(self):
        check(lib().tacs_assembler_initialize(self.ptr), "initialize")

    def setElementPrecision(self, prec):
        """Arithmetic of the element kernels of assembleJacobian: 64 or 32
        (FP32 element matrices summed into the FP64 block-CSR values;
        residuals, functions and sensitivities stay FP64).  No counterpart in
        the reference."""
        check(lib().tacs_assembler_set_element_precision(self.ptr, int(prec)),
              "setElementPrecision")

    def setJacobianKernel(self, kernel):
        """Solid Jacobian kernel: 0 measured choice, 1 node-centric rows, 2
        element blocks + row gather."""
        check(lib().tacs_assembler_set_jacobian_kernel(self.ptr, int(kernel)), "setJacobianKernel")

    def getJacobianKernel(self):
        """The solid Jacobian kernel assembleJacobian launches (1 or 2)."""
        return lib().tacs_assembler_jacobian_kernel(self.ptr)

    def autotune(self, mat, reps=3):
        """Times the solid Jacobian's launch choices (rows-kernel variants,
        and the element-block form when the Jacobian kernel setting is 0) by
        assembling alpha = 1 Jacobians into mat, and keeps the fastest.
        Returns the times per choice.  No counterpart in the reference."""
        ms = np.zeros(8)
        nc = ctypes.c_int()
        check(lib().tacs_assembler_autotune(self.ptr, mat.ptr, int(reps), ms.ctypes.data, 8,
                                            ctypes.byref(nc)), "Assembler.autotune")
        return ms[:nc.value].tolist()

    def getJacobianVariant(self):
        v = ctypes.c_int()
        check(lib().tacs_assembler_get_jacobian_variant(self.ptr, ctypes.byref(v)),
              "getJacobianVariant")
        return v.value

    def setJacobianVariant(self, variant):
        """Force the Jacobian launch choice (-1: the tuning table's)."""
        check(lib().tacs_assembler_set_jacobian_variant(self.ptr, int(variant)),
              "setJacobianVariant")

    def setAuxElements(self, elems=None):
        self._aux = elems
        check(lib().tacs_assembler_set_aux_elements(self.ptr, elems.ptr if elems else None),
              "setAuxElements")

    # -- sizes -------------------------------------------------------------
    def _info(self):
        v = [ctypes.c_int() for _ in range(4)]
        lib().tacs_assembler_get_info(self.ptr, *[ctypes.byref(x) for x in v])
        return [x.value for x in v]

    def getVarsPerNode(self):
        return self._info()[0]

    def getNumOwnedNodes(self):
        return self._info()[1]

    def getNumNodes(self):
        return self._info()[1]

    def getNumElements(self):
        return self._info()[2]

    def getNumDesignVars(self):
        return self._info()[3]

    # -- vectors and matrices ---------------------------------------------
    def createVec(self):
        return Vec(lib().tacs_assembler_create_vec(self.ptr))

    def createNodeVec(self):
        return Vec(lib().tacs_assembler_create_node_vec(self.ptr))

    def createDesignVec(self):
        return Vec(lib().tacs_assembler_create_design_vec(self.ptr))

    def createMat(self):
        return Mat(lib().tacs_assembler_create_mat(self.ptr))

    def setNodes(self, X):
        check(lib().tacs_assembler_set_nodes(self.ptr, X.ptr), "setNodes")

    def getNodes(self, X):
        check(lib().tacs_assembler_get_nodes(self.ptr, X.ptr), "getNodes")

    def setVariables(self, vec, dvec=None, ddvec=None):
        check(lib().tacs_assembler_set_variables(self.ptr, vec.ptr), "setVariables")

    def getVariables(self, vec, dvec=None, ddvec=None):
        check(lib().tacs_assembler_get_variables(self.ptr, vec.ptr), "getVariables")

    def zeroVariables(self):
        lib().tacs_assembler_zero_variables(self.ptr)

    def setDesignVars(self, vec):
        check(lib().tacs_assembler_set_design_vars(self.ptr, vec.ptr), "setDesignVars")

    def getDesignVars(self, vec):
        check(lib().tacs_assembler_get_design_vars(self.ptr, vec.ptr), "getDesignVars")

    def applyBCs(self, obj):
        if isinstance(obj, Mat):
            lib().tacs_assembler_apply_bcs_mat(self.ptr, obj.ptr)
        else:
            lib().tacs_assembler_apply_bcs_vec(self.ptr, obj.ptr)

    # -- the hot path -------------------------------------------------------
    def assembleRes(self, residual, loadScale=1.0, applyBCs=True):
        check(lib().tacs_assembler_assemble_res(self.ptr, residual.ptr, loadScale,
                                                1 if applyBCs else 0), "assembleRes")

    def assembleJacobian(self, alpha, beta, gamma, residual, A, matOr=NORMAL, loadScale=1.0,
                         applyBCs=True):
        check(lib().tacs_assembler_assemble_jacobian(
            self.ptr, alpha, beta, gamma, residual.ptr if residual is not None else None, A.ptr,
            matOr, loadScale, 1 if applyBCs else 0), "assembleJacobian")

    def assembleMatType(self, matType, A, matOr=NORMAL, loadScale=1.0, applyBCs=True):
        """Stiffness, mass or geometric stiffness matrix (Assembler.assembleMatType
        tacs/TACS.pyx:2228); Dirichlet rows zeroed with a unit diagonal."""
        check(lib().tacs_assembler_assemble_mat_type(self.ptr, int(matType), A.ptr, matOr,
                                                      loadScale, 1 if applyBCs else 0),
              "assembleMatType")

    def assembleMatCombo(self, matType1, scale1, matType2, scale2, A, matOr=NORMAL,
                         loadScale=1.0, applyBCs=True):
        """A = scale1 matType1 + scale2 matType2 (Assembler.assembleMatCombo
        tacs/TACS.pyx:2256)."""
        types = np.array([matType1, matType2], dtype=np.intc)
        scales = np.array([scale1, scale2], dtype=np.float64)
        check(lib().tacs_assembler_assemble_mat_combo(self.ptr, 2, types.ctypes.data,
                                                       scales.ctypes.data, A.ptr, matOr,
                                                       loadScale, 1 if applyBCs else 0),
              "assembleMatCombo")

    def getMatrixFreeDataSize(self, matType=JACOBIAN_MATRIX):
        """(data_size, temp_size) in doubles (TACSAssembler::getMatrixFreeDataSize)."""
        d, t = ctypes.c_int(), ctypes.c_int()
        check(lib().tacs_assembler_get_matrix_free_data_size(self.ptr, int(matType),
                                                             ctypes.byref(d), ctypes.byref(t)),
              "getMatrixFreeDataSize")
        return d.value, t.value

    def assembleMatrixFreeData(self, matType, alpha, beta, gamma, data_ptr):
        """Fill the device array at data_ptr (TACSAssembler::assembleMatrixFreeData)."""
        check(lib().tacs_assembler_assemble_matrix_free_data(self.ptr, int(matType), alpha, beta,
                                                             gamma, data_ptr),
              "assembleMatrixFreeData")

    def addMatrixFreeVecProduct(self, matType, data_ptr, temp_ptr, x, y, matOr=NORMAL,
                                loadScale=1.0, applyBCs=True):
        """y += K x from the device data (TACSAssembler::addMatrixFreeVecProduct)."""
        check(lib().tacs_assembler_add_matrix_free_vec_product(
            self.ptr, int(matType), data_ptr, temp_ptr, x.ptr, y.ptr, matOr, loadScale,
            1 if applyBCs else 0), "addMatrixFreeVecProduct")

    def evalFunctions(self, funclist):
        vals = np.zeros(len(funclist))
        check(lib().tacs_assembler_eval_functions(self.ptr, len(funclist), _ptr_array(funclist),
                                                  vals.ctypes.data), "evalFunctions")
        return vals

    def addDVSens(self, funclist, dfdxlist, alpha=1.0):
        check(lib().tacs_assembler_add_dv_sens(self.ptr, alpha, len(funclist),
                                               _ptr_array(funclist), _ptr_array(dfdxlist)),
              "addDVSens")

    def addSVSens(self, funclist, dfdulist, alpha=1.0, beta=0.0, gamma=0.0):
        check(lib().tacs_assembler_add_sv_sens(self.ptr, alpha, beta, gamma, len(funclist),
                                               _ptr_array(funclist), _ptr_array(dfdulist)),
              "addSVSens")

    def addAdjointResProducts(self, adjlist, dfdxlist, alpha=1.0, loadScale=1.0):
        check(lib().tacs_assembler_add_adjoint_res_products(
            self.ptr, alpha, len(adjlist), _ptr_array(adjlist), _ptr_array(dfdxlist), loadScale),
            "addAdjointResProducts")

    def addXptSens(self, funclist, dfdXlist, alpha=1.0):
        check(lib().tacs_assembler_add_xpt_sens(self.ptr, alpha, len(funclist),
                                                _ptr_array(funclist), _ptr_array(dfdXlist)),
              "addXptSens")

    def addAdjointResXptSensProducts(self, adjlist, dfdXlist, alpha=1.0, loadScale=1.0):
        check(lib().tacs_assembler_add_adjoint_res_xpt_sens_products(
            self.ptr, alpha, len(adjlist), _ptr_array(adjlist), _ptr_array(dfdXlist), loadScale),
            "addAdjointResXptSensProducts")


# ---------------------------------------------------------------------------
# Tuning table: measured launch variants per (kernel, shape, size class)
# ---------------------------------------------------------------------------
TUNE_SELL_SPMV, TUNE_SPMM, TUNE_JACOBIAN = 0, 1, 2
TUNING_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_table.json")


def setTuning(kernel, shape, size_class, variant):
    """Enter a launch variant in the library's tuning table (-1 restores the
    built-in default)."""
    check(lib().tacs_tune_set(int(kernel), int(shape), int(size_class), int(variant)),
          "setTuning")


def getTuning(kernel, shape, size_class):
    return lib().tacs_tune_get(int(kernel), int(shape), int(size_class))


def loadTuningTable(path=TUNING_TABLE):
    """Enter every entry of a table written by tacs_amd/autotune.py; returns
    the number of entries (0 when the file does not exist)."""
    if not os.path.exists(path):
        return 0
    with open(path) as f:
        table = json.load(f)
    n = 0
    for e in table.get("entries", []):
        setTuning(e["kernel_id"], e["shape"], e["size_class"], e["variant"])
        n += 1
    return n


def resetTuning():
    lib().tacs_tune_reset()
