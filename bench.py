"""Benchmark of the FE assembly / solve / adjoint path on one MI355X.

Workload (the headline configuration): a 128 x 128 x 128 cube of 8-node
linear hexahedra (2,097,152 elements, 2,146,689 nodes, 6,440,067 DOF, 3x3
block-CSR), E = 70e9, nu = 0.3, rho = 2700, x = 0 clamped, traction
(0, 0, 1e6) Pa on x = 1, one thickness design variable per element.

One step = the whole hot path on that input: assembleJacobian (FP64 values
plus the FP32 copy) + assembleRes, the mixed-precision solve to 1e-8
relative residual (GMRES with Krylov subspace recycling over the fixed
FP32-valued operator A32 D32^{-1}, FP64 vectors and Gram-Schmidt, FP64 true
residual at every restart), compliance, addSVSens, the adjoint solve (same
solver, reusing the recycled space), addDVSens and addAdjointResProducts.
value = elements / step time (Melem/s).

The plain FP64 GPU path (GMRES(30) with modified Gram-Schmidt -- the
reference's default orthogonalisation -- and block-Jacobi, all FP64) runs the
same step once outside the timed region for the mixed-vs-FP64 comparison and
the gradient error.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class HipEvents:
    """HIP events on the library's stream (the null stream), created through
    the HIP runtime the library is bound to."""

    def __init__(self):
        from tacs_amd._lib import hip_runtime
        self.hip = hip_runtime()
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float),
                                                 ctypes.c_void_p, ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e):
        assert self.hip.hipEventRecord(e, ctypes.c_void_p(0)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value


def spmv_traffic(kernel):
    """HBM bytes per launch of a product kernel at 128^3 from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE runs, scaled by a
    copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "spmv_traffic_hex8_128cube.json")
    try:
        return json.load(open(path))["kernels"][kernel]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def spmv_bytes(b, nrows, nnzb, word):
    """Algorithmic bytes of y = A x: values, column indices, row pointers, x
    and y once each."""
    return word * b * b * nnzb + 4 * nnzb + 4 * (nrows + 1) + 2 * word * b * nrows


def time_spmv(L, ev, p, reps=50):
    """Average duration of the matrix's products y = A x as the solvers run
    them (HIP events on the library's stream): FP64 values and vectors, FP32
    values with FP64 vectors (the recycling solver's operator) and FP32
    values and vectors.  For 3x3 blocks these are the sliced ELL kernels,
    for 6x6 blocks the block-CSR kernels."""
    b, nr, nnz = p.mat.getInfo()
    x = p.assembler.createVec()
    x.set(1.0)
    y = p.assembler.createVec()
    xd, yd = x.getDevicePointer(), y.getDevicePointer()
    n = b * nr
    hip = ev.hip
    x32, y32 = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(x32), ctypes.c_size_t(4 * n)) == 0
    assert hip.hipMalloc(ctypes.byref(y32), ctypes.c_size_t(4 * n)) == 0
    L.tacs_hip_convert_f64_f32(n, xd, x32, None)
    res = {}
    for key, pv, pw, xa, ya in (("fp64", 64, 64, xd, yd), ("mixed", 32, 64, xd, yd),
                                ("fp32", 32, 32, x32, y32)):
        for _ in range(5):
            assert L.tacs_mat_mult_prec(p.mat.ptr, pv, pw, xa, ya) == 0
        e0, e1 = ev.event(), ev.event()
        ev.record(e0)
        for _ in range(reps):
            L.tacs_mat_mult_prec(p.mat.ptr, pv, pw, xa, ya)
        ev.record(e1)
        res[key + "_ms"] = ev.elapsed_ms(e0, e1) / reps
    for q in (x32, y32):
        hip.hipFree(q)
    res["fp64_bytes"] = spmv_bytes(b, nr, nnz, 8)
    res["fp32_bytes"] = spmv_bytes(b, nr, nnz, 4)
    # FP32 values (and column indices, row pointers) with FP64 x and y
    res["mixed_bytes"] = 4 * b * b * nnz + 4 * nnz + 4 * (nr + 1) + 2 * 8 * n
    return res


def time_cheb(L, ev, p, pc, reps=50, values=16):
    """Average duration of the fused Chebyshev product of the recycling
    solver's preconditioner in the three-term form the polynomial's steps
    after the second run (values in bfloat16 or FP32, FP64 vectors; HIP
    events on the library's stream) and its algorithmic bytes: the values
    and column indices once, row pointers, z gathered once, rhs and zprev
    read, zout written (over zprev), the FP32 inverse diagonal blocks."""
    b, nr, nnz = p.mat.getInfo()
    vecs = [p.assembler.createVec() for _ in range(3)]
    for k, v in enumerate(vecs):
        v.set(1.0 + 0.1 * k)
    z, zp, rhs = [v.getDevicePointer() for v in vecs]
    for _ in range(5):
        assert L.tacs_mat_cheb_step3(p.mat.ptr, pc.ptr, values, z, zp, rhs, 0.5, 0.1) == 0
    e0, e1 = ev.event(), ev.event()
    ev.record(e0)
    for _ in range(reps):
        L.tacs_mat_cheb_step3(p.mat.ptr, pc.ptr, values, z, zp, rhs, 0.5, 0.1)
    ev.record(e1)
    ms = ev.elapsed_ms(e0, e1) / reps
    word = 2 if values == 16 else 4
    nbytes = word * b * b * nnz + 4 * nnz + 4 * (nr + 1) + 8 * b * nr * 4 + 4 * b * b * nr
    return ms, nbytes


def time_fill(L, ev, p, reps=5):
    """Average duration of the sliced ELL fill of the matrix writing all
    three copies (FP64, FP32, bfloat16) and the columns in one launch (HIP
    events on the library's stream), and its algorithmic bytes: the BCSR
    values and columns read once, two row pointers per row, every slot of
    the three copies and the columns written once."""
    b, nr, nnz = p.mat.getInfo()
    rpp, cpp, app = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    assert L.tacs_mat_get_arrays(p.mat.ptr, ctypes.byref(rpp), ctypes.byref(cpp),
                                 ctypes.byref(app)) == 0
    rowp = np.ctypeslib.as_array(ctypes.cast(rpp, ctypes.POINTER(ctypes.c_int)),
                                 shape=(nr + 1,)).copy()
    ns = (nr + 63) // 64
    width = np.zeros(ns, np.int32)
    coff = np.zeros(ns + 1, np.int64)
    slots = L.tacs_hip_sell_slots(nr, rowp.ctypes.data, width.ctypes.data, coff.ctypes.data)
    b2 = b * b
    sizes = {"w": 4 * ns, "o": 8 * (ns + 1), "cs": 4 * slots, "v64": 8 * b2 * slots,
             "v32": 4 * b2 * slots, "v16": 2 * b2 * slots}
    d = {}
    for k, n in sizes.items():
        d[k] = ctypes.c_void_p()
        assert ev.hip.hipMalloc(ctypes.byref(d[k]), ctypes.c_size_t(n)) == 0
    ev.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert ev.hip.hipMemcpy(d["w"], width.ctypes.data, sizes["w"], 1) == 0
    assert ev.hip.hipMemcpy(d["o"], coff.ctypes.data, sizes["o"], 1) == 0
    rp, cp, ap = p.mat.getDeviceArrays()
    fill = lambda: L.tacs_hip_sell_fill3(b, nr, rp, cp, d["w"], d["o"], ap, d["v64"], d["v32"],
                                         d["v16"], d["cs"], None)
    for _ in range(2):
        assert fill() == 0
    e0, e1 = ev.event(), ev.event()
    ev.record(e0)
    for _ in range(reps):
        fill()
    ev.record(e1)
    ms = ev.elapsed_ms(e0, e1) / reps
    for k in d:
        ev.hip.hipFree(d[k])
    nbytes = (8 * b2 + 4) * nnz + 8 * nr + (14 * b2 + 4) * slots
    return {"fill_ms": ms, "fill_bytes": nbytes, "fill_GBps": nbytes / ms / 1e6,
            "fill_slots": slots}


MONITOR = {"file": None, "freq": 1000}


CHEB = {"degree": 1, "ratio": 300.0, "values": 16}


def make_solver(p, solver, m, rtol, deflation=0):
    """mixed: the recycling solver on the FP32-valued operator; recycle64: the
    same recycling on the FP64 operator; fp64: GMRES(m) + MGS.  The recycling
    solvers take the Chebyshev preconditioner of --cheb-degree (1: plain
    block-Jacobi, as GMRES)."""
    from tacs_amd import TACS
    pc = TACS.Pc(p.mat)
    if solver in ("mixed", "recycle64") and CHEB["degree"] > 1:
        pc.setChebyshev(CHEB["degree"], CHEB["ratio"], CHEB["values"])
    if solver == "mixed":
        ksm = TACS.MixedKSM(p.mat, pc, m, 100000, deflation=deflation)
    elif solver == "recycle64":
        ksm = TACS.RecyclingKSM(p.mat, pc, m, 100000, deflation, prec=64)
    else:
        ksm = TACS.KSM(p.mat, pc, m, 100000)
    ksm.setTolerances(rtol, 1e-30)
    if MONITOR["file"]:
        # residual lines (every cycle start of the recycling solver, every
        # restart and freq-th step of GMRES) to a file, for long solves
        ksm.setMonitor(descript=f"{solver}", freq=MONITOR["freq"],
                       filename=f"{MONITOR['file']}.{solver}")
    return pc, ksm


def step_work(p):
    """Vectors and the function object of run_step, allocated once before
    the timed region."""
    from tacs_amd import functions
    a = p.assembler
    return dict(comp=functions.Compliance(a), dfdu=a.createVec(), psi=a.createVec(),
                dfdx=a.createDesignVec())


def run_step(p, pc, ksm, work, want_grad=False):
    """One pass of the hot path; returns timings and results."""
    from tacs_amd import lib
    L = lib()
    a = p.assembler
    comp, dfdu, psi, dfdx = work["comp"], work["dfdu"], work["psi"], work["dfdx"]
    t = {}
    t0 = time.perf_counter()
    a.zeroVariables()
    a.assembleJacobian(1.0, 0.0, 0.0, p.res, p.mat)
    L.tacs_device_synchronize()
    t["assembly_s"] = time.perf_counter() - t0
    t1 = time.perf_counter()
    pc.factor()
    flag_f = ksm.solve(p.res, p.u)
    its_f = ksm.getIterCount()
    p.u.scale(-1.0)
    a.setVariables(p.u)
    c = a.evalFunctions([comp])[0]
    dfdu.zeroEntries()
    a.addSVSens([comp], [dfdu])
    a.applyBCs(dfdu)
    flag_a = ksm.solve(dfdu, psi)
    its_a = ksm.getIterCount()
    dfdx.zeroEntries()
    a.addDVSens([comp], [dfdx], 1.0)
    a.addAdjointResProducts([psi], [dfdx], -1.0)
    L.tacs_device_synchronize()
    t["solve_adjoint_s"] = time.perf_counter() - t1
    t["step_s"] = time.perf_counter() - t0
    out = dict(t=t, compliance=c, its=(its_f, its_a), flags=(flag_f, flag_a))
    if want_grad:
        out["grad"] = dfdx.getArray().copy()
        out["u"] = p.u.getArray().copy()
    return out


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


def cpu_env():
    """Host CPU model and the threads the OpenMP baseline runs on."""
    O = _oracle()
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, O.par_threads()


def _timed(fn):
    t0 = time.perf_counter()
    out = fn()
    return time.perf_counter() - t0, out


def cpu_solid_fraction(p, t_elem, fp64_its, name, E=70e9, nu=0.3, gmres_steps=120, nrhs=1):
    """Measured multi-core CPU baseline on the same workload: the oracle's
    OpenMP path (oracle/cpu_baseline.c) runs, on all host cores, the whole
    assembly (coloured element loop, block-CSR scatter, traction loads, BC
    rows), the block-Jacobi factorisation and gmres_steps real Arnoldi steps
    of FP64 GMRES(30) + MGS on the full matrix (whole restart cycles, so the
    per-step cost averages the cycle's growing Gram-Schmidt).  The step
    estimate adds the measured per-step time times the FP64 GPU path's step
    count (forward + adjoint, per load case)."""
    O = _oracle()
    model, threads = cpu_env()
    m = p.mesh
    X = m.X.reshape(-1)
    C = O.iso_C21(E, nu)
    rowp, cols = O.node_csr(m.nnodes, m.conn, m.npe)
    t_col, colors = _timed(lambda: O.color_elements(m.nnodes, m.conn, m.npe))
    t_asm, (A, _) = _timed(lambda: O.par_assemble_solid(m.order, colors, m.conn, X, None, t_elem,
                                                        C, rowp, cols, want_res=False))
    res = np.zeros(3 * m.nnodes)
    trac = np.tile(p.traction, (p.aux_elems.size, 1))
    t_tr, _ = _timed(lambda: O.add_tractions(m.order, m.conn, X, p.aux_elems, p.aux_faces, trac,
                                             res))
    bcv = np.full(p.bc_nodes.size, 7, np.int32)
    t_bc, _ = _timed(lambda: (O.apply_bc_mat(3, p.bc_nodes, bcv, rowp, cols, A),
                              O.apply_bc_vec(3, p.bc_nodes, bcv, res)))
    t_pc, D = _timed(lambda: O.bsr_diag_inv(3, rowp, cols, A))
    steps = max(30, 30 * (gmres_steps // 30))
    t_gm, out = _timed(lambda: O.par_gmres(3, rowp, cols, A, -res, D, m=30, nrestart=100000,
                                           rtol=1e-30, max_its=steps))
    t_step = t_gm / out[2]
    its = int(sum(fp64_its))
    asm_s = t_col + t_asm + t_tr + t_bc
    step_s = asm_s + t_pc + nrhs * its * t_step
    measured_s = asm_s + t_pc + t_gm
    res_d = {
        "cores": threads, "kind": "port", "cpu_model": model,
        "sample": (f"oracle OpenMP path (oracle/cpu_baseline.c) on {threads} threads of {model}, "
                   f"the {name} workload: the whole assembly ({asm_s:.2f} s incl. colouring, "
                   f"tractions, BC rows) and block-Jacobi ({t_pc:.2f} s) run in full, then "
                   f"{out[2]} real FP64 GMRES(30)+MGS steps on the full matrix "
                   f"({1e3 * t_step:.1f} ms/step); the step is that plus the measured per-step "
                   f"time x the {its} steps" + (f" per load case x {nrhs} load cases" if nrhs > 1
                                                else "") +
                   f" the GPU FP64 path took ({measured_s:.1f} s of CPU work measured)"),
        "assembly_s": asm_s, "pc_factor_s": t_pc, "gmres_step_s": t_step,
        "gmres_steps_measured": out[2], "step_s_estimate": step_s,
    }
    del A
    if nrhs > 1:
        res_d.update(value=nrhs / step_s, unit="load cases/s")
    else:
        res_d.update(value=m.nelem / step_s / 1e6, unit="Melem/s")
    return res_d


def cpu_shell_fraction(p, t_elem, its, name, E=70e9, nu=0.3, gmres_steps=120):
    """Measured multi-core CPU baseline of the shell workload: the whole
    assembly (coloured Quad4 element loop, 6x6 block scatter, pressure loads,
    BC rows) and block-Jacobi on all host cores, then gmres_steps real FP64
    GMRES(30)+MGS steps on the full matrix; the step adds the measured
    per-step time x the GPU step count `its` (a lower bound for GMRES(30),
    which needs more steps than the recycling solver)."""
    O = _oracle()
    model, threads = cpu_env()
    m = p.mesh
    X = m.X.reshape(-1)
    rowp, cols = O.node_csr(m.nnodes, m.conn, 4)
    colors = O.color_elements(m.nnodes, m.conn, 4)
    t0 = time.perf_counter()
    A, res = O.par_assemble_shell(colors, m.conn, X, None, t_elem, E, nu, rowp, cols)
    O.shell_add_pressure(m.conn, X, p.aux_elems, np.full(p.aux_elems.size, p.pressure), res)
    mask = 0
    for v in p.bc_vars:
        mask |= 1 << v
    bm = np.full(p.bc_nodes.size, mask, np.int32)
    O.apply_bc_mat(6, p.bc_nodes, bm, rowp, cols, A)
    O.apply_bc_vec(6, p.bc_nodes, bm, res)
    asm_s = time.perf_counter() - t0
    t_pc, D = _timed(lambda: O.bsr_diag_inv(6, rowp, cols, A))
    steps = max(30, 30 * (gmres_steps // 30))
    t_gm, out = _timed(lambda: O.par_gmres(6, rowp, cols, A, -res, D, m=30, nrestart=100000,
                                           rtol=1e-30, max_its=steps))
    t_step = t_gm / out[2]
    total = int(sum(its))
    step_s = asm_s + t_pc + total * t_step
    return {"cores": threads, "kind": "port", "cpu_model": model,
            "sample": (f"oracle OpenMP path on {threads} threads of {model}, the {name} "
                       f"workload: the whole assembly ({asm_s:.2f} s) and block-Jacobi "
                       f"({t_pc:.2f} s) in full, {out[2]} real FP64 GMRES(30)+MGS steps "
                       f"({1e3 * t_step:.1f} ms/step); the step adds the per-step time x the "
                       f"{total} steps of the GPU recycling solver (GMRES(30) needs more, so "
                       f"this is a lower bound of the CPU time)"),
            "assembly_s": asm_s, "pc_factor_s": t_pc, "gmres_step_s": t_step,
            "step_s_estimate": step_s, "value": m.nelem / step_s / 1e6, "unit": "Melem/s"}


def cpu_multi(p, fp64_its, gmres_steps=30):
    """Measured multi-core CPU baseline of the many-cube workload, cube by
    cube as the reference runs them (one assembler and one GMRES per mesh):
    every cube's whole assembly and block-Jacobi on all host cores, and
    gmres_steps real FP64 GMRES(30)+MGS steps on its matrix; each cube adds
    its measured per-step time x its FP64 GPU step count (forward solve)."""
    O = _oracle()
    model, threads = cpu_env()
    C = O.iso_C21(70e9, 0.3)
    total, measured = 0.0, 0.0
    for c, m in enumerate(p.meshes):
        X = m.X.reshape(-1)
        rowp, cols = O.node_csr(m.nnodes, m.conn, 8)
        colors = O.color_elements(m.nnodes, m.conn, 8)
        t0 = time.perf_counter()
        A, _ = O.par_assemble_solid(2, colors, m.conn, X, None, np.ones(m.nelem), C, rowp, cols,
                                    want_res=False)
        res = np.zeros(3 * m.nnodes)
        ae, af = m.elems_on_x(1)
        O.add_tractions(2, m.conn, X, ae, af, np.tile(p.traction, (ae.size, 1)), res)
        bn = m.nodes_on_x(0)
        bv = np.full(bn.size, 7, np.int32)
        O.apply_bc_mat(3, bn, bv, rowp, cols, A)
        O.apply_bc_vec(3, bn, bv, res)
        D = O.bsr_diag_inv(3, rowp, cols, A)
        t_asm = time.perf_counter() - t0
        t_gm, out = _timed(lambda: O.par_gmres(3, rowp, cols, A, -res, D, m=30, nrestart=100000,
                                               rtol=1e-30, max_its=gmres_steps))
        total += t_asm + fp64_its[c] * t_gm / out[2]
        measured += t_asm + t_gm
    return {"cores": threads, "kind": "port", "cpu_model": model,
            "sample": (f"oracle OpenMP path on {threads} threads of {model}, cube by cube: every "
                       f"cube's whole assembly and block-Jacobi plus {gmres_steps} real FP64 "
                       f"GMRES(30)+MGS steps ({measured:.1f} s measured); each cube adds its "
                       f"per-step time x its GPU FP64 GMRES(30) step count"),
            "step_s_estimate": total, "value": len(p.meshes) / total, "unit": "cubes/s"}


def cfg1_cube20(L, rtol=1e-8):
    """The 20^3 cube of the reference's CPU-runnable case, run in full on the
    GPU (FP64 GMRES(30) + MGS, the reference's settings) and by the oracle's
    OpenMP path on all host cores (no extrapolation): assembly, solve,
    compliance, adjoint solve, node gradient."""
    from tacs_amd import functions
    from tacs_amd.problems import SolidProblem
    O = _oracle()
    model, threads = cpu_env()
    p = SolidProblem(20)
    a, m = p.assembler, p.mesh
    comp = functions.Compliance(a)
    dfdu, psi, g = a.createVec(), a.createVec(), a.createNodeVec()

    def gpu_step():
        t0 = time.perf_counter()
        a.zeroVariables()
        a.assembleJacobian(1.0, 0.0, 0.0, p.res, p.mat)
        flag, its, pc, ksm = p.solve_fp64(m=30, restarts=100000, rtol=rtol)
        c = a.evalFunctions([comp])[0]
        dfdu.zeroEntries()
        a.addSVSens([comp], [dfdu])
        a.applyBCs(dfdu)
        flag2 = ksm.solve(dfdu, psi)
        its2 = ksm.getIterCount()
        g.zeroEntries()
        a.addXptSens([comp], [g])
        a.addAdjointResXptSensProducts([psi], [g], -1.0)
        L.tacs_device_synchronize()
        return time.perf_counter() - t0, c, (its, its2), (flag, flag2)

    gpu_step()
    t_gpu, c_gpu, its_gpu, flags = gpu_step()
    g_gpu = g.getArray().copy()
    # the CPU: the same step by the oracle's OpenMP path
    X = m.X.reshape(-1)
    t = np.ones(m.nelem)
    C = O.iso_C21(70e9, 0.3)
    # the pattern and the colouring are set-up, as the GPU's initialize()
    rowp, cols = O.node_csr(m.nnodes, m.conn, 8)
    colors = O.color_elements(m.nnodes, m.conn, 8)
    t0 = time.perf_counter()
    A, _ = O.par_assemble_solid(2, colors, m.conn, X, None, t, C, rowp, cols, want_res=False)
    res = np.zeros(3 * m.nnodes)
    trac = np.tile(p.traction, (p.aux_elems.size, 1))
    O.add_tractions(2, m.conn, X, p.aux_elems, p.aux_faces, trac, res)
    bcv = np.full(p.bc_nodes.size, 7, np.int32)
    O.apply_bc_mat(3, p.bc_nodes, bcv, rowp, cols, A)
    O.apply_bc_vec(3, p.bc_nodes, bcv, res)
    D = O.bsr_diag_inv(3, rowp, cols, A)
    t_asm = time.perf_counter() - t0
    u, f1, it1, _ = O.par_gmres(3, rowp, cols, A, -res, D, m=30, nrestart=100000, rtol=rtol)
    c_cpu = O.compliance(2, m.conn, X, u, t, C)
    dc = O.compliance_svsens(2, m.conn, X, u, t, C)
    O.apply_bc_vec(3, p.bc_nodes, bcv, dc)
    ps, f2, it2, _ = O.par_gmres(3, rowp, cols, A, dc, D, m=30, nrestart=100000, rtol=rtol)
    g_cpu = (O.compliance_xptsens(2, m.conn, X, u, t, C)
             - O.adj_res_xpt_product(2, m.conn, X, u, ps, t, C)
             - O.traction_xpt_product(2, m.conn, X, p.aux_elems, p.aux_faces, trac, ps))
    t_cpu = time.perf_counter() - t0
    return {
        "workload": "linear hex cube 20^3 (8,000 elements, 27,783 DOF), E=70e9 nu=0.3, x=0 "
                    "clamped, traction 1e6 Pa on x=1: assembly, FP64 GMRES(30)+MGS+block-Jacobi "
                    f"to {rtol:g}, compliance, adjoint solve, node gradient",
        "gpu_step_s": t_gpu, "gpu_iterations": list(its_gpu), "gpu_flags": list(flags),
        "cpu_step_s": t_cpu, "cpu_assembly_s": t_asm, "cpu_iterations": [it1, it2],
        "cpu_flags": [f1, f2], "cpu_cores": threads, "cpu_model": model,
        "gpu_vs_cpu": t_cpu / t_gpu,
        "compliance_rel_diff": abs(c_gpu - c_cpu) / abs(c_cpu),
        "node_gradient_rel_diff": float(np.abs(g_gpu - g_cpu).max() / np.abs(g_cpu).max()),
    }


def shell_step(p, pc, ksm, ks_weight=100.0, want_grad=False):
    """One pass of the shell path: assembly, solve, KS failure, its adjoint
    w.r.t. the element thicknesses."""
    from tacs_amd import functions, lib
    L = lib()
    a = p.assembler
    t = {}
    t0 = time.perf_counter()
    a.zeroVariables()
    a.assembleJacobian(1.0, 0.0, 0.0, p.res, p.mat)
    L.tacs_device_synchronize()
    t["assembly_s"] = time.perf_counter() - t0
    t1 = time.perf_counter()
    pc.factor()
    flag_f = ksm.solve(p.res, p.u)
    its_f = ksm.getIterCount()
    p.u.scale(-1.0)
    a.setVariables(p.u)
    ks = functions.KSFailure(a, ksWeight=ks_weight)
    val = a.evalFunctions([ks])[0]
    dfdu = a.createVec()
    a.addSVSens([ks], [dfdu])
    a.applyBCs(dfdu)
    psi = a.createVec()
    flag_a = ksm.solve(dfdu, psi)
    its_a = ksm.getIterCount()
    dfdx = a.createDesignVec()
    a.addDVSens([ks], [dfdx], 1.0)
    a.addAdjointResProducts([psi], [dfdx], -1.0)
    L.tacs_device_synchronize()
    t["solve_adjoint_s"] = time.perf_counter() - t1
    t["step_s"] = time.perf_counter() - t0
    out = dict(t=t, ks=val, its=(its_f, its_a), flags=(flag_f, flag_a))
    if want_grad:
        out["grad"] = dfdx.getArray().copy()
    return out


def shell_main(args, L, ev):
    """Quad4 shell plate (512^2 by default): assembly, mixed solve and KS
    adjoint w.r.t. thickness; kernel timings.  Prints its own JSON line (not
    the headline metric)."""
    from tacs_amd.problems import ShellProblem
    n = args.n if args.n != 128 else 512
    t0 = time.perf_counter()
    p = ShellProblem(n, dv="element", pressure=1e4, bc="ss", t0=0.01)
    rng = np.random.default_rng(7)
    t_elem = rng.uniform(0.005, 0.02, p.mesh.nelem)
    p.set_thickness(t_elem)
    log(f"setup {time.perf_counter() - t0:.1f}s: {p.mesh.nelem} shells, {6 * p.mesh.nnodes} DOF")
    p.assemble()
    e0, e1 = ev.event(), ev.event()
    ev.record(e0)
    for _ in range(5):
        p.assembler.assembleJacobian(1.0, 0.0, 0.0, None, p.mat)
    ev.record(e1)
    asm_ms = ev.elapsed_ms(e0, e1) / 5
    # the same assembly with FP32 element matrices (summed in FP64)
    p.assembler.setElementPrecision(32)
    e0, e1 = ev.event(), ev.event()
    ev.record(e0)
    for _ in range(5):
        p.assembler.assembleJacobian(1.0, 0.0, 0.0, None, p.mat)
    ev.record(e1)
    asm32_ms = ev.elapsed_ms(e0, e1) / 5
    p.assembler.setElementPrecision(64)
    p.assemble()
    sp = time_spmv(L, ev, p)
    nelem = p.mesh.nelem
    kern = {"assembly_ms": asm_ms, "assembly_Melem_s": nelem / (asm_ms * 1e-3) / 1e6,
            "assembly_fp32_elements_ms": asm32_ms,
            "spmv6_fp64": {"avg_ms": sp["fp64_ms"], "bytes": sp["fp64_bytes"],
                           "GBps": sp["fp64_bytes"] / sp["fp64_ms"] / 1e6},
            "spmv6_fp32": {"avg_ms": sp["fp32_ms"], "bytes": sp["fp32_bytes"],
                           "GBps": sp["fp32_bytes"] / sp["fp32_ms"] / 1e6}}
    if args.kernels_only:
        kern.update(elements=nelem, dof=6 * p.mesh.nnodes, nnz_blocks=p.mat.getInfo()[2])
        print(json.dumps(kern))
        return
    pc, ksm = make_solver(p, "mixed", args.m, args.rtol, args.deflation)
    for w in range(args.warmup):
        shell_step(p, pc, ksm)
    steps = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        steps.append(shell_step(p, pc, ksm, want_grad=(k == args.steps - 1)))
        log(f"step {k}: {steps[-1]['t']} its={steps[-1]['its']} flags={steps[-1]['flags']}")
    wall = time.perf_counter() - t0
    # the same step with FP32 element matrices, against the FP64-element step
    fp32e = None
    if not args.no_fp64:
        p.assembler.setElementPrecision(32)
        pc32, ksm32 = make_solver(p, "mixed", args.m, args.rtol, args.deflation)
        fp32e = shell_step(p, pc32, ksm32, want_grad=True)
        p.assembler.setElementPrecision(64)
        log(f"fp32-element step: {fp32e['t']} its={fp32e['its']} flags={fp32e['flags']}")
        del pc32, ksm32
    ke_bytes = nelem * 576 * 8
    out = {
        "workload": f"Quad4 shell plate {n}x{n}, t ~ U[0.005, 0.02], E=70e9 nu=0.3, simply "
                    f"supported (u=v=w=0 on the edges), pressure 1e4 Pa; KS failure (rho=100) "
                    f"adjoint w.r.t. thickness; mixed GMRES({args.m}) recycling "
                    f"{args.deflation}, block-Jacobi, rtol {args.rtol:g}",
        "elements": nelem, "dof": 6 * p.mesh.nnodes, "nnz_blocks": p.mat.getInfo()[2],
        "ms_per_step": 1e3 * wall / args.steps,
        "Melem_per_s_step": nelem / (wall / args.steps) / 1e6,
        "assembly_ms": asm_ms, "assembly_Melem_s": nelem / (asm_ms * 1e-3) / 1e6,
        "assembly_bytes_elem_matrices": ke_bytes,
        "iterations": list(steps[-1]["its"]), "flags": list(steps[-1]["flags"]),
        "solve_adjoint_s": steps[-1]["t"]["solve_adjoint_s"], "ks": steps[-1]["ks"],
        "spmv6_fp64": {"avg_ms": sp["fp64_ms"], "bytes": sp["fp64_bytes"],
                       "GBps": sp["fp64_bytes"] / sp["fp64_ms"] / 1e6,
                       "frac": sp["fp64_bytes"] / sp["fp64_ms"] / 1e6 / 8000.0},
        "spmv6_fp32": {"avg_ms": sp["fp32_ms"], "bytes": sp["fp32_bytes"],
                       "GBps": sp["fp32_bytes"] / sp["fp32_ms"] / 1e6,
                       "frac": sp["fp32_bytes"] / sp["fp32_ms"] / 1e6 / 8000.0},
        "assembly_fp32_elements_ms": asm32_ms,
        "assembly_fp32_elements_Melem_s": nelem / (asm32_ms * 1e-3) / 1e6,
    }
    if fp32e is not None:
        g64 = steps[-1]["grad"]
        out["fp32_elements"] = {
            "solve_adjoint_s": fp32e["t"]["solve_adjoint_s"], "iterations": list(fp32e["its"]),
            "flags": list(fp32e["flags"]),
            "ks_rel_err_vs_fp64_elements": abs(fp32e["ks"] - steps[-1]["ks"]) / abs(steps[-1]["ks"]),
            "ks_grad_rel_err_vs_fp64_elements": float(np.abs(fp32e["grad"] - g64).max()
                                                      / np.abs(g64).max()),
            "note": "the same step with the element matrices computed in FP32 (summed into "
                    "FP64 block-CSR values); the errors are against the FP64-element step, "
                    "whose elements match the FP64 oracle to 1e-10"}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_shell_fraction(p, t_elem, steps[-1]["its"], f"{n}^2 Quad4")
    print(json.dumps(out))


def batched_loads(p, nrhs, seed=5):
    """Load cases on the x = 1 face: at every face node a seeded normal
    traction (mean 0, std 1e6 Pa per component), lumped to a nodal force
    (traction x face area / number of face nodes); zero on the clamped face."""
    a = p.assembler
    rng = np.random.default_rng(seed)
    face = p.mesh.nodes_on_x(1)
    vecs = []
    for _ in range(nrhs):
        v = a.createVec()
        arr = v.getArray()
        arr[:] = 0.0
        arr.reshape(-1, 3)[face] = 1e6 * rng.standard_normal((face.size, 3)) / face.size
        a.applyBCs(v)
        vecs.append(v)
    return vecs


def batched_step(p, pc, ksm, bs, xs):
    """Assembly, the batched solve of every load case and, per load case, the
    compliance and its gradient w.r.t. the element thicknesses (compliance is
    self-adjoint: psi = 2 u, so no adjoint solve)."""
    from tacs_amd import functions, lib
    L = lib()
    a = p.assembler
    t = {}
    t0 = time.perf_counter()
    a.zeroVariables()
    a.assembleJacobian(1.0, 0.0, 0.0, p.res, p.mat)
    L.tacs_device_synchronize()
    t["assembly_s"] = time.perf_counter() - t0
    t1 = time.perf_counter()
    pc.factor()
    flag = ksm.solve(bs, xs)
    L.tacs_device_synchronize()
    t["solve_s"] = time.perf_counter() - t1
    it = ksm.getStats()[0]
    comp = functions.Compliance(a)
    vals, grads = [], []
    psi = a.createVec()
    for x in xs:
        a.setVariables(x)
        vals.append(a.evalFunctions([comp])[0])
        g = a.createDesignVec()
        a.addDVSens([comp], [g], 1.0)
        psi.copyValues(x)
        psi.scale(2.0)
        a.addAdjointResProducts([psi], [g], -1.0)
        grads.append(g)
    L.tacs_device_synchronize()
    t["step_s"] = time.perf_counter() - t0
    return dict(t=t, flag=flag, its=[int(v) for v in it], compliance=vals,
                grads=[g.getArray().copy() for g in grads])


def simp_main(args, L, ev):
    """SIMP compliance on the n^3 hex cube (96^3 by default): one density DV
    per element, x ~ U[0.1, 1] (PCG64 seed 0), t = Emin + x^p (E0 - Emin)
    (p = 3, E0 = 1, Emin = 1e-9, nu = 0.3), x = 0 clamped, unit traction in z
    on the centred 8 x 8 element patch of the x = 1 face.  A step is the
    forward (assembly, recycling mixed solve, compliance) and backward
    (adjoint solve, addDVSens - addAdjointResProducts, chain rule) of the
    torch.autograd.Function.  The plain FP64 GMRES(30) path runs the same
    step once for the gradient comparison.  --fd K adds central differences
    (h = 1e-6, both sides solved to 1e-14) on K seeded DVs.  Prints its own
    JSON line."""
    import torch
    from tacs_amd.autograd import SolidComplianceProblem, simp_compliance
    from tacs_amd.problems import SolidProblem
    n = args.n if args.n != 128 else 96
    t0 = time.perf_counter()
    prob = SolidProblem(n, E=1.0, nu=0.3, traction=(0.0, 0.0, 1.0), dv="element",
                        traction_patch=8)
    rng = np.random.default_rng(0)
    x0 = rng.uniform(0.1, 1.0, prob.mesh.nelem)
    log(f"setup {time.perf_counter() - t0:.1f}s: {prob.mesh.nelem} elements, "
        f"{3 * prob.mesh.nnodes} DOF")
    sp = SolidComplianceProblem(prob, p=3.0, E0=1.0, Emin=1e-9, rtol=args.rtol, m=args.m,
                                deflation=args.deflation, restarts=2000,
                                chebyshev=(CHEB["degree"], CHEB["ratio"], CHEB["values"])
                                if CHEB["degree"] > 1 else None)
    if MONITOR["file"]:
        sp.ksm.setMonitor(descript="simp", freq=MONITOR["freq"], filename=MONITOR["file"] + ".simp")

    def step(spx):
        x = torch.tensor(x0, dtype=torch.float64, device="cuda", requires_grad=True)
        t = time.perf_counter()
        c = simp_compliance(x, spx)
        c.backward()
        torch.cuda.synchronize()
        return time.perf_counter() - t, c.item(), x.grad.detach().cpu().numpy()

    for w in range(args.warmup):
        dt, c, g = step(sp)
        log(f"warmup {w}: {dt:.3f}s its={sp.iters[-2:]}")
    times = []
    for k in range(args.steps):
        dt, c, g = step(sp)
        times.append(dt)
        log(f"step {k}: {dt:.3f}s its={sp.iters[-2:]}")
    its = sp.iters[-2:]
    out = {
        "metric": "SIMP compliance gradient (forward + adjoint) per second",
        "value": 1.0 / (sum(times) / len(times)), "unit": "gradients/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * sum(times) / len(times),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 assembly + f32/f64 mixed solve",
        "data": "synthetic structured mesh, seeded densities (no external data)",
        "config": {"workload": f"SIMP hex cube {n}^3, x ~ U[0.1,1] seed 0, p=3 E0=1 Emin=1e-9 "
                               f"nu=0.3, x=0 clamped, unit z traction on the 8x8 patch at x=1; "
                               f"mixed GMRES({args.m}) recycling {args.deflation} to "
                               f"{args.rtol:g}; plain path FP64 GMRES({args.m_fp64})+MGS",
                   "elements": prob.mesh.nelem, "dof": 3 * prob.mesh.nnodes},
        "compliance": c, "iterations_mixed": its,
    }
    if not args.no_fp64:
        sp64 = SolidComplianceProblem(prob, p=3.0, E0=1.0, Emin=1e-9, rtol=args.rtol,
                                      m=args.m_fp64, mixed=False)
        dt64, c64, g64 = step(sp64)
        log(f"fp64 step: {dt64:.3f}s its={sp64.iters[-2:]}")
        out.update(step_s_fp64=dt64, iterations_fp64=sp64.iters[-2:],
                   speedup_mixed_vs_fp64=dt64 / (sum(times) / len(times)),
                   grad_rel_err_vs_fp64=float(np.abs(g - g64).max() / np.abs(g64).max()),
                   compliance_rel_err_vs_fp64=abs(c - c64) / abs(c64))
        del sp64
    if not args.no_cpu:
        from tacs_amd.meshes import simp_thickness
        its64 = out.get("iterations_fp64", its)
        cb = cpu_solid_fraction(prob, simp_thickness(x0), its64, f"SIMP {n}^3 hex8", E=1.0,
                                nu=0.3)
        cb.update(value=1.0 / cb["step_s_estimate"], unit="gradients/s")
        out["cpu_baseline"] = cb
    if args.fd:
        # The solves reach about 1e-13 relative residual (the FP64 rounding
        # floor of this stiffness contrast), so the compliance is accurate to
        # about 1e-12 relative; a single density moves it by about
        # 1e-6 * h relative, so per-DV differences use h = 1e-3, and the
        # directional difference (all DVs, random direction) h = 1e-6.
        sp.ksm.setTolerances(1e-12, 1e-30)
        rng_fd = np.random.default_rng(1)
        idx = rng_fd.choice(prob.mesh.nelem, args.fd, replace=False)
        h = 1e-3
        fd = []
        for q, e in enumerate(idx):
            vals = []
            for sgn in (1.0, -1.0):
                xp = x0.copy()
                xp[e] += sgn * h
                vals.append(sp.forward(torch.tensor(xp, dtype=torch.float64, device="cuda")))
            fd.append((vals[0] - vals[1]) / (2 * h))
            log(f"fd {q}: element {e} fd {fd[-1]:.6e} adjoint {g[e]:.6e}")
        fd = np.array(fd)
        ge = g[idx]
        d = rng_fd.standard_normal(prob.mesh.nelem)
        hd = 1e-6
        vals = []
        for sgn in (1.0, -1.0):
            xp = np.clip(x0 + sgn * hd * d, 1e-3, None)
            vals.append(sp.forward(torch.tensor(xp, dtype=torch.float64, device="cuda")))
        dfd = (vals[0] - vals[1]) / (2 * hd)
        log(f"directional fd {dfd:.8e} adjoint {float(g @ d):.8e}")
        out["fd_check"] = {
            "h": h, "dvs": [int(e) for e in idx], "fd": fd.tolist(), "adjoint": ge.tolist(),
            "max_rel_err": float(np.max(np.abs(fd - ge) / np.abs(ge))),
            "max_err_over_max_grad": float(np.max(np.abs(fd - ge)) / np.abs(ge).max()),
            "directional": {"h": hd, "fd": dfd, "adjoint": float(g @ d),
                            "rel_err": abs(dfd - float(g @ d)) / abs(float(g @ d))},
            "solve_rtol": 1e-12}
    print(json.dumps(out))


def batched_main(args, L, ev):
    """27-node hex cube (40^3 by default) with nrhs load cases solved as one
    block of right-hand sides (batched mixed-precision GMRES with Krylov
    subspace recycling per load case); the plain FP64 GMRES(30) one load case
    at a time and the single right-hand-side recycling solver one load case
    at a time are timed beside it.  Prints its own JSON line."""
    import torch
    from tacs_amd import TACS
    from tacs_amd.problems import SolidProblem
    n = args.n if args.n != 128 else 40
    R = args.nrhs
    t0 = time.perf_counter()
    p = SolidProblem(n, order=3, dv="element")
    log(f"setup {time.perf_counter() - t0:.1f}s: {p.mesh.nelem} hex27, {3 * p.mesh.nnodes} DOF")
    p.assemble()
    if args.kernels_only:
        # the hex27 Jacobian kernels in isolation (profiler runs): both solid
        # Jacobian forms, three assemblies each, FP64 element arithmetic
        a = p.assembler
        kj = {}
        for kern in (1, 2):
            a.setJacobianKernel(kern)
            a.assembleJacobian(1.0, 0.0, 0.0, None, p.mat)
            e0, e1 = ev.event(), ev.event()
            ev.record(e0)
            for _ in range(3):
                a.assembleJacobian(1.0, 0.0, 0.0, None, p.mat)
            ev.record(e1)
            kj["rows" if kern == 1 else "element_blocks"] = ev.elapsed_ms(e0, e1) / 3
        a.setJacobianKernel(0)
        kj.update(elements=p.mesh.nelem, dof=3 * p.mesh.nnodes, nnz_blocks=p.mat.getInfo()[2],
                  flop_per_element=2283228, unit="ms per assembleJacobian (no residual)")
        # the R right-hand-side product of the batched solver (tuned variant),
        # FP32 and FP64 values, 20 launches each
        import torch
        a.assembleJacobian(1.0, 0.0, 0.0, None, p.mat)
        rp, cp, ap = p.mat.getDeviceArrays()
        b, nr, nnz = p.mat.getInfo()
        A32 = torch.empty(nnz * 9, dtype=torch.float32, device="cuda")
        L.tacs_hip_convert_f64_f32(nnz * 9, ap, A32.data_ptr(), None)
        X = torch.rand(3 * nr * R, dtype=torch.float64, device="cuda")
        Y = torch.empty_like(X)
        torch.cuda.synchronize()
        for key, fn in (("spmm_fp32_values_ms", lambda: L.tacs_hip_bsr3_spmm_mixed(
                             nr, R, rp, cp, A32.data_ptr(), X.data_ptr(), Y.data_ptr(), None)),
                        ("spmm_fp64_values_ms", lambda: L.tacs_hip_bsr3_spmm(
                             nr, R, rp, cp, ap, X.data_ptr(), Y.data_ptr(), None))):
            fn()
            e0, e1 = ev.event(), ev.event()
            ev.record(e0)
            for _ in range(20):
                fn()
            ev.record(e1)
            kj[key] = ev.elapsed_ms(e0, e1) / 20
        kj["spmm_nrhs"] = R
        kj["spmm_fp32_algorithmic_bytes"] = 36 * nnz + 4 * nnz + 4 * (nr + 1) + 2 * 8 * 3 * nr * R
        print(json.dumps(kj))
        return
    bs = batched_loads(p, R)
    xs = [p.assembler.createVec() for _ in range(R)]
    pc = TACS.Pc(p.mat)
    if CHEB["degree"] > 1:
        pc.setChebyshev(CHEB["degree"], CHEB["ratio"], CHEB["values"])
    ksm = TACS.BatchedKSM(p.mat, pc, m=args.m, restarts=100000, nrhs=R, prec=32,
                          deflation=args.deflation)
    ksm.setTolerances(args.rtol, 1e-30)
    for w in range(args.warmup):
        r = batched_step(p, pc, ksm, bs, xs)
        log(f"warmup {w}: {r['t']} flag={r['flag']} its={r['its']}")
    L.tacs_device_synchronize()
    t0 = time.perf_counter()
    steps = []
    for k in range(args.steps):
        steps.append(batched_step(p, pc, ksm, bs, xs))
        log(f"step {k}: {steps[-1]['t']} flag={steps[-1]['flag']} its={steps[-1]['its']}")
    L.tacs_device_synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    last = steps[-1]
    u_b = [x.getArray().copy() for x in xs[:2]]

    # kernels in isolation (HIP events): the multi-RHS SpMM and the DCGS2 sweeps
    rp, cp, ap = p.mat.getDeviceArrays()
    b, nr, nnz = p.mat.getInfo()
    nd = 3 * nr
    N = nd * R
    A32 = torch.empty(nnz * 9, dtype=torch.float32, device="cuda")
    L.tacs_hip_convert_f64_f32(nnz * 9, ap, A32.data_ptr(), None)
    kb = 40  # basis vectors of a typical step (20 recycled + 20 Arnoldi)
    W = torch.rand((kb + 2) * N, dtype=torch.float64, device="cuda")
    coef = torch.rand(R * (2 * kb + 2), dtype=torch.float64, device="cuda") * 1e-3
    dots = torch.zeros(2 * R * (kb + 1), dtype=torch.float64, device="cuda")
    scratch = torch.empty(L.tacs_hip_bdotq_scratch_bytes(R) // 8, dtype=torch.float64,
                          device="cuda")
    torch.cuda.synchronize()

    def timed(fn, reps=20):
        for _ in range(3):
            fn()
        e0, e1 = ev.event(), ev.event()
        ev.record(e0)
        for _ in range(reps):
            fn()
        ev.record(e1)
        return ev.elapsed_ms(e0, e1) / reps

    Wp, X, Y = W.data_ptr(), W.data_ptr() + kb * N * 8, W.data_ptr() + (kb + 1) * N * 8
    spmm_ms = timed(lambda: L.tacs_hip_bsr3_spmm_mixed(nr, R, rp, cp, A32.data_ptr(), X, Y, None))
    upd_ms = timed(lambda: L.tacs_hip_bdcgs_update(nd, R, 0, Wp, kb, Wp, N, coef.data_ptr(),
                                                   2 * kb + 2, X, Y, None))
    dot_ms = timed(lambda: L.tacs_hip_bdotq(nd, R, 2, 0, Wp, kb, Wp, N, X, Y, scratch.data_ptr(),
                                            dots.data_ptr(), dots.data_ptr() + R * (kb + 1) * 8,
                                            kb + 1, None))
    spmm_bytes = 36 * nnz + 4 * nnz + 4 * (nr + 1) + 2 * 8 * N
    upd_bytes = (kb + 4) * N * 8
    dot_bytes = (kb + 2) * N * 8
    del W, A32
    torch.cuda.empty_cache()

    # one load case at a time: plain FP64 GMRES(30) (2 cases timed) and the
    # single right-hand-side recycling solver (4 cases timed)
    seq = {}
    if not args.no_fp64:
        pc64 = TACS.Pc(p.mat)
        pc64.factor()
        k64 = TACS.KSM(p.mat, pc64, args.m_fp64, 100000)
        k64.setTolerances(args.rtol, 1e-30)
        x = p.assembler.createVec()
        tt, its64, u64 = 0.0, [], []
        for j in range(2):
            t1 = time.perf_counter()
            k64.solve(bs[j], x)
            L.tacs_device_synchronize()
            tt += time.perf_counter() - t1
            its64.append(k64.getIterCount())
            u64.append(x.getArray().copy())
        seq["fp64_gmres_s_per_case"] = tt / 2
        seq["fp64_iterations"] = its64
        seq["fp64_solve_s_16_cases_estimate"] = tt / 2 * R
        seq["u_rel_err_vs_fp64"] = max(float(np.abs(u_b[j] - u64[j]).max() / np.abs(u64[j]).max())
                                       for j in range(2))
        log(f"fp64 sequential: {tt / 2:.2f} s per case, its {its64}")
    mk = TACS.MixedKSM(p.mat, pc, args.m, 100000, deflation=args.deflation)
    mk.setTolerances(args.rtol, 1e-30)
    x = p.assembler.createVec()
    t1 = time.perf_counter()
    its1 = []
    for j in range(4):
        mk.solve(bs[j], x)
        its1.append(mk.getIterCount())
    L.tacs_device_synchronize()
    seq["single_recycling_s_per_case"] = (time.perf_counter() - t1) / 4
    seq["single_recycling_iterations"] = its1

    solve_s = last["t"]["solve_s"]
    out = {
        "metric": f"batched solve: load cases/s (hex27, {R} load cases, {args.rtol:g} relative residual)",
        "value": R / (ms_per_step * 1e-3),
        "unit": "load cases/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 assembly + f32-valued operator / f64 vectors solve",
        "data": "synthetic structured mesh and seeded load cases (no external data)",
        "config": {"workload": f"27-node hex cube {n}^3, E=70e9 nu=0.3, x=0 clamped, {R} load "
                               f"cases (seeded normal nodal traction, std 1e6 Pa, on x=1); "
                               f"assembly + batched mixed GMRES({args.m}) with per-load-case "
                               f"recycling of {args.deflation} vectors, block-Jacobi, to "
                               f"{args.rtol:g} relative residual + per-case compliance and "
                               f"thickness gradient",
                   "elements": p.mesh.nelem, "dof": nd, "nnz_blocks": nnz, "load_cases": R},
        "assembly_s": last["t"]["assembly_s"], "solve_s_batched": solve_s,
        "iterations_batched": last["its"], "flag": last["flag"],
        "roofline": {"bound": "hbm", "achieved": upd_bytes / upd_ms / 1e6, "peak": 8000.0,
                     "unit": "GB/s", "frac": upd_bytes / upd_ms / 1e6 / 8000.0, "traffic": None,
                     "kernel": f"bdcgs_update (DCGS2 update sweep, {kb} basis vectors, {R} RHS)",
                     "avg_ms": upd_ms, "algorithmic_bytes": upd_bytes},
        "roofline_dcgs_dots": {"bound": "hbm", "achieved": dot_bytes / dot_ms / 1e6,
                               "peak": 8000.0, "unit": "GB/s",
                               "frac": dot_bytes / dot_ms / 1e6 / 8000.0, "avg_ms": dot_ms,
                               "algorithmic_bytes": dot_bytes},
        "roofline_spmm": {"bound": "hbm", "achieved": spmm_bytes / spmm_ms / 1e6, "peak": 8000.0,
                          "unit": "GB/s", "frac": spmm_bytes / spmm_ms / 1e6 / 8000.0,
                          "avg_ms": spmm_ms, "algorithmic_bytes": spmm_bytes,
                          "kernel": "bsr3_spmm_quad<float> (FP32 values, FP64 block vectors)"},
        "one_case_at_a_time": seq,
    }
    if "fp64_solve_s_16_cases_estimate" in seq:
        out["speedup_batched_vs_fp64_sequential"] = seq["fp64_solve_s_16_cases_estimate"] / solve_s
    out["speedup_batched_vs_single_recycling"] = seq["single_recycling_s_per_case"] * R / solve_s
    if not args.no_cpu:
        its = seq.get("fp64_iterations", last["its"])
        out["cpu_baseline"] = cpu_solid_fraction(p, np.ones(p.mesh.nelem), [int(np.mean(its))],
                                                 f"hex27 {n}^3 x {R}", nrhs=R)
    print(json.dumps(out))


def multi_main(args, L, ev):
    """Many small uneven meshes: 32 independent 8-node hex cubes with edge
    lengths seeded-uniform in 8..24 elements, as one assembler (the
    block-diagonal union): one assembly launch for every cube, one mixed
    recycling solve (each cube's right-hand side scaled to unit norm) that
    brings every cube to its own relative residual, the
    compliance and thickness gradient of every cube.  Timed beside the same
    work cube by cube (one assembler per cube, the same solver) and plain FP64
    GMRES(30) cube by cube.  Prints its own JSON line."""
    from tacs_amd import TACS, functions
    from tacs_amd.problems import MultiCubeProblem, SolidProblem
    rng = np.random.default_rng(args.seed)
    sizes = [int(v) for v in rng.integers(8, 25, args.ncubes)]
    t0 = time.perf_counter()
    p = MultiCubeProblem(sizes, dv="element")
    log(f"setup {time.perf_counter() - t0:.1f}s: {len(sizes)} cubes {sizes}, {p.nelem} elements, "
        f"{3 * p.nnodes} DOF")
    a = p.assembler
    pc = TACS.Pc(p.mat)
    if CHEB["degree"] > 1:
        pc.setChebyshev(CHEB["degree"], CHEB["ratio"], CHEB["values"])
    union = TACS.MixedKSM(p.mat, pc, args.m, 100000, deflation=args.deflation)
    seg = TACS.SegmentedKSM(p.mat, pc, p.node_off, m=args.m, restarts=100000, prec=32,
                            deflation=args.deflation)
    rhs = a.createVec()

    def step(kind):
        t = {}
        t0 = time.perf_counter()
        a.zeroVariables()
        a.assembleJacobian(1.0, 0.0, 0.0, p.res, p.mat)
        L.tacs_device_synchronize()
        t["assembly_s"] = time.perf_counter() - t0
        pc.factor()
        if kind == "segmented":
            # every cube its own system: own Krylov and recycled space, own
            # stopping test |r_c| <= rtol |b_c|
            seg.setTolerances(args.rtol, 1e-30)
            flag = seg.solve(p.res, p.u)
            its = [int(v) for v in seg.getStats()[0]]
        else:
            # each cube's right-hand side scaled to unit norm (the cubes are
            # decoupled, so cube c's solution scales with it): |r'| <= rtol then
            # gives every cube |r_c| <= rtol |b_c|
            p.scale_cubes(p.res, rhs)
            union.setTolerances(1e-30, args.rtol)
            flag = union.solve(rhs, p.u)
            its = union.getIterCount()
            p.unscale_cubes(p.u)
        p.u.scale(-1.0)
        a.setVariables(p.u)
        comp = functions.Compliance(a)
        a.evalFunctions([comp])
        g = a.createDesignVec()
        a.addDVSens([comp], [g], 1.0)
        psi = a.createVec()
        psi.copyValues(p.u)
        psi.scale(2.0)
        a.addAdjointResProducts([psi], [g], -1.0)
        L.tacs_device_synchronize()
        t["step_s"] = time.perf_counter() - t0
        return t, flag, its

    def cube_residuals():
        # every cube's true relative residual (host, from the assembled matrix)
        rowp, cols, A = p.mat.getArrays()
        u = p.u.getArray().copy()
        b = -p.res.getArray()
        Ku = np.zeros_like(u)
        Ab = A.reshape(-1, 3, 3)
        rows = np.repeat(np.arange(rowp.size - 1), np.diff(rowp))
        np.add.at(Ku.reshape(-1, 3), rows, np.einsum("kij,kj->ki", Ab, u.reshape(-1, 3)[cols]))
        return [float(np.linalg.norm((Ku - b)[p.dof(c)]) / np.linalg.norm(b[p.dof(c)]))
                for c in range(len(sizes))]

    legs = {}
    for kind in ("segmented", "union"):
        for _ in range(args.warmup):
            step(kind)
        t0 = time.perf_counter()
        for k in range(args.steps):
            t, flag, its = step(kind)
            log(f"{kind} step {k}: {t} flag={flag} its={its}")
        wall = time.perf_counter() - t0
        legs[kind] = dict(ms=1e3 * wall / args.steps, t=t, flag=flag, its=its,
                          rel=cube_residuals(), u=p.u.getArray().copy())
    main_leg = legs[args.multi_solver]
    ms, t, its, rel = main_leg["ms"], main_leg["t"], main_leg["its"], main_leg["rel"]
    other = "union" if args.multi_solver == "segmented" else "segmented"
    u_diff = float(np.abs(legs["segmented"]["u"] - legs["union"]["u"]).max()
                   / np.abs(legs["union"]["u"]).max())

    # cube by cube: one assembler each, the same recycling solver; FP64 GMRES(30)
    t_seq, t_fp64, its_seq, its_fp64 = 0.0, 0.0, [], []
    for n in sizes:
        q = SolidProblem(n, dv="element")
        t1 = time.perf_counter()
        q.assemble()
        qpc = TACS.Pc(q.mat)
        qpc.factor()
        qk = TACS.MixedKSM(q.mat, qpc, args.m, 100000, deflation=args.deflation)
        qk.setTolerances(args.rtol, 1e-30)
        qk.solve(q.res, q.u)
        its_seq.append(qk.getIterCount())
        q.u.scale(-1.0)
        qa = q.assembler
        qa.setVariables(q.u)
        comp = functions.Compliance(qa)
        qa.evalFunctions([comp])
        g = qa.createDesignVec()
        qa.addDVSens([comp], [g], 1.0)
        L.tacs_device_synchronize()
        t_seq += time.perf_counter() - t1
        if not args.no_fp64:
            k64 = TACS.KSM(q.mat, qpc, args.m_fp64, 100000)
            k64.setTolerances(args.rtol, 1e-30)
            x = qa.createVec()
            t1 = time.perf_counter()
            k64.solve(q.res, x)
            L.tacs_device_synchronize()
            t_fp64 += time.perf_counter() - t1
            its_fp64.append(k64.getIterCount())
    out = {
        "metric": f"many small meshes: cubes/s ({len(sizes)} hex8 cubes, edges 8..24, "
                  f"{args.rtol:g} relative residual each)",
        "value": len(sizes) / (ms * 1e-3), "unit": "cubes/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 assembly + f32-valued operator / f64 vectors solve",
        "data": "synthetic structured meshes (seeded sizes, no external data)",
        "config": {"workload": f"{len(sizes)} independent 8-node hex unit cubes, edge lengths "
                               f"{sizes} elements, E=70e9 nu=0.3, x=0 clamped, traction 1e6 Pa "
                               f"on x=1; one assembler (block-diagonal union): one assembly "
                               f"launch, then " + (
                                   f"every cube its own system (own Krylov and recycled space, "
                                   f"own stopping test) in one batched mixed GMRES({args.m}) "
                                   f"recycling {args.deflation} per cube"
                                   if args.multi_solver == "segmented" else
                                   f"mixed GMRES({args.m}) recycling {args.deflation} on the "
                                   f"union with per-cube unit-norm right-hand sides") +
                               f", every cube to {args.rtol:g} relative residual, compliance "
                               f"and thickness gradient",
                   "elements": p.nelem, "dof": 3 * p.nnodes, "cubes": len(sizes),
                   "solver": args.multi_solver},
        "assembly_s": t["assembly_s"],
        "iterations_segmented_per_cube": legs["segmented"]["its"],
        "iterations_union": legs["union"]["its"],
        "max_cube_rel_residual": max(rel),
        f"ms_per_step_{other}": legs[other]["ms"],
        f"max_cube_rel_residual_{other}": max(legs[other]["rel"]),
        "segmented_vs_union_u_rel_diff": u_diff,
        "speedup_segmented_vs_union": legs["union"]["ms"] / legs["segmented"]["ms"],
        "cube_by_cube_same_solver_s": t_seq, "cube_by_cube_iterations": its_seq,
        "speedup_vs_cube_by_cube": t_seq / (ms * 1e-3),
    }
    if not args.no_fp64:
        out["cube_by_cube_fp64_gmres_solve_s"] = t_fp64
        out["cube_by_cube_fp64_iterations"] = its_fp64
        out["speedup_vs_fp64_cube_by_cube_solves"] = t_fp64 / (ms * 1e-3)
        if not args.no_cpu:
            out["cpu_baseline"] = cpu_multi(p, its_fp64)
    print(json.dumps(out))


def matfree_main(args, L, ev):
    """Matrix-free Jacobian products (TACSMatrixFreeMat: per quadrature point
    J^-1 and w det t, element products gathered per node) against the FP64
    product over the assembled sliced ELL copy, on the same cube (hex27 40^3
    by default, --order 2 for hex8), and FP64 GMRES(m_fp64) + MGS solves with
    either operator and the same block-Jacobi preconditioner.  A step is one
    operator product; prints its own JSON line."""
    from tacs_amd import TACS
    from tacs_amd.problems import SolidProblem
    order = args.order
    n = args.n if not (order == 3 and args.n == 128) else 40
    p = SolidProblem(n, order=order, dv="element")
    m = p.mesh
    log(f"matfree: order {order}, {n}^3, {m.nelem} elements, {3 * m.nnodes} DOF")
    p.assemble()
    sp = time_spmv(L, ev, p)
    a = p.assembler
    mf = TACS.MatrixFreeMat(a)
    e0, e1 = ev.event(), ev.event()
    ev.record(e0)
    mf.assembleMatrixFreeData(TACS.JACOBIAN_MATRIX, 1.0, 0.0, 0.0)
    ev.record(e1)
    data_ms = ev.elapsed_ms(e0, e1)
    x, y, y2 = a.createVec(), a.createVec(), a.createVec()
    rng = np.random.default_rng(0)
    xv = rng.standard_normal(x.getSize()).reshape(-1, 3)
    xv[p.bc_nodes] = 0.0
    x.getArray()[:] = xv.reshape(-1)
    xd, yd = x.getDevicePointer(), y.getDevicePointer()
    for _ in range(args.warmup + 4):
        assert L.tacs_matfree_mult_device(mf.ptr, xd, yd) == 0
    e0, e1 = ev.event(), ev.event()
    reps = 50
    ev.record(e0)
    for _ in range(reps):
        L.tacs_matfree_mult_device(mf.ptr, xd, yd)
    ev.record(e1)
    mf_ms = ev.elapsed_ms(e0, e1) / reps
    # parity with the assembled operator (Dirichlet rows of the product zeroed)
    mf.mult(x, y)
    p.mat.mult(x, y2)
    ym = y2.getArray().reshape(-1, 3).copy()
    ym[p.bc_nodes] = 0.0
    par = float(np.abs(y.getArray() - ym.reshape(-1)).max() / np.abs(ym).max())
    npe = order ** 3
    N = 3 * m.nnodes
    table = 8 * (4 + 10 * npe * m.nelem)
    alg = table + 4 * npe * m.nelem + 2 * 8 * N
    ev_bytes = 2 * 8 * 3 * npe * m.nelem + 4 * npe * m.nelem + 4 * (m.nnodes + 1)
    out = {
        "metric": "matrix-free Jacobian product time vs the assembled FP64 product",
        "value": mf_ms, "unit": "ms per product", "n_gpus": 1, "steps": reps, "warmup": 5,
        "ms_per_step": mf_ms, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic structured mesh (no external data)",
        "config": {"workload": f"{'hex27' if order == 3 else 'hex8'} cube {n}^3, E=70e9 nu=0.3, "
                               f"x=0 clamped; TACSMatrixFreeMat (TACS_JACOBIAN_MATRIX) vs the "
                               f"assembled 3x3 BCSR product",
                   "elements": m.nelem, "dof": N},
        "matfree_ms": mf_ms, "assembled_fp64_ms": sp["fp64_ms"],
        "speedup_vs_assembled_fp64": sp["fp64_ms"] / mf_ms,
        "assembled_fp32_values_fp64_vectors_ms": sp["mixed_ms"],
        "matfree_data_ms": data_ms, "matfree_data_bytes": table,
        "assembled_fp64_bytes": sp["fp64_bytes"],
        "roofline": {"bound": "hbm", "achieved": alg / mf_ms / 1e6, "peak": 8000.0,
                     "unit": "GB/s", "frac": alg / mf_ms / 1e6 / 8000.0, "traffic": None,
                     "kernel": f"solid_matfree_ev<{order}> + gather_elem_vectors<{3 * npe}>",
                     "algorithmic_bytes": alg,
                     "note": "table + connectivity + x + y; the element-vector round trip "
                             f"adds {ev_bytes} B"},
        "max_rel_diff_vs_assembled": par,
    }
    if not args.no_fp64:
        pc = TACS.Pc(p.mat)
        pc.factor()
        res = {}
        for key, op in (("assembled", p.mat), ("matfree", mf)):
            ksm = TACS.KSM(op, pc, args.m_fp64, 100000)
            ksm.setTolerances(args.rtol, 1e-30)
            u = a.createVec()
            L.tacs_device_synchronize()
            t0 = time.perf_counter()
            flag = ksm.solve(p.res, u)
            L.tacs_device_synchronize()
            res[key] = (time.perf_counter() - t0, ksm.getIterCount(), flag, u.getArray().copy())
            log(f"{key} GMRES: {res[key][0]:.2f}s its={res[key][1]} flag={flag}")
        ua, uf = res["assembled"][3], res["matfree"][3]
        out.update(gmres_assembled_s=res["assembled"][0], gmres_matfree_s=res["matfree"][0],
                   gmres_iterations=[res["assembled"][1], res["matfree"][1]],
                   gmres_flags=[res["assembled"][2], res["matfree"][2]],
                   gmres_speedup=res["assembled"][0] / res["matfree"][0],
                   u_rel_diff=float(np.linalg.norm(uf - ua) / np.linalg.norm(ua)),
                   gmres=f"FP64 GMRES({args.m_fp64}) + MGS + block-Jacobi to {args.rtol:g}")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_matfree(p)
    print(json.dumps(out))


def cpu_matfree(p, reps=20):
    """Measured multi-core CPU baseline of one operator product on the same
    mesh: the oracle's threaded 3x3 block-CSR product (oracle/cpu_baseline.c,
    the reference's BCSRMatVecMult3 loop over rows) on the assembled
    matrix, all host cores, best of reps products.  The reference has no
    element-level matrix-free product for solids beyond the same per-element
    loop the assembly runs, so the assembled product is its fastest form."""
    O = _oracle()
    model, threads = cpu_env()
    m = p.mesh
    rowp, cols = O.node_csr(m.nnodes, m.conn, m.npe)
    colors = O.color_elements(m.nnodes, m.conn, m.npe)
    A, _ = O.par_assemble_solid(m.order, colors, m.conn, m.X.reshape(-1), None,
                                np.ones(m.nelem), O.iso_C21(70e9, 0.3), rowp, cols,
                                want_res=False)
    x = np.random.default_rng(3).standard_normal(3 * m.nnodes)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        O.par_bsr_mult(3, rowp, cols, A, x)
        ts.append(time.perf_counter() - t0)
    t = min(ts)
    return {"cores": threads, "kind": "port", "cpu_model": model,
            "sample": (f"oracle OpenMP 3x3 block-CSR product on {threads} threads of {model}, "
                       f"the same assembled matrix ({cols.size} blocks), best of {reps}"),
            "value": 1e3 * t, "unit": "ms per product"}


def gmres_graph_main(args, L, ev):
    """FP64 GMRES(30) + MGS + block-Jacobi (the reference's solver) on the
    n^3 hex8 cubes n = 8, 12, 20, 40, step by step and with each restart
    cycle replayed as one captured graph: time per Arnoldi step.  Prints its
    own JSON line."""
    from tacs_amd import TACS
    from tacs_amd.problems import SolidProblem
    rows = {}
    for n in (8, 12, 20, 40):
        p = SolidProblem(n)
        p.assemble()
        pc = TACS.Pc(p.mat)
        pc.factor()
        x = p.assembler.createVec()
        for graph in (False, True):
            ksm = TACS.KSM(p.mat, pc, 30, 100000)
            ksm.setTolerances(args.rtol, 1e-30)
            ksm.setCycleGraph(graph)
            ksm.solve(p.res, x)
            L.tacs_device_synchronize()
            t0 = time.perf_counter()
            reps = 3
            for _ in range(reps):
                ksm.solve(p.res, x)
            L.tacs_device_synchronize()
            dt = (time.perf_counter() - t0) / reps
            its = ksm.getIterCount()
            rows[f"{n}^3 {'graph' if graph else 'step_by_step'}"] = {
                "solve_s": dt, "iterations": its, "us_per_step": 1e6 * dt / its,
                "dof": 3 * p.mesh.nnodes}
            log(f"{n}^3 graph={graph}: {dt:.4f} s, {its} steps")
    g20 = rows["20^3 graph"]["us_per_step"]
    s20 = rows["20^3 step_by_step"]["us_per_step"]
    out = {"metric": "FP64 GMRES(30) time per Arnoldi step, captured cycles vs step by step",
           "value": g20, "unit": "us per step (20^3 cube, captured)", "n_gpus": 1,
           "steps": 3, "warmup": 1, "ms_per_step": g20 / 1e3, "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic structured meshes (no external data)",
           "config": {"workload": f"hex8 cubes 8^3..40^3, E=70e9 nu=0.3, x=0 clamped, traction "
                                  f"1e6 Pa on x=1; FP64 GMRES(30)+MGS+block-Jacobi to "
                                  f"{args.rtol:g}"},
           "speedup_20cube": s20 / g20, "rows": rows}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--m", type=int, default=None,
                    help="restart length of the mixed-precision GMRES (default 60; 120 for "
                         "the shell workload)")
    ap.add_argument("--deflation", type=int, default=None,
                    help="recycled harmonic Ritz vectors of the mixed-precision GMRES (default "
                         "20; 30 for the shell workload)")
    ap.add_argument("--m-fp64", type=int, default=30,
                    help="restart length of the plain FP64 GMRES (the reference default MGS)")
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernels-only", action="store_true",
                    help="assemble and time the SpMV kernels only (for profiler runs)")
    ap.add_argument("--workload", default="hex",
                    choices=["hex", "shell", "batched", "multi", "simp", "matfree",
                             "gmres_graph"],
                    help="hex: the 128^3 headline line; shell: the Quad4 plate; batched: the "
                         "hex27 cube with a block of load cases (own JSON lines)")
    ap.add_argument("--nrhs", type=int, default=16, help="load cases of the batched workload")
    ap.add_argument("--cheb-degree", type=int, default=None,
                    help="Chebyshev preconditioner degree of the recycling solvers (1: plain "
                         "block-Jacobi; default 14 for the hex and batched workloads, 10 for "
                         "the other solid workloads, 24 for shells)")
    ap.add_argument("--cheb-values", type=int, default=16, choices=[32, 16],
                    help="precision of the matrix values the Chebyshev polynomial's products read "
                         "(16: bfloat16 copy)")
    ap.add_argument("--cheb-ratio", type=float, default=None,
                    help="interval [lmax / ratio, lmax] of the Chebyshev polynomial (default "
                         "1000 for the hex workload, 300 for the other solid workloads, 3000 for "
                         "shells)")
    ap.add_argument("--multi-solver", default="segmented", choices=["segmented", "union"],
                    help="multi workload: every cube its own system (segmented) or one union "
                         "system")
    ap.add_argument("--order", type=int, default=3, choices=[2, 3],
                    help="matfree workload: 3 (hex27, 40^3 unless --n) or 2 (hex8)")
    ap.add_argument("--ncubes", type=int, default=32, help="cubes of the multi workload")
    ap.add_argument("--seed", type=int, default=11, help="seed of the multi workload's sizes")
    ap.add_argument("--fd", type=int, default=0,
                    help="simp workload: central-difference check on this many DVs")
    ap.add_argument("--monitor", default=None,
                    help="write the solvers' residual monitor lines to FILE.<solver>")
    args = ap.parse_args()
    # the thin plate's block-Jacobi spectrum wants a longer cycle and a larger
    # recycled space (512^2: 67k + 42k steps with 120 / 30)
    shell = args.workload == "shell"
    if args.m is None:
        args.m = 120 if shell else 60
    if args.deflation is None:
        args.deflation = 30 if shell else 20
    MONITOR["file"] = args.monitor
    if args.cheb_degree is None:
        # measured optima: the 128^3 hex8 cube 14 on [lmax/1000, lmax] (solve +
        # adjoint 0.85 s against 0.90 s with 10 on [lmax/300, lmax]), the other
        # solid systems 10, the batched block of load cases 14 (its polynomial
        # steps are cheaper per Arnoldi step than the DCGS2 sweeps of 16
        # right-hand sides), shells 24
        args.cheb_degree = (24 if shell else
                            14 if args.workload in ("batched", "hex") else 10)
    if args.cheb_ratio is None:
        args.cheb_ratio = (3000.0 if shell else
                           1000.0 if args.workload == "hex" else 300.0)
    CHEB.update(degree=args.cheb_degree, ratio=args.cheb_ratio, values=args.cheb_values)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU design is "
              "described in DESIGN.md and not built", file=sys.stderr)
        sys.exit(2)

    from tacs_amd import lib
    from tacs_amd.problems import SolidProblem
    L = lib()
    ev = HipEvents()
    if args.workload == "shell":
        shell_main(args, L, ev)
        return
    if args.workload == "batched":
        batched_main(args, L, ev)
        return
    if args.workload == "multi":
        multi_main(args, L, ev)
        return
    if args.workload == "simp":
        simp_main(args, L, ev)
        return
    if args.workload == "matfree":
        matfree_main(args, L, ev)
        return
    if args.workload == "gmres_graph":
        gmres_graph_main(args, L, ev)
        return
    t0 = time.perf_counter()
    p = SolidProblem(args.n, dv="element")
    log(f"setup {time.perf_counter() - t0:.1f}s: {p.mesh.nelem} elements, {3 * p.mesh.nnodes} DOF")

    if args.kernels_only:
        p.assemble()
        ev = HipEvents()
        sp = time_spmv(L, ev, p)
        # copy kernel of known size (FP64 values -> FP32: 8 B read, 4 B
        # written per value) that scales the profiler's byte counters
        b, nr, nnz = p.mat.getInfo()
        nv = b * b * nnz
        a32 = ctypes.c_void_p()
        assert ev.hip.hipMalloc(ctypes.byref(a32), ctypes.c_size_t(4 * nv)) == 0
        rp, cp, ap = p.mat.getDeviceArrays()
        for _ in range(2):
            L.tacs_hip_convert_f64_f32(nv, ap, a32, None)
        L.tacs_device_synchronize()
        ev.hip.hipFree(a32)
        sp["copy_kernel_bytes"] = {"read": 8 * nv, "written": 4 * nv}
        # the sliced ELL fill as a reassembled step runs it: the FP64, FP32
        # and bfloat16 copies and the columns in one pass
        sp.update(time_fill(L, ev, p))
        e0, e1 = ev.event(), ev.event()
        ev.record(e0)
        for _ in range(3):
            p.assembler.assembleJacobian(1.0, 0.0, 0.0, None, p.mat)
        ev.record(e1)
        sp["assembly_ms"] = ev.elapsed_ms(e0, e1) / 3
        sp["fp64_GBps"] = sp["fp64_bytes"] / sp["fp64_ms"] / 1e6
        sp["fp32_GBps"] = sp["fp32_bytes"] / sp["fp32_ms"] / 1e6
        sp["mixed_GBps"] = sp["mixed_bytes"] / sp["mixed_ms"] / 1e6
        # the fused Chebyshev product of the preconditioner (bfloat16 values)
        from tacs_amd import TACS
        pck = TACS.Pc(p.mat)
        pck.setChebyshev(10, 300.0, 16)
        pck.factor()
        cms, cbytes = time_cheb(L, ev, p, pck)
        sp.update(cheb_bf16_ms=cms, cheb_bf16_bytes=cbytes, cheb_bf16_GBps=cbytes / cms / 1e6)
        print(json.dumps(sp))
        return
    pc, ksm = make_solver(p, "mixed", args.m, args.rtol, args.deflation)
    work = step_work(p)
    for w in range(args.warmup):
        r = run_step(p, pc, ksm, work)
        log(f"warmup {w}: {r['t']} its={r['its']}")
    L.tacs_device_synchronize()
    t0 = time.perf_counter()
    steps = []
    for k in range(args.steps):
        steps.append(run_step(p, pc, ksm, work, want_grad=(k == args.steps - 1)))
        log(f"step {k}: {steps[-1]['t']} its={steps[-1]['its']} flags={steps[-1]['flags']}")
    L.tacs_device_synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    mixed = steps[-1]

    # kernel timings (HIP events on the library's stream): the assembly as the
    # step runs it (Jacobian with its FP32 copy, residual), the Jacobian alone
    # (the roofline_assembly kernel), and the operator build of a step
    # (assembly + the FP64 and FP32 sliced ELL copies the products read)
    a = p.assembler

    def ev_ms(fn, reps=3):
        e0, e1 = ev.event(), ev.event()
        ev.record(e0)
        for _ in range(reps):
            fn()
        ev.record(e1)
        return ev.elapsed_ms(e0, e1) / reps

    asm_ms = ev_ms(lambda: a.assembleJacobian(1.0, 0.0, 0.0, p.res, p.mat))
    jac_ms = ev_ms(lambda: a.assembleJacobian(1.0, 0.0, 0.0, None, p.mat))

    def build():
        a.assembleJacobian(1.0, 0.0, 0.0, p.res, p.mat)
        p.mat.prepareProducts(64)
        p.mat.prepareProducts(32)

    build_ms = ev_ms(build)
    sp = time_spmv(L, ev, p)
    peak_bw = 8000.0
    bw64 = sp["fp64_bytes"] / sp["fp64_ms"] / 1e6
    cheb = time_cheb(L, ev, p, pc, values=CHEB["values"]) if CHEB["degree"] > 1 else None
    bw32 = sp["fp32_bytes"] / sp["fp32_ms"] / 1e6
    bwm = sp["mixed_bytes"] / sp["mixed_ms"] / 1e6
    jac_flops = 69120.0 * p.mesh.nelem  # dense B^T C B count per 8-node element
    jac_tf = jac_flops / (jac_ms * 1e-3) / 1e12

    fp64 = rec64 = None
    if not args.no_fp64:
        del pc, ksm
        # the same recycling algorithm on the FP64 operator: separates the
        # algorithm's share of the speed-up from the precision's
        pcr, ksmr = make_solver(p, "recycle64", args.m, args.rtol, args.deflation)
        run_step(p, pcr, ksmr, work)
        rec64 = run_step(p, pcr, ksmr, work, want_grad=True)
        log(f"fp64 recycling step: {rec64['t']} its={rec64['its']} flags={rec64['flags']}")
        del pcr, ksmr
        pc64, ksm64 = make_solver(p, "fp64", args.m_fp64, args.rtol)
        fp64 = run_step(p, pc64, ksm64, work, want_grad=True)
        log(f"fp64 step: {fp64['t']} its={fp64['its']} flags={fp64['flags']}")

    out = {
        "metric": "assembly Melem/s + solve+adjoint time; mixed/batched vs FP64 GPU @1e-8 resid",
        "value": p.mesh.nelem / (ms_per_step * 1e-3) / 1e6,
        "unit": "Melem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 assembly + f32/f64 mixed solve",
        "data": "synthetic structured mesh (seeded generator, no external data)",
        "config": {"workload": f"linear hex cube {args.n}^3, E=70e9 nu=0.3, x=0 clamped, "
                               f"traction 1e6 Pa on x=1, per-element thickness DVs; assembly + "
                               f"mixed GMRES({args.m}) recycling {args.deflation} vectors, "
                               + (f"Chebyshev degree {args.cheb_degree} on [lmax/"
                                  f"{args.cheb_ratio:g}, lmax] of block-Jacobi (products inside "
                                  f"the polynomial on {'bfloat16' if args.cheb_values == 16 else 'FP32'}"
                                  f" values)"
                                  if args.cheb_degree > 1 else "block-Jacobi") +
                               f", solve + compliance adjoint "
                               f"to {args.rtol:g} relative residual; plain path FP64 "
                               f"GMRES({args.m_fp64})+MGS",
                   "elements": p.mesh.nelem, "dof": 3 * p.mesh.nnodes,
                   "nnz_blocks": p.mat.getInfo()[2]},
        "assembly_Melem_s": p.mesh.nelem / (asm_ms * 1e-3) / 1e6,
        "assembly_ms": asm_ms,
        "assembly_note": "assembleJacobian as the step runs it: Jacobian (FP64 values and "
                         "the FP32 copy) and residual",
        "jacobian_ms": jac_ms,
        "operator_build_ms": build_ms,
        "operator_build_Melem_s": p.mesh.nelem / (build_ms * 1e-3) / 1e6,
        "operator_build_note": "assembly + the FP64 and FP32 sliced ELL copies the products "
                               "read (rebuilt every step, the matrix is reassembled)",
        "solve_adjoint_s_mixed": mixed["t"]["solve_adjoint_s"],
        "iterations_mixed": list(mixed["its"]),
        "roofline_fp64_spmv": {"bound": "hbm", "achieved": bw64, "peak": peak_bw,
                               "unit": "GB/s", "frac": bw64 / peak_bw,
                               "traffic": spmv_traffic("sell_spmv<3,2,double,double>"),
                               "kernel": "sell3_spmv<double,double> (FP64 3x3 block SpMV over "
                                         "the sliced ELL copy, plain path)",
                               "avg_ms": sp["fp64_ms"], "algorithmic_bytes": sp["fp64_bytes"],
                               "traffic_source": "profiles/spmv_traffic_hex8_128cube.json "
                                                 "(rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                                 "copy-kernel scaled)"},
        "roofline_mixed_spmv": {"bound": "hbm", "achieved": bwm, "peak": peak_bw,
                                "traffic": spmv_traffic("sell_spmv<3,4,float,double>"),
                                "unit": "GB/s", "frac": bwm / peak_bw, "avg_ms": sp["mixed_ms"],
                                "algorithmic_bytes": sp["mixed_bytes"],
                                "kernel": "sell3_spmv<float,double> (recycling solver operator)"},
        "roofline_fp32_spmv": {"bound": "hbm", "achieved": bw32, "peak": peak_bw,
                               "unit": "GB/s", "frac": bw32 / peak_bw, "avg_ms": sp["fp32_ms"],
                               "algorithmic_bytes": sp["fp32_bytes"]},
        "roofline_assembly": {"bound": "fp64_valu", "achieved": jac_tf, "peak": 78.6,
                              "unit": "TFLOP/s", "frac": jac_tf / 78.6, "avg_ms": jac_ms,
                              "kernel": "solid_jacobian_rows<2,true,false> (no MFMA: FP64 "
                                        "vector ALU)",
                              "note": "69,120 FLOP/element dense B^T C B count over the "
                                      "Jacobian alone; MI355X FP64 peak"},
    }
    if cheb is not None:
        # the step's dominant kernel: the fused Chebyshev product of the
        # preconditioner (about half of the step's kernel time at 128^3)
        cms, cbytes = cheb
        ckey = ("sell_spmv<3,4,unsigned short,double,true,true>" if CHEB["values"] == 16
                else "sell_spmv<3,4,float,double,true,true>")
        out["roofline"] = {
            "bound": "hbm", "achieved": cbytes / cms / 1e6, "peak": peak_bw, "unit": "GB/s",
            "frac": cbytes / cms / 1e6 / peak_bw, "traffic": spmv_traffic(ckey),
            "kernel": f"{ckey} (one Chebyshev step fused into the 3x3 block product over the "
                      f"sliced ELL copy, {'bfloat16' if CHEB['values'] == 16 else 'FP32'} "
                      f"values, FP64 vectors)",
            "avg_ms": cms, "algorithmic_bytes": cbytes,
            "traffic_source": "profiles/spmv_traffic_hex8_128cube.json (rocprofv3 FETCH_SIZE + "
                              "WRITE_SIZE, copy-kernel scaled)",
            "note": "three-term form: values and column indices once, row pointers, z gathered "
                    "once, rhs and zprev read, zout written, FP32 inverse diagonal blocks"}
    else:
        out["roofline"] = out["roofline_fp64_spmv"]
    if fp64 is not None:
        g, g64 = mixed["grad"], fp64["grad"]
        out["solve_adjoint_s_fp64"] = fp64["t"]["solve_adjoint_s"]
        out["iterations_fp64"] = list(fp64["its"])
        out["speedup_mixed_vs_fp64"] = fp64["t"]["solve_adjoint_s"] / mixed["t"]["solve_adjoint_s"]
        out["grad_rel_err_vs_fp64"] = float(np.abs(g - g64).max() / np.abs(g64).max())
        out["u_rel_err_vs_fp64"] = float(np.abs(mixed["u"] - fp64["u"]).max()
                                         / np.abs(fp64["u"]).max())
        out["compliance_rel_err_vs_fp64"] = abs(mixed["compliance"] - fp64["compliance"]) / abs(
            fp64["compliance"])
    if rec64 is not None:
        t_mix = mixed["t"]["solve_adjoint_s"]
        t_rec = rec64["t"]["solve_adjoint_s"]
        t_64 = fp64["t"]["solve_adjoint_s"]
        out["solve_adjoint_s_fp64_recycling"] = t_rec
        out["iterations_fp64_recycling"] = list(rec64["its"])
        out["speedup_split"] = {
            "algorithm": t_64 / t_rec,
            "precision": t_rec / t_mix,
            "total": t_64 / t_mix,
            "note": f"algorithm: FP64 GMRES({args.m_fp64})+MGS with block-Jacobi -> FP64 "
                    f"GCRO-DR({args.m}, {args.deflation}) with delayed CGS2" +
                    (f" and the degree-{args.cheb_degree} Chebyshev preconditioner"
                     if args.cheb_degree > 1 else "") +
                    "; precision: that -> the same solver on the FP32-valued operator",
        }
        g64r = rec64["grad"]
        out["grad_rel_err_mixed_vs_fp64_recycling"] = float(
            np.abs(mixed["grad"] - g64r).max() / np.abs(g64r).max())
    if not args.no_cpu:
        its = fp64["its"] if fp64 is not None else mixed["its"]
        out["cpu_baseline"] = cpu_solid_fraction(p, np.ones(p.mesh.nelem), its,
                                                 f"{args.n}^3 hex8")
        out["cfg1_cube20"] = cfg1_cube20(L, args.rtol)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
