"""Speed of the oracle's restatement, and of the multi-core baseline's
kernels, against the reference's own code on identical inputs, one core
each: the element Jacobian loop (TACSElement3D::addJacobian through
oracle/_ref against orc_solid_element and the baseline's
orc_iso_solid_jacobian) and the 3x3 block-CSR product (BCSRMatVecMult3
against orc_bsr_mult and the baseline's orc_par_bsr_mult).

TEST INFRASTRUCTURE: runs where the reference tree is present (this needs
oracle/_ref/libtacs_ref.so built by oracle/build_ref.sh); it tells how far
the CPU baseline that bench.py times with the oracle (on the GPU box, where
the reference cannot go) is from the reference's speed.  Writes the table to
stdout; the committed summary is profiles/cpu_oracle_vs_reference.md.

    python oracle/time_vs_reference.py
"""
import ctypes
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))
import oracle as O  # noqa: E402

REF = os.path.join(HERE, "_ref", "libtacs_ref.so")
_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)


def _dp(a):
    return a.ctypes.data_as(_d)


def _ip(a):
    return a.ctypes.data_as(_i)


def _best(fn, reps=5):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return min(ts)


def element_loop(R, order, nelem, seed=0):
    from tacs_amd.meshes import HexMesh

    npe = order ** 3
    nv = 3 * npe
    m = HexMesh(1, 1, 1, order=order)
    rng = np.random.default_rng(seed)
    Xe = np.ascontiguousarray(np.tile(m.X.reshape(-1), (nelem, 1))
                              + 0.05 * rng.standard_normal((nelem, nv)))
    C = O.iso_C21(70e9, 0.3)
    r1, m1 = np.zeros(nv), np.zeros(nv * nv)
    r2, m2 = np.zeros(nv), np.zeros(nv * nv)
    L = O.lib()
    t_ref = _best(lambda: R.ref_solid_jacobian_loop(order, nelem, _dp(Xe), ctypes.c_double(70e9),
                                                    ctypes.c_double(0.3), ctypes.c_double(1.0),
                                                    _dp(r1), _dp(m1)))
    t_orc = _best(lambda: L.orc_solid_jacobian_loop(order, nelem, _dp(Xe), _dp(C),
                                                    ctypes.c_double(1.0), _dp(r2), _dp(m2)))
    agree = np.abs(m1 - m2).max() / np.abs(m1).max()
    # the baseline's isotropic element kernel (cpu_baseline.c) on the same elements
    m3 = np.zeros(nv * nv)
    lam, mu = C[1], C[15]

    def iso_loop():
        for e in range(nelem):
            L.orc_iso_solid_jacobian(order, Xe[e].ctypes.data_as(_d), ctypes.c_double(lam),
                                     ctypes.c_double(mu), ctypes.c_double(1.0), _dp(m3))

    t_iso = _best(iso_loop, reps=3)
    agree_iso = np.abs(m1 - m3).max() / np.abs(m1).max()
    return t_ref / nelem, t_orc / nelem, agree, t_iso / nelem, agree_iso


def quad4_loop(R, nelem, seed=0):
    rng = np.random.default_rng(seed)
    base = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0], float)
    X = np.ascontiguousarray(np.tile(base, (nelem, 1)) + 0.05 * rng.standard_normal((nelem, 12)))
    r1, m1, r2, m2 = np.zeros(24), np.zeros(576), np.zeros(24), np.zeros(576)
    L = O.lib()
    c = ctypes.c_double
    t_ref = _best(lambda: R.ref_quad4_jacobian_loop(nelem, _dp(X), c(70e9), c(0.3), c(0.01),
                                                    _dp(r1), _dp(m1)))
    t_orc = _best(lambda: L.orc_quad4_jacobian_loop(nelem, _dp(X), c(70e9), c(0.3), c(0.01),
                                                    _dp(r2), _dp(m2)))
    return t_ref / nelem, t_orc / nelem, np.abs(m1 - m2).max() / np.abs(m1).max()


def spmv(R, n, reps=20):
    from tacs_amd.meshes import HexMesh

    m = HexMesh(n, n, n)
    rowp, cols = O.node_csr(m.nnodes, m.conn, 8)
    A, _ = O.assemble_solid(2, m.conn, m.X.reshape(-1), None, np.ones(m.nelem),
                            O.iso_C21(70e9, 0.3), rowp, cols, want_res=False)
    x = np.random.default_rng(1).standard_normal(3 * m.nnodes)
    y1, y2 = np.zeros_like(x), np.zeros_like(x)
    rp, cp = np.ascontiguousarray(rowp, np.int32), np.ascontiguousarray(cols, np.int32)
    L = O.lib()
    nr = m.nnodes
    t_ref = _best(lambda: R.ref_bsr_mult(3, nr, _ip(rp), _ip(cp), _dp(A), _dp(x), _dp(y1), reps))
    t_orc = _best(lambda: [L.orc_bsr_mult(3, nr, _ip(rp), _ip(cp), _dp(A), _dp(x), _dp(y2))
                           for _ in range(reps)])
    assert np.array_equal(y1, y2) or np.abs(y1 - y2).max() <= 1e-14 * np.abs(y1).max()
    y3 = np.zeros_like(x)
    t_par = _best(lambda: [L.orc_par_bsr_mult(3, nr, _ip(rp), _ip(cp), _dp(A), _dp(x), _dp(y3))
                           for _ in range(reps)])
    assert np.abs(y1 - y3).max() <= 1e-14 * np.abs(y1).max()
    return t_ref / reps, t_orc / reps, cols.size, t_par / reps


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} is missing: run oracle/build_ref.sh where the reference tree is")
    R = ctypes.CDLL(REF)
    O.lib()
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    print(f"CPU: {cpu}; one core each (OMP_NUM_THREADS=1); best of 5")
    print("| loop | reference | oracle (checker) | oracle / reference | baseline kernel "
          "(cpu_baseline.c) | baseline / reference | results |")
    print("|---|---|---|---|---|---|---|")
    for order, ne in ((2, 20000), (3, 1000)):
        tr, to, agree, ti, agree_i = element_loop(R, order, ne)
        print(f"| {'8' if order == 2 else '27'}-node element Jacobian ({ne} elements) | "
              f"{1e6 * tr:.1f} us/element | {1e6 * to:.1f} us/element | {to / tr:.2f} | "
              f"{1e6 * ti:.1f} us/element | {ti / tr:.2f} | "
              f"matrices agree to {max(agree, agree_i):.1e} |")
    tr, to, agree = quad4_loop(R, 20000)
    print(f"| Quad4 shell element Jacobian (20000 elements) | {1e6 * tr:.1f} us/element | "
          f"{1e6 * to:.1f} us/element | {to / tr:.2f} | the oracle's (no separate kernel) | "
          f"{to / tr:.2f} | matrices agree to {agree:.1e} |")
    tr, to, nnzb, tp = spmv(R, 40)
    print(f"| 3x3 block-CSR product (40^3 hex8, {nnzb} blocks) | {1e3 * tr:.2f} ms | "
          f"{1e3 * to:.2f} ms | {to / tr:.2f} | {1e3 * tp:.2f} ms | {tp / tr:.2f} | equal |")


if __name__ == "__main__":
    main()
