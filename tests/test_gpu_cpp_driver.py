"""The C++ boundary from the consumer side: tests/cpp/tacs_cube_driver.cpp,
written against the reference's class names (TACSAssembler with MPI_COMM_WORLD,
TACSElement3D, GMRES with and without isFlexible, TACSCompliance, ...), built
against include/TACSHip.h, runs on the GPU and its compliance, solutions and
thickness gradient are checked against the oracle and closed forms."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from tacs_amd.meshes import HexMesh

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "tacs_cube_driver")


def test_cpp_driver_against_oracle():
    assert os.path.exists(DRIVER), "build it with tests/cpp/build.sh"
    n = 4
    out = subprocess.run([DRIVER, str(n)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    r = json.loads(out.stdout.strip().splitlines()[-1])
    assert r["flag"] == 1 and r["flexible_flag"] == 1 and r["adjoint_flag"] == 1
    # a fixed preconditioner: flexible and plain GMRES give the same solution
    assert r["flexible_vs_plain"] <= 1e-9
    assert abs(r["flexible_iterations"] - r["iterations"]) <= 1
    # oracle compliance of the same cube (dense solve on the free dofs)
    m = HexMesh(n, n, n)
    C = O.iso_C21(70e9, 0.3)
    X = m.X.reshape(-1)
    rowp, cols = O.node_csr(m.nnodes, m.conn, 8)
    A, res = O.assemble_solid(2, m.conn, X, np.zeros(3 * m.nnodes), np.ones(m.nelem), C, rowp,
                              cols)
    ae, af = m.elems_on_x(1)
    O.add_tractions(2, m.conn, X, ae, af, np.tile([0.0, 0.0, 1e6], (ae.size, 1)), res)
    K = np.zeros((3 * m.nnodes, 3 * m.nnodes))
    Ab = A.reshape(-1, 3, 3)
    for i in range(m.nnodes):
        for j in range(rowp[i], rowp[i + 1]):
            K[3 * i:3 * i + 3, 3 * cols[j]:3 * cols[j] + 3] = Ab[j]
    free = np.ones(3 * m.nnodes, bool)
    for b in m.nodes_on_x(0):
        free[3 * b:3 * b + 3] = False
    u = np.zeros(3 * m.nnodes)
    u[free] = -np.linalg.solve(K[np.ix_(free, free)], res[free])
    c = O.compliance(2, m.conn, X, u, np.ones(m.nelem), C)
    assert abs(r["compliance"] - c) <= 1e-9 * abs(c)
    # one shared thickness t = 1: C = f^T K0^-1 f / t, so dC/dt = -C
    assert abs(r["dcdt"] + r["compliance"]) <= 1e-8 * abs(r["compliance"])
    assert r["late_connectivity_rejected"] == 1
    assert r["assembler_refs"] == 2  # the driver's reference and the function's
