"""Error paths on the GPU: an allocation the device cannot hold comes back as
an error with NULL instead of a half-built object, and the library keeps
working afterwards; a held host view of a vector keeps reaching the device
(write, norm, write, norm)."""
import numpy as np
import pytest

from tacs_amd import TACS, lib
from tacs_amd.problems import SolidProblem

pytestmark = pytest.mark.gpu


def test_absurd_allocation_returns_null_and_library_recovers():
    L = lib()
    # 64 variables on 2^31 - 1 nodes: a 1.1 TB vector, beyond the 288 GB HBM
    a = L.tacs_assembler_create(64, (1 << 31) - 1, 1, 0)
    L.tacs_incref(a)
    assert not L.tacs_assembler_create_vec(a)
    assert b"hipMalloc" in L.tacs_last_error()
    L.tacs_decref(a)
    # the next operation starts clean
    p = SolidProblem(2)
    p.assemble()
    x = p.assembler.createVec()
    x.set(2.0)
    assert abs(x.norm() - 2.0 * np.sqrt(x.getSize())) <= 1e-12 * x.norm()


def test_held_host_view_reaches_the_device():
    p = SolidProblem(2)
    v = p.assembler.createVec()
    h = v.getArray()
    n = h.size
    h[:] = 1.0
    assert abs(v.norm() - np.sqrt(n)) <= 1e-12 * np.sqrt(n)
    h[:] = 3.0  # written after a device read: still uploaded before the next
    assert abs(v.norm() - 3.0 * np.sqrt(n)) <= 1e-12 * np.sqrt(n)
    w = p.assembler.createVec()
    w.set(1.0)
    h[:] = 2.0
    v.axpy(1.0, w)  # in-place device update reads the held values
    assert np.all(v.getArray() == 3.0)
