"""CPU-only checks of the native library and the host-side API: the C-ABI
library loads and exports every symbol include/tacs_hip.h declares, handles
are reference counted, and argument errors are reported the reference's way
(non-zero status + message) without touching a GPU."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from tacs_amd import TACS, constitutive, elements, TACSError
from tacs_amd._lib import lib, parse_header, LIB_PATH
from tacs_amd.meshes import HexMesh


def test_library_exports_every_declared_symbol():
    decls = parse_header()
    assert len(decls) > 90
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB_PATH]).decode()
    syms = {l.split()[-1] for l in out.splitlines()}
    missing = [k for k in decls if k not in syms]
    assert not missing, missing
    L = lib()
    for k in decls:
        assert getattr(L, k) is not None


def test_library_is_gfx950_code():
    # the offload bundle carries gfx950 code objects
    data = open(LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_refcounting_follows_tacsobject():
    L = lib()
    props = constitutive.MaterialProperties(E=70e9, nu=0.3)
    assert L.tacs_refcount(props.ptr) == 1  # held by the Python wrapper
    con = constitutive.SolidConstitutive(props, t=1.0, tNum=-1)
    assert L.tacs_refcount(props.ptr) == 2  # the constitutive object holds one
    del con
    assert L.tacs_refcount(props.ptr) == 1


def test_assembler_argument_errors():
    m = HexMesh(2, 2, 2)
    a = TACS.Assembler.create(None, 3, m.nnodes, m.nelem)
    with pytest.raises(ValueError):
        a.setElementConnectivity(np.arange(0, 8 * 3 + 1, 8), m.conn.reshape(-1))
    bad = m.conn.reshape(-1).copy()
    bad[0] = m.nnodes + 5
    with pytest.raises(TACSError):
        a.setElementConnectivity(np.arange(0, 8 * m.nelem + 1, 8), bad)
    with pytest.raises(TACSError):
        a.addBCs(np.array([m.nnodes + 1], dtype=np.intc))
    with pytest.raises(ValueError):
        a.setElements([])
    # a traction on a non-existent face is rejected
    with pytest.raises(TACSError):
        elements.Traction3D(3, 7, elements.LinearHexaBasis(), [0, 0, 1])
    assert a.getNumElements() == m.nelem and a.getVarsPerNode() == 3


def test_ksm_monitor_objects(tmp_path):
    """KSMPrintStdout / KSMPrintFile (src/bpmat/KSM.cpp:239-330): reference
    counted handles; a file that cannot be opened is an error; set_monitor
    rejects a handle that is not a Krylov solver or not a monitor."""
    L = lib()
    m = L.tacs_ksm_print_stdout_create(b"GMRES", 0, 10)
    assert m and L.tacs_refcount(m) == 0
    L.tacs_incref(m)
    assert L.tacs_refcount(m) == 1
    f = L.tacs_ksm_print_file_create(str(tmp_path / "mon.txt").encode(), b"x", 0, 5)
    assert f and (tmp_path / "mon.txt").exists()
    L.tacs_incref(f)
    assert not L.tacs_ksm_print_file_create(str(tmp_path / "no" / "dir.txt").encode(), b"x", 0, 5)
    assert b"cannot open" in L.tacs_last_error()
    props = constitutive.MaterialProperties(E=70e9, nu=0.3)
    assert L.tacs_ksm_set_monitor(props.ptr, m) != 0
    L.tacs_decref(f)
    L.tacs_decref(m)


def test_header_parser_types():
    d = parse_header()
    res, args = d["tacs_hip_bsr_spmv"]
    assert res is ctypes.c_int and len(args) == 9
    assert args[0] is ctypes.c_int and args[4] is ctypes.c_void_p
    res, args = d["tacs_vec_norm"]
    assert res is ctypes.c_double


def test_mesh_numbering_matches_reference_test_convention():
    # tests/integration_tests/test_elast_linhexa_element_3d.py:82-104
    nx = ny = nz = 3
    ids = np.arange((nx + 1) * (ny + 1) * (nz + 1)).reshape(nx + 1, ny + 1, nz + 1)
    conn = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                conn.append([ids[i, j, k], ids[i + 1, j, k], ids[i, j + 1, k], ids[i + 1, j + 1, k],
                             ids[i, j, k + 1], ids[i + 1, j, k + 1], ids[i, j + 1, k + 1],
                             ids[i + 1, j + 1, k + 1]])
    m = HexMesh(nx, ny, nz, L=(10.0, 10.0, 10.0))
    assert np.array_equal(m.conn, np.array(conn))
    x = np.linspace(0, 10.0, 4)
    xyz = np.zeros((4, 4, 4, 3))
    xyz[..., 0], xyz[..., 1], xyz[..., 2] = np.meshgrid(x, x, x, indexing="ij")
    assert np.array_equal(m.X, xyz.reshape(-1, 3))
    assert np.array_equal(m.nodes_on_x(0), np.sort(ids[0].reshape(-1)))


@pytest.mark.parametrize("n,k", [(12, 4), (81, 20), (151, 30)])
def test_dense_smallest_eigvecs(n, k):
    """Host eigen-solver of the recycling GMRES (Hessenberg reduction,
    Francis QR, inverse iteration on the Hessenberg form): the returned
    columns span an invariant subspace whose eigenvalues are the k of
    smallest modulus (numpy's), complex pairs kept whole."""
    import ctypes

    from tacs_amd import lib

    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n)) + 0.3 * n * np.eye(n)
    V = np.zeros((k + 1) * n)
    wr, wi = np.zeros(n), np.zeros(n)
    kout = lib().tacs_dense_smallest_eigvecs(n, G.ctypes.data, k, V.ctypes.data, wr.ctypes.data,
                                             wi.ctypes.data)
    assert k <= kout <= k + 1
    Vm = V[:kout * n].reshape(kout, n).T
    T = np.linalg.lstsq(Vm, G @ Vm, rcond=None)[0]
    assert np.linalg.norm(G @ Vm - Vm @ T) <= 1e-12 * np.linalg.norm(G @ Vm)
    ev = np.linalg.eigvals(G)
    ev = ev[np.argsort(np.abs(ev), kind="stable")]
    got = np.sort_complex(np.linalg.eigvals(T))
    assert np.abs(got - np.sort_complex(ev[:kout])).max() <= 1e-10 * np.abs(ev).max()
    # all eigenvalues, ascending modulus
    w = wr + 1j * wi
    assert np.all(np.diff(np.abs(w)) >= -1e-12 * np.abs(w).max())
    assert np.abs(np.sort_complex(w) - np.sort_complex(ev)).max() <= 1e-10 * np.abs(ev).max()


def _has_gpu():
    import torch

    return torch.cuda.device_count() > 0


def test_failed_device_allocation_is_an_error_with_no_launch():
    """With no GPU every hipMalloc fails: initialize() reports it, leaves the
    assembler uninitialised, creators return NULL instead of half-built
    objects, and device work on the assembler is refused before any launch."""
    if _has_gpu():
        pytest.skip("allocations succeed on a GPU (tests/test_gpu_errors.py covers this there)")
    L = lib()
    m = HexMesh(2, 2, 2)
    a = TACS.Assembler.create(None, 3, m.nnodes, m.nelem)
    a.setElementConnectivity(np.arange(0, 8 * m.nelem + 1, 8), m.conn.reshape(-1))
    props = constitutive.MaterialProperties(E=70e9, nu=0.3)
    el = elements.Element3D(elements.LinearElasticity3D(constitutive.SolidConstitutive(props)),
                            elements.LinearHexaBasis())
    a.setElements([el] * m.nelem)
    with pytest.raises(TACSError, match="hipMalloc"):
        a.initialize()
    assert not L.tacs_assembler_create_vec(a.ptr)
    assert b"hipMalloc" in L.tacs_last_error()
    assert not L.tacs_assembler_create_mat(a.ptr)
    assert b"initialize" in L.tacs_last_error()
    assert L.tacs_assembler_assemble_jacobian(a.ptr, 1.0, 0.0, 0.0, None, None, 0, 1.0, 1) != 0
    assert L.tacs_assembler_zero_variables(a.ptr) != 0
    assert b"not initialized" in L.tacs_last_error()
    assert not L.tacs_matfree_create(a.ptr)


def test_element_index_packing_limit():
    """The node-to-element CSR packs (element << 5 | local node) in int32:
    2^26 elements or more are rejected before anything is read or built."""
    L = lib()
    a = L.tacs_assembler_create(3, 8, 1 << 26, 0)
    L.tacs_incref(a)
    ptr = np.arange(0, 9, 8, dtype=np.intc)
    conn = np.arange(8, dtype=np.intc)
    assert L.tacs_assembler_set_element_connectivity(a, ptr.ctypes.data, conn.ctypes.data) != 0
    assert b"2^26" in L.tacs_last_error()
    L.tacs_decref(a)
    b = L.tacs_assembler_create(3, 8, (1 << 26) - 1, 0)
    L.tacs_incref(b)
    L.tacs_decref(b)


def test_cpp_driver_builds_against_the_reference_names():
    """tests/cpp/tacs_cube_driver.cpp (reference class names, MPI_COMM_WORLD,
    GMRES isFlexible) compiles against include/TACSHip.h and links the
    library; without a GPU its set-up stops at initialize() with the
    allocation error, the reference's int/stderr convention."""
    import os
    import subprocess

    drv = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "tacs_cube_driver")
    assert os.path.exists(drv), "tests/cpp/build.sh builds it (part of __graft_entry__.build)"
    libs = subprocess.check_output(["ldd", drv]).decode()
    assert "libtacs_hip.so" in libs and "libmpi" in libs
    if _has_gpu():
        pytest.skip("the GPU run is tests/test_gpu_cpp_driver.py")
    r = subprocess.run([drv, "2"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "set-up failed" in r.stdout and "hipMalloc" in r.stderr


def test_tuning_table_host_semantics(tmp_path):
    """The tuning table is host state: set / get / reset, size classes,
    argument checks and the JSON loader (no GPU involved)."""
    import json

    from tacs_amd import TACS

    L = TACS.lib()
    assert [L.tacs_tune_size_class(r) for r in (1, 32767, 32768, 524287, 524288)] == [0, 0, 1, 1, 2]
    # built-in defaults: 3x3 FP64 U = 2 (variant 1), FP32 U = 4 (variant 2),
    # 6x6 U = 1 (0); 27-node Jacobian the element-block form (4)
    saved = [(k, sh, c, TACS.getTuning(k, sh, c)) for k, sh in ((0, 30), (0, 31), (2, 3))
             for c in range(3)]
    TACS.resetTuning()
    assert L.tacs_tune_variant(0, 30, 10) == 1
    assert L.tacs_tune_variant(0, 31, 10**6) == 2
    assert L.tacs_tune_variant(0, 60, 10) == 0
    assert L.tacs_tune_variant(2, 3, 10) == 4
    assert L.tacs_tune_variant(2, 2, 10) == 0
    TACS.setTuning(0, 30, 2, 5)
    assert TACS.getTuning(0, 30, 2) == 5 and TACS.getTuning(0, 30, 1) == -1
    assert L.tacs_tune_variant(0, 30, 10**6) == 5 and L.tacs_tune_variant(0, 30, 10) == 1
    with pytest.raises(TACS.TACSError):
        TACS.setTuning(0, 30, 3, 1)
    with pytest.raises(TACS.TACSError):
        TACS.setTuning(7, 30, 0, 1)
    path = tmp_path / "table.json"
    path.write_text(json.dumps({"entries": [
        {"kernel_id": 0, "shape": 31, "size_class": 1, "variant": 6},
        {"kernel_id": 1, "shape": 116, "size_class": 2, "variant": 3}]}))
    assert TACS.loadTuningTable(str(path)) == 2
    assert TACS.getTuning(0, 31, 1) == 6 and TACS.getTuning(1, 116, 2) == 3
    assert TACS.loadTuningTable(str(tmp_path / "missing.json")) == 0
    TACS.resetTuning()
    assert TACS.getTuning(0, 31, 1) == -1
    for k, sh, c, v in saved:
        TACS.setTuning(k, sh, c, v)


def test_committed_tuning_table_is_well_formed():
    """Every entry of the committed table names a valid variant of its
    kernel (variant counts as the launchers define them)."""
    import json

    from tacs_amd import TACS

    if not os.path.exists(TACS.TUNING_TABLE):
        pytest.skip("no tuning table committed")
    table = json.load(open(TACS.TUNING_TABLE))
    L = TACS.lib()
    for e in table["entries"]:
        k, sh, v = e["kernel_id"], e["shape"], e["variant"]
        if k == TACS.TUNE_SELL_SPMV:
            nv = L.tacs_hip_sell_spmv_nvariants(sh // 10)
        elif k == TACS.TUNE_SPMM:
            nv = L.tacs_hip_bsr3_spmm_nvariants(sh % 100)
        else:
            nv = L.tacs_hip_solid_jacobian_nvariants(sh) + 1
        assert 0 <= v < nv and 0 <= e["size_class"] < 3, e
        assert len(e["ms"]) == nv and min(range(nv), key=lambda i: e["ms"][i]) == v, e
