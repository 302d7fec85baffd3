"""Structured synthetic meshes used by the benchmark and the tests.

Node and element numbering follow the reference's integration tests
(tests/integration_tests/test_elast_linhexa_element_3d.py:82-104): node (i,j,k)
of an (nx+1) x (ny+1) x (nz+1) grid has id (i*(ny+1)+j)*(nz+1)+k (z fastest);
element (i,j,k) has id (i*ny+j)*nz+k and local node order x fastest
(TACSHexaBasis.cpp:137-146).
"""
import numpy as np


class HexMesh:
    """A structured box of hexahedra.

    Attributes: order (2 = 8-node, 3 = 27-node), shape (nx, ny, nz) in
    elements, X (nnodes, 3), conn (nelem, order**3) int32, plus the index
    helpers used to set boundary conditions and tractions.
    """

    def __init__(self, nx, ny, nz, order=2, L=(1.0, 1.0, 1.0)):
        self.order = order
        self.shape = (nx, ny, nz)
        p = order - 1
        gx, gy, gz = p * nx + 1, p * ny + 1, p * nz + 1
        self.grid = (gx, gy, gz)
        x = np.linspace(0.0, L[0], gx)
        y = np.linspace(0.0, L[1], gy)
        z = np.linspace(0.0, L[2], gz)
        X = np.empty((gx, gy, gz, 3))
        X[..., 0], X[..., 1], X[..., 2] = np.meshgrid(x, y, z, indexing="ij")
        self.X = X.reshape(-1, 3)
        self.L = L
        ids = np.arange(gx * gy * gz, dtype=np.int64).reshape(gx, gy, gz)
        self.node_ids = ids
        ei, ej, ek = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                                 indexing="ij")
        ei, ej, ek = ei.reshape(-1), ej.reshape(-1), ek.reshape(-1)
        cols = []
        for c in range(order):
            for b in range(order):
                for a in range(order):
                    cols.append(ids[p * ei + a, p * ej + b, p * ek + c])
        self.conn = np.stack(cols, axis=1).astype(np.int32)
        self.elem_ijk = np.stack([ei, ej, ek], axis=1)

    @property
    def nnodes(self):
        return self.X.shape[0]

    @property
    def nelem(self):
        return self.conn.shape[0]

    @property
    def npe(self):
        return self.order ** 3

    def nodes_on_x(self, side):
        """Node ids on the x=0 (side 0) or x=L (side 1) face, ascending."""
        gi = 0 if side == 0 else self.grid[0] - 1
        return np.sort(self.node_ids[gi].reshape(-1)).astype(np.int32)

    def elems_on_x(self, side):
        """Element ids adjacent to the x=0 / x=L face and the matching face
        index (0 = -x, 1 = +x in the reference's face numbering)."""
        i = 0 if side == 0 else self.shape[0] - 1
        sel = np.nonzero(self.elem_ijk[:, 0] == i)[0].astype(np.int32)
        return sel, np.full(sel.size, side, dtype=np.int32)

    def elems_on_x1_patch(self, n):
        """Elements on the x=L face whose (j,k) indices lie in the n x n
        patch centred on the face (used for point-like tractions)."""
        nx, ny, nz = self.shape
        j0, k0 = (ny - n) // 2, (nz - n) // 2
        e = self.elem_ijk
        sel = np.nonzero((e[:, 0] == nx - 1) & (e[:, 1] >= j0) & (e[:, 1] < j0 + n)
                         & (e[:, 2] >= k0) & (e[:, 2] < k0 + n))[0]
        return sel.astype(np.int32), np.ones(sel.size, dtype=np.int32)


def simp_thickness(x, p=3.0, E0=1.0, Emin=1e-9):
    """SIMP stiffness scale t(x) = (Emin + x^p (E0 - Emin)) / E0 applied on top
    of the reference's linear-in-t solid constitutive model."""
    return (Emin + x ** p * (E0 - Emin)) / E0


def simp_dthickness(x, p=3.0, E0=1.0, Emin=1e-9):
    return p * x ** (p - 1.0) * (E0 - Emin) / E0


class PlateMesh:
    """A flat structured plate of 4-node shell elements in the z = 0 plane.

    Node (i, j) (i along x, 0..nx; j along y, 0..ny) has id j*(nx+1) + i;
    element (i, j) has id j*nx + i and local nodes (i,j), (i+1,j), (i,j+1),
    (i+1,j+1) -- the tensor-product order of TACSShellQuadBasis<2>
    (src/elements/shell/TACSShellElementQuadBasis.h:147-150).
    """

    def __init__(self, nx, ny, Lx=1.0, Ly=1.0, origin=(0.0, 0.0)):
        self.shape = (nx, ny)
        x = origin[0] + np.linspace(0.0, Lx, nx + 1)
        y = origin[1] + np.linspace(0.0, Ly, ny + 1)
        Y, Xg = np.meshgrid(y, x, indexing="ij")
        self.X = np.stack([Xg.reshape(-1), Y.reshape(-1), np.zeros(Xg.size)], axis=1)
        ids = np.arange((nx + 1) * (ny + 1), dtype=np.int64).reshape(ny + 1, nx + 1)
        self.node_ids = ids
        ej, ei = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        ei, ej = ei.reshape(-1), ej.reshape(-1)
        self.conn = np.stack([ids[ej, ei], ids[ej, ei + 1], ids[ej + 1, ei],
                              ids[ej + 1, ei + 1]], axis=1).astype(np.int32)
        self.elem_ij = np.stack([ei, ej], axis=1)

    @property
    def nnodes(self):
        return self.X.shape[0]

    @property
    def nelem(self):
        return self.conn.shape[0]

    npe = 4

    def edge_nodes(self):
        """Ids of the nodes on the plate's perimeter, ascending."""
        ids = self.node_ids
        edge = np.concatenate([ids[0], ids[-1], ids[:, 0], ids[:, -1]])
        return np.unique(edge).astype(np.int32)

    def center_node(self):
        nx, ny = self.shape
        return int(self.node_ids[ny // 2, nx // 2])
