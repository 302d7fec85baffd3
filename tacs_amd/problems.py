"""Builders for the structured solid problems of the benchmark and tests, using
only the reference-style API (TACS.Assembler, elements, constitutive,
functions), as a pyTACS user script would (cf.
tests/integration_tests/test_elast_linhexa_element_3d.py:62-120)."""
import numpy as np

from . import TACS, constitutive, elements, functions
from .meshes import HexMesh


class SolidProblem:
    """A clamped box of hexahedra with a traction on the x = L face.

    dv: None (fixed t), "shared" (one DV for every element, tNum = 0) or
    "element" (one DV per element, tNum = element index).
    traction_patch: None (whole x = L face) or n (the centred n x n patch).
    materials: a list of dicts (E, nu, rho, ys); element e takes
    materials[e % len] (one constitutive object per element).
    """

    def __init__(self, n, order=2, L=1.0, E=70e9, nu=0.3, rho=2700.0, ys=270e6,
                 traction=(0.0, 0.0, 1e6), t0=1.0, dv=None, traction_patch=None, shape=None,
                 materials=None):
        nx, ny, nz = shape if shape is not None else (n, n, n)
        self.mesh = m = HexMesh(nx, ny, nz, order=order, L=(L * nx / nz, L * ny / nz, L)
                                if shape is not None else (L, L, L))
        self.order = order
        self.props = constitutive.MaterialProperties(rho=rho, E=E, nu=nu, ys=ys)
        self.basis = elements.LinearHexaBasis() if order == 2 else elements.QuadraticHexaBasis()
        a = self.assembler = TACS.Assembler.create(None, 3, m.nnodes, m.nelem)
        ptr = np.arange(0, m.npe * m.nelem + 1, m.npe, dtype=np.intc)
        a.setElementConnectivity(ptr, m.conn.reshape(-1))
        if materials is not None:
            self.mat_props = [constitutive.MaterialProperties(rho=d["rho"], E=d["E"], nu=d["nu"],
                                                              ys=d["ys"]) for d in materials]
            elems = []
            for e in range(m.nelem):
                tnum = e if dv == "element" else (0 if dv == "shared" else -1)
                con = constitutive.SolidConstitutive(self.mat_props[e % len(materials)], t=t0,
                                                     tNum=tnum)
                elems.append(elements.Element3D(elements.LinearElasticity3D(con), self.basis))
        elif dv == "element":
            elems = []
            for e in range(m.nelem):
                con = constitutive.SolidConstitutive(self.props, t=t0, tNum=e)
                elems.append(elements.Element3D(elements.LinearElasticity3D(con), self.basis))
        else:
            con = constitutive.SolidConstitutive(self.props, t=t0, tNum=0 if dv == "shared" else -1)
            elem = elements.Element3D(elements.LinearElasticity3D(con), self.basis)
            elems = [elem] * m.nelem
        a.setElements(elems)
        self.bc_nodes = m.nodes_on_x(0)
        a.addBCs(self.bc_nodes)
        a.initialize()
        self.X = a.createNodeVec()
        self.X.getArray()[:] = m.X.reshape(-1)
        a.setNodes(self.X)
        # traction aux elements on the x = L face (face index 1 = +x)
        if traction_patch is None:
            self.aux_elems, self.aux_faces = m.elems_on_x(1)
        else:
            self.aux_elems, self.aux_faces = m.elems_on_x1_patch(traction_patch)
        self.traction = np.asarray(traction, dtype=float)
        self.aux = TACS.AuxElements()
        tr = elements.Traction3D(3, 1, self.basis, list(self.traction))
        for e in self.aux_elems:
            self.aux.addElement(int(e), tr)
        a.setAuxElements(self.aux)
        self.res = a.createVec()
        self.u = a.createVec()
        self.mat = a.createMat()

    def set_thickness(self, t):
        """Design vector values (one per DV)."""
        x = self.assembler.createDesignVec()
        x.getArray()[:] = t
        self.assembler.setDesignVars(x)
        return x

    def assemble(self):
        a = self.assembler
        a.zeroVariables()
        a.assembleJacobian(1.0, 0.0, 0.0, self.res, self.mat)

    def solve_fp64(self, m=30, restarts=500, rtol=1e-8, ortho=TACS.MODIFIED_GRAM_SCHMIDT):
        pc = TACS.Pc(self.mat)
        pc.factor()
        ksm = TACS.KSM(self.mat, pc, m, restarts)
        ksm.setTolerances(rtol, 1e-30)
        if ortho != TACS.MODIFIED_GRAM_SCHMIDT:
            ksm.setOrthoType(ortho)
        flag = ksm.solve(self.res, self.u)
        self.u.scale(-1.0)
        self.assembler.setVariables(self.u)
        return flag, ksm.getIterCount(), pc, ksm

    def solve_mixed(self, m=30, restarts=1000, rtol=1e-8, deflation=0):
        pc = TACS.Pc(self.mat)
        pc.factor()
        ksm = TACS.MixedKSM(self.mat, pc, m, restarts, deflation=deflation)
        ksm.setTolerances(rtol, 1e-30)
        flag = ksm.solve(self.res, self.u)
        self.u.scale(-1.0)
        self.assembler.setVariables(self.u)
        return flag, ksm.getIterCount(), pc, ksm

    def compliance(self):
        f = functions.Compliance(self.assembler)
        return self.assembler.evalFunctions([f])[0], f


class ShellProblem:
    """A flat square plate of Quad4 shells (6 dof per node) under uniform
    pressure, as tests/integration_tests/test_shell_plate_quad.py builds it
    with pyTACS (TACSQuad4Shell, TACSIsoShellConstitutive, natural transform).

    bc: "ss" (simply supported with restrained in-plane motion: u = v = w = 0
    on the perimeter, rotations free) or "clamped" (all six dofs of the
    perimeter).  dv: None, "shared" or "element" as for SolidProblem.
    warp: amplitude of a seeded random z perturbation of the interior nodes
    (0 keeps the plate flat).  axis: reference axis of a
    TACSShellRefAxisTransform (None: the natural transform).  toff: the
    constitutive model's thickness offset (tOffset).  materials: a list of
    dicts (E, nu, rho, ys, toff); element e takes materials[e % len] (one
    constitutive object per element, dv as above).
    """

    def __init__(self, n, L=1.0, E=70e9, nu=0.3, rho=2700.0, ys=270e6, t0=0.01, dv=None,
                 pressure=1e4, bc="ss", warp=0.0, seed=0, axis=None, toff=0.0, materials=None):
        from .meshes import PlateMesh
        self.mesh = m = PlateMesh(n, n, L, L)
        if warp:
            rng = np.random.default_rng(seed)
            interior = np.setdiff1d(np.arange(m.nnodes), m.edge_nodes())
            m.X[interior, 2] += rng.uniform(-warp, warp, interior.size)
        self.props = constitutive.MaterialProperties(rho=rho, E=E, nu=nu, ys=ys)
        self.transform = (elements.ShellNaturalTransform() if axis is None
                          else elements.ShellRefAxisTransform(axis))
        a = self.assembler = TACS.Assembler.create(None, 6, m.nnodes, m.nelem)
        ptr = np.arange(0, 4 * m.nelem + 1, 4, dtype=np.intc)
        a.setElementConnectivity(ptr, m.conn.reshape(-1))
        if materials is not None:
            self.mat_props = [(constitutive.MaterialProperties(rho=d["rho"], E=d["E"], nu=d["nu"],
                                                               ys=d["ys"]), d.get("toff", 0.0))
                              for d in materials]
            elems = []
            for e in range(m.nelem):
                pm, to = self.mat_props[e % len(materials)]
                tnum = e if dv == "element" else (0 if dv == "shared" else -1)
                elems.append(elements.Quad4Shell(
                    self.transform, constitutive.IsoShellConstitutive(pm, t=t0, tNum=tnum,
                                                                      tOffset=to)))
        elif dv == "element":
            elems = [elements.Quad4Shell(self.transform,
                                         constitutive.IsoShellConstitutive(self.props, t=t0, tNum=e,
                                                                           tOffset=toff))
                     for e in range(m.nelem)]
        else:
            con = constitutive.IsoShellConstitutive(self.props, t=t0,
                                                    tNum=0 if dv == "shared" else -1,
                                                    tOffset=toff)
            elems = [elements.Quad4Shell(self.transform, con)] * m.nelem
        self.elems = elems
        a.setElements(elems)
        self.bc_nodes = m.edge_nodes()
        self.bc_vars = [0, 1, 2] if bc == "ss" else [0, 1, 2, 3, 4, 5]
        a.addBCs(self.bc_nodes, self.bc_vars)
        a.initialize()
        self.X = a.createNodeVec()
        self.X.getArray()[:] = m.X.reshape(-1)
        a.setNodes(self.X)
        self.pressure = pressure
        self.aux = TACS.AuxElements()
        pr = elems[0].createElementPressure(0, pressure)
        self.aux_elems = np.arange(m.nelem, dtype=np.int32)
        for e in range(m.nelem):
            self.aux.addElement(e, pr)
        a.setAuxElements(self.aux)
        self.res = a.createVec()
        self.u = a.createVec()
        self.mat = a.createMat()

    set_thickness = SolidProblem.set_thickness
    assemble = SolidProblem.assemble
    solve_fp64 = SolidProblem.solve_fp64
    solve_mixed = SolidProblem.solve_mixed
    compliance = SolidProblem.compliance


class MultiCubeProblem:
    """Many independent 8-node hex cubes in one assembler: the disjoint union
    of the meshes (node and element ids offset cube by cube), so one
    assembly launch covers every cube and the block-diagonal stiffness is one
    BCSR matrix.  Cube c has sizes[c] elements per edge, unit edge length,
    x = 0 clamped and the traction `traction` on its x = 1 face (the load of
    the 128^3 workload).  dv: None, "shared" or "element" as for SolidProblem.

    node_off / elem_off give each cube's first node and element; dof(c) the
    slice of its dofs in assembler vectors.
    """

    def __init__(self, sizes, E=70e9, nu=0.3, rho=2700.0, ys=270e6,
                 traction=(0.0, 0.0, 1e6), t0=1.0, dv=None):
        self.sizes = [int(n) for n in sizes]
        self.meshes = [HexMesh(n, n, n) for n in self.sizes]
        nn = np.array([m.nnodes for m in self.meshes])
        ne = np.array([m.nelem for m in self.meshes])
        self.node_off = np.concatenate([[0], np.cumsum(nn)]).astype(np.int64)
        self.elem_off = np.concatenate([[0], np.cumsum(ne)]).astype(np.int64)
        nnodes, nelem = int(self.node_off[-1]), int(self.elem_off[-1])
        self.nnodes, self.nelem = nnodes, nelem
        X = np.concatenate([m.X + np.array([2.0 * c, 0.0, 0.0])
                            for c, m in enumerate(self.meshes)])
        conn = np.concatenate([m.conn + self.node_off[c] for c, m in enumerate(self.meshes)])
        self.X_all, self.conn = X, conn.astype(np.int32)
        self.props = constitutive.MaterialProperties(rho=rho, E=E, nu=nu, ys=ys)
        self.basis = elements.LinearHexaBasis()
        a = self.assembler = TACS.Assembler.create(None, 3, nnodes, nelem)
        ptr = np.arange(0, 8 * nelem + 1, 8, dtype=np.intc)
        a.setElementConnectivity(ptr, self.conn.reshape(-1))
        if dv == "element":
            elems = []
            for e in range(nelem):
                con = constitutive.SolidConstitutive(self.props, t=t0, tNum=e)
                elems.append(elements.Element3D(elements.LinearElasticity3D(con), self.basis))
        else:
            con = constitutive.SolidConstitutive(self.props, t=t0,
                                                 tNum=0 if dv == "shared" else -1)
            elems = [elements.Element3D(elements.LinearElasticity3D(con), self.basis)] * nelem
        a.setElements(elems)
        self.bc_nodes = np.concatenate([m.nodes_on_x(0) + self.node_off[c]
                                        for c, m in enumerate(self.meshes)]).astype(np.int32)
        a.addBCs(self.bc_nodes)
        a.initialize()
        self.X = a.createNodeVec()
        self.X.getArray()[:] = X.reshape(-1)
        a.setNodes(self.X)
        self.traction = np.asarray(traction, dtype=float)
        ae = [m.elems_on_x(1)[0] + self.elem_off[c] for c, m in enumerate(self.meshes)]
        self.aux_elems = np.concatenate(ae).astype(np.int32)
        self.aux_faces = np.ones(self.aux_elems.size, dtype=np.int32)
        self.aux = TACS.AuxElements()
        tr = elements.Traction3D(3, 1, self.basis, list(self.traction))
        for e in self.aux_elems:
            self.aux.addElement(int(e), tr)
        a.setAuxElements(self.aux)
        self.res = a.createVec()
        self.u = a.createVec()
        self.mat = a.createMat()

    def dof(self, c):
        """Slice of cube c's dofs in an assembler vector."""
        return slice(3 * int(self.node_off[c]), 3 * int(self.node_off[c + 1]))

    def elems(self, c):
        return slice(int(self.elem_off[c]), int(self.elem_off[c + 1]))

    def scale_cubes(self, b, out):
        """out = b with each cube's part scaled to unit norm (zero parts kept);
        the scales are kept for unscale_cubes."""
        bv = b.getArray()
        ov = out.getArray()
        self._scales = []
        for c in range(len(self.sizes)):
            sl = self.dof(c)
            nrm = np.linalg.norm(bv[sl])
            s = 1.0 / nrm if nrm > 0 else 1.0
            ov[sl] = s * bv[sl]
            self._scales.append(s)

    def unscale_cubes(self, x):
        """Undo scale_cubes on a solution of the scaled system."""
        xv = x.getArray()
        for c, s in enumerate(self._scales):
            xv[self.dof(c)] /= s

    def cube_atol(self, b, rtol):
        """Absolute tolerance on the union residual that brings every cube to
        |r_c| <= rtol |b_c| (since |r_c| <= |r|)."""
        bv = b.getArray()
        return rtol * min(np.linalg.norm(bv[self.dof(c)]) for c in range(len(self.sizes)))

    set_thickness = SolidProblem.set_thickness
    assemble = SolidProblem.assemble
