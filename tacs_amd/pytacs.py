"""Nastran-model front end in the shape of the reference's pyTACS
(tacs/pytacs.py) and StaticProblem (tacs/problems/static.py), for the solid
and shell elements of this path.

    fea = pyTACS("model.bdf")
    fea.initialize()                       # or initialize(elemCallBack)
    probs = fea.createTACSProbsFromBDF()   # one StaticProblem per SUBCASE
    for p in probs.values():
        p.addFunction("compliance", functions.Compliance)
        p.solve()
        funcs = {}
        p.evalFunctions(funcs)             # keys "<problem>_<function>"

What the reference's loader does that is kept (tacs/pymeshloader.py):
node, element and component numbers follow the file order of GRID, element
and property cards; CHEXA and CQUAD4 connectivity is reordered from Nastran's
counter-clockwise order to the tensor-product order ([0, 1, 3, 2, 4, 5, 7, 6],
[0, 1, 3, 2]); every SPC / SPC1 set is applied to the assembler; the default
element callback makes SolidConstitutive(t = 1) / Element3D for PSOLID and
IsoShellConstitutive(t = T, kcorr = TS/T) / Quad4Shell for PSHELL from MAT1
(tacs/pytacs.py:885).  Loads from the bulk data (tacs/problems/base.py:716):
FORCE into the load vector, PLOAD4 on CHEXA as a TACSPressure3D on the face
holding G1 and G34 with the sign reversed (Nastran's positive pressure points
into the solid), PLOAD4 / PLOAD2 on CQUAD4 as TACSShellPressure.

The solve is GMRES with block-Jacobi on the GPU (the reference defaults to a
direct Schur factorisation as preconditioner); options L2Convergence,
L2ConvergenceRel, subSpaceSize and nRestarts keep their meaning.
"""
import numbers

import numpy as np

from . import TACS, constitutive, elements
from .bdf import BDF

_HEXA_TACS = [0, 1, 3, 2, 4, 5, 7, 6]
_QUAD_TACS = [0, 1, 3, 2]
# faces of an 8-node hex in tensor-product node order (TACSHexaBasis faces 0..5)
_HEX_FACES = [[0, 2, 4, 6], [1, 3, 5, 7], [0, 1, 4, 5], [2, 3, 6, 7], [0, 1, 2, 3], [4, 5, 6, 7]]


def mesh_arrays(b):
    """(element type, vars per node, conn [nelem, npe] in TACS node numbers and
    tensor-product order, X [nnodes, 3], component of each element) of a
    parsed BDF (pymeshloader.py:197-226)."""
    kinds = {t for t, _, _ in b.elements.values()}
    if not kinds:
        raise ValueError(f"{b.path}: no CHEXA or CQUAD4 elements")
    if len(kinds) > 1:
        raise ValueError(f"{b.path}: mixed solid and shell models are not supported")
    etype = kinds.pop()
    order = _HEXA_TACS if etype == "CHEXA" else _QUAD_TACS
    comp = {pid: i for i, pid in enumerate(b.properties)}
    conn = np.empty((len(b.elem_ids), len(order)), dtype=np.intc)
    comps = np.empty(len(b.elem_ids), dtype=np.intc)
    for i, eid in enumerate(b.elem_ids):
        _, pid, nids = b.elements[eid]
        conn[i] = [b.nid_to_tacs[nids[k]] for k in order]
        comps[i] = comp[pid]
    X = np.array([b.nodes[n] for n in b.node_ids])
    return etype, (3 if etype == "CHEXA" else 6), conn, X, comps


def bc_groups(b, vpn):
    """Every SPC / SPC1 set as [(TACS nodes, dofs, values or None)], nodes
    grouped by their constrained dofs (pymeshloader.py:909-960)."""
    bcs = {}
    for sets in b.spcs.values():
        for nids, comps, val in sets:
            dofs = sorted({int(ch) - 1 for ch in comps if ch.isdigit() and 0 < int(ch) <= vpn})
            for n in nids:
                if n not in b.nid_to_tacs or not dofs:
                    continue
                d = bcs.setdefault(b.nid_to_tacs[n], {})
                for k in dofs:
                    d[k] = val
    groups = {}
    for node, d in bcs.items():
        groups.setdefault(tuple(sorted(d.items())), []).append(node)
    out = []
    for key, nodes in groups.items():
        vals = [v for _, v in key]
        out.append((np.array(nodes, dtype=np.intc), [k for k, _ in key],
                    vals if any(v != 0.0 for v in vals) else None))
    return out


def load_arrays(b, loadID, vpn, scale=1.0):
    """The load set loadID as (nodal load vector F [vpn * nnodes], pressures
    [(TACS element, face, p)]) with the reference's conversions
    (tacs/problems/base.py:716-860): FORCE into F; PLOAD4 on CHEXA on the face
    holding G1 and G34 with the sign reversed; PLOAD4 / PLOAD2 on CQUAD4 on
    face 0 (the average of P1..P4)."""
    F = np.zeros(vpn * len(b.node_ids))
    pres = []
    for ld, s in b.reduced_loads(loadID, scale):
        if ld["type"] == "FORCE":
            n = b.nid_to_tacs[ld["node"]]
            F[vpn * n:vpn * n + 3] += s * ld["vector"]
        elif ld["type"] == "PLOAD2":
            for eid in ld["eids"]:
                pres.append((b.eid_to_tacs[eid], 0, s * ld["pressure"]))
        elif ld["type"] == "PLOAD4":
            p = s * float(np.mean(ld["pressures"]))
            for eid in ld["eids"]:
                etype, _, nids = b.elements[eid]
                if etype == "CQUAD4":
                    pres.append((b.eid_to_tacs[eid], 0, p))
                    continue
                if ld["g1"] is None:
                    raise ValueError(f"PLOAD4 {loadID} on CHEXA {eid}: G1 and G34 are needed")
                tn = [nids[k] for k in _HEXA_TACS]
                face = [f for f, loc in enumerate(_HEX_FACES)
                        if ld["g1"] in [tn[k] for k in loc] and ld["g34"] in [tn[k] for k in loc]]
                if len(face) != 1:
                    raise ValueError(f"PLOAD4 {loadID}: G1, G34 do not name a face of {eid}")
                # Nastran's positive pressure points into the solid
                pres.append((b.eid_to_tacs[eid], face[0], -p))
    return F, pres


class pyTACS:
    def __init__(self, fileName, comm=None, dvNum=0, scaleList=None, options=None):
        self.bdfName = fileName
        self.bdfInfo = BDF(fileName)
        self.comm = comm
        self.options = dict(options or {})
        self.dvNum = dvNum
        self.scaleList = list(scaleList or [])
        self.assembler = None
        b = self.bdfInfo
        self.elemType, self.varsPerNode, self.conn, self.X, self.compOfElem = mesh_arrays(b)
        self.compIDs = {pid: i for i, pid in enumerate(b.properties)}
        self.compDescripts = [b.comments.get(pid, f"PROP{pid}") for pid in b.properties]

    # -- model set-up --------------------------------------------------------------
    def selectCompIDs(self, include=None, exclude=None):
        """Component ids whose description contains one of the include
        strings (all when None) and none of the exclude strings."""
        def hit(name, pats):
            if isinstance(pats, str):
                pats = [pats]
            return any(p.upper() in name.upper() for p in pats)
        out = []
        for i, d in enumerate(self.compDescripts):
            if include is not None and not hit(d, include):
                continue
            if exclude is not None and hit(d, exclude):
                continue
            out.append(i)
        return out

    def _default_element(self, pid):
        b = self.bdfInfo
        prop = b.properties[pid]
        mat = b.materials[prop["mid"]]
        props = constitutive.MaterialProperties(rho=mat["rho"], E=mat["E"], nu=mat["nu"],
                                                ys=mat["St"], alpha=mat["alpha"])
        if prop["type"] == "PSOLID":
            con = constitutive.SolidConstitutive(props, t=1.0, tlb=0.0, tub=10.0, tNum=-1)
            return elements.Element3D(elements.LinearElasticity3D(con),
                                      elements.LinearHexaBasis())
        con = constitutive.IsoShellConstitutive(props, t=prop["t"], tlb=0.0, tub=1e20, tNum=-1,
                                                kcorr=prop["tst"])
        return elements.Quad4Shell(None, con)

    def initialize(self, elemCallBack=None):
        b = self.bdfInfo
        vpn = self.varsPerNode
        npe = self.conn.shape[1]
        comp_elems = []
        for pid, cid in self.compIDs.items():
            if elemCallBack is None:
                elem = self._default_element(pid)
            else:
                res = elemCallBack(self.dvNum, cid, self.compDescripts[cid], [self.elemType],
                                   {}, propID=pid)
                scales = None
                if isinstance(res, tuple):
                    res, scales = res
                    if isinstance(scales, numbers.Number):
                        scales = [scales]
                elem = res[0] if isinstance(res, (list, tuple)) else res
                if scales:
                    self.dvNum += len(scales)
                    self.scaleList.extend(scales)
            comp_elems.append(elem)
        self.elemObjects = comp_elems
        ne = self.conn.shape[0]
        a = self.assembler = TACS.Assembler.create(self.comm, vpn, len(b.node_ids), ne)
        a.setElementConnectivity(np.arange(0, npe * ne + 1, npe, dtype=np.intc),
                                 self.conn.reshape(-1))
        a.setElements([comp_elems[c] for c in self.compOfElem])
        for nodes, dofs, vals in bc_groups(b, vpn):
            a.addBCs(nodes, dofs, vals)
        a.initialize()
        self.Xpts0 = a.createNodeVec()
        self.Xpts0.getArray()[:] = self.X.reshape(-1)
        a.setNodes(self.Xpts0)
        return a

    # -- problems --------------------------------------------------------------------
    def createStaticProblem(self, name, options=None):
        if self.assembler is None:
            raise RuntimeError("pyTACS.initialize() must be called first")
        return StaticProblem(name, self, options)

    def createTACSProbsFromBDF(self):
        """One StaticProblem per SUBCASE with its LOAD set (named by SUBTITLE,
        else LABEL, else load_set_<id>), as pyTACS.createTACSProbsFromBDF
        (tacs/pytacs.py:1708)."""
        probs = {}
        for sc in self.bdfInfo.subcases:
            if sc["load"] is None:
                continue
            name = sc["name"] or sc["label"] or f"load_set_{sc['load']:03d}"
            p = self.createStaticProblem(name)
            p.addLoadFromBDF(sc["load"])
            probs[name] = p
        return probs


class StaticProblem:
    """Linear static problem (tacs/problems/static.py): R(u) - F = 0 solved
    for u from zero, functions at the solution."""

    DEFAULTS = {"L2Convergence": 1e-12, "L2ConvergenceRel": 1e-12, "subSpaceSize": 10,
                "nRestarts": 15}

    def __init__(self, name, fea, options=None):
        self.name = name
        self.fea = fea
        self.assembler = a = fea.assembler
        self.options = dict(self.DEFAULTS)
        for k, v in (options or {}).items():
            self.setOption(k, v)
        self.F = a.createVec()
        self.F.zeroEntries()
        self.aux = TACS.AuxElements()
        self.auxItems = []
        self.functionList = {}
        self.u = a.createVec()
        self.res = a.createVec()
        self.mat = None

    def setOption(self, name, value):
        if name not in self.DEFAULTS:
            raise ValueError(f"StaticProblem: option {name!r} is not supported")
        self.options[name] = value

    def addFunction(self, funcName, funcHandle, compIDs=None, **kwargs):
        if compIDs is not None:
            raise ValueError("addFunction: component subsets are not supported")
        self.functionList[funcName] = funcHandle(self.assembler, **kwargs)

    def getFunctionKeys(self):
        return list(self.functionList)

    # -- loads ---------------------------------------------------------------------------
    def addLoadToNodes(self, nodeIDs, F, nastranOrdering=False):
        b = self.fea.bdfInfo
        nodeIDs = np.atleast_1d(nodeIDs)
        F = np.atleast_2d(np.asarray(F, dtype=float))
        vpn = self.fea.varsPerNode
        arr = self.F.getArray()
        for k, n in enumerate(nodeIDs):
            node = b.nid_to_tacs[int(n)] if nastranOrdering else int(n)
            f = F[k if F.shape[0] > 1 else 0]
            arr[vpn * node:vpn * node + vpn] += f[:vpn]

    def _element(self, eid, nastranOrdering):
        b = self.fea.bdfInfo
        return b.eid_to_tacs[int(eid)] if nastranOrdering else int(eid)

    def addPressureToElements(self, elemIDs, pressures, faceIndex=0, nastranOrdering=False):
        elemIDs = np.atleast_1d(elemIDs)
        pressures = np.broadcast_to(np.asarray(pressures, dtype=float), elemIDs.shape)
        for e, p in zip(elemIDs, pressures):
            elem = self._element(e, nastranOrdering)
            obj = self.fea.elemObjects[self._comp_of(elem)].createElementPressure(int(faceIndex),
                                                                                 float(p))
            self.aux.addElement(elem, obj)
            self.auxItems.append(obj)

    def addPressureToComponents(self, compIDs, pressures, faceIndex=0):
        compIDs = np.atleast_1d(compIDs)
        pressures = np.broadcast_to(np.asarray(pressures, dtype=float), compIDs.shape)
        for c, p in zip(compIDs, pressures):
            elems = np.nonzero(self.fea.compOfElem == int(c))[0]
            self.addPressureToElements(elems, p, faceIndex)

    def _comp_of(self, elem):
        return int(self.fea.compOfElem[elem])

    def addLoadFromBDF(self, loadID, scale=1.0):
        F, pres = load_arrays(self.fea.bdfInfo, loadID, self.fea.varsPerNode, scale)
        self.F.getArray()[:] += F
        for elem, face, p in pres:
            self.addPressureToElements(elem, p, face)

    # -- solve and functions -------------------------------------------------------------
    def _ensure_matrix(self):
        if self.mat is None:
            self.mat = self.assembler.createMat()
            self.pc = TACS.Pc(self.mat)
            o = self.options
            self.ksm = TACS.KSM(self.mat, self.pc, int(o["subSpaceSize"]), int(o["nRestarts"]))

    def solve(self):
        """u = -K^{-1} (R(0) - F), R with this problem's pressure loads
        (StaticProblem.solve tacs/problems/static.py:1010)."""
        a = self.assembler
        self._ensure_matrix()
        a.setAuxElements(self.aux)
        a.zeroVariables()
        a.assembleJacobian(1.0, 0.0, 0.0, self.res, self.mat)
        self.pc.factor()
        self.res.axpy(-1.0, self.F)
        a.applyBCs(self.res)
        self.ksm.setTolerances(float(self.options["L2ConvergenceRel"]),
                               float(self.options["L2Convergence"]))
        du = a.createVec()
        flag = self.ksm.solve(self.res, du)
        self.u.copyValues(du)
        self.u.scale(-1.0)
        a.setVariables(self.u)
        return flag

    def getVariables(self, states=None):
        if states is not None:
            states.copyValues(self.u)
            return states
        return self.u.getArray().copy()

    def evalFunctions(self, funcs, evalFuncs=None):
        names = list(self.functionList) if evalFuncs is None else list(evalFuncs)
        a = self.assembler
        a.setAuxElements(self.aux)
        a.setVariables(self.u)
        vals = a.evalFunctions([self.functionList[n] for n in names])
        for n, v in zip(names, vals):
            funcs[f"{self.name}_{n}"] = v
        return funcs
