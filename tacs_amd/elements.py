"""Mirror of tacs.elements for the solid path (tacs/elements.pyx:
LinearHexaBasis :297, QuadraticHexaBasis :308, LinearElasticity3D :498,
Element3D :564, Traction3D :597)."""
import ctypes

import numpy as np

from ._lib import lib
from ._obj import TACSHandle

LINEAR = 0
NONLINEAR = 1


class LinearHexaBasis(TACSHandle):
    order = 2

    def __init__(self):
        super().__init__(lib().tacs_linear_hexa_basis_create())


class QuadraticHexaBasis(TACSHandle):
    order = 3

    def __init__(self):
        super().__init__(lib().tacs_quadratic_hexa_basis_create())


class LinearElasticity3D(TACSHandle):
    def __init__(self, con, elem_type=LINEAR):
        super().__init__(lib().tacs_linear_elasticity3d_create(con.ptr, elem_type), keep=(con,))

    def getVarsPerNode(self):
        return 3


class Element3D(TACSHandle):
    def __init__(self, model, basis):
        self.basis = basis
        super().__init__(lib().tacs_element3d_create(model.ptr, basis.ptr), keep=(model, basis))

    def getVarsPerNode(self):
        return 3

    def getNumNodes(self):
        return self.basis.order ** 3

    def createElementTraction(self, faceIndex, trac):
        return Traction3D(3, faceIndex, self.basis, trac)

    def createElementPressure(self, faceIndex, p):
        """TACSElement3D::createElementPressure (src/elements/TACSElement3D.cpp:54)."""
        return Pressure3D(3, faceIndex, self.basis, p)


class Traction3D(TACSHandle):
    def __init__(self, varsPerNode, faceIndex, basis, traction, coordComp=True):
        arr = (ctypes.c_double * 24)(*([float(v) for v in traction][:24]
                                       + [0.0] * max(0, 24 - len(traction))))
        super().__init__(lib().tacs_traction3d_create(varsPerNode, faceIndex, basis.ptr, arr,
                                                      1 if coordComp else 0), keep=(basis,))


class Pressure3D(TACSHandle):
    """TACSPressure3D (src/elements/TACSPressure3D.h): constant pressure p on
    one face of a solid element (traction -p n, n the outward unit normal)."""

    def __init__(self, varsPerNode, faceIndex, basis, p):
        super().__init__(lib().tacs_pressure3d_create(varsPerNode, faceIndex, basis.ptr,
                                                      float(p)), keep=(basis,))


class ShellNaturalTransform(TACSHandle):
    """TACSShellNaturalTransform (tacs/elements.pyx:646)."""

    def __init__(self):
        super().__init__(lib().tacs_shell_natural_transform_create())


class ShellRefAxisTransform(TACSHandle):
    """TACSShellRefAxisTransform(axis) (tacs/elements.pyx:659): the local 1-axis
    is the reference axis projected on the tangent plane."""

    def __init__(self, axis):
        a = np.ascontiguousarray(axis, dtype=np.float64).reshape(3)
        super().__init__(lib().tacs_shell_ref_axis_transform_create(a.ctypes.data))
        self.axis = a / np.linalg.norm(a) if np.linalg.norm(a) != 0.0 else a

    def getRefAxis(self):
        return self.axis.copy()


class Quad4Shell(TACSHandle):
    """TACSQuad4Shell (tacs/elements.pyx:694): 4-node MITC shell, 6 dof per
    node; transform None selects the natural transform."""

    def __init__(self, transform, con):
        if transform is None:
            transform = ShellNaturalTransform()
        self.transform = transform
        super().__init__(lib().tacs_quad4_shell_create(transform.ptr, con.ptr),
                         keep=(transform, con))

    def getVarsPerNode(self):
        return 6

    def getNumNodes(self):
        return 4

    def createElementPressure(self, faceIndex, p):
        """Uniform-pressure aux element (TACSShellElement::createElementPressure)."""
        return _Aux(lib().tacs_element_create_pressure(self.ptr, faceIndex, float(p)))


class _Aux(TACSHandle):
    def __init__(self, ptr):
        super().__init__(ptr)
