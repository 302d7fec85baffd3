"""Mirror of tacs.functions (tacs/functions.pyx: Compliance :172,
KSFailure :358)."""
from ._lib import lib
from ._obj import TACSHandle


class Compliance(TACSHandle):
    def __init__(self, assembler):
        super().__init__(lib().tacs_compliance_create(assembler.ptr), keep=(assembler,))

    def getFunctionValue(self):
        return lib().tacs_function_get_value(self.ptr)


class KSFailure(TACSHandle):
    def __init__(self, assembler, ksWeight=80.0, alpha=1.0, safetyFactor=1.0):
        super().__init__(lib().tacs_ksfailure_create(assembler.ptr, ksWeight, alpha,
                                                     safetyFactor), keep=(assembler,))

    def getFunctionValue(self):
        return lib().tacs_function_get_value(self.ptr)

    def getMaximumFailure(self):
        return lib().tacs_ksfailure_get_max_failure(self.ptr)
