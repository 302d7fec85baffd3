"""Mirror of tacs.constitutive for the solid path (tacs/constitutive.pyx:
MaterialProperties :89, SolidConstitutive :875)."""
from ._lib import lib
from ._obj import TACSHandle


class MaterialProperties(TACSHandle):
    """Isotropic material (rho, specific_heat, E, nu, ys, alpha, kappa)."""

    def __init__(self, rho=2700.0, specific_heat=921.0, E=70e9, nu=0.3, ys=270e6,
                 alpha=24e-6, kappa=230.0):
        self.props = dict(rho=rho, specific_heat=specific_heat, E=E, nu=nu, ys=ys)
        super().__init__(lib().tacs_material_create(rho, specific_heat, E, nu, ys, alpha,
                                                    kappa))


class SolidConstitutive(TACSHandle):
    def __init__(self, props, t=1.0, tNum=-1, tlb=0.0, tub=10.0):
        self.t, self.tNum = t, tNum
        super().__init__(lib().tacs_solid_constitutive_create(props.ptr, t, tNum, tlb, tub),
                         keep=(props,))


class IsoShellConstitutive(TACSHandle):
    """TACSIsoShellConstitutive (tacs/constitutive.pyx:1109): isotropic shell
    with thickness t (design variable tNum), shear correction kcorr."""

    def __init__(self, props, t=1.0, tNum=-1, tlb=0.0, tub=10.0, tOffset=0.0, kcorr=5.0 / 6.0):
        self.t, self.tNum = t, tNum
        super().__init__(lib().tacs_iso_shell_constitutive_create(props.ptr, t, tNum, tlb, tub,
                                                                  tOffset, kcorr),
                         keep=(props,))
