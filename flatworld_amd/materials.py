"""Material descriptors (flatworld/materials/materialsclass.py:5-47).

These are plain host records; the constitutive laws themselves run in the HIP
kernels (flatworld_amd/csrc/fw_kernels.hip).
"""

from .definitions import MaterialType


class Elastic:
    """St.Venant-Kirchhoff material (flatworld/materials/materialsclass.py:5-13)."""

    def __init__(self, E, nu, rho):
        self.type = MaterialType.ELASTIC
        self.E = E
        self.nu = nu
        self.rho = rho

    def getRepresentativeModulus(self):
        return self.E


class HyperElastic:
    """Compressible Neo-Hookean material (flatworld/materials/materialsclass.py:16-24)."""

    def __init__(self, E, nu, rho):
        self.type = MaterialType.NEOHOOKEAN
        self.E = E
        self.nu = nu
        self.rho = rho

    def getRepresentativeModulus(self):
        return self.E


class MisesPlastic:
    """J2 plasticity with linear isotropic hardening (flatworld/materials/materialsclass.py:27-47)."""

    def __init__(self, E, nu, rho, sigma_y, H=0.0):
        self.type = MaterialType.MISES
        self.E = E
        self.nu = nu
        self.rho = rho
        self.sigma_y = sigma_y
        self.H = H

    def getRepresentativeModulus(self):
        return self.E


class SolidProp:
    """Material holder (flatworld/solidprop.py:1-6)."""

    def __init__(self, mat):
        self.mat = mat

    def getMaterial(self):
        return self.mat
