"""Materials (material.py of the reference).

``compute_stress`` is the reference's pure function of a strain increment;
the time step applies the same update inside the HIP stress kernel using the
coefficients from ``_kernel_coefficients``.
"""
import abc

import torch

from . import _lib


class Material(abc.ABC):
    """material.py:8-51."""

    _props = ()

    def __init__(self, material_properties):
        self.properties = material_properties

    @abc.abstractmethod
    def __repr__(self):
        ...

    @abc.abstractmethod
    def compute_stress(self, dstrain):
        ...

    def validate_props(self, material_properties):
        for key in self._props:
            if key not in material_properties:
                raise KeyError(
                    f"'{key}' should be present in `material_properties` "
                    f"for {self.__class__.__name__} materials.")


class LinearElastic(Material):
    """material.py:54-116: isotropic linear elasticity, dsigma = De @ dstrain."""

    _props = ("density", "youngs_modulus", "poisson_ratio")

    def __init__(self, material_properties):
        self.validate_props(material_properties)
        E = material_properties["youngs_modulus"]
        nu = material_properties["poisson_ratio"]
        density = material_properties["density"]
        bulk_modulus = E / (3 * (1 - 2 * nu))
        constrained_modulus = E * (1 - nu) / ((1 + nu) * (1 - 2 * nu))
        shear_modulus = E / (2 * (1 + nu))
        self.properties = {
            **material_properties,
            "bulk_modulus": bulk_modulus,
            "pwave_velocity": (constrained_modulus / density) ** 0.5,
            "swave_velocity": (shear_modulus / density) ** 0.5,
        }
        self._compute_elastic_tensor()

    def __repr__(self):
        return f"LinearElastic(props={self.properties})"

    @staticmethod
    def _coefficients(E, nu):
        """(a1, a2, G) of material.py:94-101 in double precision."""
        G = E / (2 * (1 + nu))
        K = E / (3 * (1 - 2 * nu))
        return K + (4 * G / 3), K - (2 * G / 3), G

    def _compute_elastic_tensor(self):
        a1, a2, G = self._coefficients(self.properties["youngs_modulus"],
                                       self.properties["poisson_ratio"])
        self.de = torch.tensor([[a1, a2, a2, 0, 0, 0],
                                [a2, a1, a2, 0, 0, 0],
                                [a2, a2, a1, 0, 0, 0],
                                [0, 0, 0, G, 0, 0],
                                [0, 0, 0, 0, G, 0],
                                [0, 0, 0, 0, 0, G]], dtype=torch.float64)

    def compute_stress(self, dstrain):
        dstrain = torch.as_tensor(dstrain)
        if not dstrain.is_floating_point():
            dstrain = dstrain.float()
        return self.de.to(dstrain.dtype).to(dstrain.device) @ dstrain

    def _kernel_coefficients(self, per_particle_E=False):
        if per_particle_E:
            return _lib.DMPM_LINEAR_ELASTIC, self._coefficients(
                1.0, self.properties["poisson_ratio"])
        return _lib.DMPM_LINEAR_ELASTIC, self._coefficients(
            self.properties["youngs_modulus"], self.properties["poisson_ratio"])


class SimpleMaterial(Material):
    """material.py:119-131: dsigma = E * dstrain."""

    _props = ("E", "density")

    def __init__(self, material_properties):
        self.validate_props(material_properties)
        self.properties = material_properties

    def __repr__(self):
        return f"SimpleMaterial(props={self.properties})"

    def compute_stress(self, dstrain):
        return torch.as_tensor(dstrain) * self.properties["E"]

    def _kernel_coefficients(self, per_particle_E=False):
        # with a per-particle modulus (Particles.youngs_modulus), dsigma =
        # E_p dstrain is the LinearElastic update with a1 = 1, a2 = 0, G = 1
        # scaled by E_p: (E_p d0 + 0 d1) + 0 d2 = E_p d0, E_p d1, E_p d3
        # (equal to E_p d_k up to the sign of a zero)
        if per_particle_E:
            return _lib.DMPM_LINEAR_ELASTIC, (1.0, 0.0, 1.0)
        return _lib.DMPM_SIMPLE_MATERIAL, (float(self.properties["E"]), 0.0, 0.0)
