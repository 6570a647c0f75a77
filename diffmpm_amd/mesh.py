"""Mesh containers and the method-name fan-out (mesh.py of the reference)."""
import abc
from typing import Sequence, Tuple

__all__ = ["_MeshBase", "Mesh1D", "Mesh2D"]


class _MeshBase(abc.ABC):
    """mesh.py:13-92."""

    ndim: int

    def __init__(self, config: dict):
        self.particles = config["particles"]
        self.elements = config["elements"]
        self.particle_tractions = config.get("particle_surface_traction", [])
        for i, ps in enumerate(self.particles):
            ps._traction_specs = [pt for pt in self.particle_tractions
                                  if i in [int(v) for v in pt.pset]]

    def apply_on_elements(self, function: str, args: Tuple = ()):
        """mesh.py:31-44: calls elements.<function>(pset, *args) per set."""
        f = getattr(self.elements, function)
        for particle_set in self.particles:
            f(particle_set, *args)

    def apply_on_particles(self, function: str, args: Tuple = ()):
        """mesh.py:46-58: calls pset.<function>(elements, *args) per set."""
        for particle_set in self.particles:
            getattr(particle_set, function)(self.elements, *args)

    def apply_traction_on_particles(self, curr_time: float):
        """mesh.py:60-77: zero, assign (per the traction table of each set) and
        map the particle tractions to the nodes."""
        from ._plan import plan_for
        for ps in self.particles:
            plan_for(self.elements, ps).assign_tractions(curr_time)
        self.apply_on_elements("apply_particle_traction_forces")


class Mesh1D(_MeshBase):
    """mesh.py:95-110."""

    def __init__(self, config: dict):
        self.ndim = 1
        super().__init__(config)


class Mesh2D(_MeshBase):
    """mesh.py:112-126."""

    def __init__(self, config: dict):
        self.ndim = 2
        super().__init__(config)
