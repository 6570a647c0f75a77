"""USF / USL phase ordering (scheme.py of the reference).

A mesh with one particle set runs each phase as one fused C-ABI call
(dmpm_phase_*); with several sets the phases fan out method by method over
the sets exactly as the reference does (mesh.py:31-58), which keeps its
per-set zeroing behaviour."""
from __future__ import annotations

import abc

_schemes = ("usf", "usl")


class _MPMScheme(abc.ABC):
    """scheme.py:16-75."""

    def __init__(self, mesh, dt, velocity_update):
        self.mesh = mesh
        self.velocity_update = velocity_update
        self.dt = dt

    def _single(self):
        if len(self.mesh.particles) != 1:
            return None
        from ._plan import plan_for
        return plan_for(self.mesh.elements, self.mesh.particles[0])

    def compute_nodal_kinematics(self):
        """scheme.py:22-29."""
        p = self._single()
        if p is not None:
            return p.phase_nodal_kinematics()
        m = self.mesh
        m.apply_on_elements("set_particle_element_ids")
        m.apply_on_particles("update_natural_coords")
        m.apply_on_elements("compute_nodal_mass")
        m.apply_on_elements("compute_nodal_momentum")
        m.apply_on_elements("compute_velocity")
        m.apply_on_elements("apply_boundary_constraints")

    def compute_stress_strain(self):
        """scheme.py:31-35."""
        p = self._single()
        if p is not None:
            return p.phase_stress_strain(self.dt)
        self.mesh.apply_on_particles("compute_strain", args=(self.dt,))
        self.mesh.apply_on_particles("update_volume")
        self.mesh.apply_on_particles("compute_stress")

    def compute_forces(self, gravity, step: int):
        """scheme.py:37-56 at t = step * dt."""
        p = self._single()
        if p is not None:
            return p.phase_forces(gravity, step, self.dt)
        import numpy as np
        import torch
        # t = step * dt in the working precision (the fused phases do the same)
        wt = np.float64 if self.mesh.particles[0].dtype == torch.float64 else np.float32
        t = float(wt(step) * wt(self.dt))
        m = self.mesh
        m.apply_on_elements("compute_external_force")
        m.apply_on_elements("compute_body_force", args=(gravity,))
        m.apply_traction_on_particles(t)
        m.apply_on_elements("apply_concentrated_nodal_forces", args=(t,))
        m.apply_on_elements("compute_internal_force")

    def compute_particle_kinematics(self):
        """scheme.py:58-67."""
        p = self._single()
        if p is not None:
            return p.phase_particle_kinematics(self.dt, self.velocity_update)
        self.mesh.apply_on_elements("update_nodal_acceleration_velocity", args=(self.dt,))
        self.mesh.apply_on_particles("update_position_velocity",
                                     args=(self.dt, self.velocity_update))

    @abc.abstractmethod
    def precompute_stress_strain(self):
        ...

    @abc.abstractmethod
    def postcompute_stress_strain(self):
        ...


class USF(_MPMScheme):
    """scheme.py:78-104: stress update before the force computation."""

    def precompute_stress_strain(self):
        self.compute_stress_strain()

    def postcompute_stress_strain(self):
        pass


class USL(_MPMScheme):
    """scheme.py:107-133: stress update after the particle update."""

    def precompute_stress_strain(self):
        pass

    def postcompute_stress_strain(self):
        self.compute_stress_strain()
