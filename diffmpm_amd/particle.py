"""Particle state (particle.py of the reference) stored as one SoA device
block; the step methods call the HIP kernels through the C ABI."""
from typing import Optional, Tuple

import torch

from . import _lib
from ._soa import Field


class Particles:
    """particle.py:12-347.

    Field shapes follow the reference: vectors ``(N, 1, ndim)``, tensors
    ``(N, 6, 1)``, scalars ``(N, 1, 1)``, ``dvolumetric_strain`` and
    ``volumetric_strain_centroid`` ``(N, 1)``, ``element_ids`` ``(N,)`` int32.
    All are views of one ``(DMPM_P_ROWS, ld)`` block, ``self._buf``.

    Extension: ``youngs_modulus`` (``(N,)``, optional) holds a per-particle
    Young's modulus used instead of the material's scalar one."""

    loc = Field(_lib.P_X, "vec")
    velocity = Field(_lib.P_V, "vec")
    mass = Field(_lib.P_MASS, "s3")
    volume = Field(_lib.P_VOLUME, "s3")
    density = Field(_lib.P_DENSITY, "s3")
    size = Field(_lib.P_SIZE, "vec")
    momentum = Field(_lib.P_MOMENTUM, "vec")
    acceleration = Field(_lib.P_ACCEL, "vec")
    strain = Field(_lib.P_STRAIN, "tensor")
    stress = Field(_lib.P_STRESS, "tensor")
    strain_rate = Field(_lib.P_STRAIN_RATE, "tensor")
    dstrain = Field(_lib.P_DSTRAIN, "tensor")
    f_ext = Field(_lib.P_FEXT, "vec")
    traction = Field(_lib.P_TRACTION, "vec")
    reference_loc = Field(_lib.P_XI, "vec")
    dvolumetric_strain = Field(_lib.P_DVOLSTRAIN, "s2")
    volumetric_strain_centroid = Field(_lib.P_VOLSTRAIN_C, "s2")

    def __init__(self, loc, material, element_ids, initialized: Optional[bool] = None,
                 data: Optional[Tuple] = None, *, device=None, dtype=None):
        dev = torch.device(device) if device is not None else (
            loc.device if isinstance(loc, torch.Tensor) else _lib.default_device())
        dtype = dtype or torch.float32
        self.material = material
        loc = torch.as_tensor(loc, device=dev).to(dtype)  # particle.py:46
        if loc.ndim != 3:
            raise ValueError(
                f"`loc` should be of size (nparticles, 1, ndim); found {tuple(loc.shape)}")
        self._count = loc.shape[0]
        self._ndim = loc.shape[-1]
        if self._ndim not in (1, 2):
            raise ValueError("only 1D and 2D particles are supported")
        self._ld = _lib.round_up(self._count)
        self._buf = torch.zeros((_lib.P_ROWS, self._ld), dtype=dtype, device=dev)
        self._eid = torch.zeros(self._count, dtype=torch.int32, device=dev)
        self._has_fext = False
        self._has_youngs = False
        self._plans = {}
        self.loc = loc
        self.element_ids = element_ids
        if initialized is None or data is None:
            self._buf[_lib.P_MASS, : self._count] = 1
            self._buf[_lib.P_DENSITY, : self._count] = material.properties["density"]
            self._buf[_lib.P_VOLUME, : self._count] = 1
        else:
            (self.mass, self.density, self.volume, self.size, self.velocity,
             self.acceleration, self.momentum, self.strain, self.stress,
             self.strain_rate, self.dstrain, self.f_ext, self.traction,
             self.reference_loc, self.dvolumetric_strain,
             self.volumetric_strain_centroid) = data
        self.initialized = True

    # -- bookkeeping -----------------------------------------------------------
    def _on_field_set(self, name):
        if name == "f_ext":
            self._has_fext = True

    @property
    def element_ids(self):
        return self._eid

    @element_ids.setter
    def element_ids(self, value):
        v = torch.as_tensor(value, device=self._eid.device).to(torch.int32).reshape(-1)
        if v.numel() == 1 and self._count != 1:
            v = v.expand(self._count)
        self._eid.copy_(v)

    @property
    def youngs_modulus(self):
        return self._buf[_lib.P_YOUNGS, : self._count] if self._has_youngs else None

    @youngs_modulus.setter
    def youngs_modulus(self, value):
        if value is None:
            self._has_youngs = False
        else:
            self._buf[_lib.P_YOUNGS, : self._count] = torch.as_tensor(
                value, device=self._buf.device).to(self._buf.dtype).reshape(-1)
            self._has_youngs = True
        self._plans.clear()

    @property
    def dtype(self):
        return self._buf.dtype

    @property
    def device(self):
        return self._buf.device

    def __len__(self):
        return self._count

    def __repr__(self):
        return f"Particles(nparticles={len(self)})"

    def _plan(self, elements):
        from ._plan import plan_for
        return plan_for(elements, self)

    # -- reference methods -----------------------------------------------------
    def set_mass_volume(self, m):
        """particle.py:138-156."""
        m = torch.as_tensor(m, device=self.device, dtype=self.dtype)
        if m.dim() == 0:
            self._buf[_lib.P_MASS, : self._count] = m
        elif tuple(m.shape) == (self._count, 1, 1):
            self.mass = m
        else:
            raise ValueError(f"Incompatible shapes. Expected {(self._count, 1, 1)}, "
                             f"found {tuple(m.shape)}.")
        self.volume = self.mass / self.material.properties["density"]

    def compute_volume(self, elements, total_elements):
        """particle.py:158-178 (kernel: dmpm_compute_particle_volume)."""
        self._plan(elements).compute_particle_volume(elements)

    def update_natural_coords(self, elements):
        """particle.py:180-206."""
        self._plan(elements).update_natural_coords()

    def update_position_velocity(self, elements, dt: float, velocity_update: bool):
        """particle.py:208-247."""
        self._plan(elements).update_position_velocity(elements, dt, velocity_update)

    def compute_strain(self, elements, dt: float):
        """particle.py:249-279."""
        self._plan(elements).compute_strain(elements, dt)

    def compute_stress(self, *args):
        """particle.py:315-322."""
        self._any_plan().compute_stress()

    def update_volume(self, *args):
        """particle.py:324-327."""
        self._any_plan().update_volume()

    def assign_traction(self, pids, dir: int, traction_: float):
        """particle.py:329-343: traction[pids, dir] += t * V / size[dir]."""
        _lib.require_gpu(self._buf)
        pids = torch.as_tensor(pids, device=self.device).long().reshape(-1)
        row = self._buf[_lib.P_TRACTION + dir]
        vals = (torch.as_tensor(traction_, dtype=self.dtype, device=self.device)
                * self._buf[_lib.P_VOLUME][pids]) / self._buf[_lib.P_SIZE + dir][pids]
        for k in range(pids.numel()):  # sequential, like .at[pids].add
            row[pids[k]] += vals[k]

    def zero_traction(self, *args):
        """particle.py:345-347."""
        self._buf[_lib.P_TRACTION:_lib.P_TRACTION + 2].zero_()

    def _any_plan(self):
        """Plan for the per-particle methods that take no elements: any plan
        of this set, else one on a single-cell mesh."""
        if not self._plans:
            from .element import Linear1D, Quadrilateral4Node
            if self._ndim == 1:
                el = Linear1D(1, 1, 1.0, [], device=self.device, dtype=self.dtype)
            else:
                el = Quadrilateral4Node((1, 1), 1, 1.0, [], device=self.device, dtype=self.dtype)
            self._standalone_elements = el
            return self._plan(el)
        return next(iter(self._plans.values()))
