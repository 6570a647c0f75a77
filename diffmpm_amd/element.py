"""Elements (element.py of the reference): mesh connectivity, shape functions
and the grid-side methods of the step, which run as HIP kernels.

``shapefn``, ``_shapefn_natural_grad`` and ``shapefn_grad`` are the
reference's array utilities, evaluated with torch on whatever device their
argument lives on; the step itself never calls them.
"""
from __future__ import annotations

import abc
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .node import Nodes

__all__ = ["_Element", "Linear1D", "Quadrilateral4Node"]


def _plan(elements, particles):
    from ._plan import plan_for
    return plan_for(elements, particles)


class _Element(abc.ABC):
    """element.py:21-340."""

    nodes: Nodes
    total_elements: int
    ndim: int
    _element_code: int

    @abc.abstractmethod
    def id_to_node_ids(self, id):
        ...

    def id_to_node_loc(self, id):
        """element.py:47-62."""
        node_ids = self.id_to_node_ids(id).reshape(-1).to(self.nodes.device).long()
        return self.nodes.loc[node_ids]

    def id_to_node_vel(self, id):
        """element.py:64-79."""
        node_ids = self.id_to_node_ids(id).reshape(-1).to(self.nodes.device).long()
        return self.nodes.velocity[node_ids]

    @abc.abstractmethod
    def shapefn(self, xi):
        ...

    @abc.abstractmethod
    def shapefn_grad(self, xi, coords):
        ...

    # -- step methods (HIP) ----------------------------------------------------
    def set_particle_element_ids(self, particles):
        """element.py:517-541 / 820-840."""
        _plan(self, particles).set_particle_element_ids()

    def compute_nodal_mass(self, particles):
        """element.py:122-152."""
        _plan(self, particles).compute_nodal_mass()

    def compute_nodal_momentum(self, particles):
        """element.py:154-189."""
        _plan(self, particles).compute_nodal_momentum()

    def compute_velocity(self, particles):
        """element.py:191-202."""
        _plan(self, particles).compute_velocity()

    def compute_external_force(self, particles):
        """element.py:204-234."""
        _plan(self, particles).compute_external_force()

    def compute_body_force(self, particles, gravity):
        """element.py:236-268."""
        _plan(self, particles).compute_body_force(gravity)

    def apply_concentrated_nodal_forces(self, particles, curr_time):
        """element.py:270-284."""
        _plan(self, particles).apply_concentrated_nodal_forces(curr_time)

    def apply_particle_traction_forces(self, particles):
        """element.py:286-303."""
        _plan(self, particles).apply_particle_traction_forces()

    def update_nodal_acceleration_velocity(self, particles, dt, *args):
        """element.py:305-329."""
        _plan(self, particles).update_nodal_acceleration_velocity(dt)

    def apply_boundary_constraints(self, *args):
        """element.py:331-334 (constraint.py:29-44 on the nodes)."""
        self._grid_plan().apply_boundary_constraints()

    def apply_force_boundary_constraints(self, *args):
        """element.py:336-340 (unused by the schemes)."""
        ids = torch.as_tensor(self.constraints[0][0], device=self.nodes.device).long()
        for f in (self.nodes.f_int, self.nodes.f_ext, self.nodes.f_damp):
            f[ids] = 0

    def compute_internal_force(self, particles):
        """element.py:548-600 / 842-901."""
        _plan(self, particles).compute_internal_force()

    def compute_volume(self, *args):
        """element.py:543-546 / 903-909."""
        code = _lib.dtype_code(self.nodes.dtype)
        v = _lib.element_volume(self._element_code, code, self._hx, self._hy)
        self.volume = torch.full((self.total_elements, 1, 1), v, dtype=self.nodes.dtype,
                                 device=self.nodes.device)

    # -- helpers ----------------------------------------------------------------
    def _grid_plan(self):
        """A plan with no particles, for the node-only methods called without a
        particle set (apply_boundary_constraints)."""
        p = getattr(self, "_gplan", None)
        if p is None:
            from .particle import Particles
            from .material import SimpleMaterial
            from ._plan import Plan
            dummy = Particles(torch.zeros((0, 1, self.ndim)), SimpleMaterial({"E": 1, "density": 1}),
                              torch.zeros(0, dtype=torch.int32), device=self.nodes.device,
                              dtype=self.nodes.dtype)
            p = Plan(self, dummy)
            self._gplan = p
        p.refresh_tables()
        return p

    def _setup_common(self, constraints, nodes, concentrated_nodal_forces, initialized, volume,
                      node_loc, device, dtype):
        if nodes is None:
            self.nodes = Nodes(node_loc.shape[0], torch.from_numpy(node_loc),
                               device=device, dtype=dtype)
        else:
            self.nodes = nodes
        self.constraints = constraints
        self.concentrated_nodal_forces = concentrated_nodal_forces
        if initialized is None:
            self.volume = torch.ones((self.total_elements, 1, 1), dtype=self.nodes.dtype,
                                     device=self.nodes.device)
        else:
            self.volume = torch.as_tensor(volume, device=self.nodes.device)
        self.initialized = True


def _np_dtype(dtype):
    return np.float64 if dtype == torch.float64 else np.float32


class Linear1D(_Element):
    """element.py:343-600: 2-node line elements; node i at fl(i*h)."""

    ndim = 1
    _element_code = _lib.DMPM_LINEAR1D

    def __init__(self, nelements: int, total_elements: int, el_len: float,
                 constraints: Sequence[Tuple], nodes: Optional[Nodes] = None,
                 concentrated_nodal_forces: Sequence = [], initialized: Optional[bool] = None,
                 volume=None, *, device=None, dtype=None):
        self.nelements = int(nelements)
        self.total_elements = int(nelements)
        self.el_len = el_len
        dtype = dtype or torch.float32
        nd = _np_dtype(dtype)
        self._nx, self._ny = self.nelements, 1
        self._hx = float(nd(el_len))
        self._hy = self._hx
        loc = (np.arange(self.nelements + 1).astype(nd) * nd(el_len)).reshape(-1, 1, 1)
        self._setup_common(constraints, nodes, concentrated_nodal_forces, initialized, volume,
                           loc, device or _lib.default_device(), dtype)

    def id_to_node_ids(self, id):
        """element.py:414-428."""
        id = torch.as_tensor(id)
        return torch.stack([id, id + 1]).reshape(2, 1)

    def shapefn(self, xi):
        """element.py:430-454: (N,1,1) -> (N,2,1,1)."""
        xi = torch.as_tensor(xi)
        if xi.ndim != 3:
            raise ValueError(f"`xi` should be of size (npoints, 1, ndim); found {tuple(xi.shape)}")
        return torch.stack([0.5 * (1 - xi), 0.5 * (1 + xi)]).permute(1, 0, 2, 3)

    def _shapefn_natural_grad(self, xi):
        """element.py:456-486."""
        xi = torch.as_tensor(xi)
        dt = xi.dtype if xi.is_floating_point() else torch.float32
        return torch.tensor([[-0.5], [0.5]], dtype=dt, device=xi.device)

    def shapefn_grad(self, xi, coords):
        """element.py:488-515."""
        xi = torch.as_tensor(xi)
        coords = torch.as_tensor(coords)
        if xi.ndim != 3:
            raise ValueError(f"`x` should be of size (npoints, 1, ndim); found {tuple(xi.shape)}")
        g = self._shapefn_natural_grad(xi)
        coords = coords.to(g.dtype).reshape(2, 1)
        jac = g.T @ coords
        return g @ torch.linalg.inv(jac).T


class Quadrilateral4Node(_Element):
    """element.py:603-909: bilinear quadrilaterals on a structured grid.
    Node n = j*(nx+1) + i at (fl(i*hx), fl(j*hy)); element e = J*nx + I."""

    ndim = 2
    _element_code = _lib.DMPM_QUAD4

    def __init__(self, nelements, total_elements: int, el_len, constraints: Sequence[Tuple],
                 nodes: Optional[Nodes] = None, concentrated_nodal_forces: Sequence = [],
                 initialized: Optional[bool] = None, volume=None, *, device=None, dtype=None):
        ne = [int(v) for v in np.asarray(nelements).reshape(-1)]
        if len(ne) != 2:
            raise ValueError("nelements must be (nx, ny)")
        self.nelements = tuple(ne)
        self.total_elements = int(np.asarray(total_elements))
        self.el_len = el_len
        dtype = dtype or torch.float32
        nd = _np_dtype(dtype)
        h = np.broadcast_to(np.asarray(el_len, dtype=np.float64), (2,)).astype(nd)
        self._nx, self._ny = ne
        self._hx, self._hy = float(h[0]), float(h[1])
        ii, jj = np.meshgrid(np.arange(ne[0] + 1), np.arange(ne[1] + 1))
        loc = np.stack([ii.reshape(-1).astype(nd) * h[0], jj.reshape(-1).astype(nd) * h[1]],
                       axis=1).reshape(-1, 1, 2)
        self._setup_common(constraints, nodes, concentrated_nodal_forces, initialized, volume,
                           loc, device or _lib.default_device(), dtype)

    def id_to_node_ids(self, id):
        """element.py:692-723."""
        id = torch.as_tensor(id)
        nx = self.nelements[0]
        ll = torch.div(id, nx, rounding_mode="floor") * (nx + 1) + torch.remainder(id, nx)
        return torch.stack([ll, ll + 1, ll + nx + 2, ll + nx + 1]).reshape(4, 1)

    def shapefn(self, xi):
        """element.py:725-757: (N,1,2) -> (N,4,1,1)."""
        xi = torch.as_tensor(xi)
        if xi.ndim != 3:
            raise ValueError(f"`xi` should be of size (npoints, 1, ndim); found {tuple(xi.shape)}")
        x, y = xi[:, :, 0], xi[:, :, 1]
        r = torch.stack([0.25 * (1 - x) * (1 - y), 0.25 * (1 + x) * (1 - y),
                         0.25 * (1 + x) * (1 + y), 0.25 * (1 - x) * (1 + y)])
        return r.permute(1, 0, 2).unsqueeze(-1)

    def _shapefn_natural_grad(self, xi):
        """element.py:759-789."""
        xi = torch.as_tensor(xi)
        if not xi.is_floating_point():
            xi = xi.float()
        xi = xi.reshape(-1)
        x, y = xi[0], xi[1]
        return torch.stack([
            torch.stack([-0.25 * (1 - y), -0.25 * (1 - x)]),
            torch.stack([0.25 * (1 - y), -0.25 * (1 + x)]),
            torch.stack([0.25 * (1 + y), 0.25 * (1 + x)]),
            torch.stack([-0.25 * (1 + y), 0.25 * (1 - x)]),
        ])

    def shapefn_grad(self, xi, coords):
        """element.py:791-818."""
        xi = torch.as_tensor(xi)
        coords = torch.as_tensor(coords)
        if xi.ndim != 3:
            raise ValueError(f"`x` should be of size (npoints, 1, ndim); found {tuple(xi.shape)}")
        g = self._shapefn_natural_grad(xi)
        jac = g.T @ coords.to(g.dtype).reshape(4, 2)
        return g @ torch.linalg.inv(jac).T
