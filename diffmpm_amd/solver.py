"""Explicit MPM solver (solver.py of the reference)."""
from __future__ import annotations

from collections import defaultdict
from typing import Callable, Optional

import torch

from .scheme import USF, USL, _schemes


class MPMExplicit:
    """solver.py:18-221.

    ``solve_jit`` runs the steps as enqueued HIP work with no host
    synchronisation between steps (for one particle set, one dmpm_run call per
    output interval).  ``keep_temporaries`` (extension, default True) keeps
    the reference's per-step temporaries (strain_rate, dstrain,
    dvolumetric_strain, reference_loc, particle momentum) up to date; False
    skips writing them, which the reference's outputs do not need."""

    __particle_props = ("loc", "velocity", "stress", "strain")

    def __init__(self, mesh, dt: float, scheme: str = "usf", velocity_update: bool = False,
                 sim_steps: int = 1, out_steps: int = 1, out_dir: str = "results/",
                 writer_func: Optional[Callable] = None, keep_temporaries: bool = True):
        if scheme == "usf":
            self.mpm_scheme = USF(mesh, dt, velocity_update)
        elif scheme == "usl":
            self.mpm_scheme = USL(mesh, dt, velocity_update)
        else:
            raise ValueError(f"Please select scheme from {_schemes}. Found {scheme}")
        self.mesh = mesh
        self.dt = dt
        self.scheme = scheme
        self.velocity_update = velocity_update
        self.sim_steps = sim_steps
        self.out_steps = out_steps
        self.out_dir = out_dir
        self.writer_func = writer_func
        self.keep_temporaries = keep_temporaries
        # solver.py:81-85
        self.mpm_scheme.mesh.apply_on_elements("set_particle_element_ids")
        self.mpm_scheme.mesh.apply_on_elements("compute_volume")
        self.mpm_scheme.mesh.apply_on_particles(
            "compute_volume", args=(self.mesh.elements.total_elements,))

    def _step(self, i, gravity):
        s = self.mpm_scheme
        s.compute_nodal_kinematics()
        s.precompute_stress_strain()
        s.compute_forces(gravity, i)
        s.compute_particle_kinematics()
        s.postcompute_stress_strain()

    def solve(self, gravity):
        """solver.py:116-156: returns per-step stacks of position, velocity,
        stress[:, :2] and strain[:, :2]."""
        result = defaultdict(list)
        for step in range(self.sim_steps):
            self._step(step, gravity)
            for pset in self.mesh.particles:
                result["position"].append(pset.loc.clone())
                result["velocity"].append(pset.velocity.clone())
                result["stress"].append(pset.stress[:, :2, 0].clone())
                result["strain"].append(pset.strain[:, :2, 0].clone())
        return {k: torch.stack(v) for k, v in result.items()}

    def _arrays(self, squeeze_each):
        out = {}
        for name in self.__particle_props:
            vals = [getattr(p, name) for p in self.mesh.particles]
            if squeeze_each:
                vals = [v.squeeze() for v in vals]
            out[name] = torch.stack(vals).squeeze() if not squeeze_each else torch.stack(vals)
        return out

    def solve_jit(self, gravity, start_step: int = 0) -> dict:
        """solver.py:159-221: runs sim_steps steps; the writer (if any) gets the
        state after every step i with i % out_steps == 0.  Snapshots are
        copied to pinned host memory asynchronously in stream order
        (``output.SnapshotQueue``) and written by a writer thread; the call
        waits for the writes only before it returns."""
        from .output import SnapshotQueue
        single = len(self.mesh.particles) == 1
        i, end = start_step, start_step + self.sim_steps
        out = max(int(self.out_steps), 1)
        snaps = None
        if self.writer_func is not None:
            snaps = SnapshotQueue(self.mesh.particles, self.writer_func, self.out_dir,
                                  self.sim_steps)
        try:
            while i < end:
                if snaps is None:
                    nxt = end
                else:  # run through the next step index divisible by out_steps
                    k = i if i % out == 0 else (i // out + 1) * out
                    nxt = min(end, k + 1)
                if single:
                    from ._plan import plan_for
                    plan_for(self.mesh.elements, self.mesh.particles[0]).run(
                        self.scheme, self.velocity_update, gravity, self.dt, i, nxt - i,
                        self.keep_temporaries)
                else:
                    for k in range(i, nxt):
                        self._step(k, gravity)
                if snaps is not None and (nxt - 1) % out == 0:
                    snaps.snapshot(nxt - 1)
                i = nxt
        finally:
            if snaps is not None:
                snaps.close()
        return self._arrays(False)
