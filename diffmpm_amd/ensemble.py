"""Batched ensembles: many independent simulations stepped together.

The reference runs one simulation per ``MPMExplicit`` (diffmpm/solver.py);
an ensemble of them costs one launch sequence per simulation and step.
``Ensemble`` stacks the particle and node blocks of several 2D
Quadrilateral4Node simulations (different meshes, sizes, materials,
constraints) end to end and drives them with one batch plan: each step is
one launch for all simulations together (every member's tiles of cells in
one grid, after the re-sort), or one launch per phase (``autotune`` picks).

Each simulation keeps its own particle and node objects; ``gather()`` copies
their state into the ensemble's blocks (done at construction) and
``scatter()`` copies the ensemble's state back.  All simulations share dtype,
scheme, velocity_update, time step and gravity, and have one particle set
with a LinearElastic (or SimpleMaterial) material, and no particle tractions.
"""
from ctypes import byref, c_void_p

import torch

from . import _lib
from ._lib import call, ptr
from ._plan import _as_list, plan_for


class BatchPlan:
    """The C-ABI batch plan of an ``Ensemble`` (interface of ``Plan`` used by
    the run and autodiff helpers)."""

    def __init__(self, ens):
        self.ens = ens
        descs = (_lib.PlanDesc * len(ens.sims))(*[p.desc for p in ens.plans])
        h = c_void_p()
        with torch.cuda.device(ens.device):
            call("dmpm_batch_plan_create", descs, len(ens.sims), int(ens.ld), int(ens.nld),
                 byref(h))
        self.h = h
        self._tables()

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            try:
                _lib.lib().dmpm_plan_destroy(h)
            except Exception:
                pass
            self.h = None

    def _tables(self):
        ens = self.ens
        nodes, dirs, vals = [], [], []
        f_nodes, f_dirs, f_fn, f_vals = [], [], [], []
        fns, fn_index = [], {}
        for sim, off in zip(ens.sims, ens.node_off):
            el = sim.mesh.elements
            for ids, c in el.constraints:
                for i in _as_list(ids):
                    nodes.append(i + off), dirs.append(int(c.dir)), vals.append(float(c.velocity))
            for cnf in el.concentrated_nodal_forces:
                f = cnf.function
                if f is None:
                    k = -1
                else:
                    if id(f) not in fn_index:
                        fn_index[id(f)] = len(fns)
                        fns.append(f)
                    k = fn_index[id(f)]
                for i in _as_list(cnf.node_ids):
                    f_nodes.append(i + off), f_dirs.append(int(cnf.dir)), f_fn.append(k)
                    f_vals.append(float(cnf.force))
        call("dmpm_plan_set_constraints", self.h, len(nodes), _lib.i32_array(nodes),
             _lib.i32_array(dirs), _lib.f64_array(vals))
        kinds, offs, cnts, xs, fs = [], [], [], [], []
        from .functions import Linear
        for f in fns:
            if isinstance(f, Linear):
                kinds.append(_lib.DMPM_FN_LINEAR)
                offs.append(len(xs)), cnts.append(len(f.xvalues))
                xs.extend(float(v) for v in f.xvalues)
                fs.extend(float(v) for v in f.fxvalues)
            else:
                kinds.append(_lib.DMPM_FN_UNIT)
                offs.append(len(xs)), cnts.append(0)
        call("dmpm_plan_set_functions", self.h, len(fns), _lib.i32_array(kinds),
             _lib.i32_array(offs), _lib.i32_array(cnts), _lib.f64_array(xs), _lib.f64_array(fs))
        call("dmpm_plan_set_nodal_forces", self.h, len(f_nodes), _lib.i32_array(f_nodes),
             _lib.i32_array(f_dirs), _lib.i32_array(f_fn), _lib.f64_array(f_vals))

    def refresh_tables(self):
        pass

    @property
    def s(self):
        return _lib.stream_of(self.ens.pbuf)

    def _gravity(self, gravity):
        return self.ens.plans[0]._gravity(gravity)

    def set_block(self, cells_per_block):
        call("dmpm_plan_set_block", self.h, int(cells_per_block))

    def set_option(self, option, value):
        call("dmpm_plan_set_option", self.h, int(option), int(value))

    def run(self, scheme, velocity_update, gravity, dt, step0, nsteps, keep_temporaries=False,
            per_phase=False, pbuf=None, eid=None, nbuf=None):
        if per_phase:
            raise ValueError("batch plans run the fused 2D step only")
        ens = self.ens
        code = _lib.DMPM_USF if scheme == "usf" else _lib.DMPM_USL
        call("dmpm_run", self.h, ptr(ens.pbuf if pbuf is None else pbuf),
             ptr(ens.eid if eid is None else eid), ptr(ens.nbuf if nbuf is None else nbuf), code,
             1 if velocity_update else 0, self._gravity(gravity), float(dt), int(step0),
             int(nsteps), 1 if keep_temporaries else 0, self.s)


class Ensemble:
    """Independent simulations ``sims`` (``MPMExplicit``) stepped together."""

    def __init__(self, sims, per_particle_E=False):
        """per_particle_E: store each member's Young's modulus per particle
        (needed for log E gradients through ``autodiff``)."""
        if not sims:
            raise ValueError("empty ensemble")
        s0 = sims[0]
        for s in sims:
            if len(s.mesh.particles) != 1 or s.mesh.elements.ndim != 2:
                raise ValueError("ensemble members have one particle set on a 2D mesh")
            if getattr(s.mesh.particles[0], "_traction_specs", []):
                raise ValueError("ensemble members have no particle tractions")
            if (s.scheme, s.velocity_update, float(s.dt)) != (s0.scheme, s0.velocity_update,
                                                              float(s0.dt)):
                raise ValueError("ensemble members share scheme, velocity_update and dt")
            if s.mesh.particles[0].dtype != s0.mesh.particles[0].dtype:
                raise ValueError("ensemble members share a dtype")
        self.sims = sims
        if per_particle_E:
            for s in sims:
                ps = s.mesh.particles[0]
                if not ps._has_youngs:
                    E = float(ps.material.properties["youngs_modulus"])
                    ps.youngs_modulus = torch.full((len(ps),), E, dtype=ps.dtype,
                                                   device=ps.device)
                    ps._plans.clear()
        self.plans = [plan_for(s.mesh.elements, s.mesh.particles[0]) for s in sims]
        self.counts = [len(s.mesh.particles[0]) for s in sims]
        self.node_counts = [s.mesh.elements.nodes._count for s in sims]
        self.part_off, self.node_off = [], []
        po = no = 0
        for n, m in zip(self.counts, self.node_counts):
            self.part_off.append(po), self.node_off.append(no)
            po += n
            no += m
        self.n, self.nnodes = po, no
        ps0 = s0.mesh.particles[0]
        self.dtype, self.device = ps0.dtype, ps0.device
        self.ld, self.nld = _lib.round_up(max(po, 1)), _lib.round_up(max(no, 1))
        self.pbuf = torch.zeros((_lib.P_ROWS, self.ld), dtype=self.dtype, device=self.device)
        self.nbuf = torch.zeros((_lib.N_ROWS, self.nld), dtype=self.dtype, device=self.device)
        self.eid = torch.zeros(self.ld, dtype=torch.int32, device=self.device)
        self.scheme, self.velocity_update, self.dt = s0.scheme, s0.velocity_update, s0.dt
        self.gather()
        self.plan = BatchPlan(self)

    def gather(self):
        """Copies every member's particle and node state into the ensemble."""
        for s, po, n, no, m in zip(self.sims, self.part_off, self.counts, self.node_off,
                                   self.node_counts):
            ps, nodes = s.mesh.particles[0], s.mesh.elements.nodes
            self.pbuf[:, po:po + n] = ps._buf[:, :n]
            self.eid[po:po + n] = ps._eid[:n]
            self.nbuf[:, no:no + m] = nodes._buf[:, :m]

    def scatter(self):
        """Copies the ensemble's state back into every member."""
        for s, po, n, no, m in zip(self.sims, self.part_off, self.counts, self.node_off,
                                   self.node_counts):
            ps, nodes = s.mesh.particles[0], s.mesh.elements.nodes
            ps._buf[:, :n] = self.pbuf[:, po:po + n]
            ps._eid[:n] = self.eid[po:po + n]
            nodes._buf[:, :m] = self.nbuf[:, no:no + m]

    def run(self, nsteps, gravity, start_step=0, keep_temporaries=False):
        """Steps start_step .. start_step + nsteps - 1 of every member."""
        self.plan.run(self.scheme, self.velocity_update, gravity, self.dt, start_step, nsteps,
                      keep_temporaries)

    def autotune(self, gravity, steps=4, start_step=0, blocks=(64, 128, 256)):
        """Chooses how the batch plan steps: the one-launch step (every
        member's tiles of 16 x 16 cells in one launch per step, key "step")
        or one launch per phase with ``blocks`` cells per workgroup of the
        cell passes (keys 64, 128, 256), by timing ``steps`` steps (even, so
        that the one-launch step runs every one of them) after a warm-up of
        two, on scratch copies of the ensemble's state with a scratch plan.
        The ensemble's state and sort are untouched, and no choice changes
        results.  Applies the fastest (and, for odd-length runs, whose first
        step runs per phase, the fastest cells per workgroup) and returns
        {candidate: ms per step}."""
        if steps < 2 or steps % 2:
            raise ValueError("steps must be even and at least 2")
        scratch = BatchPlan(self)
        times = {}
        stream = torch.cuda.current_stream(self.device)
        for cand in ("step",) + tuple(blocks):
            scratch.set_option(_lib.DMPM_OPT_STEP_KERNEL, 1 if cand == "step" else 0)
            if cand != "step":
                scratch.set_block(cand)
            pbuf, eid, nbuf = self.pbuf.clone(), self.eid.clone(), self.nbuf.clone()
            scratch.run(self.scheme, self.velocity_update, gravity, self.dt, start_step, 2,
                        pbuf=pbuf, eid=eid, nbuf=nbuf)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            scratch.run(self.scheme, self.velocity_update, gravity, self.dt, start_step + 2,
                        steps, pbuf=pbuf, eid=eid, nbuf=nbuf)
            e1.record(stream)
            torch.cuda.synchronize(self.device)
            times[cand] = e0.elapsed_time(e1) / steps
        del scratch
        best_block = min((c for c in times if c != "step"), key=times.get)
        best = min(times, key=times.get)
        self.plan.set_option(_lib.DMPM_OPT_STEP_KERNEL, 1 if best == "step" else 0)
        self.plan.set_block(best_block)
        self.step_kernel = best == "step"
        self.cells_per_block = best_block
        return times

    def sim_index(self):
        """(N,) simulation index of every particle of the ensemble."""
        return torch.repeat_interleave(
            torch.arange(len(self.sims), device=self.device),
            torch.tensor(self.counts, device=self.device))

    def sim_slice(self, i):
        """Particle columns of member i in the ensemble's block."""
        return slice(self.part_off[i], self.part_off[i] + self.counts[i])
