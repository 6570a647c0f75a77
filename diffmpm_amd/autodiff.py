"""Differentiable rollouts of the fused 2D explicit step (torch.autograd).

The reference differentiates its rollouts with ``jax.grad`` through
``MPMExplicit.solve_jit`` (diffmpm/solver.py:179-222).  Here each step is a
``torch.autograd.Function`` whose backward calls the hand-written adjoint
kernels (``dmpm_step_adjoint``), which recompute the forward step from the
step's saved input state.  ``rollout(..., checkpoint_every=K)`` keeps only
every K-th state during the forward pass and recomputes the states inside a
segment when the backward pass reaches it.

Differentiable state, one ``(8, N)`` tensor: rows x, y, vx, vy, sigma_xx,
sigma_yy, sigma_xy, volume (``STATE_ROWS``).  Optional parameter: log of
the Young's modulus, per particle ``(N,)`` or one scalar for the whole set.
Mass, mesh, constraints and nodal forces are constants; strain, density,
volumetric strain and sigma_zz do not feed back into the 2D dynamics and get
no gradient.  The covered configurations: one particle set without particle
tractions or particle external forces, LinearElastic or SimpleMaterial (log E
for both: a per-particle E replaces the material's), on a 2D
Quadrilateral4Node mesh (the fused step) or a Linear1D mesh (the per-phase
step; the y and vy rows of the state are constants there, and the adjoint of
a 1D step recomputes the step).

Example::

    d = Differentiable(sim)
    log_e = torch.full((N,), math.log(1e6), device="cuda", requires_grad=True)
    final = rollout(d, 100, gravity, log_e=log_e, checkpoint_every=10)
    loss = ((final[0:2] - target) ** 2).sum()
    loss.backward()          # log_e.grad
"""
import math

import torch

from . import _lib
from ._lib import call, ptr
from ._plan import Plan, plan_for

STATE_ROWS = (_lib.P_X, _lib.P_X + 1, _lib.P_V, _lib.P_V + 1, _lib.P_STRESS + 0,
              _lib.P_STRESS + 1, _lib.P_STRESS + 3, _lib.P_VOLUME)
C_ROWS = 8
# (state rows, block rows) of the contiguous runs of STATE_ROWS
_RUNS = ((slice(0, 4), slice(_lib.P_X, _lib.P_X + 4)),
         (slice(4, 6), slice(_lib.P_STRESS, _lib.P_STRESS + 2)),
         (slice(6, 7), slice(_lib.P_STRESS + 3, _lib.P_STRESS + 4)),
         (slice(7, 8), slice(_lib.P_VOLUME, _lib.P_VOLUME + 1)))
assert [_lib.P_X + 2 == _lib.P_V, _lib.P_VOLUME not in range(_lib.P_X, _lib.P_X + 4)] == [True] * 2


class Differentiable:
    """Scratch buffers and plans of the differentiable step of ``sim`` (an
    ``MPMExplicit`` with one particle set on a Quadrilateral4Node mesh) or of
    an ``ensemble.Ensemble`` (``Differentiable.of_ensemble``; per-simulation
    parameters map to particles through ``Ensemble.sim_index()``)."""

    @classmethod
    def of_ensemble(cls, ens):
        from .ensemble import BatchPlan
        d = cls.__new__(cls)
        d.sim, d.ps, d.el = None, None, None
        d._blocks, d._nodes_last = None, None
        d.dim = 2
        d.per_particle_E = all(s.mesh.particles[0]._has_youngs for s in ens.sims)
        d.n = ens.n
        d._pbuf, d._nbuf, d._eid_f = ens.pbuf, ens.nbuf, ens.eid
        d.fplan, d.aplan = ens.plan, BatchPlan(ens)
        # forward steps between two private blocks, with their own sort state
        d.bplan = BatchPlan(ens)
        d._eid_b = d._eid_f.clone()
        d._ens = ens
        d._finish(ens.scheme, ens.velocity_update, ens.dt)
        return d

    # keep each step's forward quantities for its adjoint (about 20 + 15 *
    # nodes/particles values per particle of extra memory per step) instead of
    # recomputing the step in the backward pass
    save_forward = True

    def __init__(self, sim, per_particle_E=True):
        mesh = sim.mesh
        if len(mesh.particles) != 1:
            raise ValueError("the differentiable step covers one particle set")
        ps, el = mesh.particles[0], mesh.elements
        if el.ndim not in (1, 2):
            raise ValueError("the differentiable step covers Linear1D and Quadrilateral4Node "
                             "meshes")
        if ps._has_fext:
            raise ValueError("the differentiable step covers particle sets without particle "
                             "external forces")
        if getattr(ps, "_traction_specs", []):
            raise ValueError("the differentiable step covers particle sets without particle "
                             "tractions")
        if per_particle_E and not ps._has_youngs:
            props = ps.material.properties
            key = next((k for k in ("youngs_modulus", "E") if k in props), None)
            if key is None:
                raise ValueError("log E gradients need a LinearElastic or SimpleMaterial "
                                 "material")
            ps.youngs_modulus = torch.full((len(ps),), float(props[key]), dtype=ps.dtype,
                                           device=ps.device)
            ps._plans.clear()
        self.dim = el.ndim
        if self.dim == 1:
            self.save_forward = False  # the 1D adjoint recomputes its step
        self.sim, self.ps, self.el = sim, ps, el
        self.per_particle_E = per_particle_E
        self.n = len(ps)
        self._pbuf, self._nbuf, self._eid_f = ps._buf, el.nodes._buf, ps._eid
        self.fplan = plan_for(el, ps)
        self.aplan = Plan(el, ps)  # own sort state for the recomputed steps
        # forward steps alternate between two private blocks (one-launch step
        # from one to the other, no copy of the state); own sort state
        self.bplan = Plan(el, ps)
        self._blocks = None
        self._nodes_last = None  # node block of the last private-block step
        self._eid_b = self._eid_f.clone()
        self._finish(sim.scheme, sim.velocity_update, sim.dt)

    def _finish(self, scheme, velocity_update, dt):
        # step-input block (mass, E and the other constant rows set once),
        # recomputation scratch blocks
        self.pin = self._pbuf.clone()
        self.pwork = self._pbuf.clone()
        self.nwork = self._nbuf.clone()
        self.eid = torch.empty_like(self._eid_f)
        self.scheme = scheme
        self.velocity_update = velocity_update
        self.code = _lib.DMPM_USF if scheme == "usf" else _lib.DMPM_USL
        self.vu = 1 if velocity_update else 0
        self.dt = float(dt)

    # -- state <-> particle block ---------------------------------------------
    def get_state(self):
        """The current dynamic state as a new (8, N) tensor."""
        out = torch.empty((C_ROWS, self.n), dtype=self._pbuf.dtype, device=self._pbuf.device)
        for a, b in _RUNS:
            out[a] = self._pbuf[b, : self.n]
        self._last = (out, out._version)
        return out

    def _write(self, buf, state, log_e):
        for a, b in _RUNS:
            buf[b, : self.n] = state[a]
        if log_e is not None:
            self._write_log_e(buf, log_e)
            self.__dict__.setdefault("_e_written", {})[buf.data_ptr()] = (log_e, log_e._version)

    def set_state(self, state, log_e=None):
        last = getattr(self, "_last", None)
        if last is not None and last[0] is state and state._version == last[1]:
            # the block still holds this state (the output of the last step)
            if log_e is not None:
                self._write_log_e(self._pbuf, log_e)
            return
        self._write(self._pbuf, state.detach(), log_e)

    def _write_log_e(self, buf, log_e):
        buf[_lib.P_YOUNGS, : self.n] = torch.exp(log_e.detach()).to(buf.dtype).expand(self.n)

    def _write_log_e_private(self, buf, log_e):
        """_write_log_e for a block only this object writes (the private
        step blocks, the adjoint's block): skipped when the block already
        holds exp of this very log_e (same tensor, unchanged since)."""
        cache = self.__dict__.setdefault("_e_written", {})
        key = buf.data_ptr()
        hit = cache.get(key)
        if hit is not None and hit[0] is log_e and hit[1] == log_e._version:
            return
        self._write_log_e(buf, log_e)
        cache[key] = (log_e, log_e._version)

    # -- one step ---------------------------------------------------------------
    def forward_step(self, state, log_e, gravity, i, save=False, prev=None):
        """One forward step; with save=True also returns what its adjoint
        needs (node block, kinematic nodal velocity, sort by cell).  A 2D
        simulation or ensemble steps between two private particle blocks (its
        own block is not written; a 1D one on one private block).  prev: the
        node block whose velocity rows (what a massless node keeps) and
        damping-force rows the step reads (default: the last step's block when
        the step continues it, else the simulation's node block)."""
        if self.dim == 1:
            return self._forward_1d(state, log_e, gravity, i), None
        if self._blocks is None:
            self._blocks = (self._pbuf.clone(), self._pbuf.clone())
        # the state rows are read from `state` and written to `out`; the two
        # blocks carry the other rows (strain, density, sigma_zz, volumetric
        # strain) and the constants
        last = getattr(self, "_last_b", None)
        chained = self._chained(state)
        src = last[2] if last is not None else 0
        dst = 1 - src
        if log_e is not None:
            self._write_log_e_private(self._blocks[dst], log_e)
        plan = self.bplan
        plan.refresh_tables()
        sin = state.detach()
        if not sin.is_contiguous():
            sin = sin.contiguous()
        out = torch.empty((C_ROWS, self.n), dtype=self._pbuf.dtype, device=self._pbuf.device)
        if save:
            # a fresh node block per saved step, kept for its adjoint: the step
            # writes mass, velocity, acceleration, momentum and the forces of
            # every node, reads the velocity a massless node keeps from the
            # previous step's block and copies its damping force; its
            # coordinate rows are not set (neither the step nor the adjoint
            # reads them).  The kinematic nodal velocity and the sort by cell
            # go straight to the adjoint's buffers.
            if prev is None:
                prev = self._prev_node_block(chained)
            nb = torch.empty_like(self._nbuf)
            vkin, perm, cs = self._export_buffers()
            call("dmpm_step_between_saved", plan.h, ptr(self._blocks[src]),
                 ptr(self._blocks[dst]), ptr(self._eid_b), ptr(prev), ptr(nb), self.code,
                 self.vu, plan._gravity(gravity), self.dt, int(i), 2 if chained else 0,
                 ptr(sin), ptr(out), self.n, ptr(vkin), ptr(perm), ptr(cs), plan.s)
            self._last_b = (out, out._version, dst)
            self._nodes_last = nb
            return out, (nb, vkin, perm, cs)
        # a plain step writes a private node block and reads the velocity a
        # massless node keeps (and the damping force) from that block: bring
        # in those rows of the previous step's block when it is another one
        nb = self.__dict__.get("_nplain")
        if nb is None:
            nb = self._nplain = self._nbuf.clone()
        if prev is None:
            prev = self._prev_node_block(chained)
        if prev is not nb:
            self._prev_block(self._prev_rows(prev), nb)
        call("dmpm_step_between", plan.h, ptr(self._blocks[src]), ptr(self._blocks[dst]),
             ptr(self._eid_b), ptr(nb), self.code, self.vu, plan._gravity(gravity), self.dt,
             int(i), 2 if chained else 0, ptr(sin), ptr(out), self.n, plan.s)
        self._last_b = (out, out._version, dst)
        self._nodes_last = nb
        return out, None

    def _forward_1d(self, state, log_e, gravity, i):
        """One per-phase step of a Linear1D simulation on a private particle
        block, node block and element ids."""
        if self._blocks is None:
            self._blocks = (self._pbuf.clone(),)
            self._nodes_last = self._nbuf.clone()
        blk = self._blocks[0]
        last = getattr(self, "_last_1d", None)
        if not (last is not None and last[0] is state and state._version == last[1]):
            for a, b in _RUNS:
                blk[b, : self.n] = state[a].detach()
        if log_e is not None:
            self._write_log_e_private(blk, log_e)
        plan = self.bplan
        plan.refresh_tables()
        call("dmpm_run", plan.h, ptr(blk), ptr(self._eid_b), ptr(self._nodes_last), self.code,
             self.vu, plan._gravity(gravity), self.dt, int(i), 1, 0, plan.s)
        out = torch.empty((C_ROWS, self.n), dtype=self._pbuf.dtype, device=self._pbuf.device)
        for a, b in _RUNS:
            out[a] = blk[b, : self.n]
        self._last_1d = (out, out._version)
        return out

    def _export_buffers(self):
        """Buffers for a step's kinematic nodal velocity ([2][nnodes] in the
        first 2 nnodes elements) and its sort by cell."""
        vkin = torch.empty((2, self._nbuf.shape[1]), dtype=self._nbuf.dtype,
                           device=self._nbuf.device)
        perm = torch.empty(max(self.n, 1), dtype=torch.int32, device=self._nbuf.device)
        cs = torch.empty(self._ncells() + 2, dtype=torch.int32, device=self._nbuf.device)
        return vkin, perm, cs

    # node rows a step reads from the previous node block: the velocity a
    # massless node keeps and the damping force
    _PREV_ROWS = ((_lib.N_V, _lib.N_V + 2), (_lib.N_FDAMP, _lib.N_FDAMP + 2))

    def _chained(self, state):
        """Whether a step from state continues the last private-block step
        (its output, unchanged since)."""
        last = getattr(self, "_last_b", None)
        return last is not None and last[0] is state and state._version == last[1]

    def _prev_node_block(self, chained):
        """The node block a step reads the massless-node velocity and damping
        force from: the last step's when it continues it, else the
        simulation's node block (a rollout starts from the simulation's
        nodes)."""
        if chained and self._nodes_last is not None:
            return self._nodes_last
        return self._nbuf

    def _prev_rows(self, nb):
        """The rows of node block nb that the next step reads from it."""
        return torch.cat([nb[a:b] for a, b in self._PREV_ROWS])

    def _prev_block(self, rows, out=None):
        """A node block holding rows (from _prev_rows) where a step reads them."""
        if out is None:
            out = torch.empty_like(self._nbuf)
        k = 0
        for a, b in self._PREV_ROWS:
            out[a:b].copy_(rows[k:k + b - a])
            k += b - a
        return out

    def _new_plan(self):
        """A new plan of the simulation or ensemble (its own sort state and
        workspaces)."""
        if self.sim is None:
            from .ensemble import BatchPlan
            return BatchPlan(self._ens)
        return Plan(self.el, self.ps)

    def _segment_graphs(self, kmax):
        """The captured segments of up to kmax steps (kept, with their
        buffers, while this object lives)."""
        gr = self.__dict__.get("_graphs")
        if gr is None or gr.kmax < kmax:
            if gr is not None:
                del self._graphs
                del gr
                torch.cuda.empty_cache()
            gr = self._graphs = _SegmentGraphs(self, kmax)
        return gr

    def _time_dependent(self):
        """Whether a step's result depends on its index (nodal forces with a
        time function)."""
        sims = self._ens.sims if self.sim is None else [self.sim]
        return any(c.function is not None for s in sims
                   for c in s.mesh.elements.concentrated_nodal_forces)

    def _ncells(self):
        if self.el is not None:
            return self.el._nx * self.el._ny
        return sum(s.mesh.elements._nx * s.mesh.elements._ny for s in self._ens.sims)

    def _cot_buffer(self, gout):
        """An (8, ld) cotangent block holding gout in its first n columns: the
        block itself when gout is the (8, n) view of one that an earlier
        adjoint step returned (the kernels update it in place), else a new
        one."""
        ld = self._pbuf.shape[1]
        if (gout.dim() == 2 and gout.shape == (C_ROWS, self.n) and gout.stride() == (ld, 1)
                and gout.storage_offset() == 0 and gout.dtype == self._pbuf.dtype
                and gout.untyped_storage().nbytes() >= C_ROWS * ld * gout.element_size()):
            return torch.as_strided(gout, (C_ROWS, ld), (ld, 1))
        cot = torch.empty((C_ROWS, ld), dtype=self._pbuf.dtype, device=self._pbuf.device)
        cot[:, : self.n] = gout
        return cot

    def _lebar(self, lebar):
        """The dL/dlog(E_p) buffer the adjoint kernels add into: the caller's
        (an accumulation over several steps) or a new zero one."""
        if lebar is not None:
            return lebar
        return torch.zeros(self._pbuf.shape[1], dtype=self._pbuf.dtype, device=self._pbuf.device)

    def adjoint_step_saved(self, state, log_e, out, saved, gravity, gout, lebar=None):
        """Like adjoint_step, from the quantities forward_step(save=True)
        returned (no recomputation of the step)."""
        nsaved, vkin, perm, cs = saved
        st_in, st_out = state.detach(), out.detach()
        ld = self._pbuf.shape[1]
        cot = self._cot_buffer(gout)
        lebar = self._lebar(lebar)
        if st_in.is_contiguous() and st_out.is_contiguous():
            # the kernels read the state tensors; mass and E from pin
            if log_e is not None:
                self._write_log_e_private(self.pin, log_e)
            call("dmpm_step_adjoint_saved_state", self.fplan.h, ptr(self.pin), ptr(st_in),
                 ptr(st_out), self.n, ptr(nsaved), ptr(vkin), ptr(perm), ptr(cs), ptr(cot),
                 ptr(lebar), self.code, self.vu, self.fplan._gravity(gravity), self.dt,
                 self.fplan.s)
            return cot[:, : self.n], lebar[: self.n]
        self._write(self.pin, st_in, log_e)
        if self.code == _lib.DMPM_USF:
            # the kernels read the stress and volume after the step (USF only)
            for a, b in _RUNS[1:]:
                self.pwork[b, : self.n] = st_out[a]
        call("dmpm_step_adjoint_saved", self.fplan.h, ptr(self.pin), ptr(self.pwork),
             ptr(nsaved), ptr(vkin), ptr(perm), ptr(cs), ptr(cot), ptr(lebar), self.code, self.vu,
             self.fplan._gravity(gravity), self.dt, self.fplan.s)
        return cot[:, : self.n], lebar[: self.n]

    def adjoint_step(self, state, log_e, gravity, i, gout, lebar=None):
        """Cotangents (state, log E per particle) of step i's input from the
        cotangent of its output; the log E terms are added to lebar when
        given."""
        self.aplan.refresh_tables()
        self._write(self.pin, state.detach(), log_e)
        cot = self._cot_buffer(gout)
        lebar = self._lebar(lebar)
        call("dmpm_step_adjoint", self.aplan.h, ptr(self.pin), ptr(self.pwork), ptr(self.eid),
             ptr(self.nwork), ptr(cot), ptr(lebar), self.code, self.vu,
             self.aplan._gravity(gravity), self.dt, int(i), self.aplan.s)
        return cot[:, : self.n], lebar[: self.n]


class _LogEGrad:
    """dL/dlog(E_p) of one rollout: the adjoint kernels of all its steps add
    into one buffer (in the order autograd would sum the per-step terms), and
    the rollout's first step, the last one the backward pass reaches, hands
    the sum to autograd."""

    def __init__(self):
        self.buf = None


class _StepFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state, log_e, d, gravity, i, acc=None, emit=False):
        save = d.save_forward
        out, saved = d.forward_step(state, log_e, gravity, i, save=save)
        ctx.d, ctx.gravity, ctx.i = d, gravity, i
        ctx.has_log_e = log_e is not None
        ctx.saved_fwd = save
        ctx.acc, ctx.emit = acc, emit
        extra = (out,) + saved if save else ()
        ctx.save_for_backward(state, log_e if log_e is not None else state.new_empty(0), *extra)
        return out

    @staticmethod
    def backward(ctx, gout):
        state, log_e, *extra = ctx.saved_tensors
        d = ctx.d
        le = log_e if ctx.has_log_e else None
        want_e = ctx.has_log_e and ctx.needs_input_grad[1]
        acc = ctx.acc if want_e else None
        lebar = None
        if acc is not None:
            if acc.buf is None:
                acc.buf = d._lebar(None)
            lebar = acc.buf
        if ctx.saved_fwd:
            gs, ge = d.adjoint_step_saved(state, le, extra[0], tuple(extra[1:]), ctx.gravity,
                                          gout, lebar)
        else:
            gs, ge = d.adjoint_step(state, le, ctx.gravity, ctx.i, gout, lebar)
        g_log_e = None
        if want_e and (acc is None or ctx.emit):
            if acc is not None:
                acc.buf = None  # a later backward pass starts a new sum
            g_log_e = ge if log_e.dim() > 0 else ge.sum()
            g_log_e = g_log_e.to(log_e.dtype).reshape(log_e.shape)
        return gs, g_log_e, None, None, None, None, None


def step(d, state, gravity, i, log_e=None):
    """One differentiable step i (time i*dt) of ``d`` (a ``Differentiable``)."""
    return _StepFn.apply(state, log_e, d, gravity, int(i))


class _SegmentGraphs:
    """The steps of checkpointed rollout segments of a 2D ``Differentiable``
    captured into HIP graphs (``rollout(..., graphs=True)``): per segment
    length (and start step, when nodal forces depend on time) one graph of
    the plain forward steps, and one of the recomputation with saved adjoint
    inputs followed by the adjoint steps in reverse order.  The calls are
    those of the eager segments, on plans, particle blocks and buffers that
    only these graphs use (their workspaces stay where the graphs recorded
    them), so the results are bit-identical to the eager segments; the host
    work per segment is a few copies and one replay.  Memory: kmax + 1 states,
    kmax node blocks, kinematic nodal velocities and sorts."""

    def __init__(self, d, kmax):
        if d.dim != 2:
            raise ValueError("captured segments cover 2D meshes")
        self.d, self.kmax = d, int(kmax)
        self.fp, self.ap = d._new_plan(), d._new_plan()
        pb, nbuf = d._pbuf, d._nbuf
        self.blocks = (pb.clone(), pb.clone())
        self.pin = pb.clone()
        self.eid = d._eid_f.clone()
        self.nbf = nbuf.clone()  # node block of the plain steps
        self.nb0 = nbuf.clone()  # the recomputation's previous node block
        n, ld, dev, dt = d.n, pb.shape[1], pb.device, pb.dtype
        k = self.kmax
        self.S = torch.empty((k + 1, C_ROWS, n), dtype=dt, device=dev)
        self.NB = torch.empty((k,) + tuple(nbuf.shape), dtype=dt, device=dev)
        self.VK = torch.empty((k, 2, nbuf.shape[1]), dtype=dt, device=dev)
        self.PERM = torch.empty((k, max(n, 1)), dtype=torch.int32, device=dev)
        self.CS = torch.empty((k, d._ncells() + 2), dtype=torch.int32, device=dev)
        self.cot = torch.zeros((C_ROWS, ld), dtype=dt, device=dev)
        self.lebar = torch.zeros(ld, dtype=dt, device=dev)
        self.timed = d._time_dependent()
        self.graphs = {}
        self._warm = False

    def set_input(self, state, log_e):
        """Segment input state and log E into the buffers the graphs read."""
        d = self.d
        self.S[0].copy_(state.detach())
        if log_e is not None:
            for b in self.blocks + (self.pin,):
                d._write_log_e_private(b, log_e)

    def _fwd(self, j, i, grav):
        d = self.d
        call("dmpm_step_between", self.fp.h, ptr(self.blocks[j % 2]),
             ptr(self.blocks[1 - j % 2]), ptr(self.eid), ptr(self.nbf), d.code, d.vu, grav, d.dt,
             int(i), 2 if j else 0, ptr(self.S[j]), ptr(self.S[j + 1]), d.n, self.fp.s)

    def _saved(self, j, i, grav):
        d = self.d
        prev = self.nb0 if j == 0 else self.NB[j - 1]
        call("dmpm_step_between_saved", self.fp.h, ptr(self.blocks[j % 2]),
             ptr(self.blocks[1 - j % 2]), ptr(self.eid), ptr(prev), ptr(self.NB[j]), d.code,
             d.vu, grav, d.dt, int(i), 2 if j else 0, ptr(self.S[j]), ptr(self.S[j + 1]), d.n,
             ptr(self.VK[j]), ptr(self.PERM[j]), ptr(self.CS[j]), self.fp.s)

    def _adj(self, j, grav):
        d = self.d
        call("dmpm_step_adjoint_saved_state", self.ap.h, ptr(self.pin), ptr(self.S[j]),
             ptr(self.S[j + 1]), d.n, ptr(self.NB[j]), ptr(self.VK[j]), ptr(self.PERM[j]),
             ptr(self.CS[j]), ptr(self.cot), ptr(self.lebar), d.code, d.vu, grav, d.dt,
             self.ap.s)

    def _graph(self, kind, i0, k, gravity):
        grav = self.fp._gravity(gravity)
        key = (kind, i0 if self.timed else None, k, tuple(grav))
        g = self.graphs.get(key)
        if g is not None:
            return g
        self.fp.refresh_tables()
        self.ap.refresh_tables()
        if not self._warm:
            # one eager call of each kind allocates the plans' workspaces
            # before any capture records their addresses
            self._fwd(0, i0, grav)
            self._saved(0, i0, grav)
            self._adj(0, grav)
            self.lebar.zero_()
            self._warm = True
        torch.cuda.synchronize(self.S.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            if kind == "fwd":
                for j in range(k):
                    self._fwd(j, i0 + j, grav)
            else:
                for j in range(k):
                    self._saved(j, i0 + j, grav)
                for j in range(k - 1, -1, -1):
                    self._adj(j, grav)
        self.graphs[key] = g
        return g

    def forward(self, state, log_e, gravity, i0, k, first=False):
        """The state after steps i0 .. i0 + k - 1 from ``state``, and the node
        rows its first step read from the previous node block (_prev_rows).
        first: the rollout's first segment, which starts from the
        simulation's node block (later ones from the last segment's)."""
        self.set_input(state, log_e)
        g = self._graph("fwd", i0, k, gravity)
        if first:
            self.d._prev_block(self.d._prev_rows(self.d._nbuf), self.nbf)
        rows = self.d._prev_rows(self.nbf)
        g.replay()
        return self.S[k].clone(), rows

    def backward(self, state, log_e, gravity, i0, k, gout, first, prev_rows):
        """Cotangent of the segment's input state from ``gout``; the log E
        terms are added to ``self.lebar`` (cleared first when ``first``).
        prev_rows: the node rows the segment's first step read (_prev_rows)."""
        self.set_input(state, log_e)
        self.d._prev_block(prev_rows, self.nb0)
        g = self._graph("bwd", i0, k, gravity)
        self.cot[:, : self.d.n].copy_(gout)
        if first:
            self.lebar.zero_()
        g.replay()
        return self.cot[:, : self.d.n].clone()


class _SegmentFn(torch.autograd.Function):
    """k steps of a checkpointed rollout as one autograd node.  The forward
    pass runs the plain steps (nothing saved, no autograd node per step) and
    keeps the segment's input state; the backward pass recomputes the
    segment's steps from it with their adjoint quantities saved, then runs
    the k adjoint steps in reverse order.  The adjoint kernels see the same
    inputs in the same order as in an uncheckpointed rollout, so the
    gradients are bit-identical to it."""

    @staticmethod
    def forward(ctx, state, log_e, d, gravity, i0, k, acc=None, emit=False, gr=None):
        # the node rows the segment's first step reads from the previous node
        # block, so that its recomputation repeats it exactly
        if gr is not None:
            st, ctx.prev_rows = gr.forward(state, log_e, gravity, i0, k, first=emit)
        else:
            ctx.prev_rows = d._prev_rows(d._prev_node_block(d._chained(state)))
            st = state
            for i in range(i0, i0 + k):
                st, _ = d.forward_step(st, log_e, gravity, i, save=False)
        ctx.d, ctx.gravity, ctx.i0, ctx.k, ctx.gr = d, gravity, i0, k, gr
        ctx.has_log_e = log_e is not None
        ctx.acc, ctx.emit = acc, emit
        ctx.save_for_backward(state, log_e if log_e is not None else state.new_empty(0))
        return st

    @staticmethod
    def backward(ctx, gout):
        state, log_e = ctx.saved_tensors
        d, gravity, i0, k = ctx.d, ctx.gravity, ctx.i0, ctx.k
        le = log_e if ctx.has_log_e else None
        want_e = ctx.has_log_e and ctx.needs_input_grad[1]
        acc = ctx.acc if want_e else None
        gr = ctx.gr
        if gr is not None:
            # the graphs add the log E terms into their own buffer
            first = acc is None or acc.buf is None
            if acc is not None:
                acc.buf = gr.lebar
            with torch.no_grad():
                g = gr.backward(state, le, gravity, i0, k, gout, first, ctx.prev_rows)
            ge = gr.lebar[: d.n]
        else:
            g, ge = _SegmentFn._eager_backward(d, state, le, gravity, i0, k, gout, acc,
                                               ctx.prev_rows)
        ctx.prev_rows = None
        g_log_e = None
        if want_e and (acc is None or ctx.emit):
            if acc is not None:
                acc.buf = None  # a later backward pass starts a new sum
            g_log_e = ge if log_e.dim() > 0 else ge.sum()
            g_log_e = g_log_e.to(log_e.dtype).reshape(log_e.shape).clone()
        return g, g_log_e, None, None, None, None, None, None, None

    @staticmethod
    def _eager_backward(d, state, le, gravity, i0, k, gout, acc, prev_rows):
        lebar = None
        if acc is not None:
            if acc.buf is None:
                acc.buf = d._lebar(None)
            lebar = acc.buf
        save = d.save_forward
        with torch.no_grad():
            states, saved = [state.detach()], []
            prev = d._prev_block(prev_rows) if save else None
            for i in range(i0, i0 + k):
                out, sv = d.forward_step(states[-1], le, gravity, i, save=save,
                                         prev=prev if i == i0 else None)
                states.append(out)
                saved.append(sv)
            g, ge = gout, None
            for j in range(k - 1, -1, -1):
                if save:
                    g, ge = d.adjoint_step_saved(states[j], le, states[j + 1], saved[j], gravity,
                                                 g, lebar)
                else:
                    g, ge = d.adjoint_step(states[j], le, gravity, i0 + j, g, lebar)
                saved[j] = None  # free each step's node block once its adjoint ran
        return g, ge


def rollout(d, nsteps, gravity, state=None, log_e=None, start_step=0, checkpoint_every=0,
            graphs=False):
    """Differentiable rollout of ``nsteps`` steps from ``state`` (default: the
    current state of the simulation).  Returns the final (8, N) state.  With
    ``checkpoint_every=K`` only segment-start states are kept for backward:
    the forward pass runs K plain steps per segment, and the backward pass
    recomputes a segment's steps (saving what their adjoints read) when it
    reaches it.  With ``graphs=True`` (2D, checkpointed) each segment's steps
    run as HIP graph replays, captured on first use per segment length (and
    start step, when nodal forces depend on time), with the same results.  A
    rollout writes neither the simulation's (or the ensemble's) particle block
    nor its node block: its steps alternate between private particle blocks
    and write private node blocks (its first step reads the velocity a
    massless node keeps and the damping force from the simulation's nodes).
    Use the returned tensors."""
    if state is None:
        state = d.get_state()
    if not torch.is_grad_enabled() or not (
            state.requires_grad or (log_e is not None and log_e.requires_grad)):
        # nothing to differentiate: plain steps, nothing saved
        with torch.no_grad():
            if graphs and checkpoint_every and checkpoint_every < nsteps:
                gr = d._segment_graphs(checkpoint_every)
                for i0 in range(start_step, start_step + nsteps, checkpoint_every):
                    k = min(checkpoint_every, start_step + nsteps - i0)
                    state, _ = gr.forward(state, log_e, gravity, i0, k, first=i0 == start_step)
                return state
            for i in range(start_step, start_step + nsteps):
                state, _ = d.forward_step(state, log_e, gravity, i, save=False)
            return state
    # log E (per particle or one scalar): the adjoint kernels of every step
    # add dL/dlog(E_p) into one per-particle buffer, summed once for a scalar
    acc = _LogEGrad() if log_e is not None else None

    if checkpoint_every and checkpoint_every < nsteps:
        gr = d._segment_graphs(checkpoint_every) if graphs else None
        for i0 in range(start_step, start_step + nsteps, checkpoint_every):
            k = min(checkpoint_every, start_step + nsteps - i0)
            state = _SegmentFn.apply(state, log_e, d, gravity, i0, k, acc, i0 == start_step, gr)
        return state
    if graphs:
        raise ValueError("graphs=True needs checkpoint_every below nsteps")
    st = state
    for i in range(start_step, start_step + nsteps):
        st = _StepFn.apply(st, log_e, d, gravity, i, acc, i == start_step)
    return st
