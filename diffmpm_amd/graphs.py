"""Rollouts captured into HIP graphs.

The reference compiles its step loop once (``jax.lax.fori_loop`` /
``lax.scan`` in diffmpm/solver.py:179-213) and replays it.  ``StepGraph`` is
the HIP counterpart: ``nsteps`` steps of a simulation (one particle set on a
2D Quadrilateral4Node mesh, the fused path) or of an ``Ensemble`` are
captured once into a HIP graph through ``torch.cuda.CUDAGraph``, and every
``replay()`` launches the whole sequence (per step: the re-sort stages and
the one-launch step) with one call.  The steps' indices live in the plan's
device step counter (``DMPM_RUN_CLOCK``), so each replay continues in time
where the previous one ended (time functions of nodal forces see the right
t), and results equal the same steps run eagerly, bit for bit.

The graph holds the addresses of the simulation's particle and node blocks:
it stays valid while they live and the plan's tables (constraints, nodal
forces, functions) are not changed.  Replays run on the current stream.
"""
from ctypes import c_int64, byref

import torch

from . import _lib
from ._lib import call, ptr


class StepGraph:
    """``nsteps`` steps of ``target`` (``MPMExplicit`` or ``Ensemble``)
    captured into one HIP graph; the first replay runs steps start_step ..
    start_step + nsteps - 1, the next one the following nsteps, and so on.

    With ``out_steps`` (an ``MPMExplicit`` with a ``writer_func``) the graph
    also takes the reference's output snapshots (after every step i with
    i % out_steps == 0, solver.py:200-210): the captured run stops after each
    such step and copies the output rows into pinned host memory
    (``dmpm_snapshot_rows``), and after each replay a writer thread hands the
    snapshots to ``writer_func`` in step order.  nsteps must then be a
    multiple of out_steps (every replay has the same layout); a replay first
    waits until the previous replay's snapshots are written, since it
    overwrites their host buffers."""

    def __init__(self, target, gravity, nsteps, start_step=0, keep_temporaries=None,
                 out_steps=None):
        from .ensemble import Ensemble
        if nsteps < 1:
            raise ValueError("nsteps must be positive")
        if isinstance(target, Ensemble):
            if out_steps:
                raise ValueError("captured output covers single simulations")
            plan = target.plan
            pbuf, eid, nbuf = target.pbuf, target.eid, target.nbuf
            temps = bool(keep_temporaries)
            fext = False
        else:
            if len(target.mesh.particles) != 1:
                raise ValueError("a captured rollout holds one particle set")
            from ._plan import plan_for
            ps = target.mesh.particles[0]
            plan = plan_for(target.mesh.elements, ps)
            plan.refresh_tables()
            pbuf, eid, nbuf = ps._buf, ps._eid, target.mesh.elements.nodes._buf
            temps = target.keep_temporaries if keep_temporaries is None else keep_temporaries
            fext = ps._has_fext
        self.target, self.plan, self.nsteps = target, plan, int(nsteps)
        self._next = int(start_step)  # first step of the next replay (host count)
        self._keep = (pbuf, eid, nbuf)  # the graph holds their addresses
        code = _lib.DMPM_USF if target.scheme == "usf" else _lib.DMPM_USL
        flags = (1 if temps else 0) | (2 if fext else 0) | _lib.DMPM_RUN_CLOCK
        grav = plan._gravity(gravity)
        # chunks of the captured run: (steps, snapshot after them?)
        chunks = [(self.nsteps, False)]
        self._snaps = None
        if out_steps:
            out = int(out_steps)
            if target.writer_func is None:
                raise ValueError("out_steps needs the simulation's writer_func")
            if self.nsteps % out:
                raise ValueError("nsteps must be a multiple of out_steps")
            first = (-int(start_step)) % out  # offset of the first output step
            offs = list(range(first, self.nsteps, out))
            chunks, done = [], 0
            for o in offs:
                chunks.append((o + 1 - done, True))
                done = o + 1
            if done < self.nsteps:
                chunks.append((self.nsteps - done, False))
            self._offsets = offs
            from .output import CapturedSnapshots
            self._snaps = CapturedSnapshots(target.mesh.particles, len(offs),
                                            target.writer_func, target.out_dir,
                                            target.sim_steps)
        device = pbuf.device
        with torch.cuda.device(device):
            stream = _lib.stream_of(pbuf)
            call("dmpm_prepare_run", plan.h, code, flags, stream)
            call("dmpm_plan_set_clock", plan.h, int(start_step), stream)
            torch.cuda.synchronize(device)
            # the captured graph is kept (not only its executable form) so
            # that kernel_nodes() can count its nodes
            self.graph = torch.cuda.CUDAGraph(keep_graph=True)
            with torch.cuda.graph(self.graph):
                k = 0
                for n, snap in chunks:
                    if n:
                        call("dmpm_run", plan.h, ptr(pbuf), ptr(eid), ptr(nbuf), code,
                             1 if target.velocity_update else 0, grav, float(target.dt), 0,
                             n, flags, _lib.stream_of(pbuf))
                    if snap:
                        self._snaps.enqueue(k)
                        k += 1
            self.graph.instantiate()

    def kernel_nodes(self):
        """The number of kernel nodes of the captured graph (all nsteps steps)."""
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        g = ctypes.c_void_p(self.graph.raw_cuda_graph())
        n = ctypes.c_size_t(0)
        if hip.hipGraphGetNodes(g, None, ctypes.byref(n)) != 0:
            raise RuntimeError("hipGraphGetNodes failed")
        nodes = (ctypes.c_void_p * n.value)()
        if hip.hipGraphGetNodes(g, nodes, ctypes.byref(n)) != 0:
            raise RuntimeError("hipGraphGetNodes failed")
        kinds = ctypes.c_int(0)
        count = 0
        for i in range(n.value):
            if hip.hipGraphNodeGetType(ctypes.c_void_p(nodes[i]), ctypes.byref(kinds)) != 0:
                raise RuntimeError("hipGraphNodeGetType failed")
            count += kinds.value == 0  # hipGraphNodeTypeKernel
        return count

    def replay(self):
        """Runs the next ``nsteps`` steps (enqueued on the current stream)."""
        if self._snaps is not None:
            self._snaps.wait()
        self.graph.replay()
        if self._snaps is not None:
            self._snaps.submit([self._next + o for o in self._offsets])
        self._next += self.nsteps

    def wait_output(self):
        """Waits until every replay's snapshots are written."""
        if self._snaps is not None:
            self._snaps.wait()

    @property
    def step(self):
        """Index of the next step a replay runs (waits for the stream)."""
        v = c_int64()
        call("dmpm_plan_get_clock", self.plan.h, byref(v), _lib.stream_of(self._keep[0]))
        return int(v.value)
