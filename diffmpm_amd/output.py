"""Asynchronous, graph-safe output of the particle state.

The reference taps its state out of the compiled step loop with a host
callback every ``out_steps`` steps (diffmpm/solver.py:113-114, 187-210).
Here a snapshot is a set of device-to-host copies of the output rows of the
particle blocks (loc, velocity, stress, strain of every particle set) into
pinned host memory, enqueued on the simulation's stream in order with the
steps, and followed by an event.  A writer thread waits for that event and
calls ``writer_func``; the step loop never waits for the host, and the GPU
keeps stepping while files are written.  The copies are ordinary stream
operations on fixed pinned buffers, so a snapshot can also be captured into
a HIP graph (``graphs.StepGraph(..., out_steps=...)``): the event is then
recorded after each replay and the writer thread processes the replay's
snapshots in step order.

Snapshots occupy a ring of pinned slots; when every slot is still waiting to
be written, taking a new one waits for the oldest write to finish.
"""
import queue
import threading

import numpy as np
import torch

from . import _lib

# Rows of a snapshot (dmpm_snapshot_rows): block rows 0..3 (loc, velocity),
# then 13..24 (strain, stress).  name -> (first snapshot row, rows: None = ndim)
_SNAP_ROWS = 16
# (strain: block rows 13..18 -> snapshot rows 4..9; stress: 19..24 -> 10..15)
_LAYOUT = (("loc", 0, None), ("velocity", 2, None), ("stress", 10, 6), ("strain", 4, 6))


class _Slot:
    """Pinned host rows for one snapshot of every particle set."""

    def __init__(self, sets):
        self.host = [torch.empty((_SNAP_ROWS, len(ps)), dtype=ps.dtype, pin_memory=True)
                     for ps in sets]
        self.event = torch.cuda.Event()
        self.step = None

    def enqueue_copies(self, sets):
        """Device-to-host copies of the output rows on the current stream
        (two 2D copies per set, capturable)."""
        for buf, ps in zip(self.host, sets):
            _lib.call("dmpm_snapshot_rows", _lib.ptr(ps._buf), ps._ld,
                      _lib.dtype_code(ps.dtype), len(ps), _lib.ptr(buf),
                      _lib.stream_of(ps._buf))

    def arrays(self, sets):
        """The reference's output arrays (solver.py:191-198): per field, the
        squeezed per-set arrays stacked over the sets."""
        out = {name: [] for name, _, _ in _LAYOUT}
        for buf, ps in zip(self.host, sets):
            for name, r, k in _LAYOUT:
                kk = k or ps._ndim
                v = buf[r:r + kk].t().numpy()
                shape = (len(ps), 1, kk) if k is None else (len(ps), kk, 1)
                out[name].append(v.reshape(shape).squeeze())
        return {name: np.stack(v) for name, v in out.items()}


class SnapshotQueue:
    """Snapshots of ``sets`` (particle sets) handed to ``writer_func`` by a
    writer thread, in submission order; ``slots`` pinned host slots."""

    def __init__(self, sets, writer_func, out_dir, max_steps, slots=2):
        self.sets = list(sets)
        self.writer_func = writer_func
        self.out_dir, self.max_steps = out_dir, max_steps
        self._free = queue.Queue()
        for _ in range(max(int(slots), 1)):
            self._free.put(_Slot(self.sets))
        self._todo = queue.Queue()
        self._error = None
        self._thread = threading.Thread(target=self._loop, daemon=True)
        self._thread.start()

    def _loop(self):
        while True:
            item = self._todo.get()
            if item is None:
                self._todo.task_done()
                return
            slot, recycle = item
            try:
                slot.event.synchronize()
                if self._error is None:
                    self.writer_func((slot.arrays(self.sets), slot.step), None,
                                     out_dir=self.out_dir, max_steps=self.max_steps)
            except BaseException as e:  # reported by drain()
                self._error = e
            finally:
                if recycle is not None:
                    recycle.put(slot)
                self._todo.task_done()

    def snapshot(self, step):
        """Enqueues a snapshot of the state after step ``step`` on the current
        stream (waits only if every slot is still being written)."""
        slot = self._free.get()
        slot.step = int(step)
        slot.enqueue_copies(self.sets)
        slot.event.record()
        self._todo.put((slot, self._free))

    def drain(self):
        """Waits until every submitted snapshot is written; re-raises a
        writer error."""
        self._todo.join()
        if self._error is not None:
            e, self._error = self._error, None
            raise e

    def close(self):
        self.drain()
        self._todo.put(None)
        self._thread.join()


class CapturedSnapshots:
    """Fixed pinned slots for the snapshots that a captured rollout takes
    (one per output step of a replay): ``enqueue(k)`` inside the capture,
    ``submit(steps)`` after each replay."""

    def __init__(self, sets, count, writer_func, out_dir, max_steps):
        self.sets = list(sets)
        self.slots = [_Slot(self.sets) for _ in range(count)]
        self.q = SnapshotQueue(self.sets, writer_func, out_dir, max_steps, slots=1)

    def enqueue(self, k):
        self.slots[k].enqueue_copies(self.sets)

    def submit(self, steps):
        """Hands the replay's snapshots (taken after steps ``steps``, in slot
        order) to the writer after an event recorded now on the current
        stream."""
        ev = torch.cuda.Event()
        ev.record()
        for slot, step in zip(self.slots, steps):
            slot.event = ev
            slot.step = int(step)
            self.q._todo.put((slot, None))

    def wait(self):
        """Waits until the submitted snapshots are written (before the slots
        are overwritten by the next replay)."""
        self.q.drain()
