"""Snapshot writers for ``MPMExplicit(writer_func=...)``.

A writer is called as ``write((arrays, step), transforms, out_dir=...,
max_steps=...)``, the call the reference's solver makes through its host
callback (diffmpm/solver.py:187-210): ``arrays`` maps loc, velocity, stress
and strain to host arrays stacked over the particle sets, ``step`` is the
index of the step just taken.  ``NPZWriter`` stores them as
``<out_dir>/particles_<step>.npz`` with the step zero-padded to as many
digits as ``max_steps`` has (the reference's file naming,
diffmpm/writers.py:30-54).  The solver hands the writer host copies made
asynchronously (``diffmpm_amd.output``), so writing never stalls the GPU.
"""
import abc
from pathlib import Path

import numpy as np

__all__ = ["_Writer", "EmptyWriter", "NPZWriter", "npz_name"]


def npz_name(step, max_steps):
    """File name of the snapshot after step ``step`` of a ``max_steps`` run."""
    width = len(str(int(max_steps)))
    return f"particles_{int(step):0{width}d}.npz"


class _Writer(abc.ABC):
    @abc.abstractmethod
    def write(self, args, transforms, **kwargs):
        ...


class EmptyWriter(_Writer):
    """Discards every snapshot."""

    def write(self, args, transforms, **kwargs):
        return None


class NPZWriter(_Writer):
    """One ``.npz`` file per snapshot (keys loc, velocity, stress, strain)."""

    def write(self, args, transforms, *, out_dir, max_steps, **kwargs):
        arrays, step = args
        target = Path(out_dir) / npz_name(step, max_steps)
        target.parent.mkdir(parents=True, exist_ok=True)
        np.savez(target, **{k: np.asarray(v) for k, v in arrays.items()})
        return target
