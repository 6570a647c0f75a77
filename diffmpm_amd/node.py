"""Nodal state (node.py of the reference) stored as one SoA device block."""
from typing import Optional, Tuple

import torch

from . import _lib
from ._soa import Field


class Nodes:
    """node.py:8-125.  Fields keep the reference shapes: vectors
    ``(nnodes, 1, ndim)``, mass ``(nnodes, 1, 1)``; they are views of
    ``self._buf`` (rows of include/diffmpm_amd.h's node layout)."""

    loc = Field(_lib.N_LOC, "vec")
    velocity = Field(_lib.N_V, "vec")
    acceleration = Field(_lib.N_A, "vec")
    mass = Field(_lib.N_MASS, "s3")
    momentum = Field(_lib.N_P, "vec")
    f_int = Field(_lib.N_FINT, "vec")
    f_ext = Field(_lib.N_FEXT, "vec")
    f_damp = Field(_lib.N_FDAMP, "vec")

    def __init__(self, nnodes: int, loc, initialized: Optional[bool] = None,
                 data: Tuple = tuple(), *, device=None, dtype=None):
        dev = torch.device(device) if device is not None else (
            loc.device if isinstance(loc, torch.Tensor) else _lib.default_device())
        dtype = dtype or torch.float32
        loc = torch.as_tensor(loc, device=dev).to(dtype)
        if loc.ndim != 3:
            raise ValueError(f"`loc` should be of size (nnodes, 1, ndim); found {tuple(loc.shape)}")
        self.nnodes = int(nnodes)
        self._count = loc.shape[0]
        self._ndim = loc.shape[-1]
        self._ld = _lib.round_up(self._count)
        self._buf = torch.zeros((_lib.N_ROWS, self._ld), dtype=dtype, device=dev)
        self.loc = loc
        if initialized is None or not data:
            self._buf[_lib.N_MASS, : self._count] = 1  # node.py:67
        else:
            (self.velocity, self.acceleration, self.mass, self.momentum,
             self.f_int, self.f_ext, self.f_damp) = data
        self.initialized = True

    @property
    def dtype(self):
        return self._buf.dtype

    @property
    def device(self):
        return self._buf.device

    def reset_values(self):
        """node.py:105-113."""
        loc = self._buf[_lib.N_LOC:_lib.N_LOC + 2].clone()
        self._buf.zero_()
        self._buf[_lib.N_LOC:_lib.N_LOC + 2] = loc

    def __len__(self):
        return self.nnodes

    def __repr__(self):
        return f"Nodes(n={self.nnodes})"

    def get_total_force(self):
        """node.py:123-125."""
        return self.f_int + self.f_ext + self.f_damp
