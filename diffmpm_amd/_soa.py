"""Field views over the SoA blocks that the HIP kernels read and write.

A block is a ``(rows, ld)`` tensor; a field of ``k`` rows starting at row
``r`` is exposed in the reference's shape as a strided view, e.g. loc
``(N, 1, ndim)`` is ``block[r:r+ndim, :N].T[:, None, :]``.  Writing into a view
writes the block; assigning the attribute copies the value into the block.
"""
import torch


def vec_view(block, row, n, ndim):
    """(n, 1, ndim) view of ``ndim`` rows."""
    return block[row:row + ndim, :n].t().unsqueeze(1)


def tensor_view(block, row, n):
    """(n, 6, 1) view of 6 rows (Voigt order)."""
    return block[row:row + 6, :n].t().unsqueeze(-1)


def scalar3_view(block, row, n):
    """(n, 1, 1) view of one row."""
    return block[row, :n].view(n, 1, 1)


def scalar2_view(block, row, n):
    """(n, 1) view of one row."""
    return block[row, :n].view(n, 1)


def assign(view, value):
    """Copies ``value`` (broadcast to the view's shape) into the view."""
    if isinstance(value, torch.Tensor) and value.data_ptr() == view.data_ptr() \
            and value.shape == view.shape and value.stride() == view.stride():
        return
    v = torch.as_tensor(value, device=view.device)
    if v.dtype != view.dtype:
        v = v.to(view.dtype)
    view.copy_(v.expand_as(view) if v.dim() <= view.dim() else v.reshape(view.shape))


class Field:
    """Descriptor exposing a row range of ``obj._buf`` in a reference shape."""

    def __init__(self, row, kind):
        self.row = row
        self.kind = kind

    def __set_name__(self, owner, name):
        self.name = name

    def view(self, obj):
        n = obj._count
        if self.kind == "vec":
            return vec_view(obj._buf, self.row, n, obj._ndim)
        if self.kind == "tensor":
            return tensor_view(obj._buf, self.row, n)
        if self.kind == "s3":
            return scalar3_view(obj._buf, self.row, n)
        return scalar2_view(obj._buf, self.row, n)

    def __get__(self, obj, owner=None):
        if obj is None:
            return self
        return self.view(obj)

    def __set__(self, obj, value):
        assign(self.view(obj), value)
        hook = getattr(obj, "_on_field_set", None)
        if hook is not None:
            hook(self.name)
