"""Velocity constraint (constraint.py of the reference)."""
import torch


class Constraint:
    """constraint.py:4-44: fixes velocity component ``dir`` to ``velocity``.

    Element constraints are applied by the HIP grid kernels; ``apply`` is the
    generic form for any object with velocity/momentum/acceleration/mass
    fields (constraint.py:29-44)."""

    def __init__(self, dir: int, velocity: float):
        self.dir = dir
        self.velocity = velocity

    def __repr__(self):
        return f"Constraint(dir={self.dir}, velocity={self.velocity})"

    def apply(self, obj, ids):
        ids = torch.as_tensor(ids, device=obj.velocity.device).long().reshape(-1)
        v = obj.velocity
        v[ids, :, self.dir] = self.velocity
        obj.momentum[ids, :, self.dir] = obj.mass[ids, :, 0] * self.velocity
        obj.acceleration[ids, :, self.dir] = 0
