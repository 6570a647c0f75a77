"""Loading records (forces.py of the reference)."""
from typing import Any, NamedTuple, Sequence


class NodalForce(NamedTuple):
    """forces.py:9-18: force applied on a set of nodes, scaled by a function
    of time."""

    node_ids: Any
    function: Any
    dir: int
    force: float


class ParticleTraction(NamedTuple):
    """forces.py:26-41: traction applied on particles ``pids`` of the particle
    sets ``pset``, scaled by a function of time."""

    pset: Sequence[int]
    pids: Any
    function: Any
    dir: int
    traction: float
