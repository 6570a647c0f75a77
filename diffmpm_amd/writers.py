"""Output writers (writers.py of the reference): ``.npz`` snapshots."""
import abc
import logging
from pathlib import Path

import numpy as np

logger = logging.getLogger(__file__)

__all__ = ["_Writer", "EmptyWriter", "NPZWriter"]


class _Writer(abc.ABC):
    @abc.abstractmethod
    def write(self, *args, **kwargs):
        ...


class EmptyWriter(_Writer):
    """writers.py:20-25."""

    def write(self, args, transforms, **kwargs):
        pass


class NPZWriter(_Writer):
    """writers.py:28-54: particles_{step:0Nd}.npz with N digits of max_steps."""

    def write(self, args, transforms, **kwargs):
        arrays, step = args
        max_digits = int(np.log10(kwargs["max_steps"])) + 1
        if step == 0:
            req_zeros = max_digits - 1
        else:
            req_zeros = max_digits - (int(np.log10(step)) + 1)
        fileno = f"{'0' * req_zeros}{step}"
        filepath = Path(kwargs["out_dir"]).joinpath(f"particles_{fileno}.npz")
        filepath.parent.mkdir(parents=True, exist_ok=True)
        np.savez(filepath, **arrays)
        logger.info(f"Saved particle data for step {step} at {filepath}")
