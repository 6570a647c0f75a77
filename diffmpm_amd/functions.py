"""Time functions for load ramps (functions.py of the reference).

The step evaluates them on the device (``mpm_common.h: interp``); ``value``
here is the host-side equivalent for callers.
"""
import abc

import numpy as np


class Function(abc.ABC):
    """functions.py:4-13."""

    def __init__(self, id):
        self.id = id

    @abc.abstractmethod
    def value(self, x):
        ...


class Unit(Function):
    """functions.py:16-31: value(t) = 1."""

    def __init__(self, id):
        super().__init__(id)

    def value(self, x):
        return 1.0


class Linear(Function):
    """functions.py:34-49: piecewise-linear ``interp(t, xvalues, fxvalues)``,
    holding the end values outside [x_0, x_n]."""

    def __init__(self, id, xvalues, fxvalues):
        self.xvalues = np.asarray(xvalues, dtype=np.float64).reshape(-1)
        self.fxvalues = np.asarray(fxvalues, dtype=np.float64).reshape(-1)
        if self.xvalues.shape != self.fxvalues.shape or self.xvalues.size == 0:
            raise ValueError("xvalues and fxvalues must be non-empty and of equal length")
        super().__init__(id)

    def value(self, x):
        return np.float32(np.interp(np.float32(x), self.xvalues.astype(np.float32),
                                    self.fxvalues.astype(np.float32)))
