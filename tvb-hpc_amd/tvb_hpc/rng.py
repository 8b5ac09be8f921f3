"""
Normal random numbers (mirror of tvb_hpc/rng.py).

The reference fills a float32 array with Philox4x32 + Box-Muller normals
from a g++-compiled Random123 template (rng.py:29-96; its key is left
uninitialised, rng.py:41).  ``RNG.fill`` here runs the same generator on the
device: Philox4x32-10 with key = (seed lo, seed hi), counter (i, i >> 32, 0,
0) for out[4i .. 4i+3], two Box-Muller pairs per counter.  The integrator's
own noise uses the same generator with counter (step, node | svar/4 << 24,
instance, step >> 32).
"""

import numpy as np

from .hip import Context


class RNG:

    def __init__(self, context=None):
        self.ctx = context or Context.default()

    def fill(self, array, seed=42):
        "rng.py:90-96: fill a float32 array with standard normals"
        if array.dtype != np.float32:
            raise TypeError('array must be float32')
        flat = self.ctx.rng_normals(array.size, seed)
        array.reshape(-1)[:] = flat
        return array
