"""
The drop-in entry point (mirror of tvb_hpc/transforms.py:23-49).

``network_time_step(model, coupling, scheme, target=None)`` fuses model,
network and scheme into one time-stepping kernel and returns an executor.
With the default HIP target the executor is a
``tvb_hpc.hip.NetworkStepExecutor``: call it with the reference kernel's
arguments (nstep, nnode, ntime, state, input, param, drift, diffs, obsrv,
nnz, delays, row, col, weights, <coupling params>, i_step_0); the ring slot of
step n is (n) % ntime as in transforms.py:45-46.
"""

import logging

from . import utils
from .coupling import BaseCoupling
from .model import BaseModel
from .scheme import TimeStepScheme

LOG = logging.getLogger('tvb_hpc')


def network_time_step(model: BaseModel, coupling: BaseCoupling,
                      scheme: TimeStepScheme, target=None):
    target = target or utils.default_target()
    if hasattr(target, 'get_kernel_executor'):
        return target.get_kernel_executor(model, coupling, scheme)
    raise TypeError('unsupported target %r' % (target, ))
