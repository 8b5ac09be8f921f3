"""
tvb_hpc on MI355X: the delayed-coupling stochastic network integrator of
the-virtual-brain/tvb-hpc, re-implemented with hand-written HIP kernels for
gfx950 behind the reference's model / coupling / network / scheme / bold /
metric / transforms Python API.

Logging follows the reference (tvb_hpc/__init__.py:15-22): the level is read
from the ``TVB_LOG`` environment variable.
"""

import logging
import os

_level = os.environ.get('TVB_LOG', 'INFO').upper()
logging.basicConfig(level=getattr(logging, _level, logging.INFO))

__version__ = '0.1.0'
