"""
Helpers with the names of tvb_hpc/utils.py (reference): symbolic ``vars`` /
``exprs``, variable substitution, a wall-clock ``timer``.

The reference's ``default_target`` (utils.py:32) is an OpenCL target; here the
default target is the HIP executor (``tvb_hpc.hip.HipTarget``).
"""

import logging
import time

import numpy as np

from . import expr as _e


def default_target():
    from .hip import HipTarget
    return HipTarget()


def parse(text):
    "Parse an expression string (stands in for ``pymbolic.parse``)."
    return _e.parse(text)


def simplify(expr):
    "Light algebraic simplification (utils.py:59-64 used sympy)."
    return _e.simplify(_e.as_expr(expr))


def vars(svars: str):
    """Array of symbolic variables from a space-separated string
    (utils.py:67-73)."""
    return np.array([_e.Var(v) for v in svars.split()], dtype=object)


def exprs(sexprs):
    """Array of expressions from a sequence of strings / numbers
    (utils.py:76-90); numbers are kept as numbers."""
    out = []
    for s in sexprs:
        if isinstance(s, (int, float)):
            out.append(s)
        else:
            out.append(_e.parse(s))
    arr = np.empty(len(out), dtype=object)
    arr[:] = out
    return arr


class timer:
    """Measures the wall time of a ``with`` block with the monotonic
    performance counter; the seconds are kept in ``elapsed`` and reported
    at DEBUG level on the 'tvb_hpc' logger."""

    def __init__(self, msg=''):
        self.msg = msg
        self.elapsed = None

    def __enter__(self):
        self._t0 = time.perf_counter()
        return self

    def __exit__(self, *exc):
        self.elapsed = time.perf_counter() - self._t0
        logging.getLogger('tvb_hpc').debug('%s: %.6f s', self.msg or 'timer',
                                           self.elapsed)
        return False


class VarSubst:
    "Substitute variables by name for expressions (utils.py:110-120)."

    def __init__(self, verbose=False, **substs):
        self.verbose = verbose
        self.substs = substs

    def __call__(self, expr):
        expr = _e.parse(expr)
        if self.verbose:
            print(sorted(_e.free_vars(expr)))
        return _e.subst(expr, self.substs)


def subst_vars(expr, **var_name_to_exprs):
    "utils.py:123-125"
    return VarSubst(**var_name_to_exprs)(_e.parse(str(expr)))
