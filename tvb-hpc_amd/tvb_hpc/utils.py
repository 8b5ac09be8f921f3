"""
Helpers mirroring tvb_hpc/utils.py (reference): symbolic ``vars`` / ``exprs``,
variable substitution, a wall-clock ``timer`` and logging.

The reference's ``default_target`` (utils.py:32) is an OpenCL target; here the
default target is the HIP executor (``tvb_hpc.hip.HipTarget``).
"""

import logging
import os
import time

import numpy as np

from . import expr as _e

here = os.path.dirname(os.path.abspath(__file__))


def default_target():
    from .hip import HipTarget
    return HipTarget()


def can_bcast(n, m):
    return (n == 1) or (m == 1) or (n == m)


def getLogger(name):
    return logging.getLogger(name)


class NoSuchExecutable(RuntimeError):
    "Exception raised when ``which`` does not find its argument on $PATH."


def which(exe):
    "Find exe name on $PATH (utils.py:43-56)."
    if os.path.exists(exe):
        return exe
    for path in os.environ['PATH'].split(os.path.pathsep):
        maybe_path = os.path.join(path, exe)
        if os.path.exists(maybe_path):
            return maybe_path
    raise NoSuchExecutable(exe)


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
    "Wall-clock context manager logging elapsed time (utils.py:93-107)."
    logger = getLogger('tvb_hpc.utils.timer')

    def __init__(self, msg=''):
        self.msg = msg

    def __enter__(self, *args):
        self.tic = time.time()
        return self

    def __exit__(self, *args):
        self.elapsed = time.time() - self.tic
        self.logger.info('%s elapsed %.3fs', self.msg, self.elapsed)


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
