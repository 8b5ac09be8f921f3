"""
Coupling functions (mirror of tvb_hpc/coupling.py).

A coupling is a pre-summation expression of ``pre_syn`` (the delayed
observable of the source node) and ``post_syn`` (the current observable of the
target node), a weighted sum over the incoming edges, and a post-summation
expression of ``sum`` or ``mean`` (coupling.py:42-121).  ``mean`` divides the
sum by the number of nodes, not by the in-degree (network.py:129-131).
"""

import enum
import logging

import numpy as np

from . import expr as E
from .model import BaseModel
from .utils import exprs


class PostSumStat(enum.Enum):
    sum = 'sum'
    mean = 'mean'


class BaseCoupling:
    "coupling.py:42-100"

    param = {}
    pre_sum = ''
    post_sum = ''

    def pre_bounds(self):
        """(lo, hi) bounding every pre_sum value, or None when unbounded;
        an 8-bit fixed-point history needs it."""
        return None

    def __init__(self, model: BaseModel):
        self.model = model
        self.param = dict(type(self).param)
        self.param_sym = np.array([E.Var(n) for n in self.param.keys()],
                                  dtype=object)
        self.pre_sum_sym = exprs(self.pre_sum)
        self.post_sum_sym = exprs(self.post_sum)
        lname = '%s<%s>' % (type(self).__name__, type(model).__name__)
        self.logger = logging.getLogger(lname)
        self._check_io()

    def _check_io(self):
        """Broadcast obsrv / pre / post / input; obsrv is truncated to the
        number of inputs (coupling.py:68-84)."""
        obsrv_sym = self.model.obsrv_sym
        if self.model.input_sym.size < obsrv_sym.size:
            self.logger.debug('input shorter than obsrv, truncating obsrv '
                              'used for cfun.')
            obsrv_sym = obsrv_sym[:self.model.input_sym.size]
        n = max(len(obsrv_sym), len(self.pre_sum_sym),
                len(self.post_sum_sym), len(self.model.input_sym))
        terms = []
        for t in (obsrv_sym, self.pre_sum_sym, self.post_sum_sym,
                  self.model.input_sym):
            if len(t) not in (1, n):
                raise ValueError('cannot broadcast coupling terms')
            terms.append(list(t) * (n if len(t) == 1 else 1))
        self.io = list(zip(*terms))
        for i, parts in enumerate(self.io):
            self.logger.debug('io[%d] (%s) -> (%s) -> (%s) -> (%s)', i,
                              *parts)

    def post_stat(self, i: int) -> PostSumStat:
        "'mean' if the i'th post expression uses mean, else 'sum' (86-100)."
        deps = E.free_vars(E.as_expr(self.post_sum_sym[i]))
        if 'mean' in deps:
            return PostSumStat('mean')
        if 'sum' in deps:
            return PostSumStat('sum')
        raise ValueError('unknown stat in %r' % (self.post_sum, ))

    # -- helpers for the emitter ---------------------------------------------

    def pre_expr(self, k):
        return E.as_expr(self.io[k][1])

    def post_expr(self, k):
        return E.as_expr(self.io[k][2])

    def obsrv_index(self, k):
        "Index into model.obsrv of the k'th coupled observable."
        o = E.as_expr(self.io[k][0])
        for i, e in enumerate(self.model.obsrv_exprs()):
            if e == o:
                return i
        return k

    def pre_uses_post_syn(self):
        return any('post_syn' in E.free_vars(self.pre_expr(k))
                   for k in range(len(self.io)))

    def __repr__(self):
        return '%s(%r)' % (type(self).__name__, self.model)


class Linear(BaseCoupling):
    "coupling.py:103-106"
    param = {'a': 1e-3, 'b': 0}
    pre_sum = 'pre_syn',
    post_sum = 'a * sum + b',


class Sigmoidal(BaseCoupling):
    "coupling.py:109-112"
    param = {'cmin': 0, 'cmax': 0.005, 'midpoint': 6, 'r': 1, 'a': 0.56}
    pre_sum = 'cmax / (1 + exp(r * (midpoint - pre_syn)))',
    post_sum = 'a * sum',

    def pre_bounds(self):
        "pre_sum lies in [0, cmax] (for cmax >= 0 a literal)"
        cmax = self.param.get('cmax')
        if isinstance(cmax, (int, float)) and cmax >= 0:
            return 0.0, float(cmax)
        return None


class Diff(Linear):
    "coupling.py:115-116"
    pre_sum = 'pre_syn - post_syn',


class Kuramoto(Diff):
    "coupling.py:119-121"
    pre_sum = 'sin(pre_syn - post_syn)',
    post_sum = 'a * mean',


BUILTIN_COUPLINGS = {c.__name__: c for c in (Linear, Sigmoidal, Diff,
                                              Kuramoto)}
