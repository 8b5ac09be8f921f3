"""
Neural mass model definitions (mirror of tvb_hpc/model.py).

A model is described by string attributes on a class, exactly as in the
reference (model.py:31-154): ``state``, ``input``, ``param`` (space separated
names), ``drift``, ``diffs``, ``obsrv`` (tuples of expressions), ``const``
(defaults), ``auxex`` (ordered auxiliary assignments) and ``limit`` (per state
variable ranges, used for initial conditions; Kuramoto also wraps its phase).

Where the reference turns these into loopy instructions (model.py:105-140),
this package feeds them to the HIP emitter (``tvb_hpc.emit``).  The order of
evaluation per node is the reference's: constants, unpack state / input /
param, auxex in order, then drift, diffs and obsrv.
"""

from typing import List

import numpy as np

from . import expr as E
from .utils import vars, exprs


class BaseModel:
    """Parses attributes on subclasses defining a networked SDE
    (model.py:31-55)."""

    dt = 1e-3
    state = ''
    input = ''
    param = ''
    drift = '',
    diffs = '',
    obsrv = '',
    const = {}
    auxex = []
    limit = []
    # indices of state variables wrapped into their limit after each step
    wrap = ()

    def __init__(self):
        # per-instance copies: the reference mutates class-level dicts
        # (examples/hackathon.py:27); copying keeps instances independent.
        self.const = dict(type(self).const)
        self.auxex = list(type(self).auxex)
        self.state_sym = vars(self.state)
        self.input_sym = vars(self.input)
        self.param_sym = vars(self.param)
        self.drift_sym = exprs(self.drift)
        self.diffs_sym = exprs(self.diffs)
        self.obsrv_sym = exprs(self.obsrv)

    # -- reference API -----------------------------------------------------

    @property
    def indvars(self):
        return np.r_[self.state_sym, self.input_sym, self.param_sym]

    def partial(self, expr):
        "Partial derivatives of expr w.r.t. state, input, param (61-65)."
        out = [E.diff(E.as_expr(expr), v.name) for v in self.indvars]
        arr = np.empty(len(out), dtype=object)
        arr[:] = out
        return arr

    def prep_arrays(self, nnode: int) -> List[np.ndarray]:
        """state, input, param, drift, diffs, obsrv arrays for one instance
        (model.py:67-84): state uniform in ``limit``, param from ``const``."""
        dtype = np.float32
        arrs = []
        for key in 'state input param drift diffs obsrv'.split():
            shape = nnode, len(getattr(self, key + '_sym'))
            arrs.append(np.zeros(shape, dtype))
        state = arrs[0]
        for i, (lo, hi) in enumerate(self.limit):
            state[:, i] = np.random.uniform(float(lo), float(hi),
                                            size=(nnode, ))
        param = arrs[2]
        for i, psym in enumerate(self.param_sym):
            param[:, i] = self.const[psym.name]
        return arrs

    # -- helpers for the emitter -------------------------------------------

    @property
    def names(self):
        "state, input, param names"
        return ([v.name for v in self.state_sym],
                [v.name for v in self.input_sym],
                [v.name for v in self.param_sym])

    def auxex_exprs(self):
        return [(lhs, E.parse(rhs)) for lhs, rhs in self.auxex]

    def drift_exprs(self):
        return [E.as_expr(e) for e in self.drift_sym]

    def diffs_exprs(self):
        return [E.as_expr(e) for e in self.diffs_sym]

    def obsrv_exprs(self):
        return [E.as_expr(e) for e in self.obsrv_sym]

    def wrap_ranges(self):
        "[(state index, lo, hi)] for state variables wrapped after a step."
        return [(i, float(self.limit[i][0]), float(self.limit[i][1]))
                for i in self.wrap]

    def initial_state(self, nnode, ninst, rng=None, dtype=np.float32):
        "Uniform draws within ``limit``, layout (nsvar, nnode, ninst)."
        rng = rng or np.random.default_rng(0)
        S = len(self.state_sym)
        x = np.zeros((S, nnode, ninst), dtype)
        for i, (lo, hi) in enumerate(self.limit):
            x[i] = rng.uniform(float(lo), float(hi), size=(nnode, ninst))
        return x

    def __repr__(self):
        return '%s()' % type(self).__name__


class _TestModel(BaseModel):
    "model.py:157-166"
    state = 'y1 y2'
    input = 'i1'
    param = 'a b c'
    auxex = [('y1_3', 'y1 * y1 * y1')]
    drift = '(y1 - y1_3/3 + y2 + e)*b', '(a - y1 + i1 + d)/b'
    diffs = 'c', 'c'
    obsrv = 'y1',
    const = {'d': 3.0, 'e': -12.23904e-2, 'a': 1.05, 'b': 3.0, 'c': 1e-5}
    limit = (-1, 1), (-1, 1)


class Kuramoto(BaseModel):
    "Kuramoto model of phase synchronization (model.py:169-182)."
    state = 'theta'
    limit = (0, 2 * np.pi),
    input = 'I'
    param = 'omega'
    drift = 'omega + I',
    diffs = 0,
    obsrv = 'theta', 'sin(theta)'
    const = {'omega': 1.0}
    # model.py:180-182 wraps theta into [0, 2pi) with _wrap_limit(0)
    wrap = (0, )


class HMJE(BaseModel):
    "Hindmarsh-Rose-Jirsa Epileptor (model.py:185-206)."
    state = 'x1 y1 z x2 y2 g'
    limit = (-2, 1), (20, 2), (2, 5), (-2, 0), (0, 2), (-1, 1)
    input = 'c1 c2'
    param = 'x0 Iext r'
    drift = (
        'tt * (y1 - z + Iext + Kvf * c1 + ('
        '  (x1 <  0)*(-a * x1 * x1 + b * x1)'
        '+ (x1 >= 0)*(slope - x2 + 0.6 * (z - 4)**2)'
        ') * x1)',
        'tt * (c - d * x1 * x1 - y1)',
        'tt * (r * (4 * (x1 - x0) - z +  Ks * c1))',
        'tt * (-y2 + x2 - x2*x2*x2 + Iext2 + 2 * g - 0.3 * (z - 3.5) + Kf * c2)',
        'tt * ((-y2 + (x2 >= (-0.25)) * (aa * (x2 + 0.25))) / tau)',
        'tt * (-0.01 * (g - 0.1 * x1))'
    )
    diffs = 0, 0, 0, 0.0003, 0.0003, 0
    obsrv = 'x1', 'x2', 'z', '-x1 + x2'
    const = {'Iext2': 0.45, 'a': 1.0, 'b': 3.0, 'slope': 0.0, 'tt': 1.0, 'c':
             1.0, 'd': 5.0, 'Kvf': 0.0, 'Ks': 0.0, 'Kf': 0.0, 'aa': 6.0, 'tau':
             10.0, 'x0': -1.6, 'Iext': 3.1, 'r': 0.00035}


class RWW(BaseModel):
    "Reduced Wong-Wang firing rate model (model.py:209-223)."
    state = 'S'
    limit = (0, 1),
    input = 'c'
    param = 'w io'
    const = {'a': 0.270, 'b': 0.108, 'd': 154.0, 'g': 0.641,
             'ts': 100.0, 'J': 0.2609, 'w': 0.6, 'io': 0.33}
    auxex = [
        ('x', 'w * J * S + io + J * c'),
        ('h', '(a * x - b) / (1 - exp(-d*(a*x - b)))')
    ]
    drift = '- (S / ts) + (1 - S) * h * g',
    diffs = 0.01,
    obsrv = 'S',


class ReducedWongWang(RWW):
    """RWW with h(u) = u / (1 - exp(-d u)), u = a x - b, evaluated without
    its removable singularity.

    The reference's RWW (model.py:219) evaluates h as written, which is 0/0
    whenever a*x rounds exactly to b; in float32 that happens in practice
    (the CPU oracle hits it within 1000 steps of a 24-node, 96-instance
    run).  Here 1 - exp(-d u) is computed as -expm1(-d u) (no cancellation
    near u = 0) and the limit 1/d is used at u == 0."""
    auxex = [
        ('x', 'w * J * S + io + J * c'),
        ('u', 'a * x - b'),
        ('h', 'u / (-expm1(-d * u) + (u == 0)) + (u == 0) / d'),
    ]


class JansenRit(BaseModel):
    "Jansen-Rit model of visual evoked potentials (model.py:226-247)."
    state = 'y0 y1 y2 y3 y4 y5'
    limit = (-1, 1), (-500, 500), (-50, 50), (-6, 6), (-20, 20), (-500, 500)
    const = {'A': 3.25, 'B': 22.0, 'a': 0.1, 'b': 0.05, 'v0': 5.52,
             'nu_max': 0.0025, 'r': 0.56, 'J': 135.0, 'a_1': 1.0, 'a_2': 0.8,
             'a_3': 0.25, 'a_4': 0.25, 'mu': 0.22}
    input = 'lrc'
    param = 'v0 r'
    auxex = [
        ('sigm_y1_y2', '2 * nu_max / (1 + exp(r * (v0 - (y1 - y2))))'),
        ('sigm_y0_1', '2 * nu_max / (1 + exp(r * (v0 - (a_1 * J * y0))))'),
        ('sigm_y0_3', '2 * nu_max / (1 + exp(r * (v0 - (a_3 * J * y0))))'),
    ]
    drift = (
        'y3', 'y4', 'y5',
        'A * a * sigm_y1_y2 - 2 * a * y3 - a**2 * y0',
        'A * a * (mu + a_2 * J * sigm_y0_1 + lrc) - 2 * a * y4 - a**2 * y1',
        'B * b * (a_4 * J * sigm_y0_3) - 2 * b * y5 - b**2 * y2'
    )
    diffs = 0, 0, 0, 0, 0, 0
    obsrv = 'y0 - y1',


class Linear(BaseModel):
    "Linear differential equation (model.py:250-259)."
    state = 'x'
    limit = (-10, 10),
    input = 'c'
    param = 'lam'
    const = {'lam': -1}
    drift = 'lam * x + c',
    diffs = 1e-2,
    obsrv = 'x',


class G2DO(BaseModel):
    """Generic nonlinear 2-D (phase plane) oscillator (model.py:262-276).

    The state order is 'W V' and drift[0] belongs to W, as written in the
    reference; obsrv[0] is W, which is therefore the coupled variable."""
    state = 'W V'
    limit = (-5, 5), (-5, 5)
    input = 'c_0'
    param = 'a'
    const = {'tau': 1.0, 'I': 0.0, 'a': -2.0, 'b': -10.0, 'c': 0.0, 'd': 0.02,
             'e': 3.0, 'f': 1.0, 'g': 0.0, 'alpha': 1.0, 'beta': 1.0,
             'gamma': 1.0}
    drift = (
        'd * tau * (alpha*W - f*V**3 + e*V**2 + g*V + gamma*I + gamma*c_0)',
        'd * (a + b*V + c*V**2 - beta*W) / tau'
    )
    diffs = 1e-3, 1e-3
    obsrv = 'W', 'V'


class Generic2dOscillator(G2DO):
    """G2DO with the state order corrected to 'V W'.

    The reference's G2DO (model.py:262-276) declares state 'W V' but writes
    drift[0] as TVB's dV/dt (cubic in V, driven by alpha*W and the coupling)
    and drift[1] as dW/dt.  Integrated as declared, the cubic term sits in
    the other variable's equation and every trajectory diverges; with the
    order 'V W' the same expressions are TVB's generic 2-D oscillator.  The
    coupled observable is V, as in TVB and in
    dsl/dsl_cuda/CUDAmodels/oscillator.c:141."""
    state = 'V W'
    obsrv = 'V', 'W'


class MontbrioPazoRoxin(BaseModel):
    """Montbrio-Pazo-Roxin (2015) next-generation neural mass.

    Not present in the reference (only named in
    dsl/dsl_cuda/example/cuda_setup.py:156-157); written in the reference's
    attribute grammar.  tau*dr/dt = Delta/(pi*tau) + 2 r V,
    tau*dV/dt = V^2 + eta + J tau r + I + c - (pi tau r)^2.  ``eta`` is a
    per-node parameter (the excitability that the FC-fit gradient targets).

    ``limit`` (the box of random initial conditions) is the down-state
    basin: from r near 0 with V > 1 the V^2 term makes explicit steps of
    0.1 ms overshoot to infinity (the float64 oracle diverges within 1000
    steps for most such starts), while every start in this box stays finite
    over 20000 steps for coupling 0 .. 1."""
    state = 'r V'
    limit = (0.1, 1.0), (-1.5, -0.5)
    input = 'c'
    param = 'eta'
    const = {'tau': 1.0, 'Delta': 1.0, 'J': 15.0, 'I': 0.0, 'eta': -5.0}
    drift = (
        '(Delta / (pi * tau) + 2 * r * V) / tau',
        '(V**2 + eta + J * tau * r + I + c - (pi * tau * r)**2) / tau',
    )
    diffs = 0, 0
    obsrv = 'r', 'V'


BUILTIN_MODELS = {
    cls.__name__: cls for cls in (
        _TestModel, Kuramoto, HMJE, RWW, ReducedWongWang, JansenRit, Linear,
        G2DO, Generic2dOscillator, MontbrioPazoRoxin)
}
