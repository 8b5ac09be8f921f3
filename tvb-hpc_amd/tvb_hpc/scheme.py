"""
Integration schemes (mirror of tvb_hpc/scheme.py).

Symbolic one-step constructors (``euler``, ``euler_maruyama``,
``euler_maruyama_logp``; scheme.py:29-48) and the step descriptors
``EulerStep`` / ``EulerMaryuyamaStep`` (scheme.py:55-107).  The step the
device kernel performs is, per state variable s,

    x(n+1) = x(n) + dt*f_s + sqrt_dt*dWt*g_s        (EM, scheme.py:88-97)
    x(n+1) = x(n) + dt*f_s                          (Euler, scheme.py:82-85)

evaluated left to right, with ``sqrt_dt = sqrtf(dt)`` computed in float32.
"""

from typing import Union

import numpy as np

from . import expr as E


def euler(x, f, dt=None):
    "Construct standard Euler step (scheme.py:29-32)."
    dt = E.as_expr(dt) if dt is not None else E.Var('dt')
    return x + dt * f


def euler_maruyama(x, f, g, dt=None, dWt=None):
    "Construct SDE Euler-Maruyama step (scheme.py:35-39)."
    dt = E.as_expr(dt) if dt is not None else E.Var('dt')
    dWt = E.as_expr(dWt) if dWt is not None else E.Var('dWt')
    return x + dt * f + dWt * dt ** 0.5 * g


def euler_maruyama_logp(x, f, g, xn=None, dt=None, step=euler):
    "Construct normal log p of an EM transition (scheme.py:42-48)."
    dt = E.as_expr(dt) if dt is not None else E.Var('dt')
    xs = [E.as_expr(v) for v in np.atleast_1d(x)]
    fs = [E.as_expr(v) for v in np.atleast_1d(f)]
    gs = [E.as_expr(v) for v in np.atleast_1d(g)]
    if xn is None:
        xn = [E.Var(v.name + '_n') for v in xs]
    out = []
    for xi, fi, gi, xni in zip(xs, fs, gs, xn):
        mu = step(xi, fi, dt)
        sd = dt ** 0.5 * gi
        out.append(-(E.as_expr(xni) - mu) ** 2 / (2 * sd ** 2))
    arr = np.empty(len(out), dtype=object)
    arr[:] = out
    return arr


class TimeStepScheme:
    stochastic = False


class EulerStep(TimeStepScheme):
    "Deterministic Euler step (scheme.py:55-85)."

    def __init__(self, dt: Union[E.Var, float]):
        self.dt = dt

    def dt_value(self, override=None):
        if override is not None:
            return float(override)
        if isinstance(self.dt, E.Expr):
            raise ValueError('dt is symbolic; pass dt= at call time')
        return float(self.dt)

    def __repr__(self):
        return '%s(%r)' % (type(self).__name__, self.dt)


class EulerMaryuyamaStep(EulerStep):
    """Euler-Maruyama step (scheme.py:88-107).

    ``sigma``, when given, replaces every ``diffs`` expression of the model by
    an additive noise amplitude (a number, or a per-instance array supplied
    at call time)."""
    stochastic = True

    def __init__(self, dt: Union[E.Var, float], sigma=None):
        super().__init__(dt)
        self.sigma = sigma


# spelling used by some callers
EulerMaruyamaStep = EulerMaryuyamaStep
