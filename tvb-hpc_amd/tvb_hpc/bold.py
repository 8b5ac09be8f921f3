"""
Balloon-Windkessel BOLD model (mirror of tvb_hpc/bold.py).

Constants follow bold.py:24-57 (Stephan 2007 defaults, time in seconds).  The
device integrates one forward-Euler step of these equations per monitor
sample, driven by the time-averaged neural observable, as
dsl/dsl_cuda/CUDAmodels/refs/balloon.c:24-67 does.
"""

from .model import BaseModel


def _balloon_windkessel_constants():
    "bold.py:24-57"
    TAU_S = 0.65
    TAU_F = 0.41
    TAU_O = 0.98
    ALPHA = 0.32
    TE = 0.04
    V0 = 4.0
    E0 = 0.4
    EPSILON = 0.5
    NU_0 = 40.3
    R_0 = 25.0
    return {
        'RECIP_TAU_S': 1.0 / TAU_S,
        'RECIP_TAU_F': 1.0 / TAU_F,
        'RECIP_TAU_O': 1.0 / TAU_O,
        'RECIP_ALPHA': 1.0 / ALPHA,
        'E0': E0,
        'RECIP_E0': 1.0 / E0,
        'V0': V0,
        'k1': 4.3 * NU_0 * E0 * TE,
        'k2': EPSILON * R_0 * E0 * TE,
        'k3': 1.0 - EPSILON,
    }


class BalloonWindkessel(BaseModel):
    "bold.py:60-74"
    const = _balloon_windkessel_constants()
    state = 's f v q'
    input = 'x'
    drift = (
        'x - RECIP_TAU_S * s - RECIP_TAU_F * (f - 1)',
        's',
        'RECIP_TAU_O * (f - v**RECIP_ALPHA)',
        'RECIP_TAU_O * (f * (1 - (1 - E0)**(1 / f)) * RECIP_E0'
        ' - v**RECIP_ALPHA * (q / v))'
    )
    obsrv = ('V0 * (   k1 * (1 - q     ) '
             '       + k2 * (1 - q / v ) '
             '       + k3 * (1 -     v ) )', )
    diffs = 0, 0, 0, 0
    # resting state s=0, f=v=q=1 (derivatives and BOLD vanish there)
    limit = (0, 0), (1, 1), (1, 1), (1, 1)
