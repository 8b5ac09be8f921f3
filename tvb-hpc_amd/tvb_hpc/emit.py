"""
HIP emitter: model / coupling / scheme definitions -> gfx950 kernel source.

This replaces the reference's loopy code generation (tvb_hpc/base.py:22-64,
model.py:105-154, network.py:115-149, transforms.py:23-49) and the Mako
CUDA templates (dsl/dsl_cuda/tmpl8_CUDA.py) with one emitter that writes
device functions from the same string attributes:

* ``tvbx_dfun``  -- unpack state / input / parameters, auxex in order, then
  drift and diffs (model.py:113-140);
* ``tvbx_obsrv`` -- observables of the updated state;
* ``tvbx_pre`` / ``tvbx_post`` -- coupling pre- and post-summation
  expressions with ``mean = sum / nnode`` (network.py:115-136);
* ``tvbx_wrap`` -- the branchless limit wrap of model.py:142-154 (Kuramoto);
* ``tvbx_bold_step`` -- one Euler step of ``bold.BalloonWindkessel``.

The generic kernel body lives in ``csrc/step_template.hip``; the argument
block in ``csrc/tvbx_args.h``.  The resulting source is compiled by hipRTC at
run time or ahead of time by hipcc (``tvb_hpc.build``).

Parameter kinds.  Every name an expression uses that is not a state, input
or auxiliary variable is a parameter of one of these kinds:

    'const'  literal folded into the source (default for constants)
    'glob'   run-time scalar, p_glob[k]
    'inst'   per sweep instance, p_inst[k][p]
    'node'   per node, p_node[k][i]           (default for ``model.param``)
    'ni'     per node and instance, p_ni[k][i][p]
"""

import hashlib
import math
import os

import numpy as np

from . import expr as E
from .bold import BalloonWindkessel
from .coupling import BaseCoupling
from .model import BaseModel
from .scheme import EulerMaryuyamaStep, EulerStep

CSRC = os.path.normpath(os.path.join(os.path.dirname(__file__), '..', 'csrc'))
KINDS = ('const', 'glob', 'inst', 'node', 'ni')
# ring element types: float32, bfloat16, float16, 8-bit fixed point of a
# [0, 1] value (for couplings whose pre() is bounded)
RING_TYPES = {'f32': 0, 'bf16': 1, 'f16': 2, 'u8': 3}

# hipRTC / hipcc options shared by the run-time and ahead-of-time builds:
# no contraction (the only fused multiply-add is the explicit fmaf of the
# coupling sum, as in the CPU oracle), IEEE division / sqrt, no packed fp32.
DEVICE_FLAGS = ['-O3', '-std=c++17', '-ffp-contract=off',
                '-fhip-fp32-correctly-rounded-divide-sqrt',
                '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']
ARCH = 'gfx950'


def _read(name):
    with open(os.path.join(CSRC, name)) as fd:
        return fd.read()


class ParamTable:
    """Assigns each run-time parameter a slot in its kind's array."""

    def __init__(self):
        self.slots = {k: [] for k in KINDS if k != 'const'}

    def add(self, key, kind):
        if kind == 'const':
            return None
        lst = self.slots[kind]
        if key not in lst:
            lst.append(key)
        return lst.index(key)

    def counts(self):
        return {k: len(v) for k, v in self.slots.items()}


class KernelSpec:
    """Everything that determines the emitted kernel source.

    model, cfun, scheme: the reference-style objects.
    kinds: {name: kind} for model constants / params (keys 'm:<name>') and
        coupling params (keys 'c:<name>'); unspecified model ``param`` names
        are 'node', other model constants 'const', coupling params 'const'
        unless their value is symbolic (then 'glob'); 'sigma' is the kind of
        the EM additive noise override (None: the model's diffs).
    ring: 'coupled' keeps the coupled observables only; 'all' keeps every
        observable (what the reference's obsrv array holds).
    ring_pre: store pre(y) in the ring (only when pre ignores post_syn and
        ring == 'coupled'); None chooses automatically.
    """

    def __init__(self, model: BaseModel, cfun: BaseCoupling, scheme: EulerStep,
                 kinds=None, sigma_kind=None, ring='coupled', ring_pre=None,
                 order='cuda', ipl=1, ring_dtype='f32', adjoint=False,
                 unroll=8, min_waves=1, pipe=False, prio=True, emaj=0,
                 adj_unroll=0):
        self.model, self.cfun, self.scheme = model, cfun, scheme
        self.adjoint = bool(adjoint)
        if unroll not in (4, 8, 16, 32):
            raise ValueError('unroll must be 4, 8, 16 or 32')
        self.unroll = int(unroll)
        # gathers in flight per group of the adjoint's transposed rows
        # (0: as many as the forward's)
        if adj_unroll not in (0, 4, 8, 16, 32):
            raise ValueError('adj_unroll must be 0, 4, 8, 16 or 32')
        self.adj_unroll = int(adj_unroll)
        # waves per SIMD the blocked kernel's register budget must allow
        if min_waves not in (1, 2, 3, 4, 5, 6, 8):
            raise ValueError('min_waves must be 1 .. 8')
        self.min_waves = int(min_waves)
        # gathers of the next group issued before the current one is
        # consumed (two groups in flight)
        self.pipe = bool(pipe)
        # wave priority falling with a multi-step launch's progress
        self.prio = bool(prio)
        # blocked launches of up to `emaj` steps form their coupling sums
        # first, edges outer and steps inner (0: step by step)
        if emaj not in (0, 4, 8, 16):
            raise ValueError('emaj must be 0, 4, 8 or 16')
        self.emaj = int(emaj)
        if order not in ('cuda', 'reference'):
            raise ValueError('order must be cuda or reference')
        self.order = order
        if ipl not in (1, 2, 4):
            raise ValueError('ipl must be 1, 2 or 4')
        self.ipl = int(ipl)
        if ring_dtype not in RING_TYPES:
            raise ValueError('ring_dtype must be one of %s' % list(RING_TYPES))
        self.ring_dtype = ring_dtype
        self.S = len(model.state_sym)
        self.C = len(cfun.io)
        self.NOBS = len(model.obsrv_sym)
        if ring not in ('coupled', 'all'):
            raise ValueError(ring)
        self.ring = ring
        self.R = self.C if ring == 'coupled' else self.NOBS
        if self.C > self.NOBS:
            raise ValueError('more coupled terms than observables')
        self.em = isinstance(scheme, EulerMaryuyamaStep)
        if sigma_kind is None and self.em and scheme.sigma is not None:
            sigma_kind = 'const'
        self.sigma_kind = sigma_kind if self.em else None
        uses_post = cfun.pre_uses_post_syn()
        if ring_pre is None:
            ring_pre = (not uses_post) and ring == 'coupled'
        if ring_pre and (uses_post or ring != 'coupled'):
            raise ValueError('ring_pre needs a post_syn-free pre and '
                             "ring='coupled'")
        self.ring_pre = bool(ring_pre)
        self.uses_post = uses_post and not self.ring_pre
        if self.emaj and (self.uses_post or order != 'cuda'):
            raise ValueError('emaj needs the CUDA ordering and a pre() that '
                             'does not read post_syn')
        if ring_dtype == 'u8' and not (self.ring_pre and
                                       cfun.pre_bounds() is not None):
            raise ValueError('an 8-bit history needs a ring holding a '
                             'bounded pre(y) (e.g. Sigmoidal coupling)')
        self.kinds = {}
        kinds = dict(kinds or {})
        mstate, minput, mparam = model.names
        for name, val in model.const.items():
            kd = kinds.pop('m:' + name, None) or kinds.pop(name, None)
            if kd is None:
                if isinstance(val, E.Expr):
                    kd = 'glob'
                else:
                    kd = 'node' if name in mparam else 'const'
            self.kinds['m:' + name] = kd
        for name in mparam:
            if 'm:' + name not in self.kinds:
                self.kinds['m:' + name] = kinds.pop('m:' + name, 'node')
        for name, val in cfun.param.items():
            kd = kinds.pop('c:' + name, None)
            if kd is None:
                kd = 'glob' if isinstance(val, E.Expr) else 'const'
            if kd in ('node', 'ni'):
                raise ValueError('coupling parameters are per instance at '
                                 'most (%s)' % name)
            self.kinds['c:' + name] = kd
        if kinds:
            raise ValueError('unknown parameters %s' % sorted(kinds))
        for k, v in self.kinds.items():
            if v not in KINDS:
                raise ValueError('bad kind %r for %s' % (v, k))
        self.table = ParamTable()
        for key in sorted(self.kinds):
            self.table.add(key, self.kinds[key])
        if self.sigma_kind is not None:
            if self.sigma_kind in ('node', 'ni'):
                raise ValueError('sigma is per instance at most')
            self.table.add('sigma', self.sigma_kind)

    # -- values of constants -------------------------------------------------

    def const_value(self, key):
        if key == 'sigma':
            return float(self.scheme.sigma)
        scope, name = key.split(':', 1)
        src = self.model.const if scope == 'm' else self.cfun.param
        v = src[name]
        if isinstance(v, E.Expr):
            raise ValueError('%s is symbolic; give it a run-time kind' % key)
        return float(v)

    def slot(self, key):
        kind = self.kinds[key] if key != 'sigma' else self.sigma_kind
        return kind, self.table.add(key, kind)

    # -- C name mapping --------------------------------------------------------

    def _lane_fields(self):
        out = []
        for key in sorted(self.kinds) + (['sigma'] if self.sigma_kind else []):
            kind, idx = self.slot(key)
            if kind in ('inst', 'glob'):
                out.append((key, kind, idx))
        return out

    @staticmethod
    def cname(key):
        return 'v_' + key.replace(':', '_')

    def _param_expr(self, key, in_node_scope):
        kind, idx = self.slot(key)
        if kind == 'const':
            return E.c_literal(self.const_value(key))
        if kind in ('inst', 'glob'):
            return 'lp.' + self.cname(key)
        if not in_node_scope:
            raise ValueError('%s (kind %s) used outside node scope'
                             % (key, kind))
        if kind == 'node':
            return 'a.p_node[%d * a.n_node + i]' % idx
        return 'a.p_ni[((size_t)%d * a.n_node + i) * a.pitch + p]' % idx

    # -- emission ----------------------------------------------------------------

    def _model_names(self, in_node_scope=True):
        mstate, minput, mparam = self.model.names
        names = {}
        for name in self.model.const:
            names[name] = self._param_expr('m:' + name, in_node_scope)
        for name in mparam:
            names[name] = self._param_expr('m:' + name, in_node_scope)
        for s, name in enumerate(mstate):
            names[name] = 'x[%d]' % s
        for k, name in enumerate(minput):
            names[name] = 'c[%d]' % k
        for lhs, _ in self.model.auxex:
            names[lhs] = 'aux_' + lhs
        return names

    def _cfun_names(self):
        names = {}
        for name in self.cfun.param:
            names[name] = self._param_expr('c:' + name, False)
        return names

    def _check_names(self, e, names, what):
        missing = E.free_vars(e) - set(names) - {'pi'}
        if missing:
            raise ValueError('%s uses undefined names %s' % (what,
                                                              sorted(missing)))

    def emit_model(self):
        lines = []
        S, C, NOBS, R = self.S, self.C, self.NOBS, self.R
        lines.append('#define TVBX_S %d' % S)
        lines.append('#define TVBX_C %d' % C)
        lines.append('#define TVBX_NOBS %d' % NOBS)
        lines.append('#define TVBX_R %d' % R)
        lines.append('#define TVBX_EM %d' % int(self.em))
        lines.append('#define TVBX_RING_PRE %d' % int(self.ring_pre))
        lines.append('#define TVBX_USES_POST %d' % int(self.uses_post))
        lines.append('#define TVBX_ORDER_REF %d' % int(self.order ==
                                                     'reference'))
        lines.append('#define TVBX_IPL %d' % self.ipl)
        lines.append('#define TVBX_RING_T %d' % RING_TYPES[self.ring_dtype])
        lines.append('#define TVBX_UNROLL %d' % self.unroll)
        lines.append('#define TVBX_MIN_WAVES %d' % self.min_waves)
        if self.pipe:
            lines.append('#define TVBX_PIPE 1')
        lines.append('#define TVBX_PRIO %d' % int(self.prio))
        lines.append('#define TVBX_EMAJ %d' % self.emaj)
        # lane parameters
        fields = self._lane_fields()
        lines.append('struct tvbx_lane_params {')
        for key, kind, idx in fields:
            lines.append('  float %s;' % self.cname(key))
        if not fields:
            lines.append('  int unused;')
        lines.append('};')
        lines.append('__device__ __forceinline__ void tvbx_load_lane('
                     'tvbx_lane_params &lp, const tvbx_step_args &a, int p) {')
        for key, kind, idx in fields:
            src = ('a.p_inst[(size_t)%d * a.param_pitch + a.inst_off + p]'
                   % idx if kind == 'inst'
                   else 'a.p_glob[%d]' % idx)
            lines.append('  lp.%s = %s;' % (self.cname(key), src))
        lines.append('}')
        # coupling: pre / post
        cn = self._cfun_names()
        pre_c, post_c, ringv, pre_nb = [], [], [], []
        for k in range(C):
            pre = self.cfun.pre_expr(k)
            self._check_names(pre, dict(cn, pre_syn=1, post_syn=1),
                              'pre_sum[%d]' % k)
            names = dict(cn, pre_syn='pre_syn', post_syn='post_syn')
            pre_src = E.to_c(pre, names)
            post = self.cfun.post_expr(k)
            post = E.subst(post, {'mean': E.Bin('/', E.Var('sum'),
                                                E.Var('nnode'))})
            self._check_names(post, dict(cn, sum=1, nnode=1),
                              'post_sum[%d]' % k)
            post_src = E.to_c(post, dict(cn, sum='sum',
                                         nnode='(float)a.n_node'))
            post_c.append(post_src)
            if self.ring_pre:
                pre_c.append('pre_syn')
                ringv.append(E.to_c(pre, dict(cn, pre_syn='y',
                                              post_syn='0.0f')))
            else:
                pre_c.append(pre_src)
                pre_nb.append(E.to_c(pre, names, unchecked_trig=True))
                ringv.append('y')
        for r in range(C, R):
            ringv.append('y')
        sig = ('int kk, float pre_syn, float post_syn, '
               'const tvbx_lane_params &lp, const tvbx_step_args &a')
        lines.append('__device__ __forceinline__ float tvbx_pre(%s) {' % sig)
        for k in range(C):
            lines.append('  if (kk == %d) return %s;' % (k, pre_c[k]))
        lines.append('  return 0.0f;')
        lines.append('}')
        # with sines or cosines in pre: a branch-free version whose gather
        # groups are checked once (NaN outside the reduced range)
        nb = bool(pre_nb) and pre_nb != pre_c[:len(pre_nb)]
        lines.append('#define TVBX_PRE_NB %d' % int(nb))
        if nb:
            lines.append('__device__ __forceinline__ float tvbx_pre_nb(%s) {'
                         % sig)
            for k in range(C):
                lines.append('  if (kk == %d) return %s;' % (k, pre_nb[k]))
            lines.append('  return 0.0f;')
            lines.append('}')
        lines.append('__device__ __forceinline__ float tvbx_post(int kk, '
                     'float sum, const tvbx_lane_params &lp, '
                     'const tvbx_step_args &a) {')
        for k in range(C):
            lines.append('  if (kk == %d) return %s;' % (k, post_c[k]))
        lines.append('  return 0.0f;')
        lines.append('}')
        lines.append('__device__ __forceinline__ float tvbx_ring_value(int r, '
                     'float y, const tvbx_lane_params &lp, '
                     'const tvbx_step_args &a) {')
        for r in range(R):
            lines.append('  if (r == %d) return %s;' % (r, ringv[r]))
        lines.append('  return y;')
        lines.append('}')
        # dfun
        mn = self._model_names(True)
        lines.append('__device__ __forceinline__ void tvbx_dfun(const float *x, '
                     'const float *c, float *f, float *g, '
                     'const tvbx_lane_params &lp, const tvbx_step_args &a, '
                     'int i, int p) {')
        known = dict(mn)
        for lhs, rhs in self.model.auxex_exprs():
            self._check_names(rhs, known, 'auxex %s' % lhs)
            lines.append('  const float aux_%s = %s;' % (lhs, E.to_c(rhs, mn)))
        for s, e in enumerate(self.model.drift_exprs()):
            self._check_names(e, known, 'drift[%d]' % s)
            lines.append('  f[%d] = %s;' % (s, E.to_c(e, mn)))
        if self.sigma_kind is not None:
            sg = self._param_expr('sigma', True)
            for s in range(S):
                lines.append('  g[%d] = %s;' % (s, sg))
        else:
            for s, e in enumerate(self.model.diffs_exprs()):
                self._check_names(e, known, 'diffs[%d]' % s)
                lines.append('  g[%d] = %s;' % (s, E.to_c(e, mn)))
        lines.append('}')
        # obsrv
        lines.append('__device__ __forceinline__ void tvbx_obsrv(const float *x, '
                     'float *y, const tvbx_lane_params &lp, '
                     'const tvbx_step_args &a, int i, int p) {')
        on = {k: v for k, v in mn.items() if not k.startswith('aux')}
        for lhs, rhs in self.model.auxex_exprs():
            on.pop(lhs, None)
        for k, e in enumerate(self.model.obsrv_exprs()):
            self._check_names(e, on, 'obsrv[%d]' % k)
            lines.append('  y[%d] = %s;' % (k, E.to_c(e, on)))
        lines.append('}')
        # wrap
        lines.append('__device__ __forceinline__ void tvbx_wrap(float *x) {')
        for idx, lo, hi in self.model.wrap_ranges():
            lo_c, hi_c = E.c_literal(lo), E.c_literal(hi)
            dx_c = E.c_literal(hi - lo)
            lines.append('  { const float v = x[%d]; x[%d] = (v < %s) ? v + %s : '
                         '((v > %s) ? v - %s : v); }'
                         % (idx, idx, lo_c, dx_c, hi_c, dx_c))
        lines.append('}')
        lines.extend(self.emit_bold())
        return '\n'.join(lines) + '\n'

    def emit_bold(self):
        "One Euler step of bold.BalloonWindkessel (bold.py:60-74)."
        bw = BalloonWindkessel()
        names = {k: E.c_literal(v) for k, v in bw.const.items()}
        sv = [v.name for v in bw.state_sym]
        for s in sv:
            names[s] = 'b_' + s
        names['x'] = 'xin'
        out = ['__device__ __forceinline__ float tvbx_bold_step('
               'const tvbx_step_args &a, size_t ip, float xin) {',
               '  const size_t NP = (size_t)a.n_node * a.pitch;']
        for k, s in enumerate(sv):
            out.append('  float b_%s = a.bold_state[%d * NP + ip];' % (s, k))
        for k, e in enumerate(bw.drift_exprs()):
            out.append('  const float d_%d = %s;' % (k, E.to_c(e, names)))
        for k, s in enumerate(sv):
            out.append('  b_%s = b_%s + a.bold_dt * d_%d;' % (s, s, k))
            out.append('  a.bold_state[%d * NP + ip] = b_%s;' % (k, s))
        out.append('  return %s;' % E.to_c(bw.obsrv_exprs()[0], names))
        out.append('}')
        return out

    # -- adjoint ---------------------------------------------------------------

    def grad_params(self):
        """Parameters the adjoint differentiates: coupling parameters of kind
        'inst' (e.g. the global coupling G) and model parameters of kind
        'inst' or 'ni' (e.g. a per-node excitability)."""
        gc = [k for k in sorted(self.kinds) if k.startswith('c:')
              and self.kinds[k] == 'inst']
        gm = [k for k in sorted(self.kinds) if k.startswith('m:')
              and self.kinds[k] in ('inst', 'ni')]
        return gc, gm

    def _drift_inlined(self):
        "drift expressions with the auxiliary variables substituted"
        aux = {}
        for lhs, rhs in self.model.auxex_exprs():
            aux[lhs] = E.subst(rhs, aux)
        return [E.subst(e, aux) for e in self.model.drift_exprs()]

    def check_adjoint(self):
        """The hand-written adjoint covers: one coupled observable, a
        post-summation linear in the sum, pre free of post_syn, additive
        noise (diffs independent of the state), the dsl_cuda step order and
        a float32 history."""
        mstate, minput, _ = self.model.names
        if self.C != 1:
            raise ValueError('adjoint: one coupled observable only')
        if self.order != 'cuda':
            raise ValueError("adjoint: order='cuda' only")
        if self.ring_dtype != 'f32':
            raise ValueError('adjoint: float32 history only')
        if self.cfun.pre_uses_post_syn():
            raise ValueError('adjoint: pre_sum must not use post_syn')
        post = E.subst(self.cfun.post_expr(0), {'mean': E.Bin(
            '/', E.Var('sum'), E.Var('nnode'))})
        if 'sum' in E.free_vars(E.diff(post, 'sum')):
            raise ValueError('adjoint: post_sum must be linear in the sum')
        if self.sigma_kind is None:
            for g in self.model.diffs_exprs():
                if E.free_vars(g) & set(mstate):
                    raise ValueError('adjoint: multiplicative noise')

    def emit_adjoint(self):
        self.check_adjoint()
        mstate, minput, _ = self.model.names
        S = self.S
        mn = self._model_names(True)
        cn = self._cfun_names()
        gc, gm = self.grad_params()
        drift = self._drift_inlined()
        lines = ['#define TVBX_ADJ_UNROLL %d' % (self.adj_unroll or
                                                 self.unroll),
                 '#define TVBX_NGC %d' % max(len(gc), 1),
                 '#define TVBX_NGM %d' % max(len(gm), 1),
                 '#define TVBX_HAS_NGC %d' % len(gc),
                 '#define TVBX_HAS_NGM %d' % len(gm)]

        def c_of(e, names):
            return E.to_c(E.simplify(e), names)

        # vector-Jacobian products of the drift
        lines.append('__device__ __forceinline__ void tvbx_dfun_vjp('
                     'const float *x, const float *c, const float *lam, '
                     'float *jt, float *mu, float *gm, '
                     'const tvbx_lane_params &lp, const tvbx_step_args &a, '
                     'int i, int p) {')
        for r, xr in enumerate(mstate):
            terms = []
            for s in range(S):
                d = E.diff(drift[s], xr)
                if not (isinstance(d, E.Num) and float(d.value) == 0):
                    terms.append('lam[%d] * %s' % (s, c_of(d, mn)))
            lines.append('  jt[%d] = %s;' % (r, ' + '.join(terms) or '0.0f'))
        for k, ck in enumerate(minput[:self.C]):
            terms = []
            for s in range(S):
                d = E.diff(drift[s], ck)
                if not (isinstance(d, E.Num) and float(d.value) == 0):
                    terms.append('lam[%d] * %s' % (s, c_of(d, mn)))
            lines.append('  mu[%d] = %s;' % (k, ' + '.join(terms) or '0.0f'))
        for q, key in enumerate(gm):
            name = key.split(':', 1)[1]
            terms = []
            for s in range(S):
                d = E.diff(drift[s], name)
                if not (isinstance(d, E.Num) and float(d.value) == 0):
                    terms.append('lam[%d] * %s' % (s, c_of(d, mn)))
            lines.append('  gm[%d] = %s;' % (q, ' + '.join(terms) or '0.0f'))
        if not gm:
            lines.append('  gm[0] = 0.0f;')
        lines.append('}')
        # d h_k / d x
        on = {k: v for k, v in mn.items() if not k.startswith('aux')}
        for lhs, _ in self.model.auxex_exprs():
            on.pop(lhs, None)
        lines.append('__device__ __forceinline__ void tvbx_dobsrv(int k, '
                     'const float *x, float *dh, const tvbx_lane_params &lp, '
                     'const tvbx_step_args &a, int i, int p) {')
        for k, h in enumerate(self.model.obsrv_exprs()):
            lines.append('  if (k == %d) {' % k)
            for r, xr in enumerate(mstate):
                lines.append('    dh[%d] = %s;' % (r, c_of(E.diff(h, xr), on)))
            lines.append('    return;')
            lines.append('  }')
        lines.append('}')
        # pre' and the post derivatives
        pre = self.cfun.pre_expr(0)
        lines.append('__device__ __forceinline__ float tvbx_dpre(float y, '
                     'const tvbx_lane_params &lp, const tvbx_step_args &a) {')
        lines.append('  return %s;' % c_of(E.diff(pre, 'pre_syn'),
                                           dict(cn, pre_syn='y')))
        lines.append('}')
        post = E.subst(self.cfun.post_expr(0), {'mean': E.Bin(
            '/', E.Var('sum'), E.Var('nnode'))})
        pn = dict(cn, sum='sum', nnode='(float)a.n_node')
        lines.append('__device__ __forceinline__ float tvbx_dpost_dsum('
                     'const tvbx_lane_params &lp, const tvbx_step_args &a) {')
        lines.append('  const float sum = 0.0f; (void)sum;')
        lines.append('  return %s;' % c_of(E.diff(post, 'sum'), pn))
        lines.append('}')
        lines.append('__device__ __forceinline__ void tvbx_dpost_dparam('
                     'float sum, float *g, const tvbx_lane_params &lp, '
                     'const tvbx_step_args &a) {')
        for q, key in enumerate(gc):
            name = key.split(':', 1)[1]
            lines.append('  g[%d] = %s;' % (q, c_of(E.diff(post, name), pn)))
        if not gc:
            lines.append('  g[0] = 0.0f;')
        lines.append('}')
        return '\n'.join(lines) + '\n'

    def source(self):
        head = ('// generated by tvb_hpc.emit for %s / %s / %s\n'
                % (type(self.model).__name__, type(self.cfun).__name__,
                   type(self.scheme).__name__))
        src = (head + _read('tvbx_args.h') + '\n' + _read('tvbx_philox.h')
               + '\n' + _read('tvbx_math.h')
               + '\n' + self.emit_model() + '\n'
               + _read('step_template.hip'))
        if self.adjoint:
            src += '\n' + self.emit_adjoint() + '\n' + \
                _read('adjoint_template.hip')
        return src

    def key(self):
        h = hashlib.sha256(self.source().encode())
        h.update(' '.join(DEVICE_FLAGS + [ARCH]).encode())
        return h.hexdigest()[:24]

    def param_layout(self):
        """[(key, kind, index)] for every run-time parameter."""
        out = []
        for key in sorted(self.kinds) + (['sigma'] if self.sigma_kind else []):
            kind, idx = self.slot(key)
            if kind != 'const':
                out.append((key, kind, idx))
        return out

    def counts(self):
        return self.table.counts()
