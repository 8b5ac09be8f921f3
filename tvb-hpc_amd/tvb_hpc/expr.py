"""
Small symbolic expression core.

The reference builds every kernel from pymbolic expressions parsed out of the
strings on model / coupling classes (tvb_hpc/utils.py:67-90, tvb_hpc/base.py:22-64).
pymbolic is not a dependency here: this module parses the same expression
grammar with Python's ``ast`` module into a tiny immutable tree that supports

* substitution of variables by expressions (utils.py:110-125, ``VarSubst``),
* symbolic differentiation (model.py:61-65 uses pymbolic's DifferentiationMapper),
* emission as C / HIP source with a fixed operation order (the code generator
  for the device step kernel and nothing else),
* vectorised evaluation with NumPy (host-side helpers).

Operation order is part of the contract: ``a + b + c`` is ``(a + b) + c`` and
integer powers ``x**2`` / ``x**3`` are emitted as ``x*x`` / ``x*x*x`` so that
a float32 evaluation is reproducible from the expression text alone.
"""

import ast
import math
from typing import Dict, Iterable, Union

import numpy as np

__all__ = [
    'Expr', 'Num', 'Var', 'Bin', 'Neg', 'Call', 'Cmp', 'Sub',
    'parse', 'var', 'subst', 'diff', 'free_vars', 'to_c', 'evaluate',
    'simplify',
]

_FUNCS = ('sin', 'cos', 'tan', 'exp', 'expm1', 'log', 'sqrt', 'tanh', 'abs',
          'fabs', 'pow', 'sinh', 'cosh', 'atan', 'asin', 'acos',
          'sqrtf', 'expf', 'sinf', 'cosf', 'logf')


class Expr:
    """Immutable expression node; arithmetic operators build new nodes."""

    __slots__ = ()
    prec = 100

    def __add__(self, o):
        return Bin('+', self, as_expr(o))

    def __radd__(self, o):
        return Bin('+', as_expr(o), self)

    def __sub__(self, o):
        return Bin('-', self, as_expr(o))

    def __rsub__(self, o):
        return Bin('-', as_expr(o), self)

    def __mul__(self, o):
        return Bin('*', self, as_expr(o))

    def __rmul__(self, o):
        return Bin('*', as_expr(o), self)

    def __truediv__(self, o):
        return Bin('/', self, as_expr(o))

    def __rtruediv__(self, o):
        return Bin('/', as_expr(o), self)

    def __pow__(self, o):
        return Bin('**', self, as_expr(o))

    def __rpow__(self, o):
        return Bin('**', as_expr(o), self)

    def __neg__(self):
        return Neg(self)

    def __repr__(self):
        return '%s(%r)' % (type(self).__name__, str(self))

    def __hash__(self):
        return hash(self._key())

    def __eq__(self, other):
        return isinstance(other, Expr) and self._key() == other._key()

    def _key(self):
        raise NotImplementedError


class Num(Expr):
    __slots__ = ('value',)

    def __init__(self, value):
        if isinstance(value, bool):
            value = int(value)
        self.value = value

    def _key(self):
        return ('num', float(self.value))

    def __str__(self):
        v = self.value
        if isinstance(v, float) and v.is_integer() and abs(v) < 1e15:
            return repr(v)
        return repr(v)


class Var(Expr):
    __slots__ = ('name',)

    def __init__(self, name: str):
        self.name = name

    def _key(self):
        return ('var', self.name)

    def __str__(self):
        return self.name


class Bin(Expr):
    __slots__ = ('op', 'a', 'b')
    _prec = {'+': 10, '-': 10, '*': 20, '/': 20, '**': 30}

    def __init__(self, op, a, b):
        self.op, self.a, self.b = op, a, b

    @property
    def prec(self):
        return self._prec[self.op]

    def _key(self):
        return ('bin', self.op, self.a._key(), self.b._key())

    def __str__(self):
        # pymbolic-style spacing: 'a + b', 'a*b', 'a / b', 'a**b'
        p = self.prec
        la = _paren(self.a, p, right=False, op=self.op)
        rb = _paren(self.b, p, right=True, op=self.op)
        sep = {'+': ' + ', '-': ' - ', '*': '*', '/': ' / ', '**': '**'}[self.op]
        return la + sep + rb


class Neg(Expr):
    __slots__ = ('a',)
    prec = 25

    def __init__(self, a):
        self.a = a

    def _key(self):
        return ('neg', self.a._key())

    def __str__(self):
        return '-' + _paren(self.a, self.prec, right=True, op='neg')


class Call(Expr):
    __slots__ = ('fn', 'args')

    def __init__(self, fn, args):
        self.fn, self.args = fn, tuple(args)

    def _key(self):
        return ('call', self.fn, tuple(a._key() for a in self.args))

    def __str__(self):
        return '%s(%s)' % (self.fn, ', '.join(str(a) for a in self.args))


class Cmp(Expr):
    """Comparison evaluating to 1.0 (true) or 0.0 (false)."""
    __slots__ = ('op', 'a', 'b')
    prec = 5

    def __init__(self, op, a, b):
        self.op, self.a, self.b = op, a, b

    def _key(self):
        return ('cmp', self.op, self.a._key(), self.b._key())

    def __str__(self):
        return '%s %s %s' % (_paren(self.a, 6, False, 'cmp'), self.op,
                             _paren(self.b, 6, True, 'cmp'))


class Sub(Expr):
    """Subscript ``name[i, j]``; only used for textual substitution."""
    __slots__ = ('base', 'index')

    def __init__(self, base, index):
        self.base, self.index = base, tuple(index)

    def _key(self):
        return ('sub', self.base._key(), tuple(i._key() for i in self.index))

    def __str__(self):
        return '%s[%s]' % (self.base, ', '.join(str(i) for i in self.index))


def _paren(e, p, right, op):
    s = str(e)
    ep = e.prec
    need = ep < p
    if not need and ep == p and right and op in ('-', '/', '**'):
        need = True
    if not need and ep == p and not right and op == '**':
        need = True
    if isinstance(e, Num) and isinstance(e.value, (int, float)) and e.value < 0:
        need = True
    return '(' + s + ')' if need else s


def as_expr(x) -> Expr:
    if isinstance(x, Expr):
        return x
    if isinstance(x, (bool, int, float, np.integer, np.floating)):
        return Num(x.item() if hasattr(x, 'item') else x)
    if isinstance(x, str):
        return parse(x)
    raise TypeError('cannot convert %r to an expression' % (x, ))


_BINOPS = {ast.Add: '+', ast.Sub: '-', ast.Mult: '*', ast.Div: '/',
           ast.Pow: '**'}
_CMPOPS = {ast.Lt: '<', ast.Gt: '>', ast.LtE: '<=', ast.GtE: '>=',
           ast.Eq: '==', ast.NotEq: '!='}


def _from_ast(node) -> Expr:
    if isinstance(node, ast.Expression):
        return _from_ast(node.body)
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
        return Num(node.value)
    if isinstance(node, ast.Name):
        return Var(node.id)
    if isinstance(node, ast.BinOp) and type(node.op) in _BINOPS:
        return Bin(_BINOPS[type(node.op)], _from_ast(node.left),
                   _from_ast(node.right))
    if isinstance(node, ast.UnaryOp):
        if isinstance(node.op, ast.USub):
            inner = _from_ast(node.operand)
            if isinstance(inner, Num):
                return Num(-inner.value)
            return Neg(inner)
        if isinstance(node.op, ast.UAdd):
            return _from_ast(node.operand)
    if isinstance(node, ast.Compare) and len(node.ops) == 1:
        return Cmp(_CMPOPS[type(node.ops[0])], _from_ast(node.left),
                   _from_ast(node.comparators[0]))
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Name):
        return Call(node.func.id, [_from_ast(a) for a in node.args])
    if isinstance(node, ast.Subscript):
        sl = node.slice
        if isinstance(sl, ast.Index):  # pragma: no cover (py<3.9)
            sl = sl.value
        idx = sl.elts if isinstance(sl, ast.Tuple) else [sl]
        return Sub(_from_ast(node.value), [_from_ast(i) for i in idx])
    raise ValueError('unsupported expression element %s' % ast.dump(node))


def parse(text) -> Expr:
    """Parse an expression string (pymbolic grammar subset) or number.

    Raises ``Exception(repr(text))`` for unparsable input, like
    tvb_hpc/utils.py:86-89."""
    if isinstance(text, Expr):
        return text
    if not isinstance(text, str):
        return as_expr(text)
    try:
        return _from_ast(ast.parse(text.strip(), mode='eval'))
    except Exception:
        raise Exception(repr(text))


def var(name: str) -> Var:
    return Var(name)


def free_vars(e: Expr) -> set:
    out = set()

    def walk(x):
        if isinstance(x, Var):
            out.add(x.name)
        elif isinstance(x, Bin) or isinstance(x, Cmp):
            walk(x.a)
            walk(x.b)
        elif isinstance(x, Neg):
            walk(x.a)
        elif isinstance(x, Call):
            for a in x.args:
                walk(a)
        elif isinstance(x, Sub):
            walk(x.base)
            for a in x.index:
                walk(a)
    walk(e)
    return out


def subst(e: Expr, mapping: Dict[str, object]) -> Expr:
    """Replace variables by name (tvb_hpc/utils.py:110-125)."""
    m = {k: as_expr(v) for k, v in mapping.items()}

    def go(x):
        if isinstance(x, Var):
            return m.get(x.name, x)
        if isinstance(x, Bin):
            return Bin(x.op, go(x.a), go(x.b))
        if isinstance(x, Cmp):
            return Cmp(x.op, go(x.a), go(x.b))
        if isinstance(x, Neg):
            return Neg(go(x.a))
        if isinstance(x, Call):
            return Call(x.fn, [go(a) for a in x.args])
        if isinstance(x, Sub):
            return Sub(go(x.base), [go(a) for a in x.index])
        return x
    return go(as_expr(e))


# ---------------------------------------------------------------------------
# simplification (only the identities needed to keep derivatives small)

def _isnum(x, v=None):
    return isinstance(x, Num) and (v is None or float(x.value) == v)


def simplify(e: Expr) -> Expr:
    e = as_expr(e)
    if isinstance(e, Bin):
        a, b = simplify(e.a), simplify(e.b)
        if _isnum(a) and _isnum(b) and e.op != '**':
            return Num(_fold(e.op, float(a.value), float(b.value)))
        if e.op == '+':
            if _isnum(a, 0):
                return b
            if _isnum(b, 0):
                return a
        elif e.op == '-':
            if _isnum(b, 0):
                return a
            if _isnum(a, 0):
                return simplify(Neg(b))
        elif e.op == '*':
            if _isnum(a, 0) or _isnum(b, 0):
                return Num(0)
            if _isnum(a, 1):
                return b
            if _isnum(b, 1):
                return a
        elif e.op == '/':
            if _isnum(a, 0):
                return Num(0)
            if _isnum(b, 1):
                return a
        elif e.op == '**':
            if _isnum(b, 1):
                return a
            if _isnum(b, 0):
                return Num(1)
        return Bin(e.op, a, b)
    if isinstance(e, Neg):
        a = simplify(e.a)
        if _isnum(a):
            return Num(-a.value)
        if isinstance(a, Neg):
            return a.a
        return Neg(a)
    if isinstance(e, Call):
        return Call(e.fn, [simplify(x) for x in e.args])
    if isinstance(e, Cmp):
        return Cmp(e.op, simplify(e.a), simplify(e.b))
    return e


def _fold(op, a, b):
    return {'+': a + b, '-': a - b, '*': a * b, '/': a / b if b else
            math.inf}[op]


# ---------------------------------------------------------------------------
# differentiation

def diff(e: Expr, name: str) -> Expr:
    """d e / d name, simplified (model.py:61-65 role)."""
    return simplify(_d(as_expr(e), name))


def _d(e, n):
    if isinstance(e, Num):
        return Num(0)
    if isinstance(e, Var):
        return Num(1 if e.name == n else 0)
    if isinstance(e, Cmp):
        return Num(0)
    if isinstance(e, Neg):
        return Neg(_d(e.a, n))
    if isinstance(e, Bin):
        a, b = e.a, e.b
        if e.op in '+-':
            return Bin(e.op, _d(a, n), _d(b, n))
        if e.op == '*':
            return Bin('+', Bin('*', _d(a, n), b), Bin('*', a, _d(b, n)))
        if e.op == '/':
            # (a'b - ab') / b**2
            return Bin('/', Bin('-', Bin('*', _d(a, n), b),
                                Bin('*', a, _d(b, n))), Bin('*', b, b))
        if e.op == '**':
            if n not in free_vars(b):
                # a**k -> k * a**(k-1) * a'
                if _isnum(b):
                    k = b.value
                    km1 = Num(k - 1)
                    pw = a if k - 1 == 1 else Bin('**', a, km1)
                    return Bin('*', Bin('*', Num(k), pw), _d(a, n))
                return Bin('*', Bin('*', b, Bin('**', a, Bin('-', b, Num(1)))),
                           _d(a, n))
            # general: a**b * (b' log a + b a'/a)
            return Bin('*', e, Bin('+', Bin('*', _d(b, n), Call('log', [a])),
                                   Bin('/', Bin('*', b, _d(a, n)), a)))
    if isinstance(e, Call):
        f = e.fn
        if len(e.args) == 1:
            a = e.args[0]
            da = _d(a, n)
            if f in ('sin', 'sinf'):
                g = Call('cos', [a])
            elif f in ('cos', 'cosf'):
                g = Neg(Call('sin', [a]))
            elif f in ('exp', 'expf'):
                g = e
            elif f == 'expm1':
                g = Call('exp', [a])
            elif f in ('log', 'logf'):
                g = Bin('/', Num(1), a)
            elif f in ('sqrt', 'sqrtf'):
                g = Bin('/', Num(0.5), e)
            elif f == 'tanh':
                g = Bin('-', Num(1), Bin('*', e, e))
            elif f in ('abs', 'fabs'):
                g = Bin('-', Bin('*', Num(2), Cmp('>=', a, Num(0))), Num(1))
            else:
                raise ValueError('no derivative rule for %s' % f)
            return Bin('*', g, da)
        if f == 'pow' and len(e.args) == 2:
            return _d(Bin('**', e.args[0], e.args[1]), n)
    raise ValueError('cannot differentiate %s' % e)


# ---------------------------------------------------------------------------
# C / HIP emission

# sin and cos: tvbx_math.h (a bounded-argument path before the library's)
_C_FUN32 = {'sin': 'tvbx_sinf', 'cos': 'tvbx_cosf', 'tan': 'tanf',
            'exp': 'expf',
            'expm1': 'expm1f',
            'log': 'logf', 'sqrt': 'sqrtf', 'tanh': 'tanhf', 'abs': 'fabsf',
            'fabs': 'fabsf', 'pow': 'powf', 'sinh': 'sinhf', 'cosh': 'coshf',
            'atan': 'atanf', 'asin': 'asinf', 'acos': 'acosf',
            'sinf': 'tvbx_sinf', 'cosf': 'tvbx_cosf', 'expf': 'expf',
            'logf': 'logf',
            'sqrtf': 'sqrtf'}
_C_FUN64 = {k: v.rstrip('f') if v.endswith('f') and v != 'fabsf' else v
            for k, v in _C_FUN32.items()}
_C_FUN64.update({'sin': 'sin', 'cos': 'cos',
                 'abs': 'fabs', 'fabs': 'fabs', 'sinf': 'sin', 'cosf': 'cos',
                 'expf': 'exp', 'logf': 'log', 'sqrtf': 'sqrt'})


def c_literal(value, ctype='float') -> str:
    """Literal whose value is ``(ctype)(double)value`` exactly.

    Hex-float literals make the emitted constant bit-identical to a C cast of
    the double constant, independent of the compiler's decimal rounding."""
    v = float(value)
    if ctype == 'float':
        f = float(np.float32(v))
        if math.isinf(f):
            return '(__builtin_inff())' if f > 0 else '(-__builtin_inff())'
        return '(%sf)' % f.hex() if f < 0 else '%sf' % f.hex()
    if math.isinf(v):
        return '(__builtin_inf())' if v > 0 else '(-__builtin_inf())'
    return '(%s)' % v.hex() if v < 0 else v.hex()


def to_c(e: Expr, names: Dict[str, str] = None, ctype='float',
         unchecked_trig=False) -> str:
    """Emit ``e`` as a C expression.

    ``names`` maps variable names to C expressions (default: identity).
    Integer powers 2..4 become multiplication chains; other powers call
    ``pow``.  Comparisons yield ``(T)(a < b)``.  ``unchecked_trig`` (float
    only) calls the branch-free sine and cosine, NaN outside their reduced
    range, for code that checks a group of results at once."""
    names = names or {}
    funs = _C_FUN32 if ctype == 'float' else _C_FUN64
    if unchecked_trig and ctype == 'float':
        funs = dict(funs, sin='tvbx_sinf_nb', sinf='tvbx_sinf_nb',
                    cos='tvbx_cosf_nb', cosf='tvbx_cosf_nb')

    def go(x):
        if isinstance(x, Num):
            return c_literal(x.value, ctype)
        if isinstance(x, Var):
            if x.name == 'pi':
                return names.get('pi', c_literal(math.pi, ctype))
            return '(%s)' % names.get(x.name, x.name)
        if isinstance(x, Neg):
            return '(-%s)' % go(x.a)
        if isinstance(x, Cmp):
            return '((%s)(%s %s %s))' % (ctype, go(x.a), x.op, go(x.b))
        if isinstance(x, Bin):
            if x.op == '**':
                if _isnum(x.b) and float(x.b.value) in (1.0, 2.0, 3.0, 4.0):
                    k = int(x.b.value)
                    base = go(x.a)
                    return '(' + '*'.join([base] * k) + ')'
                return '%s(%s, %s)' % (funs['pow'], go(x.a), go(x.b))
            if x.op == '/' and ctype == 'float':
                # within one ulp, in a third of the correctly rounded
                # division's instructions (tvbx_math.h)
                return 'tvbx_divf(%s, %s)' % (go(x.a), go(x.b))
            return '(%s %s %s)' % (go(x.a), x.op, go(x.b))
        if isinstance(x, Call):
            fn = funs.get(x.fn)
            if fn is None:
                raise ValueError('unknown function %r' % x.fn)
            return '%s(%s)' % (fn, ', '.join(go(a) for a in x.args))
        raise ValueError('cannot emit %s' % x)
    return go(as_expr(e))


# ---------------------------------------------------------------------------
# NumPy evaluation

_NP_FUN = {'sin': np.sin, 'cos': np.cos, 'tan': np.tan, 'exp': np.exp,
           'expm1': np.expm1,
           'log': np.log, 'sqrt': np.sqrt, 'tanh': np.tanh, 'abs': np.abs,
           'fabs': np.abs, 'pow': np.power, 'sinh': np.sinh, 'cosh': np.cosh,
           'atan': np.arctan, 'asin': np.arcsin, 'acos': np.arccos,
           'sinf': np.sin, 'cosf': np.cos, 'expf': np.exp, 'logf': np.log,
           'sqrtf': np.sqrt}


def evaluate(e: Expr, env: Dict[str, object], dtype=np.float64):
    """Evaluate ``e`` with NumPy; missing names raise KeyError."""
    def go(x):
        if isinstance(x, Num):
            return dtype(x.value)
        if isinstance(x, Var):
            if x.name == 'pi' and 'pi' not in env:
                return dtype(math.pi)
            return env[x.name]
        if isinstance(x, Neg):
            return -go(x.a)
        if isinstance(x, Cmp):
            a, b = go(x.a), go(x.b)
            r = {'<': np.less, '>': np.greater, '<=': np.less_equal,
                 '>=': np.greater_equal, '==': np.equal,
                 '!=': np.not_equal}[x.op](a, b)
            return np.asarray(r).astype(dtype)
        if isinstance(x, Bin):
            a, b = go(x.a), go(x.b)
            if x.op == '**':
                if _isnum(x.b) and float(x.b.value) in (2.0, 3.0, 4.0):
                    k = int(x.b.value)
                    out = a
                    for _ in range(k - 1):
                        out = out * a
                    return out
                return np.power(a, b)
            return {'+': np.add, '-': np.subtract, '*': np.multiply,
                    '/': np.divide}[x.op](a, b)
        if isinstance(x, Call):
            return _NP_FUN[x.fn](*[go(a) for a in x.args])
        raise ValueError('cannot evaluate %s' % x)
    with np.errstate(all='ignore'):
        return go(as_expr(e))


def exprs_of(items: Iterable[Union[str, int, float]]):
    return [as_expr(i) for i in items]
