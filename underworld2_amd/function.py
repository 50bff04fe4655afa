"""A small function graph with the semantics of ``underworld.function`` used on this path.

Reference: PY/function/_function.py (operators), PY/function/branching.py (``map`` :19,
``conditional`` :144), PY/function/math.py; C++ closures in LU/Underworld/Function/src
(Map.cpp:18-127, Conditional.cpp, Binary.cpp).

Functions are evaluated vectorised over particles or points on the host to produce the
per-particle (or per-material) viscosity and body-force values that the GPU assembly
consumes.  When every input of a function is a single integer swarm variable, the Stokes
system evaluates it once per distinct value and hands the device a material table.
"""
import numpy as np

__all__ = ["Function", "input", "coord", "misc", "math", "branching", "convert"]


class _Ctx:
    def __init__(self, coords=None, swarm=None, override=None):
        self._coords = coords
        self.swarm = swarm
        self.override = override or {}

    @property
    def coords(self):
        if self._coords is None:
            self._coords = self.swarm.particleCoordinates.data
        return self._coords

    @property
    def n(self):
        if self._coords is not None:
            return self._coords.shape[0]
        if self.override:
            return len(next(iter(self.override.values())))
        return self.swarm.particleLocalCount


def convert(obj):
    if obj is None:
        return None
    if isinstance(obj, Function):
        return obj
    if hasattr(obj, "_fn_eval"):
        return _Wrap(obj)
    if isinstance(obj, (int, float, np.integer, np.floating)):
        return Constant(obj)
    if isinstance(obj, (tuple, list, np.ndarray)):
        items = [convert(o) for o in obj]
        if all(isinstance(i, Constant) for i in items):
            return Constant(np.concatenate([np.atleast_1d(i.value) for i in items]))
        return _Stack(items)
    raise TypeError(f"Cannot convert {type(obj)} to a Function")


class Function:
    def _fn_inputs(self):
        return set()

    def _fn_eval(self, ctx):
        raise NotImplementedError

    def evaluate(self, inp):
        """Evaluate at a swarm (its particles) or at an array of coordinates."""
        if hasattr(inp, "particleCoordinates"):
            ctx = _Ctx(swarm=inp)
        else:
            ctx = _Ctx(coords=np.atleast_2d(np.asarray(inp, dtype=np.float64)))
        return np.asarray(self._fn_eval(ctx))

    # arithmetic
    def _bin(self, other, op, rev=False):
        o = convert(other)
        return _Binary(o, self, op) if rev else _Binary(self, o, op)

    def __add__(self, o): return self._bin(o, np.add)
    def __radd__(self, o): return self._bin(o, np.add, True)
    def __sub__(self, o): return self._bin(o, np.subtract)
    def __rsub__(self, o): return self._bin(o, np.subtract, True)
    def __mul__(self, o): return self._bin(o, np.multiply)
    def __rmul__(self, o): return self._bin(o, np.multiply, True)
    def __truediv__(self, o): return self._bin(o, np.divide)
    def __rtruediv__(self, o): return self._bin(o, np.divide, True)
    def __pow__(self, o): return self._bin(o, np.power)
    def __neg__(self): return _Unary(self, np.negative)
    def __lt__(self, o): return self._bin(o, np.less)
    def __le__(self, o): return self._bin(o, np.less_equal)
    def __gt__(self, o): return self._bin(o, np.greater)
    def __ge__(self, o): return self._bin(o, np.greater_equal)
    def __and__(self, o): return self._bin(o, np.logical_and)
    def __or__(self, o): return self._bin(o, np.logical_or)

    def __getitem__(self, i):
        return _Index(self, i)


def _col(a, n):
    a = np.asarray(a)
    if a.ndim == 0:
        return np.full((n, 1), a)
    if a.ndim == 1:
        return np.broadcast_to(a[None, :], (n, a.shape[0]))
    return a


class Constant(Function):
    def __init__(self, value):
        self.value = np.asarray(value, dtype=np.float64 if not isinstance(value, (bool, np.bool_)) else bool)

    def _fn_eval(self, ctx):
        return _col(self.value, ctx.n)


class _Wrap(Function):
    def __init__(self, obj):
        self.obj = obj

    def _fn_inputs(self):
        return self.obj._fn_inputs()

    def _fn_eval(self, ctx):
        return self.obj._fn_eval(ctx)


class _Stack(Function):
    def __init__(self, items):
        self.items = items

    def _fn_inputs(self):
        s = set()
        for i in self.items:
            s |= i._fn_inputs()
        return s

    def _fn_eval(self, ctx):
        cols = [np.asarray(i._fn_eval(ctx), dtype=np.float64).reshape(ctx.n, -1) for i in self.items]
        return np.concatenate(cols, axis=1)


class _Binary(Function):
    def __init__(self, a, b, op):
        self.a, self.b, self.op = a, b, op

    def _fn_inputs(self):
        return self.a._fn_inputs() | self.b._fn_inputs()

    def _fn_eval(self, ctx):
        x = np.asarray(self.a._fn_eval(ctx)).reshape(ctx.n, -1)
        y = np.asarray(self.b._fn_eval(ctx)).reshape(ctx.n, -1)
        return self.op(x, y)


class _Unary(Function):
    def __init__(self, a, op):
        self.a, self.op = a, op

    def _fn_inputs(self):
        return self.a._fn_inputs()

    def _fn_eval(self, ctx):
        return self.op(np.asarray(self.a._fn_eval(ctx)).reshape(ctx.n, -1))


class _Index(Function):
    def __init__(self, a, i):
        self.a, self.i = a, i

    def _fn_inputs(self):
        return self.a._fn_inputs()

    def _fn_eval(self, ctx):
        return np.asarray(self.a._fn_eval(ctx)).reshape(ctx.n, -1)[:, self.i:self.i + 1]


class _Input(Function):
    def _fn_inputs(self):
        return {("coord",)}

    def _fn_eval(self, ctx):
        return ctx.coords


def input():
    """The evaluation coordinate (``fn.input()``)."""
    return _Input()


def coord():
    """Alias of ``input()`` (``fn.coord()``)."""
    return _Input()


class _Math:
    @staticmethod
    def dot(a, b):
        class _Dot(Function):
            def __init__(s, a, b):
                s.a, s.b = convert(a), convert(b)

            def _fn_inputs(s):
                return s.a._fn_inputs() | s.b._fn_inputs()

            def _fn_eval(s, ctx):
                x = np.asarray(s.a._fn_eval(ctx)).reshape(ctx.n, -1)
                y = np.asarray(s.b._fn_eval(ctx)).reshape(ctx.n, -1)
                return np.sum(x * y, axis=1, keepdims=True)

        return _Dot(a, b)

    @staticmethod
    def exp(a): return _Unary(convert(a), np.exp)

    @staticmethod
    def log(a): return _Unary(convert(a), np.log)

    @staticmethod
    def sqrt(a): return _Unary(convert(a), np.sqrt)

    @staticmethod
    def abs(a): return _Unary(convert(a), np.abs)

    @staticmethod
    def sin(a): return _Unary(convert(a), np.sin)

    @staticmethod
    def cos(a): return _Unary(convert(a), np.cos)


math = _Math()


class _Map(Function):
    """``branching.map(fn_key, mapping, fn_default)`` (PY/function/branching.py:19; Map.cpp)."""

    def __init__(self, fn_key, mapping, fn_default=None):
        self.key = convert(fn_key)
        self.mapping = {int(k): convert(v) for k, v in mapping.items()}
        self.default = convert(fn_default)

    def _fn_inputs(self):
        s = self.key._fn_inputs()
        for v in self.mapping.values():
            s |= v._fn_inputs()
        if self.default is not None:
            s |= self.default._fn_inputs()
        return s

    def _fn_eval(self, ctx):
        keys = np.asarray(self.key._fn_eval(ctx)).reshape(ctx.n).astype(np.int64)
        out = None
        done = np.zeros(ctx.n, dtype=bool)
        for k, f in self.mapping.items():
            m = keys == k
            v = np.asarray(f._fn_eval(ctx), dtype=np.float64).reshape(ctx.n, -1)
            if out is None:
                out = np.zeros((ctx.n, v.shape[1]))
            out[m] = v[m]
            done |= m
        if not done.all():
            if self.default is None:
                raise RuntimeError("branching.map: key not found in mapping and no default provided")
            v = np.asarray(self.default._fn_eval(ctx), dtype=np.float64).reshape(ctx.n, -1)
            out[~done] = v[~done]
        return out


class _Conditional(Function):
    """``branching.conditional([(cond, value), ...])`` (PY/function/branching.py:144):
    first true clause wins; an error if none is true."""

    def __init__(self, clauses):
        self.clauses = [(convert(c), convert(v)) for c, v in clauses]

    def _fn_inputs(self):
        s = set()
        for c, v in self.clauses:
            s |= c._fn_inputs() | v._fn_inputs()
        return s

    def _fn_eval(self, ctx):
        out = None
        done = np.zeros(ctx.n, dtype=bool)
        for c, v in self.clauses:
            cv = np.asarray(c._fn_eval(ctx)).reshape(ctx.n).astype(bool) & ~done
            vv = np.asarray(v._fn_eval(ctx), dtype=np.float64).reshape(ctx.n, -1)
            if out is None:
                out = np.zeros((ctx.n, vv.shape[1]))
            out[cv] = vv[cv]
            done |= cv
        if not done.all():
            raise RuntimeError("branching.conditional: no clause was true for some inputs")
        return out


class _Branching:
    @staticmethod
    def map(fn_key=None, mapping=None, fn_default=None, keyFunc=None, mappingDict=None):
        return _Map(fn_key if fn_key is not None else keyFunc, mapping if mapping is not None else mappingDict,
                    fn_default)

    @staticmethod
    def conditional(clauses):
        return _Conditional(clauses)


branching = _Branching()


class _Misc:
    @staticmethod
    def constant(v):
        return Constant(v)


misc = _Misc()
