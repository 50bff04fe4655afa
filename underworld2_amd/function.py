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
        """Evaluate at a swarm (its particles), a mesh (its nodes) or an array of coordinates."""
        if hasattr(inp, "particleCoordinates"):
            ctx = _Ctx(swarm=inp)
        elif hasattr(inp, "elementRes") and hasattr(inp, "data"):
            ctx = _Ctx(coords=np.array(inp.data, dtype=np.float64))
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


# ------------------------------------------------------------------------------------------
# device programs: the graph compiled for per-particle evaluation on the GPU
# (the instruction set of fn_kernels.hip)
# ------------------------------------------------------------------------------------------
class NotCompilable(Exception):
    """The graph holds a node the device evaluator does not provide."""


_BIN = {np.add: 10, np.subtract: 11, np.multiply: 12, np.divide: 13, np.power: 14, np.less: 15,
        np.less_equal: 16, np.greater: 17, np.greater_equal: 18, np.logical_and: 19, np.logical_or: 20}
_UN = {np.negative: 30, np.exp: 31, np.log: 32, np.sqrt: 33, np.abs: 34, np.sin: 35, np.cos: 36}
_NREG = 48


class DeviceProgram:
    """Register program for a set of functions evaluated at every particle of a swarm."""

    def __init__(self, dim):
        self.dim = dim
        self.ins = []
        self.consts = []
        self.fields = []      # MeshVariable objects, by field index
        self.swarmvars = []   # SwarmVariable objects, by slot
        self.uses_material = None
        self._memo = {}
        self._nreg = 0

    # registers and constants
    def _reg(self):
        r = self._nreg
        self._nreg += 1
        if r >= _NREG:
            raise NotCompilable("function graph too large for the device evaluator")
        return r

    def _emit(self, op, dst, a=-1, b=-1, n=1):
        self.ins.append((op, dst, a, b, n))
        return dst

    def _const(self, values):
        values = [float(v) for v in np.atleast_1d(np.asarray(values, dtype=np.float64)).ravel()]
        if not 1 <= len(values) <= 3:
            raise NotCompilable("constants of 1-3 components only")
        off = len(self.consts)
        self.consts.extend(values)
        r = self._reg()
        self._emit(0, r, off, -1, len(values))
        return r, len(values)

    def compile(self, f):
        """(register, components) holding f at the particle."""
        key = id(f)
        if key in self._memo:
            return self._memo[key]
        out = self._compile(f)
        self._memo[key] = out
        return out

    def _compile(self, f):
        from .mesh import MeshVariable
        from .swarm import SwarmVariable
        if isinstance(f, _Wrap):
            return self.compile(f.obj)
        if isinstance(f, Constant):
            return self._const(f.value.astype(np.float64))
        if isinstance(f, _Input):
            r = self._reg()
            return self._emit(1, r, n=self.dim), self.dim
        if isinstance(f, MeshVariable):
            if f._on_elements or not 1 <= f.nodeDofCount <= 3:
                raise NotCompilable("only nodal mesh variables of 1-3 components")
            if len(self.fields) >= 8 and f not in self.fields:
                raise NotCompilable("at most 8 mesh variables")
            if f not in self.fields:
                self.fields.append(f)
            r = self._reg()
            return self._emit(2, r, self.fields.index(f), -1, f.nodeDofCount), f.nodeDofCount
        if isinstance(f, SwarmVariable):
            if not 1 <= f.count <= 3:
                raise NotCompilable("only swarm variables of 1-3 components")
            if f not in self.swarmvars:
                if len(self.swarmvars) >= 8:
                    raise NotCompilable("at most 8 swarm variables")
                self.swarmvars.append(f)
            r = self._reg()
            return self._emit(3, r, self.swarmvars.index(f), -1, f.count), f.count
        if isinstance(f, _Binary):
            if f.op not in _BIN:
                raise NotCompilable(f"binary operation {f.op}")
            ra, na = self.compile(f.a)
            rb, nb = self.compile(f.b)
            if na != nb and 1 not in (na, nb):
                raise NotCompilable("operands of incompatible shapes")
            r = self._reg()
            n = max(na, nb)
            return self._emit(_BIN[f.op], r, ra, rb, n), n
        if isinstance(f, _Unary):
            if f.op not in _UN:
                raise NotCompilable(f"unary operation {f.op}")
            ra, na = self.compile(f.a)
            r = self._reg()
            return self._emit(_UN[f.op], r, ra, -1, na), na
        if isinstance(f, _Index):
            ra, na = self.compile(f.a)
            if not 0 <= f.i < na:
                raise NotCompilable("component index out of range")
            r = self._reg()
            return self._emit(41, r, ra, f.i, 1), 1
        if isinstance(f, _Stack):
            ra, na = self.compile(f.items[0])
            for item in f.items[1:]:
                rb, nb = self.compile(item)
                if na + nb > 3:
                    raise NotCompilable("vectors of more than 3 components")
                r = self._reg()
                self._emit(42, r, ra, rb, na + nb)
                ra, na = r, na + nb
            return ra, na
        if type(f).__name__ == "_Dot":
            ra, na = self.compile(f.a)
            rb, nb = self.compile(f.b)
            if na != nb and nb != 1:
                raise NotCompilable("dot of incompatible shapes")
            r = self._reg()
            return self._emit(40, r, ra, rb, 1), 1
        if isinstance(f, _Map):
            rk, nk = self.compile(f.key)
            vals = [(k, self.compile(v)) for k, v in f.mapping.items()]
            dflt = self.compile(f.default) if f.default is not None else None
            n = max([nv for _, (_, nv) in vals] + ([dflt[1]] if dflt else []))
            if any(nv != n for _, (_, nv) in vals) or (dflt and dflt[1] != n):
                raise NotCompilable("map values of different shapes")
            res, _ = self._const([0.0] * n)
            found, _ = self._const([0.0])
            for k, (rv, _) in vals:
                kc = len(self.consts)
                self.consts.append(float(k))
                m = self._reg()
                self._emit(51, m, rk, kc, 1)
                self._emit(52, res, rv, m, n)
                f2 = self._reg()
                self._emit(20, f2, found, m, 1)
                found = f2
            if dflt is not None:
                nf = self._reg()
                self._emit(37, nf, found, -1, 1)
                self._emit(52, res, dflt[0], nf, n)
            else:
                self._emit(53, 0, found, -1, 1)
            return res, n
        if isinstance(f, _Conditional):
            clauses = [(self.compile(c), self.compile(v)) for c, v in f.clauses]
            n = max(nv for _, (_, nv) in clauses)
            if any(nv != n for _, (_, nv) in clauses) or any(nc != 1 for (_, nc), _ in clauses):
                raise NotCompilable("conditional clauses of different shapes")
            res, _ = self._const([0.0] * n)
            done, _ = self._const([0.0])
            for (rc, _), (rv, _) in clauses:
                nd = self._reg()
                self._emit(37, nd, done, -1, 1)
                take = self._reg()
                self._emit(19, take, rc, nd, 1)
                self._emit(52, res, rv, take, n)
                d2 = self._reg()
                self._emit(20, d2, done, rc, 1)
                done = d2
            self._emit(53, 0, done, -1, 1)
            return res, n
        raise NotCompilable(f"{type(f).__name__} is evaluated on the host")

    def packed(self):
        return (np.ascontiguousarray(np.array(self.ins, dtype=np.int32).reshape(-1, 5)),
                np.ascontiguousarray(np.array(self.consts if self.consts else [0.0], dtype=np.float64)))


def compile_particle_program(fn_viscosity, fn_bodyforce, dim):
    """Compile viscosity (scalar) and body force (dim-vector) for device evaluation.
    Returns (program, eta_register, force_register); raises NotCompilable."""
    prog = DeviceProgram(dim)
    re, ne = prog.compile(fn_viscosity)
    if ne != 1:
        raise NotCompilable("viscosity must be a scalar")
    rf, nf = prog.compile(fn_bodyforce)
    if nf not in (1, dim):
        raise NotCompilable("fn_bodyforce must be a vector of the mesh dimension")
    return prog, re, rf
