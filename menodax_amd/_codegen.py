"""Right-hand sides into HIP C++: tracing, forward-mode derivatives, emission.

The reference hands its callback to numba and differentiates the compiled code
with Enzyme (modax/_sensitivity.py:148-179, modax/rodas5P.py:251-303). Here the
callback is *traced*: it is called once with symbolic arguments, which records
every arithmetic operation into a hash-consed expression graph. From that graph
this module emits

  - ``rhs``: the right-hand side as straight-line double-precision code;
  - ``neg_jac``: one forward-mode sweep per colour of the Jacobian's column
    intersection graph, the colour's seed spelled out so every zero tangent is
    pruned when the code is generated rather than by a compiler, writing
    ``-J`` straight into the slots of the iteration-matrix buffer, plus the
    ``t``-seeded sweep that gives ``df/dt`` (rodas5P.py:445-500);
  - flop counts of both, used for the roofline.

Evaluation order follows Python's: ``a + b + c`` is ``(a + b) + c``, an integer
power is exponentiation by squaring as numba lowers it, and the kernels compile
with ``-ffp-contract=off``, so the GPU rounds every operation exactly where the
Python expression does. ``evaluate`` interprets the same graph on the host with
the same rounding, which is how tests hold the emitted derivatives against
closed forms bit for bit.

The callback language is the reference's (docs/guide/callbacks.md): a function
``f(y, t, p)`` returning a tuple or list of scalars, constant indices into ``y``
and ``p``, arithmetic, ``**``, ``abs``/``min``/``max`` and ``math`` functions,
closed-over constants and tables. Branching on a state-dependent value is not
traceable and raises.
"""

from __future__ import annotations

import builtins
import functools
import math
import struct
import types
from dataclasses import dataclass

import numpy as np

# ---------------------------------------------------------------------------
# Expression graph
# ---------------------------------------------------------------------------
_UNARY_CALLS = {
    "exp": math.exp,
    "log": math.log,
    "sqrt": math.sqrt,
    "sin": math.sin,
    "cos": math.cos,
    "tan": math.tan,
    "tanh": math.tanh,
    "sinh": math.sinh,
    "cosh": math.cosh,
    "atan": math.atan,
    "asin": math.asin,
    "acos": math.acos,
    "log10": math.log10,
    "log2": math.log2,
    "expm1": math.expm1,
    "log1p": math.log1p,
    "fabs": math.fabs,
    "erf": math.erf,
}
_BINARY_CALLS = {"pow": math.pow, "atan2": math.atan2, "copysign": math.copysign}
_CMP = {"lt": "<", "le": "<=", "gt": ">", "ge": ">=", "eq": "==", "ne": "!="}


class Node:
    """One vertex of the expression graph. ``op`` is one of ``y``, ``p``,
    ``t``, ``in`` (element ``value[1]`` of a named extra input vector
    ``value[0]``, e.g. a derivative direction), ``c`` (constant), ``add``,
    ``sub``, ``mul``, ``div``, ``neg``, ``call`` (``fn`` names it), ``cmp``
    (``fn`` is the relation) or ``select``."""

    __slots__ = ("op", "args", "value", "fn", "id")

    def __init__(self, op, args, value, fn, nid):
        self.op = op
        self.args = args
        self.value = value
        self.fn = fn
        self.id = nid

    def __repr__(self):
        return f"Node#{self.id}({self.op}{'' if self.fn is None else ':' + self.fn})"


class Graph:
    """Hash-consed expression graph: structurally equal nodes are one node, so
    a subexpression the callback repeats is computed once (with the same value
    it would have had each time)."""

    def __init__(self):
        self.nodes: list[Node] = []
        self._index: dict = {}

    def _make(self, op, args=(), value=None, fn=None):
        if op == "c":
            value = float(value)
            key = ("c", struct.pack("<d", value))
        else:
            key = (op, tuple(a.id for a in args), value, fn)
        hit = self._index.get(key)
        if hit is not None:
            return hit
        node = Node(op, tuple(args), value, fn, len(self.nodes))
        self.nodes.append(node)
        self._index[key] = node
        return node

    def const(self, v):
        return self._make("c", value=v)

    def y(self, i):
        return self._make("y", value=int(i))

    def p(self, i):
        return self._make("p", value=int(i))

    def t(self):
        return self._make("t")

    def inp(self, name, i):
        return self._make("in", value=(str(name), int(i)))

    # Arithmetic with constant folding. Folding evaluates the same IEEE
    # operation the device would, so it never changes a value.
    def add(self, a, b):
        if a.op == "c" and b.op == "c":
            return self.const(a.value + b.value)
        return self._make("add", (a, b))

    def sub(self, a, b):
        if a.op == "c" and b.op == "c":
            return self.const(a.value - b.value)
        return self._make("sub", (a, b))

    def mul(self, a, b):
        if a.op == "c" and b.op == "c":
            return self.const(a.value * b.value)
        return self._make("mul", (a, b))

    def div(self, a, b):
        if a.op == "c" and b.op == "c" and b.value != 0.0:
            return self.const(a.value / b.value)
        return self._make("div", (a, b))

    def neg(self, a):
        if a.op == "c":
            return self.const(-a.value)
        return self._make("neg", (a,))

    def call(self, fn, *args):
        if all(a.op == "c" for a in args):
            impl = _UNARY_CALLS.get(fn) or _BINARY_CALLS.get(fn)
            try:
                return self.const(impl(*(a.value for a in args)))
            except (ValueError, OverflowError):
                pass
        return self._make("call", tuple(args), fn=fn)

    def cmp(self, rel, a, b):
        return self._make("cmp", (a, b), fn=rel)

    def select(self, cond, a, b):
        return self._make("select", (cond, a, b))


class TraceError(TypeError):
    """The callback did something the tracer cannot record."""


class Sym:
    """A traced scalar. Arithmetic on it records nodes in the graph."""

    __slots__ = ("g", "n")
    __array_priority__ = 1000

    def __init__(self, g: Graph, n: Node):
        self.g = g
        self.n = n

    # -- conversion -------------------------------------------------------
    def _lift(self, other):
        if isinstance(other, Sym):
            return other.n
        if isinstance(other, (bool, int, float, np.floating, np.integer)):
            return self.g.const(float(other))
        return NotImplemented

    def __bool__(self):
        raise TraceError(
            "the right-hand side branches on a traced value; a state- or "
            "time-dependent condition must be written with min/max/abs or "
            "menodax_amd.where(cond, a, b)"
        )

    def __float__(self):
        raise TraceError("a traced value cannot be converted to a Python float")

    __int__ = __index__ = __float__

    # -- arithmetic -------------------------------------------------------
    def __add__(self, o):
        o = self._lift(o)
        return NotImplemented if o is NotImplemented else Sym(self.g, self.g.add(self.n, o))

    def __radd__(self, o):
        o = self._lift(o)
        return NotImplemented if o is NotImplemented else Sym(self.g, self.g.add(o, self.n))

    def __sub__(self, o):
        o = self._lift(o)
        return NotImplemented if o is NotImplemented else Sym(self.g, self.g.sub(self.n, o))

    def __rsub__(self, o):
        o = self._lift(o)
        return NotImplemented if o is NotImplemented else Sym(self.g, self.g.sub(o, self.n))

    def __mul__(self, o):
        o = self._lift(o)
        return NotImplemented if o is NotImplemented else Sym(self.g, self.g.mul(self.n, o))

    def __rmul__(self, o):
        o = self._lift(o)
        return NotImplemented if o is NotImplemented else Sym(self.g, self.g.mul(o, self.n))

    def __truediv__(self, o):
        o = self._lift(o)
        return NotImplemented if o is NotImplemented else Sym(self.g, self.g.div(self.n, o))

    def __rtruediv__(self, o):
        o = self._lift(o)
        return NotImplemented if o is NotImplemented else Sym(self.g, self.g.div(o, self.n))

    def __neg__(self):
        return Sym(self.g, self.g.neg(self.n))

    def __pos__(self):
        return self

    def __abs__(self):
        return Sym(self.g, self.g.call("fabs", self.n))

    def __pow__(self, o):
        if isinstance(o, (int, np.integer)) or (
            isinstance(o, (float, np.floating)) and float(o).is_integer() and abs(o) <= 64
        ):
            return Sym(self.g, _int_power(self.g, self.n, int(o)))
        o = self._lift(o)
        if o is NotImplemented:
            return NotImplemented
        return Sym(self.g, self.g.call("pow", self.n, o))

    def __rpow__(self, o):
        o = self._lift(o)
        return NotImplemented if o is NotImplemented else Sym(self.g, self.g.call("pow", o, self.n))

    def _rel(self, rel, o):
        o = self._lift(o)
        return NotImplemented if o is NotImplemented else Cond(self.g, self.g.cmp(rel, self.n, o))

    def __lt__(self, o):
        return self._rel("lt", o)

    def __le__(self, o):
        return self._rel("le", o)

    def __gt__(self, o):
        return self._rel("gt", o)

    def __ge__(self, o):
        return self._rel("ge", o)

    # numpy ufuncs on object scalars call the method of the same name
    def exp(self):
        return _call("exp", self)

    def log(self):
        return _call("log", self)

    def sqrt(self):
        return _call("sqrt", self)

    def sin(self):
        return _call("sin", self)

    def cos(self):
        return _call("cos", self)

    def tan(self):
        return _call("tan", self)

    def tanh(self):
        return _call("tanh", self)

    def sinh(self):
        return _call("sinh", self)

    def cosh(self):
        return _call("cosh", self)

    def arctan(self):
        return _call("atan", self)

    def log10(self):
        return _call("log10", self)

    def expm1(self):
        return _call("expm1", self)

    def log1p(self):
        return _call("log1p", self)


class Cond:
    """A traced comparison: the first argument of ``where``, or, in arithmetic,
    the number 1.0 or 0.0 that Python makes of a bool (``a + (b - a) * (x > c)``,
    the branchless select the reference's callbacks write); it is recorded as
    ``select(cond, 1.0, 0.0)`` and then multiplied or added exactly as the
    number would be."""

    __slots__ = ("g", "n")

    def __init__(self, g, n):
        self.g = g
        self.n = n

    def __bool__(self):
        raise TraceError(
            "the right-hand side branches on a traced comparison; use "
            "menodax_amd.where(cond, a, b), min, max or abs instead"
        )

    def _num(self):
        return Sym(self.g, self.g.select(self.n, self.g.const(1.0), self.g.const(0.0)))

    def __add__(self, o):
        return self._num() + (o._num() if isinstance(o, Cond) else o)

    def __radd__(self, o):
        return o + self._num()

    def __sub__(self, o):
        return self._num() - (o._num() if isinstance(o, Cond) else o)

    def __rsub__(self, o):
        return o - self._num()

    def __mul__(self, o):
        return self._num() * (o._num() if isinstance(o, Cond) else o)

    def __rmul__(self, o):
        return o * self._num()

    def __truediv__(self, o):
        return self._num() / (o._num() if isinstance(o, Cond) else o)

    def __rtruediv__(self, o):
        return o / self._num()

    def __neg__(self):
        return -self._num()


def _int_power(g: Graph, a: Node, e: int) -> Node:
    """``a ** e`` by squaring, in the order numba's integer power takes it."""
    invert = e < 0
    e = abs(e)
    r = None
    base = a
    if e == 0:
        r = g.const(1.0)
    while e:
        if e & 1:
            r = base if r is None else g.mul(r, base)
        e >>= 1
        if e:
            base = g.mul(base, base)
    if invert:
        r = g.div(g.const(1.0), r)
    return r


def _call(fn, *args):
    syms = [a for a in args if isinstance(a, Sym)]
    if not syms:
        impl = _UNARY_CALLS.get(fn) or _BINARY_CALLS.get(fn)
        return impl(*args)
    g = syms[0].g
    nodes = [a.n if isinstance(a, Sym) else g.const(float(a)) for a in args]
    return Sym(g, g.call(fn, *nodes))


def where(cond, a, b):
    """``a if cond else b`` for traced values (both branches are evaluated)."""
    if isinstance(cond, Cond):
        g = cond.g
        na = a.n if isinstance(a, Sym) else g.const(float(a))
        nb = b.n if isinstance(b, Sym) else g.const(float(b))
        return Sym(g, g.select(cond.n, na, nb))
    return a if cond else b


def _sym_max(*args, **kw):
    if len(args) == 1:
        args = tuple(args[0])
    if not any(isinstance(a, Sym) for a in args) or kw:
        return builtins.max(*args, **kw)
    best = args[0]
    for x in args[1:]:
        # Python keeps the first maximal element: replace only on x > best.
        best = where(_as_sym(x, best) > best, x, best)
    return best


def _sym_min(*args, **kw):
    if len(args) == 1:
        args = tuple(args[0])
    if not any(isinstance(a, Sym) for a in args) or kw:
        return builtins.min(*args, **kw)
    best = args[0]
    for x in args[1:]:
        best = where(_as_sym(x, best) < best, x, best)
    return best


def _as_sym(x, like):
    if isinstance(x, Sym):
        return x
    g = like.g if isinstance(like, Sym) else None
    return Sym(g, g.const(float(x)))


class _MathShim(types.ModuleType):
    """Stands in for ``math`` while a callback is traced."""

    def __init__(self):
        super().__init__("math")
        for name in dir(math):
            setattr(self, name, getattr(math, name))
        for name in _UNARY_CALLS:
            setattr(self, name, functools.partial(_call, name))
        self.pow = lambda a, b: a**b if isinstance(a, Sym) and not isinstance(b, Sym) else _call(
            "pow", a, b
        )
        self.atan2 = functools.partial(_call, "atan2")
        self.copysign = functools.partial(_call, "copysign")


_MATH_SHIM = _MathShim()
_PATCHED_MATH = tuple(_UNARY_CALLS) + ("pow", "atan2", "copysign")


@dataclass
class Traced:
    graph: Graph
    outputs: list  # one Node per component
    n_vars: int
    n_params: int
    uses_t: bool


def trace(ode_fn, n_vars: int, n_params: int) -> Traced:
    """Record ``ode_fn(y, t, p)`` into a graph."""
    g = Graph()
    ys = [Sym(g, g.y(i)) for i in range(n_vars)]
    ps = [Sym(g, g.p(j)) for j in range(n_params)]
    ts = Sym(g, g.t())
    fn = _with_traceable_globals(ode_fn)
    saved = {name: getattr(math, name) for name in _PATCHED_MATH}
    try:
        for name in _PATCHED_MATH:
            setattr(math, name, getattr(_MATH_SHIM, name))
        out = fn(_Vector(ys), ts, _Vector(ps))
    finally:
        for name, value in saved.items():
            setattr(math, name, value)
    if isinstance(out, (Sym, int, float)):
        out = (out,)
    out = list(out)
    if len(out) != n_vars:
        raise ValueError(f"ode_fn returned {len(out)} components for a {n_vars}-variable state")
    nodes = []
    for v in out:
        if isinstance(v, Sym):
            nodes.append(v.n)
        elif isinstance(v, (int, float, np.floating, np.integer)):
            nodes.append(g.const(float(v)))
        else:
            raise TraceError(f"ode_fn returned a component of type {type(v).__name__}")
    uses_t = any(n.op == "t" for n in g.nodes)
    return Traced(g, nodes, n_vars, n_params, uses_t)


class _Vector:
    """Constant-index view over traced components, as the callback sees ``y``/``p``."""

    __slots__ = ("_items",)

    def __init__(self, items):
        self._items = items

    def __getitem__(self, k):
        if isinstance(k, slice):
            return self._items[k]
        if isinstance(k, Sym):
            raise TraceError("y and p must be indexed with constants")
        return self._items[int(k)]

    def __len__(self):
        return len(self._items)

    def __iter__(self):
        return iter(self._items)


class _OutRecorder:
    """The ``out`` array an array-form right-hand side writes into while traced."""

    def __init__(self, g: Graph, n: int):
        self._g = g
        self._n = n
        self.values: dict[int, Node] = {}

    def __setitem__(self, k, v):
        if isinstance(k, Sym):
            raise TraceError("out must be indexed with constants")
        k = int(k)
        if not 0 <= k < self._n:
            raise IndexError(f"out index {k} outside the {self._n}-variable state")
        if isinstance(v, Sym):
            self.values[k] = v.n
        elif isinstance(v, (bool, int, float, np.floating, np.integer)):
            self.values[k] = self._g.const(float(v))
        else:
            raise TraceError(f"out[{k}] was given a value of type {type(v).__name__}")

    def __getitem__(self, k):
        k = int(k)
        if k not in self.values:
            raise TraceError(f"out[{k}] read before it was written")
        return Sym(self._g, self.values[k])

    def __len__(self):
        return self._n


def trace_array_rhs(array_rhs, n_vars: int, n_params: int) -> Traced:
    """Record an array-form right-hand side ``array_rhs(y, t, p, out)``
    (modax/rodas5P.py:382-392): every component of ``out`` must be written."""
    g = Graph()
    ys = [Sym(g, g.y(i)) for i in range(n_vars)]
    ps = [Sym(g, g.p(j)) for j in range(n_params)]
    ts = Sym(g, g.t())
    out = _OutRecorder(g, n_vars)
    fn = _with_traceable_globals(array_rhs)
    saved = {name: getattr(math, name) for name in _PATCHED_MATH}
    try:
        for name in _PATCHED_MATH:
            setattr(math, name, getattr(_MATH_SHIM, name))
        fn(_Vector(ys), ts, _Vector(ps), out)
    finally:
        for name, value in saved.items():
            setattr(math, name, value)
    missing = [j for j in range(n_vars) if j not in out.values]
    if missing:
        raise ValueError(f"array_rhs left out components {missing} of the {n_vars}-variable state")
    nodes = [out.values[j] for j in range(n_vars)]
    return Traced(g, nodes, n_vars, n_params, any(n.op == "t" for n in g.nodes))


class _HookAcc:
    """The hook's accumulator row as the hook sees it while traced: reads of a
    slot not yet written this call are inputs, writes are recorded in order."""

    def __init__(self, g: Graph, size: int):
        self._g = g
        self._size = size
        self.writes: dict[int, Node] = {}

    def _slot(self, k):
        if isinstance(k, Sym):
            raise TraceError("the hook's accumulator must be indexed with constants or save_idx")
        k = int(k)
        if not 0 <= k < self._size:
            raise IndexError(f"hook accumulator index {k} outside hook_size {self._size}")
        return k

    def __getitem__(self, k):
        k = self._slot(k)
        node = self.writes.get(k)
        return Sym(self._g, self._g.inp("acc", k) if node is None else node)

    def __setitem__(self, k, v):
        k = self._slot(k)
        if isinstance(v, Sym):
            self.writes[k] = v.n
        elif isinstance(v, (bool, int, float, np.floating, np.integer)):
            self.writes[k] = self._g.const(float(v))
        else:
            raise TraceError(f"the hook stored a value of type {type(v).__name__}")

    def __len__(self):
        return self._size


@dataclass
class TracedHook:
    """A save hook traced once per save index: ``cases[s]`` is the list of
    ``(slot, node)`` stores the call at save ``s`` makes."""

    graph: Graph
    cases: list
    n_vars: int
    n_params: int
    hook_size: int


def trace_hook(hook, n_vars: int, n_params: int, n_save: int, hook_size: int) -> TracedHook:
    """Record ``hook(save_idx, y, t, p_row, acc)`` (modax/rodas5P.py:157-164,
    976-1002) for every save index: the index is a constant in each trace, so
    ``acc[save_idx]`` and host arrays indexed by it resolve at trace time."""
    g = Graph()
    ys = _Vector([Sym(g, g.y(i)) for i in range(n_vars)])
    ps = _Vector([Sym(g, g.p(j)) for j in range(n_params)])
    ts = Sym(g, g.t())
    fn = _with_traceable_globals(hook)
    saved = {name: getattr(math, name) for name in _PATCHED_MATH}
    cases = []
    try:
        for name in _PATCHED_MATH:
            setattr(math, name, getattr(_MATH_SHIM, name))
        for sv in range(n_save):
            acc = _HookAcc(g, hook_size)
            fn(sv, ys, ts, ps, acc)
            cases.append(sorted(acc.writes.items()))
    finally:
        for name, value in saved.items():
            setattr(math, name, value)
    return TracedHook(g, cases, n_vars, n_params, hook_size)


def generate_hook(th: TracedHook) -> str:
    """``save_hook(save_idx, y, t, p, acc)`` as a switch over the save index,
    save indices whose traces are identical sharing one case."""
    groups: dict = {}
    for sv, writes in enumerate(th.cases):
        groups.setdefault(tuple((k, n.id) for k, n in writes), []).append(sv)
    lines = ["template <class YA, class PA, class AC>",
             "__device__ __forceinline__ void save_hook(int save_idx, const YA& y, const double t,",
             "                                          const PA& p, const AC& acc) {",
             "  (void)y;", "  (void)t;", "  (void)p;", "  (void)acc;", "  switch (save_idx) {"]
    for key, svs in groups.items():
        if not key:
            continue
        nodes = [th.graph.nodes[nid] for _, nid in key]
        body, _ = Emitter(th.graph).body(nodes, indent="      ")
        lines.append("    " + " ".join(f"case {sv}:" for sv in svs) + " {")
        lines += body
        lines += [f"      acc.put({k}, v{nid});" for k, nid in key]
        lines += ["      break;", "    }"]
    lines += ["    default:", "      break;", "  }", "}"]
    return "\n".join(lines) + "\n"


def _traceable_value(value, depth):
    """The traceable stand-in for a value a callback reaches through a global
    or a closure cell: math functions and the math module map to the shim,
    builtin ``max``/``min`` to their traced forms, and a plain Python helper
    function is itself rewritten (the pattern of a callback built by a
    factory that receives its math operations as arguments,
    examples/dual_backend.py:78-111)."""
    if value is math:
        return _MATH_SHIM
    if value is max:
        return _sym_max
    if value is min:
        return _sym_min
    if callable(value) and getattr(value, "__module__", None) == "math":
        name = getattr(value, "__name__", "")
        if name in _PATCHED_MATH:
            return getattr(_MATH_SHIM, name)
    if isinstance(value, types.FunctionType) and depth < 8:
        return _with_traceable_globals(value, depth + 1)
    return value


def _with_traceable_globals(fn, depth=0):
    """A copy of ``fn`` whose globals and closure see ``min``/``max``/``where``
    and the math shim."""
    if not isinstance(fn, types.FunctionType):
        return fn
    closure = fn.__closure__
    if closure:
        cells = []
        for cell in closure:
            try:
                content = cell.cell_contents
            except ValueError:  # an empty cell
                cells.append(cell)
                continue
            new = _traceable_value(content, depth)
            cells.append(cell if new is content else types.CellType(new))
        closure = tuple(cells)
    glb = dict(fn.__globals__)
    glb["max"] = _sym_max
    glb["min"] = _sym_min
    for name, value in list(glb.items()):
        if value is math:
            glb[name] = _MATH_SHIM
        elif callable(value) and getattr(value, "__module__", None) == "math" and name in _PATCHED_MATH:
            glb[name] = getattr(_MATH_SHIM, name)
    for name in _PATCHED_MATH:
        if glb.get(name) is getattr(math, name, None):
            glb[name] = getattr(_MATH_SHIM, name)
    copy = types.FunctionType(fn.__code__, glb, fn.__name__, fn.__defaults__, closure)
    copy.__kwdefaults__ = fn.__kwdefaults__
    return copy


# ---------------------------------------------------------------------------
# Forward-mode tangents
# ---------------------------------------------------------------------------
def tangents(g: Graph, outputs, seed) -> list:
    """Directional derivative of ``outputs`` for ``seed = {input node id: d}``,
    ``d`` either ``1.0`` or a node (a symbolic direction, e.g. an ``in`` input).

    Returns one node per output, or ``None`` where the derivative is
    structurally zero (``None`` outputs stay ``None``). The rules are the
    textbook ones; a zero tangent is dropped and a unit one folded, so the
    emitted code holds only what the seeded inputs actually reach.
    """
    tan: dict[int, Node | None] = {}
    one = g.const(1.0)

    def d(node):
        return tan.get(node.id)

    present = [n for n in outputs if n is not None]
    needed = _ancestors(present)
    for node in g.nodes[: max((n.id for n in present), default=-1) + 1]:
        if node.id not in needed:
            continue
        op = node.op
        if op in ("y", "p", "t", "in"):
            sd = seed.get(node.id)
            tan[node.id] = sd if isinstance(sd, Node) else (one if sd else None)
            continue
        if op in ("c", "cmp"):
            tan[node.id] = None
            continue
        args = node.args
        if op == "add":
            da, db = d(args[0]), d(args[1])
            tan[node.id] = _add(g, da, db)
        elif op == "sub":
            da, db = d(args[0]), d(args[1])
            if db is None:
                tan[node.id] = da
            elif da is None:
                tan[node.id] = g.neg(db)
            else:
                tan[node.id] = g.sub(da, db)
        elif op == "mul":
            a, b = args
            da, db = d(a), d(b)
            left = None if da is None else _mul(g, da, b)
            right = None if db is None else _mul(g, a, db)
            tan[node.id] = _add(g, left, right)
        elif op == "div":
            a, b = args
            da, db = d(a), d(b)
            if db is None:
                tan[node.id] = None if da is None else g.div(da, b)
            else:
                # q = a/b, dq = (da - q*db) / b
                num = _mul(g, node, db)
                num = g.neg(num) if da is None else g.sub(da, num)
                tan[node.id] = g.div(num, b)
        elif op == "neg":
            da = d(args[0])
            tan[node.id] = None if da is None else g.neg(da)
        elif op == "select":
            _, a, b = args
            da, db = d(a), d(b)
            if da is None and db is None:
                tan[node.id] = None
            else:
                zero = g.const(0.0)
                tan[node.id] = g.select(args[0], da or zero, db or zero)
        elif op == "call":
            tan[node.id] = _call_tangent(g, node, [d(a) for a in args])
        else:  # pragma: no cover - every op is handled above
            raise AssertionError(op)
    return [None if n is None else tan.get(n.id) for n in outputs]


def adjoints(g: Graph, seeds, wrt) -> list:
    """Reverse mode: for ``seeds = [(node, weight node), ...]`` the gradient of
    ``sum weight * node`` with respect to each input node in ``wrt`` (one node
    per entry, ``None`` where structurally zero). The local partials are the
    forward rules' (a call's partial is its tangent seeded with 1)."""
    adj: dict[int, Node | None] = {}
    one = g.const(1.0)
    zero = g.const(0.0)

    def acc(node, v):
        if v is not None:
            adj[node.id] = _add(g, adj.get(node.id), v)

    present = [(n, w) for n, w in seeds if n is not None and w is not None]
    for n, w in present:
        acc(n, w)
    needed = _ancestors([n for n, _ in present])
    top = max((n.id for n, _ in present), default=-1)
    for node in reversed(g.nodes[: top + 1]):
        if node.id not in needed:
            continue
        A = adj.get(node.id)
        if A is None or node.op in ("y", "p", "t", "in", "c", "cmp"):
            continue
        op, args = node.op, node.args
        if op == "add":
            acc(args[0], A)
            acc(args[1], A)
        elif op == "sub":
            acc(args[0], A)
            acc(args[1], g.neg(A))
        elif op == "mul":
            acc(args[0], _mul(g, A, args[1]))
            acc(args[1], _mul(g, A, args[0]))
        elif op == "div":
            a, b = args
            acc(a, g.div(A, b))
            acc(b, g.neg(g.div(_mul(g, A, node), b)))
        elif op == "neg":
            acc(args[0], g.neg(A))
        elif op == "select":
            c, x, y = args
            acc(x, g.select(c, A, zero))
            acc(y, g.select(c, zero, A))
        elif op == "call":
            for k, a in enumerate(args):
                part = _call_tangent(g, node, [one if j == k else None for j in range(len(args))])
                if part is not None:
                    acc(a, _mul(g, part, A))
        else:  # pragma: no cover
            raise AssertionError(op)
    return [adj.get(w.id) for w in wrt]


def _ancestors(outputs):
    seen = set()
    stack = list(outputs)
    while stack:
        n = stack.pop()
        if n.id in seen:
            continue
        seen.add(n.id)
        stack.extend(n.args)
    return seen


def _add(g, a, b):
    if a is None:
        return b
    if b is None:
        return a
    return g.add(a, b)


def _mul(g, a, b):
    """a*b with a unit factor folded (exact)."""
    if a.op == "c" and a.value == 1.0:
        return b
    if b.op == "c" and b.value == 1.0:
        return a
    return g.mul(a, b)


def _call_tangent(g, node, dargs):
    fn = node.fn
    a = node.args[0]
    da = dargs[0]
    if fn in _UNARY_CALLS:
        if da is None:
            return None
        if fn == "exp":
            return _mul(g, node, da)
        if fn == "log":
            return g.div(da, a)
        if fn == "sqrt":
            return g.div(da, g.mul(g.const(2.0), node))
        if fn == "sin":
            return _mul(g, g.call("cos", a), da)
        if fn == "cos":
            return _mul(g, g.neg(g.call("sin", a)), da)
        if fn == "tan":
            return _mul(g, g.add(g.const(1.0), g.mul(node, node)), da)
        if fn == "tanh":
            return _mul(g, g.sub(g.const(1.0), g.mul(node, node)), da)
        if fn == "sinh":
            return _mul(g, g.call("cosh", a), da)
        if fn == "cosh":
            return _mul(g, g.call("sinh", a), da)
        if fn == "atan":
            return g.div(da, g.add(g.const(1.0), g.mul(a, a)))
        if fn == "asin":
            return g.div(da, g.call("sqrt", g.sub(g.const(1.0), g.mul(a, a))))
        if fn == "acos":
            return g.neg(g.div(da, g.call("sqrt", g.sub(g.const(1.0), g.mul(a, a)))))
        if fn == "log10":
            return g.div(da, g.mul(a, g.const(math.log(10.0))))
        if fn == "log2":
            return g.div(da, g.mul(a, g.const(math.log(2.0))))
        if fn == "expm1":
            return _mul(g, g.add(node, g.const(1.0)), da)
        if fn == "log1p":
            return g.div(da, g.add(g.const(1.0), a))
        if fn == "fabs":
            return _mul(g, g.call("copysign", g.const(1.0), a), da)
        if fn == "erf":
            k = g.const(2.0 / math.sqrt(math.pi))
            return _mul(g, g.mul(k, g.call("exp", g.neg(g.mul(a, a)))), da)
        raise AssertionError(fn)
    b = node.args[1]
    db = dargs[1]
    if fn == "pow":
        out = None
        if da is not None:
            # d a^b / da = b * a^(b-1)
            if b.op == "c":
                coeff = g.mul(b, g.call("pow", a, g.const(b.value - 1.0)))
            else:
                coeff = g.mul(b, g.call("pow", a, g.sub(b, g.const(1.0))))
            out = _mul(g, coeff, da)
        if db is not None:
            out = _add(g, out, _mul(g, g.mul(node, g.call("log", a)), db))
        return out
    if fn == "atan2":
        # d atan2(a, b) = (b da - a db) / (a^2 + b^2)
        den = g.add(g.mul(a, a), g.mul(b, b))
        num = None
        if da is not None:
            num = _mul(g, b, da)
        if db is not None:
            t2 = _mul(g, a, db)
            num = g.neg(t2) if num is None else g.sub(num, t2)
        return None if num is None else g.div(num, den)
    if fn == "copysign":
        if da is None:
            return None
        return _mul(g, g.call("copysign", g.const(1.0), g.mul(a, b)), da)
    raise AssertionError(fn)


# ---------------------------------------------------------------------------
# Host interpreter (same rounding as the emitted code)
# ---------------------------------------------------------------------------
def evaluate(g: Graph, nodes, y, t, p, extra=None):
    """Evaluate ``nodes`` on the host in IEEE double, one rounding per operation
    exactly as the emitted code (compiled with contraction off) performs them.
    ``None`` entries evaluate to 0.0. Transcendentals use the host libm.
    ``extra`` maps the names of ``in`` inputs to their vectors."""
    vals: dict[int, float] = {}
    live = _ancestors([n for n in nodes if n is not None])
    for node in g.nodes:
        if node.id not in live:
            continue
        op = node.op
        if op == "in":
            v = float(extra[node.value[0]][node.value[1]])
        elif op == "y":
            v = float(y[node.value])
        elif op == "p":
            v = float(p[node.value])
        elif op == "t":
            v = float(t)
        elif op == "c":
            v = node.value
        else:
            a = [vals[x.id] for x in node.args]
            if op == "add":
                v = a[0] + a[1]
            elif op == "sub":
                v = a[0] - a[1]
            elif op == "mul":
                v = a[0] * a[1]
            elif op == "div":
                v = a[0] / a[1] if a[1] != 0.0 else math.copysign(math.inf, a[0]) * math.copysign(1.0, a[1]) if a[0] != 0.0 else math.nan
            elif op == "neg":
                v = -a[0]
            elif op == "cmp":
                v = float(eval(f"a[0] {_CMP[node.fn]} a[1]"))  # noqa: S307
            elif op == "select":
                v = a[1] if a[0] else a[2]
            elif op == "call":
                impl = _UNARY_CALLS.get(node.fn) or _BINARY_CALLS.get(node.fn)
                v = impl(*a)
            else:  # pragma: no cover
                raise AssertionError(op)
        vals[node.id] = v
    return [0.0 if n is None else vals[n.id] for n in nodes]


# ---------------------------------------------------------------------------
# Emission
# ---------------------------------------------------------------------------
def _lit(v: float) -> str:
    """An exact double literal."""
    if math.isnan(v):
        return "__builtin_nan(\"\")"
    if math.isinf(v):
        return "__builtin_inf()" if v > 0 else "(-__builtin_inf())"
    return float(v).hex() if v != 0.0 else ("-0.0" if math.copysign(1.0, v) < 0 else "0.0")


_CALL_NAMES = {
    "exp": "exp", "log": "log", "sqrt": "sqrt", "sin": "sin", "cos": "cos", "tan": "tan",
    "tanh": "tanh", "sinh": "sinh", "cosh": "cosh", "atan": "atan", "asin": "asin",
    "acos": "acos", "log10": "log10", "log2": "log2", "expm1": "expm1", "log1p": "log1p",
    "fabs": "fabs", "erf": "erf", "pow": "pow", "atan2": "atan2", "copysign": "copysign",
}


@dataclass
class FlopCount:
    """Operations in emitted code: ``+ - * /`` one each (a negation is a sign
    flip the hardware folds into an operand, counted as zero), transcendental
    calls counted apart."""

    flops: int = 0
    transcendentals: int = 0

    def add(self, other):
        return FlopCount(self.flops + other.flops, self.transcendentals + other.transcendentals)


class Emitter:
    """Emits straight-line C++ for a set of nodes of one graph.

    Inputs (``y``, ``p`` through the accessor, ``t``) are read once, just
    before their first use; every other node becomes one ``const double``
    temporary in graph order, so the source is SSA and the compiler keeps it in
    registers.
    """

    def __init__(self, g: Graph):
        self.g = g

    def body(self, roots, *, indent="  "):
        live = _ancestors([n for n in roots if n is not None])
        lines = []
        flops = 0
        trans = 0
        # Inputs are read where first used, not all at the top: in a long body
        # that keeps a state or parameter read from holding a register across
        # code that does not need it yet.
        inputs = {"y", "p", "t", "in"}
        pending = {n.id for n in self.g.nodes if n.id in live and n.op in inputs}

        def read_inputs(ids):
            for a in ids:
                if a in pending:
                    pending.discard(a)
                    inp = self.g.nodes[a]
                    if inp.op == "in":
                        expr = f"{inp.value[0]}[{inp.value[1]}]"
                    else:
                        expr = {"y": f"y[{inp.value}]", "p": f"p[{inp.value}]", "t": "t"}[inp.op]
                    lines.append(f"{indent}const double v{a} = {expr};")

        for node in self.g.nodes:
            if node.id not in live or node.op in inputs:
                continue
            read_inputs([a.id for a in node.args])
            name = f"v{node.id}"
            op = node.op
            if op == "c":
                expr = _lit(node.value)
            elif op in ("add", "sub", "mul", "div"):
                sym = {"add": "+", "sub": "-", "mul": "*", "div": "/"}[op]
                a, b = node.args
                expr = f"v{a.id} {sym} v{b.id}"
                flops += 1
            elif op == "neg":
                expr = f"-v{node.args[0].id}"
            elif op == "call":
                args = ", ".join(f"v{a.id}" for a in node.args)
                expr = f"{_CALL_NAMES[node.fn]}({args})"
                if node.fn in ("fabs", "copysign"):
                    pass
                else:
                    trans += 1
            elif op == "cmp":
                a, b = node.args
                lines.append(f"{indent}const bool {name} = v{a.id} {_CMP[node.fn]} v{b.id};")
                continue
            elif op == "select":
                c, a, b = node.args
                expr = f"v{c.id} ? v{a.id} : v{b.id}"
            else:  # pragma: no cover
                raise AssertionError(op)
            lines.append(f"{indent}const double {name} = {expr};")
        read_inputs(sorted(pending))  # roots that are inputs themselves
        return lines, FlopCount(flops, trans)


def value_of(node, zero="0.0"):
    return zero if node is None else f"v{node.id}"


@dataclass
class GeneratedSystem:
    """Device source for one callback (and, for Rodas5P, one Jacobian layout)."""

    source: str
    rhs_flops: FlopCount
    jac_flops: FlopCount
    n_colours: int
    uses_t: bool
    dt_zero: bool = False  # df/dt is identically zero (the callback never reads t)


def generate_rhs(traced: Traced):
    """``rhs(y, t, p, out)`` as straight-line double code; returns (source, flops)."""
    em = Emitter(traced.graph)
    lines, fc = em.body(traced.outputs)
    out_lines = [f"  out[{j}] = {value_of(n)};" for j, n in enumerate(traced.outputs)]
    src = "\n".join(
        [
            "template <class PA, class YA>",
            "__device__ __forceinline__ void rhs(const YA& y, const double t,",
            "                                    const PA& p, double* __restrict__ out) {",
            "  (void)t;",
            "  (void)p;",
        ]
        + lines
        + out_lines
        + ["}"]
    )
    return src, fc


def generate_neg_jac(traced: Traced, *, colour, n_colours, store, skip_zero=False):
    """``neg_jac(y, t, p, lu, dT)``: one forward sweep per colour writing
    ``lu[slot] = -J`` for every ``(row, colour, slot)`` in ``store`` (the order
    the reference writes them, modax/rodas5P.py:278-296), then the ``t``-seeded
    sweep into ``dT``. Returns (source, flops)."""
    g = traced.graph
    y_nodes = [g.y(i) for i in range(traced.n_vars)]
    sweeps = []
    for col in range(n_colours):
        seed = {y_nodes[c].id: 1.0 for c in range(traced.n_vars) if colour[c] == col}
        sweeps.append(tangents(g, traced.outputs, seed))
    dt_tan = tangents(g, traced.outputs, {g.t().id: 1.0})
    roots = [n for sw in sweeps for n in sw if n is not None] + [n for n in dt_tan if n is not None]
    lines, fc = Emitter(g).body(roots)
    # skip_zero: a slot whose derivative is structurally zero is left to a
    # clear of the whole buffer to -0.0 (the value its store would write)
    stores = [f"  lu[{slot}] = static_cast<LUT>(-{value_of(sweeps[col][row])});"
              for row, col, slot in store if not (skip_zero and sweeps[col][row] is None)]
    dT_lines = [f"  dT[{j}] = {value_of(n)};" for j, n in enumerate(dt_tan)]
    src = "\n".join(
        [
            "template <class PA, class LM, class YA>",
            "__device__ __forceinline__ void neg_jac(const YA& y, const double t,",
            "                                        const PA& p, LM lu,",
            "                                        double* __restrict__ dT) {",
            "  typedef typename ::mdx::lu_elem<LM>::type LUT;",
            "  (void)t;",
            "  (void)p;",
        ]
        + lines
        + stores
        + dT_lines
        + ["}"]
    )
    return src, fc


@dataclass
class SensitivityNodes:
    """The forward-sensitivity right-hand side as graph nodes.

    With the direction inputs ``s`` (state, n_vars) and ``pd`` (parameters,
    n_params) and the state increment ``v`` (n_vars):

      rhs[i]      = sum_j df_i/dy_j s_j + sum_m df_i/dp_m pd_m     (J_y s + J_p pd)
      dt[i]       = d rhs[i] / dt                                   (its df/dt)
      coupling[i] = sum_j d rhs[i] / dy_j v_j                        (L v)

    ``rhs`` is the sensitivity block of the joint system (modax/_sensitivity.py:
    198-243), ``dt`` its rows of the joint df/dt (rodas5P.py:400-436) and
    ``coupling`` the block-triangular stage correction (rodas5P.py:512-561).
    ``None`` marks a structurally zero entry.
    """

    rhs: list
    dt: list
    coupling: list


def sensitivity_nodes(traced: Traced) -> SensitivityNodes:
    g = traced.graph
    n, m = traced.n_vars, traced.n_params
    s = [g.inp("s", j) for j in range(n)]
    pd = [g.inp("pd", k) for k in range(m)]
    v = [g.inp("v", j) for j in range(n)]
    seed = {g.y(j).id: s[j] for j in range(n)}
    seed.update({g.p(k).id: pd[k] for k in range(m)})
    rhs = tangents(g, traced.outputs, seed)
    dt = tangents(g, rhs, {g.t().id: 1.0})
    coupling = tangents(g, rhs, {g.y(j).id: v[j] for j in range(n)})
    return SensitivityNodes(rhs, dt, coupling)


def augmented_trace(traced: Traced, direction_columns) -> Traced:
    """The joint ``[y, S]`` system as one traced callback of
    ``n_vars * (1 + n_dirs)`` components (modax/_sensitivity.py:198-243): for
    direction ``k`` the block ``z[n + k n : n + (k+1) n]`` is ``S_k`` and its
    right-hand side ``J_y S_k + J_p e_col`` (no parameter term for ``col < 0``).
    An explicit method integrates it as it stands."""
    g = traced.graph
    n = traced.n_vars
    outs = list(traced.outputs)
    zero = g.const(0.0)
    for k, col in enumerate(direction_columns):
        seed = {g.y(j).id: g.y(n + k * n + j) for j in range(n)}
        if col >= 0:
            seed[g.p(col).id] = 1.0
        outs += [zero if x is None else x for x in tangents(g, traced.outputs, seed)]
    return Traced(g, outs, n * (1 + len(direction_columns)), traced.n_params, traced.uses_t)


def generate_sensitivity(traced: Traced):
    """Device source of ``sens_rhs``, ``sens_dt`` and ``sens_coupling`` (see
    ``SensitivityNodes``); each takes the state through an accessor, ``s``,
    ``pd`` (and ``v``) through accessors, and writes ``out[0..n_vars)``
    (for ``sens_coupling`` a pointer or an accessor whose elements take an
    assignment of a double)."""
    sn = sensitivity_nodes(traced)
    g = traced.graph
    parts = []
    for name, nodes, extra in (("sens_rhs", sn.rhs, ""), ("sens_dt", sn.dt, ""),
                               ("sens_coupling", sn.coupling, ", const VA& v")):
        lines, _ = Emitter(g).body(nodes)
        outs = [f"  out[{j}] = {value_of(x)};" for j, x in enumerate(nodes)]
        tmpl = "template <class PA, class YA, class SA, class DA" + (", class VA, class OA>"
                                                                    if extra else ">")
        out_t = "OA out" if extra else "double* __restrict__ out"
        parts.append("\n".join(
            [tmpl,
             f"__device__ __forceinline__ void {name}(const YA& y, const double t, const PA& p,",
             f"    const SA& s, const DA& pd{extra}, {out_t}) {{",
             "  (void)t;", "  (void)p;", "  (void)s;", "  (void)pd;"]
            + lines + outs + ["}"]))
    return "\n\n".join(parts) + "\n"


def generate_jvp(traced: Traced):
    """``jvp(y, t, p, v, out)``: ``out[i] = sum_j df_i/dy_j v_j``, the Jacobian
    times a vector by one forward sweep seeded with ``v`` (read through an
    accessor), each ``out[i]`` assigned a double (a pointer or an accessor
    whose elements take the assignment). The refined FP32 stage solve uses it
    for its residual; returns (source, flops)."""
    g = traced.graph
    n = traced.n_vars
    nodes = tangents(g, traced.outputs, {g.y(j).id: g.inp("v", j) for j in range(n)})
    lines, fc = Emitter(g).body(nodes)
    outs = [f"  out[{j}] = {value_of(x)};" for j, x in enumerate(nodes)]
    src = "\n".join(
        ["template <class PA, class YA, class VA, class OA>",
         "__device__ __forceinline__ void jvp(const YA& y, const double t, const PA& p,",
         "                                    const VA& v, OA out) {",
         "  (void)t;", "  (void)p;", "  (void)v;"]
        + lines + outs + ["}"])
    return src + "\n", fc


@dataclass
class AdjointNodes:
    """The derivative contractions a discrete adjoint of Rodas5P needs, as
    gradients with respect to (y, p) -- ``y_*`` and ``p_*`` lists, ``None``
    where structurally zero -- of

      f_vjp:   w . f(y, t, p)                (VJP of a stage evaluation)
      jac_vjp: a . (J_y(y, t, p) b)          (the iteration matrix's dependence on y, p)
      dt_vjp:  c . df/dt(y, t, p)            (the df/dt term's dependence)

    with ``w``, ``a``, ``b``, ``c`` the named input vectors of n_vars entries."""

    f_y: list
    f_p: list
    jac_y: list
    jac_p: list
    dt_y: list
    dt_p: list


def adjoint_nodes(traced: Traced) -> AdjointNodes:
    g = traced.graph
    n, m = traced.n_vars, traced.n_params
    ys = [g.y(j) for j in range(n)]
    ps = [g.p(k) for k in range(m)]
    w = [g.inp("w", i) for i in range(n)]
    a = [g.inp("a", i) for i in range(n)]
    b = [g.inp("b", i) for i in range(n)]
    c = [g.inp("c", i) for i in range(n)]
    wrt = ys + ps
    f = adjoints(g, list(zip(traced.outputs, w)), wrt)
    jb = tangents(g, traced.outputs, {ys[j].id: b[j] for j in range(n)})
    jv = adjoints(g, list(zip(jb, a)), wrt)
    ft = tangents(g, traced.outputs, {g.t().id: 1.0})
    dv = adjoints(g, list(zip(ft, c)), wrt)
    return AdjointNodes(f[:n], f[n:], jv[:n], jv[n:], dv[:n], dv[n:])


def generate_adjoint(traced: Traced):
    """Device source of ``f_vjp``, ``jac_vjp`` and ``dt_vjp`` (see
    ``AdjointNodes``); each adds its gradient into ``ybar`` and ``pbar``."""
    an = adjoint_nodes(traced)
    g = traced.graph
    parts = []
    for name, extra, ys, ps in (("f_vjp", "const WA& w", an.f_y, an.f_p),
                                ("jac_vjp", "const WA& a, const WA& b", an.jac_y, an.jac_p),
                                ("dt_vjp", "const WA& c", an.dt_y, an.dt_p)):
        lines, _ = Emitter(g).body(ys + ps)
        outs = [f"  ybar[{j}] += {value_of(x)};" for j, x in enumerate(ys) if x is not None]
        outs += [f"  pbar[{k}] += {value_of(x)};" for k, x in enumerate(ps) if x is not None]
        parts.append("\n".join(
            ["template <class PA, class YA, class WA>",
             f"__device__ __forceinline__ void {name}(const YA& y, const double t, const PA& p,",
             f"    {extra}, double* __restrict__ ybar, double* __restrict__ pbar) {{",
             "  (void)t;", "  (void)p;", "  (void)ybar;", "  (void)pbar;"]
            + lines + outs + ["}"]))
    return "\n\n".join(parts) + "\n"


def adjoint_from_graph(traced: Traced, y, t, p, *, w=None, a=None, b=None, c=None):
    """Host evaluation of the three contractions' (ybar, pbar), for tests."""
    an = adjoint_nodes(traced)
    z = np.zeros(traced.n_vars)
    extra = {"w": z if w is None else w, "a": z if a is None else a,
             "b": z if b is None else b, "c": z if c is None else c}
    ev = lambda nodes: np.array(evaluate(traced.graph, nodes, y, t, p, extra))  # noqa: E731
    return {"f": (ev(an.f_y), ev(an.f_p)), "jac": (ev(an.jac_y), ev(an.jac_p)),
            "dt": (ev(an.dt_y), ev(an.dt_p))}


def sensitivity_from_graph(traced: Traced, y, t, p, s, pd, v=None):
    """Host evaluation of (rhs, dt, coupling), for tests."""
    sn = sensitivity_nodes(traced)
    extra = {"s": s, "pd": pd, "v": np.zeros(traced.n_vars) if v is None else v}
    ev = lambda nodes: np.array(evaluate(traced.graph, nodes, y, t, p, extra))  # noqa: E731
    return ev(sn.rhs), ev(sn.dt), ev(sn.coupling)


def generate(traced: Traced, *, colour, n_colours, store, skip_zero=False):
    """``rhs`` and ``neg_jac`` for ``traced`` (see the two generators)."""
    rhs_src, rhs_fc = generate_rhs(traced)
    jac_src, jac_fc = generate_neg_jac(traced, colour=colour, n_colours=n_colours, store=store,
                                       skip_zero=skip_zero)
    g = traced.graph
    dt_zero = all(n is None for n in tangents(g, traced.outputs, {g.t().id: 1.0}))
    return GeneratedSystem(
        source=rhs_src + "\n\n" + jac_src + "\n",
        rhs_flops=rhs_fc,
        jac_flops=jac_fc,
        n_colours=n_colours,
        uses_t=traced.uses_t,
        dt_zero=dt_zero,
    )


def structural_pattern(traced: Traced) -> np.ndarray:
    """``(n, n)`` mask of the state components each output reads."""
    n = traced.n_vars
    mask = np.zeros((n, n), dtype=bool)
    for r, out in enumerate(traced.outputs):
        for nid in _ancestors([out]):
            node = traced.graph.nodes[nid]
            if node.op == "y":
                mask[r, node.value] = True
    return mask


def jacobian_from_graph(traced: Traced, y, t, p, *, colour=None, n_colours=None):
    """Host evaluation of the coloured sweeps: dense ``(J, dT)``, for tests.

    Entry ``(r, c)`` is read from sweep ``colour[c]``, row ``r``, as the kernel
    reads it; the colouring must be structurally orthogonal for the pattern.
    """
    n = traced.n_vars
    g = traced.graph
    if colour is None:
        colour = list(range(n))
        n_colours = n
    reads = structural_pattern(traced)
    y_nodes = [g.y(i) for i in range(n)]
    J = np.zeros((n, n))
    for col in range(n_colours):
        seed = {y_nodes[c].id: 1.0 for c in range(n) if colour[c] == col}
        vals = evaluate(g, tangents(g, traced.outputs, seed), y, t, p)
        for c in range(n):
            if colour[c] == col:
                J[:, c] = np.where(reads[:, c], vals, 0.0)
    dT = np.array(evaluate(g, tangents(g, traced.outputs, {g.t().id: 1.0}), y, t, p))
    return J, dT


def rhs_from_graph(traced: Traced, y, t, p):
    return np.array(evaluate(traced.graph, traced.outputs, y, t, p))
