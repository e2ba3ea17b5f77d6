"""Frontend for the gotran ``.ode`` model language.

gotran itself is not a dependency: the ``.ode`` language is valid Python
syntax, so a file is read with :mod:`ast` and each expression is rebuilt as a
sympy expression over real symbols.  The result is an :class:`ODEModel` that
the emitters turn into C++ (CPU model plugin, ``ParameterizedODE`` contract of
``python/goss/codegeneration.py:51-114``) and HIP device functions.

Ordering rules (they fix state and parameter indices, which are part of the
plugin ABI):

* states are numbered in the order their time derivatives ``d<X>_dt`` are
  defined in the file.  This reproduces the gotran-generated orders held by the
  reference: ``cpp/tests/fhn.h:22-23``, ``cpp/tests/beeler_reuter_1977.h:25-32``
  and the TP06 state tuple at
  ``benchmarks/tentusscher_panfilov_2006_M_cell.py:13-33`` (``Ca_i`` before
  ``R_prime`` although ``R_prime`` is declared first);
* parameters are numbered in the order of their ``parameters(...)`` calls and,
  inside one call, sorted by name (ASCII), as in ``cpp/tests/fhn.h:27-37`` and
  ``benchmarks/tentusscher_panfilov_2006_M_cell.py:36-90``.

Supported language: ``parameters``, ``states``, ``expressions``,
``component``, ``ScalarParam``, ``Conditional``, ``ContinuousConditional``,
``Lt/Le/Gt/Ge/Eq/Ne``, ``And/Or/Not``, ``exp/log/sqrt/abs/floor/ceil``,
trigonometric and hyperbolic functions, ``**`` and the ``time`` (alias ``t``)
symbol, and two forms that gotran offers besides ``d<X>_dt = ...``:

* algebraic equations of a DAE, ``alg_<X>_0 = expr`` (the state ``X`` is
  defined by ``0 = expr``).  Such a state is not differential: its entry of the
  diagonal mass matrix is 0 (``ODE::_differential_states``, ``ODE.h:80-102``,
  which the generated plugin sets as ``codegeneration.py:452-459`` does).
  Explicit solvers refuse such a model (``GRL1.cpp:56-58`` and the other
  explicit solvers); the implicit ones put the mass matrix into the Newton
  residual and the iteration matrix (``ImplicitODESolver.cpp:140,258-262``);
* Markov chains, ``rates[A, B] = expr``: a transition from state ``A`` to
  state ``B`` at rate ``expr``.  Each state with rates and no derivative of its
  own gets ``dX_dt = sum_Y rates[Y, X] * Y - (sum_Y rates[X, Y]) * X`` (inflows
  in the order the rates are written, then the outflow), numbered after the
  explicitly defined derivatives in the order the states were declared.
"""
from __future__ import annotations

import ast
import math
import operator
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import sympy as sp

TIME = sp.Symbol("time", real=True)


class ODEParseError(ValueError):
    """Raised for ``.ode`` content outside the supported language."""


@dataclass
class Param:
    name: str
    value: float
    component: str


@dataclass
class State:
    name: str
    init: float
    component: str


@dataclass
class ODEModel:
    """A parsed cell model.

    ``body`` lists every intermediate and derivative assignment in file order
    as ``(name, expr)``; ``derivatives`` maps a state name to the name of its
    derivative intermediate.
    """

    name: str
    states: List[State]
    parameters: List[Param]
    body: List[Tuple[str, sp.Expr]]
    derivatives: Dict[str, str]
    symbols: Dict[str, sp.Symbol] = field(default_factory=dict)
    algebraic: List[str] = field(default_factory=list)  # states defined by alg_<X>_0 = expr

    @property
    def is_dae(self) -> bool:
        return bool(self.algebraic)

    @property
    def differential_states(self) -> List[int]:
        """Diagonal of the mass matrix: 1 for a differential state, 0 for an algebraic one."""
        return [0 if s.name in self.algebraic else 1 for s in self.states]

    @property
    def num_states(self) -> int:
        return len(self.states)

    @property
    def num_parameters(self) -> int:
        return len(self.parameters)

    @property
    def state_names(self) -> List[str]:
        return [s.name for s in self.states]

    @property
    def parameter_names(self) -> List[str]:
        return [p.name for p in self.parameters]

    def state_symbol(self, name: str) -> sp.Symbol:
        return self.symbols[name]

    def body_dict(self) -> Dict[str, sp.Expr]:
        return dict(self.body)

    def rhs_exprs(self) -> List[sp.Expr]:
        """Derivative of each state, written over states, parameters, time and
        intermediates (not substituted)."""
        bd = self.body_dict()
        return [bd[self.derivatives[s.name]] for s in self.states]


def _logistic(lhs, rhs, sigma, greater):
    # gotran's ContinuousConditional: a logistic switch of width ``sigma``.
    # Pinned by cpp/tests/fhn.h:83-85, which prints Ge(time, a) as
    # 1 - 1/(1 + exp((time - a)/sigma)) and Le(time, b) as
    # 1/(1 + exp((time - b)/sigma)).
    s = 1 / (1 + sp.exp((lhs - rhs) / sigma))
    return 1 - s if greater else s


class _Cond:
    """A relational or boolean condition kept apart from arithmetic."""

    def __init__(self, kind: str, args):
        self.kind = kind
        self.args = args

    def to_sympy(self):
        k, a = self.kind, self.args
        if k == "Lt":
            return sp.Lt(a[0], a[1], evaluate=False)
        if k == "Le":
            return sp.Le(a[0], a[1], evaluate=False)
        if k == "Gt":
            return sp.Gt(a[0], a[1], evaluate=False)
        if k == "Ge":
            return sp.Ge(a[0], a[1], evaluate=False)
        if k == "Eq":
            return sp.Eq(a[0], a[1], evaluate=False)
        if k == "Ne":
            return sp.Ne(a[0], a[1], evaluate=False)
        if k == "And":
            return sp.And(*[c.to_sympy() for c in a], evaluate=False)
        if k == "Or":
            return sp.Or(*[c.to_sympy() for c in a], evaluate=False)
        if k == "Not":
            return sp.Not(a[0].to_sympy(), evaluate=False)
        raise ODEParseError(k)


_BINOPS = {
    ast.Add: operator.add,
    ast.Sub: operator.sub,
    ast.Mult: operator.mul,
    ast.Div: operator.truediv,
    ast.Pow: operator.pow,
}

_FUNCS = {
    "exp": sp.exp,
    "log": sp.log,
    "ln": sp.log,
    "sqrt": sp.sqrt,
    "abs": sp.Abs,
    "Abs": sp.Abs,
    "fabs": sp.Abs,
    "floor": sp.floor,
    "ceil": sp.ceiling,
    "sin": sp.sin,
    "cos": sp.cos,
    "tan": sp.tan,
    "asin": sp.asin,
    "acos": sp.acos,
    "atan": sp.atan,
    "sinh": sp.sinh,
    "cosh": sp.cosh,
    "tanh": sp.tanh,
}


def _num(value) -> sp.Expr:
    if isinstance(value, bool):
        return sp.Integer(int(value))
    if isinstance(value, int):
        return sp.Integer(value)
    return sp.Float(float(value))


class _ExprBuilder:
    """Turns a Python ``ast`` expression into sympy, resolving names."""

    def __init__(self, names: Dict[str, sp.Expr]):
        self.names = names

    def build(self, node):
        if isinstance(node, ast.Constant):
            if isinstance(node.value, (int, float)):
                return _num(node.value)
            raise ODEParseError(f"unsupported constant {node.value!r}")
        if isinstance(node, ast.Name):
            if node.id in self.names:
                return self.names[node.id]
            if node.id in ("time", "t"):
                return TIME
            if node.id == "pi":
                return sp.pi
            if node.id == "E":
                return sp.E
            raise ODEParseError(f"undefined name {node.id!r} (line {node.lineno})")
        if isinstance(node, ast.UnaryOp):
            v = self.build(node.operand)
            if isinstance(node.op, ast.USub):
                return -v
            if isinstance(node.op, ast.UAdd):
                return v
            raise ODEParseError("unsupported unary operator")
        if isinstance(node, ast.BinOp):
            op = _BINOPS.get(type(node.op))
            if op is None:
                raise ODEParseError("unsupported binary operator")
            a, b = self.build(node.left), self.build(node.right)
            if isinstance(a, _Cond) or isinstance(b, _Cond):
                raise ODEParseError("arithmetic on a condition")
            return op(a, b)
        if isinstance(node, ast.Call):
            if not isinstance(node.func, ast.Name):
                raise ODEParseError("only plain function calls are supported")
            fname = node.func.id
            args = [self.build(a) for a in node.args]
            if fname in ("Lt", "Le", "Gt", "Ge", "Eq", "Ne"):
                if len(args) != 2:
                    raise ODEParseError(f"{fname} takes two arguments")
                return _Cond(fname, args)
            if fname in ("And", "Or"):
                return _Cond(fname, [a for a in args])
            if fname == "Not":
                return _Cond("Not", args)
            if fname == "Conditional":
                cond, tv, fv = args
                if not isinstance(cond, _Cond):
                    raise ODEParseError("Conditional needs a condition")
                return sp.Piecewise((tv, cond.to_sympy()), (fv, True))
            if fname == "ContinuousConditional":
                cond, tv, fv = args[:3]
                sigma = args[3] if len(args) > 3 else sp.Float(0.1)
                if not isinstance(cond, _Cond) or cond.kind not in ("Lt", "Le", "Gt", "Ge"):
                    raise ODEParseError("ContinuousConditional needs Lt/Le/Gt/Ge")
                s = _logistic(cond.args[0], cond.args[1], sigma, cond.kind in ("Gt", "Ge"))
                if fv == 0:
                    return tv * s
                return fv + (tv - fv) * s
            if fname in ("pow", "Pow"):
                return args[0] ** args[1]
            if fname in ("max", "Max"):
                return sp.Max(*args)
            if fname in ("min", "Min"):
                return sp.Min(*args)
            f = _FUNCS.get(fname)
            if f is None:
                raise ODEParseError(f"unsupported function {fname!r} (line {node.lineno})")
            return f(*args)
        if isinstance(node, ast.Compare):
            if len(node.ops) != 1:
                raise ODEParseError("chained comparisons are not supported")
            kinds = {ast.Lt: "Lt", ast.LtE: "Le", ast.Gt: "Gt", ast.GtE: "Ge", ast.Eq: "Eq", ast.NotEq: "Ne"}
            k = kinds.get(type(node.ops[0]))
            if k is None:
                raise ODEParseError("unsupported comparison")
            return _Cond(k, [self.build(node.left), self.build(node.comparators[0])])
        if isinstance(node, ast.Tuple):
            return tuple(self.build(e) for e in node.elts)
        raise ODEParseError(f"unsupported syntax {type(node).__name__} (line {getattr(node, 'lineno', '?')})")


def _literal_value(node) -> float:
    """Evaluate a parameter/state initial value: a number, arithmetic on
    numbers (``8/3``) or ``ScalarParam(value, ...)``."""
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id == "ScalarParam":
        if not node.args:
            raise ODEParseError("ScalarParam without a value")
        return _literal_value(node.args[0])
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
        return float(node.value)
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _literal_value(node.operand)
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.BinOp) and type(node.op) in _BINOPS:
        return float(_BINOPS[type(node.op)](_literal_value(node.left), _literal_value(node.right)))
    raise ODEParseError(f"unsupported initial value at line {getattr(node, 'lineno', '?')}")


def _component_path(call: ast.Call) -> str:
    parts = []
    for a in call.args:
        if isinstance(a, ast.Constant) and isinstance(a.value, str):
            parts.append(a.value)
        else:
            raise ODEParseError("component names must be string literals")
    return "/".join(parts)


def parse_ode_text(text: str, name: str) -> ODEModel:
    tree = ast.parse(text)
    params: List[Param] = []
    state_decl: Dict[str, State] = {}
    state_decl_order: List[str] = []
    body: List[Tuple[str, sp.Expr]] = []
    names: Dict[str, sp.Expr] = {}
    symbols: Dict[str, sp.Symbol] = {}
    derivative_order: List[str] = []
    derivatives: Dict[str, str] = {}
    algebraic: List[str] = []
    rates: Dict[str, List[Tuple[str, str]]] = {}  # state -> [(other state, rate name)], in and out
    component = name

    def declare(kind, call: ast.Call):
        comp = _component_path(call) or component
        entries = []
        for kw in call.keywords:
            if kw.arg is None:
                raise ODEParseError("**kwargs not supported")
            entries.append((kw.arg, _literal_value(kw.value)))
        if kind == "parameters":
            # gotran sorts the keyword arguments of one call by name.
            entries.sort(key=lambda e: e[0])
        for nm, val in entries:
            if nm in names or nm in ("time", "t"):
                raise ODEParseError(f"name {nm!r} declared twice")
            sym = sp.Symbol(nm, real=True)
            names[nm] = sym
            symbols[nm] = sym
            if kind == "parameters":
                params.append(Param(nm, val, comp))
            else:
                state_decl[nm] = State(nm, val, comp)
                state_decl_order.append(nm)

    for stmt in tree.body:
        if isinstance(stmt, ast.Expr) and isinstance(stmt.value, ast.Call) and isinstance(stmt.value.func, ast.Name):
            fn = stmt.value.func.id
            if fn in ("parameters", "states"):
                declare(fn, stmt.value)
            elif fn in ("expressions", "component"):
                component = _component_path(stmt.value) or name
            elif fn == "comment":
                pass
            else:
                raise ODEParseError(f"unsupported statement {fn!r} (line {stmt.lineno})")
        elif (isinstance(stmt, ast.Assign) and len(stmt.targets) == 1
              and isinstance(stmt.targets[0], ast.Subscript)):
            tgt = stmt.targets[0]
            sl = tgt.slice
            if not (isinstance(tgt.value, ast.Name) and tgt.value.id == "rates" and isinstance(sl, ast.Tuple)
                    and len(sl.elts) == 2 and all(isinstance(e, ast.Name) for e in sl.elts)):
                raise ODEParseError(f"unsupported assignment (line {stmt.lineno}); rates[A, B] = ... is the only "
                                    "subscripted form")
            a, b = sl.elts[0].id, sl.elts[1].id
            for x in (a, b):
                if x not in state_decl:
                    raise ODEParseError(f"rates[{a}, {b}]: {x!r} is not a state (line {stmt.lineno})")
            if a == b:
                raise ODEParseError(f"rates[{a}, {b}]: a transition needs two states (line {stmt.lineno})")
            rname = f"rates_{a}_{b}"
            if rname in names:
                raise ODEParseError(f"rates[{a}, {b}] given twice (line {stmt.lineno})")
            expr = _ExprBuilder(names).build(stmt.value)
            if isinstance(expr, (_Cond, tuple)):
                raise ODEParseError(f"rates[{a}, {b}] is not an arithmetic expression")
            body.append((rname, sp.sympify(expr)))
            sym = sp.Symbol(rname, real=True)
            names[rname] = sym
            symbols[rname] = sym
            rates.setdefault(a, []).append((b, rname, "out"))
            rates.setdefault(b, []).append((a, rname, "in"))
        elif isinstance(stmt, ast.Assign):
            if len(stmt.targets) != 1 or not isinstance(stmt.targets[0], ast.Name):
                raise ODEParseError(f"unsupported assignment (line {stmt.lineno})")
            lhs = stmt.targets[0].id
            if lhs in names:
                raise ODEParseError(f"{lhs!r} assigned twice (line {stmt.lineno})")
            expr = _ExprBuilder(names).build(stmt.value)
            if isinstance(expr, (_Cond, tuple)):
                raise ODEParseError(f"{lhs!r} is not an arithmetic expression")
            expr = sp.sympify(expr)
            body.append((lhs, expr))
            sym = sp.Symbol(lhs, real=True)
            names[lhs] = sym
            symbols[lhs] = sym
            st = None
            if lhs.startswith("d") and lhs.endswith("_dt") and lhs[1:-3] in state_decl:
                st = lhs[1:-3]
            elif lhs.startswith("alg_") and lhs.endswith("_0") and lhs[4:-2] in state_decl:
                st = lhs[4:-2]
                algebraic.append(st)
            if st is not None:
                if st in derivatives:
                    raise ODEParseError(f"second derivative or algebraic equation for {st!r}")
                derivatives[st] = lhs
                derivative_order.append(st)
        else:
            raise ODEParseError(f"unsupported statement at line {getattr(stmt, 'lineno', '?')}")

    # Markov chains: a derivative for every state with rates and none of its own.
    for st in state_decl_order:
        if st not in rates:
            continue
        if st in derivatives:
            raise ODEParseError(f"{st!r} has both rates and its own derivative")
        ins = [names[rn] * names[other] for other, rn, d in rates[st] if d == "in"]
        outs = [names[rn] for other, rn, d in rates[st] if d == "out"]
        expr = sp.Add(*ins, evaluate=False) if ins else sp.Integer(0)
        if outs:
            expr = expr - sp.Add(*outs, evaluate=False) * names[st]
        lhs = f"d{st}_dt"
        if lhs in names:
            raise ODEParseError(f"{lhs!r} is already a name")
        body.append((lhs, expr))
        sym = sp.Symbol(lhs, real=True)
        names[lhs] = sym
        symbols[lhs] = sym
        derivatives[st] = lhs
        derivative_order.append(st)

    missing = [s for s in state_decl_order if s not in derivatives]
    if missing:
        raise ODEParseError(f"states without a derivative: {missing}")
    states = [state_decl[s] for s in derivative_order]
    return ODEModel(name=name, states=states, parameters=params, body=body,
                    derivatives=derivatives, symbols=symbols, algebraic=algebraic)


def load_ode(path: str, name: Optional[str] = None) -> ODEModel:
    """Parse a ``.ode`` file; the model name defaults to the file stem."""
    with open(path) as f:
        text = f.read()
    if name is None:
        name = os.path.splitext(os.path.basename(path))[0]
    return parse_ode_text(text, name)
