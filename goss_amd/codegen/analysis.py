"""Derived quantities of an :class:`ODEModel` needed by the solvers.

* ``linearized``: the diagonal of the Jacobian, ``linear[i] = d f_i / d x_i``,
  computed by forward-mode chain rule through the intermediates, so each
  derivative is written over named intermediates and named derivative
  intermediates (``d<name>_d<state>``) instead of a fully substituted tree.
  This is what gotran's ``linearized_derivatives`` emits for goss
  (``python/goss/codegeneration.py:584-620``; example ``cpp/tests/fhn.h:92-126``).
* ``linear_terms``: ``linear_terms[i] = 1`` when ``d f_i / d x_i`` is not
  zero and does not depend on ``x_i`` (``f_i`` affine in ``x_i`` with a
  nonzero slope).  Pinned by ``cpp/tests/fhn.h:54``,
  ``cpp/tests/beeler_reuter_1977.h:60-65`` and, for a derivative that does
  not contain its state (x' = y), ``cpp/tests/VDP.h:59-63``: the Rush-Larsen
  update divides by the slope (``RL1.h:76``).
* ``jacobian``: the full analytic Jacobian (same chain rule, all states).
* ``census``: operation counts of an emitted body.
"""
from __future__ import annotations

from collections import Counter
from dataclasses import dataclass
from typing import Dict, List, Sequence, Set, Tuple

import sympy as sp

from .ode_model import ODEModel, TIME


@dataclass
class Body:
    """An ordered list of assignments ``name = expr``; output slots are named
    ``rhs[i]``/``linear[i]``/``jac[i]`` by the emitters."""

    stmts: List[Tuple[str, sp.Expr]]


def _dependencies(model: ODEModel) -> Dict[str, Set[str]]:
    """State names each body entry depends on, transitively."""
    state_set = set(model.state_names)
    deps: Dict[str, Set[str]] = {}
    for name, expr in model.body:
        d: Set[str] = set()
        for s in expr.free_symbols:
            if s.name in state_set:
                d.add(s.name)
            elif s.name in deps:
                d |= deps[s.name]
        deps[name] = d
    return deps


def needed_names(model: ODEModel, roots: Sequence[sp.Expr]) -> List[str]:
    """Body entries needed to evaluate ``roots``, in body order."""
    bd = model.body_dict()
    need: Set[str] = set()
    stack = [s.name for r in roots for s in r.free_symbols]
    while stack:
        n = stack.pop()
        if n in need or n not in bd:
            continue
        need.add(n)
        stack.extend(s.name for s in bd[n].free_symbols)
    return [n for n, _ in model.body if n in need]


def _dname(name: str, state: str) -> str:
    return f"d{name}_d{state}"


@dataclass
class Linearized:
    """Diagonal derivatives.

    ``rhs_names``: intermediates needed by the RHS (body order).
    ``lin_stmts``: derivative intermediates ``(name, expr)`` in evaluation
    order, with the subset needed by the linear (affine) states first.
    ``linear``: the expression of ``linear[i]`` per state.
    ``linear_terms``: 0/1 flags.
    ``n_linear_stmts``: how many of ``lin_stmts`` precede the early return.
    """

    rhs_names: List[str]
    lin_stmts: List[Tuple[str, sp.Expr]]
    linear: List[sp.Expr]
    linear_terms: List[int]
    n_linear_stmts: int


def _chain_derivative(model: ODEModel, deps, state: str, root: str,
                      cache: Dict[Tuple[str, str], sp.Expr]) -> Tuple[sp.Expr, List[Tuple[str, sp.Expr]]]:
    """d(root)/d(state) through the intermediates.

    Returns the expression for the root derivative and the new derivative
    intermediates it references (in dependency order)."""
    bd = model.body_dict()
    x = model.symbols[state]
    order: List[Tuple[str, sp.Expr]] = []

    def visit(name: str) -> sp.Expr:
        key = (name, state)
        if key in cache:
            return cache[key]
        expr = bd[name]
        total = sp.diff(expr, x)
        for s in sorted(expr.free_symbols, key=lambda s: s.name):
            if s.name in bd and state in deps[s.name]:
                dsub = visit(s.name)
                if dsub == 0:
                    continue
                total = total + sp.diff(expr, s) * dsub
        total = _clean(total)
        if total.is_number or total.is_Symbol:
            res = total
        else:
            res = sp.Symbol(_dname(name, state), real=True)
            order.append((res.name, total))
        cache[key] = res
        return res

    expr = visit(root)
    return expr, order


def _clean(e: sp.Expr) -> sp.Expr:
    """Remove derivative artefacts sympy leaves for piecewise-smooth
    functions: d|x|/dx = sign(x), d floor/dx = 0 almost everywhere."""
    if e.has(sp.Derivative):
        e = e.replace(lambda a: isinstance(a, sp.Derivative) and isinstance(a.expr, (sp.floor, sp.ceiling)),
                      lambda a: sp.Integer(0))
    if e.has(sp.DiracDelta):
        e = e.replace(lambda a: isinstance(a, sp.DiracDelta), lambda a: sp.Integer(0))
    return e


def _transitive_state_deps(expr: sp.Expr, model: ODEModel, deps, extra: Dict[str, Set[str]]) -> Set[str]:
    state_set = set(model.state_names)
    out: Set[str] = set()
    for s in expr.free_symbols:
        if s.name in state_set:
            out.add(s.name)
        elif s.name in deps:
            out |= deps[s.name]
        elif s.name in extra:
            out |= extra[s.name]
    return out


def linearize(model: ODEModel) -> Linearized:
    deps = _dependencies(model)
    rhs = model.rhs_exprs()
    rhs_names = needed_names(model, [sp.Symbol(model.derivatives[s], real=True) for s in model.state_names])
    cache: Dict[Tuple[str, str], sp.Expr] = {}
    per_state: List[Tuple[sp.Expr, List[Tuple[str, sp.Expr]]]] = []
    extra_deps: Dict[str, Set[str]] = {}
    flags: List[int] = []
    for st in model.state_names:
        lin, stmts = _chain_derivative(model, deps, st, model.derivatives[st], cache)
        for nm, ex in stmts:
            extra_deps[nm] = _transitive_state_deps(ex, model, deps, extra_deps)
        per_state.append((lin, stmts))
        zero = lin.is_number and lin == 0
        flags.append(0 if zero or st in _transitive_state_deps(lin, model, deps, extra_deps) else 1)

    # Derivative intermediates: those reachable from linear (flag 1) states
    # first, then the rest; each list in creation (dependency) order.
    all_stmts: List[Tuple[str, sp.Expr]] = []
    for _, stmts in per_state:
        all_stmts.extend(stmts)
    by_name = dict(all_stmts)
    create_order = [n for n, _ in all_stmts]

    def closure(exprs):
        need: Set[str] = set()
        stack = [s.name for e in exprs for s in e.free_symbols]
        while stack:
            n = stack.pop()
            if n in need or n not in by_name:
                continue
            need.add(n)
            stack.extend(s.name for s in by_name[n].free_symbols)
        return need

    lin_need = closure([per_state[i][0] for i in range(len(flags)) if flags[i]])
    non_need = closure([per_state[i][0] for i in range(len(flags)) if not flags[i]]) - lin_need
    first = [(n, by_name[n]) for n in create_order if n in lin_need]
    second = [(n, by_name[n]) for n in create_order if n in non_need]
    return Linearized(rhs_names=rhs_names, lin_stmts=first + second,
                      linear=[p[0] for p in per_state], linear_terms=flags,
                      n_linear_stmts=len(first))


@dataclass
class Jacobian:
    """Full analytic Jacobian ``jac[i*S + j] = d f_i / d x_j``."""

    stmts: List[Tuple[str, sp.Expr]]
    entries: List[List[sp.Expr]]


def jacobian(model: ODEModel) -> Jacobian:
    deps = _dependencies(model)
    cache: Dict[Tuple[str, str], sp.Expr] = {}
    stmts: List[Tuple[str, sp.Expr]] = []
    S = model.num_states
    entries = [[sp.Integer(0)] * S for _ in range(S)]
    for i, si in enumerate(model.state_names):
        root = model.derivatives[si]
        for j, sj in enumerate(model.state_names):
            if sj not in deps[root]:
                continue
            e, new = _chain_derivative(model, deps, sj, root, cache)
            stmts.extend(new)
            entries[i][j] = e
    return Jacobian(stmts=stmts, entries=entries)


# ---------------------------------------------------------------------------
# Operation census
# ---------------------------------------------------------------------------

_CENSUS_KEYS = ("add", "mul", "div", "exp", "log", "sqrt", "pow", "floor", "abs", "select", "cmp")


def census_expr(e: sp.Expr, c: Counter, keep=None) -> None:
    """Count the operations the C printer emits for ``e``.

    The counting mirrors :mod:`goss_amd.codegen.printer`: a ``Mul`` with
    ``k`` numerator and ``m`` denominator factors costs ``k-1`` multiplies,
    ``m-1`` multiplies and one divide; integer powers 2 and 3 are multiplies;
    ``x**0.5`` is a square root."""
    from .printer import split_mul

    if e.is_Atom:
        return
    if keep is not None and not keep(e):
        return
    if isinstance(e, sp.Add):
        c["add"] += len(e.args) - 1
        for a in e.args:
            census_expr(a, c, keep)
        return
    if isinstance(e, sp.Mul):
        sign, num, den = split_mul(e)
        nn = len(num) + (0 if sign == 1 else 0)
        if num:
            c["mul"] += max(len(num) - 1, 0)
        if den:
            c["mul"] += max(len(den) - 1, 0)
            c["div"] += 1
        for a in num + den:
            census_expr(a, c, keep)
        return
    if isinstance(e, sp.Pow):
        b, x = e.args
        census_expr(b, c, keep)
        if x == -1:
            c["div"] += 1
        elif x.is_Integer and 2 <= abs(int(x)) <= 4:
            c["mul"] += abs(int(x)) - 1
            if x < 0:
                c["div"] += 1
        elif x == sp.Rational(1, 2):
            c["sqrt"] += 1
        elif x == sp.Rational(-1, 2):
            c["sqrt"] += 1
            c["div"] += 1
        else:
            c["pow"] += 1
            census_expr(x, c, keep)
        return
    if isinstance(e, sp.exp):
        c["exp"] += 1
    elif isinstance(e, sp.log):
        c["log"] += 1
    elif isinstance(e, (sp.floor, sp.ceiling)):
        c["floor"] += 1
    elif isinstance(e, sp.Abs):
        c["abs"] += 1
    elif isinstance(e, sp.Piecewise):
        c["select"] += len(e.args) - 1
        for (v, cond) in e.args:
            census_expr(v, c, keep)
            if cond is not sp.true:
                census_expr(cond, c, keep)
        return
    elif isinstance(e, sp.core.relational.Relational):
        c["cmp"] += 1
    for a in e.args:
        if isinstance(a, sp.Basic):
            census_expr(a, c, keep)


def census_body(exprs: Sequence[sp.Expr]) -> Dict[str, int]:
    c: Counter = Counter()
    for e in exprs:
        census_expr(e, c)
    return {k: int(c.get(k, 0)) for k in _CENSUS_KEYS}


def census_state_ops(stmts: Sequence[Tuple[str, sp.Expr]], outputs: Sequence[sp.Expr],
                     state_names: Sequence[str]) -> Dict[str, int]:
    """Operation counts per evaluation after common-subexpression elimination,
    counting only operations whose operands depend on a state.

    Subexpressions of parameters and time alone are excluded: they are
    constant per node (or per step) and any implementation evaluates them
    outside the per-node-step work.  This is the algorithmic work of one
    node-step that a roofline is priced on."""
    dep: Set[str] = set(state_names)
    for name, e in stmts:
        if any(s.name in dep for s in e.free_symbols):
            dep.add(name)
    exprs = [e for _, e in stmts] + list(outputs)
    rep, red = sp.cse(exprs)
    for sym, e in rep:
        if any(s.name in dep for s in e.free_symbols):
            dep.add(sym.name)
    c: Counter = Counter()

    def keep(e):
        return any(s.name in dep for s in e.free_symbols)

    for _, e in rep:
        census_expr(e, c, keep)
    for e in red:
        census_expr(e, c, keep)
    return {k: int(c.get(k, 0)) for k in _CENSUS_KEYS}


def dependence_pattern(stmts: Sequence[Tuple[str, sp.Expr]], outputs: Sequence[sp.Expr],
                       state_names: Sequence[str]) -> List[Set[int]]:
    """rows[j] = indices of the states that output j reads, through the
    statements as emitted (every operand of every statement counts, both arms
    of a piecewise included).  A forward difference in state i changes output
    j only if i is in rows[j]; every other entry of a difference-quotient
    Jacobian is exactly zero."""
    index = {s: k for k, s in enumerate(state_names)}
    deps: Dict[str, Set[int]] = {}

    def of(e: sp.Expr) -> Set[int]:
        d: Set[int] = set()
        for s in e.free_symbols:
            if s.name in index:
                d.add(index[s.name])
            elif s.name in deps:
                d |= deps[s.name]
        return d

    for name, e in stmts:
        deps[name] = of(e)
    return [of(e) for e in outputs]


def lu_fill_pattern(rows: Sequence[Set[int]]) -> List[Set[int]]:
    """Structure of the no-pivot LU factors of a matrix with structure rows
    plus the diagonal: entry (i, j) is nonzero if it is in the matrix or
    some r < min(i, j) has (i, r) and (r, j) nonzero (symbolic Gaussian
    elimination in the natural order)."""
    S = len(rows)
    pat = [set(r) | {i} for i, r in enumerate(rows)]
    for k in range(S):
        for i in range(k + 1, S):
            if k in pat[i]:
                pat[i] |= {j for j in pat[k] if j > k}
    return pat
