"""Reverse-mode differentiation of a generated statement list.

The discrete adjoint of a GRL1 step (``GRL1.cpp:61-82``, ``GRL1.h:72-84``)

    y'_i = y_i + b_i * h * phi1(a_i h)    if |a_i| > delta,   y_i + b_i h otherwise,
    b = rhs(y, p),  a_i = d rhs_i / d y_i (the linearized diagonal),
    phi1(x) = (exp(x) - 1) / x,

needs the vector-Jacobian product of (y, p) -> (rhs, a) with the seeds
``bbar_i = lam_i * h * phi1(a_i h)`` and ``abar_i = lam_i * b_i * h^2 * phi1'(a_i h)``.
:func:`reverse_sweep` writes that product as straight-line accumulations over
the same named intermediates the forward code computes: walking the
statements in reverse order, each statement ``n = e`` adds ``n_b * de/dv`` to
the adjoint ``v_b`` of every differentiable symbol ``v`` in ``e``.  Symbols
that are shared parameters, host-computed constants, time, or intermediates
that depend on time only carry no adjoint.
"""
from __future__ import annotations

from typing import Dict, Iterable, List, Sequence, Set, Tuple

import sympy as sp

from .analysis import _clean


def adj_name(name: str) -> str:
    return f"{name}__b"


def _is_call(e: sp.Expr) -> bool:
    return isinstance(e, (sp.exp, sp.log)) or (isinstance(e, sp.Pow) and not e.exp.is_Integer)


def _neg_int_pow(f: sp.Expr) -> bool:
    return isinstance(f, sp.Pow) and f.exp.is_Integer and f.exp < 0


def lift_calls(stmts: Sequence[Tuple[str, sp.Expr]], prefix: str = "tr__") -> List[Tuple[str, sp.Expr]]:
    """Give every exp, log, non-integer power and reciprocal its own statement.

    The reverse sweep differentiates statement by statement; with the
    transcendental calls named, a derivative reuses the call's value
    (d exp(u)/du is the statement itself, d x^y/dx = y*n/x) instead of
    evaluating the call again inside every partial.  The denominator of a
    product (all its negative-integer-power factors together, as the C++
    printer divides once by their product) becomes ``dn = d1*d2*...`` and
    ``rc = 1/dn``, so the quotient is ``num*rc`` and every partial that
    needs ``1/dn^2`` reads ``rc*rc``: the reverse sweep takes no reciprocal
    of its own.  Identical calls and denominators share one statement."""
    seen: Dict[sp.Expr, sp.Symbol] = {}
    out: List[Tuple[str, sp.Expr]] = []

    def name(e: sp.Expr) -> sp.Symbol:
        sym = seen.get(e)
        if sym is None:
            sym = sp.Symbol(f"{prefix}{len(seen)}", real=True)
            seen[e] = sym
            out.append((sym.name, e))
        return sym

    def recip(den: sp.Expr) -> sp.Symbol:
        base = den if den.is_Symbol else name(den)
        return name(sp.Pow(base, -1))

    def lift(e):
        if not isinstance(e, sp.Basic) or e.is_Atom:
            return e
        if isinstance(e, sp.Mul):
            den = [f for f in e.args if _neg_int_pow(f)]
            num = [lift(f) for f in e.args if not _neg_int_pow(f)]
            if den:
                num.append(recip(sp.Mul(*[lift(f.base) ** (-f.exp) for f in den])))
            return sp.Mul(*num)
        if _neg_int_pow(e):
            return recip(lift(e.base) ** (-e.exp))
        new = e.func(*[lift(a) for a in e.args]) if e.args else e
        if isinstance(new, sp.Expr) and _is_call(new):
            return name(new)
        return new

    for n, e in stmts:
        out.append((n, lift(e)))
    return out


def reverse_sweep(stmts: Sequence[Tuple[str, sp.Expr]], seeds: Dict[str, str],
                  independent: Iterable[str],
                  initial: Iterable[str] = ()) -> Tuple[List[Tuple[str, str, sp.Expr]], Set[str]]:
    """Reverse accumulation over ``stmts`` (evaluation order).

    ``seeds`` maps statement names to the names of their seed adjoints (the
    outputs).  ``initial`` names statements whose adjoint the caller starts
    at a nonzero value (and which are swept like any other).
    ``independent`` are the leaf symbols to differentiate with respect to
    (states, field parameters).  Returns the accumulation list
    ``(target_adjoint, source_adjoint, partial)`` meaning
    ``target += source * partial``, in execution order, and the set of
    adjoint names that receive contributions."""
    independent = set(independent)
    names = {n for n, _ in stmts}
    diffable: Set[str] = set(independent)
    # Statements that depend on an independent symbol, transitively.
    for n, e in stmts:
        if any(s.name in diffable for s in e.free_symbols):
            diffable.add(n)
    has_adj: Dict[str, bool] = {n: True for n in seeds}
    for n in initial:
        has_adj[n] = True
    out: List[Tuple[str, str, sp.Expr]] = []
    for n, e in reversed(stmts):
        if not has_adj.get(n):
            continue
        src = seeds.get(n, adj_name(n))
        me = sp.Symbol(n, real=True)
        for v in sorted(e.free_symbols, key=lambda s: s.name):
            if v.name not in diffable:
                continue
            if v.name not in names and v.name not in independent:
                continue
            d = _clean(sp.diff(e, v))
            if d == 0:
                continue
            # Reuse the statement's own value inside its derivative
            # (d exp(u)/du = exp(u) is the statement itself).
            if not e.is_Symbol and not e.is_Number:
                d = d.subs(e, me)
            out.append((adj_name(v.name), src, d))
            has_adj[v.name] = True
    targets = {t for t, _, _ in out}
    return out, targets


def gate_states(body: Sequence[Tuple[str, sp.Expr]], rhs: Sequence[sp.Expr], state_names: Sequence[str],
                linear_terms: Sequence[int]) -> Dict[int, sp.Symbol]:
    """States whose derivative is ``(A - y_i) * r`` with A and r independent of
    y_i (Hodgkin-Huxley gates): index -> the symbol A.

    For such a state the GRL1 step is y' = A + (y - A) exp(a h) with
    a = -r the linearized diagonal (GRL1.h:72-84 with b = a (y - A)), so its
    adjoint seeds follow from that closed form.  ``body`` are the statements
    ``rhs`` refers to, in evaluation order."""
    deps: Dict[str, Set[str]] = {}
    states = set(state_names)
    for n, e in body:
        d: Set[str] = set()
        for sym in e.free_symbols:
            if sym.name in states:
                d.add(sym.name)
            d |= deps.get(sym.name, set())
        deps[n] = d
    out: Dict[int, sp.Symbol] = {}
    for i, (name, e) in enumerate(zip(state_names, rhs)):
        if not linear_terms[i]:
            continue
        y = sp.Symbol(name, real=True)
        if any(name in deps.get(sym.name, set()) for sym in e.free_symbols if sym != y):
            continue
        c1 = sp.diff(e, y)
        if c1 == 0 or c1.has(y):
            continue
        a = sp.simplify(-e.subs(y, 0) / c1)
        if isinstance(a, sp.Symbol) and a != y:
            out[i] = a
    return out
