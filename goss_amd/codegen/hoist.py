"""Hoisting of parameter-only subexpressions.

Every subexpression whose operands are shared (non-field) parameters only --
``R*T/F``, ``exp(-0.056*72)``, ``sqrt(K_o/5.4)`` -- is the same for every node
and every step.  This pass replaces each maximal such subexpression by a
named constant ``c<k>``; the constants are computed once per launch on the
host (in FP64, with the host math library) and passed to the kernels in the
kernel-argument block.  Inside a sum or a product that also has varying
operands, the constant operands are grouped into one constant.

The rewritten statement text is used by every target (device kernels, CPU
model class, CPU oracle), so the transformation does not change agreement
between them; it only removes work and live registers from the per-node
code.
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Set, Tuple

import sympy as sp

from .ode_model import TIME


class ConstantTable:
    def __init__(self, shared_params: Set[str]):
        self.shared = set(shared_params)
        self.by_expr: Dict[sp.Expr, sp.Symbol] = {}
        self.items: List[Tuple[sp.Symbol, sp.Expr]] = []
        # intermediates that became constants: name -> constant symbol
        self.alias: Dict[str, sp.Expr] = {}

    def is_const(self, e: sp.Expr) -> bool:
        if e.is_Number:
            return True
        fs = e.free_symbols
        return all(s.name in self.shared or s.name.startswith("c__") for s in fs) and TIME not in fs

    def constant(self, e: sp.Expr) -> sp.Expr:
        """Return a symbol for the constant expression e (deduplicated)."""
        if e.is_Number or isinstance(e, sp.Symbol):
            return e
        if isinstance(e, sp.Mul):
            coeff, rest = e.as_coeff_Mul()
            if coeff != 1 and coeff.is_Number and (rest.is_Symbol):
                return e  # a literal times one parameter: cheap, keep inline
        if e in self.by_expr:
            return self.by_expr[e]
        sym = sp.Symbol(f"c__{len(self.items)}", real=True)
        self.by_expr[e] = sym
        self.items.append((sym, e))
        return sym

    def rewrite(self, e: sp.Expr) -> sp.Expr:
        e = e.xreplace({sp.Symbol(k, real=True): v for k, v in self.alias.items()}) if self.alias else e
        return self._rw(e)

    def _rw(self, e: sp.Expr) -> sp.Expr:
        if e.is_Atom:
            return e
        if self.is_const(e):
            return self.constant(e)
        if isinstance(e, (sp.Add, sp.Mul)):
            consts = [a for a in e.args if self.is_const(a)]
            others = [a for a in e.args if not self.is_const(a)]
            new_others = [self._rw(a) for a in others]
            if len(consts) >= 2 or (len(consts) == 1 and not consts[0].is_Number and not consts[0].is_Symbol):
                c = self.constant(e.func(*consts))
                return e.func(c, *new_others)
            return e.func(*(consts + new_others))
        if isinstance(e, sp.Piecewise):
            return sp.Piecewise(*[(self._rw(v), self._rw_cond(c)) for v, c in e.args])
        return e.func(*[self._rw(a) if isinstance(a, sp.Basic) else a for a in e.args])

    def _rw_cond(self, c):
        if c is sp.true or c is sp.false:
            return c
        if isinstance(c, (sp.And, sp.Or, sp.Not)):
            return c.func(*[self._rw_cond(a) for a in c.args])
        return c.func(self._rw(c.lhs), self._rw(c.rhs), evaluate=False)


def hoist_statements(stmts: Sequence[Tuple[str, sp.Expr]], table: ConstantTable) -> List[Tuple[str, sp.Expr]]:
    """Rewrite a statement list; statements that are entirely constant become
    aliases of constants and are dropped."""
    out = []
    for name, e in stmts:
        r = table.rewrite(e)
        if table.is_const(r) and not r.is_Number:
            table.alias[name] = table.constant(r)
            continue
        if r.is_Number:
            table.alias[name] = r
            continue
        out.append((name, r))
    return out
