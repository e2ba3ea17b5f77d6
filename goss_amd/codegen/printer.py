"""Deterministic C++ printer for model expressions.

One printer serves every target: the CPU model plugin, the CPU oracle and the
HIP device functions all print the same statement list with the same text, so
the only differences between targets are the binding prologue and the
implementation behind the ``m_*`` math helpers.  Literals are wrapped as
``real(<double>)`` so that the same text computes in ``float`` or ``double``
depending on what ``real`` names.

Printing rules (they fix the floating-point evaluation order):

* ``Add``: terms in sympy's canonical order, left to right;
* ``Mul``: numeric coefficient (one literal) first, then numerator factors;
  a denominator makes it ``m_div(numerator, denominator)`` (``m_rcp`` when
  the numerator is 1), IEEE division in FP64 and a hardware reciprocal in
  the FP32 kernels;
* ``x**2``, ``x**3``, ``x**4``: ``m_sq``/``m_cube``/``m_sq(m_sq(x))``;
  ``x**0.5``: ``m_sqrt``; other powers: ``m_pow``;
* ``exp(x)``: ``m_exp_l2(x log2(e), x)``, the first argument with log2(e) folded into the
  literals (the FP32 kernels' exp2 form; double uses x);
* ``Piecewise``: nested ``GOSS_SELECT(cond, a, b)`` (a ``?:`` select, or both
  arms evaluated and one selected, as goss_model_support.h defines it).
"""
from __future__ import annotations

from typing import Callable, Dict, List, Optional, Sequence, Tuple

import sympy as sp
from sympy.core.relational import Relational

RESERVED = {
    # C/C++ keywords and names that the generated code defines itself.
    "auto", "break", "case", "char", "const", "continue", "default", "do", "double", "else", "enum",
    "extern", "float", "for", "goto", "if", "int", "long", "register", "return", "short", "signed",
    "sizeof", "static", "struct", "switch", "typedef", "union", "unsigned", "void", "volatile",
    "while", "real", "states", "values", "linear", "rhs", "jac", "sp", "fp", "new", "delete",
    "this", "class", "template", "typename", "namespace", "operator", "private", "public",
    "protected", "virtual", "friend", "inline", "bool", "true", "false", "and", "or", "not",
}


def c_name(name: str) -> str:
    return name + "_" if name in RESERVED else name


def fmt_double(x: float) -> str:
    r = repr(float(x))
    if r in ("inf", "-inf", "nan"):
        raise ValueError("non-finite literal in model")
    if "e" not in r and "." not in r:
        r += ".0"
    return r


def split_mul(e: sp.Mul) -> Tuple[int, List[sp.Expr], List[sp.Expr]]:
    """Split a product into (sign, numerator factors, denominator factors).

    A numeric coefficient becomes one numerator literal (its absolute value)
    unless it is 1."""
    sign = 1
    num: List[sp.Expr] = []
    den: List[sp.Expr] = []
    coeff, factors = e.as_coeff_mul()
    if coeff.is_negative:
        sign = -1
        coeff = -coeff
    if coeff != 1:
        num.append(sp.Float(float(coeff)) if not coeff.is_Float else coeff)
    for f in factors:
        if isinstance(f, sp.Pow) and f.exp.is_number and f.exp.is_negative:
            ex = -f.exp
            den.append(f.base if ex == 1 else sp.Pow(f.base, ex, evaluate=False))
        else:
            num.append(f)
    return sign, num, den


LOG2E = 1.4426950408889634


def log2e_scaled(arg: sp.Expr) -> sp.Expr:
    """arg * log2(e) with the constant folded into arg's numeric coefficients.

    exp(x) = exp2(x log2 e).  The FP32 kernels evaluate exp2 in hardware; with log2 e folded into the
    literals of x (term by term over a sum) the scaling costs no multiply of its own.  Double and
    host code use x itself (``m_exp_l2``)."""
    if arg.is_Number:
        return sp.Float(float(arg) * LOG2E)
    if isinstance(arg, sp.Add):
        return sp.Add(*[log2e_scaled(t) for t in arg.args], evaluate=False)
    coeff, rest = arg.as_coeff_Mul()
    if coeff.is_Number:
        return sp.Mul(sp.Float(float(coeff) * LOG2E), rest, evaluate=False)
    return sp.Mul(sp.Float(LOG2E), arg, evaluate=False)


class CPrinter:
    """Print sympy expressions as C++ using ``real`` and ``m_*`` helpers."""

    PREC = {"atom": 100, "call": 100, "pow": 60, "mul": 50, "neg": 45, "add": 40, "cmp": 30,
            "and": 20, "or": 10, "select": 5}

    def __init__(self, name_map: Optional[Dict[str, str]] = None, lit_type: str = "real"):
        self.name_map = name_map or {}
        self.lit_type = lit_type

    def sym(self, s: sp.Symbol) -> str:
        if s.name in self.name_map:
            return self.name_map[s.name]
        return c_name(s.name)

    def lit(self, x) -> str:
        return f"{self.lit_type}({fmt_double(float(x))})"

    def p(self, e) -> str:
        return self._p(e)[0]

    def _wrap(self, e, min_prec: int, strict: bool = False) -> str:
        s, pr = self._p(e)
        if pr < min_prec or (strict and pr == min_prec):
            return f"({s})"
        return s

    def _p(self, e) -> Tuple[str, int]:
        P = self.PREC
        if e is sp.true or e == sp.true:
            return "true", P["atom"]
        if e is sp.false or e == sp.false:
            return "false", P["atom"]
        if isinstance(e, sp.Symbol):
            return self.sym(e), P["atom"]
        if e is sp.pi:
            return self.lit(float(sp.pi)), P["atom"]
        if e is sp.E:
            return self.lit(float(sp.E)), P["atom"]
        if e.is_Number:
            v = float(e)
            if v < 0:
                return f"-{self.lit(-v)}", P["neg"]
            return self.lit(v), P["atom"]
        if isinstance(e, sp.Add):
            terms = list(e.args)
            # Put a leading positive term first when the canonical first term
            # is negative (prints "a - b" rather than "-b + a").  The order is
            # fixed and the same for every target.
            out = ""
            for i, t in enumerate(terms):
                neg = t.could_extract_minus_sign() and (t.is_Mul or t.is_Number)
                body = -t if neg else t
                s = self._wrap(body, P["add"] + 1) if neg else self._wrap(body, P["add"])
                if i == 0:
                    out = f"-{self._wrap(body, P['mul'])}" if neg else s
                else:
                    out += f" - {s}" if neg else f" + {s}"
            return out, P["add"]
        if isinstance(e, sp.Mul):
            sign, num, den = split_mul(e)
            if not num:
                ns = self.lit(1.0)
            else:
                ns = "*".join(self._wrap(f, P["mul"]) for f in num)
            if den:
                ds = "*".join(self._wrap(f, P["mul"]) for f in den)
                s = f"m_rcp({ds})" if not num else f"m_div({ns}, {ds})"
                pr = P["call"]
            else:
                s = ns
                pr = P["mul"]
            if sign < 0:
                return f"-{s}", P["neg"]
            return s, pr
        if isinstance(e, sp.Pow):
            b, x = e.args
            bs = self.p(b)
            if x.is_Integer:
                n = int(x)
                if n == 2:
                    return f"m_sq({bs})", P["call"]
                if n == 3:
                    return f"m_cube({bs})", P["call"]
                if n == 4:
                    return f"m_sq(m_sq({bs}))", P["call"]
                if n == -1:
                    return f"m_rcp({bs})", P["call"]
                if n in (-2, -3, -4):
                    inner = {2: f"m_sq({bs})", 3: f"m_cube({bs})", 4: f"m_sq(m_sq({bs}))"}[-n]
                    return f"m_rcp({inner})", P["call"]
            if x == sp.Rational(1, 2):
                return f"m_sqrt({bs})", P["call"]
            if x == sp.Rational(-1, 2):
                return f"m_rcp(m_sqrt({bs}))", P["call"]
            return f"m_pow({bs}, {self.p(x)})", P["call"]
        if isinstance(e, sp.exp):
            arg = e.args[0]
            return f"m_exp_l2({self.p(log2e_scaled(arg))}, {self.p(arg)})", P["call"]
        if isinstance(e, sp.log):
            if len(e.args) != 1:
                raise ValueError("log with base")
            return f"m_log({self.p(e.args[0])})", P["call"]
        if isinstance(e, sp.Abs):
            return f"m_fabs({self.p(e.args[0])})", P["call"]
        if isinstance(e, sp.floor):
            return f"m_floor({self.p(e.args[0])})", P["call"]
        if isinstance(e, sp.ceiling):
            return f"m_ceil({self.p(e.args[0])})", P["call"]
        if isinstance(e, sp.sign):
            return f"m_sign({self.p(e.args[0])})", P["call"]
        if isinstance(e, sp.Heaviside):
            a = self.p(e.args[0])
            return f"m_heaviside({a})", P["call"]
        if isinstance(e, (sp.Max, sp.Min)):
            fn = "m_fmax" if isinstance(e, sp.Max) else "m_fmin"
            args = list(e.args)
            s = self.p(args[0])
            for a in args[1:]:
                s = f"{fn}({s}, {self.p(a)})"
            return s, P["call"]
        for cls, fn in ((sp.sin, "m_sin"), (sp.cos, "m_cos"), (sp.tan, "m_tan"), (sp.asin, "m_asin"),
                        (sp.acos, "m_acos"), (sp.atan, "m_atan"), (sp.sinh, "m_sinh"),
                        (sp.cosh, "m_cosh"), (sp.tanh, "m_tanh")):
            if isinstance(e, cls):
                return f"{fn}({self.p(e.args[0])})", P["call"]
        if isinstance(e, sp.Piecewise):
            pieces = list(e.args)
            # Last piece with a True condition is the else branch.
            val, cond = pieces[-1]
            if cond is not sp.true and cond != True:  # noqa: E712
                raise ValueError("Piecewise without an otherwise branch")
            s = self.p(val)
            for val, cond in reversed(pieces[:-1]):
                s = f"GOSS_SELECT({self.p(cond)}, {self.p(val)}, {s})"
            return s, P["atom"]
        if isinstance(e, Relational):
            ops = {"<": "<", "<=": "<=", ">": ">", ">=": ">=", "==": "==", "!=": "!="}
            op = ops[e.rel_op]
            return f"{self._wrap(e.lhs, P['add'])} {op} {self._wrap(e.rhs, P['add'])}", P["cmp"]
        if isinstance(e, sp.And):
            return " && ".join(self._wrap(a, P["and"] + 1) for a in e.args), P["and"]
        if isinstance(e, sp.Or):
            return " || ".join(self._wrap(a, P["or"] + 1) for a in e.args), P["or"]
        if isinstance(e, sp.Not):
            return f"!({self.p(e.args[0])})", P["call"]
        raise ValueError(f"cannot print {type(e).__name__}: {e}")


def assignments(printer: CPrinter, stmts: Sequence[Tuple[str, sp.Expr]], indent: str,
                decl: str = "const real ") -> List[str]:
    return [f"{indent}{decl}{c_name(n)} = {printer.p(e)};" for n, e in stmts]
