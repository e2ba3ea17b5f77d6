"""Plain-PyTorch restatement of a model's GRL1 step, for autograd references.

The model's right-hand side and linearized diagonal come from the same sympy
statements the code generator emits (goss_amd.codegen: the .ode frontend and
``analysis.linearize``); they are printed here as torch expressions over
node-batched FP64 tensors, and the GRL1 step (GRL1.h:72-84)

    y' = y + rhs * expm1(linear * h) / linear   if |linear| > delta,   y + rhs * h otherwise

is written with torch ops, so torch.autograd differentiates the scheme with
no hand-written derivative.  This is the reference the device adjoint is
checked against; it is test infrastructure.
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import sympy as sp
import torch
from sympy.core.relational import Relational

from goss_amd.codegen.analysis import linearize
from goss_amd.codegen.generate import make_variant
from goss_amd.codegen.ode_model import TIME


def _py(e, names: Dict[str, str]) -> str:
    """Python source of a sympy expression over torch tensors (fully parenthesised)."""
    if e is sp.true or e == sp.true:
        return "True"
    if e is sp.false or e == sp.false:
        return "False"
    if isinstance(e, sp.Symbol):
        return names.get(e.name, e.name)
    if e.is_Number:
        return repr(float(e))
    if e is sp.pi:
        return repr(float(sp.pi))
    if isinstance(e, sp.Add):
        return "(" + " + ".join(_py(a, names) for a in e.args) + ")"
    if isinstance(e, sp.Mul):
        return "(" + " * ".join(_py(a, names) for a in e.args) + ")"
    if isinstance(e, sp.Pow):
        b, x = e.args
        return f"({_py(b, names)} ** {_py(x, names)})"
    simple = {sp.exp: "torch.exp", sp.log: "torch.log", sp.Abs: "torch.abs", sp.floor: "torch.floor",
              sp.ceiling: "torch.ceil", sp.sign: "torch.sign", sp.sin: "torch.sin", sp.cos: "torch.cos",
              sp.tanh: "torch.tanh"}
    for cls, fn in simple.items():
        if isinstance(e, cls):
            return f"{fn}({_py(e.args[0], names)})"
    if isinstance(e, (sp.Max, sp.Min)):
        fn = "torch.maximum" if isinstance(e, sp.Max) else "torch.minimum"
        s = _py(e.args[0], names)
        for a in e.args[1:]:
            s = f"{fn}({s}, {_py(a, names)})"
        return s
    if isinstance(e, sp.Piecewise):
        pieces = list(e.args)
        s = _py(pieces[-1][0], names)
        for val, cond in reversed(pieces[:-1]):
            s = f"torch.where({_py(cond, names)}, {_py(val, names)}, {s})"
        return s
    if isinstance(e, Relational):
        return f"({_py(e.lhs, names)} {e.rel_op} {_py(e.rhs, names)})"
    if isinstance(e, sp.And):
        return "(" + " & ".join(_py(a, names) for a in e.args) + ")"
    if isinstance(e, sp.Or):
        return "(" + " | ".join(_py(a, names) for a in e.args) + ")"
    if isinstance(e, sp.Not):
        return f"(~{_py(e.args[0], names)})"
    raise ValueError(f"cannot print {type(e).__name__}")


class TorchModel:
    """rhs / linearized diagonal of one variant as batched torch code.

    ``linearized(y, t, p)``: y [N, S], p [N, P] (all parameters per node), t a
    float; returns (linear [N, S], rhs [N, S])."""

    def __init__(self, variant: str):
        var = make_variant(variant)
        model = var.model
        lin = linearize(model)
        self.S = model.num_states
        self.P = model.num_parameters
        self.linear_terms = torch.tensor([bool(x) for x in lin.linear_terms])
        self.field_parameter_indices = var.field_parameter_indices
        names: Dict[str, str] = {}
        for i, s in enumerate(model.state_names):
            names[s] = f"y[:, {i}]"
        for k, p in enumerate(model.parameter_names):
            names[p] = f"p[:, {k}]"
        names[TIME.name] = "tt"
        lines: List[str] = ["def f(y, tt, p):"]
        for n, e in model.body:
            lines.append(f"    {n}_ = {_py(e, names)}")
            names[n] = f"{n}_"
        for n, e in lin.lin_stmts:
            lines.append(f"    {n}_ = {_py(e, names)}")
            names[n] = f"{n}_"
        rhs = [names[model.derivatives[s]] for s in model.state_names]
        linear = [_py(e, names) for e in lin.linear]

        def as_col(x):
            return f"(({x}) + torch.zeros_like(y[:, 0]))"

        lines.append("    rhs = torch.stack([" + ", ".join(as_col(r) for r in rhs) + "], dim=1)")
        lines.append("    lin = torch.stack([" + ", ".join(as_col(a) for a in linear) + "], dim=1)")
        lines.append("    return lin, rhs")
        ns = {"torch": torch}
        exec("\n".join(lines), ns)
        self._f = ns["f"]
        self.defaults = torch.tensor([p.value for p in model.parameters], dtype=torch.float64)

    def linearized(self, y, t: float, p):
        tt = torch.tensor(float(t), dtype=torch.float64)
        return self._f(y, tt, p)

    def grl1_increment(self, at, t: float, h: float, p, delta: float = 1e-8):
        """GRL1 increment evaluated at `at` (GRL1.h:72-84): b/a (exp(a h) - 1) where |a| > delta, else b h."""
        lin, rhs = self.linearized(at, t, p)
        big = lin.abs() > delta
        safe = torch.where(big, lin, torch.ones_like(lin))
        return torch.where(big, rhs * torch.expm1(safe * h) / safe, rhs * h)

    def grl1_step(self, y, t: float, h: float, p, delta: float = 1e-8):
        return y + self.grl1_increment(y, t, h, p, delta)

    def grl2_step(self, y, t: float, h: float, p, delta: float = 1e-8):
        """GRL2 (GRL2.cpp:63-85): half step to y2, then the increment at y2 over h applied to y, both at t."""
        y2 = y + self.grl1_increment(y, t, 0.5 * h, p, delta)
        return y + self.grl1_increment(y2, t, h, p, delta)

    def rhs(self, y, t: float, p):
        return self.linearized(y, t, p)[1]

    def rl1_increment(self, at, t: float, h: float, p):
        """RL1 increment evaluated at `at` (RL1.h:68-77): exponential for the linear_terms states, Euler for
        the others."""
        lin, rhs = self.linearized(at, t, p)
        lt = self.linear_terms[None, :].expand_as(lin)
        safe = torch.where(lt, lin, torch.ones_like(lin))
        return torch.where(lt, rhs * torch.expm1(safe * h) / safe, rhs * h)

    def rl1_step(self, y, t: float, h: float, p):
        return y + self.rl1_increment(y, t, h, p)

    def rl2_step(self, y, t: float, h: float, p):
        """RL2 (RL2.cpp:75-85): as GRL2 with the RL1 increment."""
        y2 = y + self.rl1_increment(y, t, 0.5 * h, p)
        return y + self.rl1_increment(y2, t, h, p)

    def ee_step(self, y, t: float, h: float, p):
        """ExplicitEuler (ExplicitEuler.cpp:76-85)."""
        return y + h * self.rhs(y, t, p)

    def rk2_step(self, y, t: float, h: float, p):
        """RK2, the midpoint rule (RK2.cpp:73-91)."""
        k1 = self.rhs(y, t, p)
        return y + h * self.rhs(y + 0.5 * h * k1, t + 0.5 * h, p)

    def rk4_step(self, y, t: float, h: float, p):
        """RK4 (RK4.cpp:79-101)."""
        k1 = self.rhs(y, t, p)
        k2 = self.rhs(y + 0.5 * h * k1, t + 0.5 * h, p)
        k3 = self.rhs(y + 0.5 * h * k2, t + 0.5 * h, p)
        k4 = self.rhs(y + h * k3, t + h, p)
        return y + h * (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0

    def step(self, solver: str, y, t: float, h: float, p, delta: float = 1e-8):
        if solver == "GRL1":
            return self.grl1_step(y, t, h, p, delta)
        if solver == "GRL2":
            return self.grl2_step(y, t, h, p, delta)
        return {"RL1": self.rl1_step, "RL2": self.rl2_step, "ExplicitEuler": self.ee_step, "RK2": self.rk2_step,
                "RK4": self.rk4_step}[solver](y, t, h, p)

    def params(self, fp):
        """Per-node parameter matrix [N, P] with the field parameters fp [N, NFP] in their slots."""
        N = fp.shape[0]
        cols = [self.defaults[k].expand(N) for k in range(self.P)]
        for j, k in enumerate(self.field_parameter_indices):
            cols[k] = fp[:, j]
        return torch.stack(cols, dim=1)


def sampled_trace(model: TorchModel, y0, fp, t, sample_every: int, field_index: int, delta: float = 1e-8,
                  solver: str = "GRL1"):
    """`solver` over the times t (one step per interval); returns [n_samples, N] of state field_index."""
    p = model.params(fp)
    y = y0
    out = []
    for k in range(len(t) - 1):
        y = model.step(solver, y, float(t[k]), float(t[k + 1] - t[k]), p, delta)
        if (k + 1) % sample_every == 0:
            out.append(y[:, field_index])
    return torch.stack(out, dim=0)
