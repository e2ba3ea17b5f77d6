"""Emitters: one model variant -> C++ source text.

A *variant* is a model plus the choice of field states and field parameters
(the same choice the reference makes when it generates a model class,
``python/goss/codegeneration.py:224-262``).  Field parameters are read per
node; every other parameter is shared by all nodes.

``emit_model_header`` writes the expression code once, as templates over the
arithmetic type ``real`` and a parameter accessor ``PA`` (``pa.template
get<k>()`` returns parameter ``k``).  The HIP kernels instantiate it with
``float``/``double`` and an accessor that serves field parameters from
per-node registers and shared ones from the kernel-argument block; the CPU
model class instantiates it with ``double`` and a member array.

``emit_parameterized_ode`` writes the C++ ``goss::ParameterizedODE`` subclass
with the class contract of ``python/goss/codegeneration.py:51-114`` (names,
``_param_to_value`` map, ``_linear_terms``, ``get_ic``, ``copy``,
``set_field_parameters`` and the ``extern "C" create_<Name>()`` factory).
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import sympy as sp

from .adjoint import gate_states, lift_calls, reverse_sweep
from .analysis import (Jacobian, Linearized, census_body, census_state_ops, dependence_pattern, jacobian, linearize,
                       needed_names)
from .hoist import ConstantTable, hoist_statements
from .ode_model import ODEModel, TIME
from .printer import CPrinter, assignments, c_name, fmt_double


@dataclass
class Variant:
    model: ODEModel
    name: str
    field_states: List[str] = field(default_factory=list)
    field_parameters: List[str] = field(default_factory=list)

    def __post_init__(self):
        for s in self.field_states:
            if s not in self.model.state_names:
                raise ValueError(f"{s} is not a state in the {self.model.name} ODE")
        for p in self.field_parameters:
            if p not in self.model.parameter_names:
                raise ValueError(f"{p} is not a parameter in the {self.model.name} ODE")

    @property
    def field_state_indices(self) -> List[int]:
        return [self.model.state_names.index(s) for s in self.field_states]

    @property
    def field_parameter_indices(self) -> List[int]:
        return [self.model.parameter_names.index(p) for p in self.field_parameters]

    @property
    def class_name(self) -> str:
        # gotran capitalises the model name (codegeneration.py:318).
        n = self.name
        return n[:1].upper() + n[1:]


def _used_symbols(exprs) -> set:
    u = set()
    for e in exprs:
        u |= {s.name for s in e.free_symbols}
    return u


@dataclass
class Emitted:
    header: str
    census: Dict[str, Dict[str, int]]
    linear_terms: List[int]
    jac_nnz: int
    num_constants: int
    structure_nnz: Tuple[int, int] = (0, 0)  # Jacobian and LU structure sizes (implicit solvers)


def _closure(stmts: Dict[str, sp.Expr], roots: Sequence[sp.Expr]) -> set:
    need: set = set()
    stack = [s.name for r in roots for s in r.free_symbols]
    while stack:
        n = stack.pop()
        if n in need or n not in stmts:
            continue
        need.add(n)
        stack.extend(s.name for s in stmts[n].free_symbols)
    return need


def _time_only(stmts: Sequence[Tuple[str, sp.Expr]], state_names: Sequence[str]) -> set:
    """Statements that depend on time (and parameters/constants) but on no
    state.  They are evaluated in the type of ``time`` (``treal``, FP64 on
    every target) so that single-precision kernels see the same time
    discontinuities (stimulus on/off) as the FP64 path."""
    state_set = set(state_names)
    ds: Dict[str, bool] = {}
    dt: Dict[str, bool] = {}
    out = set()
    for n, e in stmts:
        a = b = False
        for s in e.free_symbols:
            if s.name in state_set:
                a = True
            elif s == TIME:
                b = True
            elif s.name in ds:
                a = a or ds[s.name]
                b = b or dt[s.name]
        ds[n], dt[n] = a, b
        if b and not a:
            out.add(n)
    return out


def time_only_names(model: ODEModel) -> List[str]:
    """Intermediates of the (unhoisted) model that depend on time but on no
    state."""
    t = _time_only(model.body, model.state_names)
    return [n for n, _ in model.body if n in t and n not in model.derivatives.values()]


class _Gen:
    """Prints the hoisted statement lists of one variant."""

    def __init__(self, var: Variant, table: ConstantTable, stmts: Sequence[Tuple[str, sp.Expr]]):
        self.var = var
        self.model = var.model
        self.table = table
        self.pidx = {p.name: k for k, p in enumerate(self.model.parameters)}
        self.cidx = {sym.name: k for k, (sym, _) in enumerate(table.items)}
        self.time_only = _time_only(stmts, self.model.state_names)
        main_map = {n: f"real({c_name(n)})" for n in self.time_only}
        main_map["time"] = "real(time)"
        self.main = CPrinter(main_map)
        tmap = {p: f"pa.template getd<{k}>()" for p, k in self.pidx.items()}
        tmap.update({c: f"pa.template cstd<{k}>()" for c, k in self.cidx.items()})
        self.time = CPrinter(tmap, lit_type="treal")

    def bind(self, used: set, ind: str) -> List[str]:
        out = []
        for i, s in enumerate(self.model.states):
            if s.name in used:
                out.append(f"{ind}const real {c_name(s.name)} = states[{i}];")
        for p, k in self.pidx.items():
            if p in used:
                out.append(f"{ind}const real {c_name(p)} = pa.template get<{k}>();")
        for c, k in self.cidx.items():
            if c in used:
                out.append(f"{ind}const real {c} = pa.template cst<{k}>();")
        return out

    def lines(self, stmts: Sequence[Tuple[str, sp.Expr]], ind: str) -> List[str]:
        out = []
        for n, e in stmts:
            if n in self.time_only:
                out += assignments(self.time, [(n, e)], ind, decl="const treal ")
            else:
                out += assignments(self.main, [(n, e)], ind)
        return out

    def used_main(self, stmts, outputs) -> set:
        """Symbols referenced from `real` context (time-only statements read
        parameters and constants through getd/cstd)."""
        u = set()
        for n, e in stmts:
            if n not in self.time_only:
                u |= {s.name for s in e.free_symbols}
        for e in outputs:
            u |= {s.name for s in e.free_symbols}
        return u


def emit_model_header(var: Variant, with_jacobian: bool = True) -> Emitted:
    model = var.model
    lin = linearize(model)
    jac = jacobian(model) if with_jacobian else None
    S, P = model.num_states, model.num_parameters
    ind = "  "
    shared = {p.name for p in model.parameters} - set(var.field_parameters)

    # -- hoisting (one constant table for every function of the variant) ----
    table = ConstantTable(shared)
    hbody = hoist_statements(model.body, table)
    hbd = dict(hbody)

    def out_expr(name: str) -> sp.Expr:
        return hbd[name] if name in hbd else table.alias[name]

    rhs_out = [out_expr(model.derivatives[s]) for s in model.state_names]
    n_first = lin.n_linear_stmts
    hlin_first = hoist_statements(lin.lin_stmts[:n_first], table)
    hlin_second = hoist_statements(lin.lin_stmts[n_first:], table)
    hlinear = [table.rewrite(e) for e in lin.linear]
    if jac is not None:
        hjac_stmts = hoist_statements(jac.stmts, table)
        hjac = [[table.rewrite(e) for e in row] for row in jac.entries]
    derivs = set(model.derivatives.values())
    derivs_names = set(model.derivatives.values())

    all_stmts = list(hbody) + hlin_first + hlin_second + (hjac_stmts if jac is not None else [])
    g = _Gen(var, table, all_stmts)
    pr = g.main
    body_inter = [(n, e) for n, e in hbody if n not in derivs]

    def select(stmts, roots):
        need = _closure(dict(stmts), roots)
        return [(n, e) for n, e in stmts if n in need]

    L: List[str] = []
    L.append(f"// Generated by goss_amd.codegen from {model.name}.ode -- do not edit.")
    L.append(f"// Variant {var.name}: field states {var.field_states}, field parameters {var.field_parameters}.")
    L.append("#pragma once")
    L.append('#include "goss_model_support.h"')
    L.append("")
    L.append(f"namespace goss_models {{ namespace {var.name} {{")
    L.append("")
    L.append(f"constexpr int num_states = {S};")
    L.append(f"constexpr int num_parameters = {P};")
    L.append(f"constexpr int num_constants = {len(table.items)};")
    L.append(f"constexpr int num_field_states = {len(var.field_states)};")
    L.append(f"constexpr int num_field_parameters = {len(var.field_parameters)};")
    fsi = ", ".join(str(i) for i in var.field_state_indices) or "-1"
    fpi = ", ".join(str(i) for i in var.field_parameter_indices) or "-1"
    L.append(f"constexpr int field_state_indices[] = {{{fsi}}};")
    L.append(f"constexpr int field_parameter_indices[] = {{{fpi}}};")
    L.append("static const char* const state_names[] = {" + ", ".join(f'"{s}"' for s in model.state_names) + "};")
    L.append("static const char* const parameter_names[] = {" + ", ".join(f'"{p}"' for p in model.parameter_names) + "};")
    fsn = ", ".join(f'"{s}"' for s in var.field_states) or "nullptr"
    fpn = ", ".join(f'"{p}"' for p in var.field_parameters) or "nullptr"
    L.append(f"static const char* const field_state_names[] = {{{fsn}}};")
    L.append(f"static const char* const field_parameter_names[] = {{{fpn}}};")
    L.append("constexpr double default_parameters[] = {" + ", ".join(fmt_double(p.value) for p in model.parameters) + "};")
    L.append("constexpr double initial_conditions[] = {" + ", ".join(fmt_double(s.init) for s in model.states) + "};")
    L.append("constexpr unsigned char linear_terms[] = {" + ", ".join(str(f) for f in lin.linear_terms) + "};")
    L.append("// Diagonal of the mass matrix (ODE::_differential_states): 0 for the algebraic states of a DAE.")
    L.append("constexpr unsigned char differential_states[] = {"
             + ", ".join(str(f) for f in model.differential_states) + "};")
    L.append(f"constexpr bool is_dae = {'true' if model.is_dae else 'false'};")
    L.append("")

    # -- constants ------------------------------------------------------------
    cprinter = CPrinter({c: f"c[{k}]" for c, k in g.cidx.items()})
    cused = _used_symbols([e for _, e in table.items])
    L.append("// Subexpressions of shared (non-field) parameters only, evaluated once per launch.")
    L.append("// c[k] may use c[j] for j < k.  Called with real = double on the host.")
    L.append("template <typename real, typename PA>")
    L.append("GOSS_HD GOSS_INLINE void compute_constants(const PA& pa, real* __restrict__ c)")
    L.append("{")
    L.append(f"{ind}(void)pa; (void)c;")
    for p, k in g.pidx.items():
        if p in cused:
            L.append(f"{ind}const real {c_name(p)} = pa.template getd<{k}>();")
    for k, (sym, e) in enumerate(table.items):
        L.append(f"{ind}c[{k}] = {cprinter.p(e)};")
    L.append("}")
    L.append("")

    # -- eval ---------------------------------------------------------------
    eval_stmts = select(body_inter, rhs_out)
    used = g.used_main(eval_stmts, rhs_out)
    L.append("// Right-hand side: values[i] = d states[i] / dt.  Intermediates that depend on time but on")
    L.append("// no state are computed in the type of `time` (treal).")
    L.append("template <typename real, typename PA, typename treal>")
    L.append("GOSS_HD GOSS_INLINE void eval(const real* __restrict__ states, const treal time, const PA& pa,")
    L.append("                              real* __restrict__ values)")
    L.append("{")
    L.append(f"{ind}(void)time; (void)pa;")
    L += g.bind(used, ind)
    L += g.lines(eval_stmts, ind)
    for i in range(S):
        L.append(f"{ind}values[{i}] = {pr.p(rhs_out[i])};")
    L.append("}")
    L.append("")

    # -- eval_intermediates: every named intermediate (monitored values) -------
    inter_names = [n for n, _ in model.body if n not in derivs_names]
    inter_out = [out_expr(n) for n in inter_names]
    mon_stmts = select(body_inter, inter_out)
    mused = g.used_main(mon_stmts, inter_out)
    L.append(f"constexpr int num_intermediates = {len(inter_names)};")
    L.append("static const char* const intermediate_names[] = {" +
             (", ".join(f'"{n}"' for n in inter_names) or "nullptr") + "};")
    L.append("// Every named intermediate of the model at (states, time): the values a caller may monitor")
    L.append("// (ParameterizedODE::eval_monitored, ParameterizedODE.h:90-105, selects among them).")
    L.append("template <typename real, typename PA, typename treal>")
    L.append("GOSS_HD GOSS_INLINE void eval_intermediates(const real* __restrict__ states, const treal time, const PA& pa,")
    L.append("                                            real* __restrict__ out)")
    L.append("{")
    L.append(f"{ind}(void)time; (void)pa; (void)states; (void)out;")
    L += g.bind(mused, ind)
    L += g.lines(mon_stmts, ind)
    for k, e in enumerate(inter_out):
        L.append(f"{ind}out[{k}] = real({pr.p(e)});")
    L.append("}")
    L.append("")

    # -- structure of the Jacobian and of its LU factors (implicit solvers) ----
    struct_rows = dependence_pattern(eval_stmts, rhs_out, model.state_names)
    sL, struct_jnnz, struct_lnnz = emit_sparse_structure(S, struct_rows, ind, model.differential_states)
    L += sL

    # -- linearized_eval ----------------------------------------------------
    lin_roots = [e for _, e in hlin_first + hlin_second] + hlinear
    lin_inter = select(body_inter, rhs_out + lin_roots)
    lused = g.used_main(lin_inter + hlin_first + hlin_second, rhs_out + hlinear)
    L.append("// Linearized right-hand side: rhs[i] as in eval and linear[i] = d rhs[i] / d states[i].")
    L.append("// Entries with linear_terms[i] == 1 are written before the only_linear return.")
    L.append("template <typename real, typename PA, typename treal>")
    L.append("GOSS_HD GOSS_INLINE void linearized_eval(const real* __restrict__ states, const treal time, const PA& pa,")
    L.append("                                         real* __restrict__ linear, real* __restrict__ rhs,")
    L.append("                                         const bool only_linear)")
    L.append("{")
    L.append(f"{ind}(void)time; (void)pa;")
    L += g.bind(lused, ind)
    L += g.lines(lin_inter, ind)
    for i in range(S):
        L.append(f"{ind}rhs[{i}] = {pr.p(rhs_out[i])};")
    L += g.lines(hlin_first, ind)
    for i in range(S):
        if lin.linear_terms[i]:
            L.append(f"{ind}linear[{i}] = {pr.p(hlinear[i])};")
    L.append(f"{ind}if (only_linear) return;")
    L += g.lines(hlin_second, ind)
    for i in range(S):
        if not lin.linear_terms[i]:
            L.append(f"{ind}linear[{i}] = {pr.p(hlinear[i])};")
    L.append("}")
    L.append("")

    # Gates (rhs_i = (A - y_i) r): y'_i = A + (y_i - A) exp(linear_i h) under GRL1/RL1.
    gates = gate_states(hbody, rhs_out, model.state_names, lin.linear_terms)

    # -- linearized_eval_gates ------------------------------------------------
    gate_roots = [gates[i] for i in sorted(gates)]
    lg_inter = select(body_inter, rhs_out + lin_roots + gate_roots)
    lgused = g.used_main(lg_inter + hlin_first + hlin_second, rhs_out + hlinear + gate_roots)
    L.append("// linearized_eval (all entries) plus, for the gate states (gate_states[i] = 1: rhs[i] = (A - y_i) r),")
    L.append("// gate_target[i] = A.  A GRL stage of a gate is then y0_i + (y_i - A) expm1(linear[i] h).")
    L.append("constexpr unsigned char gate_states[] = {" + ", ".join("1" if i in gates else "0" for i in range(S)) + "};")
    L.append("template <typename real, typename PA, typename treal>")
    L.append("GOSS_HD GOSS_INLINE void linearized_eval_gates(const real* __restrict__ states, const treal time, const PA& pa,")
    L.append("                                               real* __restrict__ linear, real* __restrict__ rhs,")
    L.append("                                               real* __restrict__ gate_target)")
    L.append("{")
    L.append(f"{ind}(void)time; (void)pa; (void)gate_target;")
    L += g.bind(lgused, ind)
    L += g.lines(lg_inter, ind)
    for i in range(S):
        L.append(f"{ind}rhs[{i}] = {pr.p(rhs_out[i])};")
    L += g.lines(hlin_first + hlin_second, ind)
    for i in range(S):
        L.append(f"{ind}linear[{i}] = {pr.p(hlinear[i])};")
    for i in sorted(gates):
        L.append(f"{ind}gate_target[{i}] = {pr.p(gates[i])};")
    L.append("}")
    L.append("")

    # -- step adjoints: grl1_vjp, rl1_vjp, eval_vjp (and their tape forms) ----
    # The gates' seeds go to A, to linear_i and straight to lam_prev_i, and rhs_i is not needed.
    independent = list(model.state_names) + list(var.field_parameters)
    n_first = len(hlin_first)
    vjp_census: List[sp.Expr] = []
    tape_slots: Dict[str, int] = {}

    def emit_vjp(kind: str) -> None:
        """kind 'grl1': GRL1 step (GRL1.h:72-84); 'rl1': RL1 step (RL1.h:68-77: exponential update for
        the linear_terms states, Euler for the others, no delta test); 'eval': y_bar += (d rhs/d y)^T seed."""
        linear_states = [i for i in range(S) if lin.linear_terms[i]]
        gts = {} if kind == "eval" else gates
        if kind == "eval":
            exp_states: List[int] = []
        elif kind == "rl1":
            exp_states = linear_states
        else:
            exp_states = list(range(S))
        rhs_needed = [i for i in range(S) if i not in gts]
        lin_needed = exp_states
        roots = [rhs_out[i] for i in rhs_needed] + [hlinear[i] for i in lin_needed]
        lin_part = (hlin_first + hlin_second) if kind == "grl1" else (hlin_first if kind == "rl1" else [])
        lin_part = select(lin_part, roots)
        inter = select(body_inter, roots + [e for _, e in lin_part] + list(gts.values()))
        out_stmts = [(f"rhs__{i}", rhs_out[i]) for i in rhs_needed] + [(f"lin__{i}", hlinear[i]) for i in lin_needed]
        fwd = lift_calls(inter + lin_part + out_stmts)
        if kind == "eval":
            seeds = {f"rhs__{i}": f"seed__[{i}]" for i in range(S)}
        else:
            seeds = {f"rhs__{i}": f"bbar__{i}" for i in rhs_needed}
            seeds.update({f"lin__{i}": (f"gc__{i}" if i in gts else f"abar__{i}") for i in lin_needed})
        fwd_set = {n for n, _ in fwd}
        init_adj: Dict[str, List[str]] = {}
        for i, a_sym in gts.items():
            if a_sym.name in fwd_set or a_sym.name in independent:
                init_adj.setdefault(a_sym.name, []).append(f"ga__{i}")
        acc_lines, targets = reverse_sweep(fwd, seeds, independent, initial=[n for n in init_adj if n in fwd_set])
        targets = set(targets) | {f"{n}__b" for n in init_adj}
        for _, _, d in acc_lines:
            if d.has(sp.Derivative):
                raise ValueError(f"{model.name}: cannot differentiate {d}")
        vjp_census.extend([e for _, e in fwd] + [d for _, _, d in acc_lines])
        vused = g.used_main(fwd, [d for _, _, d in acc_lines] + list(gts.values()))
        gv = _Gen(var, table, fwd)
        delta_txt = "real(-1.0)" if kind == "rl1" else "delta__"
        fwd_def = dict(fwd)

        def inverse_of(e, depth=0):
            # 1/linear_i from the statements when linear_i = -1/tau: -tau, no reciprocal.
            if depth > 8:
                return None
            if isinstance(e, sp.Symbol) and e.name in fwd_def:
                return inverse_of(fwd_def[e.name], depth + 1)
            if isinstance(e, sp.Mul) and len(e.args) == 2 and e.args[0] == -1:
                inv = inverse_of(e.args[1], depth + 1)
                return None if inv is None else -inv
            if isinstance(e, sp.Pow) and e.exp == -1:
                return e.base
            return None

        def seed_lines(val, scalar_y) -> List[str]:
            """Seed computations; val(name) prints a forward value, scalar_y(i) a state."""
            out: List[str] = []
            for i in range(S):
                if i in gts:
                    out.append(f"{ind}real ey__{i}, ga__{i}, gc__{i};")
                    out.append(f"{ind}m_grl1_gate_seeds(real({val(f'lin__{i}')}), {scalar_y(i)}, real({val(gts[i].name, gts[i])}), "
                               f"lam__[{i}], h__, {delta_txt}, ey__{i}, ga__{i}, gc__{i});")
                elif i in exp_states:
                    inv = inverse_of(sp.Symbol(f"lin__{i}", real=True)) if val is None_val else None
                    inv_txt = gv.main.p(inv) if inv is not None else f"m_rcp(real({val(f'lin__{i}')}))"
                    out.append(f"{ind}real bbar__{i}, abar__{i};")
                    out.append(f"{ind}m_grl1_seeds(real({val(f'lin__{i}')}), {inv_txt}, real({val(f'rhs__{i}')}), lam__[{i}], "
                               f"h__, hinv__, {delta_txt}, bbar__{i}, abar__{i});")
                elif kind != "eval":
                    out.append(f"{ind}const real bbar__{i} = lam__[{i}]*h__;")
            return out

        def None_val(name, sym=None):
            return c_name(name) if sym is None else gv.main.p(sym)

        if kind == "eval":
            sig = ("GOSS_HD GOSS_INLINE void eval_vjp{T}(const real* __restrict__ states, const treal time, const PA& pa,\n"
                   "                                  const real* __restrict__ seed__, real* __restrict__ y_bar__,\n"
                   "                                  real* __restrict__ fp_bar__{TP})")
        else:
            sig = ("GOSS_HD GOSS_INLINE void " + kind + "_vjp{T}(const real* __restrict__ states, const treal time, "
                   "const PA& pa,\n                                  const real h__, const real delta__, "
                   "const real* __restrict__ lam__,\n                                  real* __restrict__ lam_prev__, "
                   "real* __restrict__ fp_bar__{TP})")

        def outputs(adj) -> List[str]:
            out: List[str] = []
            for i, st in enumerate(model.state_names):
                a_ = f"{st}__b"
                if kind == "eval":
                    if a_ in targets:
                        out.append(f"{ind}y_bar__[{i}] += {adj(a_)};")
                else:
                    direct = f"ey__{i}" if i in gts else f"lam__[{i}]"
                    out.append(f"{ind}lam_prev__[{i}] = {direct}" + (f" + {adj(a_)};" if a_ in targets else ";"))
            for j, fpn in enumerate(var.field_parameters):
                a_ = f"{fpn}__b"
                if a_ in targets:
                    out.append(f"{ind}fp_bar__[{j}] += {adj(a_)};")
            return out

        # Register form.
        L.append("template <typename real, typename PA, typename treal>")
        L.append(sig.format(T="", TP=""))
        L.append("{")
        L.append(f"{ind}(void)time; (void)pa; (void)fp_bar__;" + ("" if kind == "eval" else " (void)delta__;"))
        L.extend(gv.bind(vused | {s_.name for s_ in model.states}, ind))
        L.extend(gv.lines(fwd, ind))
        if any(i in exp_states and i not in gts for i in range(S)):
            L.append(f"{ind}const real hinv__ = m_rcp(h__);")
        L.extend(seed_lines(None_val, lambda i: gv.main.p(sp.Symbol(model.state_names[i], real=True))))
        for t_ in sorted(targets):
            init = " + ".join(init_adj.get(t_[:-3], [])) or "real(0.0)"
            L.append(f"{ind}real {t_} = {init};")
        # Common subexpressions of the partials become temporaries, each declared just before its first use.
        cse_rep, cse_red = sp.cse([d for _, _, d in acc_lines], symbols=sp.numbered_symbols("cs__", real=True))
        cse_def = {sym.name: e for sym, e in cse_rep}
        cse_order = [sym.name for sym, _ in cse_rep]
        emitted: set = set()

        def emit_temps(expr):
            need = [n for n in cse_order if n not in emitted and sp.Symbol(n, real=True) in expr.free_symbols]
            for n in need:
                emit_temps(cse_def[n])
                if n not in emitted:
                    emitted.add(n)
                    L.append(f"{ind}const real {n} = {gv.main.p(cse_def[n])};")

        for (tgt, src, _), d in zip(acc_lines, cse_red):
            emit_temps(d)
            L.append(f"{ind}{tgt} += {src}*({gv.main.p(d)});")
        L.extend(outputs(lambda a_: a_))
        L.append("}")
        L.append("")

        # Tape form: every intermediate and adjoint in a caller-provided tape (tape.get/put/add of
        # slot k); only the states, parameters and seeds stay in registers.
        fwd_names = [n for n, _ in fwd]
        slot = {n: k for k, n in enumerate(fwd_names)}
        adj_list = sorted(targets)
        aslot = {a_: len(fwd_names) + k for k, a_ in enumerate(adj_list)}
        tape_slots[kind] = len(fwd_names) + len(adj_list)
        tape_map = {n: f"tape.get({k})" for n, k in slot.items()}
        tp = CPrinter(tape_map)
        if kind == "eval":
            # seed__ and y_bar__ may be views (v[i] reads / v[i] += x), e.g. rows of the tape.
            L.append("template <typename real, typename PA, typename treal, typename TP, typename SV, typename YV>")
            L.append("GOSS_HD GOSS_INLINE void eval_vjp_tape(const real* __restrict__ states, const treal time, const PA& pa,\n"
                     "                                       const SV& seed__, const YV& y_bar__, real* __restrict__ fp_bar__,\n"
                     "                                       const TP& tape)")
        else:
            # lam__ and lam_prev__ may be views as well (rows of the tape).
            L.append("template <typename real, typename PA, typename treal, typename TP, typename LV, typename OV>")
            L.append("GOSS_HD GOSS_INLINE void " + kind + "_vjp_tape(const real* __restrict__ states, const treal time, "
                     "const PA& pa,\n                                  const real h__, const real delta__, const LV& lam__,\n"
                     "                                  const OV& lam_prev__, real* __restrict__ fp_bar__,\n"
                     "                                  const TP& tape)")
        L.append("{")
        L.append(f"{ind}(void)time; (void)pa; (void)fp_bar__;" + ("" if kind == "eval" else " (void)delta__;"))
        L.extend(gv.bind(vused | {s_.name for s_ in model.states}, ind))
        fwd_pr = CPrinter({"time": "real(time)", **tape_map})
        for n, e in fwd:
            if n in gv.time_only:
                L.append(f"{ind}const treal {c_name(n)} = {gv.time.p(e)};")
                L.append(f"{ind}tape.put({slot[n]}, real({c_name(n)}));")
            else:
                L.append(f"{ind}tape.put({slot[n]}, {fwd_pr.p(e)});")

        def tape_val(name, sym=None):
            if name in slot:
                return f"tape.get({slot[name]})"
            return gv.main.p(sym) if sym is not None else c_name(name)

        if any(i in exp_states and i not in gts for i in range(S)):
            L.append(f"{ind}const real hinv__ = m_rcp(h__);")
        L.extend(seed_lines(tape_val, lambda i: gv.main.p(sp.Symbol(model.state_names[i], real=True))))
        for a_ in adj_list:
            init = " + ".join(init_adj.get(a_[:-3], [])) or "real(0.0)"
            L.append(f"{ind}tape.put({aslot[a_]}, {init});")
        for tgt, src, d in acc_lines:
            src_txt = f"tape.get({aslot[src]})" if src in aslot else src
            L.append(f"{ind}tape.add({aslot[tgt]}, {src_txt}*({tp.p(d)}));")
        L.extend(outputs(lambda a_: f"tape.get({aslot[a_]})"))
        L.append("}")
        L.append("")

    L.append("// Discrete adjoints of one step of size h at `time` (one step per interval):")
    L.append("//   *_vjp:  lam_prev = (d y'/d y)^T lam,  fp_bar += (d y'/d p_field)^T lam;")
    L.append("//   grl1: y'_i = y_i + rhs_i*h*phi1(linear_i*h) when |linear_i| > delta, else y_i + rhs_i*h (GRL1.h:72-84);")
    L.append("//   rl1:  the same with the exponential update for the linear_terms states only (RL1.h:68-77);")
    L.append("//   for a gate (rhs_i = (A - y_i) r) the exponential update is A + (y_i - A) exp(linear_i*h).")
    L.append("//   eval_vjp: y_bar += (d rhs/d y)^T seed, fp_bar += (d rhs/d p_field)^T seed (the explicit schemes).")
    L.append("// Reverse accumulation over the generated statements; the *_tape forms keep every intermediate")
    L.append("// and adjoint in a caller-provided tape (for targets whose registers cannot hold the sweep).")
    for kind in ("grl1", "rl1", "eval"):
        emit_vjp(kind)
    L.append(f"constexpr int grl1_vjp_tape_slots = {tape_slots['grl1']};")
    L.append(f"constexpr int rl1_vjp_tape_slots = {tape_slots['rl1']};")
    L.append(f"constexpr int eval_vjp_tape_slots = {tape_slots['eval']};")
    L.append("")

    # Census of the (unhoisted) model: state-dependent work after CSE, and
    # the printed operations of the hoisted text.
    bd = model.body_dict()
    rhs_names = needed_names(model, [sp.Symbol(model.derivatives[s], real=True) for s in model.state_names])
    extra = [n for n in needed_names(model, [e for _, e in lin.lin_stmts] + lin.linear) if n not in rhs_names]
    rhs_stmts = [(n, bd[n]) for n, _ in model.body if n in set(rhs_names)]
    lin_stmts_all = [(n, bd[n]) for n, _ in model.body if n in set(rhs_names) | set(extra)] + lin.lin_stmts
    census = {
        # Per evaluation, after common-subexpression elimination, operations
        # whose operands depend on a state (the algorithmic work).
        "eval_state": census_state_ops(rhs_stmts, [], model.state_names),
        "linearized_eval_state": census_state_ops(lin_stmts_all, lin.linear, model.state_names),
        # Operations of the emitted (hoisted) text, before compiler CSE.
        "eval": census_body([e for _, e in eval_stmts] + rhs_out),
        "linearized_eval": census_body([e for _, e in lin_inter + hlin_first + hlin_second] + rhs_out + hlinear),
        "constants": census_body([e for _, e in table.items]),
        # Forward statements plus the reverse accumulations of grl1_vjp.
        "step_vjps": census_body(vjp_census),
    }

    jac_nnz = 0
    if jac is not None:
        jroots = [e for _, e in hjac_stmts] + [e for row in hjac for e in row]
        jinter = select(body_inter, jroots)
        jstmts = select(hjac_stmts, jroots)
        jused = g.used_main(jinter + jstmts, [e for row in hjac for e in row])
        L.append("// Analytic Jacobian: jac[i*num_states + j] = d values[i] / d states[j].")
        L.append("// `jac` is a pointer or any object whose jac[k] = value stores entry k.")
        L.append("template <typename real, typename PA, typename treal, typename JAC>")
        L.append("GOSS_HD GOSS_INLINE void compute_jacobian(const real* __restrict__ states, const treal time,")
        L.append("                                          const PA& pa, JAC jac)")
        L.append("{")
        L.append(f"{ind}(void)time; (void)pa;")
        L += g.bind(jused, ind)
        L += g.lines(jinter, ind)
        L += g.lines(jstmts, ind)
        for i in range(S):
            for j in range(S):
                e = hjac[i][j]
                if e != 0:
                    jac_nnz += 1
                L.append(f"{ind}jac[{i * S + j}] = {pr.p(e)};")
        L.append("}")
        L.append("")
        census["jacobian"] = census_body([e for _, e in jinter + jstmts] + [e for row in hjac for e in row])

        # The same entries over the structure of the Jacobian (column-major, the storage of the implicit
        # kernels), each stored right after the last statement it reads, so that few temporaries stay live.
        cols = [sorted(j for j in range(S) if i in struct_rows[j]) for i in range(S)]
        slot = {}
        for i in range(S):
            for j in cols[i]:
                slot[(j, i)] = len(slot)
        for i in range(S):
            for j in range(S):
                if hjac[i][j] != 0 and (i, j) not in slot:
                    raise ValueError(f"{var.name}: analytic Jacobian entry ({i}, {j}) outside the structure")
        stmts = jinter + jstmts
        pos = {n: k for k, (n, _) in enumerate(stmts)}
        after: Dict[int, List[str]] = {}
        for (i, j), k in sorted(slot.items(), key=lambda kv: kv[1]):
            e = hjac[i][j]
            last = max([pos[s.name] for s in e.free_symbols if s.name in pos], default=-1)
            after.setdefault(last, []).append(f"{ind}J[{k}] = {pr.p(e) if e != 0 else 'real(0.0)'};")
        L.append("// The analytic Jacobian over its structure: J[slot] for the entries of jac_cols / jac_col_start")
        L.append("// (d values[j] / d states[i] at jac_col_start[i] + rank of j in jac_cols[i]).  `J` is a pointer")
        L.append("// or any object whose J[k] = value stores entry k.")
        L.append("template <typename real, typename PA, typename treal, typename JS>")
        L.append("GOSS_HD GOSS_INLINE void compute_jacobian_structured(const real* __restrict__ states, const treal time,")
        L.append("                                                     const PA& pa, const JS& J)")
        L.append("{")
        L.append(f"{ind}(void)time; (void)pa;")
        L += g.bind(jused, ind)
        L += after.get(-1, [])
        for k, st in enumerate(stmts):
            L += g.lines([st], ind)
            L += after.get(k, [])
        L.append("}")
        L.append("")

    L.append("}}  // namespace goss_models::" + var.name)
    L.append("")
    return Emitted(header="\n".join(L), census=census, linear_terms=lin.linear_terms, jac_nnz=jac_nnz,
                   num_constants=len(table.items), structure_nnz=(struct_jnnz, struct_lnnz))


def emit_parameterized_ode(var: Variant, header_include: str) -> str:
    """C++ ``goss::ParameterizedODE`` subclass over the generated templates.

    Follows the class contract of ``python/goss/codegeneration.py:51-114``;
    parameters live in one member array whose slots ``_param_to_value``
    points at."""
    model = var.model
    N = var.class_name
    ns = f"goss_models::{var.name}"
    S, P = model.num_states, model.num_parameters
    L = []
    L.append(f"// Generated by goss_amd.codegen from {model.name}.ode -- do not edit.")
    L.append(f"#ifndef {N.upper()}_PLUGIN_H_IS_INCLUDED")
    L.append(f"#define {N.upper()}_PLUGIN_H_IS_INCLUDED")
    L.append("#include <memory>")
    L.append("#include <stdexcept>")
    L.append("#include <cmath>")
    L.append('#include "goss/ParameterizedODE.h"')
    L.append(f'#include "{header_include}"')
    L.append("")
    L.append("namespace goss {")
    L.append("")
    L.append(f"class {N} : public ParameterizedODE")
    L.append("{")
    L.append("  struct Acc {")
    L.append("    const double* v;")
    L.append("    const double* c;")
    L.append("    template <int k> double get() const { return v[k]; }")
    L.append("    template <int k> double getd() const { return v[k]; }")
    L.append("    template <int k> double cst() const { return c[k]; }")
    L.append("    template <int k> double cstd() const { return c[k]; }")
    L.append("  };")
    L.append(f"  static constexpr int K = {ns}::num_constants > 0 ? {ns}::num_constants : 1;")
    L.append("public:")
    L.append(f"  {N}() : ParameterizedODE({S}, {P}, {len(var.field_states)}, {len(var.field_parameters)}, 0)")
    L.append("  {")
    L.append(f"    for (int i = 0; i < {S}; ++i) _state_names[i] = {ns}::state_names[i];")
    L.append(f"    for (int i = 0; i < {P}; ++i) {{")
    L.append(f"      _parameter_names[i] = {ns}::parameter_names[i];")
    L.append(f"      _p[i] = {ns}::default_parameters[i];")
    L.append("    }")
    for i, s in enumerate(var.field_states):
        L.append(f'    _field_state_names[{i}] = "{s}";')
        L.append(f"    _field_state_indices[{i}] = {var.field_state_indices[i]};")
    for i, p in enumerate(var.field_parameters):
        L.append(f'    _field_parameter_names[{i}] = "{p}";')
    L.append("    _bind();")
    L.append(f"    for (int i = 0; i < {S}; ++i) _linear_terms[i] = {ns}::linear_terms[i];")
    if model.is_dae:
        # codegeneration.py:452-459: the algebraic states are not differential.
        for i, d in enumerate(model.differential_states):
            if not d:
                L.append(f"    _differential_states[{i}] = 0;")
        L.append("    _is_dae = true;")
    L.append("  }")
    L.append(f"  {N}(const {N}& ode) : ParameterizedODE(ode)")
    L.append("  {")
    L.append(f"    for (int i = 0; i < {P}; ++i) _p[i] = ode._p[i];")
    L.append("    analytic_jacobian = ode.analytic_jacobian;")
    L.append("    _bind();")
    L.append("  }")
    L.append("  // Parameters can change through _param_to_value at any time, so the")
    L.append("  // parameter-only constants are recomputed on every evaluation.")
    L.append("  void eval(const double* states, double time, double* values)")
    L.append("  {")
    L.append("    double c[K];")
    L.append(f"    {ns}::compute_constants<double>(Acc{{_p, nullptr}}, c);")
    L.append(f"    {ns}::eval<double>(states, time, Acc{{_p, c}}, values);")
    L.append("  }")
    L.append("  void linearized_eval(const double* states, double time, double* linear, double* rhs,")
    L.append("                       bool only_linear) const")
    L.append("  {")
    L.append("    double c[K];")
    L.append(f"    {ns}::compute_constants<double>(Acc{{_p, nullptr}}, c);")
    L.append(f"    {ns}::linearized_eval<double>(states, time, Acc{{_p, c}}, linear, rhs, only_linear);")
    L.append("  }")
    L.append("  void get_ic(goss::DoubleVector* values) const")
    L.append("  {")
    L.append("    values->n = _num_states;")
    L.append("    values->data.reset(new double[_num_states]);")
    L.append(f"    for (int i = 0; i < {S}; ++i) values->data[i] = {ns}::initial_conditions[i];")
    L.append("  }")
    L.append("  // ODE::compute_jacobian (ODE.h:71): the analytic Jacobian when analytic_jacobian is set, as a plugin")
    L.append("  // generated with generate_jacobian would have it (codegeneration.py:186, 622-641); otherwise the")
    L.append("  // forward differences of the base class, as gotran's default plugin.")
    L.append("  void compute_jacobian(double* states, double time, double* jac)")
    L.append("  {")
    L.append("    if (!analytic_jacobian) {")
    L.append("      ODE::compute_jacobian(states, time, jac);")
    L.append("      return;")
    L.append("    }")
    L.append("    double c[K];")
    L.append(f"    {ns}::compute_constants<double>(Acc{{_p, nullptr}}, c);")
    L.append(f"    {ns}::compute_jacobian<double>(states, time, Acc{{_p, c}}, jac);")
    L.append("  }")
    L.append("  bool analytic_jacobian = false;")
    L.append("  std::shared_ptr<ODE> copy() const")
    L.append(f"  {{ return std::make_shared<{N}>(*this); }}")
    L.append("  void eval_monitored(const double*, double, double*) const")
    L.append(f"  {{ throw std::runtime_error(\"No monitored in the '{N}' model.\"); }}")
    L.append("  void set_field_parameters(const double* field_params)")
    L.append("  {")
    for i, k in enumerate(var.field_parameter_indices):
        L.append(f"    _p[{k}] = field_params[{i}];")
    if not var.field_parameters:
        L.append("    (void)field_params;")
    L.append("  }")
    L.append("private:")
    L.append(f"  double _p[{P}];")
    L.append("  void _bind()")
    L.append("  {")
    L.append(f"    for (int i = 0; i < {P}; ++i) _param_to_value[_parameter_names[i]] = &_p[i];")
    L.append("  }")
    L.append("};")
    L.append("")
    L.append("}  // namespace goss")
    L.append("")
    L.append(f'extern "C" __attribute__((visibility("default"))) goss::ParameterizedODE* create_{N}()')
    L.append("{")
    L.append(f"  return new goss::{N};")
    L.append("}")
    L.append("#endif")
    L.append("")
    return "\n".join(L)


def variant_metadata(var: Variant, em: Emitted) -> dict:
    m = var.model
    return {
        "variant": var.name,
        "model": m.name,
        "class_name": var.class_name,
        "num_states": m.num_states,
        "num_parameters": m.num_parameters,
        "state_names": m.state_names,
        "parameter_names": m.parameter_names,
        "default_parameters": [p.value for p in m.parameters],
        "initial_conditions": [s.init for s in m.states],
        "field_states": var.field_states,
        "field_state_indices": var.field_state_indices,
        "field_parameters": var.field_parameters,
        "field_parameter_indices": var.field_parameter_indices,
        "linear_terms": em.linear_terms,
        "differential_states": m.differential_states,
        "is_dae": m.is_dae,
        "census": em.census,
        "jacobian_nnz": em.jac_nnz,
        "jacobian_structure_nnz": em.structure_nnz[0],
        "lu_structure_nnz": em.structure_nnz[1],
        "num_constants": em.num_constants,
    }


def emit_sparse_structure(S: int, rows: Sequence[set], ind: str = "  ",
                          differential: Sequence[int] = None) -> Tuple[List[str], int, int]:
    """Structure of the Jacobian (rows[j] = states output j reads) and of its
    no-pivot LU factors, with straight-line LU factorisation and substitution
    over that structure.  The code follows ODE::lu_factorize (ODE.cpp:69-100)
    and ODE::forward_backward_subst (ODE.cpp:102-132) term by term, leaving out
    only products with a structurally zero factor: for finite entries the
    results equal the dense routines' (a skipped term adds an exact zero)."""
    from .analysis import lu_fill_pattern

    lu = lu_fill_pattern(rows)
    # Jacobian storage: column-major over the structure.
    cols = [sorted(j for j in range(S) if i in rows[j]) for i in range(S)]
    jac_slot = {}
    for i in range(S):
        for j in cols[i]:
            jac_slot[(j, i)] = len(jac_slot)
    # LU storage: row-major over the fill structure.
    lu_slot = {}
    for j in range(S):
        for i in sorted(lu[j]):
            lu_slot[(j, i)] = len(lu_slot)
    jn, ln = len(jac_slot), len(lu_slot)
    L: List[str] = []
    L.append("// Structure of the Jacobian d values[j] / d states[i]: values[j] reads states[i] only for the")
    L.append("// entries below (through any operand of the emitted right-hand side), so every other entry of a")
    L.append("// forward-difference Jacobian is exactly zero.  And the structure of its no-pivot LU factors")
    L.append("// (the entries, the diagonal and the fill).")
    L.append(f"constexpr int jac_nnz = {jn};")
    L.append(f"constexpr int lu_nnz = {ln};")
    L.append("// jac_cols[i]: bit j set when values[j] reads states[i].  Jacobian storage is column-major over the")
    L.append("// structure: entry (j, i) at jac_col_start[i] + (number of bits of jac_cols[i] below j).")
    masks = [sum(1 << j for j in cols[i]) for i in range(S)]
    starts, acc = [], 0
    for i in range(S):
        starts.append(acc)
        acc += len(cols[i])
    L.append("constexpr unsigned jac_cols[] = {" + ", ".join(f"0x{m:x}u" for m in masks) + "};")
    L.append("constexpr int jac_col_start[] = {" + ", ".join(str(s) for s in starts) + "};")
    L.append("// LU storage is row-major over its structure: lu_slot[j * num_states + i], -1 outside.")
    L.append("constexpr short lu_slot[] = {" + ", ".join(str(lu_slot.get((j, i), -1)) for j in range(S)
                                                       for i in range(S)) + "};")
    L.append("")
    diff = list(differential) if differential is not None else [1] * S
    L.append("// a = M + scale * J on the LU structure (fill entries zero), M the diagonal mass matrix")
    L.append("// (differential_states): the dense scale and add_mass_matrix of")
    L.append("// ImplicitODESolver::compute_factorized_jacobian (ImplicitODESolver.cpp:92-105, 258-262).  The")
    L.append("// algebraic states of a DAE add 0.0, as the reference's weight * 0 does.")
    L.append("template <typename real>")
    L.append("GOSS_HD GOSS_INLINE void lu_from_jacobian(const real* __restrict__ J, const real scale, real* __restrict__ a)")
    L.append("{")
    for j in range(S):
        for i in sorted(lu[j]):
            s = lu_slot[(j, i)]
            m = "real(1.0)" if diff[i] else "real(0.0)"
            if (j, i) in jac_slot:
                e = f"J[{jac_slot[(j, i)]}] * scale"
                if i == j:
                    e += f" + {m}"
            else:
                e = m if i == j else "real(0.0)"
            L.append(f"{ind}a[{s}] = {e};")
    L.append("}")
    L.append("")
    L.append("// ODE::lu_factorize (ODE.cpp:69-100) in place on a, over the structure.")
    L.append("template <typename real>")
    L.append("GOSS_HD GOSS_INLINE void lu_factorize_sparse(real* __restrict__ a)")
    L.append("{")
    L.append(f"{ind}real acc;")

    def A(r, c):
        return f"a[{lu_slot[(r, c)]}]"

    def has(r, c):
        return (r, c) in lu_slot

    for k in range(1, S):
        for i in range(k):
            if has(i, k):
                terms = [r for r in range(i) if has(i, r) and has(r, k)]
                if terms:
                    L.append(f"{ind}acc = {A(i, terms[0])} * {A(terms[0], k)};")
                    for r in terms[1:]:
                        L.append(f"{ind}acc = acc + {A(i, r)} * {A(r, k)};")
                    L.append(f"{ind}{A(i, k)} = {A(i, k)} - acc;")
            if has(k, i):
                terms = [r for r in range(i) if has(k, r) and has(r, i)]
                if terms:
                    L.append(f"{ind}acc = {A(k, terms[0])} * {A(terms[0], i)};")
                    for r in terms[1:]:
                        L.append(f"{ind}acc = acc + {A(k, r)} * {A(r, i)};")
                    L.append(f"{ind}{A(k, i)} = ({A(k, i)} - acc) / {A(i, i)};")
                else:
                    L.append(f"{ind}{A(k, i)} = {A(k, i)} / {A(i, i)};")
        terms = [r for r in range(k) if has(k, r) and has(r, k)]
        if terms:
            L.append(f"{ind}acc = {A(k, terms[0])} * {A(terms[0], k)};")
            for r in terms[1:]:
                L.append(f"{ind}acc = acc + {A(k, r)} * {A(r, k)};")
            L.append(f"{ind}{A(k, k)} = {A(k, k)} - acc;")
    L.append(f"{ind}(void)acc;")
    L.append("}")
    L.append("")
    L.append("// ODE::forward_backward_subst (ODE.cpp:102-132) in place on x, over the structure.")
    L.append("template <typename real>")
    L.append("GOSS_HD GOSS_INLINE void lu_solve_sparse(const real* __restrict__ a, real* __restrict__ x)")
    L.append("{")
    L.append(f"{ind}real acc;")
    for i in range(1, S):
        terms = [j for j in range(i) if has(i, j)]
        if terms:
            L.append(f"{ind}acc = {A(i, terms[0])} * x[{terms[0]}];")
            for j in terms[1:]:
                L.append(f"{ind}acc = acc + {A(i, j)} * x[{j}];")
            L.append(f"{ind}x[{i}] = x[{i}] - acc;")
    L.append(f"{ind}x[{S - 1}] = x[{S - 1}] / {A(S - 1, S - 1)};")
    for i in range(S - 2, -1, -1):
        terms = [j for j in range(i + 1, S) if has(i, j)]
        if terms:
            L.append(f"{ind}acc = {A(i, terms[0])} * x[{terms[0]}];")
            for j in terms[1:]:
                L.append(f"{ind}acc = acc + {A(i, j)} * x[{j}];")
            L.append(f"{ind}x[{i}] = (x[{i}] - acc) / {A(i, i)};")
        else:
            L.append(f"{ind}x[{i}] = x[{i}] / {A(i, i)};")
    L.append(f"{ind}(void)acc;")
    L.append("}")
    L.append("")
    return L, jn, ln
