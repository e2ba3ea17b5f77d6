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

from .analysis import Jacobian, Linearized, census_body, census_state_ops, jacobian, linearize, needed_names
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


def _bind_states(model: ODEModel, used: set, arr: str, ind: str) -> List[str]:
    out = []
    for i, s in enumerate(model.states):
        if s.name in used:
            out.append(f"{ind}const real {c_name(s.name)} = {arr}[{i}];")
    return out


def _bind_params(model: ODEModel, used: set, ind: str) -> List[str]:
    out = []
    for k, p in enumerate(model.parameters):
        if p.name in used:
            out.append(f"{ind}const real {c_name(p.name)} = pa.template get<{k}>();")
    return out


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


def time_only_names(model: ODEModel) -> List[str]:
    """Intermediates that depend on time (and parameters) but on no state.

    They are evaluated in the type of ``time`` (``treal``, double on every
    target) so that a single-precision kernel sees the same discontinuities
    in time (stimulus on/off) as the double-precision oracle."""
    state_set = set(model.state_names)
    dep_state: Dict[str, bool] = {}
    dep_time: Dict[str, bool] = {}
    out = []
    for n, e in model.body:
        ds = dt = False
        for s in e.free_symbols:
            if s.name in state_set:
                ds = True
            elif s == TIME:
                dt = True
            elif s.name in dep_state:
                ds = ds or dep_state[s.name]
                dt = dt or dep_time[s.name]
        dep_state[n], dep_time[n] = ds, dt
        if dt and not ds and n not in model.derivatives.values():
            out.append(n)
    return out


class _Printers:
    """The main printer (type ``real``) and the printer of time-only
    statements (type ``treal``)."""

    def __init__(self, model: ODEModel):
        self.time_only = set(time_only_names(model))
        main_map = {n: f"real({c_name(n)})" for n in self.time_only}
        main_map["time"] = "real(time)"
        self.main = CPrinter(main_map)
        tmap = {p.name: f"treal({c_name(p.name)})" for p in model.parameters}
        self.time = CPrinter(tmap, lit_type="treal")

    def lines(self, model: ODEModel, names: Sequence[str], ind: str) -> List[str]:
        bd = model.body_dict()
        out = []
        for n in names:
            if n in self.time_only:
                out += assignments(self.time, [(n, bd[n])], ind, decl="const treal ")
            else:
                out += assignments(self.main, [(n, bd[n])], ind)
        return out


def emit_model_header(var: Variant, with_jacobian: bool = True) -> Emitted:
    model = var.model
    prs = _Printers(model)
    pr = prs.main
    lin = linearize(model)
    S, P = model.num_states, model.num_parameters
    ind = "  "
    rhs_syms = [sp.Symbol(model.derivatives[s], real=True) for s in model.state_names]
    rhs_names = needed_names(model, rhs_syms)
    bd = model.body_dict()
    rhs_exprs = [bd[n] for n in rhs_names]

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
    L.append("")

    # -- eval ---------------------------------------------------------------
    used = _used_symbols(rhs_exprs + [bd[model.derivatives[s]] for s in model.state_names])
    L.append("// Right-hand side: values[i] = d states[i] / dt.  Intermediates that depend on time but on")
    L.append("// no state are computed in the type of `time` (treal).")
    L.append("template <typename real, typename PA, typename treal>")
    L.append("GOSS_HD GOSS_INLINE void eval(const real* __restrict__ states, const treal time, const PA& pa,")
    L.append("                              real* __restrict__ values)")
    L.append("{")
    L.append(f"{ind}(void)time; (void)pa;")
    L += _bind_states(model, used, "states", ind)
    L += _bind_params(model, used, ind)
    L += prs.lines(model, [n for n in rhs_names if n not in model.derivatives.values()], ind)
    for i, s in enumerate(model.state_names):
        L.append(f"{ind}values[{i}] = {pr.p(bd[model.derivatives[s]])};")
    L.append("}")
    L.append("")

    # -- linearized_eval ----------------------------------------------------
    lin_first = lin.lin_stmts[: lin.n_linear_stmts]
    lin_second = lin.lin_stmts[lin.n_linear_stmts:]
    used_l = used | _used_symbols([e for _, e in lin.lin_stmts] + lin.linear)
    # Intermediates referenced by derivative expressions but not by the RHS.
    extra = needed_names(model, [e for _, e in lin.lin_stmts] + lin.linear)
    extra = [n for n in extra if n not in rhs_names]
    L.append("// Linearized right-hand side: rhs[i] as in eval and linear[i] = d rhs[i] / d states[i].")
    L.append("// Entries with linear_terms[i] == 1 are written before the only_linear return.")
    L.append("template <typename real, typename PA, typename treal>")
    L.append("GOSS_HD GOSS_INLINE void linearized_eval(const real* __restrict__ states, const treal time, const PA& pa,")
    L.append("                                         real* __restrict__ linear, real* __restrict__ rhs,")
    L.append("                                         const bool only_linear)")
    L.append("{")
    L.append(f"{ind}(void)time; (void)pa;")
    used_all = used_l | _used_symbols([bd[n] for n in extra])
    L += _bind_states(model, used_all, "states", ind)
    L += _bind_params(model, used_all, ind)
    all_inter = [n for n, _ in model.body if n in set(rhs_names) | set(extra) and n not in model.derivatives.values()]
    L += prs.lines(model, all_inter, ind)
    for i, s in enumerate(model.state_names):
        L.append(f"{ind}rhs[{i}] = {pr.p(bd[model.derivatives[s]])};")
    L += assignments(pr, lin_first, ind)
    for i in range(S):
        if lin.linear_terms[i]:
            L.append(f"{ind}linear[{i}] = {pr.p(lin.linear[i])};")
    L.append(f"{ind}if (only_linear) return;")
    L += assignments(pr, lin_second, ind)
    for i in range(S):
        if not lin.linear_terms[i]:
            L.append(f"{ind}linear[{i}] = {pr.p(lin.linear[i])};")
    L.append("}")
    L.append("")

    rhs_stmts = [(n, bd[n]) for n, _ in model.body if n in set(rhs_names)]
    lin_stmts_all = [(n, bd[n]) for n, _ in model.body if n in set(rhs_names) | set(extra)] + lin.lin_stmts
    census = {
        # Per evaluation, after common-subexpression elimination, operations
        # whose operands depend on a state (the algorithmic work).
        "eval_state": census_state_ops(rhs_stmts, [], model.state_names),
        "linearized_eval_state": census_state_ops(lin_stmts_all, lin.linear, model.state_names),
        # Operations as printed (before compiler CSE), all operations.
        "eval": census_body(rhs_exprs),
        "linearized_eval": census_body(
            [bd[n] for n in all_inter] + [bd[model.derivatives[s]] for s in model.state_names]
            + [e for _, e in lin.lin_stmts] + lin.linear),
    }

    jac_nnz = 0
    if with_jacobian:
        jac = jacobian(model)
        jused = _used_symbols([e for _, e in jac.stmts] + [e for row in jac.entries for e in row])
        jextra = needed_names(model, [e for _, e in jac.stmts] + [e for row in jac.entries for e in row])
        jinter = [n for n, _ in model.body if n in set(jextra) and n not in model.derivatives.values()]
        jused |= _used_symbols([bd[n] for n in jinter])
        L.append("// Analytic Jacobian: jac[i*num_states + j] = d values[i] / d states[j].")
        L.append("template <typename real, typename PA, typename treal>")
        L.append("GOSS_HD GOSS_INLINE void compute_jacobian(const real* __restrict__ states, const treal time,")
        L.append("                                          const PA& pa, real* __restrict__ jac)")
        L.append("{")
        L.append(f"{ind}(void)time; (void)pa;")
        L += _bind_states(model, jused, "states", ind)
        L += _bind_params(model, jused, ind)
        L += prs.lines(model, jinter, ind)
        L += assignments(pr, jac.stmts, ind)
        for i in range(S):
            for j in range(S):
                e = jac.entries[i][j]
                if e != 0:
                    jac_nnz += 1
                L.append(f"{ind}jac[{i * S + j}] = {pr.p(e)};")
        L.append("}")
        L.append("")
        census["jacobian"] = census_body([bd[n] for n in jinter] + [e for _, e in jac.stmts]
                                         + [e for row in jac.entries for e in row])

    L.append("}}  // namespace goss_models::" + var.name)
    L.append("")
    return Emitted(header="\n".join(L), census=census, linear_terms=lin.linear_terms, jac_nnz=jac_nnz)


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
    L.append("  struct Acc { const double* v; template <int k> double get() const { return v[k]; } };")
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
    L.append("  }")
    L.append(f"  {N}(const {N}& ode) : ParameterizedODE(ode)")
    L.append("  {")
    L.append(f"    for (int i = 0; i < {P}; ++i) _p[i] = ode._p[i];")
    L.append("    _bind();")
    L.append("  }")
    L.append("  void eval(const double* states, double time, double* values)")
    L.append(f"  {{ {ns}::eval<double>(states, time, Acc{{_p}}, values); }}")
    L.append("  void linearized_eval(const double* states, double time, double* linear, double* rhs,")
    L.append("                       bool only_linear) const")
    L.append(f"  {{ {ns}::linearized_eval<double>(states, time, Acc{{_p}}, linear, rhs, only_linear); }}")
    L.append("  void get_ic(goss::DoubleVector* values) const")
    L.append("  {")
    L.append("    values->n = _num_states;")
    L.append("    values->data.reset(new double[_num_states]);")
    L.append(f"    for (int i = 0; i < {S}; ++i) values->data[i] = {ns}::initial_conditions[i];")
    L.append("  }")
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
        "census": em.census,
        "jacobian_nnz": em.jac_nnz,
    }
