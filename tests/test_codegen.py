"""The .ode frontend and code generator against gotran's own output.

Pins (none of them retypes a formula of the generator):
* state/parameter order against the gotran-generated headers and the
  gotran-generated TP06 Python function held by the reference;
* eval / linearized_eval / linear_terms against cpp/tests/fhn.h and
  cpp/tests/beeler_reuter_1977.h (vectors in tests/golden/gotran_kat_*.npz);
* the TP06 right-hand side against the gotran-generated Python rhs
  (benchmarks/tentusscher_panfilov_2006_M_cell.py:7-356, vectors in
  tests/golden/tp06_rhs_vectors.npz);
* linearized diagonals and the analytic Jacobian against central finite
  differences of the generated right-hand side (an invariant);
* ContinuousConditional against the logistic closed form.
"""
import numpy as np
import pytest
import sympy as sp

import oracle
from goss_amd.codegen.ode_model import ODEParseError, load_ode, parse_ode_text
from goss_amd.codegen.generate import ODE_DIR, VARIANTS
from conftest import golden

TP06_STATES = ["Xr1", "Xr2", "Xs", "m", "h", "j", "d", "f", "f2", "fCass", "s", "r", "Ca_i", "R_prime", "Ca_SR",
               "Ca_ss", "Na_i", "V", "K_i"]


def test_tp06_orders_match_gotran():
    m = load_ode(f"{ODE_DIR}/tentusscher_panfilov_2006_M_cell.ode")
    # benchmarks/tentusscher_panfilov_2006_M_cell.py:13-33
    assert m.state_names == TP06_STATES
    params = m.parameter_names
    assert len(params) == 53
    # benchmarks/tentusscher_panfilov_2006_M_cell.py:36-90 (first and a sorted block)
    assert params[:12] == ["P_kna", "g_K1", "g_Kr", "g_Ks", "g_Na", "g_bna", "g_CaL", "g_bca", "g_to", "K_mNa",
                           "K_mk", "P_NaK"]
    assert params[-10:] == ["Cm", "F", "R", "T", "V_c", "stim_amplitude", "stim_duration", "stim_period",
                            "stim_start", "K_o"]
    d = np.load(golden("tp06_rhs_vectors.npz"))
    assert np.array_equal(d["parameters"], oracle.Variant("tp06_m").default_parameters)


def test_fhn_and_br_orders_match_gotran_headers():
    fhn = load_ode(f"{ODE_DIR}/fhn.ode")
    assert fhn.state_names == ["v", "s"]  # cpp/tests/fhn.h:22-23
    assert fhn.parameter_names == ["a", "b", "c_1", "c_2", "c_3", "stim_amplitude", "stim_duration",
                                   "stim_period", "stim_start", "v_peak", "v_rest"]  # fhn.h:27-37
    br = load_ode(f"{ODE_DIR}/beeler_reuter_1977.ode")
    assert br.state_names == ["m", "h", "j", "Cai", "d", "f", "x1", "V"]  # beeler_reuter_1977.h:25-32
    assert br.parameter_names == ["E_Na", "g_Na", "g_Nac", "g_s", "IstimAmplitude", "IstimEnd", "IstimPeriod",
                                  "IstimPulseDuration", "IstimStart", "C"]  # beeler_reuter_1977.h:35-44


@pytest.mark.parametrize("model", ["fhn", "beeler_reuter_1977"])
def test_gotran_header_kat(model):
    d = np.load(golden(f"gotran_kat_{model}.npz"))
    v = oracle.Variant(model)
    assert list(d["linear_terms"]) == v.linear_terms
    for k in range(len(d["time"])):
        y, t = d["states"][k], d["time"][k]
        ev = v.eval(y, t)
        lin, rhs = v.linearized_eval(y, t)
        np.testing.assert_allclose(ev, d["eval"][k], rtol=1e-11, atol=1e-13)
        np.testing.assert_allclose(rhs, d["rhs"][k], rtol=1e-11, atol=1e-13)
        if model == "fhn":
            np.testing.assert_allclose(lin, d["linear"][k], rtol=1e-11, atol=1e-13)
        else:
            # gotran's linear[7] (dV/dt) counts the path through the
            # conditional V_eff_1 twice (beeler_reuter_1977.h:264-286:
            # di_K1_dV already contains dV_eff_1_dV*di_K1_dV_eff_1, which is
            # added again).  Every other entry must match; entry 7 is pinned
            # by finite differences in test_linearized_matches_fd.
            np.testing.assert_allclose(lin[:7], d["linear"][k][:7], rtol=1e-11, atol=1e-13)


def test_br_gotran_double_count_is_the_chain_term():
    # The gotran discrepancy in linear[7] equals exactly one extra copy of
    # d i_K1 / d V_eff_1 * d V_eff_1 / dV / C, computed here from the .ode
    # closed form of i_K1 (beeler_reuter_1977.ode:88-89) and C = 0.01.
    d = np.load(golden("gotran_kat_beeler_reuter_1977.npz"))
    v = oracle.Variant("beeler_reuter_1977")
    V = d["states"][:, 7]
    veff = np.where(np.abs(V + 23) < 0.01, 0.01, V + 23)
    dveff = np.where(np.abs(V + 23) < 0.01, 0.0, 1.0)
    e = np.exp(-0.04 * veff)
    di_dveff = 0.0035 * 0.2 * (V + 23) * (-0.04 * e) / (1 - e) ** 2
    extra = -(dveff * di_dveff) / 0.01
    ours = np.array([v.linearized_eval(d["states"][k], d["time"][k])[0][7] for k in range(len(V))])
    np.testing.assert_allclose(d["linear"][:, 7] - ours, extra, rtol=1e-7, atol=1e-9)


def test_tp06_rhs_matches_gotran_python():
    d = np.load(golden("tp06_rhs_vectors.npz"))
    v = oracle.Variant("tp06_m")
    for k in range(len(d["time"])):
        ours = v.eval(d["states"][k], d["time"][k], d["parameters"])
        np.testing.assert_allclose(ours, d["values"][k], rtol=1e-11, atol=1e-14)


def _fd_column(v, y, t, i, h):
    yp, ym = y.copy(), y.copy()
    yp[i] += h
    ym[i] -= h
    return (v.eval(yp, t) - v.eval(ym, t)) / (2 * h)


@pytest.mark.parametrize("variant", ["tp06_m", "tp06_epi", "beeler_reuter_1977", "fhn", "lorentz"])
def test_linearized_matches_fd(variant):
    v = oracle.Variant(variant)
    rng = np.random.Generator(np.random.PCG64(3))
    ic = v.initial_conditions
    vidx = {"tp06_m": 17, "tp06_epi": 17, "beeler_reuter_1977": 7}.get(variant)
    for _ in range(20):
        y = ic * (1 + rng.uniform(-0.2, 0.2, ic.size)) + rng.uniform(-1e-3, 1e-3, ic.size) * (ic == 0)
        if vidx is not None:
            y[vidx] = rng.uniform(-90, 40)
        t = rng.uniform(0, 100)
        lin, rhs = v.linearized_eval(y, t)
        np.testing.assert_allclose(rhs, v.eval(y, t), rtol=0, atol=0)
        for i in range(v.num_states):
            h = 1e-6 * max(abs(y[i]), 1e-3)
            fd = _fd_column(v, y, t, i, h)[i]
            assert abs(lin[i] - fd) <= 1e-5 * max(abs(fd), 1.0) + 1e-4 * abs(fd), (variant, i, lin[i], fd)
        # linear_terms == 1 means f_i is affine in x_i: the diagonal derivative
        # does not change when x_i moves.
        for i in range(v.num_states):
            if v.linear_terms[i]:
                y2 = y.copy()
                y2[i] *= 1.3
                y2[i] += 1e-3
                assert v.linearized_eval(y2, t)[0][i] == pytest.approx(lin[i], rel=1e-12, abs=1e-15)


@pytest.mark.parametrize("variant", ["tp06_m", "beeler_reuter_1977", "fhn", "lorentz", "oscilator"])
def test_analytic_jacobian_matches_fd(variant):
    v = oracle.Variant(variant)
    rng = np.random.Generator(np.random.PCG64(4))
    ic = v.initial_conditions
    for _ in range(5):
        y = ic * (1 + rng.uniform(-0.1, 0.1, ic.size)) + 1e-3 * (ic == 0)
        t = 5.0
        J = v.analytic_jacobian(y, t)
        for j in range(v.num_states):
            h = 1e-6 * max(abs(y[j]), 1e-3)
            col = _fd_column(v, y, t, j, h)
            np.testing.assert_allclose(J[:, j], col, rtol=2e-4, atol=1e-6 * max(1.0, np.abs(col).max()))
        # The diagonal equals the linearized diagonal.
        np.testing.assert_allclose(np.diag(J), v.linearized_eval(y, t)[0], rtol=1e-12, atol=1e-14)


def test_continuous_conditional_closed_form():
    m = parse_ode_text("parameters(a=1.0, s=0.2)\nstates(x=0.0)\ndx_dt = ContinuousConditional(Ge(time, a), 3, 1, s)\n",
                       "cc")
    expr = m.body_dict()["dx_dt"]
    T = [s for s in expr.free_symbols if s.name == "time"][0]
    f = sp.lambdify([T], expr.subs({m.symbols["a"]: 1.0, m.symbols["s"]: 0.2}))
    for t in (0.0, 0.9, 1.0, 1.1, 5.0):
        logistic = 1.0 / (1.0 + np.exp(-(t - 1.0) / 0.2))
        assert f(t) == pytest.approx(1 + 2 * logistic, rel=1e-14)


def test_parse_errors():
    with pytest.raises(ODEParseError):
        parse_ode_text("states(x=1)\n", "m")  # no derivative
    with pytest.raises(ODEParseError):
        parse_ode_text("states(x=1)\ndx_dt = foo(x)\n", "m")
    with pytest.raises(ODEParseError):
        parse_ode_text("states(x=1)\ndx_dt = x\ndx_dt = 2*x\n", "m")


def test_all_variants_generate():
    for name in VARIANTS:
        v = oracle.Variant(name)
        assert v.num_states > 0
        y = v.initial_conditions
        assert np.all(np.isfinite(v.eval(y, 0.0)))


def test_log2e_scaled_is_the_argument_times_log2e():
    """exp(x) prints as m_exp_l2(x log2(e), x); the FP32 kernels evaluate exp2 of the first argument.
    The fold into literal coefficients must leave the value x * log2(e)."""
    from goss_amd.codegen.printer import LOG2E, log2e_scaled

    V, c, y = sp.symbols("V c__5 y", real=True)
    args = [-sp.Float(0.1) * V - 4.5, sp.Float(0.3485) * V, c * V, -(V + 45) / 10, V, -V * c + 2 * y,
            sp.Float(0.0438) * (V - y) * c, sp.Float(3.0)]
    rng = np.random.Generator(np.random.PCG64(3))
    for a in args:
        s = log2e_scaled(a)
        for _ in range(3):
            sub = {V: rng.uniform(-90, 40), c: rng.uniform(0.01, 0.1), y: rng.uniform(-2, 2)}
            want = float(a.subs(sub)) * LOG2E
            assert abs(float(s.subs(sub)) - want) <= 1e-12 * max(1.0, abs(want)), (a, s)
    assert abs(LOG2E - 1.0 / np.log(2.0)) < 1e-15


@pytest.mark.parametrize("variant", ["tp06_m_fp6", "tp06_epi", "lorentz"])
def test_gate_states_are_the_closed_form_the_fp32_stage_uses(variant):
    """For every state that the generator marks as a gate (gate_states = 1), rhs_i = (A - y_i) * (-linear_i)
    with A = the gate target written by linearized_eval_gates: then rhs expm1(linear h) / linear =
    (y - A) expm1(linear h), the FP32 GRL stage's gate form.  TP06 has 12 gates (the Hodgkin-Huxley
    gating variables); Lorenz's x' = sigma (y - x) is one."""
    from goss_amd.codegen.adjoint import gate_states
    from goss_amd.codegen.analysis import linearize
    from goss_amd.codegen.generate import make_variant

    var = make_variant(variant)
    model = var.model
    lin = linearize(model)
    body = dict(model.body)
    rhs = [body[model.derivatives[s]] for s in model.state_names]
    gates = gate_states(model.body, rhs, model.state_names, lin.linear_terms)
    assert len(gates) == {"tp06_m_fp6": 12, "tp06_epi": 12, "lorentz": 1}[variant]

    def expand(e):
        # substitute intermediates until only states, parameters and time remain
        for _ in range(50):
            names = {s.name for s in e.free_symbols}
            sub = {sp.Symbol(n, real=True): body[n] for n in names if n in body and n not in model.derivatives.values()}
            if not sub:
                return e
            e = e.xreplace(sub)
        raise AssertionError("intermediate expansion did not terminate")

    rng = np.random.Generator(np.random.PCG64(4))
    ic = {sp.Symbol(s.name, real=True): s.init * (1 + 0.01 * rng.uniform(-1, 1)) for s in model.states}
    prm = {sp.Symbol(p.name, real=True): p.value for p in model.parameters}
    for i, a_sym in gates.items():
        y = sp.Symbol(model.state_names[i], real=True)
        r = expand(rhs[i])
        a = expand(a_sym)
        d = sp.diff(r, y)
        assert sp.simplify(sp.diff(a, y)) == 0
        for V in (-85.0, -41.0, -39.0, 20.0):
            sub = dict(ic)
            sub.update(prm)
            vs = sp.Symbol("V", real=True)
            if vs in sub:
                sub[vs] = V
            sub[sp.Symbol("time", real=True)] = 1.0
            lhs = float(r.subs(sub))
            rhs_closed = float(((a - y) * (-d)).subs(sub))
            assert abs(lhs - rhs_closed) <= 1e-12 * max(1.0, abs(lhs)), (variant, model.state_names[i], V)
