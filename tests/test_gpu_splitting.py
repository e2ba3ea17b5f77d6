"""The operator-splitting exchange on the device (goss_amd.splitting and the
C ABI's gather / scatter / device-state entry points) against the reference's
DOLFINODESystemSolver algorithm (python/goss/dolfinutils.py:670-810)
restated on the host with numpy index arrays and the oracle's
ODESystemSolver."""
import numpy as np
import pytest

import oracle
from test_gpu_explicit import assert_parity, perturbed

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


@pytest.mark.parametrize("precision", ["fp64", "fp32c", "fp32"])
def test_gather_scatter_equal_the_field_state_accessors(precision):
    """gather(values) sets the same states as set_field_states(values[dofs]) bit for bit;
    scatter writes get_field_states() to values[dofs] and leaves every other entry alone."""
    torch = _torch()
    from goss_amd.system import DeviceSystem

    N, NV = 1000, 3001
    rng = np.random.Generator(np.random.PCG64(3))
    dofs = rng.permutation(NV)[:N].astype(np.int32)
    a = DeviceSystem("tp06_m_fp3", "GRL1", N, precision)
    b = DeviceSystem("tp06_m_fp3", "GRL1", N, precision)
    a.set_field_dofs(dofs[None, :], NV)
    values = rng.uniform(-90, 40, NV)
    dv = torch.from_numpy(values).cuda()
    torch.cuda.synchronize()
    a.gather_field_states_device(dv.data_ptr(), NV)
    b.set_field_states(values[dofs][:, None].copy())
    assert np.array_equal(a.states(), b.states())
    out = torch.full((NV,), -7.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    a.forward(0.0, 0.1)
    a.scatter_field_states_device(out.data_ptr(), NV)
    fs = np.zeros((N, 1))
    a.get_field_states(fs)
    host = out.cpu().numpy()
    assert np.array_equal(host[dofs], fs[:, 0])
    rest = np.ones(NV, dtype=bool)
    rest[dofs] = False
    assert np.all(host[rest] == -7.0)


def test_dof_map_is_checked_on_the_host():
    torch = _torch()
    from goss_amd.system import DeviceSystem

    N = 16
    s = DeviceSystem("oscilator", "RK4", N, "fp64")  # two field states: dofs [2][N]
    v = torch.zeros(4 * N, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="no dof map"):
        s.gather_field_states_device(v.data_ptr(), 4 * N)
    good = np.arange(2 * N, dtype=np.int32).reshape(2, N) * 2
    bad = good.copy()
    bad[1, 3] = 4 * N
    with pytest.raises(RuntimeError, match="out of range"):
        s.set_field_dofs(bad, 4 * N)
    bad = good.copy()
    bad[1, 3] = bad[0, 5]
    with pytest.raises(RuntimeError, match="appears twice"):
        s.set_field_dofs(bad, 4 * N)
    bad = good.copy()
    bad[0, 0] = -1
    with pytest.raises(RuntimeError, match="out of range"):
        s.set_field_dofs(bad, 4 * N)
    with pytest.raises(ValueError):
        s.set_field_dofs(good[:1], 4 * N)
    s.set_field_dofs(good, 4 * N)
    with pytest.raises(RuntimeError, match="length differs"):
        s.gather_field_states_device(v.data_ptr(), 4 * N - 1)
    torch.cuda.synchronize()
    s.scatter_field_states_device(v.data_ptr(), 4 * N)
    host = v.cpu().numpy()
    st = s.states()
    assert np.array_equal(host[good[0]], st[:, 0]) and np.array_equal(host[good[1]], st[:, 1])


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_device_states_save_and_restore(precision):
    """get_states_device equals get_states bit for bit; set_states_device restores
    a saved copy exactly (both parts of the FP32C states)."""
    torch = _torch()
    from goss_amd.system import DeviceSystem

    N = 777
    s = DeviceSystem("tp06_m", "GRL1", N, precision)
    s.set_states(perturbed(s.states(), 5))
    s.forward_times(np.arange(11) * 0.02)
    saved = torch.empty((N, 19), dtype=torch.float64, device="cuda")
    s.get_states_device(saved.data_ptr())
    at_save = s.states()
    assert np.array_equal(saved.cpu().numpy(), at_save)
    s.forward_times(0.2 + np.arange(11) * 0.02)
    assert not np.array_equal(s.states(), at_save)
    s.set_states_device(saved.data_ptr())
    assert np.array_equal(s.states(), at_save)
    # The restored system continues exactly as the saved one would have.
    ref = DeviceSystem("tp06_m", "GRL1", N, precision)
    ref.set_states_device(saved.data_ptr())
    s.forward_times(0.2 + np.arange(11) * 0.02)
    ref.forward_times(0.2 + np.arange(11) * 0.02)
    assert np.array_equal(s.states(), ref.states())


def _host_splitting(num_entities, odes, domains, entity_to_dof, solver, steps, dt, pde, field_params):
    """DOLFINODESystemSolver.step (dolfinutils.py:750-780) restated with numpy and the oracle."""
    F = 1
    systems, dof_maps = {}, {}
    values = np.zeros(num_entities * F)
    for label, variant in odes.items():
        ents = np.nonzero(domains == label)[0]
        dof_maps[label] = entity_to_dof[ents * F]
        o = oracle.OracleSystem(variant, solver, len(ents))
        for name, (col, per_entity) in field_params.get(label, {}).items():
            fp = o.field_parameters()
            fp[:, col] = per_entity[ents]
            o.set_field_parameters(fp)
        systems[label] = o
        values[dof_maps[label]] = o.field_states()[:, 0]
    vs_ = values.copy()
    for k in range(steps):
        vs = vs_.copy()
        for label, o in systems.items():
            o.set_field_states(vs[dof_maps[label]][:, None].copy())
            o.forward(k * dt, dt)
            vs[dof_maps[label]] = o.field_states()[:, 0]
        vs_ = pde(vs)
    return vs, vs_


def test_two_domain_splitting_matches_host_restatement():
    """Two cell models on a 1,200-entity mesh (TP06 M-cell with per-node g_Kr, Beeler-Reuter),
    a permuted dof numbering, and a diffusion-like PDE update of vs_ between steps: the
    device-resident splitting solver equals the host restatement to 1e-10 relative."""
    torch = _torch()
    import goss_amd as goss
    from goss_amd.splitting import DeviceODESystemSolver

    n = 1200
    rng = np.random.Generator(np.random.PCG64(11))
    domains = np.where(rng.uniform(size=n) < 0.6, 1, 2)
    entity_to_dof = rng.permutation(n)
    g_kr = 0.153 * np.exp(rng.normal(0, 0.2, n))
    tp06 = goss.ParameterizedODE("tentusscher_panfilov_2006_M_cell.ode", field_states=["V"],
                                 field_parameters=["g_Kr", "g_Ks", "g_CaL"])
    br = goss.ParameterizedODE("beeler_reuter_1977")
    solver = DeviceODESystemSolver(n, {1: tp06, 2: br}, domains, entity_to_dof, params={"solver": "GRL1"})
    assert solver.num_distinct_domains == 2 and solver.field_names == ["V"]
    solver.set_field_parameter("g_Kr", g_kr)
    # A smoothing step standing in for the PDE solve, on the dof vector.
    nbr = np.roll(np.arange(n), 1)

    def pde_np(v):
        return v + 0.05 * (v[nbr] - v)

    nbr_t = torch.from_numpy(nbr).cuda()
    steps, dt = 120, 0.1
    vs_, vs = solver.solution_fields()
    for k in range(steps):
        solver.step((k * dt, (k + 1) * dt))
        vs_.copy_(vs + 0.05 * (vs[nbr_t] - vs))
    got = vs.cpu().numpy()
    ref, _ = _host_splitting(n, {1: "tp06_m_fp3", 2: "beeler_reuter_1977"}, domains, entity_to_dof, "GRL1",
                             steps, dt, pde_np, {1: {"g_Kr": (0, g_kr)}})
    assert got.max() > 0  # both stimuli fired the upstroke
    assert_parity(got[:, None], ref[:, None], rtol=1e-10, what="splitting vs")


def test_splitting_save_restore_and_solve():
    torch = _torch()
    import goss_amd as goss
    from goss_amd.splitting import DeviceODESystemSolver

    n = 300
    s = DeviceODESystemSolver(n, goss.ParameterizedODE("tp06_m"), params={"solver": "RL1", "precision": "fp32c"})
    steps = list(s.solve((0.0, 1.0), 0.25))
    assert [iv for iv, _ in steps] == [(0.0, 0.25), (0.25, 0.5), (0.5, 0.75), (0.75, 1.0)]
    s.save_states()
    v_saved = s.vs.clone()
    before = s.ode_system_solvers[0].states
    s.vs_.copy_(s.vs)
    s.step((1.0, 1.5))
    assert not torch.equal(s.vs, v_saved)
    s.restore_states()
    assert np.array_equal(s.ode_system_solvers[0].states, before)
    with pytest.raises(RuntimeError, match="none are saved"):
        DeviceODESystemSolver(n, goss.ParameterizedODE("tp06_m")).restore_states()
