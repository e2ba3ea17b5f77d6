// C++ host API (include/goss_hip.hpp) exercised the way goss's own C++ tests
// use ODESystemSolver (cpp/tests/odesystem.cpp): build a system from a
// solver and a model, set field parameters, step with forward and solve,
// read field states back.  Test infrastructure, run by tests/test_cpp_api.py.
//
//   test_cpp_api info            model metadata only (no GPU needed)
//   test_cpp_api create          builds a system; without a GPU this must throw
//   test_cpp_api run OUT.bin     the GPU case: writes its results to OUT.bin
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <vector>

#include "goss_hip.hpp"

namespace {

int info()
{
  const auto names = goss_hip::model_names();
  bool found = false;
  for (const auto& n : names) found = found || n == "tp06_m_fp3";
  goss_hip::ParameterizedODE ode("tp06_m_fp3");
  std::cout << "models " << names.size() << " states " << ode.num_states() << " parameters "
            << ode.num_parameters() << " field_states " << ode.num_field_states() << " field_parameters "
            << ode.num_field_parameters() << " first_state " << ode.get_state_names()[0] << " g_Kr "
            << ode.get_parameter("g_Kr") << "\n";
  try {
    ode.get_parameter("no_such_parameter");
    return 3;
  } catch (const goss_hip::Error& e) {
    std::cout << "error " << e.what() << "\n";
  }
  goss_hip::GRL1 grl1;
  goss_hip::RKF32 rkf;
  std::cout << "adaptive " << grl1.is_adaptive() << rkf.is_adaptive() << " delta " << grl1.parameters().delta
            << "\n";
  // Public tunables with the reference's defaults (GRL1.h:39, ImplicitODESolver.h:36-41,
  // ImplicitEuler.h:36-37, ThetaSolver.h:37, ESDIRK23a.h:37-38) reach the parameter block.
  grl1.delta = 1e-7;
  goss_hip::ThetaSolver theta;
  goss_hip::ESDIRK23a esdirk;
  std::cout << "tunables " << grl1.parameters().delta << " " << theta.theta << " " << theta.min_dt << " "
            << theta.eta_0 << " " << theta.kappa << " " << theta.max_iterations << " " << esdirk.min_dt << " "
            << rkf.get_atol() << " " << rkf.get_iord() << "\n";
  theta.theta = 0.75;
  theta.max_iterations = 12;
  theta.always_recompute_jacobian = true;
  theta.set_internal_time_step(0.02);
  const goss_hip_solver_params p = theta.parameters();
  rkf.set_tol(1e-6, 1e-9);
  std::cout << "block " << p.theta << " " << p.max_iterations << " " << p.always_recompute_jacobian << " "
            << p.internal_time_step << " " << rkf.parameters().atol << " " << rkf.parameters().rtol << "\n";
  std::cout << "solver " << (goss_hip::make_solver("ThetaSolver")->kind() == GOSS_HIP_THETA) << "\n";
  return found ? 0 : 2;
}

int create()
{
  auto ode = std::make_shared<goss_hip::ParameterizedODE>("tp06_m_fp3");
  auto solver = std::make_shared<goss_hip::GRL1>();
  try {
    goss_hip::ODESystemSolver system(64, solver, ode);
    std::cout << "created\n";
    return 0;
  } catch (const goss_hip::Error& e) {
    std::cout << "error " << e.code() << " " << e.what() << "\n";
    return 4;
  }
}

// TP06 M-cell (g_Kr, g_Ks, g_CaL per node), GRL1 in FP64, 1,000 nodes:
// 100 forward steps of 0.01 ms from t = 0, then solve over 50 more with the
// field states of every step; stimulus period changed through the model.
int run(const char* out_path)
{
  const goss_hip::uint N = 1000;
  auto ode = std::make_shared<goss_hip::ParameterizedODE>("tp06_m_fp3");
  ode->set_parameter("stim_start", 0.5);
  auto solver = std::make_shared<goss_hip::GRL1>();
  goss_hip::ODESystemSolver system(N, solver, ode);
  if (system.get_parameter("stim_start") != 0.5) return 5;
  const goss_hip::uint NFP = ode->num_field_parameters(), S = ode->num_states(), F = ode->num_field_states();
  std::vector<double> fp(size_t(N) * NFP);
  for (goss_hip::uint n = 0; n < N; ++n)
    for (goss_hip::uint j = 0; j < NFP; ++j)
      fp[size_t(n) * NFP + j] = ode->get_parameter(ode->get_field_parameter_names()[j]) * (1.0 + 0.001 * (n % 17) - 0.004 * j);
  system.set_field_parameters(fp.data());  // tangled, the reference's default
  system.reset_default();
  double t = 0.0;
  const double dt = 0.01;
  for (int k = 0; k < 100; ++k) {
    system.forward(t, dt);
    t += dt;
  }
  std::vector<double> times(51);
  for (int k = 0; k <= 50; ++k) times[k] = t + k * dt;
  std::vector<double> rows(size_t(51) * N * F, 0.0);
  system.solve(rows.data(), times.data(), times.size());
  const double* st = system.states();

  // ODESolver::solve on one node (ODESolver.h:78-98) against a one-node system over the same times:
  // the same kernel, so the same states bit for bit.
  {
    auto single = std::make_shared<goss_hip::GRL1>();
    single->attach(ode);
    std::vector<double> y0(S), traj(size_t(41) * S, 0.0), times1(41);
    ode->get_ic(y0.data());
    for (int k = 0; k <= 40; ++k) times1[k] = k * dt;
    single->solve(traj.data(), y0.data(), times1.data(), times1.size());
    goss_hip::ODESystemSolver one(1, solver, ode);
    for (int k = 0; k < 40; ++k) one.forward(times1[k], times1[k + 1] - times1[k]);
    const double* s1 = one.states();
    for (goss_hip::uint j = 0; j < S; ++j)
      if (traj[size_t(40) * S + j] != s1[j] || y0[j] != s1[j]) return 7;
    std::cout << "single-node solve equals the one-node system\n";
  }

  // The system and a one-node solver read the solver's tunables and the model's parameter
  // values again at every call: changes made between two calls equal a fresh object created
  // with them.
  {
    auto m1 = std::make_shared<goss_hip::ParameterizedODE>("tp06_m");
    auto m2 = std::make_shared<goss_hip::ParameterizedODE>("tp06_m");
    auto a = std::make_shared<goss_hip::GRL1>();
    a->attach(m1);
    std::vector<double> ya(S), yb(S);
    m1->get_ic(ya.data());
    a->forward(ya.data(), 0.0, 1.0);
    a->set_internal_time_step(0.01);
    a->delta = 1e-6;
    m1->set_parameter("stim_amplitude", 60.0);
    m1->set_parameter("stim_start", 1.5);
    a->forward(ya.data(), 1.0, 4.0);
    auto b = std::make_shared<goss_hip::GRL1>();
    b->attach(m2);
    m2->get_ic(yb.data());
    b->forward(yb.data(), 0.0, 1.0);
    auto c = std::make_shared<goss_hip::GRL1>();
    c->set_internal_time_step(0.01);
    c->delta = 1e-6;
    m2->set_parameter("stim_amplitude", 60.0);
    m2->set_parameter("stim_start", 1.5);
    c->attach(m2);
    c->forward(yb.data(), 1.0, 4.0);
    for (goss_hip::uint j = 0; j < S; ++j)
      if (ya[j] != yb[j]) return 8;
    if (!(ya[17] > 0)) return 9;  // the stimulus moved to 1.5 ms fired the upstroke

    auto sa = std::make_shared<goss_hip::GRL1>();
    auto m3 = std::make_shared<goss_hip::ParameterizedODE>("tp06_m");
    goss_hip::ODESystemSolver s1(8, sa, m3);
    s1.forward(0.0, 1.0);
    sa->set_internal_time_step(0.01);
    m3->set_parameter("stim_start", 1.5);
    s1.forward(1.0, 4.0);
    auto sb = std::make_shared<goss_hip::GRL1>();
    auto m4 = std::make_shared<goss_hip::ParameterizedODE>("tp06_m");
    goss_hip::ODESystemSolver s2(8, sb, m4);
    s2.forward(0.0, 1.0);
    std::vector<double> mid(size_t(8) * S), e1(size_t(8) * S), e2(size_t(8) * S);
    s2.get_states(mid.data());
    auto sc = std::make_shared<goss_hip::GRL1>();
    sc->set_internal_time_step(0.01);
    auto m5 = std::make_shared<goss_hip::ParameterizedODE>("tp06_m");
    m5->set_parameter("stim_start", 1.5);
    goss_hip::ODESystemSolver s3(8, sc, m5);
    s3.set_states(mid.data());
    s3.forward(1.0, 4.0);
    s1.get_states(e1.data());
    s3.get_states(e2.data());
    if (e1 != e2) return 10;
    // Writes through the states() mirror reach the device before the next step.
    double* w = s1.states();
    w[17] = -80.0;
    s1.forward(5.0, 0.1);
    std::vector<double> chk(size_t(8) * S);
    s1.get_states(chk.data());
    for (size_t k = 0; k < chk.size(); ++k)
      if (chk[k] != w[k]) return 11;
    std::cout << "tunables and parameters changed between calls apply\n";
  }

  // solve_async into pinned host memory returns before the rows are there; after synchronize they
  // equal solve's rows into a std::vector bit for bit (200,000 nodes, 40 steps: 3 chunks of rows).
  {
    const goss_hip::uint M = 200000;
    auto m = std::make_shared<goss_hip::ParameterizedODE>("tp06_m");
    auto sa = std::make_shared<goss_hip::GRL1>();
    auto sb = std::make_shared<goss_hip::GRL1>();
    goss_hip::ODESystemSolver a(M, sa, m), b(M, sb, m);
    std::vector<double> tt(41);
    for (int k = 0; k <= 40; ++k) tt[k] = k * 0.05;
    std::vector<double> ra(size_t(41) * M, 0.0);
    a.solve(ra.data(), tt.data(), tt.size());
    void* p = nullptr;
    goss_hip::check(goss_hip_host_alloc(&p, ra.size() * sizeof(double)));
    double* rb = static_cast<double*>(p);
    std::fill(rb, rb + ra.size(), 0.0);
    b.solve_async(rb, tt.data(), tt.size());
    b.synchronize();
    const bool same = std::equal(ra.begin() + M, ra.end(), rb + M) && rb[0] == 0.0;
    goss_hip::check(goss_hip_host_free(p));
    if (!same) return 12;
    std::cout << "solve_async into pinned memory equals solve\n";
  }

  // ImplicitEuler with the analytic Jacobian (ImplicitODESolver::jacobian = GOSS_HIP_JACOBIAN_ANALYTIC),
  // through an upstroke (stimulus at 1 ms): other bits than with the forward differences, the same
  // states to within 1e-4 mV.
  {
    auto m = std::make_shared<goss_hip::ParameterizedODE>("tp06_m");
    m->set_parameter("stim_start", 1.0);
    auto fd = std::make_shared<goss_hip::ImplicitEuler>();
    auto an = std::make_shared<goss_hip::ImplicitEuler>();
    an->jacobian = GOSS_HIP_JACOBIAN_ANALYTIC;
    goss_hip::ODESystemSolver a(64, fd, m), b(64, an, m);
    for (int k = 0; k < 100; ++k) {
      a.forward(k * 0.05, 0.05);
      b.forward(k * 0.05, 0.05);
    }
    std::vector<double> ya(size_t(64) * S), yb(size_t(64) * S);
    a.get_states(ya.data());
    b.get_states(yb.data());
    double dv = 0.0;
    for (int n = 0; n < 64; ++n) dv = std::max(dv, std::fabs(ya[size_t(n) * S + 17] - yb[size_t(n) * S + 17]));
    if (ya == yb) return 13;
    if (!(dv < 1e-4) || !(ya[17] > 0)) return 14;
    std::cout << "analytic Jacobian: max |dV| vs forward differences " << dv << " mV\n";
  }

  std::ofstream f(out_path, std::ios::binary);
  const double header[3] = {double(N), double(S), double(F)};
  f.write(reinterpret_cast<const char*>(header), sizeof(header));
  f.write(reinterpret_cast<const char*>(fp.data()), fp.size() * sizeof(double));
  f.write(reinterpret_cast<const char*>(st), size_t(N) * S * sizeof(double));
  f.write(reinterpret_cast<const char*>(rows.data()), rows.size() * sizeof(double));
  std::cout << "ran " << N << " nodes; V[0] " << st[size_t(0) * S + 17] << "\n";
  return f.good() ? 0 : 6;
}

}  // namespace

int main(int argc, char** argv)
{
  try {
    if (argc >= 2 && std::strcmp(argv[1], "info") == 0) return info();
    if (argc >= 2 && std::strcmp(argv[1], "create") == 0) return create();
    if (argc >= 3 && std::strcmp(argv[1], "run") == 0) return run(argv[2]);
  } catch (const std::exception& e) {
    std::cerr << "failed: " << e.what() << "\n";
    return 1;
  }
  std::cerr << "usage: test_cpp_api info | create | run OUT.bin\n";
  return 64;
}
