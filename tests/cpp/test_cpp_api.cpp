// C++ host API (include/goss_hip.hpp) exercised the way goss's own C++ tests
// use ODESystemSolver (cpp/tests/odesystem.cpp): build a system from a
// solver and a model, set field parameters, step with forward and solve,
// read field states back.  Test infrastructure, run by tests/test_cpp_api.py.
//
//   test_cpp_api info            model metadata only (no GPU needed)
//   test_cpp_api create          builds a system; without a GPU this must throw
//   test_cpp_api run OUT.bin     the GPU case: writes its results to OUT.bin
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
  goss_hip::ODESolver grl1("GRL1");
  goss_hip::ODESolver rkf("RKF32");
  std::cout << "adaptive " << grl1.is_adaptive() << rkf.is_adaptive() << " delta " << grl1.parameters().delta
            << "\n";
  return found ? 0 : 2;
}

int create()
{
  auto ode = std::make_shared<goss_hip::ParameterizedODE>("tp06_m_fp3");
  auto solver = std::make_shared<goss_hip::ODESolver>("GRL1");
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
  auto solver = std::make_shared<goss_hip::ODESolver>("GRL1");
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
    auto single = std::make_shared<goss_hip::ODESolver>("GRL1");
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
