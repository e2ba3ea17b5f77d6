// C entry points over the reference goss C++ library (test infrastructure).
//
// This file is linked with the reference's own sources, cpp/goss/*.cpp, read
// where they lie (see oracle/Makefile), and with the generated model plugins
// (goss_amd/csrc/generated/*_plugin.h, the goss::ParameterizedODE contract of
// python/goss/codegeneration.py:51-114).  It gives tests and the CPU baseline
// a C ABI to the reference path: ODESystemSolver (ODESystemSolver.h:38-186)
// driving the reference solvers.  Nothing in the product links it.
#include <memory>
#include <stdexcept>
#include <string>

#include "goss/ParameterizedODE.h"

#include "tp06_m_plugin.h"
#include "tp06_m_fp3_plugin.h"
#include "tp06_m_fp6_plugin.h"
#include "tp06_epi_plugin.h"
#include "fitzhughnagumo_plugin.h"
#include "lorentz_plugin.h"
#include "fhn_plugin.h"
#include "beeler_reuter_1977_plugin.h"
#include "oscilator_plugin.h"
#include "linear_decay_plugin.h"
#include "van_der_pol_dae_plugin.h"
#include "markov_three_state_plugin.h"

namespace {

using namespace goss;

std::shared_ptr<ParameterizedODE> make_model(const std::string& v)
{
  if (v == "tp06_m") return std::make_shared<Tp06_m>();
  if (v == "tp06_m_fp3") return std::make_shared<Tp06_m_fp3>();
  if (v == "tp06_m_fp6") return std::make_shared<Tp06_m_fp6>();
  if (v == "tp06_epi") return std::make_shared<Tp06_epi>();
  if (v == "fitzhughnagumo") return std::make_shared<Fitzhughnagumo>();
  if (v == "lorentz") return std::make_shared<Lorentz>();
  if (v == "fhn") return std::make_shared<Fhn>();
  if (v == "beeler_reuter_1977") return std::make_shared<Beeler_reuter_1977>();
  if (v == "oscilator") return std::make_shared<Oscilator>();
  if (v == "linear_decay") return std::make_shared<Linear_decay>();
  if (v == "van_der_pol_dae") return std::make_shared<Van_der_pol_dae>();
  if (v == "markov_three_state") return std::make_shared<Markov_three_state>();
  throw std::runtime_error("unknown model variant " + v);
}

}  // namespace

#include "ref_system_abi.h"
