// Generated by goss_amd.codegen.generate -- do not edit.
// X-macro list of the model variants compiled into the library.
#pragma once
#define GOSS_VARIANT_LIST(X) \
  X(tp06_m) \
  X(tp06_m_fp3) \
  X(tp06_m_fp6) \
  X(tp06_epi) \
  X(fitzhughnagumo) \
  X(lorentz) \
  X(fhn) \
  X(beeler_reuter_1977) \
  X(oscilator) \
  X(linear_decay) \
  X(van_der_pol_dae) \
  X(markov_three_state) \

