// Evaluates the gotran-generated model classes shipped with the reference
// (cpp/tests/fhn.h, cpp/tests/beeler_reuter_1977.h) at given points.
// Input  (binary, doubles): n, then n rows of [time, states...].
// Output (binary, doubles): n rows of [eval..., linear..., rhs...], then the
// linear_terms flags as doubles.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "fhn.h"
#include "beeler_reuter_1977.h"

template <class M>
static int run(const char* in_path, const char* out_path)
{
  M model;
  const unsigned S = model.num_states();
  FILE* fi = std::fopen(in_path, "rb");
  if (!fi) return 2;
  double nd = 0;
  if (std::fread(&nd, sizeof(double), 1, fi) != 1) return 3;
  const size_t n = (size_t)nd;
  std::vector<double> row(S + 1), ev(S), lin(S), rhs(S), out;
  for (size_t k = 0; k < n; ++k) {
    if (std::fread(row.data(), sizeof(double), S + 1, fi) != S + 1) return 4;
    model.eval(row.data() + 1, row[0], ev.data());
    model.linearized_eval(row.data() + 1, row[0], lin.data(), rhs.data(), false);
    out.insert(out.end(), ev.begin(), ev.end());
    out.insert(out.end(), lin.begin(), lin.end());
    out.insert(out.end(), rhs.begin(), rhs.end());
  }
  std::fclose(fi);
  for (unsigned i = 0; i < S; ++i) out.push_back((double)model.linear_term(i));
  FILE* fo = std::fopen(out_path, "wb");
  if (!fo) return 5;
  std::fwrite(out.data(), sizeof(double), out.size(), fo);
  std::fclose(fo);
  return 0;
}

int main(int argc, char** argv)
{
  if (argc != 4) return 1;
  const std::string m = argv[1];
  if (m == "fhn") return run<goss::Fhn>(argv[2], argv[3]);
  if (m == "beeler_reuter_1977") return run<goss::Beeler_reuter_1977>(argv[2], argv[3]);
  return 1;
}
