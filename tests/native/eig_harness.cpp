// Test harness: C entry points over the host linear algebra of the operator
// set-up (channel_flow_amd/csrc/eig.h), loaded by tests/test_eig_host.py.
#include "../../channel_flow_amd/csrc/eig.h"

extern "C" int eig_real_c(int n, const double* a, double* lam, double* v) {
  chf::la::Mat A(a, a + (size_t)n * n), V;
  std::vector<double> l;
  if (!chf::la::eig_real(A, n, l, V)) return 1;
  for (int i = 0; i < n; ++i) lam[i] = l[i];
  for (size_t i = 0; i < V.size(); ++i) v[i] = V[i];
  return 0;
}

extern "C" int inverse_c(int n, const double* a, double* out) {
  chf::la::Mat A(a, a + (size_t)n * n), B;
  if (!chf::la::inverse(A, n, B)) return 1;
  for (size_t i = 0; i < B.size(); ++i) out[i] = B[i];
  return 0;
}
