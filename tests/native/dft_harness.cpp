// Host build of the register DFTs of csrc/fftreg.h for tests/test_fftreg_host.py:
// one line of R complex values (interleaved re, im) transformed in place by
// Dft<R, S> in double or float.
#include "../../channel_flow_amd/csrc/fftreg.h"

namespace {
template <int R, int S, class V, class T>
void run(const double* in, double* out) {
  V v[R];
  for (int i = 0; i < R; ++i) {
    v[i].x = (T)in[2 * i];
    v[i].y = (T)in[2 * i + 1];
  }
  chf::Dft<R, S, V>::run(v);
  for (int i = 0; i < R; ++i) {
    out[2 * i] = (double)v[i].x;
    out[2 * i + 1] = (double)v[i].y;
  }
}

template <int S, class V, class T>
int dispatch(int r, const double* in, double* out) {
  switch (r) {
    case 2: run<2, S, V, T>(in, out); return 0;
    case 3: run<3, S, V, T>(in, out); return 0;
    case 4: run<4, S, V, T>(in, out); return 0;
    case 6: run<6, S, V, T>(in, out); return 0;
    case 8: run<8, S, V, T>(in, out); return 0;
    case 12: run<12, S, V, T>(in, out); return 0;
    case 16: run<16, S, V, T>(in, out); return 0;
    case 24: run<24, S, V, T>(in, out); return 0;
    case 32: run<32, S, V, T>(in, out); return 0;
    default: return -1;
  }
}
}  // namespace

// X[k] = sum_n x[n] e^{s 2 pi i n k / r}, s = +1 or -1; fp32 != 0 computes in float.
extern "C" int dft_c(int r, int s, int fp32, const double* in, double* out) {
  if (fp32) return s > 0 ? dispatch<1, float2, float>(r, in, out) : dispatch<-1, float2, float>(r, in, out);
  return s > 0 ? dispatch<1, double2, double>(r, in, out) : dispatch<-1, double2, double>(r, in, out);
}
