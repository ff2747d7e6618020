// Register-resident FFT building blocks.
//
// A line of length L = N1 * N2 is transformed by a group of threads in two
// steps with one LDS exchange (Cooley-Tukey, x[n2 N1 + n1] -> X[k1 N2 + k2]):
//   step A: thread n1 (< N1) holds x[n2 N1 + n1], n2 < N2, does an N2-point
//           DFT in registers, multiplies by W_L^(n1 k2) and writes row n1 of a
//           padded [N1][N2+1] LDS tile;
//   step B: thread k2 (< N2) reads column k2, does an N1-point DFT and holds
//           X[k1 N2 + k2], k1 < N1.
// The small DFTs are unrolled at compile time with constant twiddles.
#pragma once
#include <hip/hip_runtime.h>

namespace chf {

// ---------------------------------------------------- compile-time twiddles
namespace ct {
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {
  double x2 = x * x, term = x, sum = x;
  for (int n = 1; n < 14; ++n) {
    term *= -x2 / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int n = 1; n < 14; ++n) {
    term *= -x2 / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}
// cos and sin of 2 pi j / R, reduced exactly to |r| <= pi/4 around a multiple of pi/2
constexpr double cos2pi(long j, long R) {
  long jj = ((j % R) + R) % R;
  long q = (4 * jj + R / 2) / R;                 // nearest quarter turn
  double r = 2.0 * kPi * (double)(4 * jj - q * R) / (4.0 * (double)R);
  switch (q & 3) {
    case 0: return taylor_cos(r);
    case 1: return -taylor_sin(r);
    case 2: return -taylor_cos(r);
    default: return taylor_sin(r);
  }
}
constexpr double sin2pi(long j, long R) {
  long jj = ((j % R) + R) % R;
  long q = (4 * jj + R / 2) / R;
  double r = 2.0 * kPi * (double)(4 * jj - q * R) / (4.0 * (double)R);
  switch (q & 3) {
    case 0: return taylor_sin(r);
    case 1: return taylor_cos(r);
    case 2: return -taylor_sin(r);
    default: return -taylor_cos(r);
  }
}
}  // namespace ct

template <class T>
struct C2;  // complex of T
template <>
struct C2<double> {
  using type = double2;
  static __device__ __forceinline__ double2 make(double a, double b) { return make_double2(a, b); }
};
template <>
struct C2<float> {
  using type = float2;
  static __device__ __forceinline__ float2 make(float a, float b) { return make_float2(a, b); }
};

// Stores marked non-temporal, for outputs that the next pass reads and this
// one does not: streaming them past L2 leaves it to the rows the pass itself
// reads again.
__device__ __forceinline__ void st_nt(double2* p, double2 v) {
  typedef double e2 __attribute__((ext_vector_type(2)));
  const e2 t = {v.x, v.y};
  __builtin_nontemporal_store(t, reinterpret_cast<e2*>(p));
}
__device__ __forceinline__ void st_nt(float2* p, float2 v) {
  typedef float e2 __attribute__((ext_vector_type(2)));
  const e2 t = {v.x, v.y};
  __builtin_nontemporal_store(t, reinterpret_cast<e2*>(p));
}

template <class V>
__host__ __device__ __forceinline__ V cadd(V a, V b) {
  a.x += b.x;
  a.y += b.y;
  return a;
}
template <class V>
__host__ __device__ __forceinline__ V csub(V a, V b) {
  a.x -= b.x;
  a.y -= b.y;
  return a;
}
template <class V, class T>
__host__ __device__ __forceinline__ V cmulc(V a, T c, T s) {  // a * (c + i s)
  V r;
  r.x = a.x * c - a.y * s;
  r.y = a.x * s + a.y * c;
  return r;
}
template <class V>
__host__ __device__ __forceinline__ V cmulv(V a, V b) {
  V r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// a (S i)^q: a twiddle that is a whole number of quarter turns, applied as a
// swap and sign changes (a product with the exact 0 and +-1 of the table costs
// four arithmetic instructions, which IEEE semantics keep the compiler from
// removing, and gives the same values).
template <int S, class V>
__host__ __device__ __forceinline__ V quarter_turns(V a, int q) {
  V r;
  switch (((S > 0 ? q : -q) % 4 + 4) % 4) {
    case 0: r = a; break;
    case 1: r.x = -a.y; r.y = a.x; break;   // i a
    case 2: r.x = -a.x; r.y = -a.y; break;  // -a
    default: r.x = a.y; r.y = -a.x; break;  // -i a
  }
  return r;
}

// In-place, natural-order DFT of R points, X[k] = sum_n x[n] e^{S 2 pi i n k / R}.
template <int R, int S, class V>
struct Dft;

template <int S, class V>
struct Dft<1, S, V> {
  static __host__ __device__ __forceinline__ void run(V*) {}
};

template <int S, class V>
struct Dft<2, S, V> {
  static __host__ __device__ __forceinline__ void run(V* v) {
    V a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int S, class V>
struct Dft<3, S, V> {
  static __host__ __device__ __forceinline__ void run(V* v) {
    using T = decltype(v[0].x);
    const T h = (T)0.86602540378443864676 * (T)S;
    V s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    V m;
    m.x = v[0].x - (T)0.5 * s.x;
    m.y = v[0].y - (T)0.5 * s.y;
    V r;  // i * S * sqrt(3)/2 * d
    r.x = -h * d.y;
    r.y = h * d.x;
    v[0] = cadd(v[0], s);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
};

template <int S, class V>
struct Dft<4, S, V> {
  static __host__ __device__ __forceinline__ void run(V* v) {
    V a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    V a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
    V ia3;  // i * S * a3
    ia3.x = -(decltype(a3.x))S * a3.y;
    ia3.y = (decltype(a3.x))S * a3.x;
    v[0] = cadd(a0, a2);
    v[1] = cadd(a1, ia3);
    v[2] = csub(a0, a2);
    v[3] = csub(a1, ia3);
  }
};

template <int R>
struct Factor {
  static constexpr int F = (R % 4 == 0) ? 4 : ((R % 2 == 0) ? 2 : 3);
};

// R = F * M: x[F m + f] -> DFT_M over m for each f, twiddle W_R^(f k'), DFT_F over f.
template <int R, int S, class V>
struct Dft {
  static_assert(R > 4, "base cases cover R <= 4");
  static __host__ __device__ __forceinline__ void run(V* v) {
    using T = decltype(v[0].x);
    constexpr int F = Factor<R>::F, M = R / F;
    V y[F][M];
#pragma unroll
    for (int f = 0; f < F; ++f)
#pragma unroll
      for (int m = 0; m < M; ++m) y[f][m] = v[F * m + f];
#pragma unroll
    for (int f = 0; f < F; ++f) Dft<M, S, V>::run(y[f]);
#pragma unroll
    for (int k = 0; k < M; ++k) {
      V z[F];
#pragma unroll
      for (int f = 0; f < F; ++f) {
        if (f == 0 || k == 0) {
          z[f] = y[f][k];
        } else if ((4 * f * k) % R == 0) {
          z[f] = quarter_turns<S>(y[f][k], (4 * f * k) / R);  // W = (S i)^q exactly
        } else {
          const T c = (T)ct::cos2pi((long)f * k, R);
          const T s = (T)(S * ct::sin2pi((long)f * k, R));
          z[f] = cmulc(y[f][k], c, s);
        }
      }
      Dft<F, S, V>::run(z);
#pragma unroll
      for (int q = 0; q < F; ++q) v[k + M * q] = z[q];
    }
  }
};

// Two-level line transform, see the header comment.  `buf` is the group's
// padded LDS tile of N1 * (N2 + 1) complex values; `tw` holds e^{-2 pi i j / L}
// for j < L (conjugated for S = +1).  `t` is the thread's index in its group.
// The caller synchronises between step A and step B.
template <int N1, int N2, int S, class V>
__device__ __forceinline__ void fft_stepA(V (&a)[N2], V* buf, const V* tw, int t) {
  using T = decltype(a[0].x);
  Dft<N2, S, V>::run(a);
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) {
    V w = tw[t * k2];
    if (S > 0) w.y = -w.y;
    buf[t * (N2 + 1) + k2] = (k2 == 0) ? a[0] : cmulv(a[k2], w);
  }
  (void)sizeof(T);
}

template <int N1, int N2, int S, class V>
__device__ __forceinline__ void fft_stepB(V (&b)[N1], const V* buf, int t) {
#pragma unroll
  for (int n1 = 0; n1 < N1; ++n1) b[n1] = buf[n1 * (N2 + 1) + t];
  Dft<N1, S, V>::run(b);
}

}  // namespace chf
