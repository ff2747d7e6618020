// Wall-normal Chebyshev contractions.  Every wall-normal operation of the
// timestep is a small real matrix applied to all (kx, kz) columns of a
// [rows][Mz][Mx] complex array; Re and Im are independent real columns.
// The symmetric collocation grids let each contraction split by parity:
// T_l(-y) = (-1)^l T_l(y), so even and odd degrees are contracted separately
// against half the points and recombined as (E + O, E - O) at mirrored points.
//
//   synth_padded : degree-N coefficients -> values at the 2N+1 points
//                  y_p = cos(pi p/2N) (plan13 synthesis, PF:1659-1680)
//   analysis     : values at the 2N+1 points -> coefficients 0..N
//                  (plan13 analysis and truncation, PF:1686-1704)
//   synth_points : coefficients -> values at the N+1 points (plan12, PF:1573-1609)
//   rhs          : explicit Crank-Nicolson operators plus AB2 forcing of the v
//                  and omega_y equations (get_coef_rhs PF:3385-3425 with
//                  PF:252-257; get_rhs_omegay PF:3458-3489 with PF:335-340),
//                  emitted as the parity-folded right-hand sides
//   hassemble    : h_v = k^2 L_y + i alpha D L_x + i beta D L_z,
//                  h_g = i alpha L_z - i beta L_x (PF:1502-1540, dz PF:570-599)
#include <type_traits>

#include "chf_internal.h"
#include "mfma.h"

namespace chf {

// CB consecutive values from dst (a multiple of CB elements from an aligned
// base) as 8- or 16-byte vector stores.
template <int CB, class T>
__device__ __forceinline__ void store_row(T* dst, const T (&v)[CB]) {
  static_assert(CB == 2 || CB == 4, "vector stores of 2 or 4 values");
  if constexpr (sizeof(T) == 4) {
    using VT = typename std::conditional<CB == 2, float2, float4>::type;
    *reinterpret_cast<VT*>(dst) = *reinterpret_cast<const VT*>(v);
  } else {
#pragma unroll
    for (int c = 0; c < CB; c += 2) *reinterpret_cast<double2*>(dst + c) = make_double2(v[c], v[c + 1]);
  }
}

// ---------------------------------------------------------------- synthesis
// out[p] = E + O, out[P-1-p] = E - O for p < Ph, E/O = even/odd degree sums,
// with A_m(p, j) = T_{2j+m}(y_p) = cos(pi f p (2j+m) / 2N) (f = 1 on the
// 2N+1-point grid, f = 2 on the N+1-point grid).
template <class TI, class TO, class TCc>
struct SynthP {
  using TC = TCc;
  static constexpr int NACC = 2, NB = 2, KC = 16;  // KC: k rows per B tile
  static constexpr bool kRecurrence = false;
  static constexpr bool kCentre = true;
  static constexpr bool kWavePerRowBlock = false;  // four-wave blocks (measured faster)
  static constexpr bool kVecCols = true;  // each lane's adjacent columns in vector stores
  // measured: FP32 five waves 0.372 -> 0.366 ms; FP64 six waves 0.640 -> 0.625 ms
  static constexpr int kMinWaves = sizeof(TCc) == 8 ? 6 : 5;
  static constexpr int bsel(int m) { return m; }
  const TC* costab;
  const double* rtab;  // FP64 cosines for the recurrence form
  int n4;        // 4N
  int f;         // grid factor
  MfmaShape sh;  // M = folded points (without the centre row when `centre`), K = ne
  const TI* inf[6];
  TO* outf[6];
  const TI* in;
  TO* out;
  int Ny, P;
  int centre;    // P odd and the centre point y = 0 summed outside the matrix cores
  long ncold;
  __device__ void select(int z) {
    in = pick6(inf, z);
    out = pick6(outf, z);
  }
  // table indices in 32-bit unsigned arithmetic (products below 2^14 for N <= 256)
  __device__ int aindex(int m, int row, int k0, int n4_) const {
    return (int)((unsigned)(f * row * (2 * k0 + m)) % (unsigned)n4_);
  }
  __device__ int astep(int, int row, int n4_) const { return (int)((unsigned)(8 * f * row) % (unsigned)n4_); }
  // B streams: even and odd coefficients (raw loads; converted when staged)
  using TL = TI;
  static constexpr int NR = 2;
  __device__ void fetchB(int j, long col, TL* r) const {
    const int l = 2 * j;
    r[0] = in[l * ncold + col];
    r[1] = (l + 1 < Ny) ? in[(l + 1) * ncold + col] : (TL)0;
  }
  __device__ void makeB(int, const TL* r, TC* v) const {
    v[0] = (TC)r[0];
    v[1] = (TC)r[1];
  }
  __device__ void store(int r, long col, const TC* acc) const {
    out[r * ncold + col] = (TO)(acc[0] + acc[1]);
    const int q = P - 1 - r;
    if (q != r) out[q * ncold + col] = (TO)(acc[0] - acc[1]);
  }
  // CB adjacent columns from col (col a multiple of CB), as vector stores
  template <int CB>
  __device__ void store_vec(int r, long col, const TC (&acc)[2][CB]) const {
    TO e[CB], o[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      e[c] = (TO)(acc[0][c] + acc[1][c]);
      o[c] = (TO)(acc[0][c] - acc[1][c]);
    }
    store_row<CB>(out + r * ncold + col, e);
    const int q = P - 1 - r;
    if (q != r) store_row<CB>(out + q * ncold + col, o);
  }
  // y = 0: T_l(0) = (-1)^(l/2) for even l, 0 for odd l; even-stream row kk of
  // tile kc is l = 2 (KC kc + kk), KC even
  static __device__ TC centre_sign(int kk) { return (kk & 1) ? (TC)-1 : (TC)1; }
  __device__ void store_centre(long col, TC v) const { out[(long)(P / 2) * ncold + col] = (TO)v; }
};

template <class TI, class TO, class TC>
static hipError_t synth_t(Ctx& c, const TC* costab, int f, int Ph, int P, const TI* const* in,
                          TO* const* out, int nf, int cb) {
  SynthP<TI, TO, TC> pb{};
  pb.costab = costab;
  pb.rtab = c.costab;
  pb.n4 = 4 * c.d.N;
  pb.f = f;
  // with P odd the centre row would be a lone row of the last 16-row block
  // (N = 192: padded grid 13 -> 12 row blocks, Gauss-Lobatto points 7 -> 6)
  pb.centre = P & 1;
  pb.sh = mfma_shape(pb.centre ? Ph - 1 : Ph, c.d.ne);
  for (int i = 0; i < nf; ++i) {
    pb.inf[i] = in[i];
    pb.outf[i] = out[i];
  }
  pb.Ny = c.d.Ny;
  pb.P = P;
  pb.ncold = 2 * c.d.ncol;
  return run_mfma_cos(c.stream, pb, nf, cb);
}

static void as_real(const double2* const* in, const double** out, int nf) {
  for (int i = 0; i < nf; ++i) out[i] = reinterpret_cast<const double*>(in[i]);
}

// The padded synthesis, the analysis and the FP64 point values below run as
// FFTs along y (ydct.hip) when 2N has an FFT split (c.twy set), as cos-table
// products on the matrix cores otherwise.  At 256x193x256, FP64: padded
// synthesis 0.625 -> 0.46 ms, analysis 0.365 -> 0.23 ms, point values 0.134 ->
// 0.089 ms; FP32 arithmetic (mixed form): synthesis 0.35 -> 0.30 ms, analysis
// 0.20 -> 0.16 ms, h point values 0.092 -> 0.082 ms.
hipError_t launch_synth_padded(Ctx& c, const double2* const* in, double2* const* out, int nf) {
  if (c.twy) return launch_ysynth(c, in, out, nf);
  const double* ri[6];
  double* ro[6];
  as_real(in, ri, nf);
  for (int i = 0; i < nf; ++i) ro[i] = reinterpret_cast<double*>(out[i]);
  return synth_t<double, double, double>(c, c.costab, 1, c.d.Ph, c.d.Py, ri, ro, nf, c.cb_choice[K_SYNTH_PADDED]);
}

hipError_t launch_synth_points(Ctx& c, const double2* const* in, double2* const* out, int nf) {
  if (c.twy) return launch_ypoints(c, in, out, nf);
  const double* ri[6];
  double* ro[6];
  as_real(in, ri, nf);
  for (int i = 0; i < nf; ++i) ro[i] = reinterpret_cast<double*>(out[i]);
  return synth_t<double, double, double>(c, c.costab, 2, c.d.Pn, c.d.Ny, ri, ro, nf, c.cb_choice[K_SYNTH_POINTS]);
}

// h_v, h_g at the N+1 points (the AB2 forcing history): FP64 arrays in both
// forms.  With the FFT transforms they are FP64 arithmetic in both forms, as
// the right-hand-side kernels that form them there (a run that forms them in
// either place stays bit-identical).  With the cos-table products the mixed
// form computes them on the FP32 matrix cores: the values enter the solve only
// through dt (3 h^n - h^{n-1}) / 2, as the FP32 Lamb vector they are formed
// from does (measured at 256x193x256 over 100 steps, against the FP64 form:
// field errors within 3% of those with FP64 point values, wall shear
// 3.1e-10 / 1.2e-9 against 3.6e-10 / 1.0e-9).
hipError_t launch_synth_points_h(Ctx& c, const double2* const* in, double2* const* out, int nf) {
  if (c.prm.precision == 0 || c.twy) return launch_synth_points(c, in, out, nf);
  const double* ri[6];
  double* ro[6];
  as_real(in, ri, nf);
  for (int i = 0; i < nf; ++i) ro[i] = reinterpret_cast<double*>(out[i]);
  return synth_t<double, double, float>(c, c.costabf, 2, c.d.Pn, c.d.Ny, ri, ro, nf, c.cb_choice[K_SYNTH_POINTS]);
}

hipError_t launch_analysis32f(Ctx& c, const float2* const* in, float2* const* out, int nf) {
  if (!c.twy) return hipErrorNotSupported;
  return launch_yanalysis32f(c, in, out, nf);
}

hipError_t launch_synth_padded32f(Ctx& c, const float2* const* in, float2* const* out, int nf) {
  if (!c.twy) return hipErrorNotSupported;
  return launch_ysynth32f(c, in, out, nf);
}

// Mixed form: FP64 coefficients -> FP32 values on the padded grid, FP32 matrix cores.
hipError_t launch_synth_padded32(Ctx& c, const double2* const* in, float2* const* out, int nf) {
  if (c.twy) return launch_ysynth32(c, in, out, nf);
  const double* ri[6];
  float* ro[6];
  as_real(in, ri, nf);
  for (int i = 0; i < nf; ++i) ro[i] = reinterpret_cast<float*>(out[i]);
  return synth_t<double, float, float>(c, c.costabf, 1, c.d.Ph, c.d.Py, ri, ro, nf, c.cb_choice[K_SYNTH_PADDED]);
}

// ----------------------------------------------------------------- analysis
// coefficient k = 2j+m: (1/(2N c_k)) sum_p w_p cos(pi k p/2N) (x_p +- x_{2N-p})
// over the folded points p = 0..N (PF:1686-1700: /2N, c_0 = 2, w_0 = 1, else 2).
// For N even the last even coefficient, k = N, has the weights cos(pi p/2) =
// 1, 0, -1, 0, ... : it would be a lone row of a 16-row block (N = 192: 97
// even rows = 6 blocks + 1), so it is summed outside the matrix cores, as the
// synthesis sums its centre point.
template <class TI, class TO, class TCc>
struct AnalP {
  using TC = TCc;
  static constexpr int NACC = 2, NB = 2, KC = 16;  // KC: k rows per B tile
  static constexpr bool kRecurrence = true;
  static constexpr bool kCentre = true;
  static constexpr bool kWavePerRowBlock = false;  // four-wave blocks (measured faster)
  static constexpr bool kVecCols = true;  // each lane's adjacent columns in vector stores
  static constexpr int kMinWaves = 1;
  static constexpr int bsel(int m) { return m; }
  const TC* costab;
  const double* rtab;  // FP64 cosines for the recurrence form
  int n4;
  MfmaShape sh;  // M = ne, K = Ph
  const TI* inf[6];
  TO* outf[6];
  const TI* in;
  TO* out;
  int P, N;
  int centre;    // N even: coefficient k = N summed outside the matrix cores
  long ncold;
  double inv2n;
  __device__ void select(int z) {
    in = pick6(inf, z);
    out = pick6(outf, z);
  }
  __device__ int aindex(int m, int row, int k0, int n4_) const {
    return (int)((unsigned)((2 * row + m) * k0) % (unsigned)n4_);
  }
  // weights cos(pi N p / 2N) = cos(pi p / 2) of the even stream, p = k mod 4
  static __device__ TC centre_sign(int kk) { return (kk & 1) ? (TC)0 : ((kk & 2) ? (TC)-1 : (TC)1); }
  __device__ void store_centre(long col, TC v) const { out[(long)N * ncold + col] = (TO)((double)v * inv2n); }
  __device__ int astep(int m, int row, int n4_) const { return (int)((unsigned)(4 * (2 * row + m)) % (unsigned)n4_); }
  // B streams: weighted folded sums w_p (x_p +- x_{P-1-p}); the loads are
  // issued ahead, the folding is done when the tile is staged
  using TL = TI;
  static constexpr int NR = 2;
  __device__ void fetchB(int p, long col, TL* r) const {
    const int q = P - 1 - p;
    r[0] = in[p * ncold + col];
    r[1] = (q != p) ? in[q * ncold + col] : (TL)0;
  }
  __device__ void makeB(int p, const TL* r, TC* v) const {
    const TC x = (TC)r[0];
    const int q = P - 1 - p;
    const TC w = (p == 0) ? (TC)1 : (TC)2;
    if (q == p) {
      v[0] = w * x;
      v[1] = 0;
    } else {
      const TC y = (TC)r[1];
      v[0] = w * (x + y);
      v[1] = w * (x - y);
    }
  }
  __device__ void store(int r, long col, const TC* acc) const {
    const int k = 2 * r;
    out[k * ncold + col] = (TO)((double)acc[0] * (k == 0 ? 0.5 * inv2n : inv2n));
    if (k + 1 <= N) out[(k + 1) * ncold + col] = (TO)((double)acc[1] * inv2n);
  }
  template <int CB>
  __device__ void store_vec(int r, long col, const TC (&acc)[2][CB]) const {
    const int k = 2 * r;
    TO e[CB], o[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      e[c] = (TO)((double)acc[0][c] * (k == 0 ? 0.5 * inv2n : inv2n));
      o[c] = (TO)((double)acc[1][c] * inv2n);
    }
    store_row<CB>(out + k * ncold + col, e);
    if (k + 1 <= N) store_row<CB>(out + (k + 1) * ncold + col, o);
  }
};

template <class TI, class TC>
static hipError_t analysis_t(Ctx& c, const TC* costab, const TI* const* in, double2* const* out, int nf) {
  const int cb = c.cb_choice[K_ANALYSIS];
  AnalP<TI, double, TC> pb{};
  pb.costab = costab;
  pb.rtab = c.costab;
  pb.n4 = 4 * c.d.N;
  pb.centre = (c.d.N % 2 == 0) ? 1 : 0;
  pb.sh = mfma_shape(c.d.ne - pb.centre, c.d.Ph);
  for (int i = 0; i < nf; ++i) {
    pb.inf[i] = in[i];
    pb.outf[i] = reinterpret_cast<double*>(out[i]);
  }
  pb.P = c.d.Py;
  pb.N = c.d.N;
  pb.ncold = 2 * c.d.ncol;
  pb.inv2n = 1.0 / (2.0 * c.d.N);
  return run_mfma_cos(c.stream, pb, nf, cb);
}

hipError_t launch_analysis(Ctx& c, const double2* const* in, double2* const* out, int nf) {
  if (c.twy) return launch_yanalysis(c, in, out, nf);
  const double* ri[6];
  as_real(in, ri, nf);
  return analysis_t<double, double>(c, c.costab, ri, out, nf);
}

// Mixed form: FP32 values on the padded grid -> FP64 coefficients, FP32 matrix cores.
hipError_t launch_analysis32(Ctx& c, const float2* const* in, double2* const* out, int nf) {
  if (c.twy) return launch_yanalysis32(c, in, out, nf);
  const float* ri[6];
  for (int i = 0; i < nf; ++i) ri[i] = reinterpret_cast<const float*>(in[i]);
  return analysis_t<float, float>(c, c.costabf, ri, out, nf);
}

// ----------------------------------------------------------------------- rhs

__device__ __forceinline__ double col_k2(long col, int Mx, const double* al, const double* be) {
  const long cm = col >> 1;
  const int kz = (int)((unsigned)cm / (unsigned)Mx);
  const int kx = (int)(cm - (long)kz * Mx);
  return al[kx] * al[kx] + be[kz] * be[kz];
}

// The explicit Crank-Nicolson operators are polynomials in D^2 evaluated at
// the collocation points: T2 = T0 D^2 and T4 = T0 D^4 as operators on
// coefficient vectors, so
//   B_v v = T0 [c D^4 v + (1 - 2 c k^2) D^2 v - k^2 (1 - c k^2) v],
//   B_w w = T0 [c D^2 w + (1 - c k^2) w],           c = dt / (2 re).
// The kinematics pass (k_kinematics) forms the bracketed coefficient vectors
// g_v, g_w in its backward derivative sweep, whenever v and omega_y change
// (after each solve, chf_set_state and chf_init_fields); the right-hand sides
// are then cos-table syntheses at the interior points plus the AB2 forcing
// 3 dt/2 h^n - dt/2 h^{n-1} (PF:252-257, 335-340), folded by parity.

// rows r = 0.. at points p = r + POFF (POFF = 2 for v, 1 for omega_y)
template <int POFF>
struct RhsSynP {
  using TC = double;
  static constexpr int NACC = 2, NB = 2, KC = 16;  // KC: k rows per B tile
  static constexpr bool kRecurrence = false;
  static constexpr bool kCentre = false;
  static constexpr bool kWavePerRowBlock = true;  // 5-6 row blocks: one wave each (measured faster)
  static constexpr bool kVecCols = false;
  static constexpr int kMinWaves = 1;
  static constexpr int bsel(int m) { return m; }
  const double* costab;
  int n4;
  MfmaShape sh;  // M = ne - POFF, K = ne
  const double* g;
  const double *hn, *hp;
  double *outE, *outO;  // omega_y (POFF = 1): [row][real column] per parity
  double* outT;         // v (POFF = 2): [complex column][Rv rows: even, then odd][Re, Im]
  int Rv;
  int N, no;
  long ncold;
  double dt;
  __device__ void select(int) {}
  __device__ int aindex(int m, int row, int k0, int n4_) const {
    return (int)((unsigned)(2 * (row + POFF) * (2 * k0 + m)) % (unsigned)n4_);  // T_l(y_p) = cos(pi p l / N)
  }
  __device__ int astep(int, int row, int n4_) const { return (int)((unsigned)(16 * (row + POFF)) % (unsigned)n4_); }
  using TL = double;
  static constexpr int NR = 2;
  __device__ void fetchB(int j, long col, TL* r) const {
    r[0] = g[(2 * j) * ncold + col];
    r[1] = (2 * j + 1 <= N) ? g[(2 * j + 1) * ncold + col] : 0.0;
  }
  __device__ void makeB(int, const TL* r, double* v) const {
    v[0] = r[0];
    v[1] = r[1];
  }
  __device__ void store(int r, long col, const double* acc) const {
    const int p = r + POFF, q = N - p;
    const double fp = 3.0 * dt / 2.0 * hn[p * ncold + col] - dt / 2.0 * hp[p * ncold + col];
    const double fq = 3.0 * dt / 2.0 * hn[q * ncold + col] - dt / 2.0 * hp[q * ncold + col];
    const double sym = (p == q) ? fp : 0.5 * (fp + fq);
    if constexpr (POFF == 2) {
      // a mode's record is [even rows][odd rows]: a block writes the rows of
      // its columns in 64-byte runs, and the solve reads each mode in one run
      double* rec = outT + (col >> 1) * (2L * Rv) + (col & 1);
      rec[2 * r] = acc[0] + sym;
      if (r < no - POFF) rec[2 * (sh.M + r)] = acc[1] + 0.5 * (fp - fq);
    } else {
      outE[r * ncold + col] = acc[0] + sym;
      if (r < no - POFF) outO[r * ncold + col] = acc[1] + 0.5 * (fp - fq);
    }
  }
};

template <int POFF>
static hipError_t rhs_syn(Ctx& c, const double2* g, const double2* hn, const double2* hp, double2* oE,
                          double2* oO, double2* oT) {
  const int cb = c.cb_choice[POFF == 2 ? K_RHS_V : K_RHS_O];
  RhsSynP<POFF> pb{};
  pb.costab = c.costab;
  pb.n4 = 4 * c.d.N;
  pb.sh = mfma_shape(c.d.ne - POFF, c.d.ne);
  pb.g = reinterpret_cast<const double*>(g);
  pb.hn = reinterpret_cast<const double*>(hn);
  pb.hp = reinterpret_cast<const double*>(hp);
  pb.outE = reinterpret_cast<double*>(oE);
  pb.outO = reinterpret_cast<double*>(oO);
  pb.outT = reinterpret_cast<double*>(oT);
  pb.Rv = c.d.ne + c.d.no - 4;
  pb.N = c.d.N;
  pb.no = c.d.no;
  pb.ncold = 2 * c.d.ncol;
  pb.dt = c.prm.dt;
  return run_mfma_cos(c.stream, pb, 1, cb);
}

hipError_t launch_rhs(Ctx& c, int which) {
  const int nxt = c.hcur, prv = 1 - c.hcur;  // h^n, h^{n-1}
  if (c.twy) {
    // FFT form (ydct.hip): the bracket's point values by the type-I DCT, and
    // h^n's from its coefficients when they are not formed yet
    if (which == 0) return launch_yrhs(c, 0, c.gv, c.hdefer ? c.hvc : nullptr, c.hv[nxt], c.hv[prv], nullptr, nullptr, c.rvT);
    return launch_yrhs(c, 1, c.go, c.hdefer ? c.hgc : nullptr, c.hg[nxt], c.hg[prv], c.roE, c.roO, nullptr);
  }
  if (which == 0) return rhs_syn<2>(c, c.gv, c.hv[nxt], c.hv[prv], nullptr, nullptr, c.rvT);
  return rhs_syn<1>(c, c.go, c.hg[nxt], c.hg[prv], c.roE, c.roO, nullptr);
}

// ---------------------------------------------------------------- h assembly

// L^n coefficients in FP64, or (mixed form with FFT transforms) in FP32: the
// FP32 analysis's results, exact in either width
struct LIn {
  const double2* d;
  const float2* f;
  // the stored value (float2 or double2); widened where it is used, so that
  // a batch of loads stays in flight
  template <bool F32>
  __device__ __forceinline__ auto raw(long i) const {
    if constexpr (F32) return f[i];
    else return d[i];
  }
  template <bool F32>
  __device__ __forceinline__ double2 at(long i) const {
    const auto v = raw<F32>(i);
    return make_double2(v.x, v.y);
  }
};

template <class V>
__device__ __forceinline__ double2 widen(V v) {
  return make_double2(v.x, v.y);
}

struct MArgs {
  LIn lx, lz;
  double2 *o1, *o3;
  const double* tpts;  // [Pn][Ny]
  int N, Ny, Pn;
  long ncol;
};

template <bool F32>
__device__ void mean_points_block(const MArgs& a) {
  __shared__ double2 s1[260], s3[260];
  for (int l = threadIdx.x; l <= a.N; l += blockDim.x) {
    s1[l] = a.lx.at<F32>((long)l * a.ncol);
    s3[l] = a.lz.at<F32>((long)l * a.ncol);
  }
  __syncthreads();
  for (int jj = threadIdx.x; jj < a.Pn; jj += blockDim.x) {
    double2 e1 = make_double2(0, 0), o1 = make_double2(0, 0);
    double2 e3 = make_double2(0, 0), o3 = make_double2(0, 0);
    for (int l = 0; l <= a.N; ++l) {
      const double t = a.tpts[jj * a.Ny + l];
      const double2 x1 = s1[l], x3 = s3[l];
      if (l & 1) {
        o1.x += t * x1.x; o1.y += t * x1.y;
        o3.x += t * x3.x; o3.y += t * x3.y;
      } else {
        e1.x += t * x1.x; e1.y += t * x1.y;
        e3.x += t * x3.x; e3.y += t * x3.y;
      }
    }
    a.o1[jj] = make_double2(e1.x + o1.x, e1.y + o1.y);
    a.o3[jj] = make_double2(e3.x + o3.x, e3.y + o3.y);
    const int q = a.N - jj;
    if (q != jj) {
      a.o1[q] = make_double2(e1.x - o1.x, e1.y - o1.y);
      a.o3[q] = make_double2(e3.x - o3.x, e3.y - o3.y);
    }
  }
}


struct HArgs {
  LIn lx, ly, lz;
  double2 *hv, *hg;
  const double *al, *be;
  int Mx, N;
  long ncol;
};

// One thread per complex column; the derivative runs the backward recurrence
// c_k b_k = b_{k+2} + 2 (k+1) a_{k+1}, which equals the reference's explicit
// sum b_k = (2/c_k) sum_{j>k, j+k odd} j a_j.
// The launch carries one extra block that computes the mean-mode Lamb point
// values (mean_points_block) from the same coefficients.
template <bool F32>
__global__ __launch_bounds__(256) void k_hassemble(HArgs a, MArgs m) {
  if (blockIdx.x == gridDim.x - 1) {
    mean_points_block<F32>(m);
    return;
  }
  const long cm = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cm >= a.ncol) return;
  const int kz = (int)((unsigned)cm / (unsigned)a.Mx);
  const int kx = (int)(cm - (long)kz * a.Mx);
  const double al = a.al[kx], be = a.be[kz];
  const double k2 = al * al + be * be;
  const int N = a.N;
  double2 bx2 = make_double2(0, 0), bx1 = make_double2(0, 0);
  double2 bz2 = make_double2(0, 0), bz1 = make_double2(0, 0);
  double2 lxn = make_double2(0, 0), lzn = make_double2(0, 0);  // a_{k+1}
  for (int k0 = N; k0 >= 0; k0 -= 4) {
    // loads issued four degrees ahead of the recurrence
    decltype(a.lx.raw<F32>(0)) pl[4][3];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = k0 - t;
      if (k >= 0) {
        const long off = (long)k * a.ncol + cm;
        pl[t][0] = a.lx.raw<F32>(off);
        pl[t][1] = a.ly.raw<F32>(off);
        pl[t][2] = a.lz.raw<F32>(off);
      }
    }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int k = k0 - t;
    if (k < 0) break;
    const long off = (long)k * a.ncol + cm;
    const double2 lx = widen(pl[t][0]), ly = widen(pl[t][1]), lz = widen(pl[t][2]);
    double2 dx, dz;
    if (k == N) {
      dx = make_double2(0, 0);
      dz = make_double2(0, 0);
    } else {
      const double f = 2.0 * (k + 1);
      dx = make_double2(bx2.x + f * lxn.x, bx2.y + f * lxn.y);
      dz = make_double2(bz2.x + f * lzn.x, bz2.y + f * lzn.y);
    }
    bx2 = bx1;
    bx1 = dx;
    bz2 = bz1;
    bz1 = dz;
    if (k == 0) {
      dx.x *= 0.5;
      dx.y *= 0.5;
      dz.x *= 0.5;
      dz.y *= 0.5;
    }
    // h_v = (k^2 L_y + i alpha D L_x) + i beta D L_z
    double hvr = k2 * ly.x - al * dx.y;
    double hvi = k2 * ly.y + al * dx.x;
    hvr = hvr - be * dz.y;
    hvi = hvi + be * dz.x;
    // h_g = i alpha L_z - i beta L_x
    const double hgr = -1.0 * al * lz.y + be * lx.y;
    const double hgi = al * lz.x - be * lx.x;
    a.hv[off] = make_double2(hvr, hvi);
    a.hg[off] = make_double2(hgr, hgi);
    lxn = lx;
    lzn = lz;
  }
  }
}

hipError_t launch_hassemble(Ctx& c) {
  HArgs a{};
  const bool f32 = c.f32_fields;  // the analysis wrote L^n in FP32 (Lcf)
  a.lx = LIn{c.Lc[0], f32 ? c.Lcf[0] : nullptr};
  a.ly = LIn{c.Lc[1], f32 ? c.Lcf[1] : nullptr};
  a.lz = LIn{c.Lc[2], f32 ? c.Lcf[2] : nullptr};
  a.hv = c.hvc;
  a.hg = c.hgc;
  a.al = c.alpha;
  a.be = c.beta;
  a.Mx = c.d.Mx;
  a.N = c.d.N;
  a.ncol = c.d.ncol;
  MArgs m{};
  m.lx = a.lx;
  m.lz = a.lz;
  m.o1 = c.cl1[1 - c.hcur];
  m.o3 = c.cl3[1 - c.hcur];
  m.tpts = c.tpts;
  m.N = c.d.N;
  m.Ny = c.d.Ny;
  m.Pn = c.d.Pn;
  m.ncol = c.d.ncol;
  const dim3 grid((unsigned)((a.ncol + kColumnThreads - 1) / kColumnThreads) + 1);
  if (f32)
    hipLaunchKernelGGL(k_hassemble<true>, grid, dim3(kColumnThreads), 0, c.stream, a, m);
  else
    hipLaunchKernelGGL(k_hassemble<false>, grid, dim3(kColumnThreads), 0, c.stream, a, m);
  return hipGetLastError();
}

// Mean-mode Lamb (column (0,0) of L_x and L_z) -> values at the N+1 points
// (PF:1592-1609), written into the "new" history slot.

}  // namespace chf
