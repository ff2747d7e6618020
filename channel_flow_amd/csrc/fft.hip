// 3/2-padded Fourier transforms of the nonlinear term (ftrthreetotwo,
// PF:1024-1185, with expand PF:721-765 and getall PF:1350-1362), written as
// LDS-resident Stockham FFTs over batches of lines.
//
//   zfft_inv : span modes -> span points, padded Mz -> Lz = 3 nny/2, Nyquist
//              zeroed (plan8, PF:1058-1082)
//   xpass    : for each (y point, span point) line: three packed c2r
//              streamwise FFTs of the six fields (plan9, Hermitian fill and
//              Re part, PF:1097-1119), the Lamb vector
//              L = omega x u - re_p x^ (PF:1438-1473, re_p at PF:1787), two
//              packed r2c FFTs truncated to Mx modes (plan10, PF:1130-1143)
//              and the 1/(Lx Lz) scaling (PF:1176-1177)
//   zfft_fwd : span points -> span modes, truncated keeping slot Mz/2
//              (plan11 and expand(-1), PF:1158-1179)
#include "chf_internal.h"
#include "fftreg.h"

namespace chf {

bool fft_length_ok(int L) {
  if (L < 2) return false;
  while (L % 2 == 0) L /= 2;
  while (L % 3 == 0) L /= 3;
  return L == 1;
}

FftPlan make_plan(int L) {
  FftPlan p{};
  p.L = L;
  int m = L;
  while (m % 4 == 0) { p.rad[p.nst++] = 4; m /= 4; }
  while (m % 2 == 0) { p.rad[p.nst++] = 2; m /= 2; }
  while (m % 3 == 0) { p.rad[p.nst++] = 3; m /= 3; }
  return p;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Batched Stockham FFT of nl lines of length L held in LDS (line stride L).
// sign = -1 forward (e^{-i}), +1 backward; unnormalised.  tw[m] = e^{-2 pi i m/L}.
// Ping-pongs between a and b; returns the buffer holding the result.
__device__ double2* fft_lds(double2* a, double2* b, const FftPlan& pl, int nl, int sign,
                            const double2* __restrict__ tw) {
  const int L = pl.L;
  int Ns = 1;
  double2* src = a;
  double2* dst = b;
  const double s3 = 0.86602540378443864676 * sign;
  for (int st = 0; st < pl.nst; ++st) {
    const int R = pl.rad[st];
    const int NB = L / R;
    const int tstep = L / (Ns * R);
    for (int t = threadIdx.x; t < nl * NB; t += blockDim.x) {
      const int line = t / NB;
      const int j = t - line * NB;
      const double2* s = src + line * L;
      double2* d = dst + line * L;
      const int k = j % Ns;
      const int base = (j / Ns) * Ns * R + k;
      double2 v0 = s[j], v1 = s[j + NB];
      if (R == 2) {
        if (k) {
          double2 w = tw[k * tstep];
          if (sign > 0) w.y = -w.y;
          v1 = cmul(v1, w);
        }
        d[base] = make_double2(v0.x + v1.x, v0.y + v1.y);
        d[base + Ns] = make_double2(v0.x - v1.x, v0.y - v1.y);
      } else if (R == 4) {
        double2 v2 = s[j + 2 * NB], v3 = s[j + 3 * NB];
        if (k) {
          int i1 = k * tstep, i2 = 2 * k * tstep, i3 = 3 * k * tstep;
          double2 w1 = tw[i1], w2 = tw[i2], w3 = tw[i3];
          if (sign > 0) { w1.y = -w1.y; w2.y = -w2.y; w3.y = -w3.y; }
          v1 = cmul(v1, w1);
          v2 = cmul(v2, w2);
          v3 = cmul(v3, w3);
        }
        double2 a0 = make_double2(v0.x + v2.x, v0.y + v2.y);
        double2 a1 = make_double2(v0.x - v2.x, v0.y - v2.y);
        double2 a2 = make_double2(v1.x + v3.x, v1.y + v3.y);
        double2 a3 = make_double2(v1.x - v3.x, v1.y - v3.y);
        // (sign i) * a3
        double2 ia3 = make_double2(-sign * a3.y, sign * a3.x);
        d[base] = make_double2(a0.x + a2.x, a0.y + a2.y);
        d[base + Ns] = make_double2(a1.x + ia3.x, a1.y + ia3.y);
        d[base + 2 * Ns] = make_double2(a0.x - a2.x, a0.y - a2.y);
        d[base + 3 * Ns] = make_double2(a1.x - ia3.x, a1.y - ia3.y);
      } else {  // R == 3
        double2 v2 = s[j + 2 * NB];
        if (k) {
          int i1 = k * tstep, i2 = 2 * k * tstep;
          double2 w1 = tw[i1], w2 = tw[i2];
          if (sign > 0) { w1.y = -w1.y; w2.y = -w2.y; }
          v1 = cmul(v1, w1);
          v2 = cmul(v2, w2);
        }
        double2 sum = make_double2(v1.x + v2.x, v1.y + v2.y);
        double2 dif = make_double2(v1.x - v2.x, v1.y - v2.y);
        double2 m = make_double2(v0.x - 0.5 * sum.x, v0.y - 0.5 * sum.y);
        // (sign i sqrt3/2) * dif
        double2 r = make_double2(-s3 * dif.y, s3 * dif.x);
        d[base] = make_double2(v0.x + sum.x, v0.y + sum.y);
        d[base + Ns] = make_double2(m.x + r.x, m.y + r.y);
        d[base + 2 * Ns] = make_double2(m.x - r.x, m.y - r.y);
      }
    }
    __syncthreads();
    double2* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}

// ----------------------------------------------------------------- span

struct ZArgs {
  const double2* in[6];
  double2* out[6];
  int Mx, Mz, Lz, TX;
  long in_plane;   // complex elements per wall-normal plane of the input
  long out_plane;  // ... of the output
  FftPlan pl;
  const double2* tw;
};

// grid: (Mx/TX, Py, nfields); one block transforms TX span lines of one plane.
__global__ __launch_bounds__(256) void k_zfft_inv(ZArgs a) {
  extern __shared__ double2 sm[];
  const int TX = a.TX, Lz = a.Lz, Mz = a.Mz, h = Mz / 2;
  double2* b0 = sm;
  double2* b1 = sm + TX * Lz;
  const int kx0 = blockIdx.x * TX;
  const int p = blockIdx.y;
  const double2* in = a.in[blockIdx.z] + (long)p * a.in_plane;
  double2* out = a.out[blockIdx.z] + (long)p * a.out_plane;
  for (int idx = threadIdx.x; idx < TX * Lz; idx += blockDim.x) {
    const int m = idx / TX, tx = idx - m * TX;
    int kz = -1;
    if (m < h) kz = m;                       // expand(+1): 0..Mz/2-1, Nyquist zeroed
    else if (m > Lz - h) kz = m - Lz + Mz;   // negative modes
    double2 val = make_double2(0.0, 0.0);
    if (kz >= 0) val = in[(long)kz * a.Mx + kx0 + tx];
    b0[tx * Lz + m] = val;
  }
  __syncthreads();
  double2* r = fft_lds(b0, b1, a.pl, TX, +1, a.tw);
  for (int idx = threadIdx.x; idx < TX * Lz; idx += blockDim.x) {
    const int m = idx / TX, tx = idx - m * TX;
    out[(long)m * a.Mx + kx0 + tx] = r[tx * Lz + m];
  }
}

__global__ __launch_bounds__(256) void k_zfft_fwd(ZArgs a) {
  extern __shared__ double2 sm[];
  const int TX = a.TX, Lz = a.Lz, Mz = a.Mz, h = Mz / 2;
  double2* b0 = sm;
  double2* b1 = sm + TX * Lz;
  const int kx0 = blockIdx.x * TX;
  const int p = blockIdx.y;
  const double2* in = a.in[blockIdx.z] + (long)p * a.in_plane;
  double2* out = a.out[blockIdx.z] + (long)p * a.out_plane;
  for (int idx = threadIdx.x; idx < TX * Lz; idx += blockDim.x) {
    const int m = idx / TX, tx = idx - m * TX;
    b0[tx * Lz + m] = in[(long)m * a.Mx + kx0 + tx];
  }
  __syncthreads();
  double2* r = fft_lds(b0, b1, a.pl, TX, -1, a.tw);
  for (int idx = threadIdx.x; idx < TX * Mz; idx += blockDim.x) {
    const int kz = idx / TX, tx = idx - kz * TX;
    const int m = (kz <= h) ? kz : kz - Mz + Lz;  // expand(-1) keeps slot Mz/2
    out[(long)kz * a.Mx + kx0 + tx] = r[tx * Lz + m];
  }
}

static int pick_tx(int Mx, int Lz) {
  int tx = 1;
  for (int c : {16, 8, 4, 2}) {
    if (Mx % c == 0 && c * Lz <= 2048) { tx = c; break; }
  }
  return tx;
}

static hipError_t zfft_reg64(Ctx& c, double2* const* in, double2* const* out, int nf, bool inv);
static bool use_zreg(const Ctx& c);

static hipError_t launch_z(Ctx& c, double2* const* in, double2* const* out, int nf, bool inv) {
  const Dims& d = c.d;
  ZArgs a{};
  for (int f = 0; f < nf; ++f) {
    a.in[f] = in[f];
    a.out[f] = out[f];
  }
  a.Mx = d.Mx;
  a.Mz = d.Mz;
  a.Lz = d.Lz;
  a.TX = pick_tx(d.Mx, d.Lz);
  a.in_plane = (long)(inv ? d.Mz : d.Lz) * d.Mx;
  a.out_plane = (long)(inv ? d.Lz : d.Mz) * d.Mx;
  a.pl = make_plan(d.Lz);
  a.tw = c.twz;
  if (use_zreg(c)) return zfft_reg64(c, in, out, nf, inv);
  dim3 grid(d.Mx / a.TX, d.Py, nf);
  size_t shm = 2ull * a.TX * d.Lz * sizeof(double2);
  if (inv)
    hipLaunchKernelGGL(k_zfft_inv, grid, dim3(256), shm, c.stream, a);
  else
    hipLaunchKernelGGL(k_zfft_fwd, grid, dim3(256), shm, c.stream, a);
  return hipGetLastError();
}

// ------------------------------------------------ unpadded physical output

// Streamwise c2r of LPB (y point, span point) lines per block:
// in [Ny][Mz][Mx] complex -> out [Ny][Mz][nnx] real.
struct XPArgs {
  const double2* in;
  double* out;
  int Mx, L, LPB;
  long nlines;
  FftPlan pl;
  const double2* tw;
};

__global__ __launch_bounds__(256) void k_xphys(XPArgs a) {
  extern __shared__ double2 sm[];
  const int L = a.L, Mx = a.Mx, LPB = a.LPB;
  double2* b0 = sm;
  double2* b1 = sm + LPB * L;
  const long line0 = (long)blockIdx.x * LPB;
  for (int idx = threadIdx.x; idx < LPB * L; idx += blockDim.x) {
    const int ln = idx / L, m = idx - ln * L;
    const long line = line0 + ln;
    double2 val = make_double2(0.0, 0.0);
    if (line < a.nlines) {
      const double2* f = a.in + line * Mx;
      if (m == 0) {
        val = make_double2(f[0].x, 0.0);            // Re of the kx = 0 entry
      } else if (m < Mx) {
        val = f[m];
      } else if (m > L - Mx) {
        const double2 g = f[L - m];                 // getall: F(-m) = conj F(m)
        val = make_double2(g.x, -g.y);
      }                                             // m = Mx: Nyquist, zero
    }
    b0[idx] = val;
  }
  __syncthreads();
  double2* r = fft_lds(b0, b1, a.pl, LPB, +1, a.tw);
  for (int idx = threadIdx.x; idx < LPB * L; idx += blockDim.x) {
    const int ln = idx / L, m = idx - ln * L;
    const long line = line0 + ln;
    if (line < a.nlines) a.out[line * L + m] = r[idx].x;
  }
}


hipError_t launch_stream_c2r_unpadded(Ctx& c, const double2* in, double* out) {
  const Dims& d = c.d;
  XPArgs a{};
  a.in = in;
  a.out = out;
  a.Mx = d.Mx;
  a.L = 2 * d.Mx;
  a.LPB = a.L >= 2048 ? 1 : 2048 / a.L;
  a.nlines = (long)d.Ny * d.Mz;
  a.pl = make_plan(a.L);
  a.tw = c.twxu;
  const size_t shm = 2ull * a.LPB * a.L * sizeof(double2);
  hipLaunchKernelGGL(k_xphys, dim3((unsigned)((a.nlines + a.LPB - 1) / a.LPB)), dim3(256), shm, c.stream, a);
  return hipGetLastError();
}

// transform(1, .) span step (fourtrx_zcomp, PF:790-803): Mz span modes ->
// Mz span points per (wall-normal point, kx), Nyquist slot zeroed, in place.
hipError_t launch_span_inverse_unpadded(Ctx& c, double2* inout) {
  const Dims& d = c.d;
  ZArgs a{};
  a.in[0] = inout;
  a.out[0] = inout;
  a.Mx = d.Mx;
  a.Mz = d.Mz;
  a.Lz = d.Mz;
  a.TX = pick_tx(d.Mx, d.Mz);
  a.in_plane = (long)d.Mz * d.Mx;
  a.out_plane = a.in_plane;
  a.pl = make_plan(d.Mz);
  a.tw = c.twzu;
  dim3 grid(d.Mx / a.TX, d.Ny, 1);
  const size_t shm = 2ull * a.TX * d.Mz * sizeof(double2);
  hipLaunchKernelGGL(k_zfft_inv, grid, dim3(256), shm, c.stream, a);
  return hipGetLastError();
}

hipError_t launch_zfft_inv(Ctx& c, double2* const* in, double2* const* out, int nf) {
  return launch_z(c, in, out, nf, true);
}
hipError_t launch_zfft_fwd(Ctx& c, double2* const* in, double2* const* out, int nf) {
  return launch_z(c, in, out, nf, false);
}

// ------------------------------------------------------------- streamwise

struct XArgs {
  const double2* in[6];  // u, v, w, omega_x, omega_y, omega_z  [Py][Lz][Mx]
  double2* out[3];       // L_x, L_y, L_z                        [Py][Lz][Mx]
  int Mx, Lx, LPB;
  long nlines;           // Py * Lz
  double re_p, scale;
  FftPlan pl;
  const double2* tw;
};

// One block handles LPB (y point, span point) lines: 3 packed inverse FFTs,
// products, 2 packed forward FFTs.
__global__ __launch_bounds__(256) void k_xpass(XArgs a) {
  extern __shared__ double2 sm[];
  const int Lx = a.Lx, Mx = a.Mx, LPB = a.LPB;
  const int nl = 3 * LPB;
  double2* b0 = sm;
  double2* b1 = sm + nl * Lx;
  const long line0 = (long)blockIdx.x * LPB;
  // build the packed padded spectra: pair q holds fields (2q, 2q+1)
  for (int idx = threadIdx.x; idx < nl * Lx; idx += blockDim.x) {
    const int ln = idx / Lx, m = idx - ln * Lx;
    const int lb = ln / 3, q = ln - lb * 3;
    const long line = line0 + lb;
    double2 val = make_double2(0.0, 0.0);
    if (line < a.nlines) {
      const double2* f1 = a.in[2 * q] + line * Mx;
      const double2* f2 = a.in[2 * q + 1] + line * Mx;
      if (m == 0) {
        // Hermitian fill keeps Re of the kx=0 entry (Re of the backward FFT, PF:1116-1118)
        val = make_double2(f1[0].x, f2[0].x);
      } else if (m < Mx) {
        double2 A = f1[m], B = f2[m];
        val = make_double2(A.x - B.y, A.y + B.x);          // F1 + i F2
      } else if (m > Lx - Mx) {
        const int mm = Lx - m;
        double2 A = f1[mm], B = f2[mm];
        val = make_double2(A.x + B.y, B.x - A.y);          // conj(F1) + i conj(F2)
      }
    }
    b0[idx] = val;
  }
  __syncthreads();
  double2* r = fft_lds(b0, b1, a.pl, nl, +1, a.tw);
  double2* o = (r == b0) ? b1 : b0;
  // products at the Lx physical points; pack (L_x + i L_y) and (L_z + 0 i)
  for (int idx = threadIdx.x; idx < LPB * Lx; idx += blockDim.x) {
    const int lb = idx / Lx, x = idx - lb * Lx;
    const double2 g0 = r[(3 * lb + 0) * Lx + x];  // u + i v
    const double2 g1 = r[(3 * lb + 1) * Lx + x];  // w + i omega_x
    const double2 g2 = r[(3 * lb + 2) * Lx + x];  // omega_y + i omega_z
    const double uu = g0.x, vv = g0.y, ww = g1.x, oxx = g1.y, oyy = g2.x, ozz = g2.y;
    const double lx = (oyy * ww - a.re_p) - ozz * vv;
    const double ly = ozz * uu - oxx * ww;
    const double lz = oxx * vv - oyy * uu;
    o[(2 * lb + 0) * Lx + x] = make_double2(lx, ly);
    o[(2 * lb + 1) * Lx + x] = make_double2(lz, 0.0);
  }
  __syncthreads();
  double2* other = (o == b0) ? b1 : b0;
  double2* f = fft_lds(o, other, a.pl, 2 * LPB, -1, a.tw);
  for (int idx = threadIdx.x; idx < LPB * Mx; idx += blockDim.x) {
    const int lb = idx / Mx, m = idx - lb * Mx;
    const long line = line0 + lb;
    if (line >= a.nlines) continue;
    const double2 G = f[(2 * lb) * Lx + m];
    const double2 Gc = f[(2 * lb) * Lx + (m == 0 ? 0 : Lx - m)];
    // F1 = (G(m) + conj(G(-m)))/2 ; F2 = (G(m) - conj(G(-m)))/(2i)
    const double2 F1 = make_double2(0.5 * (G.x + Gc.x), 0.5 * (G.y - Gc.y));
    const double2 F2 = make_double2(0.5 * (G.y + Gc.y), -0.5 * (G.x - Gc.x));
    const double2 F3 = f[(2 * lb + 1) * Lx + m];
    const double s = a.scale;
    a.out[0][line * Mx + m] = make_double2(F1.x * s, F1.y * s);
    a.out[1][line * Mx + m] = make_double2(F2.x * s, F2.y * s);
    a.out[2][line * Mx + m] = make_double2(F3.x * s, F3.y * s);
  }
}

// ------------------------------------------- register-resident variants
// Line length L = N1 * N2 handled by a group of G = max(N1, N2) threads with
// one padded LDS exchange per transform (see fftreg.h).  V is double2 (FP64
// form) or float2 (physical-space FP32 of the mixed form).

template <int N1, int N2>
struct Geo {
  static constexpr int L = N1 * N2;
  static constexpr int G = N1 > N2 ? N1 : N2;
  static constexpr int BUF_A = N1 * (N2 + 1);  // inverse/first exchange tile
  // LDS stride between the exchange tiles of the lines of one span-pass block
  // (paddings of 1, 2 and 8 elements measured: no faster)
  static constexpr int LINE_Z = BUF_A;
  static constexpr int BUF_B = N2 * (N1 + 1);  // transposed-order tile
  static constexpr int BUF = (BUF_A > BUF_B ? BUF_A : BUF_B) > L ? (BUF_A > BUF_B ? BUF_A : BUF_B) : L;
};

template <class V>
struct RegX {
  const V* in[6];  // u, v, w, omega_x, omega_y, omega_z  [Py][Lz][Mx]
  V* out[3];       // L_x, L_y, L_z                        [Py][Lz][Mx]
  int Mx;
  long nlines;     // Py * Lz
  double re_p, scale;
  const V* tw;     // e^{-2 pi i m / Lx}
};

template <class V>
struct RegZ {
  const V* in[6];
  V* out[6];
  int Mx, Mz;
  long in_plane, out_plane;
  const V* tw;     // e^{-2 pi i m / Lz}
};

// Streamwise pass: one group per (y point, span point) line.
template <int N1, int N2, int NG, class V>
__global__ __launch_bounds__((NG * Geo<N1, N2>::G)) void k_xpass_reg(RegX<V> a) {
  using Gm = Geo<N1, N2>;
  using T = decltype(V{}.x);
  constexpr int L = Gm::L, G = Gm::G;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  V* tw = sm;
  const int g = threadIdx.x / G, t = threadIdx.x - g * G;
  V* buf = sm + L + g * Gm::BUF;
  const long line = (long)blockIdx.x * NG + g;
  const bool valid = line < a.nlines;
  const int Mx = a.Mx;
  for (int i = threadIdx.x; i < L; i += NG * G) tw[i] = a.tw[i];
  __syncthreads();
  V r[3][N1];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (t < N1) {
      V x[N2];
      const V* f1 = a.in[2 * q] + line * Mx;
      const V* f2 = a.in[2 * q + 1] + line * Mx;
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) {
        const int m = n2 * N1 + t;
        V val = C2<T>::make(0, 0);
        if (valid) {
          if (m == 0) {
            val = C2<T>::make(f1[0].x, f2[0].x);            // Re of the kx = 0 entry (PF:1116-1118)
          } else if (m < Mx) {
            const V A = f1[m], B = f2[m];
            val = C2<T>::make(A.x - B.y, A.y + B.x);         // F1 + i F2
          } else if (m > L - Mx) {
            const V A = f1[L - m], B = f2[L - m];
            val = C2<T>::make(A.x + B.y, B.x - A.y);         // conj(F1) + i conj(F2)
          }
        }
        x[n2] = val;
      }
      fft_stepA<N1, N2, +1>(x, buf, tw, t);
    }
    __syncthreads();
    if (t < N2) fft_stepB<N1, N2, +1>(r[q], buf, t);
    __syncthreads();
  }
  // Lamb vector at the physical points x = k1 N2 + t: (L_x + i L_y) into r[0], (L_z + 0i) into r[1]
  const T rep = (T)a.re_p;
  if (t < N2) {
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      const T uu = r[0][k1].x, vv = r[0][k1].y, ww = r[1][k1].x, oxx = r[1][k1].y;
      const T oyy = r[2][k1].x, ozz = r[2][k1].y;
      const T lx = (oyy * ww - rep) - ozz * vv;
      const T ly = ozz * uu - oxx * ww;
      const T lz = oxx * vv - oyy * uu;
      r[0][k1] = C2<T>::make(lx, ly);
      r[1][k1] = C2<T>::make(lz, 0);
    }
  }
  const T sc = (T)a.scale;
  const T half = (T)0.5;
  // forward transforms with the transposed split: thread t holds x[n2' N2 + t], n2' < N1
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    if (t < N2) fft_stepA<N2, N1, -1>(r[q], buf, tw, t);
    __syncthreads();
    V b[N2];
    if (t < N1) fft_stepB<N2, N1, -1>(b, buf, t);     // X[k1' N1 + t], k1' < N2
    __syncthreads();
    if (q == 0) {
      // separate the packed pair: F1 = (G(m) + conj G(-m))/2, F2 = (G(m) - conj G(-m))/(2i)
      if (t < N1) {
#pragma unroll
        for (int k = 0; k < N2; ++k) buf[k * N1 + t] = b[k];
      }
      __syncthreads();
      if (valid) {
        for (int m = t; m < Mx; m += G) {
          const V Gp = buf[m];
          const V Gc = buf[m == 0 ? 0 : L - m];
          a.out[0][line * Mx + m] = C2<T>::make(half * (Gp.x + Gc.x) * sc, half * (Gp.y - Gc.y) * sc);
          a.out[1][line * Mx + m] = C2<T>::make(half * (Gp.y + Gc.y) * sc, -half * (Gp.x - Gc.x) * sc);
        }
      }
      __syncthreads();
    } else if (t < N1 && valid) {
#pragma unroll
      for (int k = 0; k < N2; ++k) {
        const int m = k * N1 + t;
        if (m < Mx) a.out[2][line * Mx + m] = C2<T>::make(b[k].x * sc, b[k].y * sc);
      }
    }
  }
}

// LDS slot of complex element a in a Stockham buffer: one spare slot every 8
// elements spreads the stride-R writes of the early stages over the banks
// (measured: every 8 beats every 16 and 32; no padding is 27% slower; per
// buffer hand-off, dropping it where both sides are contiguous, or swizzling
// with XOR instead, measured no faster at L = 384).
__device__ __forceinline__ int spad(int a) { return a + (a >> 3); }

// One Stockham autosort radix-R stage over a line of L points held in LDS
// (Ns = product of the radices already applied; butterfly j < L/R, st_stage
// runs them on the 64 lanes of a wavefront, two per lane at L/R = 96):
//   v[r] = in[j + r L/R] * W_{Ns R}^{r (j mod Ns)},  v = DFT_R(v),
//   out[(j / Ns) Ns R + (j mod Ns) + r Ns] = v[r].
// After the last stage the line is in natural order.  `tw` = e^{-2 pi i m / L}.
// ZIN: inputs r known to be zero for every butterfly (not read); WOUT: outputs
// r that are written (the others are not needed by anyone).
template <int L, int R, int Ns, int S, class V, unsigned ZIN = 0u, unsigned WOUT = ~0u>
__device__ __forceinline__ void st_stage1(const V* in, V* out, const V* tw, int j) {
  V v[R];
  const int k = j % Ns;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if ((ZIN >> r) & 1u) {
      v[r].x = 0;
      v[r].y = 0;
    } else {
      v[r] = in[spad(j + r * (L / R))];
    }
  }
  if (Ns > 1) {
#pragma unroll
    for (int r = 1; r < R; ++r) {
      if ((ZIN >> r) & 1u) continue;
      V w = tw[r * k * (L / (Ns * R))];
      if (S > 0) w.y = -w.y;
      v[r] = cmulv(v[r], w);
    }
  }
  Dft<R, S, V>::run(v);
  const int base = (j / Ns) * Ns * R + k;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if ((WOUT >> r) & 1u) out[spad(base + r * Ns)] = v[r];
}

template <int L, int R, int Ns, int S, unsigned ZIN = 0u, unsigned WOUT = ~0u, class V>
__device__ __forceinline__ void st_stage(const V* in, V* out, const V* tw, int lane) {
#pragma unroll
  for (int j = lane; j < L / R; j += 64) st_stage1<L, R, Ns, S, V, ZIN, WOUT>(in, out, tw, j);
}

// The twiddles W_{Ns R}^{r (j mod Ns)} (conjugated for S > 0), r = 1 .. R-1, of
// butterfly j = lane of a stage with L / R <= 64 butterflies: the same for
// every transform of the line, so they are read from the LDS table once per
// line (lanes past L / R read valid entries they never use).
template <int L, int R, int Ns, int S, class V>
struct StTw {
  V w[R - 1];
  __device__ __forceinline__ void load(const V* tw, int lane) {
    const int k = lane % Ns;
#pragma unroll
    for (int r = 1; r < R; ++r) {
      V t = tw[r * k * (L / (Ns * R))];
      if (S > 0) t.y = -t.y;
      w[r - 1] = t;
    }
  }
};

// st_stage with the twiddles of this lane's butterfly in registers (L / R <= 64).
template <int L, int R, int Ns, int S, unsigned ZIN = 0u, unsigned WOUT = ~0u, class V>
__device__ __forceinline__ void st_stage_r(const V* in, V* out, const StTw<L, R, Ns, S, V>& tw, int lane) {
  static_assert(L / R <= 64, "one butterfly per lane");
  const int j = lane;
  if (j >= L / R) return;
  V v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if ((ZIN >> r) & 1u) {
      v[r].x = 0;
      v[r].y = 0;
    } else {
      v[r] = in[spad(j + r * (L / R))];
    }
  }
  if (Ns > 1) {
#pragma unroll
    for (int r = 1; r < R; ++r) {
      if ((ZIN >> r) & 1u) continue;
      v[r] = cmulv(v[r], tw.w[r - 1]);
    }
  }
  Dft<R, S, V>::run(v);
  const int k = j % Ns;
  const int base = (j / Ns) * Ns * R + k;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if ((WOUT >> r) & 1u) out[spad(base + r * Ns)] = v[r];
}

// Rows of a line that the padded streamwise transforms never need, for a
// first inverse (last forward) radix-R stage with Ns = 1 (= L/R): row r holds
// the points r L/R .. r L/R + L/R - 1.  With Mx = L/3 kept modes the band
// Mx .. L - Mx of the inverse input is zero (the 3/2 padding), and the same
// band of the forward output is discarded (truncation to Mx modes).
template <int L, int R>
struct PadRows {
  static constexpr int LR = L / R, MX = L / 3;
  static constexpr int lo = (MX + LR - 1) / LR;           // first row wholly inside the band
  static constexpr int hi = (L - MX + 1) / LR - 1;        // last such row
  static constexpr unsigned band = hi >= lo ? (((1u << (hi + 1)) - 1u) & ~((1u << lo) - 1u)) : 0u;
  static constexpr unsigned below = (1u << lo) - 1u;      // rows holding modes < Mx only partly or wholly
};

// Orders the LDS accesses of one wavefront: its lanes see each other's earlier
// writes after this point (a wavefront's LDS instructions execute in order;
// the fences keep the compiler from moving accesses across it).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS buffers per line of k_xpass_st.  A stage with at most one butterfly
// per lane can run in place: each lane reads all its inputs before it writes,
// and a wavefront's LDS instructions execute in order, so every read of the
// stage precedes every write.  One buffer per line halves the LDS a line
// holds (16-byte elements at L = 384: 0.749 -> 0.737 ms); the 8-byte form's
// occupancy is set by its registers, not its LDS, and keeps two.
template <int L, int RA, int RB, class V>
constexpr int xpass_buffers() {
  return (sizeof(V) == 16 && L / RA <= 64 && L / RB <= 64) ? 1 : 2;
}

// Streamwise pass, one 64-lane wavefront per (y point, span point) line, for
// L = RA RB RC with L / RC = 64 and Mx = L / 3 spectral points per row.  Each
// wave first reads its six input rows into registers (all loads in flight at
// once), then runs the inverse transforms through its own LDS buffers with
// wave-level ordering only, so the waves of a block never wait for each other.
// The inverse transforms run the stages RA, RB, RC; the last one leaves lane j
// holding the physical points j + 64 r (r < RC) of every field, so the six
// products form the Lamb vector in registers, and the forward transforms start
// with a radix-RC stage on exactly those points (Ns = 1, no twiddle) before RB
// and RA.  Fields are packed two per complex transform as in k_xpass_reg;
// block = NL lines.
template <int L, int RA, int RB, int RC, int NL, class V>
__global__ __launch_bounds__(64 * NL) void k_xpass_st(RegX<V> a) {
  static_assert(L == RA * RB * RC && L / RC == 64 && L % 3 == 0, "one wavefront per line");
  using T = decltype(V{}.x);
  constexpr int PL = L + (L >> 3) + 1;  // padded buffer length
  constexpr int MX = L / 3;             // spectral points per row
  constexpr int MP = (MX + 63) / 64;    // of them per lane
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  V* tw = sm;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int NB = xpass_buffers<L, RA, RB, V>();
  V* b0 = sm + L + wave * NB * PL;
  V* b1 = b0 + (NB - 1) * PL;
  const long line = (long)blockIdx.x * NL + wave;
  const bool valid = line < a.nlines;
  V in[6][MP];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
#pragma unroll
    for (int u = 0; u < MP; ++u) {
      const int m = lane + 64 * u;
      in[f][u] = (valid && m < MX) ? a.in[f][line * MX + m] : C2<T>::make(0, 0);
    }
  }
  for (int i = threadIdx.x; i < L; i += 64 * NL) tw[i] = a.tw[i];
  __syncthreads();  // the only block-wide barrier: twiddles
  if (!valid) return;
  // with one butterfly per lane in the LDS stages, this lane's twiddles of
  // every stage are read once for the line's five transforms
  // (only where the registers cost no occupancy: 16-byte elements at L = 384,
  // 0.800 -> 0.743 ms; with 8-byte elements the occupancy they cost is worth
  // more than the reads they save, 0.344 -> 0.367 ms, and at L = 192 they halve
  // the 16-byte kernel's occupancy)
  constexpr bool kRegTw = sizeof(V) == 16 && L == 384 && L / RA <= 64 && L / RB <= 64;
  StTw<L, RB, RA, +1, V> twi2;            // inverse stage 2
  StTw<L, RB, RC, -1, V> twf2;            // forward stage 2
  StTw<L, RA, RC * RB, -1, V> twf3;       // forward stage 3
  V twl[RC - 1];                          // last inverse stage (Ns = 64): conj W_L^(r lane)
  if constexpr (kRegTw) {
    twi2.load(tw, lane);
    twf2.load(tw, lane);
    twf3.load(tw, lane);
#pragma unroll
    for (int rr = 1; rr < RC; ++rr) {
      V w = tw[rr * lane];
      w.y = -w.y;
      twl[rr - 1] = w;
    }
  }
  V r[3][RC];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    // natural-order spectrum: F1 + i F2 at m, conj(F1) + i conj(F2) at L - m,
    // zero on the padding band Mx <= m <= L - Mx; Re of the kx = 0 entry (PF:1116-1118)
#pragma unroll
    for (int u = 0; u < MP; ++u) {
      const int m = lane + 64 * u;
      if (m < MX) {
        const V A = in[2 * q][u], B = in[2 * q + 1][u];
        if (m == 0) {
          b0[spad(0)] = C2<T>::make(A.x, B.x);
        } else {
          b0[spad(m)] = C2<T>::make(A.x - B.y, A.y + B.x);
          b0[spad(L - m)] = C2<T>::make(A.x + B.y, B.x - A.y);
        }
      }
    }
    // the band rows PadRows::band are never read; the band points outside them are
    using PR = PadRows<L, RA>;
    for (int m = MX + lane; m < PR::lo * PR::LR; m += 64) b0[spad(m)] = C2<T>::make(0, 0);
    for (int m = (PR::hi + 1) * PR::LR + lane; m <= L - MX; m += 64) b0[spad(m)] = C2<T>::make(0, 0);
    wave_sync();
    st_stage<L, RA, 1, +1, PR::band>(b0, b1, tw, lane);
    wave_sync();
    if constexpr (kRegTw)
      st_stage_r<L, RB, RA, +1>(b1, b0, twi2, lane);
    else
      st_stage<L, RB, RA, +1>(b1, b0, tw, lane);
    wave_sync();
    {
      // last inverse stage (Ns = 64): lane j reads j + 64 r, keeps the output in registers
#pragma unroll
      for (int rr = 0; rr < RC; ++rr) r[q][rr] = b0[spad(lane + 64 * rr)];
#pragma unroll
      for (int rr = 1; rr < RC; ++rr) {
        V w;
        if constexpr (kRegTw) {
          w = twl[rr - 1];
        } else {
          w = tw[rr * lane];
          w.y = -w.y;
        }
        r[q][rr] = cmulv(r[q][rr], w);
      }
      Dft<RC, +1, V>::run(r[q]);
    }
    wave_sync();
  }
  // Lamb vector at x = lane + 64 r: (L_x + i L_y) into r[0], (L_z + 0i) into r[1]
  const T rep = (T)a.re_p;
#pragma unroll
  for (int i = 0; i < RC; ++i) {
    const T uu = r[0][i].x, vv = r[0][i].y, ww = r[1][i].x, oxx = r[1][i].y;
    const T oyy = r[2][i].x, ozz = r[2][i].y;
    const T lx = (oyy * ww - rep) - ozz * vv;
    const T ly = ozz * uu - oxx * ww;
    const T lz = oxx * vv - oyy * uu;
    r[0][i] = C2<T>::make(lx, ly);
    r[1][i] = C2<T>::make(lz, 0);
  }
  const T sc = (T)a.scale;
  const T half = (T)0.5;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    // first forward stage (R = RC, Ns = 1) on the registers
    Dft<RC, -1, V>::run(r[q]);
#pragma unroll
    for (int rr = 0; rr < RC; ++rr) b0[spad(lane * RC + rr)] = r[q][rr];
    wave_sync();
    if constexpr (kRegTw)
      st_stage_r<L, RB, RC, -1>(b0, b1, twf2, lane);
    else
      st_stage<L, RB, RC, -1>(b0, b1, tw, lane);
    wave_sync();
    // last stage: the band rows are discarded; the L_z transform (q = 1)
    // keeps modes below Mx only
    using PR = PadRows<L, RA>;
    if constexpr (kRegTw) {
      if (q == 0)
        st_stage_r<L, RA, RC * RB, -1, 0u, ~PR::band>(b1, b0, twf3, lane);
      else
        st_stage_r<L, RA, RC * RB, -1, 0u, PR::below>(b1, b0, twf3, lane);
    } else {
      if (q == 0)
        st_stage<L, RA, RC * RB, -1, 0u, ~PR::band>(b1, b0, tw, lane);
      else
        st_stage<L, RA, RC * RB, -1, 0u, PR::below>(b1, b0, tw, lane);
    }
    wave_sync();
    if (q == 0) {
      // separate the packed pair: F1 = (G(m) + conj G(-m))/2, F2 = (G(m) - conj G(-m))/(2i)
#pragma unroll
      for (int u = 0; u < MP; ++u) {
        const int m = lane + 64 * u;
        if (m < MX) {
          const V Gp = b0[spad(m)];
          const V Gc = b0[spad(m == 0 ? 0 : L - m)];
          a.out[0][line * MX + m] = C2<T>::make(half * (Gp.x + Gc.x) * sc, half * (Gp.y - Gc.y) * sc);
          a.out[1][line * MX + m] = C2<T>::make(half * (Gp.y + Gc.y) * sc, -half * (Gp.x - Gc.x) * sc);
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < MP; ++u) {
        const int m = lane + 64 * u;
        if (m < MX) {
          const V g = b0[spad(m)];
          a.out[2][line * MX + m] = C2<T>::make(g.x * sc, g.y * sc);
        }
      }
    }
    wave_sync();
  }
}

template <int L, int RA, int RB, int RC, class V>
static hipError_t run_xpass_st(Ctx& c, const RegX<V>& a) {
  if (a.Mx * 3 != L) return hipErrorInvalidValue;  // the kernel's row length is L / 3
  constexpr int NL = sizeof(V) == 16 ? 2 : 4;  // lines per block
  constexpr int PL = L + (L >> 3) + 1;
  constexpr int NB = xpass_buffers<L, RA, RB, V>();
  const long blocks = (a.nlines + NL - 1) / NL;
  const size_t shm = sizeof(V) * (L + (size_t)NL * NB * PL);
  static_assert(sizeof(V) * (L + (size_t)NL * NB * PL) <= 65536, "line buffers exceed 64 KB of LDS");
  hipLaunchKernelGGL((k_xpass_st<L, RA, RB, RC, NL, V>), dim3((unsigned)blocks), dim3(64 * NL), shm,
                     c.stream, a);
  return hipGetLastError();
}

// Span passes: a group handles TX span lines (consecutive kx) of one plane;
// thread (t, tx) = position t of line tx, tx fastest for coalesced access.
template <int N1, int N2, int TX, bool INV, class V>
__global__ __launch_bounds__((TX * Geo<N1, N2>::G)) void k_zfft_reg(RegZ<V> a) {
  using Gm = Geo<N1, N2>;
  using T = decltype(V{}.x);
  constexpr int L = Gm::L, G = Gm::G;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  V* tw = sm;
  const int t = threadIdx.x / TX, tx = threadIdx.x - t * TX;
  V* buf = sm + L + tx * Gm::LINE_Z;
  const int Mz = a.Mz, h = Mz / 2;
  // kx groups in XCD-contiguous order: the groups that share 128-byte lines of
  // narrow (FP32) rows go through the same L2
  const int kg = (int)xcd_order(blockIdx.x, gridDim.x);
  const int kx = kg * TX + tx;
  const int p = blockIdx.y;
  const V* in = a.in[blockIdx.z] + (long)p * a.in_plane;
  V* out = a.out[blockIdx.z] + (long)p * a.out_plane;
  for (int i = threadIdx.x; i < L; i += TX * G) tw[i] = a.tw[i];
  __syncthreads();
  if (t < N1) {
    V x[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
      const int m = n2 * N1 + t;
      if (INV) {
        int kz = -1;
        if (m < h) kz = m;                       // Nyquist slot Mz/2 stays zero (expand, PF:730)
        else if (m > L - h) kz = m - L + Mz;
        x[n2] = kz >= 0 ? in[(long)kz * a.Mx + kx] : C2<T>::make(0, 0);
      } else {
        x[n2] = in[(long)m * a.Mx + kx];
      }
    }
    fft_stepA<N1, N2, INV ? +1 : -1>(x, buf, tw, t);
  }
  __syncthreads();
  if (t < N2) {
    V b[N1];
    fft_stepB<N1, N2, INV ? +1 : -1>(b, buf, t);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      const int m = k1 * N2 + t;
      if (INV) {
        out[(long)m * a.Mx + kx] = b[k1];
      } else {
        int kz = -1;                              // expand(-1): keep 0..Mz/2 and the negative band
        if (m <= h) kz = m;
        else if (m > L - h) kz = m - L + Mz;
        if (kz >= 0) out[(long)kz * a.Mx + kx] = b[k1];
      }
    }
  }
}

// FP32 span pass over column pairs: thread (t, tx) transforms the lines of
// kx = 2 (kg TX + tx) and kx + 1 together, so every global access is one
// 16-byte (kx, kx + 1) pair and a row of the block's TX pairs is a whole
// 128-byte line at TX = 8.  The twiddles are loaded once for both lines, and
// the LDS exchange moves the pair as one 16-byte value.  Each line sees the
// same operations as in k_zfft_reg.
template <int N1, int N2, int S>
__device__ __forceinline__ void fft_stepA_pair(float2 (&a)[N2], float2 (&b)[N2], float4* buf, const float2* tw,
                                               int t) {
  Dft<N2, S, float2>::run(a);
  Dft<N2, S, float2>::run(b);
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) {
    float2 w = tw[t * k2];
    if (S > 0) w.y = -w.y;
    const float2 va = (k2 == 0) ? a[0] : cmulv(a[k2], w);
    const float2 vb = (k2 == 0) ? b[0] : cmulv(b[k2], w);
    buf[t * (N2 + 1) + k2] = make_float4(va.x, va.y, vb.x, vb.y);
  }
}

template <int N1, int N2, int S>
__device__ __forceinline__ void fft_stepB_pair(float2 (&a)[N1], float2 (&b)[N1], const float4* buf, int t) {
#pragma unroll
  for (int n1 = 0; n1 < N1; ++n1) {
    const float4 q = buf[n1 * (N2 + 1) + t];
    a[n1] = make_float2(q.x, q.y);
    b[n1] = make_float2(q.z, q.w);
  }
  Dft<N1, S, float2>::run(a);
  Dft<N1, S, float2>::run(b);
}

template <int N1, int N2, int TX, bool INV>
__global__ __launch_bounds__((TX * Geo<N1, N2>::G)) void k_zfft_pair(RegZ<float2> a) {
  using Gm = Geo<N1, N2>;
  constexpr int L = Gm::L;
  extern __shared__ __align__(16) unsigned char smraw[];
  float2* tw = reinterpret_cast<float2*>(smraw);
  const int t = threadIdx.x / TX, tx = threadIdx.x - t * TX;
  float4* buf = reinterpret_cast<float4*>(smraw + sizeof(float2) * L) + tx * Gm::BUF_A;
  const int Mz = a.Mz, h = Mz / 2;
  const int kg = (int)xcd_order(blockIdx.x, gridDim.x);
  const int kp = kg * TX + tx;  // column pair: kx = 2 kp, 2 kp + 1
  const int Mp = a.Mx / 2;      // pairs per row
  const int p = blockIdx.y;
  const float4* in = reinterpret_cast<const float4*>(a.in[blockIdx.z] + (long)p * a.in_plane);
  float4* out = reinterpret_cast<float4*>(a.out[blockIdx.z] + (long)p * a.out_plane);
  for (int i = threadIdx.x; i < L; i += TX * Gm::G) tw[i] = a.tw[i];
  __syncthreads();
  if (t < N1) {
    float2 x0[N2], x1[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
      const int m = n2 * N1 + t;
      float4 q = make_float4(0, 0, 0, 0);
      if (INV) {
        int kz = -1;
        if (m < h) kz = m;                       // Nyquist slot Mz/2 stays zero (expand, PF:730)
        else if (m > L - h) kz = m - L + Mz;
        if (kz >= 0) q = in[(long)kz * Mp + kp];
      } else {
        q = in[(long)m * Mp + kp];
      }
      x0[n2] = make_float2(q.x, q.y);
      x1[n2] = make_float2(q.z, q.w);
    }
    fft_stepA_pair<N1, N2, INV ? +1 : -1>(x0, x1, buf, tw, t);
  }
  __syncthreads();
  if (t < N2) {
    float2 b0[N1], b1[N1];
    fft_stepB_pair<N1, N2, INV ? +1 : -1>(b0, b1, buf, t);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      const int m = k1 * N2 + t;
      const float4 q = make_float4(b0[k1].x, b0[k1].y, b1[k1].x, b1[k1].y);
      if (INV) {
        out[(long)m * Mp + kp] = q;
      } else {
        int kz = -1;                              // expand(-1): keep 0..Mz/2 and the negative band
        if (m <= h) kz = m;
        else if (m > L - h) kz = m - L + Mz;
        if (kz >= 0) out[(long)kz * Mp + kp] = q;
      }
    }
  }
}

bool reg_fft_length(int L) {
  switch (L) {
    case 12: case 24: case 48: case 96: case 144: case 192: case 384: return true;
    default: return false;
  }
}

// Lengths of the register span FFTs (both precisions) and of the FP32
// streamwise pass: the lengths above plus 768 (the 512-mode channel).  The
// FP64 streamwise lines of 768 points would not fit the register file and
// take the LDS kernel.
bool reg_fft_length32(int L) { return reg_fft_length(L) || L == 768; }

template <int N1, int N2, class V>
static hipError_t run_xpass_reg(Ctx& c, const RegX<V>& a) {
  using Gm = Geo<N1, N2>;
  constexpr int NG = 256 / Gm::G;
  const long blocks = (a.nlines + NG - 1) / NG;
  const size_t shm = sizeof(V) * (Gm::L + (size_t)NG * Gm::BUF);
  if (shm > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)k_xpass_reg<N1, N2, NG, V>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_xpass_reg<N1, N2, NG, V>), dim3((unsigned)blocks), dim3(NG * Gm::G), shm, c.stream, a);
  return hipGetLastError();
}

template <class V>
static hipError_t xpass_reg(Ctx& c, const RegX<V>& a) {
  switch (c.d.Lx) {
    case 12: return run_xpass_reg<4, 3>(c, a);
    case 24: return run_xpass_reg<4, 6>(c, a);
    case 48: return run_xpass_reg<6, 8>(c, a);
    case 96: return run_xpass_reg<8, 12>(c, a);
    case 144: return run_xpass_reg<12, 12>(c, a);
    case 192: return run_xpass_st<192, 8, 8, 3>(c, a);
    case 384: return run_xpass_st<384, 8, 8, 6>(c, a);
    case 768:  // FP32 only (reg_fft_length32)
      if constexpr (sizeof(V) == 8) return run_xpass_st<768, 8, 8, 12>(c, a);
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

// Column-pair launch (FP32): TX pairs per block, the largest of 8, 4, 2 whose
// exchange tiles fit in 64 KB of LDS and whose 2 TX columns divide Mx.
// Returns hipErrorNotSupported when no tile fits (the caller then takes the
// single-column kernel).
template <int N1, int N2, int TX>
static hipError_t launch_zfft_pair(Ctx& c, const RegZ<float2>& a, bool inv, int nf) {
  using Gm = Geo<N1, N2>;
  const size_t shm = sizeof(float2) * Gm::L + sizeof(float4) * (size_t)TX * Gm::BUF_A;
  if (shm > 65536 || c.d.Mx % (2 * TX)) return hipErrorNotSupported;
  for (int f = 0; f < nf; ++f)
    if (((uintptr_t)a.in[f] | (uintptr_t)a.out[f]) & 15) return hipErrorNotSupported;
  dim3 grid(c.d.Mx / (2 * TX), c.d.Py, nf);
  if (inv)
    hipLaunchKernelGGL((k_zfft_pair<N1, N2, TX, true>), grid, dim3(TX * Gm::G), shm, c.stream, a);
  else
    hipLaunchKernelGGL((k_zfft_pair<N1, N2, TX, false>), grid, dim3(TX * Gm::G), shm, c.stream, a);
  return hipGetLastError();
}

template <int N1, int N2>
static hipError_t run_zfft_pair(Ctx& c, const RegZ<float2>& a, bool inv, int nf) {
  using Gm = Geo<N1, N2>;
  constexpr int TX = (256 / Gm::G) >= 8 ? 8 : ((256 / Gm::G) >= 4 ? 4 : 2);
  hipError_t e = launch_zfft_pair<N1, N2, TX>(c, a, inv, nf);
  if constexpr (TX > 2)
    if (e == hipErrorNotSupported) e = launch_zfft_pair<N1, N2, TX / 2>(c, a, inv, nf);
  if constexpr (TX > 4)
    if (e == hipErrorNotSupported) e = launch_zfft_pair<N1, N2, TX / 4>(c, a, inv, nf);
  return e;
}

template <int N1, int N2, class V>
static hipError_t run_zfft_reg(Ctx& c, const RegZ<V>& a, bool inv, int nf) {
  using Gm = Geo<N1, N2>;
  if constexpr (sizeof(V) == 8) {
    const hipError_t e = run_zfft_pair<N1, N2>(c, a, inv, nf);
    if (e != hipErrorNotSupported) return e;
  }
  constexpr int TX = (256 / Gm::G) >= 16 ? 16 : ((256 / Gm::G) >= 8 ? 8 : 4);
  if (c.d.Mx % TX) return hipErrorInvalidValue;
  const size_t shm = sizeof(V) * (Gm::L + (size_t)TX * Gm::LINE_Z);
  dim3 grid(c.d.Mx / TX, c.d.Py, nf);
  const void* fn = inv ? (const void*)k_zfft_reg<N1, N2, TX, true, V> : (const void*)k_zfft_reg<N1, N2, TX, false, V>;
  if (shm > 65536) {
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
  }
  if (inv)
    hipLaunchKernelGGL((k_zfft_reg<N1, N2, TX, true, V>), grid, dim3(TX * Gm::G), shm, c.stream, a);
  else
    hipLaunchKernelGGL((k_zfft_reg<N1, N2, TX, false, V>), grid, dim3(TX * Gm::G), shm, c.stream, a);
  return hipGetLastError();
}

// Span pass through the register kernel; hipErrorInvalidValue when the length
// has no split or Mx is not a multiple of the kernel's line tile (16).
template <class V>
static hipError_t zfft_reg(Ctx& c, const RegZ<V>& a, bool inv, int nf) {
  if (c.d.Mx % 16) return hipErrorInvalidValue;
  switch (c.d.Lz) {
    case 12: return run_zfft_reg<4, 3>(c, a, inv, nf);
    case 24: return run_zfft_reg<4, 6>(c, a, inv, nf);
    case 48: return run_zfft_reg<6, 8>(c, a, inv, nf);
    case 96: return run_zfft_reg<8, 12>(c, a, inv, nf);
    case 144: return run_zfft_reg<12, 12>(c, a, inv, nf);
    case 192: return run_zfft_reg<12, 16>(c, a, inv, nf);
    case 384: return run_zfft_reg<16, 24>(c, a, inv, nf);
    case 768: return run_zfft_reg<24, 32>(c, a, inv, nf);
    default: return hipErrorInvalidValue;
  }
}

template <class V>
static RegZ<V> regz_args(Ctx& c, const V* const* in, V* const* out, int nf, bool inv, const V* tw) {
  RegZ<V> a{};
  for (int f = 0; f < nf; ++f) {
    a.in[f] = in[f];
    a.out[f] = out[f];
  }
  a.Mx = c.d.Mx;
  a.Mz = c.d.Mz;
  a.in_plane = (long)(inv ? c.d.Mz : c.d.Lz) * c.d.Mx;
  a.out_plane = (long)(inv ? c.d.Lz : c.d.Mz) * c.d.Mx;
  a.tw = tw;
  return a;
}

static bool use_zreg(const Ctx& c) { return reg_fft_length32(c.d.Lz) && c.d.Mx % 16 == 0; }

static hipError_t zfft_reg64(Ctx& c, double2* const* in, double2* const* out, int nf, bool inv) {
  const double2* cin[6];
  for (int f = 0; f < nf; ++f) cin[f] = in[f];
  return zfft_reg(c, regz_args<double2>(c, cin, out, nf, inv, c.twz), inv, nf);
}

hipError_t launch_zfft32(Ctx& c, const float2* const* in, float2* const* out, int nf, bool inv) {
  return zfft_reg(c, regz_args<float2>(c, in, out, nf, inv, c.twzf), inv, nf);
}

hipError_t launch_xpass32(Ctx& c) {
  RegX<float2> a{};
  for (int f = 0; f < 6; ++f) a.in[f] = c.Zf[f];
  for (int f = 0; f < 3; ++f) a.out[f] = c.Zf[f];
  a.Mx = c.d.Mx;
  a.nlines = (long)c.d.Py * c.d.Lz;
  a.re_p = c.prm.re_p;
  a.scale = 1.0 / ((double)c.d.Lx * (double)c.d.Lz);
  a.tw = c.twxf;
  return xpass_reg(c, a);
}

hipError_t launch_xpass(Ctx& c) {
  const Dims& d = c.d;
  XArgs a{};
  for (int f = 0; f < 6; ++f) a.in[f] = c.Z[f];
  // outputs overwrite Z[0..2] line by line: each block reads its own lines fully
  // before writing them, and no block touches another block's lines
  for (int f = 0; f < 3; ++f) a.out[f] = c.Z[f];
  a.Mx = d.Mx;
  a.Lx = d.Lx;
  int lpb = 2048 / (3 * d.Lx);
  if (lpb < 1) lpb = 1;
  a.LPB = lpb;
  a.nlines = (long)d.Py * d.Lz;
  a.re_p = c.prm.re_p;
  a.scale = 1.0 / ((double)d.Lx * (double)d.Lz);
  a.pl = make_plan(d.Lx);
  a.tw = c.twx;
  if (reg_fft_length(d.Lx)) {
    RegX<double2> r{};
    for (int f = 0; f < 6; ++f) r.in[f] = c.Z[f];
    for (int f = 0; f < 3; ++f) r.out[f] = c.Z[f];
    r.Mx = d.Mx;
    r.nlines = a.nlines;
    r.re_p = a.re_p;
    r.scale = a.scale;
    r.tw = c.twx;
    return xpass_reg(c, r);
  }
  dim3 grid((unsigned)((a.nlines + lpb - 1) / lpb));
  size_t shm = 2ull * 3 * lpb * d.Lx * sizeof(double2);
  if (shm > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)k_xpass, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_xpass, grid, dim3(256), shm, c.stream, a);
  return hipGetLastError();
}

}  // namespace chf
