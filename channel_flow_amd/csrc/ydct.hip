// Wall-normal Chebyshev transforms of the FP64 form through FFTs along y.
//
// Gauss-Lobatto point values (the synthesis of the h-terms' point values and
// of the statistics profiles, getall's T0 products PF:1573-1609):
//
//   f_j = sum_{k=0}^{N} a_k cos(pi k j / N),   j = 0..N,
//
// a type-I DCT.  With the even extension c_m = a_m (m <= N), c_m = a_{2N-m}
// (N < m < 2N), the 2N-point DFT S_j = sum_m c_m e^{-2 pi i m j / 2N} equals
// a_0 + (-1)^j a_N + 2 sum_{k=1}^{N-1} a_k cos(pi k j / N), so
//
//   f_j = (S_j + a_0 + (-1)^j a_N) / 2.
//
// Padded synthesis (PF:1657-1706): coefficients to the 2N+1 points
// y_p = cos(pi p / 2N), f_p = sum_{k=0}^{N} a_k cos(pi k p / 2N).  The even
// points p = 2j are the type-I DCT above; the odd points p = 2l+1 are
// X_l = sum_{k<N} a_k cos(pi k (2l+1) / 2N), a type-III DCT, computed by
// Makhoul's N-point FFT: V_0 = a_0, V_k = w_k (a_k - i a_{N-k}) / 2 with
// w_k = e^{i pi k / 2N}, v_n = sum_k V_k e^{2 pi i k n / N}, and then
// X_{2n} = v_n, X_{2n+1} = v_{N-1-n} (n < N/2).
//
// Analysis (PF:1686-1700): points f_p (p <= 2N) to
// a_k = (1/2N)(k == 0 ? 1/2 : 1) sum_{p=0}^{N} w_p g_p cos(pi k p / 2N) with
// w_0 = 1, w_p = 2, g_p = f_p +- f_{2N-p} (+ for k even, - for k odd) and
// g_N = f_N (k even) or 0 (k odd).  Even k = 2m: the type-I DCT of
// G_0 = f_0 + f_2N, G_p = f_p + f_{2N-p}, G_N = 2 f_N, at m.  Odd k = 2m+1:
// the type-III DCT of H_0 = f_0 - f_2N, H_p = 2 (f_p - f_{2N-p}), at m.
//
// Every transform is real and linear, and the FFT forms above are
// complex-linear maps that agree with it on real data, so they apply to the
// complex Fourier coefficients as they are.  O(N log N) work per column
// instead of the cos-table products' O(N^2) on the FP64 matrix cores.
//
// Layout: a block owns 8 adjacent complex columns (a row of 8 16-byte values
// is one 128-byte line), and each column's line is transformed by a group of
// threads with the two-level register FFT of fftreg.h, the twiddles read from
// the (cached) global tables.
#include "chf_internal.h"
#include "fftreg.h"

namespace chf {

constexpr int kYTX = 8;  // columns per block

// LDS stride (16-byte values) between the exchange tiles of a block's
// columns.  The lanes of one LDS pass are adjacent columns (the column index
// is fastest), so the stride sets the bank pattern: a 16-byte write pass is 8
// lanes over 8 bank slots, a read pass 16 lanes over 16.  A stride of 9 mod
// 16 puts a write pass's lanes in distinct slots and leaves reads one
// collision; the unpadded tile (a multiple of 16 values) put all lanes of a
// pass in one slot (measured: 0.530 -> 0.487 ms for the padded synthesis at
// 256x193x256).
constexpr int tile_stride(int buf) { return buf + (9 - buf % 16 + 16) % 16; }

constexpr int ymax(int a, int b) { return a > b ? a : b; }

struct YArgs {
  const double2* in[6];
  double2* out[6];
  long ncol;
  const double2* twL;  // e^{-2 pi i m / 2N}, m < 2N
  const double2* twN;  // e^{-2 pi i m / N},  m < N
  const double2* wq;   // e^{i pi k / 2N},    k < N
  double inv2n;        // 1 / 2N
};

__device__ __forceinline__ double2 yadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 ysub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 yscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// Makhoul's pre-twiddle w (a - i b) / 2
__device__ __forceinline__ double2 ypre(double2 a, double2 b, double2 w) {
  return yscale(cmulv(make_double2(a.x + b.y, a.y - b.x), w), 0.5);
}

// Gauss-Lobatto point values: the 2N-point FFT (N1 x N2) of the even extension.
template <int N1, int N2>
__global__ __launch_bounds__(kYTX* ymax(N1, N2)) void k_ypoints(YArgs a) {
  constexpr int L = N1 * N2, N = L / 2, SB = tile_stride(N1 * (N2 + 1)), TX = kYTX;
  extern __shared__ __align__(16) unsigned char smraw[];
  const int t = threadIdx.x / TX, tx = threadIdx.x - t * TX;
  double2* buf = reinterpret_cast<double2*>(smraw) + tx * SB;
  const long col = xcd_order(blockIdx.x, gridDim.x) * TX + tx;
  const long nc = a.ncol;
  const double2* in = a.in[blockIdx.z];
  double2* out = a.out[blockIdx.z];
  if (t < N1) {
    double2 x[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
      const int m = n2 * N1 + t;
      x[n2] = in[(long)(m <= N ? m : L - m) * nc + col];
    }
    fft_stepA<N1, N2, -1>(x, buf, a.twL, t);
  }
  __syncthreads();
  if (t < N2) {
    double2 b[N1];
    fft_stepB<N1, N2, -1>(b, buf, t);
    const double2 a0 = in[col], aN = in[(long)N * nc + col];
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      const int j = k1 * N2 + t;
      if (j <= N) out[(long)j * nc + col] = yscale(yadd(yadd(b[k1], a0), (j & 1) ? yscale(aN, -1.0) : aN), 0.5);
    }
  }
}

// Padded synthesis (ANA = false) or analysis (ANA = true): the 2N-point FFT
// (A1 x A2), then the N-point FFT (A1 x A2/2) through the same LDS tile.  Each
// FFT reads its inputs when it starts (a variant that formed the second FFT's
// inputs from the first one's loads, so that a block waits for memory once,
// needed 222-253 VGPRs instead of 156-172 and ran slower).
template <int A1, int A2, bool ANA>
__global__ __launch_bounds__(kYTX* ymax(A1, A2)) void k_ycheb(YArgs a) {
  constexpr int L = A1 * A2, N = L / 2, B2 = A2 / 2, TX = kYTX;
  static_assert(A2 % 2 == 0, "the second FFT is A1 x A2/2");
  constexpr int SB = tile_stride(A1 * (A2 + 1));
  extern __shared__ __align__(16) unsigned char smraw[];
  const int t = threadIdx.x / TX, tx = threadIdx.x - t * TX;
  double2* buf = reinterpret_cast<double2*>(smraw) + tx * SB;
  const long col = xcd_order(blockIdx.x, gridDim.x) * TX + tx;
  const long nc = a.ncol;
  const double2* in = a.in[blockIdx.z];
  double2* out = a.out[blockIdx.z];
  auto ld = [&](int r) { return in[(long)r * nc + col]; };
  auto st = [&](int r, double2 v) { out[(long)r * nc + col] = v; };
  auto hodd = [&](int q) { return q == 0 ? ysub(ld(0), ld(2 * N)) : yscale(ysub(ld(q), ld(2 * N - q)), 2.0); };

  // 2N-point FFT of the even extension: even points / even coefficients
  if (t < A1) {
    double2 x[A2];
#pragma unroll
    for (int n2 = 0; n2 < A2; ++n2) {
      const int m = n2 * A1 + t;
      const int p = m <= N ? m : L - m;
      x[n2] = ANA ? yadd(ld(p), ld(2 * N - p)) : ld(p);  // analysis, p = N: G_N = 2 f_N
    }
    fft_stepA<A1, A2, -1>(x, buf, a.twL, t);
  }
  __syncthreads();
  if (t < A2) {
    double2 b[A1];
    fft_stepB<A1, A2, -1>(b, buf, t);
    if constexpr (!ANA) {
      const double2 a0 = ld(0), aN = ld(N);
#pragma unroll
      for (int k1 = 0; k1 < A1; ++k1) {
        const int j = k1 * A2 + t;
        if (j <= N) st(2 * j, yscale(yadd(yadd(b[k1], a0), (j & 1) ? yscale(aN, -1.0) : aN), 0.5));
      }
    } else {
#pragma unroll
      for (int k1 = 0; k1 < A1; ++k1) {
        const int m = k1 * A2 + t;
        if (2 * m <= N) st(2 * m, yscale(b[k1], m == 0 ? 0.5 * a.inv2n : a.inv2n));
      }
    }
  }
  __syncthreads();  // the tile is reused
  // N-point FFT of the type-III DCT: odd points / odd coefficients
  if (t < A1) {
    double2 y[B2];
#pragma unroll
    for (int n2 = 0; n2 < B2; ++n2) {
      const int k = n2 * A1 + t;
      if constexpr (!ANA) {
        y[n2] = (k == 0) ? ld(0) : ypre(ld(k), ld(N - k), a.wq[k]);
      } else {
        y[n2] = (k == 0) ? hodd(0) : ypre(hodd(k), hodd(N - k), a.wq[k]);
      }
    }
    fft_stepA<A1, B2, +1>(y, buf, a.twN, t);
  }
  __syncthreads();
  if (t < B2) {
    double2 v[A1];
    fft_stepB<A1, B2, +1>(v, buf, t);
#pragma unroll
    for (int k1 = 0; k1 < A1; ++k1) {
      const int n = k1 * B2 + t;
      const int l = n < N / 2 ? 2 * n : 2 * (N - 1 - n) + 1;  // v_n = X_l
      if constexpr (!ANA) {
        st(2 * l + 1, v[k1]);
      } else {
        if (2 * l + 1 <= N) st(2 * l + 1, yscale(v[k1], a.inv2n));
      }
    }
  }
}

template <class K>
static hipError_t launch_y(Ctx& c, K kernel, int nthreads, int tile, const YArgs& a, int nf) {
  if (c.d.ncol % kYTX) return hipErrorNotSupported;
  const size_t shm = sizeof(double2) * (size_t)kYTX * tile_stride(tile);
  if (shm > 65536) {
    const hipError_t e = hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kernel, dim3((unsigned)(c.d.ncol / kYTX), 1, nf), dim3(nthreads), shm, c.stream, a);
  return hipGetLastError();
}

// mode 0: point values, 1: padded synthesis, 2: analysis; A1 x A2 = 2N
template <int A1, int A2>
static hipError_t run_y(Ctx& c, int mode, const YArgs& a, int nf) {
  constexpr int nt = kYTX * ymax(A1, A2), tile = A1 * (A2 + 1);
  if (mode == 0) return launch_y(c, k_ypoints<A1, A2>, nt, tile, a, nf);
  if (mode == 1) return launch_y(c, k_ycheb<A1, A2, false>, nt, tile, a, nf);
  return launch_y(c, k_ycheb<A1, A2, true>, nt, tile, a, nf);
}

// 1 when 2N has a split below (and so an FFT form of the transforms).
int ydct_length_supported(int N) {
  switch (2 * N) {
    case 64: case 96: case 128: case 192: case 256: case 384: case 512: return 1;
    default: return 0;
  }
}

static hipError_t run_ydct(Ctx& c, int mode, const double2* const* in, double2* const* out, int nf) {
  YArgs a{};
  for (int f = 0; f < nf; ++f) {
    a.in[f] = in[f];
    a.out[f] = out[f];
  }
  const int N = c.d.N;
  a.ncol = c.d.ncol;
  a.twL = c.twy;
  a.twN = c.twy + 2 * N;
  a.wq = c.twy + 3 * N;
  a.inv2n = 1.0 / (2.0 * N);
  switch (2 * N) {
    case 64: return run_y<8, 8>(c, mode, a, nf);
    case 96: return run_y<8, 12>(c, mode, a, nf);
    case 128: return run_y<8, 16>(c, mode, a, nf);
    case 192: return run_y<12, 16>(c, mode, a, nf);
    case 256: return run_y<16, 16>(c, mode, a, nf);
    case 384: return run_y<16, 24>(c, mode, a, nf);
    case 512: return run_y<16, 32>(c, mode, a, nf);
    default: return hipErrorNotSupported;
  }
}

hipError_t launch_ypoints(Ctx& c, const double2* const* in, double2* const* out, int nf) {
  return run_ydct(c, 0, in, out, nf);
}
hipError_t launch_ysynth(Ctx& c, const double2* const* in, double2* const* out, int nf) {
  return run_ydct(c, 1, in, out, nf);
}
hipError_t launch_yanalysis(Ctx& c, const double2* const* in, double2* const* out, int nf) {
  return run_ydct(c, 2, in, out, nf);
}

}  // namespace chf
