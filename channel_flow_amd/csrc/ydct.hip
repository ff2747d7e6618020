// Wall-normal Chebyshev transforms through FFTs along y.
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
// instead of the cos-table products' O(N^2).
//
// The padded synthesis and the analysis run in FP64 arithmetic in the FP64
// form and in FP32 in the mixed form (FP64 coefficients in and out, FP32
// padded point values); the Gauss-Lobatto point values and the right-hand
// sides are FP64 arithmetic in both forms.
//
// Layout: a block owns TX adjacent complex columns (a row of a block's
// columns is one 128-byte line of the arithmetic type), and each column's line
// is transformed by a group of threads with the two-level register FFT of
// fftreg.h.  The twiddles are generated per thread from two table entries
// (FP64, also for the FP32 arithmetic): each table read is a vector-memory
// instruction, and with one read per FFT element the address unit, not the
// memory, bounded these kernels.
#include "chf_internal.h"
#include "fftreg.h"

namespace chf {

template <class V>
struct YT;
template <>
struct YT<double2> {
  using T = double;
  static constexpr int TX = 8;
};
template <>
struct YT<float2> {
  using T = float;
  static constexpr int TX = 16;
};

// LDS stride (values) between the exchange tiles of a block's columns.  The
// lanes of one LDS pass are adjacent columns (the column index is fastest), so
// the stride sets the bank pattern: 8-byte values, a write pass is 16 lanes
// over 16 8-byte bank slots and a read pass 32 lanes over 32; 16-byte values,
// 8 and 16 lanes.  A stride of 17 mod 32 (8-byte) or 9 mod 16 (16-byte) values
// puts a write pass's lanes in distinct slots and leaves reads one collision;
// the unpadded tile (a multiple of 16 values) put all lanes of a pass in one
// slot (measured: 0.530 -> 0.487 ms for the FP64 padded synthesis at
// 256x193x256).
template <class V>
constexpr int tile_stride(int buf) {
  return sizeof(V) == 8 ? buf + (17 - buf % 32 + 32) % 32 : buf + (9 - buf % 16 + 16) % 16;
}

constexpr int ymax(int a, int b) { return a > b ? a : b; }

struct YArgs {
  const void* in[6];
  void* out[6];
  long ncol;
  const double2* twL;  // e^{-2 pi i m / 2N}, m < 2N
  const double2* twN;  // e^{-2 pi i m / N},  m < N
  const double2* wq;   // e^{i pi k / 2N},    k < N
  double inv2n;        // 1 / 2N
};

template <class D, class S>
__device__ __forceinline__ D ycvt(const S& s) {
  D d;
  d.x = (decltype(d.x))s.x;
  d.y = (decltype(d.y))s.y;
  return d;
}
template <class V>
__device__ __forceinline__ V yadd(V a, V b) {
  a.x += b.x;
  a.y += b.y;
  return a;
}
template <class V>
__device__ __forceinline__ V ysub(V a, V b) {
  a.x -= b.x;
  a.y -= b.y;
  return a;
}
template <class V, class T>
__device__ __forceinline__ V yscale(V a, T s) {
  a.x *= s;
  a.y *= s;
  return a;
}
// Makhoul's pre-twiddle w (a - i b) / 2
template <class V>
__device__ __forceinline__ V ypre(V a, V b, double2 w) {
  using T = typename YT<V>::T;
  V d;
  d.x = a.x + b.y;
  d.y = a.y - b.x;
  return yscale(cmulv(d, ycvt<V>(w)), (T)0.5);
}

// Step A of fftreg.h's two-level FFT with generated twiddles: thread t needs
// W^(t k) for k < N2 (W = e^{-2 pi i / (N1 N2)}); from w = W^t and w^4 = W^(4t)
// (two table reads, 4t < N1 N2), w^k = (w^4)^(k/4) w^(k%4) in FP64, at most
// seven roundings from the table values.
template <int N1, int N2, int S, class V>
__device__ __forceinline__ void ystepA(V (&a)[N2], V* buf, const double2* tw, int t) {
  static_assert(N2 >= 4, "w^4 = W^(4t) lies in the table");
  Dft<N2, S, V>::run(a);
  double2 w1 = tw[t], w4 = tw[4 * t];
  if (S > 0) {
    w1.y = -w1.y;
    w4.y = -w4.y;
  }
  const double2 w2 = cmulv(w1, w1), w3 = cmulv(w2, w1);
  double2 q = make_double2(1.0, 0.0);
  buf[t * (N2 + 1)] = a[0];
#pragma unroll
  for (int k2 = 1; k2 < N2; ++k2) {
    const int r = k2 & 3;
    if (r == 0) q = cmulv(q, w4);
    const double2 w = r == 0 ? q : cmulv(q, r == 1 ? w1 : (r == 2 ? w2 : w3));
    buf[t * (N2 + 1) + k2] = cmulv(a[k2], ycvt<V>(w));
  }
}

// Gauss-Lobatto point values, FP64 coefficients to FP64 values in arithmetic
// V: the 2N-point FFT (N1 x N2) of the even extension.
template <int N1, int N2, class V>
__global__ __launch_bounds__(YT<V>::TX* ymax(N1, N2)) void k_ypoints(YArgs a) {
  using T = typename YT<V>::T;
  constexpr int L = N1 * N2, N = L / 2, TX = YT<V>::TX, SB = tile_stride<V>(N1 * (N2 + 1));
  extern __shared__ __align__(16) unsigned char smraw[];
  const int t = threadIdx.x / TX, tx = threadIdx.x - t * TX;
  V* buf = reinterpret_cast<V*>(smraw) + tx * SB;
  const long col = xcd_order(blockIdx.x, gridDim.x) * TX + tx;
  const long nc = a.ncol;
  const double2* in = static_cast<const double2*>(a.in[blockIdx.z]);
  double2* out = static_cast<double2*>(a.out[blockIdx.z]);
  if (t < N1) {
    V x[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
      const int m = n2 * N1 + t;
      x[n2] = ycvt<V>(in[(long)(m <= N ? m : L - m) * nc + col]);
    }
    ystepA<N1, N2, -1>(x, buf, a.twL, t);
  }
  __syncthreads();
  if (t < N2) {
    V b[N1];
    fft_stepB<N1, N2, -1>(b, buf, t);
    const V a0 = ycvt<V>(in[col]), aN = ycvt<V>(in[(long)N * nc + col]);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      const int j = k1 * N2 + t;
      if (j <= N)
        out[(long)j * nc + col] = ycvt<double2>(yscale(yadd(yadd(b[k1], a0), (j & 1) ? yscale(aN, (T)-1) : aN), (T)0.5));
    }
  }
}

// Padded synthesis (ANA = false) or analysis (ANA = true): the 2N-point FFT
// (A1 x A2), then the N-point FFT (A1 x A2/2) through the same LDS tile.
// With these splits thread t of a column holds, for the first FFT, the
// extended sequence at m = n A1 + t, which is index k = m (n < A2/2) and
// index N - k (at n + A2/2) of the second.  SINGLE: the second FFT's inputs
// are formed from the first one's loads (each input read once, more
// registers); otherwise they are read again when the second FFT starts.
template <int A1, int A2, class V, class TI, class TO, bool ANA, bool SINGLE>
__global__ __launch_bounds__(YT<V>::TX* ymax(A1, A2)) void k_ycheb(YArgs a) {
  using T = typename YT<V>::T;
  constexpr int L = A1 * A2, N = L / 2, B2 = A2 / 2, TX = YT<V>::TX;
  static_assert(A2 % 2 == 0, "the second FFT is A1 x A2/2");
  constexpr int SB = tile_stride<V>(A1 * (A2 + 1));
  extern __shared__ __align__(16) unsigned char smraw[];
  const int t = threadIdx.x / TX, tx = threadIdx.x - t * TX;
  V* buf = reinterpret_cast<V*>(smraw) + tx * SB;
  const long col = xcd_order(blockIdx.x, gridDim.x) * TX + tx;
  const long nc = a.ncol;
  const TI* in = static_cast<const TI*>(a.in[blockIdx.z]);
  TO* out = static_cast<TO*>(a.out[blockIdx.z]);
  auto ld = [&](int r) { return ycvt<V>(in[(long)r * nc + col]); };
  auto st = [&](int r, V v) { st_nt(out + (long)r * nc + col, ycvt<TO>(v)); };
  // H_q of the analysis from its two points
  auto hodd = [&](int q, V f, V g) { return q == 0 ? ysub(f, g) : yscale(ysub(f, g), (T)2); };

  V y[B2];  // the N-point FFT's inputs (threads t < A1)
  // Makhoul pre-twiddles w_k, k = n A1 + t: w_t (w_A1)^n, generated in FP64
  const double2 wstep = a.wq[A1];
  if (t < A1) {
    double2 wk = a.wq[t];
    V x[A2];
    if constexpr (SINGLE) {
#pragma unroll
      for (int n = 0; n < B2; ++n) {
        const int k = n * A1 + t;          // < N: m = k, and m = N + k at n + B2
        const int q = k == 0 ? N : N - k;  // the even extension at m = N + k
        if constexpr (!ANA) {
          x[n] = ld(k);
          x[n + B2] = ld(q);
          y[n] = (k == 0) ? x[n] : ypre(x[n], x[n + B2], wk);
        } else {
          const V f0 = ld(k), g0 = ld(2 * N - k), f1 = ld(q), g1 = ld(2 * N - q);
          x[n] = yadd(f0, g0);
          x[n + B2] = yadd(f1, g1);  // q = N: 2 f_N
          const V h0 = hodd(k, f0, g0);
          y[n] = (k == 0) ? h0 : ypre(h0, hodd(q, f1, g1), wk);
        }
        wk = cmulv(wk, wstep);
      }
    } else {
#pragma unroll
      for (int n2 = 0; n2 < A2; ++n2) {
        const int m = n2 * A1 + t;
        const int p = m <= N ? m : L - m;
        x[n2] = ANA ? yadd(ld(p), ld(2 * N - p)) : ld(p);  // analysis, p = N: G_N = 2 f_N
      }
    }
    ystepA<A1, A2, -1>(x, buf, a.twL, t);
  }
  __syncthreads();
  // 2N-point FFT of the even extension: even points / even coefficients
  if (t < A2) {
    V b[A1];
    fft_stepB<A1, A2, -1>(b, buf, t);
    if constexpr (!ANA) {
      const V a0 = ld(0), aN = ld(N);
#pragma unroll
      for (int k1 = 0; k1 < A1; ++k1) {
        const int j = k1 * A2 + t;
        if (j <= N) st(2 * j, yscale(yadd(yadd(b[k1], a0), (j & 1) ? yscale(aN, (T)-1) : aN), (T)0.5));
      }
    } else {
      const T s0 = (T)(0.5 * a.inv2n), s1 = (T)a.inv2n;
#pragma unroll
      for (int k1 = 0; k1 < A1; ++k1) {
        const int m = k1 * A2 + t;
        if (2 * m <= N) st(2 * m, yscale(b[k1], m == 0 ? s0 : s1));
      }
    }
  }
  __syncthreads();  // the tile is reused
  // N-point FFT of the type-III DCT: odd points / odd coefficients
  if (t < A1) {
    if constexpr (!SINGLE) {
      // the inputs read again (the first FFT's are no longer live)
      double2 wk = a.wq[t];
#pragma unroll
      for (int n = 0; n < B2; ++n) {
        const int k = n * A1 + t;
        if constexpr (!ANA) {
          y[n] = (k == 0) ? ld(0) : ypre(ld(k), ld(N - k), wk);
        } else {
          const V h0 = hodd(k, ld(k), ld(2 * N - k));
          y[n] = (k == 0) ? h0 : ypre(h0, hodd(N - k, ld(N - k), ld(N + k)), wk);
        }
        wk = cmulv(wk, wstep);
      }
    }
    ystepA<A1, B2, +1>(y, buf, a.twN, t);
  }
  __syncthreads();
  if (t < B2) {
    V v[A1];
    fft_stepB<A1, B2, +1>(v, buf, t);
#pragma unroll
    for (int k1 = 0; k1 < A1; ++k1) {
      const int n = k1 * B2 + t;
      const int l = n < N / 2 ? 2 * n : 2 * (N - 1 - n) + 1;  // v_n = X_l
      if constexpr (!ANA) {
        st(2 * l + 1, v[k1]);
      } else {
        if (2 * l + 1 <= N) st(2 * l + 1, yscale(v[k1], (T)a.inv2n));
      }
    }
  }
}

// Right-hand sides of the v (POFF = 2) and omega_y (POFF = 1) half systems
// (PF:252-257, 335-340; the explicit operator as the bracket g formed by the
// kinematics pass): f = sum_k g_k T_k at the Gauss-Lobatto points (the type-I
// DCT, FP64), plus the AB2 forcing F = 3 dt/2 h^n - dt/2 h^(n-1), folded by
// parity at the point pairs p = r + POFF, q = N - p: even row r
// ((f+F)_p + (f+F)_q)/2 (rows r < ne - POFF), odd row r ((f+F)_p - (f+F)_q)/2
// (r < no - POFF).  v: each mode's rows as one record, [column][Rv: even rows,
// then odd rows]; omega_y: [row][column] per parity.
//
// FUSE: h^n arrives as coefficients (the h assembly's output) and this kernel
// forms its point values, writes them (the next step's h^(n-1) and the state
// record's h^n) and uses them, in place of a separate point-value pass and a
// read of its output.  Otherwise h^n's point values are read.  Both forms
// compute every value the same way (the point values with the same FFT and
// end terms as k_ypoints, f + F in the step-B layout, the pairs from the LDS
// tile), so a run that switches between them is bit-identical to one that
// does not.
struct YRhsArgs {
  const double2* g;
  const double2* hc;  // FUSE: h^n coefficients
  double2* hn;        // h^n point values (FUSE: written, else read)
  const double2* hp;  // h^(n-1) point values
  double2 *outT, *outE, *outO;
  long ncol;
  int Rv, ne, no;
  double dt;
  const double2* twL;
};

template <int N1, int N2, int POFF, bool FUSE>
__global__ __launch_bounds__(YT<double2>::TX* ymax(N1, N2)) void k_yrhs(YRhsArgs a) {
  constexpr int L = N1 * N2, N = L / 2, TX = YT<double2>::TX, G = ymax(N1, N2);
  constexpr int SB = tile_stride<double2>(N1 * (N2 + 1));
  constexpr int KJ = N / N2 + 1;  // step-B outputs k1 with a point j = k1 N2 + t <= N
  static_assert(SB > N, "the point values fit the exchange tile");
  extern __shared__ __align__(16) unsigned char smraw[];
  const int t = threadIdx.x / TX, tx = threadIdx.x - t * TX;
  double2* buf = reinterpret_cast<double2*>(smraw) + tx * SB;
  const long col = xcd_order(blockIdx.x, gridDim.x) * TX + tx;
  const long nc = a.ncol;
  const double c1 = 1.5 * a.dt, c0 = 0.5 * a.dt;
  // step A of the type-I DCT of the coefficient column src (threads t < N1)
  auto stepA = [&](const double2* src) {
    if (t < N1) {
      double2 x[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) {
        const int m = n2 * N1 + t;
        x[n2] = src[(long)(m <= N ? m : L - m) * nc + col];
      }
      ystepA<N1, N2, -1>(x, buf, a.twL, t);
    }
  };
  // f_j = (S_j + a_0 + (-1)^j a_N) / 2 at j = k1 N2 + t, after step B
  auto endterms = [&](const double2 (&b)[N1], const double2* src, double2 (&f)[KJ]) {
    const double2 s0 = src[col], sN = src[(long)N * nc + col];
#pragma unroll
    for (int k1 = 0; k1 < KJ; ++k1) {
      const int j = k1 * N2 + t;
      f[k1] = yscale(yadd(yadd(b[k1], s0), (j & 1) ? yscale(sN, -1.0) : sN), 0.5);
    }
  };
  // AB2 forcing c1 h^n - c0 h^(n-1) at one point
  auto ab2 = [&](double2 n, double2 p) { return make_double2(c1 * n.x - c0 * p.x, c1 * n.y - c0 * p.y); };
  double2 hf[KJ], hn[KJ], hp[KJ];  // at this thread's step-B points
  if constexpr (FUSE) {
    stepA(a.hc);
    if (t < N2) {
#pragma unroll
      for (int k1 = 0; k1 < KJ; ++k1) {
        const int j = k1 * N2 + t;
        if (j <= N) hp[k1] = a.hp[(long)j * nc + col];
      }
    }
    __syncthreads();
    if (t < N2) {
      double2 b[N1];
      fft_stepB<N1, N2, -1>(b, buf, t);
      endterms(b, a.hc, hn);
#pragma unroll
      for (int k1 = 0; k1 < KJ; ++k1) {
        const int j = k1 * N2 + t;
        if (j <= N) {
          st_nt(a.hn + (long)j * nc + col, hn[k1]);
          hf[k1] = ab2(hn[k1], hp[k1]);
        }
      }
    }
    __syncthreads();  // the tile is reused
  }
  stepA(a.g);
  if constexpr (!FUSE) {
    // requested before the exchange, so that they arrive during it
    if (t < N2) {
#pragma unroll
      for (int k1 = 0; k1 < KJ; ++k1) {
        const int j = k1 * N2 + t;
        if (j <= N) {
          hn[k1] = a.hn[(long)j * nc + col];
          hp[k1] = a.hp[(long)j * nc + col];
        }
      }
    }
  }
  __syncthreads();
  double2 f[KJ];
  if (t < N2) {
    double2 b[N1];
    fft_stepB<N1, N2, -1>(b, buf, t);
    endterms(b, a.g, f);
  }
  __syncthreads();  // the tile now takes the values f + F
  if (t < N2) {
#pragma unroll
    for (int k1 = 0; k1 < KJ; ++k1) {
      const int j = k1 * N2 + t;
      if (j <= N) {
        const double2 h = FUSE ? hf[k1] : ab2(hn[k1], hp[k1]);
        buf[j] = make_double2(f[k1].x + h.x, f[k1].y + h.y);
      }
    }
  }
  __syncthreads();
  const int Me = a.ne - POFF, Mo = a.no - POFF;
  for (int r = t; r < Me; r += G) {
    const int p = r + POFF, q = N - p;
    const double2 vp = buf[p];
    double2 ev, od = make_double2(0.0, 0.0);
    if (q == p) {
      ev = vp;
    } else {
      const double2 vq = buf[q];
      ev = yscale(yadd(vp, vq), 0.5);
      od = yscale(ysub(vp, vq), 0.5);
    }
    if constexpr (POFF == 2) {
      double2* rec = a.outT + col * a.Rv;
      rec[r] = ev;
      if (r < Mo) rec[Me + r] = od;
    } else {
      a.outE[(long)r * nc + col] = ev;
      if (r < Mo) a.outO[(long)r * nc + col] = od;
    }
  }
}

template <class K, class A>
static hipError_t launch_y(Ctx& c, K kernel, int tx, int nthreads, size_t shm, const A& a, int nf) {
  if (c.d.ncol % tx) return hipErrorNotSupported;
  if (shm > 65536) {
    const hipError_t e = hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kernel, dim3((unsigned)(c.d.ncol / tx), 1, nf), dim3(nthreads), shm, c.stream, a);
  return hipGetLastError();
}

// Which form of the second FFT's inputs each kernel takes (measured at
// 256x193x256, ms, single load / read again: FP32 synthesis 0.296 / 0.336,
// analysis 0.156 / 0.195; FP64 synthesis 0.540 / 0.458, analysis 0.230 /
// 0.286).
template <class V, bool ANA>
constexpr bool single_load() {
  return sizeof(V) == 8 || ANA;
}

// mode 0: point values, 1: padded synthesis, 2: analysis; A1 x A2 = 2N
template <int A1, int A2, class V, class TI, class TO>
static hipError_t run_y(Ctx& c, int mode, const YArgs& a, int nf) {
  constexpr int TX = YT<V>::TX, nt = TX * ymax(A1, A2);
  const size_t shm = sizeof(V) * (size_t)TX * tile_stride<V>(A1 * (A2 + 1));
  if (mode == 0) {
    // point values are FP64 arithmetic only (h^n, the statistics profiles)
    if constexpr (sizeof(V) == sizeof(double2)) return launch_y(c, k_ypoints<A1, A2, V>, TX, nt, shm, a, nf);
    return hipErrorNotSupported;
  }
  if (mode == 1) return launch_y(c, k_ycheb<A1, A2, V, TI, TO, false, single_load<V, false>()>, TX, nt, shm, a, nf);
  return launch_y(c, k_ycheb<A1, A2, V, TI, TO, true, single_load<V, true>()>, TX, nt, shm, a, nf);
}

// 1 when 2N has a split below (and so an FFT form of the transforms).
int ydct_length_supported(int N) {
  switch (2 * N) {
    case 64: case 96: case 128: case 192: case 256: case 384: case 512: return 1;
    default: return 0;
  }
}

template <class V, class TI, class TO>
static hipError_t run_ydct(Ctx& c, int mode, const TI* const* in, TO* const* out, int nf) {
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
    case 64: return run_y<8, 8, V, TI, TO>(c, mode, a, nf);
    case 96: return run_y<8, 12, V, TI, TO>(c, mode, a, nf);
    case 128: return run_y<8, 16, V, TI, TO>(c, mode, a, nf);
    case 192: return run_y<12, 16, V, TI, TO>(c, mode, a, nf);
    case 256: return run_y<16, 16, V, TI, TO>(c, mode, a, nf);
    case 384: return run_y<16, 24, V, TI, TO>(c, mode, a, nf);
    case 512: return run_y<16, 32, V, TI, TO>(c, mode, a, nf);
    default: return hipErrorNotSupported;
  }
}

hipError_t launch_ypoints(Ctx& c, const double2* const* in, double2* const* out, int nf) {
  return run_ydct<double2>(c, 0, in, out, nf);
}
hipError_t launch_ysynth(Ctx& c, const double2* const* in, double2* const* out, int nf) {
  return run_ydct<double2>(c, 1, in, out, nf);
}
hipError_t launch_yanalysis(Ctx& c, const double2* const* in, double2* const* out, int nf) {
  return run_ydct<double2>(c, 2, in, out, nf);
}
template <int N1, int N2>
static hipError_t run_yrhs(Ctx& c, int which, bool fuse, const YRhsArgs& a) {
  constexpr int TX = YT<double2>::TX, nt = TX * ymax(N1, N2);
  const size_t shm = sizeof(double2) * (size_t)TX * tile_stride<double2>(N1 * (N2 + 1));
  if (which == 0)
    return fuse ? launch_y(c, k_yrhs<N1, N2, 2, true>, TX, nt, shm, a, 1)
                : launch_y(c, k_yrhs<N1, N2, 2, false>, TX, nt, shm, a, 1);
  return fuse ? launch_y(c, k_yrhs<N1, N2, 1, true>, TX, nt, shm, a, 1)
              : launch_y(c, k_yrhs<N1, N2, 1, false>, TX, nt, shm, a, 1);
}

hipError_t launch_yrhs(Ctx& c, int which, const double2* g, const double2* hc, double2* hn, const double2* hp,
                       double2* oE, double2* oO, double2* oT) {
  YRhsArgs a{};
  a.g = g;
  a.hc = hc;
  a.hn = hn;
  a.hp = hp;
  a.outT = oT;
  a.outE = oE;
  a.outO = oO;
  a.ncol = c.d.ncol;
  a.Rv = c.d.ne + c.d.no - 4;
  a.ne = c.d.ne;
  a.no = c.d.no;
  a.dt = c.prm.dt;
  a.twL = c.twy;
  const bool fuse = hc != nullptr;
  switch (2 * c.d.N) {
    case 64: return run_yrhs<8, 8>(c, which, fuse, a);
    case 96: return run_yrhs<8, 12>(c, which, fuse, a);
    case 128: return run_yrhs<8, 16>(c, which, fuse, a);
    case 192: return run_yrhs<12, 16>(c, which, fuse, a);
    case 256: return run_yrhs<16, 16>(c, which, fuse, a);
    case 384: return run_yrhs<16, 24>(c, which, fuse, a);
    case 512: return run_yrhs<16, 32>(c, which, fuse, a);
    default: return hipErrorNotSupported;
  }
}

hipError_t launch_ysynth32(Ctx& c, const double2* const* in, float2* const* out, int nf) {
  return run_ydct<float2>(c, 1, in, out, nf);
}
hipError_t launch_ysynth32f(Ctx& c, const float2* const* in, float2* const* out, int nf) {
  return run_ydct<float2>(c, 1, in, out, nf);
}
hipError_t launch_yanalysis32f(Ctx& c, const float2* const* in, float2* const* out, int nf) {
  return run_ydct<float2>(c, 2, in, out, nf);
}
hipError_t launch_yanalysis32(Ctx& c, const float2* const* in, double2* const* out, int nf) {
  return run_ydct<float2>(c, 2, in, out, nf);
}

}  // namespace chf
