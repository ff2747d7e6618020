// Generic wall-normal contraction on FP64 matrix cores.
//
//   out = epilogue( A_m . B_{s(m)} , m = 0..NACC-1 )
//
// A_m: small real matrices [M x K] (row = output row), pre-packed on the host
//      in v_mfma_f64_16x16x4f64 operand order: frag[m][rb][kb][lane] =
//      A_m[16 rb + lane%16][4 kb + lane/16];
// B_s: NB real column streams [K x ncold] produced by Prob::loadB from the
//      field arrays (parity gathers, mirrored-point folds);
// one workgroup owns all M rows of a BN-column tile, so every B element is
// read from HBM once; wave w owns row blocks w, w+4, w+8, ...
#pragma once
#include "chf_internal.h"

namespace chf {

typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef float fvec4 __attribute__((ext_vector_type(4)));

// Matrix-core traits per compute type.  Operand layouts (both types, probed):
// A lane l = A[l%16][l/16], B lane l = B[l/16][l%16]; the 16x16 result rows
// differ: FP64 lane l element e = D[l/16 + 4e][l%16], FP32 = D[4(l/16) + e][l%16].
template <class TC>
struct Mf;
template <>
struct Mf<double> {
  using acc = dvec4;
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int e) { return (lane >> 4) + 4 * e; }
};
template <>
struct Mf<float> {
  using acc = fvec4;
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int e) { return 4 * (lane >> 4) + e; }
};

// A contraction with P::kRecurrence set and FP32 arithmetic takes its A
// operand from an FP64 cosine recurrence in registers (one FMA per matrix op,
// no LDS table); the others read the LDS table.  Which is faster depends on
// the shape (measured: recurrence for the analysis, table for the synthesis).
template <class P>
constexpr bool cos_recurrence() { return sizeof(typename P::TC) == 4 && P::kRecurrence; }

// One of six per-field values, chosen by a wave-uniform index through
// constant-index reads: indexing the array with the runtime value would make
// the parameter block addressable and move it to scratch memory.
template <class T>
__device__ __forceinline__ T pick6(T const (&a)[6], int z) {
  switch (z) {
    case 0: return a[0];
    case 1: return a[1];
    case 2: return a[2];
    case 3: return a[3];
    case 4: return a[4];
    default: return a[5];
  }
}

template <class P, int RBW, int CB>
__global__ __launch_bounds__(256) void k_mfma(P pb) {
  constexpr int NA = P::NACC, NB = P::NB, KC = 16, BN = 16 * CB;
  constexpr int SB = (BN % 32 == 0) ? BN + 16 : BN;  // row stride: 2 rows per half-wave hit disjoint banks
  __shared__ double sB[NB][KC][SB];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  P p = pb;
  p.select(blockIdx.z);
  const long col0 = (long)blockIdx.x * BN;
  dvec4 acc[NA][RBW][CB];
#pragma unroll
  for (int m = 0; m < NA; ++m)
#pragma unroll
    for (int r = 0; r < RBW; ++r)
#pragma unroll
      for (int c = 0; c < CB; ++c) acc[m][r][c] = dvec4{0.0, 0.0, 0.0, 0.0};
  const int MB = p.sh.MB, KB = p.sh.KB;
  const int nkc = (KB * 4 + KC - 1) / KC;
  for (int kc = 0; kc < nkc; ++kc) {
    for (int i = threadIdx.x; i < KC * BN; i += 256) {
      const int kk = i / BN, c = i - kk * BN;
      const int k = kc * KC + kk;
      const long col = col0 + c;
      double v[NB];
      if (k < p.sh.K && col < p.ncold) {
        p.loadB(k, col, v);
      } else {
#pragma unroll
        for (int s = 0; s < NB; ++s) v[s] = 0.0;
      }
#pragma unroll
      for (int s = 0; s < NB; ++s) sB[s][kk][c] = v[s];
    }
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < KC / 4; ++kb) {
      const int kg = kc * (KC / 4) + kb;
      if (kg < KB) {
        double bv[NB][CB];
#pragma unroll
        for (int s = 0; s < NB; ++s)
#pragma unroll
          for (int c = 0; c < CB; ++c) bv[s][c] = sB[s][kb * 4 + (lane >> 4)][c * 16 + (lane & 15)];
#pragma unroll
        for (int r = 0; r < RBW; ++r) {
          const int rb = wave + 4 * r;
          if (rb < MB) {
#pragma unroll
            for (int m = 0; m < NA; ++m) {
              const double a = p.frag[m][((long)rb * KB + kg) * 64 + lane];
#pragma unroll
              for (int c = 0; c < CB; ++c)
                acc[m][r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv[P::bsel(m)][c], acc[m][r][c], 0, 0, 0);
            }
          }
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < RBW; ++r) {
    const int rb = wave + 4 * r;
    if (rb >= MB) continue;
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      const long col = col0 + c * 16 + (lane & 15);
      if (col >= p.ncold) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        // v_mfma_f64_16x16x4f64 result: lane l, element e holds D[l/16 + 4e][l%16]
        const int row = rb * 16 + (lane >> 4) + 4 * e;
        if (row >= p.sh.M) continue;
        double v[NA];
#pragma unroll
        for (int m = 0; m < NA; ++m) v[m] = acc[m][r][c][e];
        p.store(row, col, v);
      }
    }
  }
}

// Variant for the transform contractions, whose matrix entries are all of the
// form cos(pi n / 2N): A_m(r, k) = costab[(r0(m, r) + k * dk(m, r)) mod 4N],
// read from a 4N-entry table in LDS through indices that each lane advances
// incrementally per k block, so no matrix bytes move from HBM or L2.  B tiles
// are software-pipelined: the next tile is loaded into registers while the
// current one feeds the matrix cores.
template <class P, int RBW, int CB>
__global__ __launch_bounds__(256) void k_mfma_cos(P pb) {
  using TC = typename P::TC;
  using MF = Mf<TC>;
  constexpr int NA = P::NACC, NB = P::NB, KC = 16, BN = 16 * CB;
  constexpr int SB = (BN % 32 == 0) ? BN + 16 : BN;
  constexpr int PER = (KC * BN + 255) / 256;  // tile elements per thread
  constexpr bool REC = cos_recurrence<P>();
  __shared__ TC sB[NB][KC][SB];
  static_assert(NB * KC * SB >= 256 && 256 % BN == 0, "centre-row reduction layout");
  extern __shared__ __align__(16) unsigned char stabraw[];
  TC* stab = reinterpret_cast<TC*>(stabraw);  // 4N cosines (table form)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  P p = pb;
  p.select(blockIdx.z);
  const int n4 = p.n4;
  if constexpr (!REC)
    for (int i = threadIdx.x; i < n4; i += 256) stab[i] = p.costab[i];
  // Column tiles in XCD-contiguous order: the 16-column tiles that share the
  // cache lines of narrow (FP32) rows are written through the same L2.
  const long tile = (gridDim.x & 7) ? (long)blockIdx.x
                                    : (long)((blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3));
  const long col0 = tile * BN;
  const int MB = p.sh.MB, KB = p.sh.KB;
  // Table form, per (m, r): current table index and its per-k-block increment.
  // Recurrence form: the lane's A values along k are cos(phi0 + n dphi), so
  // c_{n+1} = 2 cos(dphi) c_n - c_{n-1}, run in FP64 from two table values.
  int idx[NA][RBW], stp[NA][RBW];
  double cc[NA][RBW], cp[NA][RBW], c2[NA][RBW];
#pragma unroll
  for (int m = 0; m < NA; ++m)
#pragma unroll
    for (int r = 0; r < RBW; ++r) {
      const int row = (wave + 4 * r) * 16 + (lane & 15);
      const int k0 = lane >> 4;
      idx[m][r] = p.aindex(m, row, k0, n4);
      stp[m][r] = p.astep(m, row, n4);  // index advance for k -> k + 4
      if constexpr (REC) {
        const int i0 = idx[m][r], st = stp[m][r];
        int im = i0 - st;
        im = im < 0 ? im + n4 : im;
        cc[m][r] = p.rtab[i0];
        cp[m][r] = p.rtab[im];
        c2[m][r] = 2.0 * p.rtab[st];
      }
    }
  typename MF::acc acc[NA][RBW][CB];
#pragma unroll
  for (int m = 0; m < NA; ++m)
#pragma unroll
    for (int r = 0; r < RBW; ++r)
#pragma unroll
      for (int c = 0; c < CB; ++c) acc[m][r][c] = typename MF::acc{0, 0, 0, 0};
  const int nkc = (KB * 4 + KC - 1) / KC;
  TC nxt[PER][NB];
  auto fetch = [&](int kc) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int i = threadIdx.x + q * 256;
      const int kk = i / BN, c = i - kk * BN;
      const int k = kc * KC + kk;
      const long col = col0 + c;
      if (i < KC * BN && k < p.sh.K && col < p.ncold) {
        p.loadB(k, col, nxt[q]);
      } else {
#pragma unroll
        for (int s = 0; s < NB; ++s) nxt[q][s] = 0;
      }
    }
  };
  TC cen = 0;  // this thread's part of the centre-row sum (kCentre problems)
  fetch(0);
  for (int kc = 0; kc < nkc; ++kc) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int i = threadIdx.x + q * 256;
      if (i < KC * BN) {
        const int kk = i / BN, c = i - kk * BN;
#pragma unroll
        for (int s = 0; s < NB; ++s) sB[s][kk][c] = nxt[q][s];
      }
    }
    __syncthreads();
    if (kc + 1 < nkc) fetch(kc + 1);
    if constexpr (P::kCentre) {
      // centre row: alternating sum of the even-stream rows (KC is even, so the
      // sign of row kk is (-1)^kk); every thread keeps one column (256 % BN == 0)
      if (p.centre) {
#pragma unroll
        for (int q = 0; q < CB; ++q) {
          const int i = threadIdx.x + q * 256;
          const int kk = i / BN, c = i - kk * BN;
          const TC v = sB[0][kk][c];
          cen = (kk & 1) ? cen - v : cen + v;
        }
      }
    }
#pragma unroll
    for (int kb = 0; kb < KC / 4; ++kb) {
      const int kg = kc * (KC / 4) + kb;
      if (kg < KB) {
        TC bv[NB][CB];
#pragma unroll
        for (int s = 0; s < NB; ++s)
#pragma unroll
          for (int c = 0; c < CB; ++c) bv[s][c] = sB[s][kb * 4 + (lane >> 4)][c * 16 + (lane & 15)];
#pragma unroll
        for (int r = 0; r < RBW; ++r) {
          const int rb = wave + 4 * r;
          if (rb < MB) {
#pragma unroll
            for (int m = 0; m < NA; ++m) {
              TC a;
              if constexpr (REC) {
                a = (TC)cc[m][r];
                const double nx = fma(c2[m][r], cc[m][r], -cp[m][r]);
                cp[m][r] = cc[m][r];
                cc[m][r] = nx;
              } else {
                a = stab[idx[m][r]];
                int t = idx[m][r] + stp[m][r];
                idx[m][r] = t >= n4 ? t - n4 : t;
              }
#pragma unroll
              for (int c = 0; c < CB; ++c)
                acc[m][r][c] = MF::mma(a, bv[P::bsel(m)][c], acc[m][r][c]);
            }
          }
        }
      }
    }
    __syncthreads();
  }
  if constexpr (P::kCentre) {
    if (p.centre) {
      // reduce the 256 / BN partial sums of each column through the (now idle) B tile
      TC* red = &sB[0][0][0];
      red[threadIdx.x] = cen;
      __syncthreads();
      if (threadIdx.x < BN) {
        TC v = 0;
        for (int j = 0; j < 256 / BN; ++j) v += red[threadIdx.x + j * BN];
        const long col = col0 + threadIdx.x;
        if (col < p.ncold) p.store_centre(col, v);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RBW; ++r) {
    const int rb = wave + 4 * r;
    if (rb >= MB) continue;
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      const long col = col0 + c * 16 + (lane & 15);
      if (col >= p.ncold) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = rb * 16 + MF::row(lane, e);
        if (row >= p.sh.M) continue;
        TC v[NA];
#pragma unroll
        for (int m = 0; m < NA; ++m) v[m] = acc[m][r][c][e];
        p.store(row, col, v);
      }
    }
  }
}

template <class P, int RBW, int CB>
hipError_t launch_mfma_cos(hipStream_t st, const P& p, int nz) {
  const long tiles = (p.ncold + 16 * CB - 1) / (16 * CB);
  const size_t shm = cos_recurrence<P>() ? 0 : sizeof(typename P::TC) * p.n4;
  hipLaunchKernelGGL((k_mfma_cos<P, RBW, CB>), dim3((unsigned)tiles, 1, nz), dim3(256), shm, st, p);
  return hipGetLastError();
}

// cb = column blocks (16 columns each) per wave, as picked by the autotuner;
// 0 or an unsupported value takes the default (8 accumulator blocks per wave,
// which keeps 2-3 waves per SIMD).  Choices are kept within 16 accumulator
// blocks per wave.  The tile choice does not change any output's summation
// order, so every choice gives bit-identical results.
template <class P, int RBW>
hipError_t run_mfma_cos_rbw(hipStream_t st, const P& p, int nz, int cb) {
  constexpr int budget = 8 / (P::NACC * RBW) > 0 ? 8 / (P::NACC * RBW) : 1;
  constexpr int CBH = budget >= 4 ? 4 : (budget >= 2 ? 2 : 1);
  if (cb == 1) return launch_mfma_cos<P, RBW, 1>(st, p, nz);
  if constexpr (P::NACC * RBW * 2 <= 16)
    if (cb == 2) return launch_mfma_cos<P, RBW, 2>(st, p, nz);
  if constexpr (P::NACC * RBW * 4 <= 16)
    if (cb == 4) return launch_mfma_cos<P, RBW, 4>(st, p, nz);
  return launch_mfma_cos<P, RBW, CBH>(st, p, nz);
}

template <class P>
hipError_t run_mfma_cos(hipStream_t st, const P& p, int nz, int cb = 0) {
  const int rbw = (p.sh.MB + 3) / 4;
  switch (rbw) {
    case 1: return run_mfma_cos_rbw<P, 1>(st, p, nz, cb);
    case 2: return run_mfma_cos_rbw<P, 2>(st, p, nz, cb);
    case 3: return run_mfma_cos_rbw<P, 3>(st, p, nz, cb);
    case 4: return run_mfma_cos_rbw<P, 4>(st, p, nz, cb);
    case 5: return run_mfma_cos_rbw<P, 5>(st, p, nz, cb);
    default: return hipErrorInvalidValue;
  }
}

// Picks rows-per-wave (from M) and columns-per-wave so that the accumulators
// stay within NACC * RBW * CB <= 16 (64 doubles per lane).
template <class P, int RBW>
hipError_t run_mfma_rbw(hipStream_t st, const P& p, int nz) {
  constexpr int budget = 16 / (P::NACC * RBW);
  constexpr int CB = budget >= 4 ? 4 : (budget >= 2 ? 2 : 1);
  const long tiles = (p.ncold + 16 * CB - 1) / (16 * CB);
  hipLaunchKernelGGL((k_mfma<P, RBW, CB>), dim3((unsigned)tiles, 1, nz), dim3(256), 0, st, p);
  return hipGetLastError();
}

template <class P>
hipError_t run_mfma(hipStream_t st, const P& p, int nz) {
  const int rbw = (p.sh.MB + 3) / 4;
  switch (rbw) {
    case 1: return run_mfma_rbw<P, 1>(st, p, nz);
    case 2: return run_mfma_rbw<P, 2>(st, p, nz);
    case 3: return run_mfma_rbw<P, 3>(st, p, nz);
    case 4: return run_mfma_rbw<P, 4>(st, p, nz);
    case 5: return run_mfma_rbw<P, 5>(st, p, nz);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace chf
