// Wall-normal contractions on matrix cores:
//
//   out = epilogue( A_m . B_{s(m)} , m = 0..NACC-1 )
//
// A_m: [M x K] matrices whose entries are cosines cos(pi n / 2N) (Chebyshev
//      polynomials at Gauss-Lobatto points), generated per lane;
// B_s: NB real column streams [K x ncold]: Prob::fetchB loads NR raw values
//      per (k, column) from the field arrays (issued one tile ahead) and
//      Prob::makeB forms the stream values from them when the tile is staged
//      (parity gathers, mirrored-point folds, conversion to the compute type);
// one workgroup owns all M rows of a BN-column tile, so every B element is
// read from HBM once; wave w of NWV owns row blocks w, w + NWV, w + 2 NWV, ...
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

// The matrix entries are all of the form cos(pi n / 2N): A_m(r, k) =
// costab[(r0(m, r) + k * dk(m, r)) mod 4N], read from a 4N-entry table in LDS
// through indices that each lane advances incrementally per k block, so no
// matrix bytes move from HBM or L2.  B tiles are software-pipelined: the next
// tile is loaded into registers while the current one feeds the matrix cores.
//
// Problems with P::kCentre may leave one output row out of M (p.centre set):
// a row whose matrix entries along k are P::centre_sign(k mod KC) in {1, -1, 0}
// on the even stream (the synthesis at y = 0, T_l(0) = (-1)^(l/2)).  Lanes 0-15 of each wave sum it for one
// column each from the B tiles, in k order whatever the tile shape, and hand
// it to P::store_centre, so a lone row does not cost a whole 16-row block of
// matrix-core work.
//
// FULL: every wave owns all RBW of its row blocks (MB == NWV * RBW).  The k
// step then has no row-block guards, and the compiler issues the table reads
// of all row blocks ahead of their matrix ops.
// Minimum waves per SIMD the compiler must leave room for: P::kMinWaves for
// tiles whose accumulators take at most 48 registers (wider tiles would
// spill), otherwise no constraint.
// Problems with P::kVecCols and tiles of CB >= 2 column blocks give each lane
// CB adjacent columns (CB j .. CB j + CB-1 for lane j of a block row) instead
// of columns j, j + 16, ...: the results are the same, and the epilogue writes
// a lane's columns with one vector store, so each row of a tile leaves as
// whole lines (FP32 rows of 16 columns are half a 128-byte line otherwise).
template <class P, int CB>
constexpr bool vec_cols() { return P::kVecCols && CB >= 2; }

template <class P, int RBW, int CB>
constexpr int min_waves() {
  return (P::NACC * RBW * CB * 4 * (int)sizeof(typename P::TC) / 4 <= 48) ? P::kMinWaves : 1;
}

template <class P, int RBW, int CB, int NWV, bool FULL>
__global__ __launch_bounds__(64 * NWV) __attribute__((amdgpu_waves_per_eu(min_waves<P, RBW, CB>())))
void k_mfma_cos(P pb) {
  using TC = typename P::TC;
  using MF = Mf<TC>;
  constexpr int NT = 64 * NWV;
  constexpr int NA = P::NACC, NB = P::NB, KC = P::KC, BN = 16 * CB;
  constexpr int SB = (BN % 32 == 0) ? BN + 16 : BN;
  constexpr int PER = (KC * BN + NT - 1) / NT;  // tile elements per thread
  constexpr bool REC = cos_recurrence<P>();
  __shared__ TC sB[NB][KC][SB];
  static_assert(KC % 4 == 0, "B tiles hold whole k blocks (and the centre signs repeat every 4)");
  static_assert(BN <= 16 * NWV, "one centre-row column per lane < 16 of each wave");
  extern __shared__ __align__(16) unsigned char stabraw[];
  TC* stab = reinterpret_cast<TC*>(stabraw);  // 4N cosines (table form)
  // wave index in a scalar register: the row-block guards below are then
  // uniform branches, not per-lane masks that cut the k step into pieces
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  P p = pb;
  p.select(blockIdx.z);
  const int n4 = p.n4;
  if constexpr (!REC)
    for (int i = threadIdx.x; i < n4; i += NT) stab[i] = p.costab[i];
  // Column tiles in XCD-contiguous order: the 16-column tiles that share the
  // cache lines of narrow (FP32) rows are written through the same L2.
  const long tile = xcd_order(blockIdx.x, gridDim.x);
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
      const int row = (wave + NWV * r) * 16 + (lane & 15);
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
  // raw loads of the next tile: nothing consumes them until it is staged, so
  // they stay in flight across the current tile's matrix-core work
  using TL = typename P::TL;
  constexpr int NR = P::NR;
  TL nxt[PER][NR];
  auto fetch = [&](int kc) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int i = threadIdx.x + q * NT;
      const int kk = i / BN, c = i - kk * BN;
      const int k = kc * KC + kk;
      const long col = col0 + c;
#pragma unroll
      for (int s = 0; s < NR; ++s) nxt[q][s] = 0;
      if (i < KC * BN && k < p.sh.K && col < p.ncold) p.fetchB(k, col, nxt[q]);
    }
  };
  TC cen = 0;  // centre-row sum of this lane's column (kCentre problems)
  fetch(0);
  for (int kc = 0; kc < nkc; ++kc) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int i = threadIdx.x + q * NT;
      if (i < KC * BN) {
        const int kk = i / BN, c = i - kk * BN;
        TC v[NB];
        p.makeB(kc * KC + kk, nxt[q], v);
#pragma unroll
        for (int s = 0; s < NB; ++s) sB[s][kk][c] = v[s];
      }
    }
    __syncthreads();
    if (kc + 1 < nkc) fetch(kc + 1);
    if constexpr (P::kCentre) {
      // column wave * 16 + lane (lane < 16) sums its rows in k order, so the
      // summation order does not depend on the tile shape
      if (p.centre && lane < 16 && wave * 16 < BN) {
        const int c = wave * 16 + lane;
#pragma unroll
        for (int kk = 0; kk < KC; ++kk) cen += P::centre_sign(kk) * sB[0][kk][c];
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
          for (int c = 0; c < CB; ++c)
            bv[s][c] = sB[s][kb * 4 + (lane >> 4)][vec_cols<P, CB>() ? CB * (lane & 15) + c : c * 16 + (lane & 15)];
#pragma unroll
        for (int r = 0; r < RBW; ++r) {
          if (FULL || wave + NWV * r < MB) {
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
    if (p.centre && lane < 16 && wave * 16 < BN) {
      const long col = col0 + wave * 16 + lane;
      if (col < p.ncold) p.store_centre(col, cen);
    }
  }
  if constexpr (vec_cols<P, CB>()) {
    // ncold is a multiple of 16 and so of CB: a lane's columns are all inside
#pragma unroll
    for (int r = 0; r < RBW; ++r) {
      const int rb = wave + NWV * r;
      if (!FULL && rb >= MB) continue;
      const long col = col0 + CB * (lane & 15);
      if (col >= p.ncold) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = rb * 16 + MF::row(lane, e);
        if (row >= p.sh.M) continue;
        TC v[NA][CB];
#pragma unroll
        for (int m = 0; m < NA; ++m)
#pragma unroll
          for (int c = 0; c < CB; ++c) v[m][c] = acc[m][r][c][e];
        p.template store_vec<CB>(row, col, v);
      }
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < RBW; ++r) {
    const int rb = wave + NWV * r;
    if (!FULL && rb >= MB) continue;
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

template <class P, int RBW, int CB, int NWV>
hipError_t launch_mfma_cos(hipStream_t st, const P& p, int nz) {
  const long tiles = (p.ncold + 16 * CB - 1) / (16 * CB);
  const size_t shm = cos_recurrence<P>() ? 0 : sizeof(typename P::TC) * p.n4;
  if (p.sh.MB == NWV * RBW) {
    hipLaunchKernelGGL((k_mfma_cos<P, RBW, CB, NWV, true>), dim3((unsigned)tiles, 1, nz), dim3(64 * NWV), shm, st,
                       p);
  } else {
    // blocks of more than four waves are launched with one row block per wave only
    if constexpr (NWV > 4)
      return hipErrorInvalidValue;
    else
      hipLaunchKernelGGL((k_mfma_cos<P, RBW, CB, NWV, false>), dim3((unsigned)tiles, 1, nz), dim3(64 * NWV), shm,
                         st, p);
  }
  return hipGetLastError();
}

// cb = column blocks (16 columns each) per wave, as picked by the autotuner;
// 0 or an unsupported value takes the default (8 accumulator blocks per wave,
// which keeps 2-3 waves per SIMD).  Choices are kept within 16 accumulator
// blocks per wave.  The tile choice does not change any output's summation
// order, so every choice gives bit-identical results.
template <class P, int RBW, int NWV>
hipError_t run_mfma_cos_rbw(hipStream_t st, const P& p, int nz, int cb) {
  constexpr int budget = 8 / (P::NACC * RBW) > 0 ? 8 / (P::NACC * RBW) : 1;
  constexpr int CBH = budget >= 4 ? 4 : (budget >= 2 ? 2 : 1);
  if (cb == 1) return launch_mfma_cos<P, RBW, 1, NWV>(st, p, nz);
  if constexpr (P::NACC * RBW * 2 <= 16)
    if (cb == 2) return launch_mfma_cos<P, RBW, 2, NWV>(st, p, nz);
  if constexpr (P::NACC * RBW * 4 <= 16)
    if (cb == 4) return launch_mfma_cos<P, RBW, 4, NWV>(st, p, nz);
  return launch_mfma_cos<P, RBW, CBH, NWV>(st, p, nz);
}

template <class P>
hipError_t run_mfma_cos(hipStream_t st, const P& p, int nz, int cb = 0) {
  // P::kWavePerRowBlock with 5 or 6 row blocks: one wave per row block, so
  // every wave has the same work (with four waves, 2 + 2 + 1 + 1 leaves two
  // waves idle for half the k loop).  Measured per problem: faster for the
  // right-hand-side brackets, slower for the point synthesis; seven-wave
  // blocks (7 row blocks) were slower for every problem.
  if constexpr (P::kWavePerRowBlock) {
    switch (p.sh.MB) {
      case 5: return run_mfma_cos_rbw<P, 1, 5>(st, p, nz, cb);
      case 6: return run_mfma_cos_rbw<P, 1, 6>(st, p, nz, cb);
      default: break;
    }
  }
  // otherwise four waves per block (measured: three waves for 6 row blocks,
  // 2 + 2 + 2 instead of 2 + 2 + 1 + 1, is slower: fewer waves hide less latency)
  switch ((p.sh.MB + 3) / 4) {
    case 1: return run_mfma_cos_rbw<P, 1, 4>(st, p, nz, cb);
    case 2: return run_mfma_cos_rbw<P, 2, 4>(st, p, nz, cb);
    case 3: return run_mfma_cos_rbw<P, 3, 4>(st, p, nz, cb);
    case 4: return run_mfma_cos_rbw<P, 4, 4>(st, p, nz, cb);
    case 5: return run_mfma_cos_rbw<P, 5, 4>(st, p, nz, cb);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace chf
