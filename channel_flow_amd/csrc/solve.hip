// Per-(kx,|kz|) solves of the v and omega_y equations with kinematics
// (PF:242-394, getvelocityandvorticity PF:2347-2407), the mean mode
// (PF:401-474), and the GPU set-up of the operator inverses (PF:181-204).
//
// The collocation operators of get_coef_v (PF:3348-3381) and get_coef_omegay
// (PF:3432-3455) depend on (kx, kz) through k^2 only, so the modes +kz and
// -kz share one matrix, and with the grid symmetric about y = 0 each
// operator decouples into an even-degree and an odd-degree system:
//   v   : rows (r0+r1)/2, (r2-r3)/2 and the symmetric parts at points
//         p = 2..N/2 (even);  (r0-r1)/2, (r2+r3)/2 and the antisymmetric parts
//         at p = 2..(N-1)/2 (odd);
//   w_y : row (r0+-r1)/2 and the (anti)symmetric parts at p = 1..N/2.
// Set-up inverts each half system once (Gauss-Jordan with partial pivoting in
// LDS); a step multiplies the stored inverses, as the reference multiplies the
// zgetri inverses with a_multiply_x (PF:4123-4139).
#include "chf_internal.h"
#include "fftreg.h"
#include "mfma.h"

namespace chf {

// ------------------------------------------------------------------ set-up

struct SetupArgs {
  const double *t0, *t2, *t4;
  const double *al, *be;
  double* inv;      // v halves (and omega halves when not diagonalised)
  double* inv_w;    // omega halves
  int Mx, N, Ny, ne, no;
  long stride, row, stride_w, row_w;
  int wdiag;        // omega operator diagonalised: only group 0's omega halves are stored
  const int* crep;  // representative group of each operator class
  double dt, re;
};

__device__ inline double op_v(const SetupArgs& a, double k2, int p, int l) {
  // get_coef_v interior row (PF:3361-3362)
  const int o = p * a.Ny + l;
  return -a.t4[o] * a.dt * 0.5 / a.re + (1 + k2 * a.dt / a.re) * a.t2[o] -
         k2 * (1 + k2 * a.dt * 0.5 / a.re) * a.t0[o];
}

__device__ inline double op_w(const SetupArgs& a, double k2, int p, int l) {
  // get_coef_omegay interior row (PF:3444); at k^2 = 0 this is get_coef_u0 (PF:3502)
  const int o = p * a.Ny + l;
  return -a.t2[o] * a.dt * 0.5 / a.re + (1 + k2 * a.dt * 0.5 / a.re) * a.t0[o];
}

// grid: (operator classes, 4) ; blockIdx.y = eq*2 + parity (eq 0 = v, 1 = omega_y).
// A class is the set of groups whose k^2 is bitwise equal; it is inverted once,
// at its first group's k^2.
__global__ __launch_bounds__(256) void k_setup(SetupArgs a) {
  extern __shared__ double sm[];
  __shared__ int piv[160];
  __shared__ double colk[160];
  const int cls = blockIdx.x;
  const int g = a.crep[cls];
  const int eq = blockIdx.y >> 1, par = blockIdx.y & 1;
  const int kx = g % a.Mx, kz = g / a.Mx;
  const double k2 = a.al[kx] * a.al[kx] + a.be[kz] * a.be[kz];
  const int n = par ? a.no : a.ne;
  const int nP = n + (n & 1);  // even width: 16-byte aligned row pairs
  const int neP = a.ne + (a.ne & 1);
  if (eq == 1 && g > 0 && a.wdiag) return;
  const long orow = eq ? a.row_w : a.row;
  double* out = (eq ? a.inv_w + (long)cls * a.stride_w : a.inv + (long)cls * a.stride) + par * neP;
  if (eq == 0 && g == 0) {  // k^2 = 0 has no v system (PF:184, 247)
    for (int idx = threadIdx.x; idx < a.ne * nP; idx += blockDim.x)
      out[(long)(idx / nP) * orow + idx % nP] = 0.0;
    return;
  }
  double* A = sm;  // n x n row-major
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
    const int i = idx / n, j = idx - i * n;
    const int l = 2 * j + par;
    double val;
    if (eq == 0) {
      if (i == 0) val = 1.0;
      else if (i == 1) val = (double)l * (double)l;
      else val = op_v(a, k2, i, l);                 // point p = i (rows 2..)
    } else {
      if (i == 0) val = 1.0;
      else val = op_w(a, k2, i, l);                 // point p = i (rows 1..)
    }
    A[idx] = val;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (int k = 0; k < n; ++k) {
    if (threadIdx.x < 64) {
      double best = -1.0;
      int bi = n;
      for (int i = k + lane; i < n; i += 64) {
        const double v = fabs(A[i * n + k]);
        if (v > best) { best = v; bi = i; }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
      }
      if (lane == 0) piv[k] = bi;
    }
    __syncthreads();
    const int pr = piv[k];
    if (pr != k) {
      for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const double t = A[k * n + j];
        A[k * n + j] = A[pr * n + j];
        A[pr * n + j] = t;
      }
      __syncthreads();
    }
    const double pinv = 1.0 / A[k * n + k];
    for (int i = threadIdx.x; i < n; i += blockDim.x) colk[i] = A[i * n + k];
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += blockDim.x) A[k * n + j] = (j == k ? 1.0 : A[k * n + j]) * pinv;
    __syncthreads();
    for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
      const int i = idx / n, j = idx - i * n;
      if (i == k) continue;
      const double base = (j == k) ? 0.0 : A[idx];
      A[idx] = base - colk[i] * A[k * n + j];
    }
    __syncthreads();
  }
  // undo the row interchanges as column interchanges, last first
  for (int k = n - 1; k >= 0; --k) {
    const int pr = piv[k];
    if (pr != k) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double t = A[i * n + k];
        A[i * n + k] = A[i * n + pr];
        A[i * n + pr] = t;
      }
    }
    __syncthreads();
  }
  // transposed store: row j, entry i = Ainv[i][j]; zero padding for i >= n and j >= n
  for (int idx = threadIdx.x; idx < a.ne * nP; idx += blockDim.x) {
    const int j = idx / nP, i = idx - j * nP;
    out[(long)j * orow + i] = (i < n && j < n) ? A[i * n + j] : 0.0;
  }
}

hipError_t launch_setup_inverses(Ctx& c) {
  const Dims& d = c.d;
  SetupArgs a{};
  a.t0 = c.t0;
  a.t2 = c.t2;
  a.t4 = c.t4;
  a.al = c.alpha;
  a.be = c.beta;
  a.inv = c.inv;
  a.inv_w = const_cast<double*>(c.inv_w);
  a.Mx = d.Mx;
  a.N = d.N;
  a.Ny = d.Ny;
  a.ne = d.ne;
  a.no = d.no;
  a.stride = d.inv_stride;
  a.row = d.inv_row;
  a.stride_w = c.stride_w;
  a.row_w = c.row_w;
  a.wdiag = c.wdiag ? 1 : 0;
  a.crep = c.crep;
  a.dt = c.prm.dt;
  a.re = c.prm.re;
  size_t shm = sizeof(double) * (size_t)d.ne * d.ne;
  if (shm > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)k_setup, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_setup, dim3(d.nclass, 4), dim3(256), shm, c.stream, a);
  return hipGetLastError();
}

// ------------------------------------------------------------------- solve

struct SolveArgs {
  const double2* rvT;              // v right-hand sides, [column][Rv rows] (even rows, then odd)
  const double2 *roE, *roO;        // omega_y right-hand sides, [row][column] per parity
  int Rv;                          // (ne - 2) + (no - 2)
  const double* inv;
  const double* inv_w;
  double2 *vst, *oy;  // v by group (record [l][side], see KinArgs), omega_y
  double2 *au0, *w0;
  const double2 *c1n, *c1p, *c3n, *c3p;  // mean Lamb point values, n and n-1
  const double *t0, *t2;
  int Mx, Mz, N, Ny, ne, no;
  long stride, row, stride_w, row_w, ncol;
  int wdiag;
  const int *bmat, *bg0, *bg1;  // per block: operator class, first and second group (-1: none)
  double dt, re;
};

// Block -> solve block in XCD-contiguous order: every XCD takes a contiguous
// run of blocks, so the neighbouring kx columns whose 16-byte values share a
// cache line are read and written through the same L2.
__device__ __forceinline__ int xcd_group(int b, int nb) { return (int)xcd_order(b, nb); }

// 16-byte load of inverse data that is read once per step: marked non-temporal
// so the 4.9 GB stream does not evict the right-hand sides and fields from L2.
__device__ __forceinline__ double2 stream_load(const double* p) {
  typedef double nt2 __attribute__((ext_vector_type(2)));
  const nt2 t = __builtin_nontemporal_load(reinterpret_cast<const nt2*>(p));
  return make_double2(t.x, t.y);
}

// One block per (operator class, up to two groups).  A group (kx, |kz|) holds
// the modes m0 = (kz, kx) and m1 = (Mz - kz, kx); groups with the same k^2
// share the operator, so the block streams the class's half-system inverses
// once, multiplies each by the 4 real right-hand-side vectors (Re, Im of both
// modes) of each of its groups, and writes v^{n+1} (and omega_y^{n+1} when that
// operator is not diagonalised) (PF:242-394).  The block holding group 0 is
// the mean mode (PF:401-474): U and W with the omega_y-type operator at k^2 = 0.
// Matrix columns per load batch in the streaming loop of k_solve; two batches
// alternate, so the next batch is in flight while the current one is used.
// Measured at 256x193x256: 6 (12 loads in flight per lane, 2 waves per SIMD)
// beats 4 (8 loads, 3 waves), 5 and 8.
constexpr int kSolveSU = 6;

// MEMBERS: the solve of several independent channels (an ensemble) in one
// launch.  Channel m owns the solve blocks [first[m], first[m + 1]) of the
// grid, in the same XCD-contiguous order as a launch of its own; the block
// size and LDS are the largest any channel needs.  Each block runs its
// channel's block exactly as a launch of that channel alone would, so the
// results are bit-identical to separate launches.
template <bool MEMBERS>
__global__ __launch_bounds__(512) void k_solve(SolveArgs a1, const SolveArgs* __restrict__ args,
                                               const int* __restrict__ first, int nm) {
  extern __shared__ double rb[];  // [slot][Ny][4 vec], slot = group * neq + eq
  constexpr int SU = kSolveSU;
  int b = xcd_group(blockIdx.x, gridDim.x);
  int m = 0;
  if constexpr (MEMBERS) {
    while (m + 1 < nm && b >= first[m + 1]) ++m;
    b -= first[m];
  }
  const SolveArgs& a = MEMBERS ? args[m] : a1;
  const int Ny = a.Ny, N = a.N, ne = a.ne, no = a.no;
  const int mat = a.bmat[b];
  const int gs0 = a.bg0[b], gs1 = a.bg1[b];
  const int ngr = gs1 >= 0 ? 2 : 1;
  const bool mean = (gs0 == 0);
  const int neq = a.wdiag ? 1 : 2;  // equations this kernel solves for a non-mean group
  const double c = a.dt * 0.5 / a.re;

  // ---- this thread's task: one row pair (2i, 2i+1) of one half system of one
  // equation (the launch has at least as many threads as tasks).  Its first SU
  // matrix columns are requested before the right-hand-side gather, so the
  // gather's latency overlaps the start of the stream.
  const int neP = ne + (ne & 1), noP = no + (no & 1);
  const int pe = neP / 2, po = noP / 2;  // row pairs per half system
  const int ntask = (mean ? 1 : neq) * (pe + po);
  const int t = threadIdx.x;
  const bool active = t < ntask;
  const int eq = mean ? 1 : t / (pe + po);
  const int rr = mean ? t : t - eq * (pe + po);
  const bool odd = rr >= pe;
  const int i2 = 2 * (odd ? rr - pe : rr);
  const int n = odd ? no : ne;
  const long nP = eq ? a.row_w : a.row;
  const double* M = (eq ? a.inv_w + (long)mat * a.stride_w : a.inv + (long)mat * a.stride) + (odd ? neP : 0) + i2;
  double2 ma[SU], mb[SU];
#pragma unroll
  for (int u = 0; u < SU; ++u) ma[u] = (active && u < n) ? stream_load(M + (long)u * nP) : make_double2(0, 0);

  // ---- right-hand sides, parity-folded (rows of the half systems)
  if (!mean) {
    for (int gi = 0; gi < ngr; ++gi) {
      const int g = gi ? gs1 : gs0;
      const int kx = g % a.Mx, kzab = g / a.Mx;
      const long col0 = (long)kzab * a.Mx + kx, col1 = (long)(kzab ? a.Mz - kzab : 0) * a.Mx + kx;
      for (int idx = threadIdx.x; idx < neq * Ny; idx += blockDim.x) {
        const int eq = idx / Ny, row = idx - eq * Ny;
        // row < ne: even system row; else odd system row (row - ne)
        const bool odd = row >= ne;
        const int i = odd ? row - ne : row;
        const int nbc = eq ? 1 : 2;
        double2 x0 = make_double2(0, 0), x1 = make_double2(0, 0);
        if (i >= nbc) {
          const int r = i - nbc;
          if (eq) {
            const double2* src = odd ? a.roO : a.roE;
            x0 = src[(long)r * a.ncol + col0];
            if (kzab > 0) x1 = src[(long)r * a.ncol + col1];
          } else {
            // consecutive rows are consecutive threads: one contiguous run per mode
            const int k = odd ? (ne - 2) + r : r;
            x0 = a.rvT[col0 * a.Rv + k];
            if (kzab > 0) x1 = a.rvT[col1 * a.Rv + k];
          }
        }
        double* dst = rb + ((gi * neq + eq) * Ny + row) * 4;
        dst[0] = x0.x;
        dst[1] = x0.y;
        dst[2] = x1.x;
        dst[3] = x1.y;
      }
    }
  } else {
    // mean mode: U (vectors 0,1) and W (vectors 2,3) with the omega_y-type
    // operator at k^2 = 0; forcing -dt/2 (3 L^n - L^{n-1}) (PF:403-404, 441-442)
    for (int row = threadIdx.x; row < Ny; row += blockDim.x) {
      const bool odd = row >= ne;
      const int i = odd ? row - ne : row;
      double acc[4] = {0, 0, 0, 0};
      if (i >= 1) {
        const int p = i, q = N - p;
        for (int l = odd ? 1 : 0; l <= N; l += 2) {
          const double t = a.t0[p * Ny + l] + c * a.t2[p * Ny + l];
          const double2 U = a.au0[l], W = a.w0[l];
          acc[0] += t * U.x; acc[1] += t * U.y; acc[2] += t * W.x; acc[3] += t * W.y;
        }
        const double2 f1p = a.c1n[p], f1q = a.c1n[q], g1p = a.c1p[p], g1q = a.c1p[q];
        const double2 f3p = a.c3n[p], f3q = a.c3n[q], g3p = a.c3p[p], g3q = a.c3p[q];
        const double hd = -1.0 * a.dt / 2.0;
        const double2 Fp = make_double2(hd * (3.0 * f1p.x - g1p.x), hd * (3.0 * f1p.y - g1p.y));
        const double2 Fq = make_double2(hd * (3.0 * f1q.x - g1q.x), hd * (3.0 * f1q.y - g1q.y));
        const double2 Gp = make_double2(hd * (3.0 * f3p.x - g3p.x), hd * (3.0 * f3p.y - g3p.y));
        const double2 Gq = make_double2(hd * (3.0 * f3q.x - g3q.x), hd * (3.0 * f3q.y - g3q.y));
        const double s = odd ? -1.0 : 1.0;
        if (p == q) {
          acc[0] += Fp.x; acc[1] += Fp.y; acc[2] += Gp.x; acc[3] += Gp.y;
        } else {
          acc[0] += 0.5 * (Fp.x + s * Fq.x); acc[1] += 0.5 * (Fp.y + s * Fq.y);
          acc[2] += 0.5 * (Gp.x + s * Gq.x); acc[3] += 0.5 * (Gp.y + s * Gq.y);
        }
      }
      for (int q4 = 0; q4 < 4; ++q4) rb[row * 4 + q4] = acc[q4];  // slot 0
    }
  }
  __syncthreads();
  // ---- dense half-system inverses x = Einv be, Oinv bo: 16-byte matrix loads,
  // accumulating the vectors of every group of the block.
  if (active) {
    const int ns = mean ? 1 : ngr;  // vector sets (slots) for this equation
    const double* rhs0 = rb + ((mean ? 0 : eq) * Ny + (odd ? ne : 0)) * 4;
    const double* rhs1 = rb + ((neq + eq) * Ny + (odd ? ne : 0)) * 4;
    double x[2][8] = {};
    // The right-hand sides are the same for every lane (LDS broadcasts); a
    // scheduling barrier per column keeps their reads next to their use, so
    // they hold registers for one column only and the load depth, not the
    // right-hand sides, fills the registers.
    auto column = [&](const double2 m, int jj) {
      __builtin_amdgcn_sched_barrier(0);
      const double* r0 = rhs0 + jj * 4;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        x[0][q] = fma(m.x, r0[q], x[0][q]);
        x[0][4 + q] = fma(m.y, r0[q], x[0][4 + q]);
      }
      if (ns > 1) {
        const double* r1 = rhs1 + jj * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          x[1][q] = fma(m.x, r1[q], x[1][q]);
          x[1][4 + q] = fma(m.y, r1[q], x[1][4 + q]);
        }
      }
    };
    // batches [j, j + SU) alternate between ma and mb; the first was requested
    // before the gather.  The next batch is always requested (its columns
    // clamped to n - 1), so the number of loads in flight is fixed and the
    // current batch waits for its own loads only.
    int j = 0;
    if (n < SU) {
#pragma unroll
      for (int u = 0; u < SU; ++u)
        if (u < n) column(ma[u], u);
      j = n;
    } else {
      while (true) {
#pragma unroll
        for (int u = 0; u < SU; ++u) mb[u] = stream_load(M + (long)min(j + SU + u, n - 1) * nP);
#pragma unroll
        for (int u = 0; u < SU; ++u) column(ma[u], j + u);
        j += SU;
        if (j + SU > n) break;
#pragma unroll
        for (int u = 0; u < SU; ++u) ma[u] = stream_load(M + (long)min(j + SU + u, n - 1) * nP);
#pragma unroll
        for (int u = 0; u < SU; ++u) column(mb[u], j + u);
        j += SU;
        if (j + SU > n) break;
      }
    }
    for (; j < n; ++j) {
      const double2 m = stream_load(M + (long)j * nP);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double r = rhs0[j * 4 + q];
        x[0][q] = fma(m.x, r, x[0][q]);
        x[0][4 + q] = fma(m.y, r, x[0][4 + q]);
        if (ns > 1) {
          const double r1 = rhs1[j * 4 + q];
          x[1][q] = fma(m.x, r1, x[1][q]);
          x[1][4 + q] = fma(m.y, r1, x[1][4 + q]);
        }
      }
    }
    for (int gi = 0; gi < ns; ++gi) {
      const int g = gi ? gs1 : gs0;
      const int kx = g % a.Mx, kzab = g / a.Mx;
      const long col0 = (long)kzab * a.Mx + kx, col1 = (long)(kzab ? a.Mz - kzab : 0) * a.Mx + kx;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = i2 + h;
        if (i >= n) break;
        const int l = 2 * i + (odd ? 1 : 0);
        const double* xx = x[gi];
        if (mean) {
          a.au0[l] = make_double2(xx[4 * h + 0], xx[4 * h + 1]);
          a.w0[l] = make_double2(xx[4 * h + 2], xx[4 * h + 3]);
        } else if (eq == 0) {
          // v: the group's record is written by this block alone, in full
          // lines (scattered 16-byte column stores cost a third of the solve)
          double2* rec = a.vst + (long)g * Ny * 2;
          rec[2 * l] = make_double2(xx[4 * h + 0], xx[4 * h + 1]);
          if (kzab > 0) rec[2 * l + 1] = make_double2(xx[4 * h + 2], xx[4 * h + 3]);
        } else {
          a.oy[(long)l * a.ncol + col0] = make_double2(xx[4 * h + 0], xx[4 * h + 1]);
          if (kzab > 0) a.oy[(long)l * a.ncol + col1] = make_double2(xx[4 * h + 2], xx[4 * h + 3]);
        }
      }
    }
  }
}

// ------------------------------------------------------------- kinematics

// getvelocityandvorticity (PF:2347-2407) for every mode with k^2 != 0:
//   u = (i alpha Dv - i beta omega_y)/k^2,  w = (i beta Dv + i alpha omega_y)/k^2,
//   omega_x = Dw - i beta v,                omega_z = -Du + i alpha v,
// and for the mean mode (PF:435-438, 469-472) u = U, w = W, omega_z = -DU,
// omega_x = DW.  D is the coefficient-space derivative of dz (PF:570-599),
// b_k = b_{k+2} + 2(k+1) a_{k+1}, b_0 halved.
//
// The recurrences split into two independent chains by degree parity: chain q
// reads v at degrees of parity q and omega_y at parity 1-q, and produces
// Dv, u, w at parity 1-q and Du, Dw, omega_x, omega_z at parity q.  One thread
// runs one chain of one mode, backwards over the degree, with its loads issued
// eight degrees ahead.  Its outputs are stored non-temporally: the next
// kernels read them, this one does not (the padded synthesis that reads u, w,
// omega_x, omega_z runs 5% faster, this pass 3% slower).
//
// The same pass forms the explicit-operator brackets of the next solve's
// right-hand sides (get_coef_rhs / get_rhs_omegay, PF:252-257, 335-340):
//   g_v = c D^4 v + (1 - 2 c k^2) D^2 v - k^2 (1 - c k^2) v   (at parity q)
//   g_w = c D^2 omega_y + (1 - c k^2) omega_y                  (at parity 1-q)
// whose chains read the same v and omega_y degrees as chain q.
struct KinArgs {
  const double2 *v, *oy, *au0, *w0;
  double2 *u, *w, *ox, *oz;
  float2 *uf, *wf, *oxf, *ozf;  // F32: u, w, omega_x, omega_z rounded to FP32, written instead
  double2 *gv, *go;  // brackets for the next solve
  const double *al, *be;
  int* nonfinite;    // set to 1 when v, omega_y, U or W hold a NaN or an infinity
  int Mx, Mz, N, Ny;
  long ncol;
  double c;          // dt / (2 re)
};

// u, w, omega_x, omega_z at one offset: FP64, or (F32) rounded to FP32
template <bool F32>
__device__ __forceinline__ void kin_store(const KinArgs& a, int which, long off, double2 v) {
  if constexpr (F32) {
    float2* p = which == 0 ? a.uf : which == 1 ? a.wf : which == 2 ? a.oxf : a.ozf;
    st_nt(p + off, make_float2((float)v.x, (float)v.y));
  } else {
    double2* p = which == 0 ? a.u : which == 1 ? a.w : which == 2 ? a.ox : a.oz;
    st_nt(p + off, v);
  }
}

template <bool F32>
__global__ __launch_bounds__(256) void k_kinematics(KinArgs a) {
  const int q = blockIdx.y;  // chain: parity of the v degrees it reads
  const int N = a.N;
  if (blockIdx.x == gridDim.x - 1) {
    // mean mode, one block: U, W staged in LDS, derivatives by lanes 0..3
    if (q != 0) return;
    extern __shared__ double sm[];
    double* x = sm;              // U.re U.im W.re W.im   [4][Ny]
    double* d = sm + 4 * a.Ny;   // DU.re DU.im DW.re DW.im
    bool finite = true;
    for (int k = threadIdx.x; k < a.Ny; k += blockDim.x) {
      const double2 U = a.au0[k], W = a.w0[k];
      x[0 * a.Ny + k] = U.x; x[1 * a.Ny + k] = U.y; x[2 * a.Ny + k] = W.x; x[3 * a.Ny + k] = W.y;
      finite = finite && isfinite(U.x) && isfinite(U.y) && isfinite(W.x) && isfinite(W.y);
    }
    if (!finite) a.nonfinite[0] = 1;
    __syncthreads();
    if (threadIdx.x < 4) {
      const double* xa = x + threadIdx.x * a.Ny;
      double* b = d + threadIdx.x * a.Ny;
      double b2 = 0.0, b1 = 0.0;
      b[N] = 0.0;
      for (int k = N - 1; k >= 0; --k) {
        const double bk = b2 + 2.0 * (k + 1) * xa[k + 1];
        b2 = b1;
        b1 = bk;
        b[k] = (k == 0) ? 0.5 * bk : bk;
      }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < a.Ny; k += blockDim.x) {
      const long off = (long)k * a.ncol;
      kin_store<F32>(a, 0, off, make_double2(x[0 * a.Ny + k], x[1 * a.Ny + k]));
      kin_store<F32>(a, 1, off, make_double2(x[2 * a.Ny + k], x[3 * a.Ny + k]));
      kin_store<F32>(a, 3, off, make_double2(-1.0 * d[0 * a.Ny + k], -1.0 * d[1 * a.Ny + k]));
      kin_store<F32>(a, 2, off, make_double2(d[2 * a.Ny + k], d[3 * a.Ny + k]));
    }
    return;
  }
  const long cm = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cm >= a.ncol || cm == 0) return;
  const int kz = (int)((unsigned)cm / (unsigned)a.Mx);
  const int kx = (int)(cm - (long)kz * a.Mx);
  if (kz == a.Mz / 2) return;  // span Nyquist: never solved, stays zero (PF:247)
  const double al = a.al[kx], be = a.be[kz];
  const double k2 = al * al + be * be;
  double2 vp = make_double2(0, 0);                    // v_{k+1}
  double2 dv2 = make_double2(0, 0);                   // Dv_{k+2} (unhalved)
  double2 up = make_double2(0, 0), wp = make_double2(0, 0);    // u_{k+1}, w_{k+1}
  double2 du2 = make_double2(0, 0), dw2 = make_double2(0, 0);  // Du_{k+2}, Dw_{k+2}
  // bracket chains: latest unhalved D^2 v, D^3 v, D^4 v, D omega, D^2 omega and omega_{k+1}
  double2 d2s = make_double2(0, 0), d3s = make_double2(0, 0), d4s = make_double2(0, 0);
  double2 e1s = make_double2(0, 0), e2s = make_double2(0, 0), wn = make_double2(0, 0);
  const double cc = a.c, c2 = 1.0 - 2.0 * cc * k2, c0 = k2 * (1.0 - cc * k2), cw = 1.0 - cc * k2;
  // x * 0 is 0 for every finite x and NaN for a NaN or an infinity: one
  // multiply-add per value, no branch between the batched loads
  double probe = 0.0;
  for (int k0 = N; k0 >= 0; k0 -= 8) {
    double2 in[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int k = k0 - s;
      if (k >= 0) in[s] = (((k & 1) == q) ? a.v : a.oy)[(long)k * a.ncol + cm];
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int k = k0 - s;
      if (k < 0) break;
      const long off = (long)k * a.ncol + cm;
      const double f = 2.0 * (k + 1);
      if ((k & 1) == q) {
        // Du_k, Dw_k from u_{k+1}, w_{k+1}; omega_x, omega_z with v_k
        double2 du = make_double2(du2.x + f * up.x, du2.y + f * up.y);
        double2 dw = make_double2(dw2.x + f * wp.x, dw2.y + f * wp.y);
        du2 = du;
        dw2 = dw;
        if (k == 0) {
          du.x *= 0.5; du.y *= 0.5; dw.x *= 0.5; dw.y *= 0.5;
        }
        const double2 vk = in[s];
        kin_store<F32>(a, 2, off, make_double2(dw.x + be * vk.y, dw.y - be * vk.x));
        kin_store<F32>(a, 3, off, make_double2(-du.x - al * vk.y, -du.y + al * vk.x));
        // brackets: D^2 v, D^4 v at k from D v, D^3 v at k+1; D omega at k from omega_{k+1}
        double2 q2 = make_double2(d2s.x + f * dv2.x, d2s.y + f * dv2.y);
        double2 q4 = make_double2(d4s.x + f * d3s.x, d4s.y + f * d3s.y);
        e1s = make_double2(e1s.x + f * wn.x, e1s.y + f * wn.y);
        d2s = q2;
        d4s = q4;
        if (k == 0) {
          q2.x *= 0.5; q2.y *= 0.5; q4.x *= 0.5; q4.y *= 0.5;
        }
        st_nt(&a.gv[off], make_double2(cc * q4.x + c2 * q2.x - c0 * vk.x, cc * q4.y + c2 * q2.y - c0 * vk.y));
        vp = vk;
      } else {
        // Dv_k from v_{k+1}; u_k, w_k with omega_y_k
        double2 dv = make_double2(dv2.x + f * vp.x, dv2.y + f * vp.y);
        dv2 = dv;
        if (k == 0) {
          dv.x *= 0.5; dv.y *= 0.5;
        }
        const double omr = in[s].x, omi = in[s].y;
        const double2 uk = make_double2(-1.0 * al * dv.y / k2 + be * omi / k2, al * dv.x / k2 - be * omr / k2);
        const double2 wk = make_double2(-1.0 * dv.y * be / k2 - 1.0 * omi * al / k2, be * dv.x / k2 + al * omr / k2);
        kin_store<F32>(a, 0, off, uk);
        kin_store<F32>(a, 1, off, wk);
        up = uk;
        wp = wk;
        // brackets: D^3 v at k from D^2 v at k+1; D^2 omega at k from D omega at k+1
        d3s = make_double2(d3s.x + f * d2s.x, d3s.y + f * d2s.y);
        double2 r2 = make_double2(e2s.x + f * e1s.x, e2s.y + f * e1s.y);
        e2s = r2;
        if (k == 0) {
          r2.x *= 0.5; r2.y *= 0.5;
        }
        st_nt(&a.go[off], make_double2(cc * r2.x + cw * omr, cc * r2.y + cw * omi));
        wn = in[s];
      }
      probe = fma(in[s].x, 0.0, fma(in[s].y, 0.0, probe));
    }
  }
  if (probe != probe) a.nonfinite[0] = 1;
}

// v from the solve's per-group records (record g = |kz| Mx + kx holds
// [l][side], side 0 for kz = |kz|, 1 for kz = Mz - |kz|) into the field layout
// [l][kz][kx].  Block = one |kz| and VL degrees: it reads the 256-byte record
// segments of its Mx groups whole and writes Mx-long rows of both kz.  The
// mean column and the span Nyquist row are not written.
constexpr int kVL = 8;

__global__ __launch_bounds__(256) void k_vtranspose(const double2* __restrict__ vst, double2* __restrict__ v,
                                                    int Mx, int Mz, int Ny, long ncol) {
  extern __shared__ double2 tile[];  // [VL][2][Mx]
  const int kzab = blockIdx.x, l0 = blockIdx.y * kVL;
  const int nl = min(kVL, Ny - l0);
  const int per = 2 * kVL;  // double2 per record segment
  for (int i = threadIdx.x; i < Mx * per; i += blockDim.x) {
    const int kx = i / per, e = i - kx * per;
    const int l = e >> 1, side = e & 1;
    if (l < nl) tile[(l * 2 + side) * Mx + kx] = vst[((long)(kzab * Mx + kx) * Ny + l0 + l) * 2 + side];
  }
  __syncthreads();
  const int nside = kzab ? 2 : 1;
  for (int i = threadIdx.x; i < nl * nside * Mx; i += blockDim.x) {
    const int kx = i % Mx, ls = i / Mx;
    const int l = ls / nside, side = ls - l * nside;
    const int kz = side ? Mz - kzab : kzab;
    if (kz == 0 && kx == 0) continue;  // the mean column: U, W are solved instead
    v[(long)(l0 + l) * ncol + (long)kz * Mx + kx] = tile[(l * 2 + side) * Mx + kx];
  }
}

static SolveArgs solve_args(const Ctx& c) {
  const Dims& d = c.d;
  SolveArgs a{};
  a.rvT = c.rvT;
  a.Rv = d.ne + d.no - 4;
  a.roE = c.roE;
  a.roO = c.roO;
  a.inv = c.inv;
  a.inv_w = c.inv_w;
  a.vst = c.vst;
  a.oy = c.oy;
  a.au0 = c.au0;
  a.w0 = c.w0;
  a.c1n = c.cl1[c.hcur];
  a.c1p = c.cl1[1 - c.hcur];
  a.c3n = c.cl3[c.hcur];
  a.c3p = c.cl3[1 - c.hcur];
  a.t0 = c.t0;
  a.t2 = c.t2;
  a.Mx = d.Mx;
  a.Mz = d.Mz;
  a.N = d.N;
  a.Ny = d.Ny;
  a.ne = d.ne;
  a.no = d.no;
  a.stride = d.inv_stride;
  a.row = d.inv_row;
  a.stride_w = c.stride_w;
  a.row_w = c.row_w;
  a.wdiag = c.wdiag ? 1 : 0;
  a.bmat = c.bmat;
  a.bg0 = c.bg0;
  a.bg1 = c.bg1;
  a.ncol = d.ncol;
  a.dt = c.prm.dt;
  a.re = c.prm.re;
  return a;
}

// LDS bytes (two groups per block, one or two equations each) and threads of
// one channel's solve blocks: one thread per task; with the omega_y operator
// diagonalised a block streams the two v half systems only (ne/2 + no/2 row
// pairs, 128 threads at N = 192), otherwise the omega_y half systems as well.
static void solve_shape(const Ctx& c, size_t* shm, int* threads) {
  const Dims& d = c.d;
  *shm = sizeof(double) * 4ull * d.Ny * 2 * (c.wdiag ? 1 : 2);
  const int pairs = (d.ne + (d.ne & 1)) / 2 + (d.no + (d.no & 1)) / 2;
  *threads = ((c.wdiag ? 1 : 2) * pairs + 63) / 64 * 64;
}

static hipError_t allow_shm(const void* kernel, size_t shm) {
  if (shm <= 65536) return hipSuccess;
  return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
}

hipError_t launch_vtranspose(Ctx& c) {
  const Dims& d = c.d;
  const size_t tshm = sizeof(double2) * 2ull * kVL * d.Mx;
  if (tshm > 65536) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_vtranspose, dim3(d.Mz / 2, (d.Ny + kVL - 1) / kVL), dim3(256), tshm, c.stream,
                     (const double2*)c.vst, c.v, d.Mx, d.Mz, d.Ny, d.ncol);
  return hipGetLastError();
}

hipError_t launch_solve(Ctx& c) {
  const SolveArgs a = solve_args(c);
  size_t shm;
  int threads;
  solve_shape(c, &shm, &threads);
  if (threads > 512) return hipErrorInvalidValue;
  hipError_t e = allow_shm((const void*)k_solve<false>, shm);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_solve<false>, dim3(c.d.nblocks), dim3(threads), shm, c.stream, a, nullptr, nullptr, 1);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_vtranspose(c);
}

size_t solve_members_bytes(int n) { return sizeof(SolveArgs) * (size_t)n + sizeof(int) * (size_t)(n + 1); }

hipError_t write_solve_members(Ctx* const* m, int n, void* dev) {
  std::vector<char> h(solve_members_bytes(n));
  SolveArgs* args = reinterpret_cast<SolveArgs*>(h.data());
  int* first = reinterpret_cast<int*>(h.data() + sizeof(SolveArgs) * (size_t)n);
  first[0] = 0;
  for (int i = 0; i < n; ++i) {
    args[i] = solve_args(*m[i]);
    first[i + 1] = first[i] + m[i]->d.nblocks;
  }
  return hipMemcpy(dev, h.data(), h.size(), hipMemcpyHostToDevice);
}

hipError_t launch_solve_members(hipStream_t st, Ctx* const* m, int n, const void* dev) {
  size_t shm = 0;
  int threads = 0;
  long blocks = 0;
  for (int i = 0; i < n; ++i) {
    size_t s;
    int t;
    solve_shape(*m[i], &s, &t);
    shm = s > shm ? s : shm;
    threads = t > threads ? t : threads;
    blocks += m[i]->d.nblocks;
  }
  if (threads > 512 || blocks > 0x7fffffffL) return hipErrorInvalidValue;
  hipError_t e = allow_shm((const void*)k_solve<true>, shm);
  if (e != hipSuccess) return e;
  const SolveArgs* args = static_cast<const SolveArgs*>(dev);
  const int* first = reinterpret_cast<const int*>(static_cast<const char*>(dev) + sizeof(SolveArgs) * (size_t)n);
  hipLaunchKernelGGL(k_solve<true>, dim3((unsigned)blocks), dim3(threads), shm, st, SolveArgs{}, args, first, n);
  return hipGetLastError();
}

// ---------------------------------------------- diagonalised omega_y solve
//
// For every mode but the mean, omega_y^{n+1} of parity p is
//   x = R diag(1 / (1 + c (k^2 - lam))) L b
// (L, R, lam from the set-up), i.e. two FP64 matrix-core products with
// matrices shared by all modes.  Block = one parity and 64 real columns (32
// complex modes, kx fastest): the right-hand sides b (the n-1 interior rows)
// are staged in LDS, Y = L B runs on v_mfma_f64_16x16x4f64 with the A
// fragments of L read from L2, Y is scaled per (row, mode) and written back
// to LDS as the B operand of X = R Y, and X lands in omega_y at degrees 2i + p.
// Eight wavefronts per block own the 16-row blocks, TPW each (wave w: row
// blocks w, w + 8, ...).  TPW = 1 whenever at most 8 row blocks exist (N up to
// 254): the kernel then needs 70 VGPRs, so three blocks fit per CU and its
// waves also fit beside the v solve's (measured 0.164 -> 0.143 ms against
// TPW = 2, which larger N take).
struct WArgs {
  const double *L0, *L1, *R0, *R1, *lam0, *lam1;
  const double *ro0, *ro1;  // parity-folded right-hand sides, real view [rows][2 ncol]
  double* oy;               // real view [Ny][2 ncol]
  const double *al, *be;
  int n0, n1, Mx, Mz;
  long ncold;
  double c;
};

template <int TPW>
__global__ __launch_bounds__(512) void k_wsolve(WArgs a) {
  using MF = Mf<double>;
  constexpr int CT = 64, NW = 8;  // columns, waves (<= 8 TPW row blocks)
  extern __shared__ double sB[];   // [rows][CT]
  const int par = blockIdx.y;
  const int n = par ? a.n1 : a.n0, m = n - 1;
  const double* L = par ? a.L1 : a.L0;
  const double* R = par ? a.R1 : a.R0;
  const double* lam = par ? a.lam1 : a.lam0;
  const double* ro = par ? a.ro1 : a.ro0;
  const int MB1 = (m + 15) / 16, KB = (m + 3) / 4, MB2 = (n + 15) / 16;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long col0 = (long)blockIdx.x * CT;
  // stage B: rows r < m, zero rows up to 16 MB1.  Each thread issues eight
  // loads before its first LDS store; the loads are unconditional (row and
  // column clamped into the array, the value zeroed afterwards), so no branch
  // splits them and they are in flight together.
  constexpr int SL = 8;
  const int nst = 16 * MB1 * CT;
  for (int i0 = threadIdx.x; i0 < nst; i0 += SL * blockDim.x) {
    double v[SL];
#pragma unroll
    for (int u = 0; u < SL; ++u) {
      const int idx = i0 + u * blockDim.x;
      const int r = idx / CT, cc = idx - r * CT;
      const long col = col0 + cc;
      const double x = ro[(long)min(r, m - 1) * a.ncold + min(col, a.ncold - 1)];
      v[u] = (idx < nst && r < m && col < a.ncold) ? x : 0.0;
    }
#pragma unroll
    for (int u = 0; u < SL; ++u) {
      const int idx = i0 + u * blockDim.x;
      if (idx < nst) sB[idx] = v[u];
    }
  }
  __syncthreads();
  MF::acc acc[TPW][4];
  // acc = A B over k blocks 0..KB-1 (A: packed fragments of L or R, MB row
  // blocks; B: the LDS tile).  The A fragments come from L2 in batches of four
  // k blocks, the next batch requested before the current one is used; the
  // batch indices are clamped so the loads are unconditional and the number in
  // flight is fixed.
  auto product = [&](const double* A, int MB) {
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[t][cb] = MF::acc{0, 0, 0, 0};
    double fa[TPW][4], fb[TPW][4];
    auto fetch = [&](double (&f)[TPW][4], int kb0) {
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        const int rb = min(wave + NW * t, MB - 1);
#pragma unroll
        for (int u = 0; u < 4; ++u) f[t][u] = A[((long)rb * KB + min(kb0 + u, KB - 1)) * 64 + lane];
      }
    };
    auto use = [&](const double (&f)[TPW][4], int kb0) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int kb = kb0 + u;
        if (kb >= KB) break;
        double bf[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) bf[cb] = sB[(4 * kb + (lane >> 4)) * CT + 16 * cb + (lane & 15)];
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
          if (wave + NW * t < MB) {
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) acc[t][cb] = MF::mma(f[t][u], bf[cb], acc[t][cb]);
          }
        }
      }
    };
    fetch(fa, 0);
    for (int kb0 = 0; kb0 < KB; kb0 += 8) {
      fetch(fb, kb0 + 4);
      use(fa, kb0);
      if (kb0 + 4 >= KB) break;
      fetch(fa, kb0 + 8);
      use(fb, kb0 + 4);
    }
  };
  product(L, MB1);
  __syncthreads();
  // scale rows by 1 / (1 + c (k^2 - lam)) and write Y back as the next B operand
  double s[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const long cm = (col0 + 16 * cb + (lane & 15)) >> 1;
    const int kz = (int)((unsigned)cm / (unsigned)a.Mx), kx = (int)(cm - (long)kz * a.Mx);
    const bool in = cm < (a.ncold >> 1);
    const double alx = in ? a.al[kx] : 0.0, bez = in ? a.be[kz] : 0.0;
    s[cb] = alx * alx + bez * bez;
  }
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int rb = wave + NW * t;
    if (rb >= MB1) continue;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = 16 * rb + MF::row(lane, e);
      const double lr = lam[row];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        const double y = row < m ? acc[t][cb][e] / (1.0 + a.c * (s[cb] - lr)) : 0.0;
        sB[row * CT + 16 * cb + (lane & 15)] = y;
      }
    }
  }
  __syncthreads();
  product(R, MB2);
  // store x_i -> omega_y at degree 2i + p; the mean mode and the span Nyquist
  // column are not solved here (PF:247)
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const long col = col0 + 16 * cb + (lane & 15);
    const long cm = col >> 1;
    if (col >= a.ncold || cm == 0) continue;
    const int kz = (int)((unsigned)cm / (unsigned)a.Mx);
    if (kz == a.Mz / 2) continue;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const int rb = wave + NW * t;
      if (rb >= MB2) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = 16 * rb + MF::row(lane, e);
        if (i < n) a.oy[(long)(2 * i + par) * a.ncold + col] = acc[t][cb][e];
      }
    }
  }
}

hipError_t launch_wsolve(Ctx& c) {
  const Dims& d = c.d;
  WArgs a{};
  a.L0 = c.wL[0];
  a.L1 = c.wL[1];
  a.R0 = c.wR[0];
  a.R1 = c.wR[1];
  a.lam0 = c.wlam[0];
  a.lam1 = c.wlam[1];
  a.ro0 = reinterpret_cast<const double*>(c.roE);
  a.ro1 = reinterpret_cast<const double*>(c.roO);
  a.oy = reinterpret_cast<double*>(c.oy);
  a.al = c.alpha;
  a.be = c.beta;
  a.n0 = d.ne;
  a.n1 = d.no;
  a.Mx = d.Mx;
  a.Mz = d.Mz;
  a.ncold = 2 * d.ncol;
  a.c = c.prm.dt * 0.5 / c.prm.re;
  const int m = d.ne - 1;
  const size_t shm = sizeof(double) * 64 * 16 * ((m + 15) / 16);
  if (shm > 65536) return hipErrorInvalidValue;
  const int mb = (d.ne + 15) / 16;  // row blocks of the larger (even) half system
  if (mb > 16) return hipErrorInvalidValue;
  const unsigned tiles = (unsigned)((a.ncold + 63) / 64);
  if (mb <= 8)
    hipLaunchKernelGGL(k_wsolve<1>, dim3(tiles, 2), dim3(512), shm, c.stream, a);
  else
    hipLaunchKernelGGL(k_wsolve<2>, dim3(tiles, 2), dim3(512), shm, c.stream, a);
  return hipGetLastError();
}

hipError_t launch_kinematics(Ctx& c, bool f32) {
  if (f32 && !c.uf) return hipErrorInvalidValue;
  const Dims& d = c.d;
  KinArgs a{};
  a.uf = c.uf;
  a.wf = c.wf;
  a.oxf = c.oxf;
  a.ozf = c.ozf;
  a.v = c.v;
  a.oy = c.oy;
  a.au0 = c.au0;
  a.w0 = c.w0;
  a.u = c.u;
  a.w = c.w;
  a.ox = c.ox;
  a.oz = c.oz;
  a.gv = c.gv;
  a.go = c.go;
  a.al = c.alpha;
  a.be = c.beta;
  a.nonfinite = c.nonfinite;
  a.Mx = d.Mx;
  a.Mz = d.Mz;
  a.N = d.N;
  a.Ny = d.Ny;
  a.ncol = d.ncol;
  a.c = c.prm.dt * 0.5 / c.prm.re;
  const unsigned nb = (unsigned)((d.ncol + kColumnThreads - 1) / kColumnThreads) + 1;  // + the mean-mode block
  const size_t shm = sizeof(double) * 8ull * d.Ny;
  const void* k = f32 ? (const void*)k_kinematics<true> : (const void*)k_kinematics<false>;
  if (shm > 65536) {
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
  }
  if (f32)
    hipLaunchKernelGGL(k_kinematics<true>, dim3(nb, 2), dim3(kColumnThreads), shm, c.stream, a);
  else
    hipLaunchKernelGGL(k_kinematics<false>, dim3(nb, 2), dim3(kColumnThreads), shm, c.stream, a);
  return hipGetLastError();
}

}  // namespace chf
