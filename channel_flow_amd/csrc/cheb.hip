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
#include "chf_internal.h"

namespace chf {

// Generic column kernel: each thread owns one real column and TM rows of
// NACC accumulators; the A rows of the block are staged in LDS and read as
// broadcasts.
template <class Prob, int TM>
__global__ __launch_bounds__(256) void k_cols(Prob pb) {
  constexpr int NA = Prob::NACC;
  extern __shared__ double sA[];
  Prob p = pb;
  p.select(blockIdx.z);
  const int K = p.K;
  const int r0 = blockIdx.y * TM;
  for (int i = threadIdx.x; i < NA * TM * K; i += blockDim.x) {
    const int m = i / (TM * K);
    const int rem = i - m * TM * K;
    const int r = rem / K;
    const int j = rem - r * K;
    sA[i] = p.a(m, r0 + r, j);
  }
  __syncthreads();
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= p.ncold) return;
  double acc[NA][TM];
#pragma unroll
  for (int m = 0; m < NA; ++m)
#pragma unroll
    for (int r = 0; r < TM; ++r) acc[m][r] = 0.0;
  for (int j = 0; j < K; ++j) {
    double b[NA];
    p.loadB(j, col, b);
#pragma unroll
    for (int m = 0; m < NA; ++m)
#pragma unroll
      for (int r = 0; r < TM; ++r) acc[m][r] = fma(sA[(m * TM + r) * K + j], b[m], acc[m][r]);
  }
#pragma unroll
  for (int r = 0; r < TM; ++r) {
    double v[NA];
#pragma unroll
    for (int m = 0; m < NA; ++m) v[m] = acc[m][r];
    p.store(r0 + r, col, v);
  }
}

template <class Prob, int TM>
static hipError_t run_cols(Ctx& c, const Prob& pb, int rows, int nz) {
  dim3 grid((unsigned)((pb.ncold + 255) / 256), (unsigned)((rows + TM - 1) / TM), nz);
  size_t shm = sizeof(double) * Prob::NACC * TM * pb.K;
  hipLaunchKernelGGL((k_cols<Prob, TM>), grid, dim3(256), shm, c.stream, pb);
  return hipGetLastError();
}

// ---------------------------------------------------------------- synthesis
// out[p] = E + O, out[P-1-p] = E - O for p < Ph, E/O = even/odd degree sums.
struct SynthProb {
  static constexpr int NACC = 2;
  const double* A;  // [Ph][Ny]
  const double* inf[6];
  double* outf[6];
  const double* in;
  double* out;
  int Ny, Ph, P, K;
  long ncold;
  __device__ void select(int z) {
    in = inf[z];
    out = outf[z];
  }
  __device__ double a(int m, int r, int j) const {
    const int l = 2 * j + m;
    return (r < Ph && l < Ny) ? A[r * Ny + l] : 0.0;
  }
  __device__ void loadB(int j, long col, double* b) const {
    const int l = 2 * j;
    b[0] = in[l * ncold + col];
    b[1] = (l + 1 < Ny) ? in[(l + 1) * ncold + col] : 0.0;
  }
  __device__ void store(int r, long col, const double* acc) const {
    if (r >= Ph) return;
    out[r * ncold + col] = acc[0] + acc[1];
    const int q = P - 1 - r;
    if (q != r) out[q * ncold + col] = acc[0] - acc[1];
  }
};

static hipError_t synth(Ctx& c, const double* A, int Ph, int P, const double2* const* in,
                        double2* const* out, int nf) {
  SynthProb pb{};
  pb.A = A;
  for (int f = 0; f < nf; ++f) {
    pb.inf[f] = reinterpret_cast<const double*>(in[f]);
    pb.outf[f] = reinterpret_cast<double*>(out[f]);
  }
  pb.Ny = c.d.Ny;
  pb.Ph = Ph;
  pb.P = P;
  pb.K = c.d.ne;
  pb.ncold = 2 * c.d.ncol;
  return run_cols<SynthProb, 16>(c, pb, Ph, nf);
}

hipError_t launch_synth_padded(Ctx& c, const double2* const* in, double2* const* out, int nf) {
  return synth(c, c.tsyn, c.d.Ph, c.d.Py, in, out, nf);
}

hipError_t launch_synth_points(Ctx& c, const double2* const* in, double2* const* out, int nf) {
  return synth(c, c.tpts, c.d.Pn, c.d.Ny, in, out, nf);
}

// ----------------------------------------------------------------- analysis
// coefficient 2j (2j+1) = sum over folded points of W[k][p] (x_p +- x_{P-1-p}).
struct AnalProb {
  static constexpr int NACC = 2;
  const double* A;  // [Ny][Ph]
  const double* inf[6];
  double* outf[6];
  const double* in;
  double* out;
  int Ny, Ph, P, K, N;
  long ncold;
  __device__ void select(int z) {
    in = inf[z];
    out = outf[z];
  }
  __device__ double a(int m, int r, int p) const {
    const int k = 2 * r + m;
    return (k <= N) ? A[k * Ph + p] : 0.0;
  }
  __device__ void loadB(int p, long col, double* b) const {
    const double x = in[p * ncold + col];
    const int q = P - 1 - p;
    if (q == p) {
      b[0] = x;
      b[1] = 0.0;
    } else {
      const double y = in[q * ncold + col];
      b[0] = x + y;
      b[1] = x - y;
    }
  }
  __device__ void store(int r, long col, const double* acc) const {
    const int k = 2 * r;
    if (k > N) return;
    out[k * ncold + col] = acc[0];
    if (k + 1 <= N) out[(k + 1) * ncold + col] = acc[1];
  }
};

hipError_t launch_analysis(Ctx& c, const double2* const* in, double2* const* out, int nf) {
  AnalProb pb{};
  pb.A = c.tana;
  for (int f = 0; f < nf; ++f) {
    pb.inf[f] = reinterpret_cast<const double*>(in[f]);
    pb.outf[f] = reinterpret_cast<double*>(out[f]);
  }
  pb.Ny = c.d.Ny;
  pb.Ph = c.d.Ph;
  pb.P = c.d.Py;
  pb.K = c.d.Ph;
  pb.N = c.d.N;
  pb.ncold = 2 * c.d.ncol;
  return run_cols<AnalProb, 16>(c, pb, c.d.ne, nf);
}

// ----------------------------------------------------------------------- rhs

__device__ __forceinline__ double col_k2(long col, int Mx, const double* al, const double* be) {
  const long cm = col >> 1;
  const int kz = (int)(cm / Mx);
  const int kx = (int)(cm - (long)kz * Mx);
  return al[kx] * al[kx] + be[kz] * be[kz];
}

// v equation rows at points p = 2 + r: c T4 + (1 - k^2 dt/re) T2 - k^2 (1 - k^2 c) T0
struct RhsVProb {
  static constexpr int NACC = 6;  // (E: T4, T2, T0), (O: T4, T2, T0)
  const double *t0, *t2, *t4;
  const double* vin;
  const double *hn, *hp;
  double *outE, *outO;
  const double *al, *be;
  int Ny, N, K, ne, no, Mx;
  long ncold;
  double dt, re;
  __device__ void select(int) {}
  __device__ double a(int m, int r, int j) const {
    const int par = m / 3, which = m - 3 * par;
    const int l = 2 * j + par;
    const int nrow = par ? no - 2 : ne - 2;
    if (r >= nrow || l > N) return 0.0;
    const int p = 2 + r;
    const double* T = which == 0 ? t4 : (which == 1 ? t2 : t0);
    return T[p * Ny + l];
  }
  __device__ void loadB(int j, long col, double* b) const {
    const double e = vin[(2 * j) * ncold + col];
    const double o = (2 * j + 1 <= N) ? vin[(2 * j + 1) * ncold + col] : 0.0;
    b[0] = b[1] = b[2] = e;
    b[3] = b[4] = b[5] = o;
  }
  __device__ void store(int r, long col, const double* acc) const {
    if (r >= ne - 2) return;  // ne >= no: no row of either half lies beyond
    const double k2 = col_k2(col, Mx, al, be);
    const double c = dt * 0.5 / re;
    const double c2 = 1.0 - k2 * dt / re;
    const double c0 = k2 * (1.0 - k2 * c);
    const int p = 2 + r, q = N - p;
    const double fp = 3.0 * dt / 2.0 * hn[p * ncold + col] - dt / 2.0 * hp[p * ncold + col];
    const double fq = 3.0 * dt / 2.0 * hn[q * ncold + col] - dt / 2.0 * hp[q * ncold + col];
    if (r < ne - 2) {
      const double sym = (p == q) ? fp : 0.5 * (fp + fq);
      outE[r * ncold + col] = c * acc[0] + c2 * acc[1] - c0 * acc[2] + sym;
    }
    if (r < no - 2) outO[r * ncold + col] = c * acc[3] + c2 * acc[4] - c0 * acc[5] + 0.5 * (fp - fq);
  }
};

// omega_y rows at points p = 1 + r: c T2 + (1 - k^2 c) T0
struct RhsOProb {
  static constexpr int NACC = 4;  // (E: T2, T0), (O: T2, T0)
  const double *t0, *t2;
  const double* oin;
  const double *hn, *hp;
  double *outE, *outO;
  const double *al, *be;
  int Ny, N, K, ne, no, Mx;
  long ncold;
  double dt, re;
  __device__ void select(int) {}
  __device__ double a(int m, int r, int j) const {
    const int par = m / 2, which = m - 2 * par;
    const int l = 2 * j + par;
    const int nrow = par ? no - 1 : ne - 1;
    if (r >= nrow || l > N) return 0.0;
    const int p = 1 + r;
    const double* T = which == 0 ? t2 : t0;
    return T[p * Ny + l];
  }
  __device__ void loadB(int j, long col, double* b) const {
    const double e = oin[(2 * j) * ncold + col];
    const double o = (2 * j + 1 <= N) ? oin[(2 * j + 1) * ncold + col] : 0.0;
    b[0] = b[1] = e;
    b[2] = b[3] = o;
  }
  __device__ void store(int r, long col, const double* acc) const {
    if (r >= ne - 1) return;
    const double k2 = col_k2(col, Mx, al, be);
    const double c = dt * 0.5 / re;
    const double c0 = 1.0 - k2 * c;
    const int p = 1 + r, q = N - p;
    const double fp = 3.0 * dt / 2.0 * hn[p * ncold + col] - dt / 2.0 * hp[p * ncold + col];
    const double fq = 3.0 * dt / 2.0 * hn[q * ncold + col] - dt / 2.0 * hp[q * ncold + col];
    if (r < ne - 1) {
      const double sym = (p == q) ? fp : 0.5 * (fp + fq);
      outE[r * ncold + col] = c * acc[0] + c0 * acc[1] + sym;
    }
    if (r < no - 1) outO[r * ncold + col] = c * acc[2] + c0 * acc[3] + 0.5 * (fp - fq);
  }
};

hipError_t launch_rhs(Ctx& c, int which) {
  const Dims& d = c.d;
  const int nxt = c.hcur;        // h^n
  const int prv = 1 - c.hcur;    // h^{n-1}
  if (which == 0) {
    RhsVProb pv{};
    pv.t0 = c.t0;
    pv.t2 = c.t2;
    pv.t4 = c.t4;
    pv.vin = reinterpret_cast<const double*>(c.v);
    pv.hn = reinterpret_cast<const double*>(c.hv[nxt]);
    pv.hp = reinterpret_cast<const double*>(c.hv[prv]);
    pv.outE = reinterpret_cast<double*>(c.rvE);
    pv.outO = reinterpret_cast<double*>(c.rvO);
    pv.al = c.alpha;
    pv.be = c.beta;
    pv.Ny = d.Ny;
    pv.N = d.N;
    pv.K = d.ne;
    pv.ne = d.ne;
    pv.no = d.no;
    pv.Mx = d.Mx;
    pv.ncold = 2 * d.ncol;
    pv.dt = c.prm.dt;
    pv.re = c.prm.re;
    return run_cols<RhsVProb, 8>(c, pv, d.ne - 2, 1);
  }
  RhsOProb po{};
  po.t0 = c.t0;
  po.t2 = c.t2;
  po.oin = reinterpret_cast<const double*>(c.oy);
  po.hn = reinterpret_cast<const double*>(c.hg[nxt]);
  po.hp = reinterpret_cast<const double*>(c.hg[prv]);
  po.outE = reinterpret_cast<double*>(c.roE);
  po.outO = reinterpret_cast<double*>(c.roO);
  po.al = c.alpha;
  po.be = c.beta;
  po.Ny = d.Ny;
  po.N = d.N;
  po.K = d.ne;
  po.ne = d.ne;
  po.no = d.no;
  po.Mx = d.Mx;
  po.ncold = 2 * d.ncol;
  po.dt = c.prm.dt;
  po.re = c.prm.re;
  return run_cols<RhsOProb, 8>(c, po, d.ne - 1, 1);
}

// ---------------------------------------------------------------- h assembly

struct HArgs {
  const double2 *lx, *ly, *lz;
  double2 *hv, *hg;
  const double *al, *be;
  int Mx, N;
  long ncol;
};

// One thread per complex column; the derivative runs the backward recurrence
// c_k b_k = b_{k+2} + 2 (k+1) a_{k+1}, which equals the reference's explicit
// sum b_k = (2/c_k) sum_{j>k, j+k odd} j a_j.
__global__ __launch_bounds__(256) void k_hassemble(HArgs a) {
  const long cm = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cm >= a.ncol) return;
  const int kz = (int)(cm / a.Mx);
  const int kx = (int)(cm - (long)kz * a.Mx);
  const double al = a.al[kx], be = a.be[kz];
  const double k2 = al * al + be * be;
  const int N = a.N;
  double2 bx2 = make_double2(0, 0), bx1 = make_double2(0, 0);
  double2 bz2 = make_double2(0, 0), bz1 = make_double2(0, 0);
  double2 lxn = make_double2(0, 0), lzn = make_double2(0, 0);  // a_{k+1}
  for (int k = N; k >= 0; --k) {
    const long off = (long)k * a.ncol + cm;
    const double2 lx = a.lx[off], ly = a.ly[off], lz = a.lz[off];
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

hipError_t launch_hassemble(Ctx& c) {
  HArgs a{};
  a.lx = c.Lc[0];
  a.ly = c.Lc[1];
  a.lz = c.Lc[2];
  a.hv = c.hvc;
  a.hg = c.hgc;
  a.al = c.alpha;
  a.be = c.beta;
  a.Mx = c.d.Mx;
  a.N = c.d.N;
  a.ncol = c.d.ncol;
  hipLaunchKernelGGL(k_hassemble, dim3((unsigned)((a.ncol + 255) / 256)), dim3(256), 0, c.stream, a);
  return hipGetLastError();
}

// Mean-mode Lamb (column (0,0) of L_x and L_z) -> values at the N+1 points
// (PF:1592-1609), written into the "new" history slot.
struct MArgs {
  const double2 *lx, *lz;
  double2 *o1, *o3;
  const double* tpts;  // [Pn][Ny]
  int N, Ny, Pn;
  long ncol;
};

__global__ void k_mean_points(MArgs a) {
  const int j = threadIdx.x;
  for (int jj = j; jj < a.Pn; jj += blockDim.x) {
    double2 e1 = make_double2(0, 0), o1 = make_double2(0, 0);
    double2 e3 = make_double2(0, 0), o3 = make_double2(0, 0);
    for (int l = 0; l <= a.N; ++l) {
      const double t = a.tpts[jj * a.Ny + l];
      const double2 x1 = a.lx[(long)l * a.ncol];
      const double2 x3 = a.lz[(long)l * a.ncol];
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

hipError_t launch_mean_points(Ctx& c) {
  MArgs a{};
  a.lx = c.Lc[0];
  a.lz = c.Lc[2];
  a.o1 = c.cl1[1 - c.hcur];
  a.o3 = c.cl3[1 - c.hcur];
  a.tpts = c.tpts;
  a.N = c.d.N;
  a.Ny = c.d.Ny;
  a.Pn = c.d.Pn;
  a.ncol = c.d.ncol;
  hipLaunchKernelGGL(k_mean_points, dim3(1), dim3(256), 0, c.stream, a);
  return hipGetLastError();
}

}  // namespace chf
