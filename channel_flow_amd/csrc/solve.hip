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

namespace chf {

// ------------------------------------------------------------------ set-up

struct SetupArgs {
  const double *t0, *t2, *t4;
  const double *al, *be;
  double *invV, *invW;
  int Mx, N, Ny, ne, no;
  long stride;
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

// grid: (ngroups, 4) ; blockIdx.y = eq*2 + parity (eq 0 = v, 1 = omega_y)
__global__ __launch_bounds__(256) void k_setup(SetupArgs a) {
  extern __shared__ double sm[];
  __shared__ int piv[160];
  __shared__ double colk[160];
  const int g = blockIdx.x;
  const int eq = blockIdx.y >> 1, par = blockIdx.y & 1;
  const int kx = g % a.Mx, kz = g / a.Mx;
  const double k2 = a.al[kx] * a.al[kx] + a.be[kz] * a.be[kz];
  const int n = par ? a.no : a.ne;
  const int nP = n + (n & 1);  // even leading dimension: 16-byte aligned row pairs
  const int neP = a.ne + (a.ne & 1);
  double* out = (eq ? a.invW : a.invV) + (long)g * a.stride + (par ? (long)a.ne * neP : 0);
  if (eq == 0 && g == 0) {  // k^2 = 0 has no v system (PF:184, 247)
    for (int i = threadIdx.x; i < n * nP; i += blockDim.x) out[i] = 0.0;
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
  // transposed store: out[j*nP + i] = Ainv[i][j], zero in the padding row i = n
  for (int idx = threadIdx.x; idx < n * nP; idx += blockDim.x) {
    const int j = idx / nP, i = idx - j * nP;
    out[idx] = (i < n) ? A[i * n + j] : 0.0;
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
  a.invV = c.invV;
  a.invW = c.invW;
  a.Mx = d.Mx;
  a.N = d.N;
  a.Ny = d.Ny;
  a.ne = d.ne;
  a.no = d.no;
  a.stride = d.inv_stride;
  a.dt = c.prm.dt;
  a.re = c.prm.re;
  size_t shm = sizeof(double) * (size_t)d.ne * d.ne;
  if (shm > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)k_setup, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_setup, dim3(d.ngroups, 4), dim3(256), shm, c.stream, a);
  return hipGetLastError();
}

// ------------------------------------------------------------------- solve

struct SolveArgs {
  const double2 *rvE, *rvO, *roE, *roO;  // parity-folded right-hand sides
  const double *invV, *invW;
  double2 *v, *oy, *u, *w, *ox, *oz;
  double2 *au0, *w0;
  const double2 *c1n, *c1p, *c3n, *c3p;  // mean Lamb point values, n and n-1
  const double *t0, *t2;
  const double *al, *be;
  int Mx, Mz, N, Ny, ne, no;
  long stride, ncol;
  double dt, re;
  int mode;  // 0 = solve + kinematics, 1 = kinematics only
};

// Backward recurrence for the derivative of 4 real coefficient vectors held
// in LDS (x[q*Ny + k]) into d[q*Ny + k]; lanes 0..nv-1 each own one vector.
__device__ void lds_deriv(const double* x, double* d, int nv, int N, int Ny) {
  const int q = threadIdx.x;
  if (q < nv) {
    const double* a = x + q * Ny;
    double* b = d + q * Ny;
    double b2 = 0.0, b1 = 0.0;
    b[N] = 0.0;
    b1 = 0.0;
    for (int k = N - 1; k >= 0; --k) {
      const double bk = b2 + 2.0 * (k + 1) * a[k + 1];
      b2 = b1;
      b1 = bk;
      b[k] = (k == 0) ? 0.5 * bk : bk;
    }
  }
}

// One block per (kx, |kz|) group; modes m0 = (kz, kx), m1 = (Mz - kz, kx).
__global__ __launch_bounds__(256) void k_solve(SolveArgs a) {
  extern __shared__ double sm[];
  const int Ny = a.Ny, N = a.N, ne = a.ne, no = a.no;
  const int g = blockIdx.x;
  const int kx = g % a.Mx, kzab = g / a.Mx;
  const bool mean = (g == 0);
  const bool two = (kzab > 0) || mean;
  const int kz0 = kzab, kz1 = kzab ? a.Mz - kzab : 0;
  const long col0 = (long)kz0 * a.Mx + kx, col1 = (long)kz1 * a.Mx + kx;
  // LDS, 24 Ny doubles with aliasing by live range:
  //   region A (8 Ny): rhs [2 eq][4 vec][Ny], then U, W [4 vec][Ny] each
  //   region B (8 Ny): V, O (new coefficients)
  //   region C (4 Ny): DV, then DW ; region D (4 Ny): DU
  double* rb = sm;
  double* su = sm;
  double* sw = sm + 4 * Ny;
  double* sv = sm + 8 * Ny;
  double* so = sv + 4 * Ny;
  double* sd1 = sm + 16 * Ny;
  double* sd2 = sd1;
  double* sd3 = sm + 20 * Ny;
  const double c = a.dt * 0.5 / a.re;

  if (a.mode == 0) {
    // ---- right-hand sides, parity-folded (rows of the half systems)
    if (!mean) {
      for (int idx = threadIdx.x; idx < 2 * Ny; idx += blockDim.x) {
        const int eq = idx / Ny, row = idx - eq * Ny;
        // row < ne: even system row; else odd system row (row - ne)
        const bool odd = row >= ne;
        const int i = odd ? row - ne : row;
        const int nbc = eq ? 1 : 2;
        double2 x0 = make_double2(0, 0), x1 = make_double2(0, 0);
        if (i >= nbc) {
          const int r = i - nbc;
          const double2* src = eq ? (odd ? a.roO : a.roE) : (odd ? a.rvO : a.rvE);
          x0 = src[(long)r * a.ncol + col0];
          if (kzab > 0) x1 = src[(long)r * a.ncol + col1];
        }
        double* dst = rb + eq * 4 * Ny;
        dst[0 * Ny + row] = x0.x;
        dst[1 * Ny + row] = x0.y;
        dst[2 * Ny + row] = x1.x;
        dst[3 * Ny + row] = x1.y;
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
        double* dst = rb + 4 * Ny;
        for (int q4 = 0; q4 < 4; ++q4) dst[q4 * Ny + row] = acc[q4];
      }
    }
    __syncthreads();
    // ---- dense half-system inverses: x = Einv be, Oinv bo.  Thread = one row
    // pair (2i, 2i+1) of one of the four half systems; 16-byte matrix loads.
    const int neP = ne + (ne & 1), noP = no + (no & 1);
    const int pe = neP / 2, po = noP / 2;           // row pairs per half system
    for (int t = threadIdx.x; t < 2 * (pe + po); t += blockDim.x) {
      const int eq = t / (pe + po);
      if (mean && eq == 0) continue;
      const int rr = t - eq * (pe + po);
      const bool odd = rr >= pe;
      const int i2 = 2 * (odd ? rr - pe : rr);
      const int n = odd ? no : ne, nP = odd ? noP : neP;
      const double* M = (eq ? a.invW : a.invV) + (long)g * a.stride + (odd ? (long)ne * neP : 0) + i2;
      const double* rhs = rb + eq * 4 * Ny + (odd ? ne : 0);
      double x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      int j = 0;
      for (; j + 4 <= n; j += 4) {
        double2 m4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) m4[u] = *reinterpret_cast<const double2*>(M + (long)(j + u) * nP);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double r = rhs[q * Ny + j + u];
            x[q] = fma(m4[u].x, r, x[q]);
            x[4 + q] = fma(m4[u].y, r, x[4 + q]);
          }
        }
      }
      for (; j < n; ++j) {
        const double2 m = *reinterpret_cast<const double2*>(M + (long)j * nP);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double r = rhs[q * Ny + j];
          x[q] = fma(m.x, r, x[q]);
          x[4 + q] = fma(m.y, r, x[4 + q]);
        }
      }
      double* dst = eq ? so : sv;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = i2 + h;
        if (i >= n) break;
        const int l = 2 * i + (odd ? 1 : 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) dst[q * Ny + l] = x[4 * h + q];
      }
    }
    __syncthreads();
  } else {
    // kinematics only: fields from the current v, omega_y (or U, W for the mean)
    for (int k = threadIdx.x; k < Ny; k += blockDim.x) {
      if (mean) {
        const double2 U = a.au0[k], W = a.w0[k];
        so[0 * Ny + k] = U.x; so[1 * Ny + k] = U.y; so[2 * Ny + k] = W.x; so[3 * Ny + k] = W.y;
      } else {
        const double2 v0 = a.v[(long)k * a.ncol + col0];
        const double2 o0 = a.oy[(long)k * a.ncol + col0];
        double2 v1 = make_double2(0, 0), o1 = make_double2(0, 0);
        if (two) {
          v1 = a.v[(long)k * a.ncol + col1];
          o1 = a.oy[(long)k * a.ncol + col1];
        }
        sv[0 * Ny + k] = v0.x; sv[1 * Ny + k] = v0.y; sv[2 * Ny + k] = v1.x; sv[3 * Ny + k] = v1.y;
        so[0 * Ny + k] = o0.x; so[1 * Ny + k] = o0.y; so[2 * Ny + k] = o1.x; so[3 * Ny + k] = o1.y;
      }
    }
    __syncthreads();
  }

  if (mean) {
    // U in so[0..1], W in so[2..3]:  u00 = U, w00 = W, oz00 = -DU, ox00 = DW
    lds_deriv(so, sd1, 4, N, Ny);
    __syncthreads();
    for (int k = threadIdx.x; k < Ny; k += blockDim.x) {
      const double2 U = make_double2(so[0 * Ny + k], so[1 * Ny + k]);
      const double2 W = make_double2(so[2 * Ny + k], so[3 * Ny + k]);
      const long off = (long)k * a.ncol;
      if (a.mode == 0) {
        a.au0[k] = U;
        a.w0[k] = W;
      }
      a.u[off] = U;
      a.w[off] = W;
      a.oz[off] = make_double2(-1.0 * sd1[0 * Ny + k], -1.0 * sd1[1 * Ny + k]);
      a.ox[off] = make_double2(sd1[2 * Ny + k], sd1[3 * Ny + k]);
    }
    return;
  }

  // ---- store v, omega_y and run the kinematics (getvelocityandvorticity)
  const double al = a.al[kx];
  const double be0 = a.be[kz0], be1 = a.be[kz1];
  const double k2 = al * al + be0 * be0;
  if (a.mode == 0) {
    for (int k = threadIdx.x; k < Ny; k += blockDim.x) {
      const long off = (long)k * a.ncol;
      a.v[off + col0] = make_double2(sv[0 * Ny + k], sv[1 * Ny + k]);
      a.oy[off + col0] = make_double2(so[0 * Ny + k], so[1 * Ny + k]);
      if (kzab > 0) {
        a.v[off + col1] = make_double2(sv[2 * Ny + k], sv[3 * Ny + k]);
        a.oy[off + col1] = make_double2(so[2 * Ny + k], so[3 * Ny + k]);
      }
    }
  }
  lds_deriv(sv, sd1, 4, N, Ny);  // Dv
  __syncthreads();
  for (int k = threadIdx.x; k < Ny; k += blockDim.x) {
    for (int m = 0; m < 2; ++m) {
      const double be = m ? be1 : be0;
      const double dvr = sd1[(2 * m) * Ny + k], dvi = sd1[(2 * m + 1) * Ny + k];
      const double omr = so[(2 * m) * Ny + k], omi = so[(2 * m + 1) * Ny + k];
      // PF:2368-2375: u = (i alpha Dv - i beta omega)/k^2, w = (i beta Dv + i alpha omega)/k^2
      su[(2 * m + 1) * Ny + k] = al * dvr / k2 - be * omr / k2;
      su[(2 * m) * Ny + k] = -1.0 * al * dvi / k2 + be * omi / k2;
      sw[(2 * m + 1) * Ny + k] = be * dvr / k2 + al * omr / k2;
      sw[(2 * m) * Ny + k] = -1.0 * dvi * be / k2 - 1.0 * omi * al / k2;
    }
  }
  __syncthreads();
  if (threadIdx.x < 64) lds_deriv(sw, sd2, 4, N, Ny);
  else if (threadIdx.x >= 64 && threadIdx.x < 128) {
    // second wave: Du (lane index shifted)
    const int q = threadIdx.x - 64;
    if (q < 4) {
      const double* x = su + q * Ny;
      double* b = sd3 + q * Ny;
      double b2 = 0.0, b1 = 0.0;
      b[N] = 0.0;
      for (int k = N - 1; k >= 0; --k) {
        const double bk = b2 + 2.0 * (k + 1) * x[k + 1];
        b2 = b1;
        b1 = bk;
        b[k] = (k == 0) ? 0.5 * bk : bk;
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < Ny; k += blockDim.x) {
    const long off = (long)k * a.ncol;
    for (int m = 0; m < 2; ++m) {
      if (m == 1 && kzab == 0) break;
      const long col = m ? col1 : col0;
      const double be = m ? be1 : be0;
      const double vr = sv[(2 * m) * Ny + k], vi = sv[(2 * m + 1) * Ny + k];
      // PF:2388-2389: omega_x = Dw - i beta v ; PF:2401-2402: omega_z = -Du + i alpha v
      const double2 oxv = make_double2(sd2[(2 * m) * Ny + k] + be * vi, sd2[(2 * m + 1) * Ny + k] - be * vr);
      const double2 ozv = make_double2(-sd3[(2 * m) * Ny + k] - al * vi, -sd3[(2 * m + 1) * Ny + k] + al * vr);
      a.u[off + col] = make_double2(su[(2 * m) * Ny + k], su[(2 * m + 1) * Ny + k]);
      a.w[off + col] = make_double2(sw[(2 * m) * Ny + k], sw[(2 * m + 1) * Ny + k]);
      a.ox[off + col] = oxv;
      a.oz[off + col] = ozv;
    }
  }
}

hipError_t launch_solve(Ctx& c, int mode) {
  const Dims& d = c.d;
  SolveArgs a{};
  a.rvE = c.rvE;
  a.rvO = c.rvO;
  a.roE = c.roE;
  a.roO = c.roO;
  a.invV = c.invV;
  a.invW = c.invW;
  a.v = c.v;
  a.oy = c.oy;
  a.u = c.u;
  a.w = c.w;
  a.ox = c.ox;
  a.oz = c.oz;
  a.au0 = c.au0;
  a.w0 = c.w0;
  a.c1n = c.cl1[c.hcur];
  a.c1p = c.cl1[1 - c.hcur];
  a.c3n = c.cl3[c.hcur];
  a.c3p = c.cl3[1 - c.hcur];
  a.t0 = c.t0;
  a.t2 = c.t2;
  a.al = c.alpha;
  a.be = c.beta;
  a.Mx = d.Mx;
  a.Mz = d.Mz;
  a.N = d.N;
  a.Ny = d.Ny;
  a.ne = d.ne;
  a.no = d.no;
  a.stride = d.inv_stride;
  a.ncol = d.ncol;
  a.dt = c.prm.dt;
  a.re = c.prm.re;
  a.mode = mode;
  size_t shm = sizeof(double) * 24ull * d.Ny;
  if (shm > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)k_solve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_solve, dim3(d.ngroups), dim3(256), shm, c.stream, a);
  return hipGetLastError();
}

}  // namespace chf
