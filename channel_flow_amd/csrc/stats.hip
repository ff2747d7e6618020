// Unpadded physical fields and plane statistics, computed where the fields
// live.
//
//   physical: transform(1, .) (PF:948-961 with fourtrx_zcomp(+1), PF:774-860):
//             Chebyshev coefficients -> values at y_k = cos(pi k/N), span
//             modes -> nny points (Nyquist slot zeroed), streamwise modes ->
//             nnx real points (Hermitian fill, Re of the kx = 0 entry), i.e.
//             the reference's real*8 (0:nnx-1, 0:nny-1, 0:N) == [Ny][nny][nnx].
//   plane statistics: the plane means, covariances and extrema of the
//             physical u, v, w that statistics (PF:2816-3020) reduces.
#include "chf_internal.h"

namespace chf {

// One block per wall-normal point k of three physical fields [Ny][nny][nnx].
// out (12 x Ny, quantity-major): ua va wa uu vv ww uv uw vw max|u| max|v| max|w|
__global__ __launch_bounds__(256) void k_plane_stats(const double* __restrict__ u, const double* __restrict__ v,
                                                     const double* __restrict__ w, long plane, int Ny,
                                                     double* __restrict__ out) {
  __shared__ double red[6][256];
  const int k = blockIdx.x;
  const double* pu = u + (long)k * plane;
  const double* pv = v + (long)k * plane;
  const double* pw = w + (long)k * plane;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (long i = threadIdx.x; i < plane; i += blockDim.x) {
    const double a = pu[i], b = pv[i], c = pw[i];
    s[0] += a;
    s[1] += b;
    s[2] += c;
    s[3] = fmax(s[3], fabs(a));
    s[4] = fmax(s[4], fabs(b));
    s[5] = fmax(s[5], fabs(c));
  }
  for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = s[q];
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
      for (int q = 3; q < 6; ++q) red[q][threadIdx.x] = fmax(red[q][threadIdx.x], red[q][threadIdx.x + h]);
    }
    __syncthreads();
  }
  const double coef = 1.0 / (double)plane;
  const double ma = red[0][0] * coef, mb = red[1][0] * coef, mc = red[2][0] * coef;
  const double xa = red[3][0], xb = red[4][0], xc = red[5][0];
  __syncthreads();
  double t[6] = {0, 0, 0, 0, 0, 0};
  for (long i = threadIdx.x; i < plane; i += blockDim.x) {
    const double a = pu[i] - ma, b = pv[i] - mb, c = pw[i] - mc;
    t[0] += a * a;
    t[1] += b * b;
    t[2] += c * c;
    t[3] += a * b;
    t[4] += a * c;
    t[5] += b * c;
  }
  for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = t[q];
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0 * Ny + k] = ma;
    out[1 * Ny + k] = mb;
    out[2 * Ny + k] = mc;
    for (int q = 0; q < 6; ++q) out[(3 + q) * Ny + k] = red[q][0] * coef;
    out[9 * Ny + k] = xa;
    out[10 * Ny + k] = xb;
    out[11 * Ny + k] = xc;
  }
}

hipError_t launch_physical(Ctx& c, const double2* field, double2* work, double* out) {
  // Chebyshev synthesis at the N+1 points into the complex work plane (every
  // point of every column is written, so the work needs no clearing)
  const double2* in1[1] = {field};
  double2* o1[1] = {work};
  hipError_t e = launch_synth_points(c, in1, o1, 1);
  if (e != hipSuccess) return e;
  e = launch_span_inverse_unpadded(c, work);
  if (e != hipSuccess) return e;
  return launch_stream_c2r_unpadded(c, work, out);
}

hipError_t launch_plane_stats(Ctx& c, const double* u, const double* v, const double* w, double* out) {
  const Dims& d = c.d;
  hipLaunchKernelGGL(k_plane_stats, dim3(d.Ny), dim3(256), 0, c.stream, u, v, w, (long)d.Mz * 2 * d.Mx, d.Ny,
                     out);
  return hipGetLastError();
}

}  // namespace chf
