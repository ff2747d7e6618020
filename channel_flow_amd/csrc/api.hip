// C ABI of libchf: context life cycle, state exchange in the reference
// layout, and the step sequence of the reference time loop (PF:220-532).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <vector>

#include "chf_internal.h"
#include "eig.h"

using namespace chf;

namespace chf {
const char* const kKernelNames[K_COUNT] = {
    "synth_padded", "zfft_inv", "xpass", "zfft_fwd", "analysis", "hassemble",
    "synth_points", "mean_points", "rhs_v", "rhs_o", "solve", "rhs_prep", "kinematics", "wsolve"};
}

namespace {

thread_local std::string g_create_err;

// Work memory of chf_physical_noalloc: the complex [Ny][Mz][Mx] plane of the
// unpadded physical transform, in a file-scope device array so that the
// entry point allocates nothing.  Each GPU holds its own copy for the life of
// the process; host code reaches it only through the address that
// hipGetSymbolAddress returns for the device (never through the array's
// name, which on the host is a stand-in, not a device address).
constexpr size_t kScratchBytes = size_t(128) << 20;
__device__ __align__(256) double2 scratch[kScratchBytes / sizeof(double2)];
std::mutex g_scratch_mu;
std::map<int, double2*> g_scratch_of_device;

// The device's scratch address, looked up once per device with that device
// current (chf_create runs on the caller's current device).
hipError_t resolve_scratch(int device, double2** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_of_device.find(device);
  if (it == g_scratch_of_device.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    it = g_scratch_of_device.emplace(device, static_cast<double2*>(p)).first;
  }
  *out = it->second;
  return hipSuccess;
}

#define CHF_TRY(expr)                                                   \
  do {                                                                  \
    hipError_t e_ = (expr);                                             \
    if (e_ != hipSuccess) {                                             \
      c->err = std::string(#expr) + ": " + hipGetErrorString(e_);       \
      return (int)e_;                                                   \
    }                                                                   \
  } while (0)

int fail(Ctx* c, const std::string& msg) {
  if (c) c->err = msg;
  return -1;
}

// SCHEB (PF:3272-3342): T0, T1, T2, T4 at y_j = cos(pi j/N), rows j = points.
void scheb(int n, std::vector<double>& t0, std::vector<double>& t1, std::vector<double>& t2,
           std::vector<double>& t4) {
  const int m = n + 1;
  t0.assign((size_t)m * m, 0.0);
  t1.assign((size_t)m * m, 0.0);
  t2.assign((size_t)m * m, 0.0);
  t4.assign((size_t)m * m, 0.0);
  for (int l = 0; l <= n; ++l) {
    const double L = (double)l;
    const double sg = (l % 2 == 0) ? 1.0 : -1.0;
    t0[0 * m + l] = 1.0;
    t0[(size_t)n * m + l] = sg;
    t1[0 * m + l] = L * L;
    t1[(size_t)n * m + l] = -sg * L * L;
    t2[0 * m + l] = L * L * (L * L - 1.0) / 3.0;
    t2[(size_t)n * m + l] = sg * L * L * (L * L - 1.0) / 3.0;
  }
  const double h0 = 4.0 * std::atan(1.0) / (double)n;
  for (int j = 1; j < n; ++j) {
    const double q = h0 * j;
    const double c1 = std::cos(q), c2 = c1 * c1, s1 = std::sin(q), s2 = s1 * s1;
    const double ds1 = 1.0 / s1, ds3 = ds1 / s2, ds5 = ds3 / s2, ds7 = ds5 / s2;
    for (int l = 0; l <= n; ++l) {
      const double bj = (double)l;
      const double qj = q * bj;
      const double sj = std::sin(qj), cj = std::cos(qj);
      const double bs = bj * bj * s2, bc = bj * cj * s1;
      const size_t o = (size_t)j * m + l;
      t0[o] = cj;
      t1[o] = bj * sj * ds1;
      t2[o] = bj * (sj * c1 - bc) * ds3;
      t4[o] = bj * (sj * c1 * (9.0 + 6.0 * (c2 - bs)) - bc * (4.0 + 11.0 * c2 - bs)) * ds7;
    }
  }
}

// cos(pi a / b) with the argument reduced exactly in integers.
double cospi_frac(long a, long b) {
  const long per = 2 * b;
  long r = a % per;
  if (r < 0) r += per;
  return std::cos(M_PI * (double)r / (double)b);
}

template <class T>
hipError_t dalloc(Ctx* c, T** p, size_t n) {
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e == hipSuccess) c->bytes += (long long)(n * sizeof(T));
  return e;
}

int validate(const chf_params* p, std::string& msg) {
  if (!p) { msg = "null parameters"; return -1; }
  if (p->nnx < 4 || p->nnx % 4) { msg = "nnx must be a positive multiple of 4 (PF:721-765 padding)"; return -1; }
  if (p->nny < 4 || p->nny % 4) { msg = "nny must be a positive multiple of 4 (PF:721-765 padding)"; return -1; }
  if (p->nnz < 4 || p->nnz > 256) { msg = "nnz must lie in [4, 256]"; return -1; }
  if (!fft_length_ok(3 * p->nnx / 2) || !fft_length_ok(3 * p->nny / 2)) {
    msg = "3*nnx/2 and 3*nny/2 must be of the form 2^a 3^b";
    return -1;
  }
  if (!(p->aphi > 0) || !(p->beta > 0) || !(p->dt > 0) || !(p->re > 0)) {
    msg = "aphi, beta, dt and re must be positive";
    return -1;
  }
  if (p->precision != 0 && p->precision != 1) { msg = "precision must be 0 (FP64) or 1 (mixed)"; return -1; }
  if (p->precision == 1 && (!reg_fft_length32(3 * p->nnx / 2) || !reg_fft_length32(3 * p->nny / 2) ||
                            (p->nnx / 2) % 16)) {
    msg = "mixed precision needs 3*nnx/2 and 3*nny/2 in {12,24,48,96,144,192,384,768} and nnx % 32 == 0";
    return -1;
  }
  return 0;
}

void free_all(Ctx* c) {
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(c->alpha); f(c->beta);
  f(c->v); f(c->vst); f(c->oy); f(c->u); f(c->w); f(c->ox); f(c->oz);
  f(c->uf); f(c->wf); f(c->oxf); f(c->ozf);
  for (auto& l : c->Lcf) f(l);
  for (int i = 0; i < 2; ++i) { f(c->hv[i]); f(c->hg[i]); f(c->cl1[i]); f(c->cl3[i]); }
  f(c->au0); f(c->w0);
  for (auto& y : c->Y) f(y);
  for (auto& z : c->Z) f(z);
  for (auto& y : c->Yf) f(y);
  for (auto& z : c->Zf) f(z);
  f(c->costabf); f(c->twxf); f(c->twzf);
  for (auto& l : c->Lc) f(l);
  f(c->hvc); f(c->hgc);
  f(c->rvT); f(c->roE); f(c->roO);
  f(c->tsyn); f(c->tana); f(c->tpts); f(c->t0); f(c->t2); f(c->t4);
  f(c->inv); f(c->inv0); f(c->crep); f(c->bmat); f(c->bg0); f(c->bg1);
  for (int i = 0; i < 2; ++i) { f(c->wL[i]); f(c->wR[i]); f(c->wlam[i]); }
  f(c->phys_c); f(c->phys_r[0]); f(c->phys_r[1]); f(c->phys_r[2]); f(c->stats); f(c->twxu); f(c->twzu); f(c->twx); f(c->twz); f(c->twy); f(c->costab);
  f(c->gv); f(c->go); f(c->nonfinite);
  for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
  c->event_pool.clear();
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

// Matrix-core A fragments of a rows x cols row-major matrix:
// out[(rb * KB + kb) * 64 + lane] = A[16 rb + lane % 16][4 kb + lane / 16], zero outside.
std::vector<double> pack_fragments(const std::vector<double>& a, int rows, int cols) {
  const int MB = (rows + 15) / 16, KB = (cols + 3) / 4;
  std::vector<double> out((size_t)MB * KB * 64, 0.0);
  for (int rb = 0; rb < MB; ++rb)
    for (int kb = 0; kb < KB; ++kb)
      for (int l = 0; l < 64; ++l) {
        const int r = 16 * rb + l % 16, k = 4 * kb + l / 16;
        if (r < rows && k < cols) out[((size_t)rb * KB + kb) * 64 + l] = a[(size_t)r * cols + k];
      }
  return out;
}

// The omega_y half systems (get_coef_omegay, PF:3432-3455, folded by parity)
// are [1^T ; c (sigma T0_I - T2_I)] x = [0 ; b] with c = dt/(2 re) and
// sigma = 1/c + k^2: k^2 enters through sigma only.  On the null space of the
// wall row, x = Q z with Q = [-1^T ; I], the interior rows read
// c (sigma M0 - M2) z = b with M0 = T0_I Q, M2 = T2_I Q.  With
// M0^-1 M2 = V diag(lam) V^-1 (lam: the Dirichlet second-derivative
// eigenvalues, real and distinct),
//     x = (Q V) diag(1 / (1 + c (k^2 - lam))) (V^-1 M0^-1) b,
// so every mode shares two matrices and the per-mode work is a diagonal
// scale.  The decomposition is accepted only if it reproduces direct solves
// of the half system at k^2 = 0, 1 and the largest k^2 to 1e-12.
int omega_diagonalise(Ctx* c, const std::vector<double>& t0, const std::vector<double>& t2,
                      const std::vector<double>& al, const std::vector<double>& be) {
  const Dims& d = c->d;
  const int Ny = d.Ny;
  const double cc = c->prm.dt * 0.5 / c->prm.re;
  double smax = 0.0;
  for (double a : al)
    for (double b : be) smax = std::max(smax, a * a + b * b);
  std::vector<double> Lh[2], Rh[2], lam[2];
  for (int par = 0; par < 2; ++par) {
    const int n = par ? d.no : d.ne, m = n - 1;
    if (m < 1) return 0;
    la::Mat M0((size_t)m * m), M2((size_t)m * m);
    for (int r = 0; r < m; ++r) {
      const int p = r + 1;  // interior point of row r
      for (int k = 0; k < m; ++k) {
        const int l0 = par, lk = 2 * (k + 1) + par;  // Q column k: -1 at x_0, +1 at x_{k+1}
        M0[(size_t)r * m + k] = t0[(size_t)p * Ny + lk] - t0[(size_t)p * Ny + l0];
        M2[(size_t)r * m + k] = t2[(size_t)p * Ny + lk] - t2[(size_t)p * Ny + l0];
      }
    }
    la::Mat M0i, V, Vi;
    if (!la::inverse(M0, m, M0i)) return 0;
    const la::Mat G = la::matmul(M0i, M2, m, m, m);
    std::vector<double> ev;
    if (!la::eig_real(G, m, ev, V)) return 0;
    if (!la::inverse(V, m, Vi)) return 0;
    la::Mat L = la::matmul(Vi, M0i, m, m, m);
    la::Mat R((size_t)n * m);
    for (int k = 0; k < m; ++k) {
      double s = 0.0;
      for (int r = 0; r < m; ++r) s += V[(size_t)r * m + k];
      R[k] = -s;
      for (int r = 0; r < m; ++r) R[(size_t)(r + 1) * m + k] = V[(size_t)r * m + k];
    }
    // verification against direct solves of the half system
    for (double s : {0.0, 1.0, smax}) {
      la::Mat A((size_t)n * n);
      for (int j = 0; j < n; ++j) A[j] = 1.0;
      for (int i = 1; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          const size_t o = (size_t)i * Ny + 2 * j + par;
          A[(size_t)i * n + j] = -cc * t2[o] + (1.0 + cc * s) * t0[o];
        }
      std::vector<double> b(m), x(n, 0.0), y(m, 0.0);
      for (int r = 0; r < m; ++r) b[r] = std::sin(1.7 * r + 0.3) + 0.1 * r / m;
      la::Mat rhs(n, 0.0);
      for (int r = 0; r < m; ++r) rhs[r + 1] = b[r];
      std::vector<int> piv;
      if (!la::lu(A, piv, n)) return 0;
      la::lu_solve(A, piv, n, rhs, 1);
      for (int j = 0; j < m; ++j) {
        double t = 0.0;
        for (int r = 0; r < m; ++r) t += L[(size_t)j * m + r] * b[r];
        y[j] = t / (1.0 + cc * (s - ev[j]));
      }
      double err = 0.0, ref = 0.0;
      for (int i = 0; i < n; ++i) {
        double t = 0.0;
        for (int j = 0; j < m; ++j) t += R[(size_t)i * m + j] * y[j];
        err = std::max(err, std::fabs(t - rhs[i]));
        ref = std::max(ref, std::fabs(rhs[i]));
      }
      if (!(err <= 1e-12 * ref)) return 0;
    }
    Lh[par] = pack_fragments(L, m, m);
    Rh[par] = pack_fragments(R, n, m);
    lam[par] = ev;
    lam[par].resize((size_t)((m + 15) / 16) * 16, 0.0);
  }
  for (int par = 0; par < 2; ++par) {
    CHF_TRY(dalloc(c, &c->wL[par], Lh[par].size()));
    CHF_TRY(dalloc(c, &c->wR[par], Rh[par].size()));
    CHF_TRY(dalloc(c, &c->wlam[par], lam[par].size()));
    CHF_TRY(hipMemcpy(c->wL[par], Lh[par].data(), Lh[par].size() * sizeof(double), hipMemcpyHostToDevice));
    CHF_TRY(hipMemcpy(c->wR[par], Rh[par].data(), Rh[par].size() * sizeof(double), hipMemcpyHostToDevice));
    CHF_TRY(hipMemcpy(c->wlam[par], lam[par].data(), lam[par].size() * sizeof(double), hipMemcpyHostToDevice));
  }
  c->wdiag = true;
  return 0;
}

// Autotuner: for each cos-table contraction, times 1, 2 and 4 column blocks
// per wave on this context's own shapes and buffers (a warm-up launch, then the
// mean of three) and keeps the fastest.  The state is all zeros at this point
// and every choice computes bit-identical outputs, so tuning has no effect on
// results, only on speed.
int autotune(Ctx* c) {
  struct Stage {
    int id;
    std::function<hipError_t()> run;
  };
  std::vector<Stage> st;
  if (c->twy) {
    // the padded synthesis, the analysis, the point values and the
    // right-hand sides run as FFTs, which have no tile choice
  } else if (c->prm.precision == 0) {
    st.push_back({K_SYNTH_PADDED, [c] {
                    const double2* in6[6] = {c->u, c->v, c->w, c->ox, c->oy, c->oz};
                    return launch_synth_padded(*c, in6, c->Y, 6);
                  }});
    st.push_back({K_ANALYSIS, [c] {
                    const double2* y3[3] = {c->Y[0], c->Y[1], c->Y[2]};
                    return launch_analysis(*c, y3, c->Lc, 3);
                  }});
  } else {
    st.push_back({K_SYNTH_PADDED, [c] {
                    const double2* in6[6] = {c->u, c->v, c->w, c->ox, c->oy, c->oz};
                    return launch_synth_padded32(*c, in6, c->Yf, 6);
                  }});
    st.push_back({K_ANALYSIS, [c] {
                    const float2* y3[3] = {c->Yf[0], c->Yf[1], c->Yf[2]};
                    return launch_analysis32(*c, y3, c->Lc, 3);
                  }});
  }
  if (!c->twy)
    st.push_back({K_SYNTH_POINTS, [c] {
                    const double2* hc[2] = {c->hvc, c->hgc};
                    double2* ho[2] = {c->hv[0], c->hg[0]};
                    return launch_synth_points_h(*c, hc, ho, 2);
                  }});
  if (!c->twy) {
    st.push_back({K_RHS_V, [c] { return launch_rhs(*c, 0); }});
    st.push_back({K_RHS_O, [c] { return launch_rhs(*c, 1); }});
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CHF_TRY(hipEventCreate(&e0));
  CHF_TRY(hipEventCreate(&e1));
  const int cand[3] = {1, 2, 4};
  hipError_t err = hipSuccess;
  for (auto& s : st) {
    float best = 3.0e38f;
    int bestcb = 0;
    for (int i = 0; i < 3 && err == hipSuccess; ++i) {
      c->cb_choice[s.id] = cand[i];
      err = s.run();
      if (err == hipSuccess) err = hipEventRecord(e0, c->stream);
      for (int r = 0; r < 3 && err == hipSuccess; ++r) err = s.run();
      if (err == hipSuccess) err = hipEventRecord(e1, c->stream);
      if (err == hipSuccess) err = hipEventSynchronize(e1);
      float ms = 0.f;
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
      c->tune_ms[s.id][i] = ms / 3.0f;
      if (err == hipSuccess && ms < best) {
        best = ms;
        bestcb = cand[i];
      }
    }
    c->cb_choice[s.id] = bestcb;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  CHF_TRY(err);
  return 0;
}

int build(Ctx* c) {
  Dims& d = c->d;
  const chf_params& p = c->prm;
  d.Mx = p.nnx / 2;
  d.Mz = p.nny;
  d.N = p.nnz;
  d.Ny = p.nnz + 1;
  d.Lx = 3 * p.nnx / 2;
  d.Lz = 3 * p.nny / 2;
  d.Py = 2 * d.N + 1;
  d.Ph = d.N + 1;
  d.ne = d.N / 2 + 1;
  d.no = (d.N + 1) / 2;
  d.Pn = d.N / 2 + 1;
  d.ncol = (long)d.Mz * d.Mx;
  d.ngroups = d.Mx * (d.Mz / 2);

  CHF_TRY(hipGetDevice(&c->device));
  CHF_TRY(resolve_scratch(c->device, &c->scratch));
  int max_lds = 0;
  CHF_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
  {
    // largest LDS request: the set-up inversion of the even half system
    const long need = 8L * d.ne * d.ne + 2 * 160 * 8;
    if (need > max_lds) {
      c->err = "nnz too large for the LDS-resident operator inversion on this device";
      return -1;
    }
  }
  CHF_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  CHF_TRY(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
  CHF_TRY(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  CHF_TRY(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));

  const size_t fld = (size_t)d.Ny * d.ncol;
  CHF_TRY(dalloc(c, &c->alpha, d.Mx));
  CHF_TRY(dalloc(c, &c->beta, d.Mz));
  CHF_TRY(dalloc(c, &c->v, fld));
  CHF_TRY(dalloc(c, &c->vst, (size_t)d.ngroups * d.Ny * 2));
  CHF_TRY(dalloc(c, &c->oy, fld));
  CHF_TRY(dalloc(c, &c->u, fld));
  CHF_TRY(dalloc(c, &c->w, fld));
  CHF_TRY(dalloc(c, &c->ox, fld));
  CHF_TRY(dalloc(c, &c->oz, fld));
  for (int i = 0; i < 2; ++i) {
    CHF_TRY(dalloc(c, &c->hv[i], fld));
    CHF_TRY(dalloc(c, &c->hg[i], fld));
    CHF_TRY(dalloc(c, &c->cl1[i], d.Ny));
    CHF_TRY(dalloc(c, &c->cl3[i], d.Ny));
  }
  CHF_TRY(dalloc(c, &c->au0, d.Ny));
  CHF_TRY(dalloc(c, &c->w0, d.Ny));
  const size_t ysz = (size_t)d.Py * d.ncol;
  const size_t zsz = (size_t)d.Py * d.Lz * d.Mx;
  if (p.precision == 0) {
    for (auto& y : c->Y) CHF_TRY(dalloc(c, &y, ysz));
    for (auto& z : c->Z) CHF_TRY(dalloc(c, &z, zsz));
  } else {
    for (auto& y : c->Yf) CHF_TRY(dalloc(c, &y, ysz));
    for (auto& z : c->Zf) CHF_TRY(dalloc(c, &z, zsz));
  }
  for (auto& l : c->Lc) CHF_TRY(dalloc(c, &l, fld));
  CHF_TRY(dalloc(c, &c->hvc, fld));
  CHF_TRY(dalloc(c, &c->hgc, fld));
  CHF_TRY(dalloc(c, &c->rvT, (size_t)(d.ne + d.no - 4) * d.ncol));
  CHF_TRY(dalloc(c, &c->roE, (size_t)d.ne * d.ncol));
  CHF_TRY(dalloc(c, &c->roO, (size_t)d.ne * d.ncol));
  CHF_TRY(dalloc(c, &c->gv, fld));
  CHF_TRY(dalloc(c, &c->go, fld));
  CHF_TRY(dalloc(c, &c->nonfinite, 1));

  // wavenumbers (PF:123-134, id = 0)
  std::vector<double> al(d.Mx), be(d.Mz);
  for (int i = 0; i < d.Mx; ++i) al[i] = (double)i / p.aphi;
  for (int j = 0; j < d.Mz; ++j) be[j] = (double)(j <= d.Mz / 2 ? j : j - d.Mz) / p.beta;
  CHF_TRY(hipMemcpy(c->alpha, al.data(), d.Mx * sizeof(double), hipMemcpyHostToDevice));
  CHF_TRY(hipMemcpy(c->beta, be.data(), d.Mz * sizeof(double), hipMemcpyHostToDevice));

  // operator classes: the collocation operators depend on k^2 only, so groups
  // whose k^2 (computed as the set-up and solve kernels compute it) is bitwise
  // equal share one set of inverses; solve blocks take up to two groups of a
  // class, in order of their first group
  {
    std::map<double, int> cls_of;
    std::vector<std::vector<int>> members;
    for (int g = 0; g < d.ngroups; ++g) {
      const int kx = g % d.Mx, kz = g / d.Mx;
      const double k2 = al[kx] * al[kx] + be[kz] * be[kz];
      auto it = cls_of.find(k2);
      if (it == cls_of.end()) {
        cls_of.emplace(k2, (int)members.size());
        members.push_back({g});
      } else {
        members[it->second].push_back(g);
      }
    }
    std::vector<int> crep, bmat, bg0, bg1;
    for (size_t k = 0; k < members.size(); ++k) {
      crep.push_back(members[k][0]);
      for (size_t i = 0; i < members[k].size(); i += 2) {
        bmat.push_back((int)k);
        bg0.push_back(members[k][i]);
        bg1.push_back(i + 1 < members[k].size() ? members[k][i + 1] : -1);
      }
    }
    d.nclass = (int)crep.size();
    d.nblocks = (int)bmat.size();
    CHF_TRY(dalloc(c, &c->crep, crep.size()));
    CHF_TRY(dalloc(c, &c->bmat, bmat.size()));
    CHF_TRY(dalloc(c, &c->bg0, bg0.size()));
    CHF_TRY(dalloc(c, &c->bg1, bg1.size()));
    CHF_TRY(hipMemcpy(c->crep, crep.data(), crep.size() * sizeof(int), hipMemcpyHostToDevice));
    CHF_TRY(hipMemcpy(c->bmat, bmat.data(), bmat.size() * sizeof(int), hipMemcpyHostToDevice));
    CHF_TRY(hipMemcpy(c->bg0, bg0.data(), bg0.size() * sizeof(int), hipMemcpyHostToDevice));
    CHF_TRY(hipMemcpy(c->bg1, bg1.data(), bg1.size() * sizeof(int), hipMemcpyHostToDevice));
  }

  // SCHEB operator matrices
  std::vector<double> t0, t1, t2, t4;
  scheb(d.N, t0, t1, t2, t4);
  const size_t tt = (size_t)d.Ny * d.Ny;
  CHF_TRY(dalloc(c, &c->t0, tt));
  CHF_TRY(dalloc(c, &c->t2, tt));
  CHF_TRY(dalloc(c, &c->t4, tt));
  CHF_TRY(hipMemcpy(c->t0, t0.data(), tt * sizeof(double), hipMemcpyHostToDevice));
  CHF_TRY(hipMemcpy(c->t2, t2.data(), tt * sizeof(double), hipMemcpyHostToDevice));
  CHF_TRY(hipMemcpy(c->t4, t4.data(), tt * sizeof(double), hipMemcpyHostToDevice));

  // omega_y operator: diagonalised when the decomposition verifies, else its
  // half-system inverses are stored next to the v ones.  CHF_OMEGA_INVERSES=stored
  // in the environment selects the stored inverses without trying.
  {
    const char* force = std::getenv("CHF_OMEGA_INVERSES");
    const bool stored = force && std::strcmp(force, "stored") == 0;
    const int rc = stored ? 0 : omega_diagonalise(c, t0, t2, al, be);
    if (rc) return rc;
    const long neP = d.ne + (d.ne & 1), noP = d.no + (d.no & 1);
    if (c->wdiag) {
      d.inv_row = neP + noP;
      d.inv_stride = (long)d.ne * d.inv_row;
      CHF_TRY(dalloc(c, &c->inv, (size_t)d.nclass * d.inv_stride));
      CHF_TRY(dalloc(c, &c->inv0, (size_t)d.ne * (neP + noP)));
      c->inv_w = c->inv0;
      c->stride_w = 0;
      c->row_w = neP + noP;
    } else {
      d.inv_row = 2 * (neP + noP);
      d.inv_stride = (long)d.ne * d.inv_row;
      CHF_TRY(dalloc(c, &c->inv, (size_t)d.nclass * d.inv_stride));
      c->inv_w = c->inv + (neP + noP);
      c->stride_w = d.inv_stride;
      c->row_w = d.inv_row;
    }
  }

  // transform matrices: synthesis on the 2N+1 grid (folded rows p = 0..N),
  // analysis weights (PF:1696-1700: /2N, c_0 = 2, end points weighted once),
  // synthesis on the N+1 grid (folded rows j = 0..N/2)
  const int N = d.N;
  std::vector<double> ts((size_t)d.Ph * d.Ny), ta((size_t)d.Ny * d.Ph), tp((size_t)d.Pn * d.Ny);
  for (int pp = 0; pp < d.Ph; ++pp)
    for (int l = 0; l <= N; ++l) ts[(size_t)pp * d.Ny + l] = cospi_frac((long)pp * l, 2L * N);
  for (int k = 0; k <= N; ++k)
    for (int pp = 0; pp < d.Ph; ++pp) {
      const double ck = (k == 0) ? 2.0 : 1.0;
      const double wp = (pp == 0) ? 1.0 : 2.0;
      ta[(size_t)k * d.Ph + pp] = wp * cospi_frac((long)k * pp, 2L * N) / (2.0 * N * ck);
    }
  for (int j = 0; j < d.Pn; ++j)
    for (int l = 0; l <= N; ++l) tp[(size_t)j * d.Ny + l] = cospi_frac((long)j * l, (long)N);
  CHF_TRY(dalloc(c, &c->tsyn, ts.size()));
  CHF_TRY(dalloc(c, &c->tana, ta.size()));
  CHF_TRY(dalloc(c, &c->tpts, tp.size()));
  CHF_TRY(hipMemcpy(c->tsyn, ts.data(), ts.size() * sizeof(double), hipMemcpyHostToDevice));
  CHF_TRY(hipMemcpy(c->tana, ta.data(), ta.size() * sizeof(double), hipMemcpyHostToDevice));
  CHF_TRY(hipMemcpy(c->tpts, tp.data(), tp.size() * sizeof(double), hipMemcpyHostToDevice));

  // contraction operators packed for the matrix cores (even/odd degree split)
  {
    std::vector<double> ct(4 * (size_t)N);
    for (int n = 0; n < 4 * N; ++n) ct[n] = cospi_frac(n, 2L * N);
    CHF_TRY(dalloc(c, &c->costab, ct.size()));
    CHF_TRY(hipMemcpy(c->costab, ct.data(), ct.size() * sizeof(double), hipMemcpyHostToDevice));
    std::vector<float> cf(ct.begin(), ct.end());
    CHF_TRY(dalloc(c, &c->costabf, cf.size()));
    CHF_TRY(hipMemcpy(c->costabf, cf.data(), cf.size() * sizeof(float), hipMemcpyHostToDevice));
  }


  // FFT twiddles e^{-2 pi i m/L}
  auto twiddles = [](int L) {
    std::vector<double2> t(L);
    for (int m = 0; m < L; ++m)
      t[m] = make_double2(cospi_frac(2L * m, L), -std::sin(2.0 * M_PI * (double)m / (double)L));
    return t;
  };
  {
    auto tx = twiddles(d.Lx), tz = twiddles(d.Lz);
    CHF_TRY(dalloc(c, &c->twx, d.Lx));
    CHF_TRY(dalloc(c, &c->twz, d.Lz));
    CHF_TRY(hipMemcpy(c->twx, tx.data(), d.Lx * sizeof(double2), hipMemcpyHostToDevice));
    CHF_TRY(hipMemcpy(c->twz, tz.data(), d.Lz * sizeof(double2), hipMemcpyHostToDevice));
    std::vector<float2> txf(d.Lx), tzf(d.Lz);
    for (int m = 0; m < d.Lx; ++m) txf[m] = make_float2((float)tx[m].x, (float)tx[m].y);
    for (int m = 0; m < d.Lz; ++m) tzf[m] = make_float2((float)tz[m].x, (float)tz[m].y);
    CHF_TRY(dalloc(c, &c->twxf, d.Lx));
    CHF_TRY(dalloc(c, &c->twzf, d.Lz));
    CHF_TRY(hipMemcpy(c->twxf, txf.data(), d.Lx * sizeof(float2), hipMemcpyHostToDevice));
    CHF_TRY(hipMemcpy(c->twzf, tzf.data(), d.Lz * sizeof(float2), hipMemcpyHostToDevice));
  }

  // tables of the FFT forms of the wall-normal transforms (2N-point and
  // N-point twiddles, the type-III DCT pre-twiddles e^{i pi k/2N} =
  // cos(pi k/2N) + i cos(pi (N-k)/2N)); their blocks take 8 (FP64) or 16
  // (FP32) adjacent columns.  CHF_Y_TRANSFORMS=matrix in the environment keeps
  // the cos-table products.
  const char* ytr = std::getenv("CHF_Y_TRANSFORMS");
  const bool ymatrix = ytr && std::strcmp(ytr, "matrix") == 0;
  if (!ymatrix && ydct_length_supported(d.N) && d.ncol % 16 == 0) {
    const int N = d.N;
    std::vector<double2> ty = twiddles(2 * N), tn = twiddles(N);
    ty.insert(ty.end(), tn.begin(), tn.end());
    for (int k = 0; k < N; ++k) ty.push_back(make_double2(cospi_frac(k, 2L * N), cospi_frac(N - k, 2L * N)));
    CHF_TRY(dalloc(c, &c->twy, ty.size()));
    CHF_TRY(hipMemcpy(c->twy, ty.data(), ty.size() * sizeof(double2), hipMemcpyHostToDevice));
    if (c->prm.precision == 1) {
      // u, w, omega_x, omega_z as the padded FP32 synthesis reads them; the
      // span Nyquist column is never written and stays zero
      for (float2** q : {&c->uf, &c->wf, &c->oxf, &c->ozf}) {
        CHF_TRY(dalloc(c, q, fld));
        CHF_TRY(hipMemsetAsync(*q, 0, fld * sizeof(float2), c->stream));
      }
      for (auto& l : c->Lcf) {
        CHF_TRY(dalloc(c, &l, fld));
        CHF_TRY(hipMemsetAsync(l, 0, fld * sizeof(float2), c->stream));
      }
      c->f32_fields = true;
    }
  }

  // twiddles of the unpadded physical transform (lengths nnx and nny)
  {
    auto tx = twiddles(2 * d.Mx), tz = twiddles(d.Mz);
    CHF_TRY(dalloc(c, &c->twxu, tx.size()));
    CHF_TRY(dalloc(c, &c->twzu, tz.size()));
    CHF_TRY(hipMemcpy(c->twxu, tx.data(), tx.size() * sizeof(double2), hipMemcpyHostToDevice));
    CHF_TRY(hipMemcpy(c->twzu, tz.data(), tz.size() * sizeof(double2), hipMemcpyHostToDevice));
  }

  // zero the state
  for (double2* q : {c->v, c->oy, c->u, c->w, c->ox, c->oz, c->hv[0], c->hv[1], c->hg[0], c->hg[1], c->gv, c->go})
    CHF_TRY(hipMemsetAsync(q, 0, fld * sizeof(double2), c->stream));
  for (double2* q : {c->cl1[0], c->cl1[1], c->cl3[0], c->cl3[1], c->au0, c->w0})
    CHF_TRY(hipMemsetAsync(q, 0, d.Ny * sizeof(double2), c->stream));
  CHF_TRY(hipMemsetAsync(c->vst, 0, (size_t)d.ngroups * d.Ny * 2 * sizeof(double2), c->stream));
  CHF_TRY(hipMemsetAsync(c->nonfinite, 0, sizeof(int), c->stream));

  // kernels that stage more than 64 KiB of LDS
  CHF_TRY(launch_setup_inverses(*c));
  CHF_TRY(hipStreamSynchronize(c->stream));
  return autotune(c);
}

hipEvent_t next_event(Ctx* c) {
  if (c->pool_next == c->event_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->event_pool.push_back(e);
  }
  return c->event_pool[c->pool_next++];
}

// Runs one launcher; with timing on, brackets it with events on the stream.
template <class F>
hipError_t timed(Ctx* c, int id, F&& launch) {
  if (!c->timing) return launch();
  hipEvent_t a = next_event(c), b = next_event(c);
  if (!a || !b) return hipErrorOutOfMemory;
  hipError_t e = hipEventRecord(a, c->stream);
  if (e != hipSuccess) return e;
  e = launch();
  if (e != hipSuccess) return e;
  e = hipEventRecord(b, c->stream);
  if (e != hipSuccess) return e;
  c->marks.push_back({id, a, b});
  return hipSuccess;
}

// Forms h^n's point values that nonlinear() left to the next solve, with the
// FP64 point-value kernel in both forms: the values the right-hand-side
// kernels would have formed, bit for bit.
hipError_t form_h_points(Ctx* c) {
  if (!c->hdefer) return hipSuccess;
  const double2* hc[2] = {c->hvc, c->hgc};
  double2* ho[2] = {c->hv[c->hcur], c->hg[c->hcur]};
  const hipError_t e = launch_ypoints(*c, hc, ho, 2);
  if (e == hipSuccess) c->hdefer = false;
  return e;
}

// Runs a() on the context's stream and b() on the side stream, both after the
// work queued so far, and joins them: two field groups of one transform
// stage, so that each launch's tail overlaps the other's work.
template <class A, class B>
hipError_t on_two_streams(Ctx* c, A&& a, B&& b) {
  hipError_t e = hipEventRecord(c->ev_fork, c->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(c->side, c->ev_fork, 0);
  if (e == hipSuccess) e = a();
  if (e != hipSuccess) return e;
  hipStream_t main = c->stream;
  c->stream = c->side;  // the launchers use c->stream
  e = b();
  if (e == hipSuccess) e = hipEventRecord(c->ev_join, c->side);
  c->stream = main;
  if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, c->ev_join, 0);
  return e;
}

// Mixed form: the padded FP32 synthesis of fields f0 .. f0 + nf - 1 of
// (u, v, w, omega_x, omega_y, omega_z) into Yf.  u, w, omega_x, omega_z come
// from their FP32 copies when those are current: the values the FP64 inputs
// are rounded to on load, so the result is the same either way.
hipError_t synth32_fields(Ctx* c, const double2* const* in6, int f0, int nf) {
  if (!c->f32_current) return launch_synth_padded32(*c, in6 + f0, c->Yf + f0, nf);
  const float2* f32src[6] = {c->uf, nullptr, c->wf, c->oxf, nullptr, c->ozf};
  const float2* fin[6];
  const double2* din[6];
  float2* fout[6];
  float2* dout[6];
  int nfin = 0, ndin = 0;
  for (int f = f0; f < f0 + nf; ++f) {
    if (f32src[f]) {
      fin[nfin] = f32src[f];
      fout[nfin++] = c->Yf[f];
    } else {
      din[ndin] = in6[f];
      dout[ndin++] = c->Yf[f];
    }
  }
  hipError_t e = nfin ? launch_synth_padded32f(*c, fin, fout, nfin) : hipSuccess;
  if (e == hipSuccess && ndin) e = launch_synth_padded32(*c, din, dout, ndin);
  return e;
}

// Mixed form: the FP32 analysis of L_x, L_y, L_z (f0 .. f0 + nf - 1) into
// Lcf (FP32, with the FFT transforms: the same values) or Lc, whichever the
// h assembly reads.
hipError_t analysis32_fields(Ctx* c, const float2* const* y3, int f0, int nf) {
  if (c->f32_fields) return launch_analysis32f(*c, y3 + f0, c->Lcf + f0, nf);
  return launch_analysis32(*c, y3 + f0, c->Lc + f0, nf);
}

// defer: with the FFT transforms, leave h^n's point values to the next
// solve's right-hand-side kernels (the ensemble's graphs form them here).
int nonlinear(Ctx* c, bool defer = true) {
  CHF_TRY(form_h_points(c));  // the previous step's, before the h assembly overwrites hvc, hgc
  // six fields in the order the streamwise pass pairs them: (u,v) (w,ox) (oy,oz)
  const double2* in6[6] = {c->u, c->v, c->w, c->ox, c->oy, c->oz};
  if (!c->timing && c->overlap) {
    // synthesis and span FFTs of fields 0-2 and 3-5, then of L_x, L_y and
    // L_z, on the two streams (with timing on every launch runs alone)
    if (c->prm.precision == 0) {
      CHF_TRY(on_two_streams(
          c, [&] { const hipError_t e = launch_synth_padded(*c, in6, c->Y, 3); return e ? e : launch_zfft_inv(*c, c->Y, c->Z, 3); },
          [&] { const hipError_t e = launch_synth_padded(*c, in6 + 3, c->Y + 3, 3); return e ? e : launch_zfft_inv(*c, c->Y + 3, c->Z + 3, 3); }));
      CHF_TRY(launch_xpass(*c));
      const double2* y3[3] = {c->Y[0], c->Y[1], c->Y[2]};
      CHF_TRY(on_two_streams(
          c, [&] { const hipError_t e = launch_zfft_fwd(*c, c->Z, c->Y, 2); return e ? e : launch_analysis(*c, y3, c->Lc, 2); },
          [&] { const hipError_t e = launch_zfft_fwd(*c, c->Z + 2, c->Y + 2, 1); return e ? e : launch_analysis(*c, y3 + 2, c->Lc + 2, 1); }));
    } else {
      const float2* y6[6] = {c->Yf[0], c->Yf[1], c->Yf[2], c->Yf[3], c->Yf[4], c->Yf[5]};
      const float2* z3[3] = {c->Zf[0], c->Zf[1], c->Zf[2]};
      const float2* y3[3] = {c->Yf[0], c->Yf[1], c->Yf[2]};
      CHF_TRY(on_two_streams(
          c, [&] { const hipError_t e = synth32_fields(c, in6, 0, 3); return e ? e : launch_zfft32(*c, y6, c->Zf, 3, true); },
          [&] { const hipError_t e = synth32_fields(c, in6, 3, 3); return e ? e : launch_zfft32(*c, y6 + 3, c->Zf + 3, 3, true); }));
      CHF_TRY(launch_xpass32(*c));
      CHF_TRY(on_two_streams(
          c, [&] { const hipError_t e = launch_zfft32(*c, z3, c->Yf, 2, false); return e ? e : analysis32_fields(c, y3, 0, 2); },
          [&] { const hipError_t e = launch_zfft32(*c, z3 + 2, c->Yf + 2, 1, false); return e ? e : analysis32_fields(c, y3, 2, 1); }));
    }
  } else if (c->prm.precision == 0) {
    CHF_TRY(timed(c, K_SYNTH_PADDED, [&] { return launch_synth_padded(*c, in6, c->Y, 6); }));
    CHF_TRY(timed(c, K_ZFFT_INV, [&] { return launch_zfft_inv(*c, c->Y, c->Z, 6); }));
    CHF_TRY(timed(c, K_XPASS, [&] { return launch_xpass(*c); }));
    CHF_TRY(timed(c, K_ZFFT_FWD, [&] { return launch_zfft_fwd(*c, c->Z, c->Y, 3); }));
    const double2* y3[3] = {c->Y[0], c->Y[1], c->Y[2]};
    CHF_TRY(timed(c, K_ANALYSIS, [&] { return launch_analysis(*c, y3, c->Lc, 3); }));
  } else {
    // mixed form: FP32 wall-normal values, FP32 FFTs and products, FP64 coefficients out
    const float2* y6[6] = {c->Yf[0], c->Yf[1], c->Yf[2], c->Yf[3], c->Yf[4], c->Yf[5]};
    const float2* z3[3] = {c->Zf[0], c->Zf[1], c->Zf[2]};
    const float2* y3[3] = {c->Yf[0], c->Yf[1], c->Yf[2]};
    CHF_TRY(timed(c, K_SYNTH_PADDED, [&] { return synth32_fields(c, in6, 0, 6); }));
    CHF_TRY(timed(c, K_ZFFT_INV, [&] { return launch_zfft32(*c, y6, c->Zf, 6, true); }));
    CHF_TRY(timed(c, K_XPASS, [&] { return launch_xpass32(*c); }));
    CHF_TRY(timed(c, K_ZFFT_FWD, [&] { return launch_zfft32(*c, z3, c->Yf, 3, false); }));
    CHF_TRY(timed(c, K_ANALYSIS, [&] { return analysis32_fields(c, y3, 0, 3); }));
  }
  CHF_TRY(timed(c, K_HASSEMBLE, [&] { return launch_hassemble(*c); }));
  const int nxt = 1 - c->hcur;
  if (defer && c->twy) {
    c->hdefer = true;
  } else {
    const double2* hc[2] = {c->hvc, c->hgc};
    double2* ho[2] = {c->hv[nxt], c->hg[nxt]};
    CHF_TRY(timed(c, K_SYNTH_POINTS, [&] { return launch_synth_points_h(*c, hc, ho, 2); }));
  }
  c->hcur = nxt;
  return 0;
}

int solve(Ctx* c) {
  if (c->wdiag && !c->timing && c->overlap) {
    // the omega_y path (right-hand side, diagonalised solve) does not touch
    // the v path's data until the kinematics: it runs on the side stream,
    // the matrix-core-bound omega_y solve beside the HBM-bound v solve.
    // (With timing on, every launch runs alone so that each time is its own.)
    CHF_TRY(hipEventRecord(c->ev_fork, c->stream));
    CHF_TRY(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    hipStream_t main = c->stream;
    c->stream = c->side;  // the launchers and timing marks use c->stream
    hipError_t e = timed(c, K_RHS_O, [&] { return launch_rhs(*c, 1); });
    if (e == hipSuccess) e = timed(c, K_WSOLVE, [&] { return launch_wsolve(*c); });
    if (e == hipSuccess) e = hipEventRecord(c->ev_join, c->side);
    c->stream = main;
    CHF_TRY(e);
    CHF_TRY(timed(c, K_RHS_V, [&] { return launch_rhs(*c, 0); }));
    c->hdefer = false;  // both right-hand-side kernels have formed h^n's point values
    CHF_TRY(timed(c, K_SOLVE, [&] { return launch_solve(*c); }));
    CHF_TRY(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  } else {
    CHF_TRY(timed(c, K_RHS_V, [&] { return launch_rhs(*c, 0); }));
    CHF_TRY(timed(c, K_RHS_O, [&] { return launch_rhs(*c, 1); }));
    c->hdefer = false;
    CHF_TRY(timed(c, K_SOLVE, [&] { return launch_solve(*c); }));
    if (c->wdiag) CHF_TRY(timed(c, K_WSOLVE, [&] { return launch_wsolve(*c); }));
  }
  // the mixed form with the FFT transforms keeps u, w, omega_x, omega_z in
  // FP32 only, as the next nonlinear pass reads them
  CHF_TRY(timed(c, K_KINEMATICS, [&] { return launch_kinematics(*c, c->f32_fields); }));
  c->f32_current = c->f32_fields;
  c->f64_stale = c->f32_fields;
  return 0;
}

// solve() split around k_solve, for an ensemble, whose members' k_solve blocks
// run as one launch: the right-hand sides and the diagonalised omega_y solve
// before it (the omega_y path shares no data with k_solve), v's transpose and
// the kinematics after it.  Each kernel computes what it computes in solve().
int solve_before(Ctx* c) {
  CHF_TRY(timed(c, K_RHS_V, [&] { return launch_rhs(*c, 0); }));
  CHF_TRY(timed(c, K_RHS_O, [&] { return launch_rhs(*c, 1); }));
  c->hdefer = false;
  if (c->wdiag) CHF_TRY(timed(c, K_WSOLVE, [&] { return launch_wsolve(*c); }));
  return 0;
}

int solve_after(Ctx* c) {
  CHF_TRY(launch_vtranspose(*c));
  CHF_TRY(timed(c, K_KINEMATICS, [&] { return launch_kinematics(*c); }));
  c->f32_current = false;
  c->f64_stale = false;
  return 0;
}

// Waits for the context's work and reports a state that the last kinematics
// pass found to hold a NaN or an infinity (v, omega_y, U or W).
int check_finite(Ctx* c, const char* who) {
  int flag = 0;
  CHF_TRY(hipMemcpyAsync(&flag, c->nonfinite, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  CHF_TRY(hipStreamSynchronize(c->stream));
  if (flag) {
    c->err = std::string(who) + ": the state is not finite (NaN or infinity in v, omega_y, U or W)";
    return CHF_NONFINITE;
  }
  return 0;
}

// A new v, omega_y, U, W: clear the non-finite flag, then derive u, w,
// omega_x, omega_z and the next solve's right-hand-side brackets with the
// kinematics pass that follows every solve.
hipError_t load_kinematics(Ctx* c) {
  hipError_t e = hipMemsetAsync(c->nonfinite, 0, sizeof(int), c->stream);
  if (e != hipSuccess) return e;
  c->f32_current = false;
  c->f64_stale = false;
  return launch_kinematics(*c);
}

// u, w, omega_x, omega_z in FP64 for the calls that read them, when the step
// kept them in FP32 only: the same kinematics pass with FP64 outputs (it also
// rewrites the brackets with the values they hold).
hipError_t form_fields64(Ctx* c) {
  if (!c->f64_stale) return hipSuccess;
  const hipError_t e = launch_kinematics(*c);
  if (e == hipSuccess) c->f64_stale = false;
  return e;
}

}  // namespace

extern "C" {

int chf_create(const chf_params* prm, void** out) {
  if (!out) return -1;
  *out = nullptr;
  std::string msg;
  if (validate(prm, msg)) {
    g_create_err = msg;
    return -1;
  }
  Ctx* c = new Ctx();
  c->prm = *prm;
  int rc = build(c);
  if (rc) {
    g_create_err = c->err;
    free_all(c);
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

int chf_destroy(void* ctx) {
  if (!ctx) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
  return 0;
}

const char* chf_last_error(void* ctx) {
  if (!ctx) return g_create_err.c_str();
  return static_cast<Ctx*>(ctx)->err.c_str();
}

int chf_set_state(void* ctx, const chf_state* st) {
  if (!ctx || !st) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  const Dims& d = c->d;
  const size_t fb = (size_t)d.Ny * d.ncol * sizeof(double2), mb = d.Ny * sizeof(double2);
  const int n = c->hcur, pv = 1 - c->hcur;
  struct { double2* dst; const double* src; size_t bytes; } m[] = {
      {c->cl1[n], st->clamb1n, mb}, {c->cl1[pv], st->clamb1nm1, mb},
      {c->cl3[n], st->clamb3n, mb}, {c->cl3[pv], st->clamb3nm1, mb},
      {c->au0, st->au0, mb}, {c->w0, st->w0, mb},
      {c->v, st->vnp1, fb}, {c->oy, st->omegaynp1, fb},
      {c->hv[n], st->rhsvn, fb}, {c->hg[n], st->rhsomegayn, fb},
      {c->hv[pv], st->unp1, fb}, {c->hg[pv], st->wnp1, fb}};
  for (auto& e : m)
    if (!e.src) return fail(c, "chf_set_state: null array");
  for (auto& e : m) CHF_TRY(hipMemcpyAsync(e.dst, e.src, e.bytes, hipMemcpyHostToDevice, c->stream));
  c->hdefer = false;  // h^n's point values are the record's
  // u, w, omega_x, omega_z and the brackets of the loaded v, omega_y, U, W
  CHF_TRY(load_kinematics(c));
  return check_finite(c, "chf_set_state");
}

int chf_get_state(void* ctx, const chf_state* st) {
  if (!ctx || !st) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  const Dims& d = c->d;
  const size_t fb = (size_t)d.Ny * d.ncol * sizeof(double2), mb = d.Ny * sizeof(double2);
  const int n = c->hcur, pv = 1 - c->hcur;
  struct { const double2* src; double* dst; size_t bytes; } m[] = {
      {c->cl1[n], st->clamb1n, mb}, {c->cl1[pv], st->clamb1nm1, mb},
      {c->cl3[n], st->clamb3n, mb}, {c->cl3[pv], st->clamb3nm1, mb},
      {c->au0, st->au0, mb}, {c->w0, st->w0, mb},
      {c->v, st->vnp1, fb}, {c->oy, st->omegaynp1, fb},
      {c->hv[n], st->rhsvn, fb}, {c->hg[n], st->rhsomegayn, fb},
      {c->hv[pv], st->unp1, fb}, {c->hg[pv], st->wnp1, fb}};
  for (auto& e : m)
    if (!e.dst) return fail(c, "chf_get_state: null array");
  CHF_TRY(form_h_points(c));
  for (auto& e : m) CHF_TRY(hipMemcpyAsync(e.dst, e.src, e.bytes, hipMemcpyDeviceToHost, c->stream));
  CHF_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int chf_init_fields(void* ctx, const double* vnp1, const double* omegaynp1, const double* au0,
                    const double* w0) {
  if (!ctx) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!vnp1 || !omegaynp1 || !au0 || !w0) return fail(c, "chf_init_fields: null array");
  const Dims& d = c->d;
  const size_t fb = (size_t)d.Ny * d.ncol * sizeof(double2), mb = d.Ny * sizeof(double2);
  CHF_TRY(hipMemcpyAsync(c->v, vnp1, fb, hipMemcpyHostToDevice, c->stream));
  CHF_TRY(hipMemcpyAsync(c->oy, omegaynp1, fb, hipMemcpyHostToDevice, c->stream));
  CHF_TRY(hipMemcpyAsync(c->au0, au0, mb, hipMemcpyHostToDevice, c->stream));
  CHF_TRY(hipMemcpyAsync(c->w0, w0, mb, hipMemcpyHostToDevice, c->stream));
  // h^n = 0 and L^n = 0 so that the rotation in getnonlinear leaves h^{-1} = 0 (PF:2316)
  c->hdefer = false;
  for (double2* q : {c->hv[0], c->hv[1], c->hg[0], c->hg[1]})
    CHF_TRY(hipMemsetAsync(q, 0, fb, c->stream));
  for (double2* q : {c->cl1[0], c->cl1[1], c->cl3[0], c->cl3[1]})
    CHF_TRY(hipMemsetAsync(q, 0, mb, c->stream));
  // fields with all-zero u, w, omega_x, omega_z at the Nyquist column
  for (double2* q : {c->u, c->w, c->ox, c->oz, c->gv, c->go}) CHF_TRY(hipMemsetAsync(q, 0, fb, c->stream));
  CHF_TRY(load_kinematics(c));
  int rc = nonlinear(c);
  if (rc) return rc;
  return check_finite(c, "chf_init_fields");
}

int chf_nonlinear(void* ctx) {
  if (!ctx) return -1;
  return nonlinear(static_cast<Ctx*>(ctx));
}

int chf_solve(void* ctx) {
  if (!ctx) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  const int rc = solve(c);
  return rc ? rc : check_finite(c, "chf_solve");
}

int chf_step(void* ctx, int n) {
  if (!ctx) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  if (n < 0) return fail(c, "chf_step: negative step count");
  for (int i = 0; i < n; ++i) {
    int rc = solve(c);
    if (rc) return rc;
    rc = nonlinear(c);
    if (rc) return rc;
  }
  return check_finite(c, "chf_step");
}

int chf_sync(void* ctx) {
  if (!ctx) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  CHF_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int chf_get_fields(void* ctx, double* unp1, double* wnp1, double* omegaxnp1, double* omegaznp1) {
  if (!ctx) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  const size_t fb = (size_t)c->d.Ny * c->d.ncol * sizeof(double2);
  CHF_TRY(form_fields64(c));
  struct { const double2* src; double* dst; } m[] = {
      {c->u, unp1}, {c->w, wnp1}, {c->ox, omegaxnp1}, {c->oz, omegaznp1}};
  for (auto& e : m)
    if (e.dst) CHF_TRY(hipMemcpyAsync(e.dst, e.src, fb, hipMemcpyDeviceToHost, c->stream));
  CHF_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int chf_dims(void* ctx, int* mx, int* mz, int* ny) {
  if (!ctx) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  if (mx) *mx = c->d.Mx;
  if (mz) *mz = c->d.Mz;
  if (ny) *ny = c->d.Ny;
  return 0;
}

long long chf_device_bytes(void* ctx) {
  if (!ctx) return -1;
  return static_cast<Ctx*>(ctx)->bytes;
}

int chf_set_timing(void* ctx, int on) {
  if (!ctx) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  CHF_TRY(hipStreamSynchronize(c->stream));
  c->timing = on != 0;
  c->marks.clear();
  c->pool_next = 0;
  return 0;
}

}  // extern "C"

namespace {

// Buffers of the unpadded physical transform, allocated on first use (they
// are not part of the timestep).
int ensure_physical(Ctx* c) {
  if (c->stats) return 0;
  const Dims& d = c->d;
  const size_t plane = (size_t)d.Ny * d.Mz * 2 * d.Mx;
  CHF_TRY(dalloc(c, &c->phys_c, (size_t)d.Ny * d.ncol));
  for (auto& p : c->phys_r) CHF_TRY(dalloc(c, &p, plane));
  CHF_TRY(dalloc(c, &c->stats, 12 * (size_t)d.Ny));
  return 0;
}

const double2* field_ptr(Ctx* c, int field) {
  switch (field) {
    case 0: return c->u;
    case 1: return c->v;
    case 2: return c->w;
    case 3: return c->ox;
    case 4: return c->oy;
    case 5: return c->oz;
    default: return nullptr;
  }
}

}  // namespace

extern "C" {

int chf_physical(void* ctx, int field, double* out) {
  if (!ctx) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  const double2* f = field_ptr(c, field);
  if (!f) return fail(c, "chf_physical: field must be 0..5 (u, v, w, omega_x, omega_y, omega_z)");
  if (!out) return fail(c, "chf_physical: null output");
  const int rc = ensure_physical(c);
  if (rc) return rc;
  CHF_TRY(form_fields64(c));
  CHF_TRY(launch_physical(*c, f, c->phys_c, c->phys_r[0]));
  const size_t n = (size_t)c->d.Ny * c->d.Mz * 2 * c->d.Mx;
  CHF_TRY(hipMemcpyAsync(out, c->phys_r[0], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CHF_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int chf_physical_noalloc(void* ctx, int field, double* out) {
  if (!ctx) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  const double2* f = field_ptr(c, field);
  if (!f) return fail(c, "chf_physical_noalloc: field must be 0..5 (u, v, w, omega_x, omega_y, omega_z)");
  if (!out) return fail(c, "chf_physical_noalloc: null output");
  const size_t need = (size_t)c->d.Ny * c->d.ncol * sizeof(double2);
  if (need > kScratchBytes)
    return fail(c, "chf_physical_noalloc: the transform needs " + std::to_string(need) +
                       " bytes of work memory (Ny * nny * nnx/2 complex values), more than its " +
                       std::to_string(kScratchBytes) + "-byte scratch; use chf_physical");
  CHF_TRY(form_fields64(c));
  CHF_TRY(launch_physical(*c, f, c->scratch, out));
  return 0;
}

int chf_statistics(void* ctx, double* out) {
  if (!ctx) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!out) return fail(c, "chf_statistics: null output");
  const int rc = ensure_physical(c);
  if (rc) return rc;
  CHF_TRY(form_fields64(c));
  for (int q = 0; q < 3; ++q) CHF_TRY(launch_physical(*c, field_ptr(c, q), c->phys_c, c->phys_r[q]));
  CHF_TRY(launch_plane_stats(*c, c->phys_r[0], c->phys_r[1], c->phys_r[2], c->stats));
  CHF_TRY(hipMemcpyAsync(out, c->stats, 12 * (size_t)c->d.Ny * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  CHF_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int chf_tile_choice(void* ctx, int id, double* ms_by_choice) {
  if (!ctx) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  if (id < 0 || id >= K_COUNT) return fail(c, "chf_tile_choice: kernel id out of range");
  if (ms_by_choice)
    for (int i = 0; i < 3; ++i) ms_by_choice[i] = c->tune_ms[id][i];
  return c->cb_choice[id];
}

int chf_set_tile_choice(void* ctx, int id, int cb) {
  if (!ctx) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  if (id < 0 || id >= K_COUNT || c->cb_choice[id] <= 0)
    return fail(c, "chf_set_tile_choice: kernel class has no tile choice");
  if (cb != 1 && cb != 2 && cb != 4) return fail(c, "chf_set_tile_choice: choice must be 1, 2 or 4");
  c->cb_choice[id] = cb;
  return 0;
}

int chf_kernel_count(void) { return K_COUNT; }

const char* chf_kernel_name(int id) {
  if (id < 0 || id >= K_COUNT) return "";
  return kKernelNames[id];
}

int chf_timing(void* ctx, int id, double* total_ms, int* launches) {
  if (!ctx) return -1;
  Ctx* c = static_cast<Ctx*>(ctx);
  if (id < 0 || id >= K_COUNT) return fail(c, "chf_timing: kernel id out of range");
  CHF_TRY(hipStreamSynchronize(c->stream));
  double tot = 0.0;
  int n = 0;
  for (const auto& m : c->marks) {
    if (m.kernel != id) continue;
    float ms = 0.f;
    CHF_TRY(hipEventElapsedTime(&ms, m.a, m.b));
    tot += ms;
    ++n;
  }
  if (total_ms) *total_ms = tot;
  if (launches) *launches = n;
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ ensemble
//
// Independent channels of any mix of sizes, Reynolds numbers and precisions
// advanced together.  One step of all members is one hipGraph launch of three
// stages: every member's right-hand sides and omega_y solve as parallel
// branches (each member on its own stream, forked from and joined to the
// ensemble stream); one k_solve_members launch that streams every member's
// half-system inverses (PF:242-394 over all members); then every member's
// kinematics and nonlinear terms as parallel branches again.  Small members
// fill the GPU together.  The AB2 history buffers swap every step, so a graph
// (with the solve arguments it names) is captured per combination of the
// members' history phases (normally two) and cached.

namespace {

struct EnsGraph {
  std::vector<int> phase;         // members' history phases the graph was captured at
  hipGraphExec_t exec = nullptr;
  void* solve_args = nullptr;     // device: the members' k_solve arguments for that phase
};

struct Ens {
  std::vector<Ctx*> m;
  hipStream_t origin = nullptr;
  hipEvent_t fork = nullptr, mid = nullptr;
  std::vector<hipEvent_t> ready, join;
  std::vector<EnsGraph> graphs;
  std::string err;
};

int ens_fail(Ens* e, const std::string& msg, int rc = -1) {
  e->err = msg;
  return rc;
}

#define ENS_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return ens_fail(e, std::string(#expr) + ": " + hipGetErrorString(e_), (int)e_); \
  } while (0)

void ens_free(Ens* e) {
  for (auto& g : e->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.solve_args) (void)hipFree(g.solve_args);
  }
  e->graphs.clear();
  for (Ctx* c : e->m) {
    if (!c) continue;
    (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
  }
  e->m.clear();
  for (hipEvent_t ev : e->join) (void)hipEventDestroy(ev);
  e->join.clear();
  for (hipEvent_t ev : e->ready) (void)hipEventDestroy(ev);
  e->ready.clear();
  if (e->mid) (void)hipEventDestroy(e->mid);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->origin) (void)hipStreamDestroy(e->origin);
}

// Captures one step of every member into g (solve arguments written first,
// outside the capture); member host state (history phase) is restored
// afterwards, the caller advances it per launch.
int ens_capture(Ens* e, EnsGraph* g) {
  const int n = (int)e->m.size();
  for (Ctx* c : e->m) g->phase.push_back(c->hcur);
  ENS_TRY(hipMalloc(&g->solve_args, solve_members_bytes(n)));
  ENS_TRY(write_solve_members(e->m.data(), n, g->solve_args));
  ENS_TRY(hipStreamBeginCapture(e->origin, hipStreamCaptureModeRelaxed));
  int rc = 0;
  std::string what;
  hipError_t he = hipEventRecord(e->fork, e->origin);
  for (int i = 0; i < n && he == hipSuccess && rc == 0; ++i) {
    Ctx* c = e->m[i];
    he = hipStreamWaitEvent(c->stream, e->fork, 0);
    if (he != hipSuccess) break;
    rc = solve_before(c);
    if (rc) {
      what = c->err;
      break;
    }
    he = hipEventRecord(e->ready[i], c->stream);
    if (he == hipSuccess) he = hipStreamWaitEvent(e->origin, e->ready[i], 0);
  }
  if (he == hipSuccess && rc == 0) he = launch_solve_members(e->origin, e->m.data(), n, g->solve_args);
  if (he == hipSuccess && rc == 0) he = hipEventRecord(e->mid, e->origin);
  for (int i = 0; i < n && he == hipSuccess && rc == 0; ++i) {
    Ctx* c = e->m[i];
    he = hipStreamWaitEvent(c->stream, e->mid, 0);
    if (he != hipSuccess) break;
    rc = solve_after(c);
    if (rc == 0) rc = nonlinear(c, false);
    if (rc) {
      what = c->err;
      break;
    }
    he = hipEventRecord(e->join[i], c->stream);
    if (he == hipSuccess) he = hipStreamWaitEvent(e->origin, e->join[i], 0);
  }
  hipGraph_t graph = nullptr;
  const hipError_t ee = hipStreamEndCapture(e->origin, &graph);
  for (int i = 0; i < n; ++i) e->m[i]->hcur = g->phase[i];
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return ens_fail(e, "capture of a member step failed: " + what, rc);
  }
  if (he != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return ens_fail(e, std::string("capture: ") + hipGetErrorString(he), (int)he);
  }
  ENS_TRY(ee);
  hipGraphExec_t exec = nullptr;
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  ENS_TRY(ie);
  g->exec = exec;
  return 0;
}

}  // namespace

extern "C" {

int chf_ensemble_create(const chf_params* members, int n, void** out) {
  if (!out) return -1;
  *out = nullptr;
  if (!members || n <= 0) {
    g_create_err = "chf_ensemble_create: no members";
    return -1;
  }
  for (int i = 0; i < n; ++i) {
    std::string msg;
    if (validate(&members[i], msg)) {
      g_create_err = "member " + std::to_string(i) + ": " + msg;
      return -1;
    }
  }
  Ens* e = new Ens();
  for (int i = 0; i < n; ++i) {
    Ctx* c = new Ctx();
    c->prm = members[i];
    c->overlap = false;
    const int rc = build(c);
    if (rc) {
      g_create_err = "member " + std::to_string(i) + ": " + c->err;
      free_all(c);
      delete c;
      ens_free(e);
      delete e;
      return rc;
    }
    e->m.push_back(c);
  }
  hipError_t he = hipStreamCreateWithFlags(&e->origin, hipStreamNonBlocking);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&e->mid, hipEventDisableTiming);
  e->ready.assign(n, nullptr);
  e->join.assign(n, nullptr);
  for (int i = 0; i < n && he == hipSuccess; ++i) he = hipEventCreateWithFlags(&e->ready[i], hipEventDisableTiming);
  for (int i = 0; i < n && he == hipSuccess; ++i) he = hipEventCreateWithFlags(&e->join[i], hipEventDisableTiming);
  if (he != hipSuccess) {
    g_create_err = std::string("ensemble streams/events: ") + hipGetErrorString(he);
    ens_free(e);
    delete e;
    return (int)he;
  }
  *out = e;
  return 0;
}

int chf_ensemble_destroy(void* ens) {
  if (!ens) return -1;
  Ens* e = static_cast<Ens*>(ens);
  if (e->origin) (void)hipStreamSynchronize(e->origin);
  ens_free(e);
  delete e;
  return 0;
}

const char* chf_ensemble_last_error(void* ens) {
  if (!ens) return g_create_err.c_str();
  return static_cast<Ens*>(ens)->err.c_str();
}

int chf_ensemble_size(void* ens) {
  if (!ens) return -1;
  return (int)static_cast<Ens*>(ens)->m.size();
}

int chf_ensemble_member(void* ens, int i, void** ctx) {
  if (!ens || !ctx) return -1;
  Ens* e = static_cast<Ens*>(ens);
  if (i < 0 || i >= (int)e->m.size()) return ens_fail(e, "chf_ensemble_member: index out of range");
  *ctx = e->m[i];
  return 0;
}

int chf_ensemble_step(void* ens, int n) {
  if (!ens) return -1;
  Ens* e = static_cast<Ens*>(ens);
  if (n < 0) return ens_fail(e, "chf_ensemble_step: negative step count");
  for (Ctx* c : e->m) {
    if (c->timing) return ens_fail(e, "chf_ensemble_step: turn member timing off first");
    // the graphs form h^n's point values in nonlinear(): pending ones first
    ENS_TRY(form_h_points(c));
    // member work queued outside the graph (set_state, init) must finish first
    ENS_TRY(hipStreamSynchronize(c->stream));
  }
  for (int s = 0; s < n; ++s) {
    std::vector<int> phase;
    for (Ctx* c : e->m) phase.push_back(c->hcur);
    hipGraphExec_t exec = nullptr;
    for (auto& g : e->graphs)
      if (g.phase == phase) exec = g.exec;
    if (!exec) {
      EnsGraph g;
      const int rc = ens_capture(e, &g);
      if (rc) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.solve_args) (void)hipFree(g.solve_args);
        return rc;
      }
      exec = g.exec;
      e->graphs.push_back(g);
    }
    ENS_TRY(hipGraphLaunch(exec, e->origin));
    for (Ctx* c : e->m) c->hcur = 1 - c->hcur;  // what nonlinear() does per step
  }
  ENS_TRY(hipStreamSynchronize(e->origin));
  for (size_t i = 0; i < e->m.size(); ++i) {
    const int rc = check_finite(e->m[i], "chf_ensemble_step");
    if (rc == CHF_NONFINITE) return ens_fail(e, "member " + std::to_string(i) + ": " + e->m[i]->err, rc);
    if (rc) return ens_fail(e, e->m[i]->err, rc);
  }
  return 0;
}

}  // extern "C"
