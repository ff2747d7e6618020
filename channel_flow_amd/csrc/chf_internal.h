// Internal declarations shared by the HIP translation units of libchf.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/chf.h"

namespace chf {

// Workgroups are dealt round-robin to the eight XCDs (block b to XCD b mod 8),
// each XCD with its own L2.  xcd_order(b, nb) is a permutation of 0..nb-1 that
// gives every XCD a contiguous run of logical indices, for any block count, so
// blocks that touch neighbouring data share an L2.
__device__ __forceinline__ long xcd_order(long b, long nb) {
  const long q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// Threads per block of the one-thread-per-column kernels (derivative sweeps, h
// assembly, kinematics): one wavefront, so that their 32768 columns at
// 256x193x256 spread over all 256 CUs (256-thread blocks left half of them idle).
constexpr int kColumnThreads = 64;

// Sizes derived from chf_params (reference symbols in brackets).
struct Dims {
  int Mx;     // streamwise modes kept, nnx/2 [nhx+1]
  int Mz;     // spanwise modes incl. Nyquist, nny [ny+1]
  int N;      // Chebyshev degree [nz]
  int Ny;     // N+1 coefficients / points
  int Lx;     // padded streamwise points 3 nnx/2 [plan9/10 length]
  int Lz;     // padded spanwise points 3 nny/2   [plan8/11 length]
  int Py;     // padded wall-normal points 2N+1   [plan13 grid, PF:1647-1667]
  int Ph;     // Py points folded by parity: N+1
  int ne;     // even Chebyshev degrees 0,2,..: N/2+1
  int no;     // odd degrees 1,3,..: (N+1)/2
  int Pn;     // N+1-point grid folded by parity: N/2+1
  long ncol;  // complex columns per plane: Mz*Mx
  int ngroups;  // (kx,|kz|) groups: Mx*(Mz/2)
  int nclass;   // operator classes (groups with bitwise-equal k^2 share an inverse)
  int nblocks;  // solve blocks: up to two groups of one class each
  long inv_row;     // doubles per inverse row j: 2 (neP + noP), the four half systems side by side
  long inv_stride;  // doubles per group: ne * inv_row
};

// Device-side view of the wall-normal matrices (all point-major, row = point).
struct Mats {
  const double* tsyn;   // [Ph][Ny]   T_l(y_p), y_p = cos(pi p/2N), p = 0..N
  const double* tana;   // [Ny][Ph]   analysis weights folded onto p = 0..N
  const double* tpts;   // [Pn][Ny]   T_l(y_j), y_j = cos(pi j/N), j = 0..N/2
  const double* t0;     // [Ny][Ny]   SCHEB T0 (operators)
  const double* t2;     // [Ny][Ny]   SCHEB T2
  const double* t4;     // [Ny][Ny]   SCHEB T4
};

// Shape of a contraction on FP64 matrix cores (16x16x4 blocks).
struct MfmaShape {
  int M, K;   // logical rows / reduction length
  int MB;     // row blocks of 16
  int KB;     // k blocks of 4
};

inline MfmaShape mfma_shape(int M, int K) {
  MfmaShape s{};
  s.M = M;
  s.K = K;
  s.MB = (M + 15) / 16;
  s.KB = (K + 3) / 4;
  return s;
}

// Kernel classes for per-launch timing.  K_MEAN_POINTS and K_RHS_PREP have no
// launch of their own: the mean-mode point values come from the last block of
// the h-assembly launch and are timed with it, and the right-hand-side
// brackets are formed by the kinematics pass (the ids are kept so that the
// class ids stay stable).
enum KernelId {
  K_SYNTH_PADDED = 0, K_ZFFT_INV, K_XPASS, K_ZFFT_FWD, K_ANALYSIS, K_HASSEMBLE, K_SYNTH_POINTS,
  K_MEAN_POINTS, K_RHS_V, K_RHS_O, K_SOLVE, K_RHS_PREP, K_KINEMATICS, K_WSOLVE, K_COUNT
};

struct Ctx {
  chf_params prm{};
  Dims d{};
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream of the context (default priority): with the omega_y operator
  // diagonalised, the omega_y right-hand side and solve run on it while the v
  // path runs on `stream`; the pair of events forks and joins it
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // off for ensemble members: their steps already run as parallel graph
  // branches, and the extra fork/join nodes cost more than they overlap
  bool overlap = true;
  // wavenumbers
  double* alpha = nullptr;  // [Mx]
  double* beta = nullptr;   // [Mz]
  // state (complex, [Ny][Mz][Mx] unless noted)
  double2 *v = nullptr, *oy = nullptr;              // coefficients
  double2* vst = nullptr;  // v as the solve writes it: per (kx, |kz|) group, [Ny][2 sides]
  double2 *u = nullptr, *w = nullptr, *ox = nullptr, *oz = nullptr;  // derived
  // Mixed form with the FFT transforms: u, w, omega_x, omega_z also in FP32,
  // the values the padded synthesis rounds them to on load.  The per-step
  // kinematics writes only these (f64_stale: u, w, ox, oz are not current and
  // are formed from v, omega_y, U, W when read); f32_current: they are.
  float2 *uf = nullptr, *wf = nullptr, *oxf = nullptr, *ozf = nullptr;
  // the same forms keep L^n's coefficients (the FP32 analysis's results, exact
  // in FP32) in FP32 between the analysis and the h assembly
  float2* Lcf[3] = {};
  bool f32_fields = false, f32_current = false, f64_stale = false;
  double2* hv[2] = {nullptr, nullptr};  // h_v point values; [hcur] = h^n
  double2* hg[2] = {nullptr, nullptr};  // h_g point values
  double2* cl1[2] = {nullptr, nullptr};  // mean L_x(0,0) point values [Ny]
  double2* cl3[2] = {nullptr, nullptr};  // mean L_z(0,0) point values [Ny]
  double2 *au0 = nullptr, *w0 = nullptr;  // mean U, W coefficients [Ny]
  int hcur = 0;
  // h^n's point values hv[hcur], hg[hcur] are not yet formed: hvc, hgc hold
  // h^n's coefficients and the next solve's right-hand-side kernels form them
  // (FFT transforms only)
  bool hdefer = false;
  // nonlinear work
  double2* Y[6] = {};     // [Py][Mz][Mx]   wall-normal points, Fourier modes
  double2* Z[6] = {};     // [Py][Lz][Mx]   wall-normal + span points, x modes
  float2* Yf[6] = {};     // mixed form: FP32 copies of Y and Z
  float2* Zf[6] = {};
  double2* Lc[3] = {};    // [Ny][Mz][Mx]   Lamb vector coefficients
  double2 *hvc = nullptr, *hgc = nullptr;  // h coefficients
  // solve work
  // v right-hand sides per mode: [column][even rows (ne - 2), then odd rows
  // (no - 2)], so a solve block reads each of its modes' rows contiguously;
  // omega_y right-hand sides per parity: [row][column] (the omega_y solve
  // reads 64 neighbouring columns at a time)
  double2* rvT = nullptr;
  double2 *roE = nullptr, *roO = nullptr;
  double2 *gv = nullptr, *go = nullptr;  // explicit-operator coefficient vectors
  // gv/go are written by the kinematics pass for the next solve (it runs after
  // every solve and whenever the state is replaced), so a resumed run forms
  // them with the same kernel as an uninterrupted one
  // device flag: the last kinematics pass saw a NaN or an infinity in v,
  // omega_y, U or W (cleared when a new state is loaded)
  int* nonfinite = nullptr;
  // operators
  double* host_mats = nullptr;
  double *tsyn = nullptr, *tana = nullptr, *tpts = nullptr, *t0 = nullptr, *t2 = nullptr,
         *t4 = nullptr;
  double* costab = nullptr;  // cos(pi n / 2N), n < 4N
  float* costabf = nullptr;  // the same rounded to FP32 (mixed form)
  float2 *twxf = nullptr, *twzf = nullptr;  // FP32 twiddles (mixed form)
  // Half-system inverses, transposed: group g, row j (< ne) holds
  //   [v even (neP)][v odd (noP)][omega even (neP)][omega odd (noP)]
  // entries i of column j, zero-padded (i >= n, and j >= no for the odd halves),
  // so a solve block reads one contiguous row per step.  Group 0: mean mode
  // (omega_y-type operator at k^2 = 0), its v halves are zero.
  double* inv = nullptr;
  // omega_y half systems: either their inverses sit beside the v ones in `inv`
  // (inv_w = inv + neP + noP, same strides), or the operator is diagonalised
  // (wdiag) and only the mean group's k^2 = 0 inverse is stored, in inv0.
  bool wdiag = false;
  int* crep = nullptr;                               // [nclass] first group of each class
  int *bmat = nullptr, *bg0 = nullptr, *bg1 = nullptr;  // [nblocks] class, groups (-1: none)
  double* inv0 = nullptr;
  const double* inv_w = nullptr;
  long stride_w = 0, row_w = 0;
  // diagonalised omega_y solve, per parity: packed matrix-core A fragments of
  // L = V^-1 M0^-1 ((n-1) x (n-1)) and R = Q V (n x (n-1)), eigenvalues lam
  double* wL[2] = {nullptr, nullptr};
  double* wR[2] = {nullptr, nullptr};
  double* wlam[2] = {nullptr, nullptr};
  double2* twx = nullptr;  // exp(-2 pi i m/Lx), m < Lx
  double2* twz = nullptr;  // exp(-2 pi i m/Lz), m < Lz
  // FFT forms of the wall-normal transforms (null when 2N has no split or
  // ncol is not a multiple of 16): the 2N-point twiddles exp(-2 pi i m/2N),
  // then exp(-2 pi i m/N) (m < N), then exp(i pi k/2N) (k < N), FP64
  double2* twy = nullptr;
  // unpadded physical output and plane statistics (allocated on first use)
  double2* phys_c = nullptr;                          // [Ny][Mz][Mx] work
  double* phys_r[3] = {nullptr, nullptr, nullptr};    // [Ny][Mz][2 Mx] physical u, v, w
  double* stats = nullptr;                            // [12][Ny]
  double2 *twxu = nullptr, *twzu = nullptr;           // unpadded twiddles (2 Mx, Mz), set up at creation
  double2* scratch = nullptr;  // this device's no-allocation work memory (shared by its contexts)
  long long bytes = 0;
  std::string err;
  // column blocks per wave of each cos-table contraction, picked at creation
  // by timing the candidates on the context's own shapes (0 = default)
  int cb_choice[K_COUNT] = {};
  float tune_ms[K_COUNT][3] = {};
  // optional per-launch HIP-event timing on the context stream
  bool timing = false;
  struct Mark {
    int kernel;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> event_pool;
  size_t pool_next = 0;

  Mats mats() const { return Mats{tsyn, tana, tpts, t0, t2, t4}; }
};

// FFT plan: Stockham stages with radices 4, 2, 3.
struct FftPlan {
  int L;
  int nst;
  int rad[16];
};
FftPlan make_plan(int L);
bool fft_length_ok(int L);

extern const char* const kKernelNames[K_COUNT];

// ---- kernel launchers (each returns hipGetLastError()) ----
bool reg_fft_length(int L);
bool reg_fft_length32(int L);  // register span FFTs; FP32 streamwise pass
hipError_t launch_zfft32(Ctx& c, const float2* const* in, float2* const* out, int nf, bool inv);
hipError_t launch_xpass32(Ctx& c);
hipError_t launch_synth_padded32(Ctx& c, const double2* const* in, float2* const* out, int nf);
// FP32 coefficients in (FFT transforms only): the values launch_synth_padded32 rounds its inputs to
hipError_t launch_synth_padded32f(Ctx& c, const float2* const* in, float2* const* out, int nf);
// FP32 coefficients out (FFT transforms only): launch_analysis32's values, which are FP32 numbers
hipError_t launch_analysis32f(Ctx& c, const float2* const* in, float2* const* out, int nf);
hipError_t launch_analysis32(Ctx& c, const float2* const* in, double2* const* out, int nf);
hipError_t launch_zfft_inv(Ctx& c, double2* const* in, double2* const* out, int nf);
hipError_t launch_xpass(Ctx& c);
hipError_t launch_zfft_fwd(Ctx& c, double2* const* in, double2* const* out, int nf);
hipError_t launch_synth_padded(Ctx& c, const double2* const* in, double2* const* out, int nf);
hipError_t launch_analysis(Ctx& c, const double2* const* in, double2* const* out, int nf);
hipError_t launch_hassemble(Ctx& c);
hipError_t launch_synth_points(Ctx& c, const double2* const* in, double2* const* out, int nf);
// ydct.hip: the wall-normal transforms as FFTs along y (when c.twy is set)
int ydct_length_supported(int N);  // 1 when 2N has an FFT split
hipError_t launch_ypoints(Ctx& c, const double2* const* in, double2* const* out, int nf);
hipError_t launch_ysynth(Ctx& c, const double2* const* in, double2* const* out, int nf);
hipError_t launch_yanalysis(Ctx& c, const double2* const* in, double2* const* out, int nf);
// which: 0 = v, 1 = omega_y; hc != nullptr: h^n from its coefficients hc, its
// point values written to hn; else hn holds them
hipError_t launch_yrhs(Ctx& c, int which, const double2* g, const double2* hc, double2* hn, const double2* hp,
                       double2* oE, double2* oO, double2* oT);
hipError_t launch_ysynth32(Ctx& c, const double2* const* in, float2* const* out, int nf);
hipError_t launch_ysynth32f(Ctx& c, const float2* const* in, float2* const* out, int nf);  // FP32 coefficients in
hipError_t launch_yanalysis32f(Ctx& c, const float2* const* in, float2* const* out, int nf);  // FP32 coefficients out
hipError_t launch_yanalysis32(Ctx& c, const float2* const* in, double2* const* out, int nf);
hipError_t launch_synth_points_h(Ctx& c, const double2* const* in, double2* const* out, int nf);  // FP32 compute in the mixed form
hipError_t launch_span_inverse_unpadded(Ctx& c, double2* inout);
hipError_t launch_stream_c2r_unpadded(Ctx& c, const double2* in, double* out);
// unpadded physical field: work = complex [Ny][Mz][Mx], out = real [Ny][nny][nnx]
hipError_t launch_physical(Ctx& c, const double2* field, double2* work, double* out);
hipError_t launch_plane_stats(Ctx& c, const double* u, const double* v, const double* w, double* out);
hipError_t launch_rhs(Ctx& c, int which);  // 0 = v, 1 = omega_y
hipError_t launch_solve(Ctx& c);       // v, omega_y and the mean mode (k_solve, then launch_vtranspose)
hipError_t launch_vtranspose(Ctx& c);  // v from the solve's per-group records to [Ny][Mz][Mx]
// The k_solve blocks of several channels in one launch: solve_members_bytes(n)
// of device memory hold their arguments, written by write_solve_members
// (synchronous; the arguments name the members' current AB2 history phase).
size_t solve_members_bytes(int n);
hipError_t write_solve_members(Ctx* const* m, int n, void* dev);
hipError_t launch_solve_members(hipStream_t st, Ctx* const* m, int n, const void* dev);
hipError_t launch_wsolve(Ctx& c);      // diagonalised omega_y solve (all modes but the mean)
// u, w, omega_x, omega_z (and the next RHS brackets) from v, omega_y, U, W;
// f32: u, w, omega_x, omega_z into uf, wf, oxf, ozf (rounded to FP32) instead
hipError_t launch_kinematics(Ctx& c, bool f32 = false);
hipError_t launch_setup_inverses(Ctx& c);

}  // namespace chf
