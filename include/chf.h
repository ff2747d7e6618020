/*
 * chf.h - C ABI of the MI355X per-timestep core of the Fourier-Fourier-Chebyshev
 * channel-flow DNS (v - omega_y formulation, Crank-Nicolson / Adams-Bashforth-2).
 *
 * The reference solver (poiseuilleflow.f90 = PF, paraandcons.f90 = PC) has no
 * plugin or operator API: its time loop (PF:220-532) calls Fortran external
 * subroutines that share state through module arrays (PC:28-38).  Each entry
 * point below replaces one of those call sites; the comment names it.
 *
 * Conventions
 *   - All arrays are host pointers in the reference memory layout:
 *       spectral/point field  real*8 (0:nx,0:ny,0:nz), Re/Im interleaved along
 *                             dim 1 == complex128[Ny][Mz][Mx], C order,
 *                             Mx = nnx/2, Mz = nny, Ny = nnz+1;
 *       mean-mode array       real*8 (2,0:nz) == complex128[Ny].
 *   - Return value: 0 ok; CHF_NONFINITE when the state holds a NaN or an
 *     infinity (below); other values < 0: invalid argument (message in
 *     chf_last_error); > 0 a HIP runtime error code.
 *   - One context = one channel on the current HIP device and one stream.
 *     Calls on one context must come from one host thread.
 */
#ifndef CHF_H
#define CHF_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status of chf_set_state, chf_init_fields, chf_solve, chf_step and
 * chf_ensemble_step when v, omega_y, U or W hold a NaN or an infinity after
 * the call (a blow-up, or a non-finite state loaded).  The reference has no
 * such check: it ignores every status and prints CFL numbers only
 * (PF:201-204, 2933).  The state is left as it is. */
#define CHF_NONFINITE (-2)

/* Compile-time parameters of PC:4-26 that the timestep core reads (P = 1). */
typedef struct chf_params {
  int nnx;        /* streamwise points (PC:5); nnx % 4 == 0                    */
  int nny;        /* spanwise points (PC:6); nny % 4 == 0                      */
  int nnz;        /* Chebyshev degree N (PC:7); N >= 4                         */
  double aphi;    /* Lx = 2 pi aphi (PC:21)                                    */
  double beta;    /* Lz = 2 pi beta (PC:21)                                    */
  double dt;      /* time step (PC:23)                                         */
  double re;      /* Reynolds number (PC:24)                                   */
  double re_p;    /* mean pressure gradient (PC:25)                            */
  int precision;  /* 0 = FP64 throughout; 1 = mixed (FP32 transforms/products, */
                  /*     FP64 solves and state)                                */
} chf_params;

/* Pre-solve state = the restart record of PF:2476-2479 (output) and
 * PF:1913-1917 (read back).  Field arrays are complex128[Ny][Mz][Mx], mean
 * arrays complex128[Ny].  h-arrays hold wall-normal point values y_j =
 * cos(pi j/N); v, omega_y, au0, w0 hold Chebyshev coefficients. */
typedef struct chf_state {
  double *clamb1n, *clamb1nm1, *clamb3n, *clamb3nm1, *au0, *w0;  /* record 1 */
  double *vnp1, *omegaynp1, *rhsvn, *rhsomegayn, *unp1, *wnp1;   /* record 2 */
} chf_state;

/* Builds the per-(kx,|kz|) operator inverses on the GPU (replaces SCHEB PF:156
 * and the get_coef_v/get_coef_omegay + zgetrf/zgetri set-up loop, PF:181-204).
 * The omega_y operator is diagonalised when the decomposition reproduces
 * direct solves to 1e-12, else its half-system inverses are stored;
 * CHF_OMEGA_INVERSES=stored in the environment selects the stored inverses.
 * The wall-normal transforms (padded synthesis, analysis, Gauss-Lobatto
 * point values, the solves' right-hand sides) run as FFTs along y, in the
 * form's own arithmetic, when 2N is one of 64, 96, 128, 192, 256, 384, 512 and
 * nnx/2 * nny is a multiple of 16, else as cos-table products on the matrix
 * cores; CHF_Y_TRANSFORMS=matrix selects the products. */
int chf_create(const chf_params* prm, void** ctx);
int chf_destroy(void* ctx);
const char* chf_last_error(void* ctx);

/* Host -> device copy of a pre-solve state (replaces initial(isign=1),
 * PF:1861-1924).  Derives u, w, omega_x, omega_z from the loaded v, omega_y,
 * U, W (getvelocityandvorticity, PF:2347-2407), so chf_get_fields,
 * chf_physical and chf_statistics describe the loaded state, and a run
 * resumed from a chf_get_state record continues bit for bit as the
 * uninterrupted run does. */
int chf_set_state(void* ctx, const chf_state* st);
/* Device -> host copy of the pre-solve state (replaces output, PF:2460-2504).
 * With the FFT transforms, first forms h^n's point values when the last
 * chf_nonlinear left them to the next solve (the values that solve would
 * form, bit for bit). */
int chf_get_state(void* ctx, const chf_state* st);

/* Cold start (replaces initial(isign=-1) after input, PF:1932-1946): from v,
 * omega_y (coefficients, complex128[Ny][Mz][Mx]) and the mean profiles U, W
 * (complex128[Ny]) derive u, w, omega_x, omega_z (getvelocityandvorticity,
 * PF:2347-2407), set h^{-1} = 0 and run one getnonlinear. */
int chf_init_fields(void* ctx, const double* vnp1, const double* omegaynp1,
                    const double* au0, const double* w0);

/* getnonlinear (PF:1380-1612): six dealiased products -> Lamb vector ->
 * h_v, h_g, AB2 history rotation, point values.  With the FFT transforms the
 * point values of h^n are formed by the next chf_solve's right-hand-side
 * kernels (or by chf_get_state, or a second chf_nonlinear, if one comes
 * first); results are the same either way. */
int chf_nonlinear(void* ctx);
/* Per-mode v / omega_y solves, kinematics and the mean mode (PF:242-474).
 * Returns after the solve has completed (CHF_NONFINITE on a blow-up). */
int chf_solve(void* ctx);
/* n x [chf_solve ; chf_nonlinear] from a pre-solve state (PF:220-532).
 * Returns after the steps have completed (CHF_NONFINITE when the state after
 * the last solve is not finite). */
int chf_step(void* ctx, int n);
/* Waits for all work queued on the context's stream. */
int chf_sync(void* ctx);

/* Post-solve velocity and vorticity (unp1, wnp1, omegaxnp1, omegaznp1 as the
 * reference holds them after its solve loop, PF:387-472), any pointer may be
 * NULL.  Valid after chf_solve or chf_init_fields.  The mixed form with the
 * FFT transforms keeps these four fields in FP32 during a run (the values its
 * padded synthesis reads); this call, chf_physical and chf_statistics first
 * form them in FP64 from v, omega_y, U, W (one kinematics pass). */
int chf_get_fields(void* ctx, double* unp1, double* wnp1, double* omegaxnp1,
                   double* omegaznp1);

/* Unpadded physical field (replaces transform(1, .), PF:948-961 with
 * fourtrx_zcomp(+1), PF:774-860, as statistics calls it at PF:2723-2725):
 * field 0..5 = u, v, w, omega_x, omega_y, omega_z after the last solve, out =
 * real*8 (0:nnx-1, 0:nny-1, 0:N) == double[Ny][nny][nnx] at x_i = 2 pi aphi i/nnx,
 * z_j = 2 pi beta j/nny, y_k = cos(pi k/N). */
int chf_physical(void* ctx, int field, double* out);

/* chf_physical into caller-owned device memory, allocating nothing (no
 * counterpart in the reference: it is for callers that cannot allocate, such
 * as code under HIP graph capture or a tight serving loop).  `out` is a device
 * pointer to Ny * nny * nnx doubles, laid out as chf_physical's output.  The
 * transform's complex work plane (Ny * nny * nnx/2 values, 16 bytes each)
 * lives in a 128 MiB array that the library holds on each GPU for the life of
 * the process: larger grids are rejected (< 0, use chf_physical).  The work is
 * queued on the context's stream (chf_sync waits for it).  Every call on one
 * GPU shares that GPU's array, so calls that overlap on one GPU (from
 * contexts on different streams or threads) race. */
int chf_physical_noalloc(void* ctx, int field, double* out);

/* Plane statistics of the physical u, v, w after the last solve (the
 * reductions of statistics, PF:2816-3020): out = double[12][Ny], rows
 * ua va wa uu vv ww uv uw vw max|u| max|v| max|w| per y_k (means, covariances
 * about the plane mean, and extrema over each plane). */
int chf_statistics(void* ctx, double* out);

/* Sizes derived from the parameters (for callers that allocate buffers). */
int chf_dims(void* ctx, int* mx, int* mz, int* ny);

/* Device bytes held by the context. */
long long chf_device_bytes(void* ctx);

/* Instrumentation (no counterpart in the reference, whose only timer is the
 * mpi_wtime bracket of the solve loop, PF:217-229, 481-486).  With timing on,
 * every kernel launch of chf_solve/chf_nonlinear/chf_step is bracketed by HIP
 * events and the launches run one after another (without timing the omega_y
 * solve overlaps the v solve on a second stream), so each kernel class's time
 * is its own; chf_timing sums the elapsed time of kernel class `id`
 * (0 <= id < chf_kernel_count()) since the last chf_set_timing. */
int chf_set_timing(void* ctx, int on);
int chf_kernel_count(void);
const char* chf_kernel_name(int id);
int chf_timing(void* ctx, int id, double* total_ms, int* launches);

/* Autotuner report (no counterpart in the reference): chf_create times the
 * cos-table contractions (kernel classes synth_padded, analysis, synth_points,
 * rhs_v, rhs_o) with 1, 2 and 4 column blocks per wavefront on the context's
 * own shapes and keeps the fastest.  Returns the chosen count (0 for classes
 * without a tile choice); ms_by_choice (may be NULL) receives the three
 * measured times per launch.  The choice never changes results. */
int chf_tile_choice(void* ctx, int id, double* ms_by_choice);
/* Overrides the tuned choice of kernel class `id` (a class for which
 * chf_tile_choice returns > 0) with cb in {1, 2, 4}; results do not change. */
int chf_set_tile_choice(void* ctx, int id, int cb);

/* Ensemble: n independent channels (any mix of grid sizes, Reynolds numbers
 * and precisions) advanced together, one step of every member being one
 * hipGraph launch: the members' right-hand sides as parallel branches, one
 * solve launch over every member's solve blocks, then the members'
 * kinematics and nonlinear terms as parallel branches (each member's result
 * is bit-identical to the same channel stepped alone).  The
 * reference runs one channel per job, so these have no single counterpart;
 * each member is a full context whose state is exchanged with the calls
 * above (chf_set_state, chf_init_fields, chf_get_state, chf_get_fields), and
 * one ensemble step is chf_step(member, 1) for every member (PF:220-532).
 * chf_ensemble_step returns after the steps have completed.  A member context
 * belongs to its ensemble: do not pass it to chf_destroy. */
int chf_ensemble_create(const chf_params* members, int n, void** ens);
int chf_ensemble_destroy(void* ens);
const char* chf_ensemble_last_error(void* ens);
int chf_ensemble_size(void* ens);
int chf_ensemble_member(void* ens, int i, void** ctx);
int chf_ensemble_step(void* ens, int nsteps);

#ifdef __cplusplus
}
#endif
#endif /* CHF_H */
