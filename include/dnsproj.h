/*
 * dnsproj.h -- C ABI of the MI355X (gfx950) pressure-projection library.
 *
 * Plain C types only.  Device arrays are hipMalloc'd pointers passed as
 * double*; streams are hipStream_t passed as void* (NULL = the null stream).
 * Every entry point returns 0 on success or a DNSP_ERR_* code, and leaves a
 * message for dnsp_last_error() (thread-local).  The reference aborts with
 * `error stop` in these situations; its Fortran shim does the same on a
 * nonzero status.
 *
 * Array layout is the reference's: Fortran column-major, x fastest,
 *   P(0:nx-1, 0:nz-1, 0:ny-1)  ->  P[i + nx*(j + nz*k)].
 */
#ifndef DNSPROJ_H
#define DNSPROJ_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DNSP_OK               0
#define DNSP_ERR_ARG          1  /* invalid size, NULL pointer, non-finite input */
#define DNSP_ERR_HIP          2  /* a HIP runtime call failed */
#define DNSP_ERR_NOMEM        3  /* device allocation failed */
#define DNSP_ERR_SINGULAR     4  /* a zero or non-finite tridiagonal pivot */
#define DNSP_ERR_UNSUPPORTED  5  /* FFT length with a prime factor above 13, or a form's size limits */
#define DNSP_ERR_NOT_CONVERGED 6 /* refinement stopped above its tolerance (outputs still written) */
#define DNSP_ERR_NONFINITE    7  /* NaN or Inf in a solution or residual: the right-hand side held one
                                  * (the reference stops on a NaN bulk value, src/mod_get_bulk.f90:180-186) */

typedef struct dnsp_poisson dnsp_poisson;

/* Message of the last failing call on this thread ("" if none). */
const char* dnsp_last_error(void);

/* Library version string. */
const char* dnsp_version(void);

/* Waits for the work enqueued on `stream` (the reference synchronises the
 * device after every stage; callers that need it use this). */
int dnsp_stream_synchronize(void* stream);

/*
 * Replaces init_spec_poisson_multigpu
 * (src/poisson_solver_mgpu_dtdma/poisson_solver_multigpu.f90:33-326), single
 * y-rank branch (mpi_divs_y == 1, :162-222), and the size checks of
 * alloc_poisson_solver_multigpu.f90:3-38.
 *   nx_glob, nz_glob : periodic sizes (any size >= 1 whose prime factors are in {2,3,5,7,11,13})
 *   ny               : wall-normal size (>= 2)
 *   inv_dx, inv_dz   : 1/dx, 1/dz (the reference's inv_dx, inv_dz)
 *   P_band_*         : y bands (length ny) as read from geom_data/P_band_abc.dat,
 *   P_band_*_00      : and the (0,0)-mode set from P_band_abc_00.dat
 *   precision        : DNSP_PRECISION_FP64 (0): the plain FP64 form, as the
 *                      reference; DNSP_PRECISION_MIXED (1): the same, plus the
 *                      state of the mixed-precision form (two FP32 fields;
 *                      nx, nz in {32,64,96,128,192,256,384,512,768}, ny <= 1024)
 * The eigenvalues a_x, a_z (:132-155) and the FFT twiddles are computed on the
 * host; the reference's poi_C array (:162-179) is not stored.
 * On success *out owns device memory on the current HIP device.
 */
#define DNSP_PRECISION_FP64   0
#define DNSP_PRECISION_MIXED  1
int dnsp_poisson_init(int nx_glob, int nz_glob, int ny, double inv_dx, double inv_dz,
                      const double* P_band_a, const double* P_band_b, const double* P_band_c,
                      const double* P_band_a_00, const double* P_band_b_00, const double* P_band_c_00,
                      int precision, dnsp_poisson** out);

/*
 * Replaces run_spec_poisson_multigpu (poisson_solver_multigpu.f90:328-554),
 * single y-rank branch.  P is a device array of nx*nz*ny doubles holding the
 * right-hand side on entry and the solution on exit (same pinned (0,0) mode as
 * the reference; the mean is not removed).  Enqueued on `stream`; returns
 * without synchronising.  P needs the alignment of a double only (an array
 * that is not 16-byte aligned runs runtime-sized x kernels).
 */
int dnsp_poisson_run(dnsp_poisson* h, double* P, void* stream);

/* Out-of-place form of dnsp_poisson_run: reads the RHS from `rhs` (left
 * unchanged unless rhs == P) and writes the solution to P.  Same kernels and
 * HBM traffic as the in-place form. */
int dnsp_poisson_solve(dnsp_poisson* h, const double* rhs, double* P, void* stream);

/* Same as dnsp_poisson_run for a host array: copies in, solves, copies out,
 * synchronises.  For callers whose P lives in host memory (the reference's
 * CPU test driver).  DNSP_ERR_NONFINITE when the solution holds a NaN or
 * Inf (P_host is still written). */
int dnsp_poisson_run_host(dnsp_poisson* h, double* P_host);

/* Counts the non-finite values of a solution P (nx*nz*ny device doubles)
 * after the work on `stream`: DNSP_OK when there are none, else
 * DNSP_ERR_NONFINITE with the count in *n_nonfinite (may be NULL).  One read
 * of P; synchronises `stream`.  A NaN or Inf anywhere in a right-hand side
 * reaches every point of the solution (the x and z transforms spread it over
 * a plane's modes, each mode's y solve over the column), so a solve's input
 * need not be checked separately.  No counterpart in the reference's solver. */
int dnsp_poisson_check(dnsp_poisson* h, const double* P, long long* n_nonfinite, void* stream);

/* Stage entry points of the same solve (device arrays, in place):
 *   forward  = x R2HC, z R2HC          (poisson_solver_multigpu.f90:360-393)
 *   tridiag  = per-mode y Thomas       (:395-422)
 *   backward = z HC2R, x HC2R, 1/(nx*nz) (:501-551)
 * The spectrum uses FFTW's halfcomplex order along x and along z. */
int dnsp_poisson_forward(dnsp_poisson* h, double* P, void* stream);
int dnsp_poisson_tridiag(dnsp_poisson* h, double* S, void* stream);
int dnsp_poisson_backward(dnsp_poisson* h, double* S, void* stream);

/* Stage `stage` (0..4: x forward, z forward, tridiagonal, z inverse, x
 * inverse) of dnsp_poisson_solve(h, rhs, P): running the five in order on
 * one stream equals the solve.  Where the handle runs the blocked solve
 * (below) the intermediate layouts are internal and stages 0, 3 write the
 * handle's workspace, stages 1, 2 P; elsewhere these are the in-place
 * stages of dnsp_poisson_stage.  For timing the kernels of the solve one by
 * one.  No counterpart in the reference. */
int dnsp_poisson_solve_stage(dnsp_poisson* h, const double* rhs, double* P, int stage, void* stream);

/* Blocked solve (no counterpart in the reference; same arithmetic, results
 * bit-identical to the in-place stages with the same kernel variants): for
 * nx a multiple of 16 with the compile-time FFT kernels and the partitioned
 * tridiagonal, dnsp_poisson_run / dnsp_poisson_solve go through a workspace
 * of one field held by the handle: x-transformed lines go to an x-blocked
 * layout [k][a/16][j][a%16], the z transform turns that into mode-major
 * blocks [a/16][b][k][a%16], the tridiagonal solve writes the x-blocked
 * layout back, the inverse z transform halfcomplex lines, so that every z
 * tile and tridiagonal mode group reads a contiguous block of memory.  On (1) by default where supported;
 * dnsp_poisson_set_blocked(h, 0) returns to the in-place stages.  Arrays
 * that are not 16-byte aligned take the in-place stages.  The workspace
 * makes concurrent solves on one handle race (as the reference's shared work
 * arrays do). */
int dnsp_poisson_set_blocked(dnsp_poisson* h, int on);
int dnsp_poisson_get_blocked(const dnsp_poisson* h, int* supported, int* on);
/* The variants of the five kernels dnsp_poisson_solve runs now (stage order:
 * x forward, z forward, tridiagonal, z inverse, x inverse), in the blocked
 * solve (z: 0 = 256, 1 = 512 threads) or the in-place stages. */
int dnsp_poisson_get_solve_variants(const dnsp_poisson* h, int* v);

/* One kernel of the solve, for timing and tuning:
 *   DNSP_STAGE_X_FWD (:360-375), DNSP_STAGE_Z_FWD (:377-393), DNSP_STAGE_TRIDIAG (:395-422),
 *   DNSP_STAGE_Z_BWD (:501-520), DNSP_STAGE_X_BWD (:523-551).
 * Reads src, writes P; src may differ from P only for the two x stages.
 * Running the five in order on one array equals dnsp_poisson_run. */
#define DNSP_STAGE_X_FWD    0
#define DNSP_STAGE_Z_FWD    1
#define DNSP_STAGE_TRIDIAG  2
#define DNSP_STAGE_Z_BWD    3
#define DNSP_STAGE_X_BWD    4
int dnsp_poisson_stage(dnsp_poisson* h, const double* src, double* P, int stage, void* stream);

/* Kernel strategy (also the knobs of the tuner):
 *   fft:     DNSP_FFT_AUTO (compile-time-sized kernels where the length has
 *            them, else the runtime-sized ones), DNSP_FFT_GENERIC (always the
 *            runtime-sized kernels), DNSP_FFT_FIXED (compile-time-sized only;
 *            DNSP_ERR_UNSUPPORTED if a length has none)
 *   tridiag: DNSP_TRIDIAG_AUTO (partitioned for ny >= 32), DNSP_TRIDIAG_SERIAL
 *            (one serial Thomas sweep per mode, the reference's operation order
 *            at :406-421, bit-identical to it; needs an nx*nz*ny workspace),
 *            DNSP_TRIDIAG_PARTITION (registers + parallel cyclic reduction on
 *            the chunk tips; one HBM read and write per point; ny <= 1024)
 * No counterpart in the reference, which has one kernel per stage. */
#define DNSP_FFT_AUTO            0
#define DNSP_FFT_GENERIC         1
#define DNSP_FFT_FIXED           2
#define DNSP_TRIDIAG_AUTO        0
#define DNSP_TRIDIAG_SERIAL      1
#define DNSP_TRIDIAG_PARTITION   2
int dnsp_poisson_set_strategy(dnsp_poisson* h, int fft, int tridiag);
/* What the handle currently runs: 1/0 compile-time x and z FFT kernels, and
 * the partitioned tridiagonal rows per chunk (0 = serial Thomas). */
int dnsp_poisson_get_strategy(const dnsp_poisson* h, int* fft_ct_x, int* fft_ct_z, int* part_rows);

/* Kernel variants.  xvar, zvar (0..3): the compile-time-sized FFT kernels'
 * (complex elements per workgroup tile, threads): x (768,256), (3072,512),
 * (1536,256), (6144,512); z (3072,256), (3072,512), (1536,256), (6144,512);
 * ignored where a length has no such kernels.  set_variants sets both the
 * forward and the inverse transforms (the tuner picks them separately; get
 * returns the forward ones) and the blocked solve's knobs alike.
 * tvar (0..7): partitioned tridiagonal workgroups of (chunks per mode, modes)
 * = (32, 16) with registers capped for 4 resident waves per SIMD (0),
 * (32, 8) (1), (48, 16) (2), (16, 16) for ny <= 512, else (64, 8) (3);
 * rows per chunk follow from ny.  4..7: the same shapes with two-pass chunks
 * (the chunk interior is solved again once its end values are known,
 * instead of being kept in registers across the reduced solve).  The
 * mixed-precision form has its own set; on its FP32 spectra shapes 3 and 7
 * are (16, 32), whose 32 modes make 128-byte rows.  No counterpart in the
 * reference. */
int dnsp_poisson_set_variants(dnsp_poisson* h, int xvar, int zvar, int tvar);
int dnsp_poisson_get_variants(const dnsp_poisson* h, int* xvar, int* zvar, int* tvar);
int dnsp_poisson_set_mixed_variants(dnsp_poisson* h, int xvar, int zvar, int tvar);
int dnsp_poisson_get_mixed_variants(const dnsp_poisson* h, int* xvar, int* zvar, int* tvar);

/* Autotuner: times `reps` forward+inverse pairs of every FFT kernel variant
 * of each direction, and pairs of tridiagonal solves of every partitioned
 * variant, on `work` (nx*nz*ny device doubles, overwritten; NULL = a
 * temporary allocation) and keeps the fastest in the handle; for a
 * DNSP_PRECISION_MIXED handle the same for the mixed-precision kernels.
 * times_ms (optional, DNSP_TUNE_SLOTS doubles) receives the mean ms per pair:
 * 4 x, 4 z and 8 tridiagonal variants of the FP64 form (slots 0..15), then
 * the same of the mixed form (16..31); -1 where not applicable.  Synchronises
 * `stream`. */
#define DNSP_TUNE_SLOTS 32
int dnsp_poisson_autotune(dnsp_poisson* h, double* work, int reps, void* stream, double* times_ms);

/*
 * Mixed-precision form of the solve (no counterpart in the reference, whose
 * solve is FP64 throughout; it solves the same discrete system).  Sweep 0:
 * FP32 x and z transforms and FP32 spectra, FP32 arithmetic in the y
 * tridiagonal solves of the well-conditioned modes (|lambda| >=
 * max_k(|a_k| + |c_k|) / 128, all but the lowest kx, kz), FP64 arithmetic in
 * those of the others, and the (0,0) mode (the mean profile, pinned by the
 * P_band_*_00 set) solved entirely in FP64 from FP64 plane sums.  Each
 * refinement sweep evaluates the FP64 residual r = rhs - L p inside the
 * forward x transform, solves L d = r the same way and adds d in FP64.
 * One refinement reaches a relative residual ||rhs - L p||/||rhs|| of about
 * 1e-11 on channel grids; the handle must be DNSP_PRECISION_MIXED.
 *   dnsp_poisson_solve_mixed: `refinements` sweeps (0..DNSP_MAX_REFINEMENTS)
 *     after sweep 0; rhs may equal P (a copy of rhs is then kept when
 *     refinements > 1).  rhs and P 16-byte aligned (DNSP_ERR_ARG
 *     otherwise).  Enqueued on `stream`, no synchronisation.
 *   dnsp_poisson_run_mixed: in place, P = rhs on entry.
 *   dnsp_poisson_solve_refine: refines until the relative residual is <= tol
 *     or max_refinements sweeps are done; returns the sweeps done and the
 *     residual (each check reads P and rhs once more).  Synchronises.
 *     Status DNSP_ERR_NONFINITE when the residual is NaN or Inf, and
 *     DNSP_ERR_NOT_CONVERGED when it is still above tol after
 *     max_refinements sweeps; *refinements_done and *rel_resid are written
 *     in both cases.
 */
#define DNSP_MAX_REFINEMENTS 8
/* The form dnsp_poisson_run (and dnsp_poisson_run_host) uses: -1 = FP64
 * (the default), 0..DNSP_MAX_REFINEMENTS = the mixed form with that many
 * refinements (DNSP_PRECISION_MIXED handles only).  Lets an unmodified
 * caller of the drop-in select the mixed form. */
int dnsp_poisson_set_form(dnsp_poisson* h, int refinements);
int dnsp_poisson_solve_mixed(dnsp_poisson* h, const double* rhs, double* P, int refinements, void* stream);
int dnsp_poisson_run_mixed(dnsp_poisson* h, double* P, int refinements, void* stream);
int dnsp_poisson_solve_refine(dnsp_poisson* h, const double* rhs, double* P, double tol, int max_refinements,
                              int* refinements_done, double* rel_resid, void* stream);
/* FP32 arithmetic in the sweeps' well-conditioned tridiagonal solves: on (1,
 * the default) or off (0: FP64 arithmetic for every mode). */
int dnsp_poisson_set_mixed_fp32(dnsp_poisson* h, int on);
/* The mixed form's split: whether FP32 sweeps are on, the |lambda| from which
 * a mode runs in FP32, and how many of the nx*nz modes do (for the current
 * tridiagonal variant).  Any output may be NULL. */
int dnsp_poisson_mixed_info(const dnsp_poisson* h, int* fp32_on, double* f32_min, long long* fp32_modes);

/*
 * Batched form: many independent solves of uneven sizes (no counterpart in
 * the reference, which solves one field per rank).  Each stage of the solve
 * is one launch per distinct transform length (x, z) or rows-per-chunk
 * (tridiagonal) over the tiles of all instances of that size.
 *   dnsp_batch_create: n instances; nx[i], nz[i] in {32,64,96,128,192,256,384,
 *     512,768}, 2 <= ny[i] <= 1024; bands = the instances' six band arrays
 *     (a, b, c, a00, b00, c00, ny[i] each) concatenated in instance order;
 *     precision as for dnsp_poisson_init (MIXED adds two FP32 fields per
 *     instance).
 *   dnsp_batch_bind: the device arrays of every instance (host arrays of n
 *     device pointers; rhs[i] != P[i], both 16-byte aligned); synchronous,
 *     done once.
 *   dnsp_batch_run: FP64 solves P[i] = L_i^{-1} rhs[i] of all instances.
 *   dnsp_batch_run_mixed: mixed-precision solves, `refinements` 0 or 1.
 *   Both enqueue on `stream` only (no allocation, no synchronisation), so a
 *   caller may capture them in a graph.
 *   dnsp_batch_autotune: times the whole batched solve per variant (x, z:
 *     0..3 as dnsp_poisson_set_variants; tridiagonal: 0 = 8-mode groups of
 *     32 chunks, 1 = the same with registers capped for 4 waves per SIMD, 2
 *     and 3 = 0 and 1 with two-pass chunks, 4 and 5 = 8-mode groups of 16
 *     chunks of twice the rows, one-pass and two-pass (32 chunks where the
 *     rows would exceed 32)), keeps the fastest; overwrites the bound P
 *     arrays; times_ms as for dnsp_poisson_autotune.  Synchronises.
 *   dnsp_batch_set_variants / get_variants: FP64 x, z, tridiagonal, then the
 *     mixed form's (6 ints).
 */
typedef struct dnsp_batch dnsp_batch;
int dnsp_batch_create(int n, const int* nx, const int* nz, const int* ny, const double* inv_dx,
                      const double* inv_dz, const double* bands, int precision, dnsp_batch** out);
int dnsp_batch_bind(dnsp_batch* b, const double* const* rhs, double* const* P);
int dnsp_batch_run(dnsp_batch* b, void* stream);
/* The FP64 solve of every instance in ONE persistent cooperative launch: all
 * workgroups resident, the five stages separated by grid barriers, the
 * workgroups taking every stage's tiles of all instances in turn.  x and z use variant 2 ((1536, 256) tiles), the tridiagonal
 * two-pass 8-mode groups: results equal dnsp_batch_run with
 * dnsp_batch_set_variants(b, 2, 2, 2, ...).  ny <= 384 (DNSP_ERR_UNSUPPORTED
 * otherwise, or where the device has no cooperative launch).  One kernel
 * over all sizes holds the registers of the largest body, so it is slower
 * than the per-size launches of dnsp_batch_run (measured in bench.py).
 * Not capturable in a graph.  No counterpart in the reference. */
int dnsp_batch_run_persistent(dnsp_batch* b, void* stream);
int dnsp_batch_run_mixed(dnsp_batch* b, int refinements, void* stream);
int dnsp_batch_autotune(dnsp_batch* b, int reps, void* stream, double* times_ms);
int dnsp_batch_set_variants(dnsp_batch* b, int xvar, int zvar, int tvar, int mxvar, int mzvar, int mtvar);
int dnsp_batch_get_variants(const dnsp_batch* b, int* v6);
int dnsp_batch_size(const dnsp_batch* b, int* n);
int dnsp_batch_free(dnsp_batch* b);

/* Relative residual ||rhs - L P|| / ||rhs|| (||rhs - L P|| if rhs = 0) of the
 * discrete operator the solve inverts: periodic second differences in x and
 * z, the plain y bands, and the (0,0)-mode bands on the xz-plane means.  FP64,
 * any precision of handle.  Synchronises `stream`.  DNSP_ERR_NONFINITE (with
 * *rel_resid written) when the residual is NaN or Inf. */
int dnsp_poisson_residual(dnsp_poisson* h, const double* rhs, const double* P, double* rel_resid, void* stream);

/* dnsp_poisson_residual with no device allocation anywhere on its path: its
 * scratch (line sums, plane constants, reduction partials) lives in a
 * file-scope device array of the library, 2.1 MB per GPU, held until the
 * process ends.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (tight serving loops, code run while a
 * caller captures a stream into a graph, harnesses that forbid hipMalloc).
 * Limits: nz*ny <= 262144 lines and ny <= 4096 (DNSP_ERR_ARG beyond them);
 * the handle's device must be current.  Same result as
 * dnsp_poisson_residual, bit for bit.  Every call on one GPU shares that
 * GPU's scratch: calls that overlap on one GPU (other streams, other host
 * threads) race.  Synchronises `stream`. */
int dnsp_poisson_residual_noalloc(dnsp_poisson* h, const double* rhs, const double* P, double* rel_resid,
                                  void* stream);

/* min |pivot| over all modes and rows: the init diagnostic min_abs_div_bacp
 * (poisson_solver_multigpu.f90:181-222), computed on the device at init. */
int dnsp_poisson_min_pivot(const dnsp_poisson* h, double* out);

/* Sizes of the handle: nx, nz, ny. */
int dnsp_poisson_shape(const dnsp_poisson* h, int* nx, int* nz, int* ny);

/* Releases the handle's device memory (dealloc_poisson_solver_multigpu.f90). */
int dnsp_poisson_free(dnsp_poisson* h);

/* Block partition, remainder to the first blocks
 * (poisson_solver_multigpu.f90:556-575).  Host only. */
int dnsp_get_lo_n_bounds_1d(int mpi_divs, int n_glob, int* lo_blocks_mpi, int* n_blocks_mpi);

/* Poisson y bands from the wall-normal faces yu(0:ny)
 * (src/pycode/calc_bands.py:181-201 with yp_ext from :226-227).  Host only.
 * is_dirich_first = 0 gives P_band_abc, 1 gives P_band_abc_00. */
int dnsp_mk_P_abc(int ny, const double* yu, int is_dirich_first, double* a, double* b, double* c);


/* ------------------------------------------------------------------------ */
/* Projection stencils (src/mod_cfd_subroutines.f90).  "Halo arrays" are the  */
/* reference's (-1:nx, -1:nz, -1:ny) fields: (nx+2)*(nz+2)*(ny+2) doubles, x  */
/* fastest; P is (0:nx-1, 0:nz-1, 0:ny-1).  1-D coefficient arrays, rho_min  */
/* and wall values passed by pointer are device arrays.  All entry points    */
/* enqueue on `stream` and return without synchronising.                     */
/* ------------------------------------------------------------------------ */

/* One-rank effect of diezdecomp_halos_execute_generic with the x and z
 * periodic halo objects (inc_set_alloc_arrays.f90:338-356): periodic wrap of
 * the x and z halo cells over every padded plane.  The y object is not
 * periodic and changes nothing on one rank. */
int dnsp_halo_periodic_xz(double* A, int nx, int nz, int ny, void* stream);

/* set_val_slice (src/mod_utils.f90:38-61): planes k0..k1 of the 0-based
 * padded view (0 = the k = -1 ghost; negative values count from the end)
 * set to val over the whole padded plane. */
int dnsp_set_val_slice(double* A, int nx, int nz, int ny, int k0, int k1, double val, void* stream);

/* Neumann y ghosts A(:,:,-1) = A(:,:,0), A(:,:,ny) = A(:,:,ny-1)
 * (mod_cfd_subroutines.f90:1100-1120, periodicity_pressure :1022-1038). */
int dnsp_neumann_ghosts_y(double* A, int nx, int nz, int ny, void* stream);

/* Poisson right-hand side, incompressible (mod_cfd_subroutines.f90:527-537):
 * P = inv_dtmax*((dU*inv_dx + dW*inv_dz) + dV*inv_dy_p(k)); hats are halo arrays
 * with valid x/z halos and V_hat wall planes. */
int dnsp_build_p_rhs(double* P, const double* U_hat, const double* W_hat, const double* V_hat,
                     const double* inv_dy_p, int nx, int nz, int ny, double inv_dtmax, double inv_dx,
                     double inv_dz, void* stream);

/* Poisson right-hand side, variable density with the rho_min splitting
 * (mod_cfd_subroutines.f90:538-574).  P_now, P_old, rho, div_now: halo arrays;
 * inv_dy_p_pad: ny+1 values; rho_min: one device double. */
int dnsp_build_p_rhs_varprop(double* P, const double* U_hat, const double* W_hat, const double* V_hat,
                             const double* P_now, const double* P_old, const double* rho, const double* div_now,
                             const double* inv_dy_p, const double* inv_dy_p_pad, const double* rho_min, int nx,
                             int nz, int ny, double inv_dtmax, double inv_dx, double inv_dz, double inv_dx2,
                             double inv_dz2, void* stream);

/* Minimum of rho over the interior into out (one device double)
 * (mod_cfd_subroutines.f90:470-499; the MPI_Allreduce is the identity on one GPU). */
int dnsp_rho_min(const double* rho, int nx, int nz, int ny, double* out, void* stream);

/* Gradient correction update_vars (mod_cfd_subroutines.f90:1087-1216,
 * :1225-1244; bulk forcing :1217-1221 not included):
 *   P_per <- P extended periodically in x/z and by Neumann ghosts in y (NULL to skip
 *            in the incompressible form);
 *   U, W (all k) and V (k >= kmin_Vhat) corrected from the hats;
 *   variable density (use_incompressible = 0): P_old <- P_now, P_now <- P_per;
 *   U, V, W periodic x/z halos, then wall planes from wall_bc (device, 6 doubles:
 *   U_bot, V_bot, W_bot, U_top, V_top, W_top; U, W at k = -1 and ny, V at k = -1, 0, ny).
 * dPdy_ddy_hi/lo: ny values each. */
int dnsp_update_vars(const double* P, double* P_per, double* P_now, double* P_old, double* U, double* V, double* W,
                     const double* U_hat, const double* V_hat, const double* W_hat, const double* rho,
                     const double* dPdy_ddy_hi, const double* dPdy_ddy_lo, const double* rho_min,
                     const double* wall_bc, int nx, int nz, int ny, int kmin_Vhat, double dtmax, double inv_dx,
                     double inv_dz, int use_incompressible, void* stream);

/* Geometry coefficients from the faces yu(0:ny) (src/pycode/calc_bands.py:272-277,
 * :336-343): inv_dy_p (ny), inv_dy_u (ny+1), inv_dy_p_pad (ny+1), dPdy_ddy_lo and
 * dPdy_ddy_hi (ny, as read at inc_set_alloc_arrays.f90:51-52).  Any output may be
 * NULL.  Host only. */
int dnsp_geom_coeffs(int ny, const double* yu, double* inv_dy_p, double* inv_dy_u, double* inv_dy_p_pad,
                     double* dPdy_ddy_lo, double* dPdy_ddy_hi);

/* ------------------------------------------------------------------------ */
/* Immersed-boundary forcing (src/diezdecomp_api_ibm.f90, src/ibm_setup/)     */
/* ------------------------------------------------------------------------ */
typedef struct dnsp_ibm dnsp_ibm;

/* Replaces diezdecomp_ibm_init (diezdecomp_api_ibm.f90:61-371) on one rank.
 * A field handled by exec is padded by nside = nhalo + nskip cells on every
 * side: (nx+2*nside)*(nz+2*nside)*(ny+2*nside) doubles, x fastest.
 * ijk: n records of 9 ints (target, source 1, source 2; each i, j, k),
 * wrapped mod (nx, nz, ny) here; c: n records of 3 doubles (cw, c1, c2), the
 * reference's all_c(0:2) order.  Repeated targets are rejected. */
int dnsp_ibm_create(int nx, int nz, int ny, int nhalo, int nskip, int n, const int* ijk, const double* c,
                    dnsp_ibm** out);

/* Replaces diezdecomp_ibm_exec (:374-445): A[t] = c1*A[s1] + c2*A[s2] + cw for
 * every record (all sources read first), then the new values copied into the
 * targets' periodic images inside the padded array. */
int dnsp_ibm_exec(dnsp_ibm* h, double* A, void* stream);

/* Records, halo images, and 1 when targets and sources are disjoint (single-kernel pass). */
int dnsp_ibm_info(const dnsp_ibm* h, int* n_forcings, int* n_images, int* disjoint);
int dnsp_ibm_free(dnsp_ibm* h);

/* ibm_read_coeffs (src/mod_utils.f90:181-212).  Host only.  With ijk == NULL
 * only the record count is returned in *n; else ijk (9*n ints) and c (3*n:
 * cw, c1, c2) are filled, indices wrapped mod (nx, nz, ny). */
int dnsp_ibm_read_coeffs(const char* path, int nx, int nz, int ny, int* n, int* ijk, double* c);

/* div-field extrapolation coefficients from the T records
 * (src/ibm_setup/set_ibm_bands.f90:57-77): c = (0, 1 + d_ta/d_ab, -d_ta/d_ab);
 * yp: the ny cell centres.  Host only. */
int dnsp_ibm_div_coeffs(int n, const int* ijk, const double* yp, int nx, int nz, int ny, double inv_dx,
                        double inv_dz, double* c);

/* Rough-wall surface: H(x,z) = sum_i a_i cos(t_i) + b_i sin(t_i),
 * t_i = 2 pi (m_i x/Lx + n_i z/Lz); the bottom wall is at H, the top wall at
 * Ly - H (Mod_Height_Function of the examples' rough_surf_define.f90; a_i,
 * b_i as that module holds them, i.e. after its halving of the file values). */
typedef struct {
    int n_modes;
    const double* mnab; /* 4 per mode: m, n, a, b */
    double Lx, Lz, Ly;
} dnsp_surface;

/* IBM coefficients of a rough wall (src/pycode/ibm_coeffs_engine.f90 and its
 * driver main_runner_ibm_coeffs_engine.f90).  x (nx), z (nz), y (ny): the
 * field component's coordinates over one surface tile (x and z uniform);
 * cc_x, cc_z: tile repetitions; quady = 0: ghost points with mirrored image
 * (make_rough_surface_ghost, apply_Fadlun), 1: quadratic through the wall
 * (make_rough_surface_quady); bc_bot/bc_top: the component's wall values.
 * Records are sorted by (k, j, i) of the target: ijk (9 per record: target,
 * source 1, source 2) and c (3: cw, c1, c2), as dnsp_ibm_create takes them.
 * Slabs (nx*cc_x by nz*cc_z, i fastest): bottom = highest bottom target per
 * column (-1 if none), top = lowest top target (ny if none).  Call with
 * ijk == NULL to get the record count in *n; then cap must be >= *n.
 * Host only. */
int dnsp_ibm_generate(const dnsp_surface* surf, int nx, int nz, int ny, const double* x, const double* z,
                      const double* y, int cc_x, int cc_z, int quady, double bc_bot, double bc_top, int cap, int* n,
                      int* ijk, double* c, int* slab_bot, int* slab_top);
/* Writers of the files ibm_read_coeffs and ibm_fill_slabs_2d read: 60-byte
 * records t(3 int32) c1 p1(3) c2 p2(3) cw after an int32 count; slabs: int32
 * nx, nz then nx*nz int32, i fastest.  Host only. */
int dnsp_ibm_write_coeffs(const char* path, int n, const int* ijk, const double* c);
int dnsp_ibm_write_slab(const char* path, int nx, int nz, const int* slab);

/* Rough-wall column slab (ibm_fill_slabs_2d, src/mod_utils.f90:215-239) minus
 * k_first, clipped as set_ibm_bands.f90:196-218 (bottom min(v, ny-1), top
 * max(v, 0)).  out: nx*nz ints, i fastest.  Host only. */
int dnsp_ibm_read_slab(const char* path, int nx, int nz, int ny, int k_first, int is_top, int* out);

/* ibm_part_2 for one field (src/ibm_setup/mod_cfd_ibm.f90:47-141): on the halo
 * array A, k in [0, bot(i,j)] <- wall_bot and k in [top(i,j), kmax] <- wall_top;
 * kmax = ny-1 for U, W, T and ny for V.  bot, top: device nx*nz ints. */
int dnsp_ibm_part2(double* A, const int* bot, const int* top, int nx, int nz, int ny, int kmax, double wall_bot,
                   double wall_top, void* stream);

/* ------------------------------------------------------------------------ */
/* On-disk formats (host only).                                              */
/* ------------------------------------------------------------------------ */
/* Field files, cfd_read_arr / cfd_write_arr (src/mod_io.f90:7-79, :81-150):
 * int32 nx, nz, ny then nx*nz*ny FP64 values, x fastest.  halo = 0: A is
 * (nx, nz, ny); halo = 1: A is a (nx+2)(nz+2)(ny+2) halo array and only its
 * interior is read or written.  A header that differs from (nx, nz, ny) is
 * DNSP_ERR_ARG ("Mismatch in grid params", the reference's error stop). */
int dnsp_io_field_shape(const char* path, int* nx, int* nz, int* ny);
int dnsp_io_read_field(const char* path, int nx, int nz, int ny, int halo, double* A);
int dnsp_io_write_field(const char* path, int nx, int nz, int ny, int halo, const double* A);
/* Coefficient text files: a count line, then rows of numbers
 * (calc_bands.py:358-364).  read_column: read_column_r8
 * (src/mod_utils.f90:14-35): out[0..n) = column icol (0..3) of rows
 * n_prev .. n_prev+n-1.  write_table: nrows x ncols, row-major. */
int dnsp_io_read_column(const char* path, int icol, int n_prev, int n, double* out);
int dnsp_io_write_table(const char* path, int nrows, int ncols, const double* rows);
/* yu.txt, one face per line (calc_bands.py:211): *ny = faces - 1; yu (may be
 * NULL to query the size) receives ny+1 values. */
int dnsp_io_read_faces(const char* path, int* ny, double* yu);

/* ------------------------------------------------------------------------ */
/* Projection step (no single counterpart: the reference's time loop,       */
/* src/main.f90, calls these routines one by one).  The object owns the      */
/* fields of one channel case on the device and enqueues, on one stream:     */
/* rho_min + the variable-density right-hand side (or the incompressible     */
/* one), the Poisson solve (FP64, or the mixed form with 0 or 1 refinement), */
/* update_vars (kmin_Vhat = 1), then on rough walls the opening of the next  */
/* build_vars on U, V, W (mod_cfd_subroutines.f90:98-113): where slabs are   */
/* set the column fill (ibm_part_2) and cfd_periodicity, then where handles  */
/* are set the forcing and wall planes of ibm_part_1 (mod_cfd_ibm.f90:10-45). */
/* ------------------------------------------------------------------------ */
typedef struct dnsp_step dnsp_step;

/* Field indices for dnsp_step_field: halo arrays (nx+2)(nz+2)(ny+2) except
 * DNSP_STEP_P, the (nx, nz, ny) pressure (right-hand side, then solution). */
#define DNSP_STEP_U_HAT    0
#define DNSP_STEP_V_HAT    1
#define DNSP_STEP_W_HAT    2
#define DNSP_STEP_U        3
#define DNSP_STEP_V        4
#define DNSP_STEP_W        5
#define DNSP_STEP_RHO      6
#define DNSP_STEP_P_NOW    7
#define DNSP_STEP_P_OLD    8
#define DNSP_STEP_DIV_NOW  9
#define DNSP_STEP_P_PER    10
#define DNSP_STEP_P        11
#define DNSP_STEP_NFIELDS  12

/* yu: the ny+1 wall-normal faces (bands and metric coefficients follow from
 * them); precision as for dnsp_poisson_init; fields start at zero. */
int dnsp_step_create(int nx, int nz, int ny, const double* yu, double inv_dx, double inv_dz, double dtmax,
                     int use_incompressible, int precision, dnsp_step** out);
/* Device address of a field (the caller fills the inputs). */
int dnsp_step_field(dnsp_step* s, int which, double** ptr);
/* Device-to-device copy of a whole field from src / to dst, on `stream`. */
int dnsp_step_write(dnsp_step* s, int which, const double* src, void* stream);
int dnsp_step_read(dnsp_step* s, int which, double* dst, void* stream);
/* Wall values U_bot, V_bot, W_bot, U_top, V_top, W_top (host array of 6). */
int dnsp_step_set_walls(dnsp_step* s, const double* wall_bc);
/* IBM forcing of component 0 (U), 1 (V), 2 (W): a handle created with
 * nhalo = 1, nskip = 0 (borrowed, must outlive the step); NULL removes it. */
int dnsp_step_set_ibm(dnsp_step* s, int component, dnsp_ibm* ibm);
/* Rough-wall column slabs (host nx*nz ints each, as dnsp_ibm_part2) of U,
 * V and W alike; NULL removes them. */
int dnsp_step_set_columns(dnsp_step* s, const int* bot, const int* top);
/* The same for one component 0 (U), 1 (V), 2 (W): the reference reads a
 * slab pair per component (rough_U/V/W_bot/top_inds, set_ibm_bands.f90:
 * 114-218) and ibm_part_2 fills each field from its own. */
int dnsp_step_set_component_columns(dnsp_step* s, int component, const int* bot, const int* top);
/* The embedded Poisson solver (for tuning; owned by the step). */
int dnsp_step_poisson(dnsp_step* s, dnsp_poisson** h);
/* Device bytes the step holds (fields, geometry, solver work arrays). */
int dnsp_step_footprint(const dnsp_step* s, size_t* bytes);
/* refinements: -1 = FP64 solve, 0 or 1 = mixed form (precision MIXED). */
int dnsp_step_enqueue(dnsp_step* s, int refinements, void* stream);
/* Records the same sequence into a graph (stream capture on a non-null
 * stream), kept in the step; dnsp_step_launch replays it. */
int dnsp_step_capture(dnsp_step* s, int refinements, void* stream);
int dnsp_step_launch(dnsp_step* s, void* stream);
int dnsp_step_free(dnsp_step* s);
/* Sizes and constants of a step (any output may be NULL). */
int dnsp_step_info(const dnsp_step* s, int* nx, int* nz, int* ny, double* inv_dx, double* inv_dz, double* dtmax,
                   int* use_incompressible);

/* ------------------------------------------------------------------------ */
/* Predictor: the time-step stages before the projection, in the order of    */
/* build_vars (src/mod_cfd_subroutines.f90:10-468): periodicity and walls of */
/* U, V, W, T (cfd_periodicity :860-896), div_now ghost planes (:120-144),   */
/* U_hat, V_hat, W_hat (:147-354), T (:367-407), periodicity of the hats and */
/* T, then for variable properties update_props (:776-857) and               */
/* build_rhs_energy (:618-745).  The bulk-pressure forcing                   */
/* (force_Sf_x_pressure, :184-237) is not included.  It works on a step's    */
/* U, V, W, hats, rho and div_now; dnsp_predictor_enqueue followed by         */
/* dnsp_step_enqueue is one time step (main.f90 time loop).                  */
/* ------------------------------------------------------------------------ */
typedef struct dnsp_predictor dnsp_predictor;

typedef struct {
    double nu, cond, Sf_x, Sf_z, Sq, Gb_x, Gb_y, Gb_z;
    double walls[8]; /* wall_BC_U/V/W/T_bot, then wall_BC_U/V/W/T_top */
    int kmin_Vhat;   /* first V_hat plane (1 on the bottom rank, main.f90:86-87) */
} dnsp_pred_params;

/* Predictor fields: halo arrays (nx+2)(nz+2)(ny+2) except the Adams-Bashforth
 * histories and RHS_energy (nx*nz*ny).  The property fields (DNSP_PRED_MU
 * and after) exist in the variable-property form only. */
#define DNSP_PRED_T          0
#define DNSP_PRED_RU_OLD     1
#define DNSP_PRED_RV_OLD     2
#define DNSP_PRED_RW_OLD     3
#define DNSP_PRED_RT_OLD     4
#define DNSP_PRED_MU         5
#define DNSP_PRED_COND       6
#define DNSP_PRED_BUOY       7
#define DNSP_PRED_CP         8
#define DNSP_PRED_ENTH       9
#define DNSP_PRED_DIVJAC     10
#define DNSP_PRED_T_EXT      11
#define DNSP_PRED_RHS_ENERGY 12
#define DNSP_PRED_NFIELDS    13

/* Property laws of T (the examples' expressions_rho_mu_cond_buoyancy_cp_div_jacobian.f90):
 * DNSP_LAW_CONST data = {c}; DNSP_LAW_POWER data = {c, p}: c T^p; DNSP_LAW_CUBIC
 * data = n segments of {hi, x0, c0, c1, c2, c3}, hi increasing (the last one
 * unused): T takes the first segment with T <= hi and
 * c0 + c1 (T-x0) + c2 (T-x0)^2 + c3 (T-x0)^3. */
#define DNSP_LAW_CONST 0
#define DNSP_LAW_POWER 1
#define DNSP_LAW_CUBIC 2
#define DNSP_Q_ENTH    0
#define DNSP_Q_RHO     1
#define DNSP_Q_MU      2
#define DNSP_Q_COND    3
#define DNSP_Q_CP      4
#define DNSP_Q_DRHO_DT 5

/* y-derivative coefficients of the predictor from the ny+1 faces
 * (calc_bands.py:212-215, 272-277, 285-320, 345-357), host only: out holds
 * Ruw_ddy (3 ny, a b c per row), Ruw_d2dy2 (3 ny), Rv_ddy (3 (ny+1)),
 * Rv_d2dy2 (3 (ny+1)), Rv_interp_uw bottom/top (2 (ny+1)), inv_dy_u (ny+1),
 * extrap_T_wall_0, _1, _m1, _m2: dnsp_predictor_geom_size(ny) doubles. */
size_t dnsp_predictor_geom_size(int ny);
int dnsp_predictor_geom(int ny, const double* yu, double* out);
/* Laws default to constants rho = cp = 1, mu = nu, cond = cond, enth = 0;
 * div_jacobian 0, buoyancy 0. */
int dnsp_predictor_create(dnsp_step* s, const double* yu, const dnsp_pred_params* p, dnsp_predictor** out);
int dnsp_predictor_field(dnsp_predictor* pr, int which, double** ptr, size_t* count);
/* Device-to-device copies of a whole field (count as dnsp_predictor_field). */
int dnsp_predictor_write(dnsp_predictor* pr, int which, const double* src, void* stream);
int dnsp_predictor_read(dnsp_predictor* pr, int which, double* dst, void* stream);
int dnsp_predictor_set_law(dnsp_predictor* pr, int quantity, int kind, const double* data, int n);
/* div_jacobian: formula != 0: -1/(rho^2 cp) drho_dT, else the constant;
 * buoyancy: linear != 0: b1 (rho - b0), else 0. */
int dnsp_predictor_set_derived(dnsp_predictor* pr, int divjac_formula, double divjac_const, int buoy_linear,
                               double b1, double b0);
/* Rough-wall column slabs (nx*nz ints, i fastest; both NULL: none) of U, V,
 * W and T alike.  ibm_part_2 fills the columns of each of U_hat, V_hat,
 * W_hat and T; with T's slabs (rough_T_bot/top_inds) T_ext takes the wall
 * values at k <= bot and k >= top and div_now is zero there. */
int dnsp_predictor_set_columns(dnsp_predictor* pr, const int* bot, const int* top);
/* The same for one component 0..3 = U, V, W, T (the reference's per-component
 * rough_<X>_bot/top_inds, set_ibm_bands.f90:114-218). */
int dnsp_predictor_set_component_columns(dnsp_predictor* pr, int component, const int* bot, const int* top);
/* Rough-wall forcing (borrowed handles, NULL: none) for component 0..4 =
 * U, V, W, T, div_now (the reference's obj_ibm_U/V/W/T/div).  With any
 * forcing or columns set the predictor follows build_vars' rough-wall
 * branch (use_rough_surf; quady ordering: dnsp_predictor_set_quady): ibm_part_1 (forcing of U, V, W, T,
 * wall planes, div_now forcing) instead of the opening periodicity,
 * ibm_part_2 on the hats and T after the predictor (each with its column
 * slabs), and the T forcing at the start of build_rhs_energy. */
int dnsp_predictor_set_ibm(dnsp_predictor* pr, int component, dnsp_ibm* h);
/* The quadratic-wall ordering of the rough branch (use_rough_surf_quady,
 * mod_cfd_subroutines.f90:98-108, :417-427): the opening fills the columns
 * of U, V, W, T and applies cfd_periodicity before ibm_part_1, and after
 * the predictor's column fill and cfd_periodicity the hats and T are forced
 * again (ibm_part_1, div_now forcing included), then T's x/z halos.
 * quady = 0 (default): the ghost-point ordering. */
int dnsp_predictor_set_quady(dnsp_predictor* pr, int quady);
int dnsp_predictor_enqueue(dnsp_predictor* pr, void* stream);
/* Records the predictor (and, with_step != 0, the step's sequence after
 * it: one full time step) into a graph; dnsp_predictor_launch replays it. */
int dnsp_predictor_capture(dnsp_predictor* pr, int with_step, int refinements, void* stream);
int dnsp_predictor_launch(dnsp_predictor* pr, void* stream);
int dnsp_predictor_free(dnsp_predictor* pr);

/* ------------------------------------------------------------------------ */
/* Serving engine (no counterpart in the reference): projection-step         */
/* requests for an ensemble of channel cases, continuously batched.          */
/* ------------------------------------------------------------------------ */
typedef struct dnsp_server dnsp_server;

/* One channel case.  Host arrays are copied.  IBM records per component
 * (U, V, W) as dnsp_ibm_create (n_ibm[k] = 0: none); bot/top: rough-wall
 * column slabs (nx*nz ints) or NULL.  A request's inputs are generated on
 * the device from (seed, request seed). */
typedef struct {
    int nx, nz, ny;
    const double* yu; /* ny+1 faces */
    double inv_dx, inv_dz, dtmax;
    int incompressible;
    double walls[6];
    int n_ibm[3];
    const int* ibm_ijk[3];
    const double* ibm_c[3];
    const int* bot;
    const int* top;
    unsigned long long seed;
} dnsp_case_desc;

/* hbm_budget: device bytes the resident cases may hold; max_slots (1..16):
 * requests in flight at once (one stream each); precision and refinements
 * as for dnsp_step_enqueue, or refinements = DNSP_SERVER_PICK_FORM
 * (precision MIXED): on its first admission each case times its step in
 * the FP64 form and in the mixed form with one refinement (both meet a
 * relative residual of 1e-10) and keeps the faster. */
#define DNSP_SERVER_PICK_FORM (-2)
int dnsp_server_create(size_t hbm_budget, int max_slots, int precision, int refinements, dnsp_server** out);
/* Registers a case (its device state is created on first admission). */
int dnsp_server_add_case(dnsp_server* sv, const dnsp_case_desc* d, int* case_id);
/* Device bytes the case holds while resident: an upper bound computed from
 * its description until its first admission, the bytes its step, solver,
 * IBM handles and tables hold after it (the admission test uses this). */
int dnsp_server_case_footprint(const dnsp_server* sv, int case_id, size_t* bytes);
/* The case's form (-1 FP64, 0/1 mixed refinements) and, when the server
 * picks forms, the step times (ms) it measured for FP64 and mixed (-1 before
 * the case's first admission).  Any output may be NULL. */
int dnsp_server_case_form(const dnsp_server* sv, int case_id, int* refinements, double* ms_fp64,
                          double* ms_mixed);
/* Replays n requests (case_id[i], seed[i]) arriving at arrival_s[i] seconds
 * after the call starts (non-decreasing), against the wall clock: waiting
 * requests are admitted into free slots in arrival order, skipping those
 * whose case is busy or does not fit the budget yet; a case is made
 * resident on first admission (least-recently-used idle cases evicted when
 * needed, and only when evicting every idle case would make room) and stays
 * resident otherwise.  Outputs per request: latency_ms
 * (completion - arrival), start_ms (admission, may be NULL), checksum[4]
 * (may be NULL: sum and index-weighted sum of P, then of U).  stats (may be
 * NULL, 4 doubles): admissions, evictions, peak resident bytes, wall s.
 * Synchronous. */
int dnsp_server_run(dnsp_server* sv, int n, const int* case_id, const double* arrival_s,
                    const unsigned long long* seed, double* latency_ms, double* start_ms, double* checksum,
                    double* stats);
int dnsp_server_free(dnsp_server* sv);

#ifdef __cplusplus
}
#endif

#endif /* DNSPROJ_H */
