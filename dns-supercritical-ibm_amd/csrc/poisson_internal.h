// The Poisson solver handle, shared by poisson.hip (FP64 form, tuner) and
// poisson_mixed.hip (mixed-precision form).
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"
#include "fft_lds.h"

struct dnsp_poisson {
    int nx = 0, nz = 0, ny = 0;
    double inv_dx = 0, inv_dz = 0;
    int precision = 0;
    int device = 0;
    dnsp::FftPlan px{}, pz{};
    int Gx = 1, TAz = 1, ncol = 0;
    double* d_ax = nullptr;
    double* d_az = nullptr;
    int* d_is0x = nullptr;
    int* d_is0z = nullptr;
    double* d_bands = nullptr;  // a, b, c, a00, b00, c00 (ny each)
    double2* d_twx = nullptr;
    double2* d_twz = nullptr;
    double* d_cp = nullptr;    // serial Thomas c' workspace, nx*nz*ny (allocated when used)
    int fft_strategy = DNSP_FFT_AUTO;
    int tri_strategy = DNSP_TRIDIAG_AUTO;
    bool ct_x = false, ct_z = false;  // compile-time FFT kernels in use
    int xvar = 0, zvar = 0;           // their (tile, threads) variants, forward
    int xvar_b = 0, zvar_b = 0;       // and inverse transforms
    // blocked FP64 solve (x-blocked and mode-major layouts through d_W)
    bool blk_ok = false;              // supported sizes (nx % 16 == 0, compile-time FFTs, partitioned)
    bool blk_on = true;               // used by dnsp_poisson_run / solve when supported
    double* d_W = nullptr;            // its workspace, nx*nz*ny doubles
    int bxvar = 0, bxvar_b = 2;       // its x variants (forward, inverse)
    int bzvar = 1, bzvar_b = 1;       // its z variants (0: 256, 1: 512 threads)
    int btvar = 4;                    // its tridiagonal variant
    int tvar = 1;                     // partitioned tridiagonal variant
    int part_rows = 0;                // partitioned tridiagonal rows per chunk (0 = serial)
    double* d_host = nullptr;  // staging for dnsp_poisson_run_host
    double min_pivot = 0;
    // modes with |lambda| below this solve their reduced system serially
    double pcr_min = 0;

    // form dnsp_poisson_run uses: -1 = FP64, >= 0 = mixed with that many refinements
    int form = -1;
    // mixed-precision form (precision == DNSP_PRECISION_MIXED)
    int mxvar = 0, mzvar = 0, mtvar = 0;  // its FFT and tridiagonal variants
    int mxv[6] = {0, 0, 0, 0, 0, 0};      // x variant per kernel kind (MxKind), picked by the tuner
    int mrows = 0;                         // its partitioned rows per chunk
    float2* d_twx32 = nullptr;
    float2* d_twz32 = nullptr;
    float* d_A = nullptr;       // FP32 field: sweep-0 spectrum, then the sweep-0 field
    float* d_B = nullptr;       // FP32 field: refinement spectra
    double* d_lsA = nullptr;    // FP64 line sums (nz*ny)
    double* d_lsB = nullptr;
    double* d_small = nullptr;  // msum, pbar0, pbar1, q (ny each), residual partials
    double* d_fcopy = nullptr;  // copy of an in-place right-hand side (allocated when used)
    // mode groups of the mixed tridiagonal for M = 8, 16, 32 modes per group:
    // list [m][0] (FP64 arithmetic: a mode with |lambda| < f32_min, mode 0
    // and the group of modes 1..M-1) and [m][1] (FP32 arithmetic), in d_glist
    int* d_glist = nullptr;
    int gl_off[3][2] = {}, gl_n[3][2] = {};
    size_t gl_total = 0;    // ints in d_glist
    double f32_min = 0;     // |lambda| from which a mode's sweeps run in FP32
    bool f32_sweeps = true; // FP32 arithmetic for the well-conditioned groups
};

namespace dnsp {

// Residual-norm reduction: fixed grid, two partial sums per workgroup.
constexpr int kResBlocks = 1024;

// Allocates and fills the mixed-precision state (poisson_mixed.hip).
int mixed_setup(dnsp_poisson* h);
void mixed_release(dnsp_poisson* h);
// Times the mixed-precision kernel variants and keeps the fastest.
int mixed_autotune(dnsp_poisson* h, double* f, int reps, hipStream_t st, double* times_ms);

}  // namespace dnsp
