// Launchers of the compile-time-sized kernels (poisson_ct_x.hip,
// poisson_ct_z.hip, poisson_tri.hip, poisson_mixed_x.hip).  FFT `variant` in
// [0, kCtVariants) selects the (tile, threads) pair of ct_xvar_tile /
// ct_xvar_threads (x) or ct_zvar_tile / ct_zvar_threads (z).
#pragma once

#include <hip/hip_runtime.h>

namespace dnsp {

bool ct_fft_supported(int n);
void ct_x_launch(int n, int variant, bool inv, const double* src, double* dst, long nlines, const double2* tw,
                 double scale, hipStream_t st);
void ct_z_launch(int n, int variant, bool inv, double* S, int nx, int ny, const double2* tw, hipStream_t st);
// The blocked FP64 solve (poisson_ct.h): the x forward transform from lines
// to the x-blocked layout (variants as ct_x_launch); the z transforms from
// x-blocked to mode-major (forward) and from x-blocked to halfcomplex lines
// (inverse), variant 0: 256 threads, 1: 512.
bool xb_supported(int n);
void xb_x_launch(int n, int variant, const double* src, double* dst, long nlines, int nz, const double2* tw,
                 hipStream_t st);
void mm_z_launch(int n, int variant, bool inv, const double* src, double* dst, int nx, int ny, const double2* tw,
                 hipStream_t st);
// FP32 z transforms; the forward one also turns the FP64 line sums lsum (if
// not NULL) into the FP64 plane sums msum.
void ct_z_launch(int n, int variant, bool inv, float* S, int nx, int ny, const float2* tw, const double* lsum,
                 double* msum, hipStream_t st);

// Partitioned tridiagonal solve, ny <= 32 * kPartChunks.  Variants (P chunks
// per mode of up to 32 rows, M modes per workgroup): 0 = P 32, M 16, registers
// capped for 4 resident waves per SIMD; 1 = P 32, M 8; 2 = P 64, M 4;
// 3 = P 64, M 8.
constexpr int kPartChunks = 32;
constexpr int kPartVariants = 8;
// tuner table: first slot of the x, z and tridiagonal variants of one form
constexpr int kTuneSlot[3] = {0, 4, 8};
constexpr int kTuneFormSlots = 16;
int part_rows_for(int ny);  // rows per chunk at kPartChunks chunks (0: ny too large)
// modes per workgroup (M) of a tridiagonal variant, FP32 or FP64 spectra
int part_group_modes(int variant, int ny, bool fp32);
// Ar: arithmetic type (double, or float for FP32 spectra); glist (device,
// nlist entries): solve only the listed mode groups (NULL: all of them);
// MM: S in the mode-major blocks of the blocked solve, the solution written
// to So in its x-blocked layout.
template <typename T, typename Ar = double, bool MM = false>
void part_launch(int variant, T* S, const double* ax, const double* az, const double* bands, int nx, int nz,
                 int ny, double pcr_min, const double* msum, double* pbar, double inv_nxnz, hipStream_t st,
                 const int* glist = nullptr, int nlist = 0, T* So = nullptr);

// Mixed-precision x kernels (poisson_mixed.h).
enum MxKind { MX_FWD = 0, MX_RES_F32 = 1, MX_RES_F64 = 2, MX_BWD = 3, MX_ACC_F32 = 4, MX_ACC_F64 = 5 };
struct MxArgs {
    const void* src = nullptr;      // f (FWD), iterate p (RES), FP32 spectrum (BWD, ACC)
    const double* f = nullptr;      // RES: right-hand side
    const double* q = nullptr;      // RES: per-plane constants
    const double* bands = nullptr;  // RES: plain a, b, c
    const void* base = nullptr;     // ACC: FP32 field or FP64 iterate (NULL = 0)
    const double* pbar0 = nullptr;  // ACC
    const double* pbar1 = nullptr;  // ACC
    void* dst = nullptr;            // FP32 spectrum (FWD, RES), FP32 field (BWD), FP64 P (ACC)
    double* lsum = nullptr;         // FP64 line sums of the input (FWD), of r (RES), of the output
    int nz = 0, ny = 0;
    double ix2 = 0, iz2 = 0;
    long nlines = 0;
    const float2* tw = nullptr;
    double scale = 1.0;
};
void mx_x_launch(int kind, int n, int variant, const MxArgs& a, hipStream_t st);

}  // namespace dnsp
