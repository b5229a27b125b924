// Launchers of the compile-time-sized FFT kernels (poisson_ct_x.hip,
// poisson_ct_z.hip).  variant in [0, kCtVariants) selects the (tile, threads)
// pair of ct_var_tile / ct_var_threads.
#pragma once

#include <hip/hip_runtime.h>

namespace dnsp {

bool ct_fft_supported(int n);
void ct_x_launch(int n, int variant, bool inv, const double* src, double* dst, long nlines, const double2* tw,
                 double scale, hipStream_t st);
void ct_z_launch(int n, int variant, bool inv, double* S, int nx, int ny, const double2* tw, hipStream_t st);

}  // namespace dnsp
