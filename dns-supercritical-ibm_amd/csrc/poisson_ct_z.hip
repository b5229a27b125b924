// Instantiations of the compile-time-sized z-direction FFT kernels.
#include "poisson_ct.h"
#include "poisson_ct_launch.h"

namespace dnsp {

#define DNSP_CT_SIZES(X) X(32) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768)

template <int N, int V>
static void ct_z(bool inv, double* S, int nx, int ny, const double2* tw, hipStream_t st) {
    constexpr int TA = ct_cols<N, ct_var_tile(V)>();
    constexpr int NT = ct_var_threads(V);
    const int ncol = (nx + 1) / 2;  // complex columns = pairs of adjacent x rows
    const int ntc = (ncol + TA - 1) / TA;
    const int nb = ntc * ny;
    if (!inv) kt_zfwd<N, TA, NT><<<nb, NT, 0, st>>>(S, nx, ncol, ntc, tw);
    else      kt_zbwd<N, TA, NT><<<nb, NT, 0, st>>>(S, nx, ncol, ntc, tw);
}

template <int N>
static void ct_z_v(int v, bool inv, double* S, int nx, int ny, const double2* tw, hipStream_t st) {
    switch (v) {
        case 0: ct_z<N, 0>(inv, S, nx, ny, tw, st); break;
        case 1: ct_z<N, 1>(inv, S, nx, ny, tw, st); break;
        case 2: ct_z<N, 2>(inv, S, nx, ny, tw, st); break;
        default: ct_z<N, 3>(inv, S, nx, ny, tw, st); break;
    }
}

void ct_z_launch(int n, int variant, bool inv, double* S, int nx, int ny, const double2* tw, hipStream_t st) {
    switch (n) {
#define DNSP_CASE(N) case N: ct_z_v<N>(variant, inv, S, nx, ny, tw, st); break;
        DNSP_CT_SIZES(DNSP_CASE)
#undef DNSP_CASE
        default: break;
    }
}

}  // namespace dnsp
