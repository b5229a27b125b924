// Instantiations of the compile-time-sized z-direction FFT kernels (FP64 and
// FP32 spectra).
#include "poisson_ct.h"
#include "poisson_ct_launch.h"

namespace dnsp {

#define DNSP_CT_SIZES(X) X(32) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768)

template <int N, int V, typename T>
static void ct_z(bool inv, T* S, int nx, int ny, const typename Cx<T>::V* tw, const double* lsum, double* msum,
                 hipStream_t st) {
    constexpr int TA = ct_cols<N, ct_zvar_tile(V)>();
    constexpr int NT = ct_zvar_threads(V);
    const int ncol = (nx + 1) / 2;  // complex columns = pairs of adjacent x rows
    const int ntc = (ncol + TA - 1) / TA;
    const int nb = ntc * ny;
    if (!inv) kt_zfwd<N, TA, NT, T><<<nb, NT, 0, st>>>(S, nx, ncol, ntc, tw, lsum, msum);
    else      kt_zbwd<N, TA, NT, T><<<nb, NT, 0, st>>>(S, nx, ncol, ntc, tw);
}

template <int N, typename T>
static void ct_z_v(int v, bool inv, T* S, int nx, int ny, const typename Cx<T>::V* tw, const double* lsum,
                   double* msum, hipStream_t st) {
    switch (v) {
        case 0: ct_z<N, 0, T>(inv, S, nx, ny, tw, lsum, msum, st); break;
        case 1: ct_z<N, 1, T>(inv, S, nx, ny, tw, lsum, msum, st); break;
        case 2: ct_z<N, 2, T>(inv, S, nx, ny, tw, lsum, msum, st); break;
        default: ct_z<N, 3, T>(inv, S, nx, ny, tw, lsum, msum, st); break;
    }
}

template <int N, int NT>
static void mm_z(bool inv, const double* src, double* dst, int nx, int ny, const double2* tw, hipStream_t st) {
    const int nb = (nx / 16) * ny;
    if (!inv) kt_zfwd_mm<N, NT><<<nb, NT, 0, st>>>(src, dst, nx, ny, tw);
    else      kt_zbwd_xl<N, NT><<<nb, NT, 0, st>>>(src, dst, nx, ny, tw);
}

void mm_z_launch(int n, int variant, bool inv, const double* src, double* dst, int nx, int ny, const double2* tw,
                 hipStream_t st) {
    switch (n) {
#define DNSP_CASE(N)                                                     \
    case N:                                                              \
        if (variant == 0) mm_z<N, 256>(inv, src, dst, nx, ny, tw, st);   \
        else mm_z<N, 512>(inv, src, dst, nx, ny, tw, st);                \
        break;
        DNSP_CT_SIZES(DNSP_CASE)
#undef DNSP_CASE
        default: break;
    }
}

void ct_z_launch(int n, int variant, bool inv, double* S, int nx, int ny, const double2* tw, hipStream_t st) {
    switch (n) {
#define DNSP_CASE(N) case N: ct_z_v<N, double>(variant, inv, S, nx, ny, tw, nullptr, nullptr, st); break;
        DNSP_CT_SIZES(DNSP_CASE)
#undef DNSP_CASE
        default: break;
    }
}

void ct_z_launch(int n, int variant, bool inv, float* S, int nx, int ny, const float2* tw, const double* lsum,
                 double* msum, hipStream_t st) {
    switch (n) {
#define DNSP_CASE(N) case N: ct_z_v<N, float>(variant, inv, S, nx, ny, tw, lsum, msum, st); break;
        DNSP_CT_SIZES(DNSP_CASE)
#undef DNSP_CASE
        default: break;
    }
}

}  // namespace dnsp
