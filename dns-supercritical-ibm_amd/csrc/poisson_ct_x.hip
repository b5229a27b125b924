// Instantiations of the compile-time-sized x-direction FFT kernels.
#include "poisson_ct.h"
#include "poisson_ct_launch.h"

namespace dnsp {

#define DNSP_CT_SIZES(X) X(32) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768)

bool ct_fft_supported(int n) {
    switch (n) {
#define DNSP_CASE(N) case N:
        DNSP_CT_SIZES(DNSP_CASE)
#undef DNSP_CASE
        return true;
        default: return false;
    }
}

template <int N, int V>
static void ct_x(bool inv, const double* src, double* dst, long nlines, const double2* tw, double scale,
                 hipStream_t st) {
    constexpr int G = ct_lines<N, ct_xvar_tile(V)>();
    constexpr int NT = ct_xvar_threads(V);
    const int nb = (int)((nlines + 2 * G - 1) / (2 * G));
    if (!inv) kt_xfwd<N, G, NT><<<nb, NT, 0, st>>>(src, dst, nlines, tw);
    else      kt_xbwd<N, G, NT><<<nb, NT, 0, st>>>(src, dst, nlines, tw, scale);
}

template <int N, int V>
static void ct_x_xb(const double* src, double* dst, long nlines, int nz, const double2* tw, hipStream_t st) {
    constexpr int G = ct_lines<N, ct_xvar_tile(V)>();
    constexpr int NT = ct_xvar_threads(V);
    const int nb = (int)((nlines + 2 * G - 1) / (2 * G));
    kt_xfwd_xb<N, G, NT><<<nb, NT, 0, st>>>(src, dst, nlines, nz, tw);
}

template <int N>
static void ct_x_v(int v, bool inv, const double* src, double* dst, long nlines, const double2* tw,
                   double scale, hipStream_t st) {
    switch (v) {
        case 0: ct_x<N, 0>(inv, src, dst, nlines, tw, scale, st); break;
        case 1: ct_x<N, 1>(inv, src, dst, nlines, tw, scale, st); break;
        case 2: ct_x<N, 2>(inv, src, dst, nlines, tw, scale, st); break;
        default: ct_x<N, 3>(inv, src, dst, nlines, tw, scale, st); break;
    }
}

#define DNSP_XB_SIZES(X) X(32) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768)

bool xb_supported(int n) { return ct_fft_supported(n) && n % 16 == 0; }

void xb_x_launch(int n, int variant, const double* src, double* dst, long nlines, int nz, const double2* tw,
                 hipStream_t st) {
    switch (n) {
#define DNSP_CASE(N)                                                         \
    case N:                                                                  \
        switch (variant) {                                                   \
            case 0: ct_x_xb<N, 0>(src, dst, nlines, nz, tw, st); break;      \
            case 1: ct_x_xb<N, 1>(src, dst, nlines, nz, tw, st); break;      \
            case 2: ct_x_xb<N, 2>(src, dst, nlines, nz, tw, st); break;      \
            default: ct_x_xb<N, 3>(src, dst, nlines, nz, tw, st); break;     \
        }                                                                    \
        break;
        DNSP_XB_SIZES(DNSP_CASE)
#undef DNSP_CASE
        default: break;
    }
}

void ct_x_launch(int n, int variant, bool inv, const double* src, double* dst, long nlines, const double2* tw,
                 double scale, hipStream_t st) {
    switch (n) {
#define DNSP_CASE(N) case N: ct_x_v<N>(variant, inv, src, dst, nlines, tw, scale, st); break;
        DNSP_CT_SIZES(DNSP_CASE)
#undef DNSP_CASE
        default: break;
    }
}

}  // namespace dnsp
