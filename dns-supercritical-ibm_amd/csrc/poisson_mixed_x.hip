// Instantiations of the mixed-precision x-direction kernels (poisson_mixed.h).
// The kernels on an FP64 iterate (second and later refinement sweeps) exist
// in FFT variant 0 only.
#include "poisson_ct_launch.h"
#include "poisson_mixed.h"

namespace dnsp {

#define DNSP_CT_SIZES(X) X(32) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768)

template <int N, int V>
static void mx_x(int kind, const MxArgs& a, hipStream_t st) {
    constexpr int G = ct_lines<N, ct_xvar_tile(V)>();
    constexpr int NT = ct_xvar_threads(V);
    const int nb = (int)((a.nlines + 2 * G - 1) / (2 * G));
    switch (kind) {
        case MX_FWD:
            kt_xfwd_mf<N, G, NT><<<nb, NT, 0, st>>>(static_cast<const double*>(a.src), static_cast<float*>(a.dst),
                                                    a.lsum, a.nlines, a.tw);
            break;
        case MX_RES_F32:
            kt_xfwd_mr<N, G, NT, float><<<nb, NT, 0, st>>>(static_cast<const float*>(a.src), a.f, a.q, a.bands,
                                                           static_cast<float*>(a.dst), a.lsum, a.nz, a.ny, a.ix2,
                                                           a.iz2, a.nlines, a.tw);
            break;
        case MX_BWD:
            kt_xbwd_mf<N, G, NT><<<nb, NT, 0, st>>>(static_cast<const float*>(a.src), static_cast<float*>(a.dst),
                                                    a.lsum, a.nlines, a.tw, (float)a.scale);
            break;
        case MX_ACC_F32:
            kt_xbwd_ma<N, G, NT, float><<<nb, NT, 0, st>>>(static_cast<const float*>(a.src),
                                                           static_cast<const float*>(a.base), a.pbar0, a.pbar1,
                                                           static_cast<double*>(a.dst), a.lsum, a.nz, a.nlines,
                                                           a.tw, a.scale);
            break;
        default:
            if constexpr (V == 0) {
                if (kind == MX_RES_F64)
                    kt_xfwd_mr<N, G, NT, double><<<nb, NT, 0, st>>>(
                        static_cast<const double*>(a.src), a.f, a.q, a.bands, static_cast<float*>(a.dst), a.lsum,
                        a.nz, a.ny, a.ix2, a.iz2, a.nlines, a.tw);
                else
                    kt_xbwd_ma<N, G, NT, double><<<nb, NT, 0, st>>>(
                        static_cast<const float*>(a.src), static_cast<const double*>(a.base), a.pbar0, a.pbar1,
                        static_cast<double*>(a.dst), a.lsum, a.nz, a.nlines, a.tw, a.scale);
            }
            break;
    }
}

template <int N>
static void mx_x_v(int kind, int v, const MxArgs& a, hipStream_t st) {
    if (kind == MX_RES_F64 || kind == MX_ACC_F64) v = 0;
    switch (v) {
        case 0: mx_x<N, 0>(kind, a, st); break;
        case 1: mx_x<N, 1>(kind, a, st); break;
        case 2: mx_x<N, 2>(kind, a, st); break;
        default: mx_x<N, 3>(kind, a, st); break;
    }
}

void mx_x_launch(int kind, int n, int variant, const MxArgs& a, hipStream_t st) {
    switch (n) {
#define DNSP_CASE(N) case N: mx_x_v<N>(kind, variant, a, st); break;
        DNSP_CT_SIZES(DNSP_CASE)
#undef DNSP_CASE
        default: break;
    }
}

}  // namespace dnsp
