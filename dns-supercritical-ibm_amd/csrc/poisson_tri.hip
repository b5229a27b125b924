// Instantiations and launcher of the partitioned tridiagonal kernel
// (kt_tri_part, poisson_ct.h) for FP64 and FP32 spectra.
#include <atomic>

#include "poisson_ct.h"
#include "poisson_ct_launch.h"

namespace dnsp {

#define DNSP_PART_ROWS(X) X(2) X(3) X(4) X(6) X(8) X(12) X(16) X(24) X(32)

static int rows_for(int ny, int chunks) {
    for (int mr : {2, 3, 4, 6, 8, 12, 16, 24, 32})
        if (mr * chunks >= ny) return mr;
    return 0;
}

int part_rows_for(int ny) { return rows_for(ny, kPartChunks); }

int part_group_modes(int variant, int ny, bool fp32) {
    switch (variant) {
        case 1: case 5: return 8;
        case 3: case 7: return fp32 ? (ny <= 16 * 32 ? 32 : 16) : (ny <= 16 * 32 ? 16 : 8);
        default: return 16;
    }
}

template <int MR, int P, int M, typename T, int OCC, bool TP, typename Ar, bool MM>
static void part_launch_mr(T* S, const double* ax, const double* az, const double* bands, int nx, int nz, int ny,
                           double pcr_min, const double* msum, double* pbar, double inv_nxnz, hipStream_t st,
                           const int* glist, int nlist, T* So) {
    const long plane = (long)nx * nz;
    // every group, plus the group of modes 1..M-1; or the listed ones
    const int ngroups = glist ? nlist : (int)((plane + M - 1) / M) + 1;
    if (ngroups <= 0) return;
    const size_t lds = 3 * sizeof(Ar) * MR * P;
    kt_tri_part<MR, P, M, T, OCC, TP, Ar, MM><<<ngroups, M * P, lds, st>>>(S, ax, az, bands, nx, ny, plane, pcr_min,
                                                                        msum, pbar, inv_nxnz, glist, nz, So);
}

// P chunks per mode; rows per chunk from ny (at most 32 rows, ny <= 32 * P)
template <int P, int M, typename T, int OCC, bool TP, typename Ar, bool MM>
static void part_launch_pm(T* S, const double* ax, const double* az, const double* bands, int nx, int nz, int ny,
                           double pcr_min, const double* msum, double* pbar, double inv_nxnz, hipStream_t st,
                           const int* glist, int nlist, T* So) {
    switch (rows_for(ny, P)) {
#define DNSP_CASE(MR)                                                                                          \
    case MR:                                                                                                   \
        part_launch_mr<MR, P, M, T, OCC, TP, Ar, MM>(S, ax, az, bands, nx, nz, ny, pcr_min, msum, pbar, inv_nxnz, st, \
                                                 glist, nlist, So);                                            \
        break;
        DNSP_PART_ROWS(DNSP_CASE)
#undef DNSP_CASE
        default: break;
    }
}

template <typename T, typename Ar, bool MM>
void part_launch(int variant, T* S, const double* ax, const double* az, const double* bands, int nx, int nz,
                 int ny, double pcr_min, const double* msum, double* pbar, double inv_nxnz, hipStream_t st,
                 const int* glist, int nlist, T* So) {
#define DNSP_PV(P, M, OCC, TP)                                                                                    \
    part_launch_pm<P, M, T, OCC, TP, Ar, MM>(S, ax, az, bands, nx, nz, ny, pcr_min, msum, pbar, inv_nxnz, st, glist, \
                                             nlist, So)
    if constexpr (sizeof(T) == 4) {
        // FP32 spectra: (16, 32) in place of (64, 8), whose 8-mode FP32 rows
        // are 32-byte segments (32 modes make a 128-byte row).  16 chunks of
        // at most 32 rows hold ny <= 512; above that (ny <= 32 * kPartChunks)
        // the slot runs (32, 16).
        const bool c16 = ny <= 16 * 32;
        switch (variant) {
            case 0: DNSP_PV(32, 16, 4, false); break;
            case 1: DNSP_PV(32, 8, 1, false); break;
            case 2: DNSP_PV(48, 16, 1, false); break;
            case 3:
                if (c16) DNSP_PV(16, 32, 1, false);
                else DNSP_PV(32, 16, 4, false);
                break;
            case 4: DNSP_PV(32, 16, 4, true); break;
            case 5: DNSP_PV(32, 8, 1, true); break;
            case 6: DNSP_PV(48, 16, 1, true); break;
            default:
                if (c16) DNSP_PV(16, 32, 1, true);
                else DNSP_PV(32, 16, 4, true);
                break;
        }
        return;
    }
    // FP64: 3 and 7 are (16, 16) for ny <= 512 (16 chunks of twice the
    // rows, one reduction step fewer), else (64, 8)
    const bool c16 = ny <= 16 * 32;
    switch (variant) {
        case 0: DNSP_PV(32, 16, 4, false); break;
        case 1: DNSP_PV(32, 8, 1, false); break;
        case 2: DNSP_PV(48, 16, 1, false); break;
        case 3:
            if (c16) DNSP_PV(16, 16, 1, false);
            else DNSP_PV(64, 8, 1, false);
            break;
        case 4: DNSP_PV(32, 16, 4, true); break;
        case 5: DNSP_PV(32, 8, 1, true); break;
        case 6: DNSP_PV(48, 16, 1, true); break;
        default:
            if (c16) DNSP_PV(16, 16, 1, true);
            else DNSP_PV(64, 8, 1, true);
            break;
    }
#undef DNSP_PV
}

#define DNSP_PART_INST(T, AR, MM)                                                                          \
    template void part_launch<T, AR, MM>(int, T*, const double*, const double*, const double*, int, int, int, \
                                         double, const double*, double*, double, hipStream_t, const int*, int, T*);
DNSP_PART_INST(double, double, false)
DNSP_PART_INST(double, double, true)
DNSP_PART_INST(float, double, false)
DNSP_PART_INST(float, float, false)
#undef DNSP_PART_INST

}  // namespace dnsp
