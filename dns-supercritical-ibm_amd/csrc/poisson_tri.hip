// Instantiations and launcher of the partitioned tridiagonal kernel
// (kt_tri_part, poisson_ct.h) for FP64 and FP32 spectra.
#include <atomic>

#include "poisson_ct.h"
#include "poisson_ct_launch.h"

namespace dnsp {

#define DNSP_PART_ROWS(X) X(2) X(3) X(4) X(6) X(8) X(12) X(16) X(24) X(32)

int part_rows_for(int ny) {
    for (int mr : {2, 3, 4, 6, 8, 12, 16, 24, 32})
        if (mr * kPartChunks >= ny) return mr;
    return 0;
}

static int mr_slot(int mr) {
    switch (mr) {
        case 2: return 0;
        case 3: return 1;
        case 4: return 2;
        case 6: return 3;
        case 8: return 4;
        case 12: return 5;
        case 16: return 6;
        case 24: return 7;
        default: return 8;
    }
}

// Resident workgroups per CU of the persistent (8-mode) instances, by storage
// type and rows per chunk (0 = not queried yet).
static std::atomic<int> g_occ[2][9];

template <int MR, int M, typename T, bool LOOP, int OCC>
static void part_launch_mr(T* S, const double* ax, const double* az, const int* i0x,
                           const int* i0z, const double* bands, int nx, int nz, int ny, double pcr_min, int ncu,
                           const double* msum, double* pbar, double inv_nxnz, hipStream_t st) {
    const long plane = (long)nx * nz;
    const int ngroups = (int)((plane + M - 1) / M);
    const size_t lds = 6 * sizeof(double) * ny;
    auto kern = kt_tri_part<MR, kPartChunks, M, T, LOOP, OCC>;
    int grid = ngroups;
    if (LOOP) {
        std::atomic<int>& slot = g_occ[sizeof(T) == 4 ? 1 : 0][mr_slot(MR)];
        int per = slot.load(std::memory_order_relaxed);
        if (per <= 0) {
            // sized for the largest ny of this MR, so one query serves every size
            const size_t lds_max = 6 * sizeof(double) * (size_t)(MR * kPartChunks);
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(kern),
                                                             M * kPartChunks, lds_max) != hipSuccess ||
                per <= 0)
                per = 1;
            slot.store(per, std::memory_order_relaxed);
        }
        const long want = (long)per * (ncu > 0 ? ncu : 256);
        if (want < grid) grid = (int)want;
        if (grid < 8) grid = 8;  // the kernel spreads groups over all 8 XCDs
    }
    kern<<<grid, M * kPartChunks, lds, st>>>(S, ax, az, i0x, i0z, bands, nx, ny, plane, pcr_min, ngroups, msum,
                                             pbar, inv_nxnz);
}

template <int M, typename T, bool LOOP, int OCC>
static void part_launch_m(int mr, T* S, const double* ax, const double* az, const int* i0x,
                          const int* i0z, const double* bands, int nx, int nz, int ny, double pcr_min, int ncu,
                          const double* msum, double* pbar, double inv_nxnz, hipStream_t st) {
    switch (mr) {
#define DNSP_CASE(MR)                                                                                       \
    case MR:                                                                                                \
        part_launch_mr<MR, M, T, LOOP, OCC>(S, ax, az, i0x, i0z, bands, nx, nz, ny, pcr_min, ncu, msum, pbar, \
                                 inv_nxnz, st);                                                             \
        break;
        DNSP_PART_ROWS(DNSP_CASE)
#undef DNSP_CASE
        default: break;
    }
}

template <typename T>
void part_launch(int variant, int mr, T* S, const double* ax, const double* az, const int* i0x, const int* i0z,
                 const double* bands, int nx, int nz, int ny, double pcr_min, int ncu, const double* msum,
                 double* pbar, double inv_nxnz, hipStream_t st) {
    switch (variant) {
        case 0:
            part_launch_m<16, T, false, 4>(mr, S, ax, az, i0x, i0z, bands, nx, nz, ny, pcr_min, ncu, msum, pbar,
                                        inv_nxnz, st);
            break;
        case 1:
            part_launch_m<8, T, false, 1>(mr, S, ax, az, i0x, i0z, bands, nx, nz, ny, pcr_min, ncu, msum, pbar,
                                       inv_nxnz, st);
            break;
        case 2:
            part_launch_m<8, T, false, 4>(mr, S, ax, az, i0x, i0z, bands, nx, nz, ny, pcr_min, ncu, msum, pbar,
                                       inv_nxnz, st);
            break;
        default:
            part_launch_m<8, T, true, 1>(mr, S, ax, az, i0x, i0z, bands, nx, nz, ny, pcr_min, ncu, msum, pbar,
                                      inv_nxnz, st);
            break;
    }
}

template void part_launch<double>(int, int, double*, const double*, const double*, const int*, const int*,
                                  const double*, int, int, int, double, int, const double*, double*, double,
                                  hipStream_t);
template void part_launch<float>(int, int, float*, const double*, const double*, const int*, const int*,
                                 const double*, int, int, int, double, int, const double*, double*, double,
                                 hipStream_t);

}  // namespace dnsp
