// Immersed-boundary forcing on one GPU (FP64, gfx950).
//
// The reference's diezdecomp_ibm_exec (src/diezdecomp_api_ibm.f90:374-445)
// runs two passes over a padded field A (interior n plus nhalo+nskip cells on
// every side):
//   fluids:  every target t gets A[t] = c1*A[s1] + c2*A[s2] + cw, with all
//            sources read before any target is written (pack, then apply);
//   forcing: a target with a periodic image in the halo shell copies its new
//            value into every periodic image inside the padded array.
// On one rank the MPI exchange is a local gather, so the handle keeps the
// flat padded positions of targets, sources and images, built on the host
// (the index maps of :109-371), and exec is one or two gathers plus one copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <new>
#include <unordered_set>
#include <vector>

#include "common.h"

namespace dnsp {

constexpr int kIbmThreads = 256;

// fluids pass when targets and sources are disjoint: one gather-apply kernel
__global__ void k_ibm_apply(double* A, const long long* __restrict__ pos, const double* __restrict__ c, int n) {
#pragma clang fp contract(off)
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double v = c[3 * r + 1] * A[pos[3 * r + 1]] + c[3 * r + 2] * A[pos[3 * r + 2]] + c[3 * r];
    A[pos[3 * r]] = v;
}

// fluids pass, general case: values first (k_ibm_values), then stores (k_ibm_store)
__global__ void k_ibm_values(const double* A, const long long* __restrict__ pos, const double* __restrict__ c,
                             double* vals, int n) {
#pragma clang fp contract(off)
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    vals[r] = c[3 * r + 1] * A[pos[3 * r + 1]] + c[3 * r + 2] * A[pos[3 * r + 2]] + c[3 * r];
}

__global__ void k_ibm_store(double* A, const long long* __restrict__ pos, const double* __restrict__ vals, int n) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) A[pos[3 * r]] = vals[r];
}

// forcing pass: image <- its target (images never coincide with targets)
__global__ void k_ibm_images(double* A, const long long* __restrict__ img, int m) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < m) A[img[2 * r]] = A[img[2 * r + 1]];
}

// ibm_part_2 (src/ibm_setup/mod_cfd_ibm.f90:47-141): column fills of a halo array
__global__ void k_ibm_columns(double* A, const int* __restrict__ bot, const int* __restrict__ top, int nx, int nz,
                             int ny, int kmax, double wall_bot, double wall_top) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nx * nz) return;
    const int i = (int)(idx % nx), j = (int)(idx / nx);
    const long sy = (long)(nx + 2) * (nz + 2);
    const long base = (long)(i + 1) + (long)(nx + 2) * (j + 1);
    const int b = min(bot[idx], ny), t = max(top[idx], -1);  // stay inside the k = -1..ny planes
    for (int k = 0; k <= b; ++k) A[base + (long)(k + 1) * sy] = wall_bot;
    for (int k = t; k <= kmax; ++k) A[base + (long)(k + 1) * sy] = wall_top;
}

static inline int imod(long a, int n) {
    long r = a % n;
    return (int)(r < 0 ? r + n : r);
}

}  // namespace dnsp

using namespace dnsp;

struct dnsp_ibm {
    int nx = 0, nz = 0, ny = 0, nhalo = 0, nskip = 0, n = 0, nimg = 0;
    bool disjoint = true;
    long long* d_pos = nullptr;  // 3 per forcing: target, source 1, source 2
    double* d_c = nullptr;       // 3 per forcing: cw, c1, c2
    long long* d_img = nullptr;  // 2 per image: image, target
    double* d_vals = nullptr;    // n, general fluids pass only
    int device = 0;
};

static void ibm_release(dnsp_ibm* h) {
    if (!h) return;
    (void)hipFree(h->d_pos);
    (void)hipFree(h->d_c);
    (void)hipFree(h->d_img);
    (void)hipFree(h->d_vals);
    delete h;
}

namespace dnsp {

size_t ibm_device_bytes(const dnsp_ibm* h) {
    if (!h) return 0;
    size_t b = 0;
    if (h->d_pos) b += sizeof(long long) * 3 * (size_t)h->n;
    if (h->d_c) b += sizeof(double) * 3 * (size_t)h->n;
    if (h->d_img) b += sizeof(long long) * 2 * (size_t)h->nimg;
    if (h->d_vals) b += sizeof(double) * (size_t)h->n;
    return b;
}

}  // namespace dnsp

extern "C" {

int dnsp_ibm_create(int nx, int nz, int ny, int nhalo, int nskip, int n, const int* ijk, const double* c,
                    dnsp_ibm** out) {
    clear_error();
    DNSP_ARG_CHECK(out, "out is NULL");
    *out = nullptr;
    DNSP_ARG_CHECK(nx >= 1 && nz >= 1 && ny >= 1 && nhalo >= 0 && nskip >= 0 && n >= 0, "invalid sizes");
    DNSP_ARG_CHECK(n == 0 || (ijk && c), "NULL forcing arrays");
    const int side = nhalo + nskip;
    const long A0 = nx + 2L * side, A1 = nz + 2L * side, A2 = ny + 2L * side;
    const int N[3] = {nx, nz, ny};
    auto flat = [&](long i, long j, long k) { return (k + side) * A1 * A0 + (j + side) * A0 + (i + side); };

    std::vector<long long> pos((size_t)3 * n);
    std::vector<double> cc((size_t)3 * n);
    std::unordered_set<long long> tset, sset;
    for (int r = 0; r < n; ++r) {
        for (int p = 0; p < 3; ++p) {
            // indices wrapped mod n as read (src/mod_utils.f90:197-200)
            const int i = imod(ijk[9 * r + 3 * p + 0], nx);
            const int j = imod(ijk[9 * r + 3 * p + 1], nz);
            const int k = imod(ijk[9 * r + 3 * p + 2], ny);
            pos[(size_t)3 * r + p] = flat(i, j, k);
        }
        for (int q = 0; q < 3; ++q) {
            cc[(size_t)3 * r + q] = c[3 * r + q];
            if (!std::isfinite(c[3 * r + q])) {
                set_error("forcing %d coefficient %d is not finite", r, q);
                return DNSP_ERR_ARG;
            }
        }
        if (!tset.insert(pos[(size_t)3 * r]).second) {
            set_error("forcing %d repeats a target point", r);
            return DNSP_ERR_ARG;
        }
        sset.insert(pos[(size_t)3 * r + 1]);
        sset.insert(pos[(size_t)3 * r + 2]);
    }
    bool disjoint = true;
    for (long long t : tset)
        if (sset.count(t)) {
            disjoint = false;
            break;
        }
    // forcing pass: targets with a periodic image in the halo shell
    std::vector<long long> img;
    for (int r = 0; r < n; ++r) {
        const int t[3] = {imod(ijk[9 * r], nx), imod(ijk[9 * r + 1], nz), imod(ijk[9 * r + 2], ny)};
        bool halo = false;
        std::vector<long long> cand;
        for (int si = -1; si <= 1; ++si)
            for (int sj = -1; sj <= 1; ++sj)
                for (int sk = -1; sk <= 1; ++sk) {
                    const long q[3] = {t[0] + (long)si * nx, t[1] + (long)sj * nz, t[2] + (long)sk * ny};
                    bool inner = true, in_pad = true, in_arr = true;
                    for (int d = 0; d < 3; ++d) {
                        inner = inner && q[d] >= 0 && q[d] < N[d];
                        in_pad = in_pad && q[d] >= -nhalo && q[d] <= N[d] + nhalo - 1;
                        in_arr = in_arr && q[d] + side >= 0 && q[d] + side < N[d] + 2L * side;
                    }
                    if (in_pad && !inner) halo = true;
                    if (in_arr && !inner) cand.push_back(flat(q[0], q[1], q[2]));
                }
        if (halo)
            for (long long p : cand) {
                img.push_back(p);
                img.push_back(pos[(size_t)3 * r]);
            }
    }
    (void)A2;
    dnsp_ibm* h = new (std::nothrow) dnsp_ibm();
    if (!h) {
        set_error("host allocation failed");
        return DNSP_ERR_NOMEM;
    }
    h->nx = nx; h->nz = nz; h->ny = ny; h->nhalo = nhalo; h->nskip = nskip; h->n = n;
    h->nimg = (int)(img.size() / 2);
    h->disjoint = disjoint;
    auto up = [&](void** dst, const void* src, size_t bytes) -> int {
        if (bytes == 0) return DNSP_OK;
        DNSP_HIP_TRY(hipMalloc(dst, bytes));
        DNSP_HIP_TRY(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return DNSP_OK;
    };
    int rc = hipGetDevice(&h->device) == hipSuccess ? DNSP_OK : DNSP_ERR_HIP;
    if (!rc) rc = up((void**)&h->d_pos, pos.data(), sizeof(long long) * pos.size());
    if (!rc) rc = up((void**)&h->d_c, cc.data(), sizeof(double) * cc.size());
    if (!rc) rc = up((void**)&h->d_img, img.data(), sizeof(long long) * img.size());
    if (!rc && !disjoint && n > 0) {
        if (hipMalloc(&h->d_vals, sizeof(double) * n) != hipSuccess) {
            set_error("hipMalloc of the IBM value buffer failed");
            rc = DNSP_ERR_NOMEM;
        }
    }
    if (rc) {
        if (rc == DNSP_ERR_HIP) set_error("hipGetDevice failed");
        ibm_release(h);
        return rc;
    }
    *out = h;
    return DNSP_OK;
}

int dnsp_ibm_exec(dnsp_ibm* h, double* A, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(h && A, "NULL handle or array");
    hipStream_t st = as_stream(stream);
    if (h->n > 0) {
        const int nb = (h->n + kIbmThreads - 1) / kIbmThreads;
        if (h->disjoint) {
            k_ibm_apply<<<nb, kIbmThreads, 0, st>>>(A, h->d_pos, h->d_c, h->n);
        } else {
            k_ibm_values<<<nb, kIbmThreads, 0, st>>>(A, h->d_pos, h->d_c, h->d_vals, h->n);
            k_ibm_store<<<nb, kIbmThreads, 0, st>>>(A, h->d_pos, h->d_vals, h->n);
        }
        DNSP_LAUNCH_CHECK();
    }
    if (h->nimg > 0) {
        k_ibm_images<<<(h->nimg + kIbmThreads - 1) / kIbmThreads, kIbmThreads, 0, st>>>(A, h->d_img, h->nimg);
        DNSP_LAUNCH_CHECK();
    }
    return DNSP_OK;
}

int dnsp_ibm_info(const dnsp_ibm* h, int* n_forcings, int* n_images, int* disjoint) {
    clear_error();
    DNSP_ARG_CHECK(h && n_forcings && n_images && disjoint, "NULL argument");
    *n_forcings = h->n;
    *n_images = h->nimg;
    *disjoint = h->disjoint ? 1 : 0;
    return DNSP_OK;
}

int dnsp_ibm_free(dnsp_ibm* h) {
    clear_error();
    ibm_release(h);
    return DNSP_OK;
}

// ibm_read_coeffs (src/mod_utils.f90:181-212): int32 count, then 60-byte
// records t(3 x i4) c1(f8) s1(3 x i4) c2(f8) s2(3 x i4) cw(f8); indices are
// wrapped mod (nx, nz, ny).  With ijk == NULL only *n is set.
int dnsp_ibm_read_coeffs(const char* path, int nx, int nz, int ny, int* n, int* ijk, double* c) {
    clear_error();
    DNSP_ARG_CHECK(path && n && nx >= 1 && nz >= 1 && ny >= 1, "invalid arguments");
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        set_error("cannot open %s", path);
        return DNSP_ERR_ARG;
    }
    int nb = 0;
    int rc = DNSP_OK;
    if (std::fread(&nb, 4, 1, f) != 1 || nb < 0) {
        set_error("%s: bad record count", path);
        rc = DNSP_ERR_ARG;
    }
    if (!rc && ijk) {
        DNSP_ARG_CHECK(c, "c is NULL");
        for (int r = 0; r < nb && !rc; ++r) {
            int32_t t[3], s1[3], s2[3];
            double c1, c2, cw;
            if (std::fread(t, 4, 3, f) != 3 || std::fread(&c1, 8, 1, f) != 1 || std::fread(s1, 4, 3, f) != 3 ||
                std::fread(&c2, 8, 1, f) != 1 || std::fread(s2, 4, 3, f) != 3 || std::fread(&cw, 8, 1, f) != 1) {
                set_error("%s: truncated at record %d", path, r);
                rc = DNSP_ERR_ARG;
                break;
            }
            const int32_t* pts[3] = {t, s1, s2};
            for (int p = 0; p < 3; ++p) {
                ijk[9 * r + 3 * p + 0] = imod(pts[p][0], nx);
                ijk[9 * r + 3 * p + 1] = imod(pts[p][1], nz);
                ijk[9 * r + 3 * p + 2] = imod(pts[p][2], ny);
            }
            c[3 * r + 0] = cw;
            c[3 * r + 1] = c1;
            c[3 * r + 2] = c2;
        }
    }
    std::fclose(f);
    if (!rc) *n = nb;
    return rc;
}

// set_ibm_bands.f90:57-77: extrapolation coefficients of the div field from
// the T forcing geometry: c = (0, 1 + d_ta/d_ab, -d_ta/d_ab).  yp: ny cell centres.
int dnsp_ibm_div_coeffs(int n, const int* ijk, const double* yp, int nx, int nz, int ny, double inv_dx,
                        double inv_dz, double* c) {
    clear_error();
    DNSP_ARG_CHECK(n >= 0 && (n == 0 || (ijk && yp && c)) && nx >= 1 && nz >= 1 && ny >= 1, "invalid arguments");
    const double dx = 1 / inv_dx, dz = 1 / inv_dz;
    auto qp = [](int i0, int t, int nn) {
        int i = i0;
        if (std::abs((i0 - nn) - t) < std::abs(i - t)) i = i0 - nn;
        if (std::abs((i0 + nn) - t) < std::abs(i - t)) i = i0 + nn;
        return i;
    };
    for (int r = 0; r < n; ++r) {
        const int* q = ijk + 9 * r;
        const int it = q[0], jt = q[1], kt = q[2];
        const int ia = qp(q[3], it, nx), ja = qp(q[4], jt, nz), ka = q[5];
        const int ib = qp(q[6], it, nx), jb = qp(q[7], jt, nz), kb = q[8];
        DNSP_ARG_CHECK(ka >= 0 && ka < ny && kb >= 0 && kb < ny && kt >= 0 && kt < ny, "k index out of range");
        const double d_ab = std::sqrt((dx * (ib - ia)) * (dx * (ib - ia)) + (dz * (jb - ja)) * (dz * (jb - ja)) +
                                      (yp[kb] - yp[ka]) * (yp[kb] - yp[ka]));
        const double d_ta = std::sqrt((dx * (it - ia)) * (dx * (it - ia)) + (dz * (jt - ja)) * (dz * (jt - ja)) +
                                      (yp[kt] - yp[ka]) * (yp[kt] - yp[ka]));
        c[3 * r + 0] = 0.0;
        c[3 * r + 1] = 1 + d_ta / d_ab;
        c[3 * r + 2] = -d_ta / d_ab;
    }
    return DNSP_OK;
}

// ibm_fill_slabs_2d (src/mod_utils.f90:215-239): int32 nx, nz then nx*nz int32
// (i fastest); returns the slab minus k_first clipped as set_ibm_bands.f90:196-218
// (bottom: min(v, ny-1), top: max(v, 0)).
int dnsp_ibm_read_slab(const char* path, int nx, int nz, int ny, int k_first, int is_top, int* out) {
    clear_error();
    DNSP_ARG_CHECK(path && out && nx >= 1 && nz >= 1 && ny >= 1, "invalid arguments");
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        set_error("cannot open %s", path);
        return DNSP_ERR_ARG;
    }
    int32_t hdr[2] = {0, 0};
    int rc = DNSP_OK;
    if (std::fread(hdr, 4, 2, f) != 2 || hdr[0] != nx || hdr[1] != nz) {
        set_error("%s: header (%d, %d) does not match (%d, %d)", path, hdr[0], hdr[1], nx, nz);
        rc = DNSP_ERR_ARG;
    } else if (std::fread(out, 4, (size_t)nx * nz, f) != (size_t)nx * nz) {
        set_error("%s: truncated slab", path);
        rc = DNSP_ERR_ARG;
    }
    std::fclose(f);
    if (rc) return rc;
    for (long q = 0; q < (long)nx * nz; ++q) {
        const int v = out[q] - k_first;
        out[q] = is_top ? std::max(v, 0) : std::min(v, ny - 1);
    }
    return DNSP_OK;
}

int dnsp_ibm_part2(double* A, const int* bot, const int* top, int nx, int nz, int ny, int kmax, double wall_bot,
                   double wall_top, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(A && bot && top && nx >= 1 && nz >= 1 && ny >= 1, "invalid arguments");
    DNSP_ARG_CHECK(kmax >= -1 && kmax <= ny, "kmax=%d outside [-1, ny]", kmax);
    const long n = (long)nx * nz;
    k_ibm_columns<<<(int)((n + kIbmThreads - 1) / kIbmThreads), kIbmThreads, 0, as_stream(stream)>>>(
        A, bot, top, nx, nz, ny, kmax, wall_bot, wall_top);
    DNSP_LAUNCH_CHECK();
    return DNSP_OK;
}

}  // extern "C"
