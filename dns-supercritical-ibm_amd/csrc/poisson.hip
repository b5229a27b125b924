// FP64 FFT + wall-normal tridiagonal pressure Poisson solve for gfx950.
//
// Path (single GPU, one y-rank), all in place on the caller's device array:
//   k_xfwd   x R2HC of every (j,k) line             poisson_solver_multigpu.f90:360-375
//   k_zfwd   z R2HC of every (a,k) column            :377-393 (the transpose is the tile load)
//   k_thomas per-(a,b)-mode Thomas along y           :395-422 with poi_C (:162-179) recomputed
//   k_zbwd   z HC2R                                  :501-520
//   k_xbwd   x HC2R and the 1/(nx*nz) scale          :523-551
// The spectrum is the reference FFTW branch's real-real halfcomplex layout:
// S(a,b,k) at a + nx*(b + nz*k), a the x halfcomplex index, b the z one.
//
// Real transforms are done two at a time through one complex FFT: two real
// x-lines u, v form u + i v; along z, the real and imaginary halfcomplex
// rows a and nx-a of one x wavenumber form the complex column X_a(z).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

#include "common.h"
#include "fft_lds.h"
#include "poisson_ct.h"
#include "poisson_ct_launch.h"
#include "poisson_internal.h"

namespace dnsp {

constexpr int kFftThreads = 256;
// complex elements per workgroup in the FFT kernels (LDS = 16 B each)
constexpr int kFftTileElems = 2048;

// ---------------------------------------------------------------------------
// x direction
// ---------------------------------------------------------------------------

// P (real lines of length nx, nlines of them) -> S halfcomplex, in place allowed.
__global__ __launch_bounds__(kFftThreads) void k_xfwd(const double* P, double* S, int nx, long nlines,
                                                      int G, FftPlan pl) {
    extern __shared__ double2 sm[];
    const long l0 = (long)blockIdx.x * 2 * G;
    const int tot = 2 * G * nx;
    double* smd = reinterpret_cast<double*>(sm);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e / nx;
        const int t = e - r * nx;
        const long line = l0 + r;
        const double v = (line < nlines) ? P[line * nx + t] : 0.0;
        smd[(size_t)((r >> 1) * nx + t) * 2 + (r & 1)] = v;
    }
    __syncthreads();
    fft_lds<false>(sm, G, pl);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e / nx;
        const int t = e - r * nx;
        const long line = l0 + r;
        if (line >= nlines) continue;
        const bool lo = 2 * t <= nx;
        const int u = lo ? t : nx - t;
        const double2 C = sm[(r >> 1) * nx + u];
        const double2 Cm = sm[(r >> 1) * nx + (u == 0 ? 0 : nx - u)];
        double val;
        if ((r & 1) == 0) val = lo ? 0.5 * (C.x + Cm.x) : 0.5 * (C.y - Cm.y);
        else              val = lo ? 0.5 * (C.y + Cm.y) : 0.5 * (Cm.x - C.x);
        S[line * nx + t] = val;
    }
}

// S halfcomplex lines -> P real lines, times `scale`.  In place allowed.
__global__ __launch_bounds__(kFftThreads) void k_xbwd(const double* S, double* P, int nx, long nlines,
                                                      int G, FftPlan pl, double scale) {
    extern __shared__ double2 sm[];
    const long l0 = (long)blockIdx.x * 2 * G;
    const int half = nx / 2 + 1;  // t in [0, nx/2]
    const int tot = G * half;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int g = e / half;
        const int t = e - g * half;
        const long la = l0 + 2 * g, lb = la + 1;
        const bool single = (t == 0) || (2 * t == nx);
        const double xr = (la < nlines) ? S[la * nx + t] : 0.0;
        const double xi = (!single && la < nlines) ? S[la * nx + nx - t] : 0.0;
        const double yr = (lb < nlines) ? S[lb * nx + t] : 0.0;
        const double yi = (!single && lb < nlines) ? S[lb * nx + nx - t] : 0.0;
        // C_t = X_t + i Y_t ; C_{n-t} = conj(X_t) + i conj(Y_t)
        sm[g * nx + t] = make_double2(xr - yi, xi + yr);
        if (!single) sm[g * nx + nx - t] = make_double2(xr + yi, yr - xi);
    }
    __syncthreads();
    fft_lds<true>(sm, G, pl);
    const int tot2 = 2 * G * nx;
    const double* smd = reinterpret_cast<const double*>(sm);
    for (int e = threadIdx.x; e < tot2; e += blockDim.x) {
        const int r = e / nx;
        const int t = e - r * nx;
        const long line = l0 + r;
        if (line >= nlines) continue;
        P[line * nx + t] = smd[(size_t)((r >> 1) * nx + t) * 2 + (r & 1)] * scale;
    }
}

// ---------------------------------------------------------------------------
// z direction: complex column c pairs the halfcomplex x rows
//   re_row(c) = c,  im_row(c) = nx - c (c >= 1), nx/2 (c == 0, nx even), none (c == 0, nx odd)
// ---------------------------------------------------------------------------

__device__ __forceinline__ int zcol_row(int c, int half, int nx) {
    if (half == 0) return c;
    if (c == 0) return (nx % 2 == 0 && nx > 1) ? nx / 2 : -1;
    return nx - c;
}

__global__ __launch_bounds__(kFftThreads) void k_zfwd(double* S, int nx, int nz, int ncol, int TA,
                                                      FftPlan pl) {
    extern __shared__ double2 sm[];
    const int c0 = blockIdx.x * TA;
    const size_t plane = (size_t)nx * nz;
    double* Sk = S + (size_t)blockIdx.y * plane;
    double* smd = reinterpret_cast<double*>(sm);
    const int tot = nz * TA * 2;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int tt = e % TA;
        const int rest = e / TA;
        const int hf = rest & 1;
        const int j = rest >> 1;
        const int c = c0 + tt;
        const int row = (c < ncol) ? zcol_row(c, hf, nx) : -1;
        const double v = (row >= 0) ? Sk[(size_t)j * nx + row] : 0.0;
        smd[(size_t)(tt * nz + j) * 2 + hf] = v;
    }
    __syncthreads();
    fft_lds<false>(sm, TA, pl);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int tt = e % TA;
        const int rest = e / TA;
        const int hf = rest & 1;
        const int b = rest >> 1;
        const int c = c0 + tt;
        const int row = (c < ncol) ? zcol_row(c, hf, nx) : -1;
        if (row < 0) continue;
        const bool lo = 2 * b <= nz;
        const int q = lo ? b : nz - b;
        const double2 Z = sm[tt * nz + q];
        const double2 Zm = sm[tt * nz + (q == 0 ? 0 : nz - q)];
        double val;
        if (hf == 0) val = lo ? 0.5 * (Z.x + Zm.x) : 0.5 * (Z.y - Zm.y);
        else         val = lo ? 0.5 * (Z.y + Zm.y) : 0.5 * (Zm.x - Z.x);
        Sk[(size_t)b * nx + row] = val;
    }
}

__global__ __launch_bounds__(kFftThreads) void k_zbwd(double* S, int nx, int nz, int ncol, int TA,
                                                      FftPlan pl) {
    extern __shared__ double2 sm[];
    const int c0 = blockIdx.x * TA;
    const size_t plane = (size_t)nx * nz;
    double* Sk = S + (size_t)blockIdx.y * plane;
    const int half = nz / 2 + 1;
    const int tot = half * TA;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int tt = e % TA;
        const int q = e / TA;
        const int c = c0 + tt;
        const int ru = (c < ncol) ? zcol_row(c, 0, nx) : -1;
        const int rv = (c < ncol) ? zcol_row(c, 1, nx) : -1;
        const bool single = (q == 0) || (2 * q == nz);
        const double ur = (ru >= 0) ? Sk[(size_t)q * nx + ru] : 0.0;
        const double ui = (ru >= 0 && !single) ? Sk[(size_t)(nz - q) * nx + ru] : 0.0;
        const double vr = (rv >= 0) ? Sk[(size_t)q * nx + rv] : 0.0;
        const double vi = (rv >= 0 && !single) ? Sk[(size_t)(nz - q) * nx + rv] : 0.0;
        sm[tt * nz + q] = make_double2(ur - vi, ui + vr);
        if (!single) sm[tt * nz + nz - q] = make_double2(ur + vi, vr - ui);
    }
    __syncthreads();
    fft_lds<true>(sm, TA, pl);
    const double* smd = reinterpret_cast<const double*>(sm);
    const int tot2 = nz * TA * 2;
    for (int e = threadIdx.x; e < tot2; e += blockDim.x) {
        const int tt = e % TA;
        const int rest = e / TA;
        const int hf = rest & 1;
        const int j = rest >> 1;
        const int c = c0 + tt;
        const int row = (c < ncol) ? zcol_row(c, hf, nx) : -1;
        if (row < 0) continue;
        Sk[(size_t)j * nx + row] = smd[(size_t)(tt * nz + j) * 2 + hf];
    }
}

// ---------------------------------------------------------------------------
// y direction: per-mode Thomas.  bands = [a | b | c | a00 | b00 | c00], ny each.
// Same operation order as the reference (no contraction), so this stage
// reproduces a serial FP64 Thomas bit for bit.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_thomas(double* S, double* cpw, const double* __restrict__ ax,
                                                const double* __restrict__ az,
                                                const int* __restrict__ is0x,
                                                const int* __restrict__ is0z,
                                                const double* __restrict__ bands, int nx, int nz, int ny) {
#pragma clang fp contract(off)
    const long plane = (long)nx * nz;
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= plane) return;
    const int i = (int)(m % nx);
    const int j = (int)(m / nx);
    const double a_xz = ax[i] + az[j];
    const bool is00 = is0x[i] && is0z[j];
    const double* A = bands + (is00 ? 3 * ny : 0);
    const double* B = A + ny;
    const double* C = B + ny;
    double* d = S + m;
    double* cp = cpw + m;
    double den = B[0] + a_xz;
    double c_prev = C[0] / den;
    double d_prev = d[0] / den;
    d[0] = d_prev;
    cp[0] = c_prev;
    for (int k = 1; k < ny; ++k) {
        const size_t o = (size_t)k * plane;
        den = B[k] + a_xz - A[k] * c_prev;
        d_prev = (d[o] - A[k] * d_prev) / den;
        c_prev = C[k] / den;
        d[o] = d_prev;
        cp[o] = c_prev;
    }
    double x = d_prev;
    for (int k = ny - 2; k >= 0; --k) {
        const size_t o = (size_t)k * plane;
        x = d[o] - x * cp[o];
        d[o] = x;
    }
}

// min over modes and rows of |pivot| (init diagnostic); one value per block.
__global__ __launch_bounds__(256) void k_min_pivot(double* out, const double* __restrict__ ax,
                                                   const double* __restrict__ az,
                                                   const int* __restrict__ is0x,
                                                   const int* __restrict__ is0z,
                                                   const double* __restrict__ bands, int nx, int nz,
                                                   int ny) {
#pragma clang fp contract(off)
    __shared__ double red[256];
    const long plane = (long)nx * nz;
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double mn = 1e20;
    if (m < plane) {
        const int i = (int)(m % nx);
        const int j = (int)(m / nx);
        const double a_xz = ax[i] + az[j];
        const bool is00 = is0x[i] && is0z[j];
        const double* A = bands + (is00 ? 3 * ny : 0);
        const double* B = A + ny;
        const double* C = B + ny;
        double den = B[0] + a_xz;
        mn = fabs(den);
        double cp = C[0] / den;
        for (int k = 1; k < ny; ++k) {
            den = B[k] + a_xz - A[k] * cp;
            mn = fmin(mn, fabs(den));
            cp = C[k] / den;
        }
        if (!(mn == mn)) mn = -1.0;  // NaN pivot: report as invalid
    }
    red[threadIdx.x] = mn;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static bool fft_factorize(int n, FftPlan& pl) {
    pl.n = n;
    pl.nf = 0;
    int r = n;
    auto push = [&](int f) {
        if (pl.nf >= kFftMaxFactors) return false;
        pl.fac[pl.nf++] = f;
        r /= f;
        return true;
    };
    while (r % 8 == 0 && r != 16) if (!push(8)) return false;
    while (r % 4 == 0) if (!push(4)) return false;
    while (r % 2 == 0) if (!push(2)) return false;
    for (int p : {3, 5, 7, 11, 13})
        while (r % p == 0) if (!push(p)) return false;
    return r == 1;
}

static int fft_radix_limit(const FftPlan& pl) {
    // largest G*n allowed by the register budget of each stage
    int lim = 1 << 30;
    for (int s = 0; s < pl.nf; ++s) {
        int R = pl.fac[s];
        int nb = (kFftRegBudget / R) > 0 ? (kFftRegBudget / R) : 1;
        lim = std::min(lim, nb * R * kFftThreads);
    }
    return lim;
}

static int make_twiddles(int n, double2** dev) {
    std::vector<double2> tw((size_t)n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int t = 0; t < n; ++t) {
        long double ang = 2.0L * pi * (long double)t / (long double)n;
        tw[(size_t)t] = make_double2((double)cosl(ang), (double)(-sinl(ang)));
    }
    DNSP_HIP_TRY(hipMalloc(dev, sizeof(double2) * (size_t)n));
    DNSP_HIP_TRY(hipMemcpy(*dev, tw.data(), sizeof(double2) * (size_t)n, hipMemcpyHostToDevice));
    return DNSP_OK;
}

// poisson_solver_multigpu.f90:139-152 (FFTW halfcomplex index -> wavenumber)
static int hc_wavenumber(int i, int n) {
    if (i <= n / 2) return i;
    int ii = (n + 1) / 2 - 1 - (i - (n / 2 + 1));
    return ii > 1 ? ii : 1;
}

static void eigenvalues(int n, double inv_d, std::vector<double>& lam, std::vector<int>& is00) {
    const double const_pi = 4.0 * std::atan(1.0);
    lam.resize((size_t)n);
    is00.resize((size_t)n);
    for (int i = 0; i < n; ++i) {
        int ii = hc_wavenumber(i, n);
        double s = std::sin(ii * const_pi / n);
        lam[(size_t)i] = -4 * (inv_d * inv_d) * (s * s);
        is00[(size_t)i] = (ii == 0) ? 1 : 0;
    }
}


}  // namespace dnsp

using namespace dnsp;

static void poisson_release(dnsp_poisson* h) {
    if (!h) return;
    int cur = 0;
    if (hipGetDevice(&cur) == hipSuccess && cur != h->device) (void)hipSetDevice(h->device);
    (void)hipFree(h->d_ax);
    (void)hipFree(h->d_az);
    (void)hipFree(h->d_is0x);
    (void)hipFree(h->d_is0z);
    (void)hipFree(h->d_bands);
    (void)hipFree(h->d_twx);
    (void)hipFree(h->d_twz);
    (void)hipFree(h->d_cp);
    (void)hipFree(h->d_W);
    (void)hipFree(h->d_host);
    mixed_release(h);
    if (cur != h->device) (void)hipSetDevice(cur);
    delete h;
}

static int poisson_setup(dnsp_poisson* h, const double* const bands[6]) {
    DNSP_HIP_TRY(hipGetDevice(&h->device));
    const int nx = h->nx, nz = h->nz, ny = h->ny;
    if (!fft_factorize(nx, h->px) || !fft_factorize(nz, h->pz)) {
        set_error("nx=%d / nz=%d: FFT lengths must factor into 2, 3, 5, 7, 11 and 13", nx, nz);
        return DNSP_ERR_UNSUPPORTED;
    }
    const int limx = fft_radix_limit(h->px), limz = fft_radix_limit(h->pz);
    if (nx > limx || nz > limz || nx > kFftTileElems * 2 || nz > kFftTileElems * 2) {
        set_error("nx=%d / nz=%d: FFT length too large for the LDS kernels", nx, nz);
        return DNSP_ERR_UNSUPPORTED;
    }
    h->Gx = std::max(1, std::min(kFftTileElems / nx, limx / nx));
    h->TAz = std::max(1, std::min(kFftTileElems / nz, limz / nz));
    h->ncol = (nx + 1) / 2;

    std::vector<double> ax, az;
    std::vector<int> i0x, i0z;
    eigenvalues(nx, h->inv_dx, ax, i0x);
    eigenvalues(nz, h->inv_dz, az, i0z);
    DNSP_HIP_TRY(hipMalloc(&h->d_ax, sizeof(double) * nx));
    DNSP_HIP_TRY(hipMalloc(&h->d_az, sizeof(double) * nz));
    DNSP_HIP_TRY(hipMalloc(&h->d_is0x, sizeof(int) * nx));
    DNSP_HIP_TRY(hipMalloc(&h->d_is0z, sizeof(int) * nz));
    DNSP_HIP_TRY(hipMemcpy(h->d_ax, ax.data(), sizeof(double) * nx, hipMemcpyHostToDevice));
    DNSP_HIP_TRY(hipMemcpy(h->d_az, az.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
    DNSP_HIP_TRY(hipMemcpy(h->d_is0x, i0x.data(), sizeof(int) * nx, hipMemcpyHostToDevice));
    DNSP_HIP_TRY(hipMemcpy(h->d_is0z, i0z.data(), sizeof(int) * nz, hipMemcpyHostToDevice));

    std::vector<double> hb((size_t)6 * ny);
    for (int s = 0; s < 6; ++s) std::memcpy(hb.data() + (size_t)s * ny, bands[s], sizeof(double) * ny);
    // weakest coupling a_k + c_k of the plain bands (interior rows when there are any)
    {
        double mn = 1e300;
        const int k0 = ny > 2 ? 1 : 0, k1 = ny > 2 ? ny - 1 : ny;
        for (int k = k0; k < k1; ++k) mn = std::min(mn, std::fabs(bands[0][k]) + std::fabs(bands[2][k]));
        h->pcr_min = mn;
    }
    DNSP_HIP_TRY(hipMalloc(&h->d_bands, sizeof(double) * 6 * ny));
    DNSP_HIP_TRY(hipMemcpy(h->d_bands, hb.data(), sizeof(double) * 6 * ny, hipMemcpyHostToDevice));

    int rc = make_twiddles(nx, &h->d_twx);
    if (rc) return rc;
    rc = make_twiddles(nz, &h->d_twz);
    if (rc) return rc;
    h->px.tw = h->d_twx;
    h->pz.tw = h->d_twz;


    // min |pivot| diagnostic (init_spec_poisson_multigpu :181-222)
    const long plane = (long)nx * nz;
    const int nblk = (int)((plane + 255) / 256);
    double* d_red = nullptr;
    DNSP_HIP_TRY(hipMalloc(&d_red, sizeof(double) * nblk));
    k_min_pivot<<<nblk, 256>>>(d_red, h->d_ax, h->d_az, h->d_is0x, h->d_is0z, h->d_bands, nx, nz, ny);
    hipError_t le = hipGetLastError();
    std::vector<double> red((size_t)nblk);
    hipError_t ce = hipMemcpy(red.data(), d_red, sizeof(double) * nblk, hipMemcpyDeviceToHost);
    (void)hipFree(d_red);
    DNSP_HIP_TRY(le);
    DNSP_HIP_TRY(ce);
    double mn = 1e20;
    bool bad = false;
    for (double v : red) {
        if (v < 0) bad = true;
        mn = std::min(mn, v);
    }
    h->min_pivot = mn;
    if (bad || !(mn > 0) || !std::isfinite(mn)) {
        set_error("tridiagonal pivot is zero or not finite (min |pivot| = %g)", mn);
        return DNSP_ERR_SINGULAR;
    }
    return DNSP_OK;
}

static int poisson_apply_strategy(dnsp_poisson* h) {
    const bool want_ct = h->fft_strategy != DNSP_FFT_GENERIC;
    h->ct_x = want_ct && ct_fft_supported(h->nx);
    h->ct_z = want_ct && ct_fft_supported(h->nz);
    if (h->fft_strategy == DNSP_FFT_FIXED && !(h->ct_x && h->ct_z)) {
        set_error("no compile-time FFT kernels for nx=%d / nz=%d", h->nx, h->nz);
        return DNSP_ERR_UNSUPPORTED;
    }
    const int mr = part_rows_for(h->ny);
    if (h->tri_strategy == DNSP_TRIDIAG_PARTITION && mr == 0) {
        set_error("ny=%d too large for the partitioned tridiagonal kernel", h->ny);
        return DNSP_ERR_UNSUPPORTED;
    }
    const bool part = (h->tri_strategy == DNSP_TRIDIAG_PARTITION) ||
                      (h->tri_strategy == DNSP_TRIDIAG_AUTO && mr > 0 && h->ny >= 32);
    h->part_rows = part ? mr : 0;
    if (!part && !h->d_cp) {
        const size_t N = (size_t)h->nx * h->nz * h->ny;
        DNSP_HIP_TRY(hipMalloc(&h->d_cp, sizeof(double) * N));
    }
    // the blocked solve: compile-time x and z transforms, nx a multiple of 16,
    // the partitioned tridiagonal; its workspace is one field
    h->blk_ok = h->fft_strategy == DNSP_FFT_AUTO && h->ct_x && h->ct_z && xb_supported(h->nx) && part;
    if (h->blk_ok && !h->d_W) {
        const size_t N = (size_t)h->nx * h->nz * h->ny;
        DNSP_HIP_TRY(hipMalloc(&h->d_W, sizeof(double) * N));
    }
    return DNSP_OK;
}

extern "C" {

int dnsp_poisson_set_strategy(dnsp_poisson* h, int fft, int tridiag) {
    clear_error();
    DNSP_ARG_CHECK(h, "NULL handle");
    DNSP_ARG_CHECK(fft >= DNSP_FFT_AUTO && fft <= DNSP_FFT_FIXED, "fft strategy %d", fft);
    DNSP_ARG_CHECK(tridiag >= DNSP_TRIDIAG_AUTO && tridiag <= DNSP_TRIDIAG_PARTITION, "tridiag strategy %d",
                   tridiag);
    const int of = h->fft_strategy, ot = h->tri_strategy;
    h->fft_strategy = fft;
    h->tri_strategy = tridiag;
    int rc = poisson_apply_strategy(h);
    if (rc) {
        h->fft_strategy = of;
        h->tri_strategy = ot;
        (void)poisson_apply_strategy(h);
    }
    return rc;
}

int dnsp_poisson_get_strategy(const dnsp_poisson* h, int* fft_ct_x, int* fft_ct_z, int* part_rows) {
    clear_error();
    DNSP_ARG_CHECK(h && fft_ct_x && fft_ct_z && part_rows, "NULL argument");
    *fft_ct_x = h->ct_x ? 1 : 0;
    *fft_ct_z = h->ct_z ? 1 : 0;
    *part_rows = h->part_rows;
    return DNSP_OK;
}

int dnsp_poisson_set_variants(dnsp_poisson* h, int xvar, int zvar, int tvar) {
    clear_error();
    DNSP_ARG_CHECK(h, "NULL handle");
    DNSP_ARG_CHECK(xvar >= 0 && xvar < kCtVariants && zvar >= 0 && zvar < kCtVariants,
                   "FFT kernel variants must be in [0, %d)", kCtVariants);
    DNSP_ARG_CHECK(tvar >= 0 && tvar < kPartVariants, "tridiagonal variant must be in [0, %d)", kPartVariants);
    h->xvar = h->xvar_b = xvar;
    h->zvar = h->zvar_b = zvar;
    h->tvar = tvar;
    // the blocked solve's knobs alike (its z kernels: 256 or 512 threads)
    h->bxvar = h->bxvar_b = xvar;
    h->bzvar = h->bzvar_b = ct_zvar_threads(zvar) == 512 ? 1 : 0;
    h->btvar = tvar;
    return DNSP_OK;
}

int dnsp_poisson_get_variants(const dnsp_poisson* h, int* xvar, int* zvar, int* tvar) {
    clear_error();
    DNSP_ARG_CHECK(h && xvar && zvar && tvar, "NULL argument");
    *xvar = h->xvar;
    *zvar = h->zvar;
    *tvar = h->tvar;
    return DNSP_OK;
}

// Times each FFT kernel variant per direction (`reps` forward transforms,
// then `reps` inverse ones, keeping the fastest of each) and pairs of
// tridiagonal solves per partitioned variant, on `work` (zero-filled first;
// NULL = a temporary device array); for a mixed-precision handle the same
// for its kernels.  times_ms (optional, DNSP_TUNE_SLOTS entries): mean ms of
// a forward + inverse pair of the x, z (4 slots each) and of a pair of
// tridiagonal (8 slots) variants, then the same for the mixed-precision
// kernels; -1 where not applicable.
static_assert(kCtVariants == 4 && kPartVariants == 8 && DNSP_TUNE_SLOTS == 2 * kTuneFormSlots, "tuner table layout");
int dnsp_poisson_autotune(dnsp_poisson* h, double* work, int reps, void* stream, double* times_ms) {
    clear_error();
    DNSP_ARG_CHECK(h, "NULL handle");
    DNSP_ARG_CHECK(reps >= 1 && reps <= 1000, "reps=%d", reps);
    hipStream_t st = as_stream(stream);
    const size_t N = (size_t)h->nx * h->nz * h->ny;
    double* buf = work;
    if (!buf) DNSP_HIP_TRY(hipMalloc(&buf, sizeof(double) * N));
    int rc = DNSP_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // slots: x fwd, x bwd, z fwd, z bwd, tridiagonal
    double best[5] = {1e30, 1e30, 1e30, 1e30, 1e30};
    int pick[5] = {h->xvar, h->xvar_b, h->zvar, h->zvar_b, h->tvar};
    const int keep[5] = {h->xvar, h->xvar_b, h->zvar, h->zvar_b, h->tvar};
    int* knob[5] = {&h->xvar, &h->xvar_b, &h->zvar, &h->zvar_b, &h->tvar};
    if (times_ms)
        for (int i = 0; i < DNSP_TUNE_SLOTS; ++i) times_ms[i] = -1.0;
    auto timed = [&](int stage, int n, double* per) -> int {
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < n; ++r) {
            const int c = dnsp_poisson_stage(h, buf, buf, stage, stream);
            if (c) return c;
        }
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess) {
            set_error("autotune: event synchronize failed");
            return DNSP_ERR_HIP;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *per = (double)ms / n;
        return DNSP_OK;
    };
    do {
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
            hipMemsetAsync(buf, 0, sizeof(double) * N, st) != hipSuccess) {
            set_error("autotune: event or memset failed");
            rc = DNSP_ERR_HIP;
            break;
        }
        for (int dir = 0; dir < 3 && rc == DNSP_OK; ++dir) {
            const bool has = dir == 0 ? h->ct_x : dir == 1 ? h->ct_z : h->part_rows > 0;
            if (!has) continue;
            const int nvar = dir == 2 ? kPartVariants : kCtVariants;
            const int sf = dir == 0 ? DNSP_STAGE_X_FWD : dir == 1 ? DNSP_STAGE_Z_FWD : DNSP_STAGE_TRIDIAG;
            const int sb = dir == 0 ? DNSP_STAGE_X_BWD : dir == 1 ? DNSP_STAGE_Z_BWD : DNSP_STAGE_TRIDIAG;
            const int kf = 2 * dir, kb = dir == 2 ? 4 : 2 * dir + 1;  // knob slots
            for (int v = 0; v < nvar && rc == DNSP_OK; ++v) {
                *knob[kf] = v;
                *knob[kb] = v;
                rc = dnsp_poisson_stage(h, buf, buf, sf, stream);  // warm-up
                if (!rc) rc = dnsp_poisson_stage(h, buf, buf, sb, stream);
                double tf = 0.0, tb = 0.0;
                if (!rc) rc = timed(sf, reps, &tf);
                if (!rc && dir < 2) rc = timed(sb, reps, &tb);
                if (rc) break;
                if (dir == 2) tf *= 2.0;  // a pair of tridiagonal solves, as the FFT pairs
                if (times_ms) times_ms[kTuneSlot[dir] + v] = tf + tb;
                if (tf < best[kf]) {
                    best[kf] = tf;
                    pick[kf] = v;
                }
                if (dir < 2 && tb < best[kb]) {
                    best[kb] = tb;
                    pick[kb] = v;
                }
            }
        }
        if (rc || !h->blk_ok) break;
        // the blocked solve's own knobs, timed stage by stage through
        // dnsp_poisson_solve_stage on buf (rhs and solution) and the workspace
        if (hipMemsetAsync(h->d_W, 0, sizeof(double) * N, st) != hipSuccess) {
            set_error("autotune: memset failed");
            rc = DNSP_ERR_HIP;
            break;
        }
        const bool on = h->blk_on;
        h->blk_on = true;
        auto timed_b = [&](int stage, double* per) -> int {
            int c = dnsp_poisson_solve_stage(h, buf, buf, stage, stream);  // warm-up
            (void)hipEventRecord(e0, st);
            for (int r = 0; r < reps && !c; ++r) c = dnsp_poisson_solve_stage(h, buf, buf, stage, stream);
            if (c) return c;
            (void)hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) != hipSuccess) {
                set_error("autotune: event synchronize failed");
                return DNSP_ERR_HIP;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            *per = (double)ms / reps;
            return DNSP_OK;
        };
        int* bk[5] = {&h->bxvar, &h->bzvar, &h->btvar, &h->bzvar_b, &h->bxvar_b};  // by stage
        const int nv[5] = {kCtVariants, 2, kPartVariants, 2, kCtVariants};
        for (int stage = 0; stage < 5 && !rc; ++stage) {
            double bt = 1e30;
            int bv = *bk[stage];
            for (int v = 0; v < nv[stage] && !rc; ++v) {
                *bk[stage] = v;
                double t = 0.0;
                rc = timed_b(stage, &t);
                if (!rc && t < bt) {
                    bt = t;
                    bv = v;
                }
            }
            *bk[stage] = bv;
        }
        h->blk_on = on;
    } while (0);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) {
        if (!work && buf) (void)hipFree(buf);
        for (int i = 0; i < 5; ++i) *knob[i] = keep[i];
        return rc;
    }
    for (int i = 0; i < 5; ++i) *knob[i] = pick[i];
    if (h->precision == DNSP_PRECISION_MIXED) rc = mixed_autotune(h, buf, reps, st, times_ms ? times_ms + kTuneFormSlots : nullptr);
    if (!work && buf) (void)hipFree(buf);
    return rc;
}

int dnsp_poisson_init(int nx_glob, int nz_glob, int ny, double inv_dx, double inv_dz,
                      const double* P_band_a, const double* P_band_b, const double* P_band_c,
                      const double* P_band_a_00, const double* P_band_b_00, const double* P_band_c_00,
                      int precision, dnsp_poisson** out) {
    clear_error();
    DNSP_ARG_CHECK(out != nullptr, "out is NULL");
    *out = nullptr;
    DNSP_ARG_CHECK(nx_glob >= 1 && nz_glob >= 1, "nx_glob=%d, nz_glob=%d must be >= 1", nx_glob, nz_glob);
    DNSP_ARG_CHECK(ny >= 2, "ny_glob <2 (ny=%d)", ny);  // alloc_poisson_solver_multigpu.f90:34
    DNSP_ARG_CHECK((long)nx_glob * nz_glob * ny < (1L << 40), "field too large");
    DNSP_ARG_CHECK(std::isfinite(inv_dx) && std::isfinite(inv_dz), "inv_dx / inv_dz not finite");
    DNSP_ARG_CHECK(precision == DNSP_PRECISION_FP64 || precision == DNSP_PRECISION_MIXED,
                   "precision=%d: 0 (FP64) or 1 (FP64 plus the mixed-precision form)", precision);
    const double* bands[6] = {P_band_a, P_band_b, P_band_c, P_band_a_00, P_band_b_00, P_band_c_00};
    for (int s = 0; s < 6; ++s) {
        DNSP_ARG_CHECK(bands[s] != nullptr, "band array %d is NULL", s);
        for (int k = 0; k < ny; ++k)
            DNSP_ARG_CHECK(std::isfinite(bands[s][k]), "band array %d row %d is not finite", s, k);
    }
    dnsp_poisson* h = new (std::nothrow) dnsp_poisson();
    if (!h) {
        set_error("host allocation failed");
        return DNSP_ERR_NOMEM;
    }
    h->nx = nx_glob;
    h->nz = nz_glob;
    h->ny = ny;
    h->inv_dx = inv_dx;
    h->inv_dz = inv_dz;
    h->precision = precision;
    int rc = poisson_setup(h, bands);
    if (!rc) rc = poisson_apply_strategy(h);
    if (!rc) rc = mixed_setup(h);
    if (rc) {
        poisson_release(h);
        return rc;
    }
    *out = h;
    return DNSP_OK;
}

int dnsp_poisson_stage(dnsp_poisson* h, const double* src, double* P, int stage, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(h && src && P, "NULL handle or array");
    DNSP_ARG_CHECK(src == P || stage == DNSP_STAGE_X_FWD || stage == DNSP_STAGE_X_BWD,
                   "stage %d runs in place only (src must equal dst)", stage);
    hipStream_t st = as_stream(stream);
    const long nlines = (long)h->nz * h->ny;
    const int gx = (int)((nlines + 2 * h->Gx - 1) / (2 * h->Gx));
    const dim3 gz((h->ncol + h->TAz - 1) / h->TAz, h->ny);
    const size_t ldsx = sizeof(double2) * h->Gx * h->nx;
    const size_t ldsz = sizeof(double2) * h->TAz * h->nz;
    const double inv_nxnz = 1.0 / ((double)h->nx * (double)h->nz + 0.0);
    // the compile-time x kernels move (re, im) pairs as 16-byte accesses: an
    // array that is only 8-byte aligned (a view offset by one double, a
    // Fortran section) takes the runtime-sized kernels, which access doubles
    const bool x16 = ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(P)) & 15) == 0;
    switch (stage) {
        case DNSP_STAGE_X_FWD:
            if (h->ct_x && x16) ct_x_launch(h->nx, h->xvar, false, src, P, nlines, h->d_twx, 1.0, st);
            else k_xfwd<<<gx, kFftThreads, ldsx, st>>>(src, P, h->nx, nlines, h->Gx, h->px);
            break;
        case DNSP_STAGE_Z_FWD:
            if (h->ct_z) ct_z_launch(h->nz, h->zvar, false, P, h->nx, h->ny, h->d_twz, st);
            else k_zfwd<<<gz, kFftThreads, ldsz, st>>>(P, h->nx, h->nz, h->ncol, h->TAz, h->pz);
            break;
        case DNSP_STAGE_TRIDIAG: {
            if (h->part_rows > 0) {
                part_launch<double>(h->tvar, P, h->d_ax, h->d_az, h->d_bands,
                                    h->nx, h->nz, h->ny, h->pcr_min, nullptr, nullptr, 0.0, st);
            } else {
                const long plane = (long)h->nx * h->nz;
                const int nb = (int)((plane + 255) / 256);
                k_thomas<<<nb, 256, 0, st>>>(P, h->d_cp, h->d_ax, h->d_az, h->d_is0x, h->d_is0z, h->d_bands,
                                             h->nx, h->nz, h->ny);
            }
            break;
        }
        case DNSP_STAGE_Z_BWD:
            if (h->ct_z) ct_z_launch(h->nz, h->zvar_b, true, P, h->nx, h->ny, h->d_twz, st);
            else k_zbwd<<<gz, kFftThreads, ldsz, st>>>(P, h->nx, h->nz, h->ncol, h->TAz, h->pz);
            break;
        case DNSP_STAGE_X_BWD:
            if (h->ct_x && x16) ct_x_launch(h->nx, h->xvar_b, true, src, P, nlines, h->d_twx, inv_nxnz, st);
            else k_xbwd<<<gx, kFftThreads, ldsx, st>>>(src, P, h->nx, nlines, h->Gx, h->px, inv_nxnz);
            break;
        default:
            set_error("unknown stage %d", stage);
            return DNSP_ERR_ARG;
    }
    DNSP_LAUNCH_CHECK();
    return DNSP_OK;
}

int dnsp_poisson_forward(dnsp_poisson* h, double* P, void* stream) {
    int rc = dnsp_poisson_stage(h, P, P, DNSP_STAGE_X_FWD, stream);
    return rc ? rc : dnsp_poisson_stage(h, P, P, DNSP_STAGE_Z_FWD, stream);
}

int dnsp_poisson_tridiag(dnsp_poisson* h, double* S, void* stream) {
    return dnsp_poisson_stage(h, S, S, DNSP_STAGE_TRIDIAG, stream);
}

int dnsp_poisson_backward(dnsp_poisson* h, double* S, void* stream) {
    int rc = dnsp_poisson_stage(h, S, S, DNSP_STAGE_Z_BWD, stream);
    return rc ? rc : dnsp_poisson_stage(h, S, S, DNSP_STAGE_X_BWD, stream);
}

int dnsp_poisson_set_form(dnsp_poisson* h, int refinements) {
    clear_error();
    DNSP_ARG_CHECK(h, "NULL handle");
    DNSP_ARG_CHECK(refinements >= -1 && refinements <= DNSP_MAX_REFINEMENTS, "refinements=%d not in [-1, %d]",
                   refinements, DNSP_MAX_REFINEMENTS);
    DNSP_ARG_CHECK(refinements < 0 || h->precision == DNSP_PRECISION_MIXED,
                   "the mixed form needs a handle created with DNSP_PRECISION_MIXED");
    h->form = refinements;
    return DNSP_OK;
}

int dnsp_poisson_solve_stage(dnsp_poisson* h, const double* rhs, double* P, int stage, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(h && rhs && P, "NULL handle or array");
    DNSP_ARG_CHECK(stage >= 0 && stage < 5, "stage %d not in [0, 5)", stage);
    const bool blk = h->blk_ok && h->blk_on &&
                     ((reinterpret_cast<uintptr_t>(rhs) | reinterpret_cast<uintptr_t>(P)) & 15) == 0;
    if (!blk) {
        // the in-place stages; the first one reads rhs
        const int st_of[5] = {DNSP_STAGE_X_FWD, DNSP_STAGE_Z_FWD, DNSP_STAGE_TRIDIAG, DNSP_STAGE_Z_BWD,
                              DNSP_STAGE_X_BWD};
        return dnsp_poisson_stage(h, stage == 0 ? rhs : P, P, st_of[stage], stream);
    }
    hipStream_t st = as_stream(stream);
    const long nlines = (long)h->nz * h->ny;
    const double inv_nxnz = 1.0 / ((double)h->nx * (double)h->nz + 0.0);
    // P lines -> W x-blocked -> P mode-major -> W x-blocked -> P lines -> P
    switch (stage) {
        case 0: xb_x_launch(h->nx, h->bxvar, rhs, h->d_W, nlines, h->nz, h->d_twx, st); break;
        case 1: mm_z_launch(h->nz, h->bzvar, false, h->d_W, P, h->nx, h->ny, h->d_twz, st); break;
        case 2:
            part_launch<double, double, true>(h->btvar, P, h->d_ax, h->d_az, h->d_bands, h->nx, h->nz, h->ny,
                                              h->pcr_min, nullptr, nullptr, 0.0, st, nullptr, 0, h->d_W);
            break;
        case 3: mm_z_launch(h->nz, h->bzvar_b, true, h->d_W, P, h->nx, h->ny, h->d_twz, st); break;
        default: ct_x_launch(h->nx, h->bxvar_b, true, P, P, nlines, h->d_twx, inv_nxnz, st); break;
    }
    DNSP_LAUNCH_CHECK();
    return DNSP_OK;
}

int dnsp_poisson_solve(dnsp_poisson* h, const double* rhs, double* P, void* stream) {
    for (int s = 0; s < 5; ++s) {
        const int rc = dnsp_poisson_solve_stage(h, rhs, P, s, stream);
        if (rc) return rc;
    }
    return DNSP_OK;
}

int dnsp_poisson_run(dnsp_poisson* h, double* P, void* stream) {
    if (h && h->form >= 0) return dnsp_poisson_run_mixed(h, P, h->form, stream);
    return dnsp_poisson_solve(h, P, P, stream);
}

int dnsp_poisson_set_blocked(dnsp_poisson* h, int on) {
    clear_error();
    DNSP_ARG_CHECK(h, "NULL handle");
    h->blk_on = on != 0;
    return DNSP_OK;
}

int dnsp_poisson_get_solve_variants(const dnsp_poisson* h, int* v) {
    clear_error();
    DNSP_ARG_CHECK(h && v, "NULL argument");
    const bool blk = h->blk_ok && h->blk_on;
    v[0] = blk ? h->bxvar : h->xvar;
    v[1] = blk ? h->bzvar : h->zvar;
    v[2] = blk ? h->btvar : h->tvar;
    v[3] = blk ? h->bzvar_b : h->zvar_b;
    v[4] = blk ? h->bxvar_b : h->xvar_b;
    return DNSP_OK;
}

int dnsp_poisson_get_blocked(const dnsp_poisson* h, int* supported, int* on) {
    clear_error();
    DNSP_ARG_CHECK(h, "NULL handle");
    if (supported) *supported = h->blk_ok ? 1 : 0;
    if (on) *on = (h->blk_ok && h->blk_on) ? 1 : 0;
    return DNSP_OK;
}

int dnsp_poisson_run_host(dnsp_poisson* h, double* P_host) {
    clear_error();
    DNSP_ARG_CHECK(h && P_host, "NULL handle or array");
    const size_t bytes = sizeof(double) * (size_t)h->nx * h->nz * h->ny;
    if (!h->d_host) DNSP_HIP_TRY(hipMalloc(&h->d_host, bytes));
    DNSP_HIP_TRY(hipMemcpy(h->d_host, P_host, bytes, hipMemcpyHostToDevice));
    int rc = dnsp_poisson_run(h, h->d_host, nullptr);
    if (rc) return rc;
    DNSP_HIP_TRY(hipMemcpy(P_host, h->d_host, bytes, hipMemcpyDeviceToHost));
    DNSP_HIP_TRY(hipDeviceSynchronize());
    const size_t n = bytes / sizeof(double);
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(P_host[i])) {
            set_error("the solution holds NaN or Inf (first at %zu): non-finite right-hand side?", i);
            return DNSP_ERR_NONFINITE;
        }
    return DNSP_OK;
}

int dnsp_poisson_min_pivot(const dnsp_poisson* h, double* out) {
    clear_error();
    DNSP_ARG_CHECK(h && out, "NULL handle or pointer");
    *out = h->min_pivot;
    return DNSP_OK;
}

int dnsp_poisson_shape(const dnsp_poisson* h, int* nx, int* nz, int* ny) {
    clear_error();
    DNSP_ARG_CHECK(h && nx && nz && ny, "NULL handle or pointer");
    *nx = h->nx;
    *nz = h->nz;
    *ny = h->ny;
    return DNSP_OK;
}

int dnsp_poisson_free(dnsp_poisson* h) {
    clear_error();
    poisson_release(h);
    return DNSP_OK;
}

}  // extern "C"
