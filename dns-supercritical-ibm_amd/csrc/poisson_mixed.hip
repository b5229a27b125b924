// Mixed-precision Poisson solve (FP32 transforms and spectra, FP64
// tridiagonal arithmetic and FP64 (0,0) mode, FP64 residual refinement) and
// the FP64 residual norm.  Kernels of the sweeps are in poisson_mixed.h; the
// operator is the one the FP64 solve of poisson.hip inverts
// (poisson_solver_multigpu.f90:132-155 eigenvalues, :395-422 bands).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <vector>

#include "common.h"
#include "poisson_ct.h"
#include "poisson_ct_launch.h"
#include "poisson_internal.h"

namespace dnsp {

// Per-plane constants of a refinement residual: the iterate is p = pt + pbar(k)
// with pt a field whose FP64 line sums are lsum.  With m(k) = pbar(k) +
// (sum of plane k of pt)/(nx*nz), the plane means of p,
//   q(k) = (L_y pbar)(k) + ((L00_y - L_y) m)(k)
// (L_y, L00_y: the plain and the (0,0)-mode y band operators), so that
// L p = (L_xz + L_y) pt + q(k).  pbar may be NULL (zero).  One wavefront per plane.
__global__ __launch_bounds__(64) void k_mixed_q(double* __restrict__ q, const double* __restrict__ lsum,
                                                const double* __restrict__ pbar, const double* __restrict__ bands,
                                                int nz, int ny, double inv_nxnz) {
    const int k = (int)blockIdx.x;
    const int lane = (int)threadIdx.x;
    double m[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        const int kk = k - 1 + u;
        const bool in = kk >= 0 && kk < ny;
        double acc = 0.0;
        if (in)
            for (int j = lane; j < nz; j += 64) acc += lsum[(size_t)kk * nz + j];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        m[u] = in ? acc * inv_nxnz + (pbar ? pbar[kk] : 0.0) : 0.0;
    }
    if (lane == 0) {
        const double* a = bands;
        const double* b = bands + ny;
        const double* c = bands + 2 * ny;
        const double* a0 = bands + 3 * ny;
        const double* b0 = bands + 4 * ny;
        const double* c0 = bands + 5 * ny;
        double v = 0.0;
        if (pbar) {
            v += b[k] * pbar[k];
            if (k > 0) v += a[k] * pbar[k - 1];
            if (k < ny - 1) v += c[k] * pbar[k + 1];
        }
        v += (b0[k] - b[k]) * m[1];
        if (k > 0) v += (a0[k] - a[k]) * m[0];
        if (k < ny - 1) v += (c0[k] - c[k]) * m[2];
        q[k] = v;
    }
}

// FP64 sums of the x lines of P (one wavefront per line).
__global__ __launch_bounds__(256) void k_line_sums(const double* __restrict__ P, double* __restrict__ lsum, int nx,
                                                   long nlines) {
    const long line = (long)blockIdx.x * 4 + (long)(threadIdx.x / 64);
    const int lane = (int)threadIdx.x & 63;
    if (line >= nlines) return;  // wavefront-uniform
    double acc = 0.0;
    for (int i = lane; i < nx; i += 64) acc += P[line * nx + i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) lsum[line] = acc;
}

// Per-workgroup counts of the NaN and Inf values of a (grid kResBlocks x 256).
__global__ __launch_bounds__(256) void k_count_nonfinite(const double* __restrict__ a, long n,
                                                         double* __restrict__ part) {
    __shared__ unsigned int sc[4];
    unsigned int c = 0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        c += isfinite(a[i]) ? 0u : 1u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) sc[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (double)((sc[0] + sc[1]) + (sc[2] + sc[3]));
}

// Partial sums of (f - L p)^2 and f^2: grid kResBlocks x 256, fixed order.
// L p = periodic x, z second differences + plain y bands + q(k).
__global__ __launch_bounds__(256) void k_resid_norm(const double* __restrict__ p, const double* __restrict__ f,
                                                    const double* __restrict__ q, const double* __restrict__ bands,
                                                    int nx, int nz, int ny, double ix2, double iz2,
                                                    double* __restrict__ part) {
    __shared__ double sr[4], sf[4];
    const long plane = (long)nx * nz;
    double ar = 0.0, af = 0.0;
    // workgroups stride over x lines (plane k, index j: one division per line)
    for (long line = blockIdx.x; line < (long)nz * ny; line += gridDim.x) {
        const int k = (int)(line / nz);
        const int j = (int)(line - (long)k * nz);
        const long base = line * nx;
        const long zmo = j == 0 ? (long)(nz - 1) * nx : -(long)nx;
        const long zpo = j == nz - 1 ? -(long)(nz - 1) * nx : (long)nx;
        const double bk = bands[ny + k], ak = bands[k], ck = bands[2 * ny + k], qk = q[k];
        for (int i = (int)threadIdx.x; i < nx; i += 256) {
            const long idx = base + i;
            const double c = p[idx];
            const double xm = p[idx + (i == 0 ? nx - 1 : -1)];
            const double xp = p[idx + (i == nx - 1 ? 1 - nx : 1)];
            const double zm = p[idx + zmo];
            const double zp = p[idx + zpo];
            double v = ix2 * (xp - 2 * c + xm) + iz2 * (zp - 2 * c + zm) + bk * c;
            if (k > 0) v += ak * p[idx - plane];
            if (k < ny - 1) v += ck * p[idx + plane];
            const double fv = f[idx];
            const double r = fv - (v + qk);
            ar += r * r;
            af += fv * fv;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ar += __shfl_xor(ar, o, 64);
        af += __shfl_xor(af, o, 64);
    }
    const int w = (int)threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
        sr[w] = ar;
        sf[w] = af;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = (sr[0] + sr[1]) + (sr[2] + sr[3]);
        part[2 * blockIdx.x + 1] = (sf[0] + sf[1]) + (sf[2] + sf[3]);
    }
}

static int make_twiddles32(int n, float2** dev) {
    std::vector<float2> tw((size_t)n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int t = 0; t < n; ++t) {
        const long double ang = 2.0L * pi * (long double)t / (long double)n;
        tw[(size_t)t] = make_float2((float)cosl(ang), (float)(-sinl(ang)));
    }
    DNSP_HIP_TRY(hipMalloc(dev, sizeof(float2) * (size_t)n));
    DNSP_HIP_TRY(hipMemcpy(*dev, tw.data(), sizeof(float2) * (size_t)n, hipMemcpyHostToDevice));
    return DNSP_OK;
}

static size_t small_len(const dnsp_poisson* h) { return 4 * (size_t)h->ny + 2 * (size_t)kResBlocks; }

// The FP32 sweeps keep their accuracy where a mode's y system is well
// conditioned: with |lambda| >= max_k(|a_k| + |c_k|) / kF32Cond its condition
// number is below about 2 kF32Cond, so FP32 rounding costs no more than the
// FP32 transforms do.  Lower modes (the (0,0) mode and the lowest kx, kz)
// keep FP64 arithmetic.
constexpr double kF32Cond = 128.0;

// Group lists of the mixed tridiagonal (see dnsp_poisson::d_glist): a group
// goes to the FP32 list when every valid mode in it has |lambda| >= f32_min.
static int make_group_lists(dnsp_poisson* h) {
    const int nx = h->nx, nz = h->nz, ny = h->ny;
    std::vector<double> ax((size_t)nx), az((size_t)nz), bands((size_t)3 * ny);
    DNSP_HIP_TRY(hipMemcpy(ax.data(), h->d_ax, sizeof(double) * nx, hipMemcpyDeviceToHost));
    DNSP_HIP_TRY(hipMemcpy(az.data(), h->d_az, sizeof(double) * nz, hipMemcpyDeviceToHost));
    DNSP_HIP_TRY(hipMemcpy(bands.data(), h->d_bands, sizeof(double) * 3 * ny, hipMemcpyDeviceToHost));
    double cmax = 0.0;
    for (int k = 0; k < ny; ++k) cmax = std::max(cmax, std::fabs(bands[(size_t)k]) + std::fabs(bands[(size_t)2 * ny + k]));
    h->f32_min = cmax / kF32Cond;
    const long plane = (long)nx * nz;
    std::vector<int> all;
    const int Ms[3] = {8, 16, 32};
    for (int mi = 0; mi < 3; ++mi) {
        const int M = Ms[mi];
        const long ngroups = (plane + M - 1) / M;
        std::vector<int> lo, hi;
        lo.push_back(0);
        lo.push_back((int)ngroups);  // modes 1..M-1
        for (long g = 1; g < ngroups; ++g) {
            bool f32 = true;
            for (long m = g * M; m < std::min(plane, (g + 1) * M) && f32; ++m) {
                const double lam = ax[(size_t)(m % nx)] + az[(size_t)(m / nx)];
                f32 = std::fabs(lam) >= h->f32_min;
            }
            (f32 ? hi : lo).push_back((int)g);
        }
        h->gl_off[mi][0] = (int)all.size();
        h->gl_n[mi][0] = (int)lo.size();
        all.insert(all.end(), lo.begin(), lo.end());
        h->gl_off[mi][1] = (int)all.size();
        h->gl_n[mi][1] = (int)hi.size();
        all.insert(all.end(), hi.begin(), hi.end());
    }
    h->gl_total = all.size();
    DNSP_HIP_TRY(hipMalloc(&h->d_glist, sizeof(int) * all.size()));
    DNSP_HIP_TRY(hipMemcpy(h->d_glist, all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice));
    return DNSP_OK;
}

int mixed_setup(dnsp_poisson* h) {
    const size_t nlines = (size_t)h->nz * h->ny;
    if (!h->d_small) {
        DNSP_HIP_TRY(hipMalloc(&h->d_small, sizeof(double) * small_len(h)));
        DNSP_HIP_TRY(hipMemset(h->d_small, 0, sizeof(double) * small_len(h)));
        DNSP_HIP_TRY(hipMalloc(&h->d_lsA, sizeof(double) * nlines));
        DNSP_HIP_TRY(hipMalloc(&h->d_lsB, sizeof(double) * nlines));
    }
    if (h->precision != DNSP_PRECISION_MIXED) return DNSP_OK;
    h->mrows = part_rows_for(h->ny);
    if (!ct_fft_supported(h->nx) || !ct_fft_supported(h->nz) || h->mrows == 0) {
        set_error("mixed-precision form: nx=%d and nz=%d must be in {32,64,96,128,192,256,384,512,768} "
                  "and ny=%d <= %d",
                  h->nx, h->nz, h->ny, 32 * kPartChunks);
        return DNSP_ERR_UNSUPPORTED;
    }
    const size_t N = (size_t)h->nx * h->nz * h->ny;
    int rc = make_twiddles32(h->nx, &h->d_twx32);
    if (!rc) rc = make_twiddles32(h->nz, &h->d_twz32);
    if (rc) return rc;
    DNSP_HIP_TRY(hipMalloc(&h->d_A, sizeof(float) * N));
    DNSP_HIP_TRY(hipMalloc(&h->d_B, sizeof(float) * N));
    h->mtvar = 0;
    return make_group_lists(h);
}

size_t poisson_device_bytes(const dnsp_poisson* h) {
    if (!h) return 0;
    const size_t nx = (size_t)h->nx, nz = (size_t)h->nz, ny = (size_t)h->ny, N = nx * nz * ny;
    size_t b = sizeof(double) * (nx + nz + 6 * ny) + sizeof(int) * (nx + nz);  // eigenvalues, flags, bands
    if (h->d_twx) b += sizeof(double2) * nx;
    if (h->d_twz) b += sizeof(double2) * nz;
    if (h->d_cp) b += sizeof(double) * N;
    if (h->d_W) b += sizeof(double) * N;
    if (h->d_host) b += sizeof(double) * N;
    if (h->d_small) b += sizeof(double) * small_len(h) + 2 * sizeof(double) * nz * ny;  // + line sums
    if (h->d_twx32) b += sizeof(float2) * nx;
    if (h->d_twz32) b += sizeof(float2) * nz;
    if (h->d_A) b += sizeof(float) * N;
    if (h->d_B) b += sizeof(float) * N;
    if (h->d_fcopy) b += sizeof(double) * N;
    if (h->d_glist) b += sizeof(int) * h->gl_total;
    return b;
}

void mixed_release(dnsp_poisson* h) {
    (void)hipFree(h->d_twx32);
    (void)hipFree(h->d_twz32);
    (void)hipFree(h->d_A);
    (void)hipFree(h->d_B);
    (void)hipFree(h->d_lsA);
    (void)hipFree(h->d_lsB);
    (void)hipFree(h->d_small);
    (void)hipFree(h->d_fcopy);
    (void)hipFree(h->d_glist);
}

namespace {

struct Small {
    double *msum, *pbar0, *pbar1, *q, *part;
};

Small small_of(const dnsp_poisson* h) {
    double* s = h->d_small;
    const int ny = h->ny;
    return {s, s + ny, s + 2 * ny, s + 3 * ny, s + 4 * ny};
}

double inv_nxnz_of(const dnsp_poisson* h) { return 1.0 / ((double)h->nx * (double)h->nz); }

MxArgs base_args(const dnsp_poisson* h) {
    MxArgs a;
    a.nz = h->nz;
    a.ny = h->ny;
    a.ix2 = h->inv_dx * h->inv_dx;
    a.iz2 = h->inv_dz * h->inv_dz;
    a.nlines = (long)h->nz * h->ny;
    a.tw = h->d_twx32;
    a.scale = inv_nxnz_of(h);
    return a;
}

// The tridiagonal stage of a mixed sweep on the FP32 spectrum S: the FP64
// groups, then the FP32 ones (or every group in FP64 when the FP32 sweeps
// are off).
void mixed_tri(const dnsp_poisson* h, int variant, float* S, const double* msum, double* pbar, hipStream_t st) {
    const double inv = inv_nxnz_of(h);
    if (!h->f32_sweeps) {
        part_launch<float, double>(variant, S, h->d_ax, h->d_az, h->d_bands, h->nx, h->nz, h->ny, h->pcr_min, msum,
                                   pbar, inv, st);
        return;
    }
    const int M = part_group_modes(variant, h->ny, true);
    const int mi = M == 8 ? 0 : (M == 16 ? 1 : 2);
    part_launch<float, double>(variant, S, h->d_ax, h->d_az, h->d_bands, h->nx, h->nz, h->ny, h->pcr_min, msum, pbar,
                               inv, st, h->d_glist + h->gl_off[mi][0], h->gl_n[mi][0]);
    part_launch<float, float>(variant, S, h->d_ax, h->d_az, h->d_bands, h->nx, h->nz, h->ny, h->pcr_min, msum, pbar,
                              inv, st, h->d_glist + h->gl_off[mi][1], h->gl_n[mi][1]);
}

// Sweep 0: f -> d_A holds the FP32 field pt (line sums in d_lsA) and pbar0 the
// FP64 plane-mean profile; with `out` it writes out = pt + pbar0 in FP64 instead.
void sweep0(dnsp_poisson* h, const double* f, double* out, hipStream_t st) {
    const Small sm = small_of(h);
    MxArgs a = base_args(h);
    a.src = f;
    a.dst = h->d_A;
    a.lsum = h->d_lsA;
    mx_x_launch(MX_FWD, h->nx, h->mxv[MX_FWD], a, st);
    ct_z_launch(h->nz, h->mzvar, false, h->d_A, h->nx, h->ny, h->d_twz32, h->d_lsA, sm.msum, st);
    mixed_tri(h, h->mtvar, h->d_A, sm.msum, sm.pbar0, st);
    ct_z_launch(h->nz, h->mzvar, true, h->d_A, h->nx, h->ny, h->d_twz32, nullptr, nullptr, st);
    MxArgs b = base_args(h);
    b.src = h->d_A;
    if (out) {
        b.base = nullptr;
        b.pbar0 = sm.pbar0;
        b.dst = out;
        mx_x_launch(MX_ACC_F32, h->nx, h->mxv[MX_ACC_F32], b, st);
    } else {
        b.dst = h->d_A;
        b.lsum = h->d_lsA;
        mx_x_launch(MX_BWD, h->nx, h->mxv[MX_BWD], b, st);
    }
}

// One refinement sweep.  first: the iterate is (d_A, pbar0) from sweep 0,
// else the FP64 P with its line sums in d_lsA.  Writes P; with `sums` also
// the line sums of P into d_lsA.
void refine(dnsp_poisson* h, const double* f, double* P, bool first, bool sums, hipStream_t st) {
    const Small sm = small_of(h);
    k_mixed_q<<<h->ny, 64, 0, st>>>(sm.q, h->d_lsA, first ? sm.pbar0 : nullptr, h->d_bands, h->nz, h->ny,
                                    inv_nxnz_of(h));
    MxArgs a = base_args(h);
    a.src = first ? static_cast<const void*>(h->d_A) : static_cast<const void*>(P);
    a.f = f;
    a.q = sm.q;
    a.bands = h->d_bands;
    a.dst = h->d_B;
    a.lsum = h->d_lsB;
    mx_x_launch(first ? MX_RES_F32 : MX_RES_F64, h->nx, h->mxv[first ? MX_RES_F32 : MX_RES_F64], a, st);
    ct_z_launch(h->nz, h->mzvar, false, h->d_B, h->nx, h->ny, h->d_twz32, h->d_lsB, sm.msum, st);
    mixed_tri(h, h->mtvar, h->d_B, sm.msum, sm.pbar1, st);
    ct_z_launch(h->nz, h->mzvar, true, h->d_B, h->nx, h->ny, h->d_twz32, nullptr, nullptr, st);
    MxArgs b = base_args(h);
    b.src = h->d_B;
    b.base = first ? static_cast<const void*>(h->d_A) : static_cast<const void*>(P);
    b.pbar0 = first ? sm.pbar0 : nullptr;
    b.pbar1 = sm.pbar1;
    b.dst = P;
    b.lsum = sums ? h->d_lsA : nullptr;
    mx_x_launch(first ? MX_ACC_F32 : MX_ACC_F64, h->nx, h->mxv[first ? MX_ACC_F32 : MX_ACC_F64], b, st);
}

// rel = ||f - L P|| / ||f|| (||f - L P|| when f = 0), with P's line sums in
// lsum; q (ny) and part (2 kResBlocks) are device scratch, host_part (2
// kResBlocks) host memory.  Synchronises the stream.
int resid_norm_ws(const dnsp_poisson* h, const double* f, const double* P, const double* lsum, double* q,
                  double* part, double* host_part, double* rel, hipStream_t st) {
    k_mixed_q<<<h->ny, 64, 0, st>>>(q, lsum, nullptr, h->d_bands, h->nz, h->ny, inv_nxnz_of(h));
    k_resid_norm<<<kResBlocks, 256, 0, st>>>(P, f, q, h->d_bands, h->nx, h->nz, h->ny, h->inv_dx * h->inv_dx,
                                             h->inv_dz * h->inv_dz, part);
    DNSP_LAUNCH_CHECK();
    DNSP_HIP_TRY(hipMemcpyAsync(host_part, part, sizeof(double) * 2 * kResBlocks, hipMemcpyDeviceToHost, st));
    DNSP_HIP_TRY(hipStreamSynchronize(st));
    long double sr = 0.0L, sf = 0.0L;
    for (int b = 0; b < kResBlocks; ++b) {
        sr += host_part[2 * (size_t)b];
        sf += host_part[2 * (size_t)b + 1];
    }
    *rel = (double)(sf > 0 ? sqrtl(sr / sf) : sqrtl(sr));
    return DNSP_OK;
}

int resid_norm(dnsp_poisson* h, const double* f, const double* P, const double* lsum, double* rel, hipStream_t st) {
    const Small sm = small_of(h);
    std::vector<double> part(2 * (size_t)kResBlocks);
    return resid_norm_ws(h, f, P, lsum, sm.q, sm.part, part.data(), rel, st);
}

// Scratch of dnsp_poisson_residual_noalloc in a file-scope device array: the
// line sums (up to kNaLines), the plane constants q (up to kNaRows) and the
// reduction partials.  Each GPU the library runs on holds its own copy
// (2.1 MB) until the process ends.  Host code reaches it only through the
// address hipGetSymbolAddress returns for the current device (the array's
// name is a host-side stand-in, not a device address).
constexpr long kNaLines = 1L << 18;
constexpr int kNaRows = 4096;
__device__ __align__(256) double scratch[kNaLines + kNaRows + 2 * kResBlocks];

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mutex;
double* g_na_ptr[kNaMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device.
int noalloc_scratch(int device, double** out) {
    DNSP_ARG_CHECK(device >= 0 && device < kNaMaxDevices, "device %d: at most %d devices", device,
                   kNaMaxDevices);
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_ptr[device]) {
        void* ptr = nullptr;
        DNSP_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_na_ptr[device] = static_cast<double*>(ptr);
    }
    *out = g_na_ptr[device];
    return DNSP_OK;
}

// The mixed form's x kernels move (re, im) pairs of rhs and P as 16-byte
// accesses.
bool aligned16(const void* a, const void* b) {
    return ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0;
}

int check_mixed(const dnsp_poisson* h) {
    DNSP_ARG_CHECK(h, "NULL handle");
    if (h->precision != DNSP_PRECISION_MIXED) {
        set_error("handle was created with precision %d; the mixed-precision form needs DNSP_PRECISION_MIXED",
                  h->precision);
        return DNSP_ERR_ARG;
    }
    return DNSP_OK;
}

// The right-hand side the sweeps read: rhs itself, or a copy when the solution
// overwrites it before a later sweep reads it again.
int rhs_for(dnsp_poisson* h, const double* rhs, const double* P, bool reread, hipStream_t st, const double** f) {
    *f = rhs;
    if (rhs != P || !reread) return DNSP_OK;
    const size_t bytes = sizeof(double) * (size_t)h->nx * h->nz * h->ny;
    if (!h->d_fcopy) DNSP_HIP_TRY(hipMalloc(&h->d_fcopy, bytes));
    DNSP_HIP_TRY(hipMemcpyAsync(h->d_fcopy, rhs, bytes, hipMemcpyDeviceToDevice, st));
    *f = h->d_fcopy;
    return DNSP_OK;
}

}  // namespace

int mixed_autotune(dnsp_poisson* h, double* f, int reps, hipStream_t st, double* times_ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        set_error("autotune: event creation failed");
        return DNSP_ERR_HIP;
    }
    const Small sm = small_of(h);
    // one x kernel of the sweeps, on the handle's arrays (f: the work field)
    auto x_kind = [&](int kind, int v) {
        MxArgs a = base_args(h);
        if (kind == MX_FWD) {
            a.src = f;
            a.dst = h->d_A;
            a.lsum = h->d_lsA;
        } else if (kind == MX_BWD) {
            a.src = h->d_A;
            a.dst = h->d_A;
            a.lsum = h->d_lsA;
        } else if (kind == MX_RES_F32) {
            a.src = h->d_A;
            a.f = f;
            a.q = sm.q;
            a.bands = h->d_bands;
            a.dst = h->d_B;
            a.lsum = h->d_lsB;
        } else {
            a.src = h->d_B;
            a.base = h->d_A;
            a.pbar0 = sm.pbar0;
            a.pbar1 = sm.pbar1;
            a.dst = f;
        }
        mx_x_launch(kind, h->nx, v, a, st);
    };
    auto time_of = [&](auto&& fn, double* per) -> int {
        fn();  // warm-up
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) fn();
        (void)hipEventRecord(e1, st);
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            set_error("mixed autotune: launch or synchronize failed");
            return DNSP_ERR_HIP;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *per = (double)ms / reps;
        return DNSP_OK;
    };
    int rc = DNSP_OK;
    // x: each kernel kind of the first two sweeps on its own
    const int kinds[4] = {MX_FWD, MX_BWD, MX_RES_F32, MX_ACC_F32};
    int pickx[4];
    double bestx[4];
    for (int i = 0; i < 4; ++i) {
        pickx[i] = h->mxv[kinds[i]];
        bestx[i] = 1e30;
    }
    for (int v = 0; v < kCtVariants && rc == DNSP_OK; ++v) {
        double sum = 0.0;
        for (int i = 0; i < 4 && rc == DNSP_OK; ++i) {
            double t = 0.0;
            rc = time_of([&]() { x_kind(kinds[i], v); }, &t);
            sum += t;
            if (!rc && t < bestx[i]) {
                bestx[i] = t;
                pickx[i] = v;
            }
        }
        if (times_ms && !rc) times_ms[kTuneSlot[0] + v] = sum;
    }
    // z (forward + inverse pairs) and the tridiagonal (pairs of solves)
    int pick[3] = {h->mxvar, h->mzvar, h->mtvar};
    double best[3] = {1e30, 1e30, 1e30};
    for (int dir = 1; dir < 3 && rc == DNSP_OK; ++dir) {
        const int nvar = dir == 2 ? kPartVariants : kCtVariants;
        for (int v = 0; v < nvar && rc == DNSP_OK; ++v) {
            double per = 0.0;
            rc = time_of(
                [&]() {
                    if (dir == 1) {
                        ct_z_launch(h->nz, v, false, h->d_A, h->nx, h->ny, h->d_twz32, nullptr, nullptr, st);
                        ct_z_launch(h->nz, v, true, h->d_A, h->nx, h->ny, h->d_twz32, nullptr, nullptr, st);
                    } else {
                        for (int r = 0; r < 2; ++r) mixed_tri(h, v, h->d_A, nullptr, nullptr, st);
                    }
                },
                &per);
            if (rc) break;
            if (times_ms) times_ms[kTuneSlot[dir] + v] = per;
            if (per < best[dir]) {
                best[dir] = per;
                pick[dir] = v;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    for (int i = 0; i < 4; ++i) h->mxv[kinds[i]] = pickx[i];
    h->mxvar = pickx[0];
    h->mzvar = pick[1];
    h->mtvar = pick[2];
    return DNSP_OK;
}

}  // namespace dnsp

using namespace dnsp;

extern "C" {

int dnsp_poisson_solve_mixed(dnsp_poisson* h, const double* rhs, double* P, int refinements, void* stream) {
    clear_error();
    int rc = check_mixed(h);
    if (rc) return rc;
    DNSP_ARG_CHECK(rhs && P, "NULL array");
    DNSP_ARG_CHECK(refinements >= 0 && refinements <= DNSP_MAX_REFINEMENTS, "refinements=%d not in [0, %d]",
                   refinements, DNSP_MAX_REFINEMENTS);
    DNSP_ARG_CHECK(aligned16(rhs, P), "the mixed form needs 16-byte aligned rhs and P");
    hipStream_t st = as_stream(stream);
    const double* f = rhs;
    rc = rhs_for(h, rhs, P, refinements > 1, st, &f);
    if (rc) return rc;
    sweep0(h, f, refinements == 0 ? P : nullptr, st);
    for (int r = 0; r < refinements; ++r) refine(h, f, P, r == 0, r + 1 < refinements, st);
    DNSP_LAUNCH_CHECK();
    return DNSP_OK;
}

int dnsp_poisson_run_mixed(dnsp_poisson* h, double* P, int refinements, void* stream) {
    return dnsp_poisson_solve_mixed(h, P, P, refinements, stream);
}

int dnsp_poisson_solve_refine(dnsp_poisson* h, const double* rhs, double* P, double tol, int max_refinements,
                              int* refinements_done, double* rel_resid, void* stream) {
    clear_error();
    int rc = check_mixed(h);
    if (rc) return rc;
    DNSP_ARG_CHECK(rhs && P, "NULL array");
    DNSP_ARG_CHECK(aligned16(rhs, P), "the mixed form needs 16-byte aligned rhs and P");
    DNSP_ARG_CHECK(tol >= 0 && std::isfinite(tol), "tol must be finite and >= 0");
    DNSP_ARG_CHECK(max_refinements >= 0 && max_refinements <= DNSP_MAX_REFINEMENTS,
                   "max_refinements=%d not in [0, %d]", max_refinements, DNSP_MAX_REFINEMENTS);
    hipStream_t st = as_stream(stream);
    const double* f = rhs;
    rc = rhs_for(h, rhs, P, true, st, &f);
    if (rc) return rc;
    double rel = 0.0;
    int done = 0;
    if (max_refinements == 0) {
        sweep0(h, f, P, st);
        k_line_sums<<<(int)(((long)h->nz * h->ny + 3) / 4), 256, 0, st>>>(P, h->d_lsA, h->nx, (long)h->nz * h->ny);
        rc = resid_norm(h, f, P, h->d_lsA, &rel, st);
    } else {
        sweep0(h, f, nullptr, st);
        for (int r = 0; r < max_refinements; ++r) {
            refine(h, f, P, r == 0, true, st);
            done = r + 1;
            rc = resid_norm(h, f, P, h->d_lsA, &rel, st);
            if (rc || rel <= tol) break;
        }
    }
    if (rc) return rc;
    if (refinements_done) *refinements_done = done;
    if (rel_resid) *rel_resid = rel;
    if (!std::isfinite(rel)) {
        set_error("mixed solve: the relative residual is %g after %d refinements (non-finite right-hand side?)", rel,
                  done);
        return DNSP_ERR_NONFINITE;
    }
    if (rel > tol) {
        set_error("mixed solve: relative residual %.3e above tol %.3e after %d refinements", rel, tol, done);
        return DNSP_ERR_NOT_CONVERGED;
    }
    return DNSP_OK;
}

int dnsp_poisson_residual(dnsp_poisson* h, const double* rhs, const double* P, double* rel_resid, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(h && rhs && P && rel_resid, "NULL handle, array or output");
    hipStream_t st = as_stream(stream);
    const long nlines = (long)h->nz * h->ny;
    k_line_sums<<<(int)((nlines + 3) / 4), 256, 0, st>>>(P, h->d_lsB, h->nx, nlines);
    const int rc = resid_norm(h, rhs, P, h->d_lsB, rel_resid, st);
    if (rc) return rc;
    if (!std::isfinite(*rel_resid)) {
        set_error("residual is %g: the solution or the right-hand side holds a NaN or Inf", *rel_resid);
        return DNSP_ERR_NONFINITE;
    }
    return DNSP_OK;
}

int dnsp_poisson_residual_noalloc(dnsp_poisson* h, const double* rhs, const double* P, double* rel_resid,
                                  void* stream) {
    clear_error();
    DNSP_ARG_CHECK(h && rhs && P && rel_resid, "NULL handle, array or output");
    const long nlines = (long)h->nz * h->ny;
    DNSP_ARG_CHECK(nlines <= kNaLines && h->ny <= kNaRows,
                   "nz*ny = %ld lines and ny = %d: the static scratch of dnsp_poisson_residual_noalloc holds at "
                   "most %ld lines and %d rows",
                   nlines, h->ny, kNaLines, kNaRows);
    int dev = -1;
    DNSP_HIP_TRY(hipGetDevice(&dev));
    DNSP_ARG_CHECK(dev == h->device, "the handle lives on device %d but device %d is current", h->device, dev);
    double* ws = nullptr;
    int rc = noalloc_scratch(dev, &ws);
    if (rc) return rc;
    double* lsum = ws;
    double* q = ws + kNaLines;
    double* part = q + kNaRows;
    hipStream_t st = as_stream(stream);
    k_line_sums<<<(int)((nlines + 3) / 4), 256, 0, st>>>(P, lsum, h->nx, nlines);
    double host_part[2 * kResBlocks];
    rc = resid_norm_ws(h, rhs, P, lsum, q, part, host_part, rel_resid, st);
    if (rc) return rc;
    if (!std::isfinite(*rel_resid)) {
        set_error("residual is %g: the solution or the right-hand side holds a NaN or Inf", *rel_resid);
        return DNSP_ERR_NONFINITE;
    }
    return DNSP_OK;
}

int dnsp_poisson_check(dnsp_poisson* h, const double* P, long long* n_nonfinite, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(h && P, "NULL handle or array");
    hipStream_t st = as_stream(stream);
    const Small sm = small_of(h);
    const long n = (long)h->nx * h->nz * h->ny;
    k_count_nonfinite<<<kResBlocks, 256, 0, st>>>(P, n, sm.part);
    DNSP_LAUNCH_CHECK();
    std::vector<double> part((size_t)kResBlocks);
    DNSP_HIP_TRY(hipMemcpyAsync(part.data(), sm.part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
    DNSP_HIP_TRY(hipStreamSynchronize(st));
    long long cnt = 0;
    for (double v : part) cnt += (long long)v;
    if (n_nonfinite) *n_nonfinite = cnt;
    if (cnt) {
        set_error("%lld of %ld values of the solution are NaN or Inf (non-finite right-hand side?)", cnt, n);
        return DNSP_ERR_NONFINITE;
    }
    return DNSP_OK;
}

int dnsp_poisson_set_mixed_fp32(dnsp_poisson* h, int on) {
    clear_error();
    int rc = check_mixed(h);
    if (rc) return rc;
    h->f32_sweeps = on != 0;
    return DNSP_OK;
}

int dnsp_poisson_mixed_info(const dnsp_poisson* h, int* fp32_on, double* f32_min, long long* fp32_modes) {
    clear_error();
    int rc = check_mixed(h);
    if (rc) return rc;
    if (fp32_on) *fp32_on = h->f32_sweeps ? 1 : 0;
    if (f32_min) *f32_min = h->f32_min;
    if (fp32_modes) {
        // modes of the FP32 groups of the variant in use
        const int M = part_group_modes(h->mtvar, h->ny, true);
        const int mi = M == 8 ? 0 : (M == 16 ? 1 : 2);
        const long plane = (long)h->nx * h->nz;
        long long n = 0;
        if (h->f32_sweeps) {
            std::vector<int> g((size_t)h->gl_n[mi][1]);
            if (!g.empty())
                DNSP_HIP_TRY(hipMemcpy(g.data(), h->d_glist + h->gl_off[mi][1], sizeof(int) * g.size(),
                                       hipMemcpyDeviceToHost));
            for (int q : g) n += std::min<long>(M, plane - (long)q * M);
        }
        *fp32_modes = n;
    }
    return DNSP_OK;
}

int dnsp_poisson_set_mixed_variants(dnsp_poisson* h, int xvar, int zvar, int tvar) {
    clear_error();
    int rc = check_mixed(h);
    if (rc) return rc;
    DNSP_ARG_CHECK(xvar >= 0 && xvar < kCtVariants && zvar >= 0 && zvar < kCtVariants,
                   "FFT kernel variants must be in [0, %d)", kCtVariants);
    DNSP_ARG_CHECK(tvar >= 0 && tvar < kPartVariants, "tridiagonal variant must be in [0, %d)", kPartVariants);
    h->mxvar = xvar;
    for (int k = 0; k < 6; ++k) h->mxv[k] = (k == MX_RES_F64 || k == MX_ACC_F64) ? 0 : xvar;
    h->mzvar = zvar;
    h->mtvar = tvar;
    return DNSP_OK;
}

int dnsp_poisson_get_mixed_variants(const dnsp_poisson* h, int* xvar, int* zvar, int* tvar) {
    clear_error();
    int rc = check_mixed(h);
    if (rc) return rc;
    DNSP_ARG_CHECK(xvar && zvar && tvar, "NULL argument");
    *xvar = h->mxvar;
    *zvar = h->mzvar;
    *tvar = h->mtvar;
    return DNSP_OK;
}

}  // extern "C"
