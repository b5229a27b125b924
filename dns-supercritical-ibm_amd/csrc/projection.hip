// Projection stencils around the Poisson solve (FP64, gfx950): halos and
// wall planes, the Poisson right-hand side, the rho_min reduction and the
// gradient correction, on the reference's halo arrays.
//
// Halo arrays are the reference's (-1:nx, -1:nz, -1:ny) fields, x fastest:
// element (i,j,k) at (i+1) + (nx+2)*((j+1) + (nz+2)*(k+1)).  P without halo
// is (0:nx-1, 0:nz-1, 0:ny-1).  Every kernel keeps the reference's
// evaluation order with contraction off, so results are bit-identical to a
// serial FP64 evaluation of the same expressions.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "common.h"
#include "stencil_grid.h"

namespace dnsp {

struct Halo {
    int nx, nz, ny;
    __host__ __device__ long sx() const { return 1; }
    __host__ __device__ long sz() const { return nx + 2; }
    __host__ __device__ long sy() const { return (long)(nx + 2) * (nz + 2); }
    __host__ __device__ long at(int i, int j, int k) const {
        return (long)(i + 1) + (long)(nx + 2) * ((long)(j + 1) + (long)(nz + 2) * (k + 1));
    }
    __host__ __device__ long total() const { return (long)(nx + 2) * (nz + 2) * (ny + 2); }
};

constexpr int kStencilThreads = 256;

static inline int blocks_for(long n, int t = kStencilThreads) { return (int)((n + t - 1) / t); }

// ---------------------------------------------------------------------------
// halos and wall planes
// ---------------------------------------------------------------------------

// Periodic x and z halos over the whole padded array (every k, ghosts
// included): the one-rank effect of the x pass followed by the z pass, i.e.
// A(i,j,k) = A(wrap(i), wrap(j), k) on the halo shell.
__global__ void k_halo_xz(double* A, Halo h) {
    const long per_plane = 2L * (h.nz + 2) + 2L * h.nx;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= per_plane * (h.ny + 2)) return;
    const int k = (int)(idx / per_plane) - 1;
    const long s = idx % per_plane;
    int i, j;
    if (s < 2L * (h.nz + 2)) {
        i = (s < h.nz + 2) ? -1 : h.nx;
        j = (int)(s % (h.nz + 2)) - 1;
    } else {
        const long t = s - 2L * (h.nz + 2);
        i = (int)(t % h.nx);
        j = (t < h.nx) ? -1 : h.nz;
    }
    const int wi = i < 0 ? h.nx - 1 : (i >= h.nx ? 0 : i);
    const int wj = j < 0 ? h.nz - 1 : (j >= h.nz ? 0 : j);
    A[h.at(i, j, k)] = A[h.at(wi, wj, k)];
}

// set_val_slice: padded planes kp0..kp1 (0 = the k = -1 ghost) set to val
__global__ void k_set_planes(double* A, long plane_elems, long first, long count, double val) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < count) A[first * plane_elems + idx] = val;
}

// Neumann ghosts in y: A(:,:,-1) = A(:,:,0), A(:,:,ny) = A(:,:,ny-1) over the padded plane
__global__ void k_neumann_y(double* A, Halo h) {
    const long pe = h.sy();
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 2 * pe) return;
    if (idx < pe) A[idx] = A[pe + idx];
    else A[(long)(h.ny + 1) * pe + (idx - pe)] = A[(long)h.ny * pe + (idx - pe)];
}

// ---------------------------------------------------------------------------
// Poisson right-hand side (mod_cfd_subroutines.f90:527-574)
// ---------------------------------------------------------------------------

__global__ void k_rhs_incomp(double* P, const double* __restrict__ Uh, const double* __restrict__ Wh,
                             const double* __restrict__ Vh, const double* __restrict__ inv_dy_p, Halo h,
                             BandGrid bg, double inv_dtmax, double inv_dx, double inv_dz) {
#pragma clang fp contract(off)
    int i, j, k;
    if (!band_point(bg, i, j, k)) return;
    const long idx = (long)i + (long)h.nx * ((long)j + (long)h.nz * k);
    const long c = h.at(i, j, k);
    P[idx] = inv_dtmax * (((Uh[c + 1] - Uh[c]) * inv_dx + (Wh[c + h.sz()] - Wh[c]) * inv_dz) +
                          (Vh[c + h.sy()] - Vh[c]) * inv_dy_p[k]);
}

__device__ __forceinline__ double rfac(double rho_min, double nb, double rc) {
#pragma clang fp contract(off)
    return 1.0 - rho_min * (0.5 / nb + 0.5 / rc);
}

// Variable-density right-hand side.  P_now, P_old and rho are read at seven
// points each, so they are staged in LDS: a workgroup owns a 64 x 4 (i, j)
// tile and marches through a run of kRunK planes, keeping planes k-1, k,
// k+1 of the three fields (tile plus one halo point in x and z) in a ring of
// three LDS slots and loading one new plane per step (issued before the
// step's arithmetic), so each of their values is read about once instead
// of by every neighbouring thread (0.36 -> 0.28 ms at 384x256x192; the
// per-thread version made 6.3 L1 requests per point).
__global__ __launch_bounds__(kTileX * kTileZ) void k_rhs_varprop(
    double* P, const double* __restrict__ Uh, const double* __restrict__ Wh, const double* __restrict__ Vh,
    const double* __restrict__ Pn, const double* __restrict__ Po, const double* __restrict__ rho,
    const double* __restrict__ div_now, const double* __restrict__ inv_dy_p, const double* __restrict__ inv_dy_p_pad,
    const double* __restrict__ rho_min_p, Halo h, BandGrid bg, double inv_dtmax, double inv_dx, double inv_dz,
    double inv_dx2, double inv_dz2) {
#pragma clang fp contract(off)
    __shared__ double sF[3][3][kRingP];  // P_now, P_old, rho
    int i0, j0, k0, k1;
    if (!ring_tile(bg, h.ny, i0, j0, k0, k1)) return;
    const int i = i0 + (int)threadIdx.x, j = j0 + (int)threadIdx.y;
    const bool active = i < h.nx && j < h.nz;
    const long sz = h.sz(), sy = h.sy();
    const double* fields[3] = {Pn, Po, rho};
    PlaneRing<3> ring(sF, fields, i0, j0, h.nx, h.nz);
    ring.fetch(k0 - 1);
    ring.put(k0 - 1);
    ring.fetch(k0);
    ring.put(k0);
    ring.fetch(k0 + 1);
    const double rho_min = *rho_min_p;
    const int e0 = ring_point();
    for (int k = k0; k < k1; ++k) {
        ring.put(k + 1);  // the slot of plane k - 2, free since the last barrier
        __syncthreads();
        if (k + 2 <= k1) ring.fetch(k + 2);  // in flight during this step's arithmetic
        if (active) {
            const int sm = (k + 2) % 3, s0 = k % 3, sp = (k + 1) % 3;
            const long c = h.at(i, j, k);
            auto ps = [&](int slot, int e) { return 2 * sF[0][slot][e] - sF[1][slot][e]; };
            auto rh = [&](int slot, int e) { return sF[2][slot][e]; };
            const double Ps_00 = ps(s0, e0);
            const double Ps_xp = ps(s0, e0 + 1), Ps_xm = ps(s0, e0 - 1);
            const double Ps_zp = ps(s0, e0 + kRingX), Ps_zm = ps(s0, e0 - kRingX);
            const double Ps_yp = ps(sp, e0), Ps_ym = ps(sm, e0);
            const double rc = rh(s0, e0);
            const double rho_xp = rfac(rho_min, rh(s0, e0 + 1), rc), rho_xm = rfac(rho_min, rh(s0, e0 - 1), rc);
            const double rho_zp = rfac(rho_min, rh(s0, e0 + kRingX), rc);
            const double rho_zm = rfac(rho_min, rh(s0, e0 - kRingX), rc);
            const double rho_yp = rfac(rho_min, rh(sp, e0), rc), rho_ym = rfac(rho_min, rh(sm, e0), rc);
            const double idy = inv_dy_p[k];
            const double t1 = inv_dtmax * (((Uh[c + 1] - Uh[c]) * inv_dx + (Wh[c + sz] - Wh[c]) * inv_dz) +
                                           (Vh[c + sy] - Vh[c]) * idy) * rho_min;
            const double t2 = inv_dtmax * rho_min * div_now[c];
            const double t3 = ((rho_xp * (Ps_xp - Ps_00) - rho_xm * (Ps_00 - Ps_xm)) * inv_dx2 +
                               (rho_zp * (Ps_zp - Ps_00) - rho_zm * (Ps_00 - Ps_zm)) * inv_dz2) +
                              (rho_yp * (Ps_yp - Ps_00) * inv_dy_p_pad[k + 1] -
                               rho_ym * (Ps_00 - Ps_ym) * inv_dy_p_pad[k]) * idy;
            P[(long)i + (long)h.nx * ((long)j + (long)h.nz * k)] = (t1 - t2) + t3;
        }
        __syncthreads();  // slot (k + 2) % 3 is restaged next
    }
}

// ---------------------------------------------------------------------------
// rho_min over the interior (mod_cfd_subroutines.f90:470-499): an ordered
// 64-bit key per double so a global atomicMin gives the minimum.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long dkey(double v) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double dunkey(unsigned long long k) {
    unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

__global__ void k_min_init(double* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *reinterpret_cast<unsigned long long*>(out) = 0xffffffffffffffffull;
}

// Workgroups stride over groups of kMinRows interior x rows; a group's
// kMinRows * nx values are spread over all the workgroup's threads (whole
// rows per thread left half the lanes idle at nx = 384), and each thread
// issues kMinLoads loads before comparing their keys.
constexpr int kMinRows = 8, kMinLoads = 4;
__global__ void k_min_interior(const double* __restrict__ A, Halo h, double* out) {
    __shared__ unsigned long long red[kStencilThreads];
    const int rows = h.nz * h.ny;
    const int per = kMinRows * h.nx;
    unsigned long long best = 0xffffffffffffffffull;
    for (int row0 = (int)blockIdx.x * kMinRows; row0 < rows; row0 += (int)gridDim.x * kMinRows) {
        for (int e0 = (int)threadIdx.x; e0 < per; e0 += kMinLoads * kStencilThreads) {
            double v[kMinLoads];
#pragma unroll
            for (int u = 0; u < kMinLoads; ++u) {
                const int e = e0 + u * kStencilThreads;
                const int r = (int)((unsigned)e / (unsigned)h.nx);
                const int i = e - r * h.nx;
                const int row = row0 + r;
                const int k = row / h.nz, j = row - k * h.nz;
                // past the group or the last row: a repeat of an interior value
                const long at = e < per && row < rows ? h.at(i, j, k) : h.at(0, 0, 0);
                v[u] = A[at];
            }
#pragma unroll
            for (int u = 0; u < kMinLoads; ++u) {
                const unsigned long long key = dkey(v[u]);
                best = key < best ? key : best;
            }
        }
    }
    red[threadIdx.x] = best;
    __syncthreads();
    for (int s = kStencilThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s && red[threadIdx.x + s] < red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMin(reinterpret_cast<unsigned long long*>(out), red[0]);
}

__global__ void k_min_decode(double* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = dunkey(*reinterpret_cast<unsigned long long*>(out));
}

// ---------------------------------------------------------------------------
// gradient correction (mod_cfd_subroutines.f90:1087-1216)
// P_per is the periodic/Neumann extension of P: P_per(i,j,k) = P(wrap i, wrap j, clamp k).
// ---------------------------------------------------------------------------
__device__ __forceinline__ long pidx(const Halo& h, int i, int j, int k) {
    i = i < 0 ? h.nx - 1 : (i >= h.nx ? 0 : i);
    j = j < 0 ? h.nz - 1 : (j >= h.nz ? 0 : j);
    k = k < 0 ? 0 : (k >= h.ny ? h.ny - 1 : k);
    return (long)i + (long)h.nx * ((long)j + (long)h.nz * k);
}

__global__ void k_update_incomp(const double* __restrict__ P, double* U, double* V, double* W,
                                const double* __restrict__ Uh, const double* __restrict__ Vh,
                                const double* __restrict__ Wh, const double* __restrict__ hi,
                                const double* __restrict__ lo, Halo h, BandGrid bg, int kmin_Vhat, double dtmax,
                                double inv_dx, double inv_dz) {
#pragma clang fp contract(off)
    int i, j, k;
    if (!band_point(bg, i, j, k)) return;
    const long idx = (long)i + (long)h.nx * ((long)j + (long)h.nz * k);
    const long c = h.at(i, j, k);
    const double p = P[idx];
    U[c] = Uh[c] - dtmax * (p - P[pidx(h, i - 1, j, k)]) * inv_dx;
    W[c] = Wh[c] - dtmax * (p - P[pidx(h, i, j - 1, k)]) * inv_dz;
    if (k >= kmin_Vhat) V[c] = Vh[c] - dtmax * (hi[k] * p + lo[k] * P[pidx(h, i, j, k - 1)]);
}

__global__ void k_update_varprop(const double* __restrict__ P, double* U, double* V, double* W,
                                 const double* __restrict__ Uh, const double* __restrict__ Vh,
                                 const double* __restrict__ Wh, const double* __restrict__ Pn,
                                 const double* __restrict__ Po, const double* __restrict__ rho,
                                 const double* __restrict__ hi, const double* __restrict__ lo,
                                 const double* __restrict__ rho_min_p, Halo h, BandGrid bg, int kmin_Vhat,
                                 double dtmax, double inv_dx, double inv_dz) {
#pragma clang fp contract(off)
    int i, j, k;
    if (!band_point(bg, i, j, k)) return;
    const long idx = (long)i + (long)h.nx * ((long)j + (long)h.nz * k);
    const long c = h.at(i, j, k);
    const double inv_rho_min = 1.0 / *rho_min_p;
    const double p = P[idx];
    const double ps_c = 2 * Pn[c] - Po[c];
    const double rc = rho[c];
    auto flux = [&](long cb, double pb) {
        const double inv_rho_now = 0.5 / rc + 0.5 / rho[cb];
        const double P_fwd = p * inv_rho_min + (inv_rho_now - inv_rho_min) * ps_c;
        const double P_bwd = pb * inv_rho_min + (inv_rho_now - inv_rho_min) * (2 * Pn[cb] - Po[cb]);
        return make_double2(P_fwd, P_bwd);
    };
    double2 f = flux(c - 1, P[pidx(h, i - 1, j, k)]);
    U[c] = Uh[c] - dtmax * (f.x - f.y) * inv_dx;
    f = flux(c - h.sz(), P[pidx(h, i, j - 1, k)]);
    W[c] = Wh[c] - dtmax * (f.x - f.y) * inv_dz;
    if (k >= kmin_Vhat) {
        f = flux(c - h.sy(), P[pidx(h, i, j, k - 1)]);
        V[c] = Vh[c] - dtmax * (hi[k] * f.x + lo[k] * f.y);
    }
}

// P_per over the whole padded array, and (varprop) P_old <- P_now, P_now <- P_per
// One thread per padded point, x fastest (a pure stream: no tiles, so no
// idle lanes on the nx + 2 wide rows); IDX32: the padded array has fewer
// than 2^31 points (32-bit index split).
template <bool IDX32>
__global__ void k_pper_rotate(const double* __restrict__ P, double* P_per, double* Pn, double* Po, Halo h,
                              int rotate) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= h.total()) return;
    int i, j, k;
    if constexpr (IDX32) {
        const unsigned px = (unsigned)(h.nx + 2), pe = (unsigned)h.sy(), x = (unsigned)idx;
        const unsigned kp = x / pe, r = x - kp * pe, jp = r / px;
        i = (int)(r - jp * px) - 1;
        j = (int)jp - 1;
        k = (int)kp - 1;
    } else {
        i = (int)(idx % (h.nx + 2)) - 1;
        j = (int)((idx / (h.nx + 2)) % (h.nz + 2)) - 1;
        k = (int)(idx / h.sy()) - 1;
    }
    const double v = P[pidx(h, i, j, k)];
    if (P_per) P_per[idx] = v;
    if (rotate) {
        Po[idx] = Pn[idx];
        Pn[idx] = v;
    }
}

// Velocity halos and wall planes after the correction (update_vars :1225-1244):
// ghost planes k = -1, ny (and k = 0 for V) take the wall values over the whole
// padded plane; other planes get periodic x/z halos.
__global__ void k_vel_finish(double* U, double* V, double* W, Halo h, double ub, double vb, double wb, double ut,
                             double vt, double wt) {
    const long pe = h.sy();
    const long per_plane = 2L * (h.nz + 2) + 2L * h.nx;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    // part 1: wall planes: U,W at k=-1,ny ; V at k=-1,0,ny
    const long nwall = 7 * pe;
    if (idx < nwall) {
        const int which = (int)(idx / pe);
        const long o = idx % pe;
        switch (which) {
            case 0: U[o] = ub; break;
            case 1: W[o] = wb; break;
            case 2: V[o] = vb; break;
            case 3: V[pe + o] = vb; break;
            case 4: U[(long)(h.ny + 1) * pe + o] = ut; break;
            case 5: W[(long)(h.ny + 1) * pe + o] = wt; break;
            default: V[(long)(h.ny + 1) * pe + o] = vt; break;
        }
        return;
    }
    // part 2: periodic halos of planes k = 0..ny-1 (U, W) and 1..ny-1 (V)
    const long r = idx - nwall;
    if (r >= 3 * per_plane * h.ny) return;
    const int f = (int)(r / (per_plane * h.ny));
    const long rr = r % (per_plane * h.ny);
    const int k = (int)(rr / per_plane);
    const long s = rr % per_plane;
    if (f == 1 && k == 0) return;  // V plane 0 is a wall plane
    int i, j;
    if (s < 2L * (h.nz + 2)) {
        i = (s < h.nz + 2) ? -1 : h.nx;
        j = (int)(s % (h.nz + 2)) - 1;
    } else {
        const long t = s - 2L * (h.nz + 2);
        i = (int)(t % h.nx);
        j = (t < h.nx) ? -1 : h.nz;
    }
    const int wi = i < 0 ? h.nx - 1 : (i >= h.nx ? 0 : i);
    const int wj = j < 0 ? h.nz - 1 : (j >= h.nz ? 0 : j);
    double* A = f == 0 ? U : (f == 1 ? V : W);
    A[h.at(i, j, k)] = A[h.at(wi, wj, k)];
}

static bool halo_ok(int nx, int nz, int ny) {
    return nx >= 1 && nz >= 1 && ny >= 1 && (double)(nx + 2) * (nz + 2) * (ny + 2) < 4e12;
}

}  // namespace dnsp

using namespace dnsp;

extern "C" {

int dnsp_halo_periodic_xz(double* A, int nx, int nz, int ny, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(A && halo_ok(nx, nz, ny), "invalid halo array arguments");
    Halo h{nx, nz, ny};
    const long n = (2L * (nz + 2) + 2L * nx) * (ny + 2);
    k_halo_xz<<<blocks_for(n), kStencilThreads, 0, as_stream(stream)>>>(A, h);
    DNSP_LAUNCH_CHECK();
    return DNSP_OK;
}

int dnsp_set_val_slice(double* A, int nx, int nz, int ny, int k0, int k1, double val, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(A && halo_ok(nx, nz, ny), "invalid halo array arguments");
    const int n2 = ny + 2;
    if (k0 < 0) k0 += n2;
    if (k1 < 0) k1 += n2;
    DNSP_ARG_CHECK(k0 >= 0 && k0 < n2 && k1 >= 0 && k1 < n2, "plane index out of range");
    const int kmin = k0 < k1 ? k0 : k1, kmax = k0 < k1 ? k1 : k0;
    Halo h{nx, nz, ny};
    const long count = (long)(kmax - kmin + 1) * h.sy();
    k_set_planes<<<blocks_for(count), kStencilThreads, 0, as_stream(stream)>>>(A, h.sy(), kmin, count, val);
    DNSP_LAUNCH_CHECK();
    return DNSP_OK;
}

int dnsp_neumann_ghosts_y(double* A, int nx, int nz, int ny, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(A && halo_ok(nx, nz, ny), "invalid halo array arguments");
    Halo h{nx, nz, ny};
    k_neumann_y<<<blocks_for(2 * h.sy()), kStencilThreads, 0, as_stream(stream)>>>(A, h);
    DNSP_LAUNCH_CHECK();
    return DNSP_OK;
}

int dnsp_build_p_rhs(double* P, const double* U_hat, const double* W_hat, const double* V_hat,
                     const double* inv_dy_p, int nx, int nz, int ny, double inv_dtmax, double inv_dx, double inv_dz,
                     void* stream) {
    clear_error();
    DNSP_ARG_CHECK(P && U_hat && W_hat && V_hat && inv_dy_p && halo_ok(nx, nz, ny), "invalid arguments");
    Halo h{nx, nz, ny};
    const BandGrid bg = band_grid(nx, nz, ny);
    k_rhs_incomp<<<bg.blocks(), band_block(), 0, as_stream(stream)>>>(P, U_hat, W_hat, V_hat, inv_dy_p, h, bg,
                                                                      inv_dtmax, inv_dx, inv_dz);
    DNSP_LAUNCH_CHECK();
    return DNSP_OK;
}

int dnsp_build_p_rhs_varprop(double* P, const double* U_hat, const double* W_hat, const double* V_hat,
                             const double* P_now, const double* P_old, const double* rho, const double* div_now,
                             const double* inv_dy_p, const double* inv_dy_p_pad, const double* rho_min, int nx,
                             int nz, int ny, double inv_dtmax, double inv_dx, double inv_dz, double inv_dx2,
                             double inv_dz2, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(P && U_hat && W_hat && V_hat && P_now && P_old && rho && div_now && inv_dy_p && inv_dy_p_pad &&
                       rho_min && halo_ok(nx, nz, ny),
                   "invalid arguments");
    Halo h{nx, nz, ny};
    const BandGrid bg = ring_grid(nx, nz, ny);
    k_rhs_varprop<<<bg.blocks(), band_block(), 0, as_stream(stream)>>>(
        P, U_hat, W_hat, V_hat, P_now, P_old, rho, div_now, inv_dy_p, inv_dy_p_pad, rho_min, h, bg, inv_dtmax, inv_dx,
        inv_dz, inv_dx2, inv_dz2);
    DNSP_LAUNCH_CHECK();
    return DNSP_OK;
}

int dnsp_rho_min(const double* rho, int nx, int nz, int ny, double* out, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(rho && out && halo_ok(nx, nz, ny), "invalid arguments");
    Halo h{nx, nz, ny};
    hipStream_t st = as_stream(stream);
    const int nb = std::min(2048, (nz * ny + kMinRows - 1) / kMinRows);
    k_min_init<<<1, 1, 0, st>>>(out);
    k_min_interior<<<nb, kStencilThreads, 0, st>>>(rho, h, out);
    k_min_decode<<<1, 1, 0, st>>>(out);
    DNSP_LAUNCH_CHECK();
    return DNSP_OK;
}

int dnsp_update_vars(const double* P, double* P_per, double* P_now, double* P_old, double* U, double* V, double* W,
                     const double* U_hat, const double* V_hat, const double* W_hat, const double* rho,
                     const double* dPdy_ddy_hi, const double* dPdy_ddy_lo, const double* rho_min,
                     const double* wall_bc, int nx, int nz, int ny, int kmin_Vhat, double dtmax, double inv_dx,
                     double inv_dz, int use_incompressible, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(P && U && V && W && U_hat && V_hat && W_hat && dPdy_ddy_hi && dPdy_ddy_lo && wall_bc &&
                       halo_ok(nx, nz, ny),
                   "invalid arguments");
    DNSP_ARG_CHECK(kmin_Vhat >= 0 && kmin_Vhat <= ny, "kmin_Vhat=%d", kmin_Vhat);
    DNSP_ARG_CHECK(use_incompressible || (P_now && P_old && rho && rho_min),
                   "variable-density correction needs P_now, P_old, rho and rho_min");
    Halo h{nx, nz, ny};
    hipStream_t st = as_stream(stream);
    const BandGrid bg = band_grid(nx, nz, ny);
    if (use_incompressible)
        k_update_incomp<<<bg.blocks(), band_block(), 0, st>>>(P, U, V, W, U_hat, V_hat, W_hat, dPdy_ddy_hi,
                                                              dPdy_ddy_lo, h, bg, kmin_Vhat, dtmax, inv_dx, inv_dz);
    else
        k_update_varprop<<<bg.blocks(), band_block(), 0, st>>>(P, U, V, W, U_hat, V_hat, W_hat, P_now, P_old, rho,
                                                               dPdy_ddy_hi, dPdy_ddy_lo, rho_min, h, bg, kmin_Vhat,
                                                               dtmax, inv_dx, inv_dz);
    DNSP_LAUNCH_CHECK();
    if (P_per || !use_incompressible) {
        auto kr = h.total() < (1L << 31) ? k_pper_rotate<true> : k_pper_rotate<false>;
        kr<<<blocks_for(h.total()), kStencilThreads, 0, st>>>(P, P_per, P_now, P_old, h, use_incompressible ? 0 : 1);
        DNSP_LAUNCH_CHECK();
    }
    const long nfin = 7 * h.sy() + 3 * (2L * (nz + 2) + 2L * nx) * ny;
    k_vel_finish<<<blocks_for(nfin), kStencilThreads, 0, st>>>(U, V, W, h, wall_bc[0], wall_bc[1], wall_bc[2],
                                                              wall_bc[3], wall_bc[4], wall_bc[5]);
    DNSP_LAUNCH_CHECK();
    return DNSP_OK;
}

// calc_bands.py:272-277 and :336-343 (rows 0..ny-1 of dPdy_ddy_lh, as read at
// inc_set_alloc_arrays.f90:51-52).  Host only.
int dnsp_geom_coeffs(int ny, const double* yu, double* inv_dy_p, double* inv_dy_u, double* inv_dy_p_pad,
                     double* dPdy_ddy_lo, double* dPdy_ddy_hi) {
    clear_error();
    DNSP_ARG_CHECK(ny >= 1 && yu, "invalid arguments");
    std::vector<double> yp((size_t)ny), ext((size_t)ny + 2), pad((size_t)ny + 2);
    for (int i = 0; i < ny; ++i) yp[(size_t)i] = (yu[i] + yu[i + 1]) / 2;
    ext[0] = yu[0];
    pad[0] = 2 * yu[0] - yp[0];
    for (int i = 0; i < ny; ++i) ext[(size_t)i + 1] = pad[(size_t)i + 1] = yp[(size_t)i];
    ext[(size_t)ny + 1] = yu[ny];
    pad[(size_t)ny + 1] = 2 * yu[ny] - yp[(size_t)ny - 1];
    for (int i = 0; i < ny; ++i)
        if (inv_dy_p) inv_dy_p[i] = 1. / (yu[i + 1] - yu[i]);
    for (int i = 0; i <= ny; ++i) {
        if (inv_dy_u) inv_dy_u[i] = 1. / (ext[(size_t)i + 1] - ext[(size_t)i]);
        if (inv_dy_p_pad) inv_dy_p_pad[i] = 1. / (pad[(size_t)i + 1] - pad[(size_t)i]);
    }
    for (int i = 0; i < ny; ++i) {
        double lo = 0.0, hi = 0.0;
        if (i != 0) {
            const double dy = yp[(size_t)i] - yp[(size_t)i - 1];
            lo = -1 / dy;
            hi = 1 / dy;
        }
        if (dPdy_ddy_lo) dPdy_ddy_lo[i] = lo;
        if (dPdy_ddy_hi) dPdy_ddy_hi[i] = hi;
    }
    return DNSP_OK;
}

}  // extern "C"
