// The time-step stages before the projection, on one GPU: the momentum
// predictor (U_hat, V_hat, W_hat with second-order Adams-Bashforth), the
// temperature update, the thermophysical property update and the energy
// right-hand side with div_now, in the order of the reference's build_vars
// (src/mod_cfd_subroutines.f90:10-468, update_props :776-857,
// build_rhs_energy :618-745, cfd_periodicity :860-896).
//
// Fields use the reference's halo layout (see projection.hip); the velocity,
// density and div_now fields are the projection step's own (dnsp_step), so a
// predictor enqueue followed by the step's enqueue is one full time step.
// Every kernel keeps the reference's evaluation order with contraction off:
// results are bit-identical to a serial FP64 evaluation of the same
// expressions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "common.h"
#include "stencil_grid.h"

#pragma clang fp contract(off)

using namespace dnsp;

namespace {

struct Grid {
    int nx, nz, ny;
    __host__ __device__ long sz() const { return nx + 2; }
    __host__ __device__ long sy() const { return (long)(nx + 2) * (nz + 2); }
    __host__ __device__ long at(int i, int j, int k) const {
        return (long)(i + 1) + (long)(nx + 2) * ((long)(j + 1) + (long)(nz + 2) * (k + 1));
    }
    __host__ __device__ long total() const { return (long)(nx + 2) * (nz + 2) * (ny + 2); }
    __host__ __device__ long interior() const { return (long)nx * nz * ny; }
};

constexpr int kThreads = 256;

inline int blocks_for(long n) { return (int)((n + kThreads - 1) / kThreads); }


// Coefficients of the y derivatives (device layout of dnsp_predictor_geom)
struct Geom {
    const double* ruw;   // Ruw_ddy a,b,c per row (3 ny)
    const double* ruw2;  // Ruw_d2dy2 (3 ny)
    const double* rv;    // Rv_ddy (3 (ny+1))
    const double* rv2;   // Rv_d2dy2 (3 (ny+1))
    const double* itp;   // Rv_interp_uw bottom, top (2 (ny+1))
    const double* idyu;  // inv_dy_u (ny+1)
};

struct Phys {
    double dt, idx, idz, idx2, idz2, idx__2, idz__2;
    double nu, cond, Sf_x, Sf_z, Sq, Gb_x, Gb_y, Gb_z;
};

constexpr double kThird = 0.3333333333333333;
constexpr double kTwoThirds = 0.6666666666666666;

// --- periodicity and walls ---------------------------------------------------

// cfd_periodicity on one rank for four fields: every padded plane gets its
// periodic x/z halo, then the wall planes (A, C: k = -1 and ny; B: k = -1, 0
// and ny) take the wall values over the whole padded plane.
__global__ void k_periodicity4(double* A, double* B, double* C, double* T, Grid g, double ab, double bb, double cb,
                               double tb, double at, double bt, double ct, double tt) {
    const long pe = g.sy();
    const long per_plane = 2L * (g.nz + 2) + 2L * g.nx;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nwall = 9 * pe;
    if (idx < nwall) {
        const int which = (int)(idx / pe);
        const long o = idx % pe;
        const long top = (long)(g.ny + 1) * pe + o;
        switch (which) {
            case 0: A[o] = ab; break;
            case 1: B[o] = bb; break;
            case 2: B[pe + o] = bb; break;
            case 3: C[o] = cb; break;
            case 4: T[o] = tb; break;
            case 5: A[top] = at; break;
            case 6: B[top] = bt; break;
            case 7: C[top] = ct; break;
            default: T[top] = tt; break;
        }
        return;
    }
    const long r = idx - nwall;
    if (r >= 4 * per_plane * g.ny) return;
    const int f = (int)(r / (per_plane * g.ny));
    const long rr = r % (per_plane * g.ny);
    const int k = (int)(rr / per_plane);
    const long s = rr % per_plane;
    if (f == 1 && k == 0) return;  // B plane 0 is a wall plane
    int i, j;
    if (s < 2L * (g.nz + 2)) {
        i = (s < g.nz + 2) ? -1 : g.nx;
        j = (int)(s % (g.nz + 2)) - 1;
    } else {
        const long t = s - 2L * (g.nz + 2);
        i = (int)(t % g.nx);
        j = (t < g.nx) ? -1 : g.nz;
    }
    const int wi = i < 0 ? g.nx - 1 : (i >= g.nx ? 0 : i);
    const int wj = j < 0 ? g.nz - 1 : (j >= g.nz ? 0 : j);
    double* X = f == 0 ? A : (f == 1 ? B : (f == 2 ? C : T));
    X[g.at(i, j, k)] = X[g.at(wi, wj, k)];
}

// Wall planes of ibm_part_1 (src/ibm_setup/mod_cfd_ibm.f90:26-38): k = -1
// and ny of A, B, C, T over the whole padded plane (B bottom: k = -1 only).
__global__ void k_walls4(double* A, double* B, double* C, double* T, Grid g, double ab, double bb, double cb,
                         double tb, double at, double bt, double ct, double tt) {
    const long pe = g.sy();
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 8 * pe) return;
    const int which = (int)(idx / pe);
    const long o = idx % pe;
    const long top = (long)(g.ny + 1) * pe + o;
    switch (which) {
        case 0: A[o] = ab; break;
        case 1: B[o] = bb; break;
        case 2: C[o] = cb; break;
        case 3: T[o] = tb; break;
        case 4: A[top] = at; break;
        case 5: B[top] = bt; break;
        case 6: C[top] = ct; break;
        default: T[top] = tt; break;
    }
}

// div_now ghost planes by linear extrapolation of the first two interior
// planes (mod_cfd_subroutines.f90:120-144), whole padded planes
__global__ void k_div_walls(double* D, Grid g, double e0, double e1, double em1, double em2) {
    const long pe = g.sy();
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 2 * pe) return;
    if (idx < pe) {
        D[idx] = e0 * D[pe + idx] + e1 * D[2 * pe + idx];
    } else {
        const long o = idx - pe;
        D[(long)(g.ny + 1) * pe + o] = em1 * D[(long)g.ny * pe + o] + em2 * D[(long)(g.ny - 1) * pe + o];
    }
}

// div_now after its interior update: periodic halos of every padded plane,
// then zero ghost planes (mod_cfd_subroutines.f90:719-744)
__global__ void k_div_finish(double* D, Grid g) {
    const long pe = g.sy();
    const long per_plane = 2L * (g.nz + 2) + 2L * g.nx;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < 2 * pe) {
        D[idx < pe ? idx : (long)(g.ny + 1) * pe + (idx - pe)] = 0.0;
        return;
    }
    const long r = idx - 2 * pe;
    if (r >= per_plane * g.ny) return;
    const int k = (int)(r / per_plane);
    const long s = r % per_plane;
    int i, j;
    if (s < 2L * (g.nz + 2)) {
        i = (s < g.nz + 2) ? -1 : g.nx;
        j = (int)(s % (g.nz + 2)) - 1;
    } else {
        const long t = s - 2L * (g.nz + 2);
        i = (int)(t % g.nx);
        j = (t < g.nx) ? -1 : g.nz;
    }
    const int wi = i < 0 ? g.nx - 1 : (i >= g.nx ? 0 : i);
    const int wj = j < 0 ? g.nz - 1 : (j >= g.nz ? 0 : j);
    D[g.at(i, j, k)] = D[g.at(wi, wj, k)];
}

// --- momentum predictor ---------------------------------------------------------

// One interior point: the x and z momentum right-hand sides at the U and W
// points of (i, j, k), the y one at the V point when k >= kmin_Vhat;
// X_hat = X + dt (1.5 R - 0.5 R_old), R_old = R
// (mod_cfd_subroutines.f90:147-354).  u, v, w, m(di, dj, dk) return U, V,
// W, mu at (i + di, j + dj, k + dk); r, dv, b likewise rho, div_now and the
// buoyancy field (read at the point and at (-1, 0, 0), (0, -1, 0), (0, 0, -1)).
template <bool VARPROP, class AU, class AV, class AW, class AM, class AR, class AD, class AB>
__device__ __forceinline__ void momentum_point(
    const AU& u, const AV& v, const AW& w, const AM& m, const AR& r, const AD& dv, const AB& b, int i, int j, int k,
    double* __restrict__ Ru_old, double* __restrict__ Rv_old, double* __restrict__ Rw_old,
    double* __restrict__ U_hat, double* __restrict__ V_hat, double* __restrict__ W_hat, const Grid& g,
    const Geom& G, const Phys& p, int kmin_Vhat) {
    const long c = g.at(i, j, k);
    const long o = (long)i + (long)g.nx * ((long)j + (long)g.nz * k);
    const double idx = p.idx, idz = p.idz, idx2 = p.idx2, idz2 = p.idz2, idx__2 = p.idx__2, idz__2 = p.idz__2;
    {
        const double Ra = G.ruw[3 * k], Rb = G.ruw[3 * k + 1], Rc = G.ruw[3 * k + 2];
        const double R2a = G.ruw2[3 * k], R2b = G.ruw2[3 * k + 1], R2c = G.ruw2[3 * k + 2];
        const double dudx = (u(1, 0, 0) - u(-1, 0, 0)) * idx__2;
        const double dudz = (u(0, 1, 0) - u(0, -1, 0)) * idz__2;
        const double dudy = (u(0, 0, -1) * Ra + u(0, 0, 0) * Rb) + u(0, 0, 1) * Rc;
        const double d2ux = ((u(1, 0, 0) - 2 * u(0, 0, 0)) + u(-1, 0, 0)) * idx2;
        const double d2uz = ((u(0, 1, 0) - 2 * u(0, 0, 0)) + u(0, -1, 0)) * idz2;
        const double d2uy_a = u(0, 0, -1) * R2a, d2uy_b = u(0, 0, 0) * R2b, d2uy_c = u(0, 0, 1) * R2c;
        const double dwdx = (w(1, 0, 0) - w(-1, 0, 0)) * idx__2;
        const double dwdz = (w(0, 1, 0) - w(0, -1, 0)) * idz__2;
        const double dwdy = (w(0, 0, -1) * Ra + w(0, 0, 0) * Rb) + w(0, 0, 1) * Rc;
        const double d2wx = ((w(1, 0, 0) - 2 * w(0, 0, 0)) + w(-1, 0, 0)) * idx2;
        const double d2wz = ((w(0, 1, 0) - 2 * w(0, 0, 0)) + w(0, -1, 0)) * idz2;
        const double d2wy_a = w(0, 0, -1) * R2a, d2wy_b = w(0, 0, 0) * R2b, d2wy_c = w(0, 0, 1) * R2c;
        const double v_u = 0.25 * (((v(0, 0, 0) + v(-1, 0, 0)) + v(0, 0, 1)) + v(-1, 0, 1));
        const double w_u = 0.25 * (((w(0, 0, 0) + w(-1, 0, 0)) + w(0, 1, 0)) + w(-1, 1, 0));
        const double u_w = 0.25 * (((u(0, 0, 0) + u(0, -1, 0)) + u(1, 0, 0)) + u(1, -1, 0));
        const double v_w = 0.25 * (((v(0, 0, 0) + v(0, -1, 0)) + v(0, 0, 1)) + v(0, -1, 1));
        double Ru, Rw;
        if constexpr (!VARPROP) {
            Ru = (((-(u(0, 0, 0) * dudx) - v_u * dudy) - w_u * dudz) +
                  p.nu * ((((d2ux + d2uy_a) + d2uy_b) + d2uy_c) + d2uz)) +
                 p.Sf_x;
            Rw = (((-(u_w * dwdx) - v_w * dwdy) - w(0, 0, 0) * dwdz) +
                  p.nu * ((((d2wx + d2wy_a) + d2wy_b) + d2wy_c) + d2wz)) +
                 p.Sf_z;
        } else {
            {
                const double conv = (u(0, 0, 0) * dudx + w_u * dudz) + v_u * dudy;
                const double D = ((d2ux + d2uz) + ((d2uy_a + d2uy_b) + d2uy_c)) +
                                 kThird * ((dv(0, 0, 0) - dv(-1, 0, 0)) * idx);
                const double E = 2 * dudx - kTwoThirds * (0.5 * (dv(-1, 0, 0) + dv(0, 0, 0)));
                const double Fx = (m(0, 0, 0) - m(-1, 0, 0)) * idx;
                const double Gt = dudy + ((0.5 * (v(0, 0, 0) + v(0, 0, 1))) - (0.5 * (v(-1, 0, 0) + v(-1, 0, 1)))) * idx;
                const double H = ((0.5 * (m(-1, 0, -1) + m(0, 0, -1))) * Ra + (0.5 * (m(-1, 0, 0) + m(0, 0, 0))) * Rb) +
                                 (0.5 * (m(-1, 0, 1) + m(0, 0, 1))) * Rc;
                const double I = dudz + ((0.5 * (w(0, 0, 0) + w(0, 1, 0))) - (0.5 * (w(-1, 0, 0) + w(-1, 1, 0)))) * idx;
                const double J = ((0.5 * (m(-1, 1, 0) + m(0, 1, 0))) - (0.5 * (m(-1, -1, 0) + m(0, -1, 0)))) * idz__2;
                const double visc = (((0.5 * (m(-1, 0, 0) + m(0, 0, 0))) * D + E * Fx) + Gt * H) + I * J;
                Ru = (((-conv + visc) + p.Sf_x) + p.Gb_x * (0.5 * (b(0, 0, 0) + b(-1, 0, 0)))) /
                     (0.5 * (r(-1, 0, 0) + r(0, 0, 0)));
            }
            {
                const double conv = (u_w * dwdx + w(0, 0, 0) * dwdz) + v_w * dwdy;
                const double D = ((d2wx + d2wz) + ((d2wy_a + d2wy_b) + d2wy_c)) +
                                 kThird * ((dv(0, 0, 0) - dv(0, -1, 0)) * idz);
                const double E = dwdx + ((0.5 * (u(0, 0, 0) + u(1, 0, 0))) - (0.5 * (u(0, -1, 0) + u(1, -1, 0)))) * idz;
                const double Fx = ((0.5 * (m(1, -1, 0) + m(1, 0, 0))) - (0.5 * (m(-1, -1, 0) + m(-1, 0, 0)))) * idx__2;
                const double Gt = dwdy + ((0.5 * (v(0, 0, 0) + v(0, 0, 1))) - (0.5 * (v(0, -1, 0) + v(0, -1, 1)))) * idz;
                const double H = ((0.5 * (m(0, -1, -1) + m(0, 0, -1))) * Ra + (0.5 * (m(0, -1, 0) + m(0, 0, 0))) * Rb) +
                                 (0.5 * (m(0, -1, 1) + m(0, 0, 1))) * Rc;
                const double I = 2 * dwdz - kTwoThirds * (0.5 * (dv(0, -1, 0) + dv(0, 0, 0)));
                const double J = (m(0, 0, 0) - m(0, -1, 0)) * idz;
                const double visc = (((0.5 * (m(0, -1, 0) + m(0, 0, 0))) * D + E * Fx) + Gt * H) + I * J;
                Rw = (((-conv + visc) + p.Sf_z) + p.Gb_z * (0.5 * (b(0, 0, 0) + b(0, -1, 0)))) /
                     (0.5 * (r(0, -1, 0) + r(0, 0, 0)));
            }
        }
        U_hat[c] = u(0, 0, 0) + p.dt * (1.5 * Ru - 0.5 * Ru_old[o]);
        W_hat[c] = w(0, 0, 0) + p.dt * (1.5 * Rw - 0.5 * Rw_old[o]);
        Ru_old[o] = Ru;
        Rw_old[o] = Rw;
    }
    if (k < kmin_Vhat) return;
    {
        const double Va = G.rv[3 * k], Vb = G.rv[3 * k + 1], Vc = G.rv[3 * k + 2];
        const double V2a = G.rv2[3 * k], V2b = G.rv2[3 * k + 1], V2c = G.rv2[3 * k + 2];
        const double ib = G.itp[2 * k], it = G.itp[2 * k + 1];
        const double idyu = G.idyu[k];
        const double dvdx = (v(1, 0, 0) - v(-1, 0, 0)) * idx__2;
        const double dvdz = (v(0, 1, 0) - v(0, -1, 0)) * idz__2;
        const double dvdy = (v(0, 0, -1) * Va + v(0, 0, 0) * Vb) + v(0, 0, 1) * Vc;
        const double d2vx = ((v(1, 0, 0) - 2 * v(0, 0, 0)) + v(-1, 0, 0)) * idx2;
        const double d2vz = ((v(0, 1, 0) - 2 * v(0, 0, 0)) + v(0, -1, 0)) * idz2;
        const double d2vy_a = v(0, 0, -1) * V2a, d2vy_b = v(0, 0, 0) * V2b, d2vy_c = v(0, 0, 1) * V2c;
        const double u_v = (0.5 * (u(0, 0, -1) + u(1, 0, -1))) * ib + (0.5 * (u(0, 0, 0) + u(1, 0, 0))) * it;
        const double w_v = (0.5 * (w(0, 0, -1) + w(0, 1, -1))) * ib + (0.5 * (w(0, 0, 0) + w(0, 1, 0))) * it;
        double Rv;
        if constexpr (!VARPROP) {
            Rv = ((-(u_v * dvdx) - v(0, 0, 0) * dvdy) - w_v * dvdz) +
                 p.nu * ((((d2vx + d2vy_a) + d2vy_b) + d2vy_c) + d2vz);
        } else {
            const double conv = (u_v * dvdx + w_v * dvdz) + v(0, 0, 0) * dvdy;
            const double D = ((d2vx + d2vz) + ((d2vy_a + d2vy_b) + d2vy_c)) +
                             kThird * ((dv(0, 0, 0) - dv(0, 0, -1)) * idyu);
            const double E = dvdx + ((0.5 * (u(0, 0, 0) + u(1, 0, 0))) - (0.5 * (u(0, 0, -1) + u(1, 0, -1)))) * idyu;
            const double Fx = ((0.5 * (m(1, 0, -1) + m(1, 0, 0))) - (0.5 * (m(-1, 0, -1) + m(-1, 0, 0)))) * idx__2;
            const double Gt = 2 * dvdy - kTwoThirds * (0.5 * (dv(0, 0, -1) + dv(0, 0, 0)));
            const double H = (m(0, 0, 0) - m(0, 0, -1)) * idyu;
            const double I = dvdz + ((0.5 * (w(0, 0, 0) + w(0, 1, 0))) - (0.5 * (w(0, 0, -1) + w(0, 1, -1)))) * idyu;
            const double J = ((0.5 * (m(0, 1, -1) + m(0, 1, 0))) - (0.5 * (m(0, -1, -1) + m(0, -1, 0)))) * idz__2;
            const double visc = (((0.5 * (m(0, 0, -1) + m(0, 0, 0))) * D + E * Fx) + Gt * H) + I * J;
            Rv = ((-conv + visc) + p.Gb_y * (0.5 * (b(0, 0, 0) + b(0, 0, -1)))) / (0.5 * (r(0, 0, -1) + r(0, 0, 0)));
        }
        V_hat[c] = v(0, 0, 0) + p.dt * (1.5 * Rv - 0.5 * Rv_old[o]);
        Rv_old[o] = Rv;
    }
}

// U, V, W (and mu) are read at up to 16 neighbouring points each, so a
// workgroup stages them in LDS (PlaneRing, stencil_grid.h): it owns a 64 x 4
// (i, j) tile and marches through kRunK planes with planes k-1, k, k+1 of
// each field in a ring of three slots, the next plane's loads issued before
// a step's arithmetic.  rho, div_now and buoyancy (four points each) are
// read directly.  The per-thread version made 12.5 L1 and 2.9 L2 read
// requests per point.
template <bool VARPROP>
__global__ __launch_bounds__(kThreads) void k_momentum(
    const double* __restrict__ U, const double* __restrict__ V, const double* __restrict__ W,
    const double* __restrict__ mu, const double* __restrict__ rho, const double* __restrict__ dvn,
    const double* __restrict__ buoy, double* __restrict__ Ru_old, double* __restrict__ Rv_old,
    double* __restrict__ Rw_old, double* __restrict__ U_hat, double* __restrict__ V_hat, double* __restrict__ W_hat,
    Grid g, BandGrid bg, Geom G, Phys p, int kmin_Vhat) {
    constexpr int NF = VARPROP ? 4 : 3;
    __shared__ double sF[NF][3][kRingP];
    int i0, j0, k0, k1;
    if (!ring_tile(bg, g.ny, i0, j0, k0, k1)) return;
    const int i = i0 + (int)threadIdx.x, j = j0 + (int)threadIdx.y;
    const bool active = i < g.nx && j < g.nz;
    const double* fields[4] = {U, V, W, mu};
    PlaneRing<NF> ring(sF, fields, i0, j0, g.nx, g.nz);
    ring.fetch(k0 - 1);
    ring.put(k0 - 1);
    ring.fetch(k0);
    ring.put(k0);
    ring.fetch(k0 + 1);
    const int e0 = ring_point();
    const long sz = g.sz(), sy = g.sy();
    // rho, div_now and buoy of this thread's column at plane k - 1: read at
    // the march's first plane, then carried from the step before (the plane
    // below is otherwise read again from beyond L2 once the step's other
    // streams have pushed it out)
    double rm = 0.0, dm = 0.0, bm = 0.0;
    if (VARPROP && active) {
        const long cm = g.at(i, j, k0 - 1);
        rm = rho[cm];
        dm = dvn[cm];
        bm = buoy[cm];
    }
    for (int k = k0; k < k1; ++k) {
        ring.put(k + 1);  // the slot of plane k - 2, free since the last barrier
        __syncthreads();
        if (k + 2 <= k1) ring.fetch(k + 2);  // in flight during this step's arithmetic
        if (active) {
            const int sm = (k + 2) % 3, s0 = k % 3, sp = (k + 1) % 3;
            auto slot = [&](int dk) { return dk < 0 ? sm : (dk > 0 ? sp : s0); };
            auto lu = [&](int di, int dj, int dk) { return sF[0][slot(dk)][e0 + di + dj * kRingX]; };
            auto lv = [&](int di, int dj, int dk) { return sF[1][slot(dk)][e0 + di + dj * kRingX]; };
            auto lw = [&](int di, int dj, int dk) { return sF[2][slot(dk)][e0 + di + dj * kRingX]; };
            auto lm = [&](int di, int dj, int dk) { return sF[NF - 1][slot(dk)][e0 + di + dj * kRingX]; };
            const long c = g.at(i, j, k);
            double r0 = 0.0, d0 = 0.0, b0 = 0.0;
            if constexpr (VARPROP) {
                r0 = rho[c];
                d0 = dvn[c];
                b0 = buoy[c];
            }
            // (0, 0, -1) from registers, (0, 0, 0) as loaded, x and z neighbours from memory
            auto at = [&](const double* f, double cen, double below, int di, int dj, int dk) {
                return dk < 0 ? below : ((di | dj) == 0 ? cen : f[c + di + dj * sz + dk * sy]);
            };
            auto lr = [&](int di, int dj, int dk) { return at(rho, r0, rm, di, dj, dk); };
            auto ld = [&](int di, int dj, int dk) { return at(dvn, d0, dm, di, dj, dk); };
            auto lb = [&](int di, int dj, int dk) { return at(buoy, b0, bm, di, dj, dk); };
            momentum_point<VARPROP>(lu, lv, lw, lm, lr, ld, lb, i, j, k, Ru_old, Rv_old, Rw_old, U_hat, V_hat, W_hat,
                                    g, G, p, kmin_Vhat);
            rm = r0;
            dm = d0;
            bm = b0;
        }
        __syncthreads();  // slot (k + 2) % 3 is restaged next
    }
}

// --- temperature ------------------------------------------------------------------

// Rt at every interior point into Rt_now (mod_cfd_subroutines.f90:367-396);
// the update is a separate pass because T is read at neighbours.
template <bool VARPROP>
__global__ __launch_bounds__(kThreads) void k_temp_rhs(const double* __restrict__ T, const double* __restrict__ U,
                                                       const double* __restrict__ V, const double* __restrict__ W,
                                                       const double* __restrict__ rho, const double* __restrict__ cp,
                                                       const double* __restrict__ RHS, double* __restrict__ Rt_now,
                                                       Grid g, BandGrid bg, Geom G, Phys p) {
    int i, j, k;
    if (!band_point(bg, i, j, k)) return;
    const long c = g.at(i, j, k);
    const long sz = g.sz(), sy = g.sy();
    const long o = (long)i + (long)g.nx * ((long)j + (long)g.nz * k);
    auto t = [&](int di, int dj, int dk) { return T[c + di + dj * sz + dk * sy]; };
    const double Ra = G.ruw[3 * k], Rb = G.ruw[3 * k + 1], Rc = G.ruw[3 * k + 2];
    const double dtdx = (t(1, 0, 0) - t(-1, 0, 0)) * p.idx__2;
    const double dtdz = (t(0, 1, 0) - t(0, -1, 0)) * p.idz__2;
    const double dtdy = (t(0, 0, -1) * Ra + t(0, 0, 0) * Rb) + t(0, 0, 1) * Rc;
    const double u_t = 0.5 * (U[c] + U[c + 1]);
    const double v_t = 0.5 * (V[c] + V[c + sy]);
    const double w_t = 0.5 * (W[c] + W[c + sz]);
    double Rt;
    if constexpr (!VARPROP) {
        const double R2a = G.ruw2[3 * k], R2b = G.ruw2[3 * k + 1], R2c = G.ruw2[3 * k + 2];
        const double d2tx = ((t(1, 0, 0) - 2 * t(0, 0, 0)) + t(-1, 0, 0)) * p.idx2;
        const double d2tz = ((t(0, 1, 0) - 2 * t(0, 0, 0)) + t(0, -1, 0)) * p.idz2;
        const double lap = (((d2tx + t(0, 0, -1) * R2a) + t(0, 0, 0) * R2b) + t(0, 0, 1) * R2c) + d2tz;
        Rt = (((-(u_t * dtdx) - v_t * dtdy) - w_t * dtdz) + p.cond * lap) + p.Sq;
    } else {
        Rt = -((u_t * dtdx + w_t * dtdz) + v_t * dtdy) + RHS[o] / (rho[c] * cp[c]);
    }
    Rt_now[o] = Rt;
}

__global__ void k_temp_update(double* __restrict__ T, const double* __restrict__ Rt_now, double* __restrict__ Rt_old,
                              Grid g, BandGrid bg, double dt) {
    int i, j, k;
    if (!band_point(bg, i, j, k)) return;
    const long c = g.at(i, j, k);
    const long o = (long)i + (long)g.nx * ((long)j + (long)g.nz * k);
    const double rt = Rt_now[o];
    T[c] = T[c] + dt * (1.5 * rt - 0.5 * Rt_old[o]);
    Rt_old[o] = rt;
}

// --- thermophysical properties -------------------------------------------------------

// A property law: 0 constant c; 1 power c T^p (p = -1, -1/2, 1/2, 1 as
// division, square root or product); 2 piecewise cubic over nseg segments
// of 6 values (hi, x0, c0, c1, c2, c3) from `off`: a point takes the first
// segment with T <= hi (the last one above every hi) and
// c0 + c1 d + c2 d^2 + c3 d^3 with d = T - x0, in the examples' term order.
struct Law {
    int kind, nseg, off;
    double c, p;
    // cubic: kBuckets equal buckets over [blo, blo + kBuckets / binv) map T
    // to a starting segment (the law's bucket table), refined exactly by
    // comparisons with the bounds
    double blo, binv;
};

constexpr int kBuckets = 256;

enum { Q_ENTH = 0, Q_RHO, Q_MU, Q_COND, Q_CP, Q_DRHO, Q_N };

struct PropSet {
    Law law[Q_N];
    int divjac_formula;  // 0: constant divjac_c; 1: -1/(rho^2 cp) drho_dT
    double divjac_c;
    int buoy_linear;  // 0: zero; 1: b1 (rho - b0)
    double b1, b0;
};

// tab: the law's kBuckets starting segments.  The segment is the first one
// with T <= hi (the last if none): the bucket's entry, moved down while the
// previous bound still holds T and up while T exceeds the current bound,
// which is that segment whatever bucket rounding picked.
__device__ __forceinline__ double cubic_eval(const Law& L, const double* seg, const int* tab, double T) {
    const double* sg = seg + 6 * L.off;
    const double x = (T - L.blo) * L.binv;
    const int b = !(x >= 0.0) ? 0 : (x >= (double)kBuckets ? kBuckets - 1 : (int)x);
    int lo = tab[b];
    while (lo > 0 && T <= sg[6 * (lo - 1)]) --lo;
    while (lo < L.nseg - 1 && T > sg[6 * lo]) ++lo;
    const double* s = sg + 6 * lo;
    const double d = T - s[1];
    return ((s[2] + s[3] * d) + s[4] * (d * d)) + s[5] * ((d * d) * d);
}

// ALL_CUBIC: every law is known to be a piecewise cubic (no constant or
// power-law code in the kernel).
template <bool ALL_CUBIC>
__device__ __forceinline__ double law_eval(const Law& L, const double* seg, const int* tab, double T) {
    if constexpr (!ALL_CUBIC) {
        if (L.kind == 0) return L.c;
        if (L.kind == 1) {
            if (L.p == -1.0) return L.c / T;
            if (L.p == -0.5) return L.c / sqrt(T);
            if (L.p == 0.5) return L.c * sqrt(T);
            if (L.p == 1.0) return L.c * T;
            return L.c * pow(T, L.p);
        }
    }
    return cubic_eval(L, seg, tab, T);
}

// All six laws piecewise cubic: the union of their segment bounds splits T
// into merged intervals on each of which every law keeps one segment, so one
// interval search (instead of one per law) selects all six segments.
// knots[0..m) increase; interval i holds knots[i-1] < T <= knots[i] (0: T <=
// knots[0]; m: T above every knot); mseg[6 i + q] is law q's segment there (an
// index into the segment table), mtab[kMergedBuckets] the first interval of
// each of the equal buckets over [blo, blo + kMergedBuckets / binv).  The
// segment is the per-law search's (the first with T <= its bound, the last
// above every bound): the law's bounds are among the knots, so T <= hi exactly
// when the interval's upper knot is <= hi.
constexpr int kMergedBuckets = 1024;
struct MergedLaws {
    const double* knots;
    const int* mseg;
    const int* mtab;
    int m;
    double blo, binv;
};

__device__ __forceinline__ int merged_interval(const double* knots, const int* mtab, int m, double blo,
                                               double binv, double T) {
    const double x = (T - blo) * binv;
    const int b = !(x >= 0.0) ? 0 : (x >= (double)kMergedBuckets ? kMergedBuckets - 1 : (int)x);
    int i = mtab[b];
    while (i > 0 && T <= knots[i - 1]) --i;
    while (i < m && T > knots[i]) ++i;
    return i;
}

__device__ __forceinline__ double seg_eval(const double* seg, int sgi, double T) {
    const double* s = seg + 6 * sgi;
    const double d = T - s[1];
    return ((s[2] + s[3] * d) + s[4] * (d * d)) + s[5] * ((d * d) * d);
}

// T_ext and the seven properties at every padded point (update_props): in
// rough-wall columns (bot/top given) T_ext takes the wall values for
// k <= bot(i, j) and k >= top(i, j), halo points those of their periodic
// image.  Each workgroup stages the cubic laws' segments (nseg of them) and
// bucket tables in LDS once and strides over the points, so the interval
// searches read LDS.
constexpr int kMaxSegments = 1024;  // 48 KB of LDS, plus 6 KB of bucket tables (merged tables: up to 64 KB in all)

// IDX32: the padded field has fewer than 2^31 points (32-bit index split).
template <bool ALL_CUBIC, bool IDX32>
__global__ __launch_bounds__(kThreads) void k_props(const double* __restrict__ T, double* __restrict__ T_ext,
                                                    double* __restrict__ enth, double* __restrict__ rho,
                                                    double* __restrict__ mu, double* __restrict__ cond,
                                                    double* __restrict__ cp, double* __restrict__ divjac,
                                                    double* __restrict__ buoy, const int* __restrict__ bot,
                                                    const int* __restrict__ top, double t_bot, double t_top,
                                                    const double* __restrict__ seg_g, int nseg,
                                                    const int* __restrict__ tab_g, PropSet ps, Grid g,
                                                    MergedLaws ml) {
    // LDS: the segments; then per-law bucket tables (general laws) or the
    // merged knots, segment map and buckets (all laws cubic)
    extern __shared__ double seg[];
    double* knots = seg + 6 * nseg;
    int* tab = reinterpret_cast<int*>(ALL_CUBIC ? knots + ml.m : knots);
    int* mseg = tab;
    int* mtab = mseg + 6 * (ml.m + 1);
    for (int e = (int)threadIdx.x; e < 6 * nseg; e += kThreads) seg[e] = seg_g[e];
    if constexpr (ALL_CUBIC) {
        for (int e = (int)threadIdx.x; e < ml.m; e += kThreads) knots[e] = ml.knots[e];
        for (int e = (int)threadIdx.x; e < 6 * (ml.m + 1); e += kThreads) mseg[e] = ml.mseg[e];
        for (int e = (int)threadIdx.x; e < kMergedBuckets; e += kThreads) mtab[e] = ml.mtab[e];
    } else {
        for (int e = (int)threadIdx.x; e < Q_N * kBuckets; e += kThreads) tab[e] = tab_g[e];
    }
    __syncthreads();
    // T_ext: in rough-wall columns the wall values
    auto wall_te = [&](long idx, double te) -> double {
        if (bot == nullptr) return te;
        int i, j, k;
        if constexpr (IDX32) {
            const unsigned pe = (unsigned)g.sy(), px = (unsigned)(g.nx + 2), x = (unsigned)idx;
            const unsigned kp = x / pe, r = x - kp * pe, jp = r / px;
            k = (int)kp - 1;
            j = (int)jp - 1;
            i = (int)(r - jp * px) - 1;
        } else {
            const long pe = g.sy();
            k = (int)(idx / pe) - 1;
            const long r = idx - (long)(k + 1) * pe;
            j = (int)(r / (g.nx + 2)) - 1;
            i = (int)(r - (long)(j + 1) * (g.nx + 2)) - 1;
        }
        const int wi = i < 0 ? g.nx - 1 : (i >= g.nx ? 0 : i);
        const int wj = j < 0 ? g.nz - 1 : (j >= g.nz ? 0 : j);
        if (k >= 0 && k < g.ny) {
            const long col = (long)wi + (long)g.nx * wj;
            if (k <= bot[col]) te = t_bot;
            if (k >= top[col]) te = t_top;
        }
        return te;
    };
    // the seven properties at one temperature: o = enth, rho, mu, cond, cp, divjac, buoy
    auto eval = [&](double te, double* o) {
        double h, r, m, k, c;
        const int* ms = nullptr;
        if constexpr (ALL_CUBIC) {
            ms = mseg + 6 * merged_interval(knots, mtab, ml.m, ml.blo, ml.binv, te);
            h = seg_eval(seg, ms[Q_ENTH], te);
            r = seg_eval(seg, ms[Q_RHO], te);
            m = seg_eval(seg, ms[Q_MU], te);
            k = seg_eval(seg, ms[Q_COND], te);
            c = seg_eval(seg, ms[Q_CP], te);
        } else {
            h = law_eval<false>(ps.law[Q_ENTH], seg, tab + Q_ENTH * kBuckets, te);
            r = law_eval<false>(ps.law[Q_RHO], seg, tab + Q_RHO * kBuckets, te);
            m = law_eval<false>(ps.law[Q_MU], seg, tab + Q_MU * kBuckets, te);
            k = law_eval<false>(ps.law[Q_COND], seg, tab + Q_COND * kBuckets, te);
            c = law_eval<false>(ps.law[Q_CP], seg, tab + Q_CP * kBuckets, te);
        }
        o[0] = h;
        o[1] = r;
        o[2] = m;
        o[3] = k;
        o[4] = c;
        if (ps.divjac_formula) {
            double dr;
            if constexpr (ALL_CUBIC) dr = seg_eval(seg, ms[Q_DRHO], te);
            else dr = law_eval<false>(ps.law[Q_DRHO], seg, tab + Q_DRHO * kBuckets, te);
            o[5] = (-1.0 / ((r * r) * c)) * dr;
        } else {
            o[5] = ps.divjac_c;
        }
        o[6] = ps.buoy_linear ? ps.b1 * (r - ps.b0) : 0.0;
    };
    double* const out[8] = {T_ext, enth, rho, mu, cond, cp, divjac, buoy};
    // two consecutive points per thread: 16-byte loads and stores (every
    // field is its own allocation, so pair p sits at a 16-byte boundary)
    const long n = g.total();
    const long np = n >> 1;
    const long stride = (long)gridDim.x * kThreads;
    const double2* T2 = reinterpret_cast<const double2*>(T);
    long p = (long)blockIdx.x * kThreads + threadIdx.x;
    double2 tn = p < np ? T2[p] : make_double2(0.0, 0.0);
    double2 tnn = p + stride < np ? T2[p + stride] : make_double2(0.0, 0.0);
    for (; p < np; p += stride) {
        // the next two pairs' loads are in flight during this pair's evaluation
        const double2 t = tn;
        tn = tnn;
        if (p + 2 * stride < np) tnn = T2[p + 2 * stride];
        const double te0 = wall_te(2 * p, t.x), te1 = wall_te(2 * p + 1, t.y);
        double o0[7], o1[7];
        eval(te0, o0);
        eval(te1, o1);
        reinterpret_cast<double2*>(out[0])[p] = make_double2(te0, te1);
#pragma unroll
        for (int q = 0; q < 7; ++q) reinterpret_cast<double2*>(out[q + 1])[p] = make_double2(o0[q], o1[q]);
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const double te = wall_te(n - 1, T[n - 1]);
        double o[7];
        eval(te, o);
        out[0][n - 1] = te;
        for (int q = 0; q < 7; ++q) out[q + 1][n - 1] = o[q];
    }
}

// RHS_energy and div_now = divjac RHS_energy at interior points
// (mod_cfd_subroutines.f90:651-694), zero in rough-wall columns (:697-716).
// T and cond are read at seven points each: a workgroup stages them in LDS
// (PlaneRing, stencil_grid.h) and marches through kRunK planes of its 64 x 4
// tile, as k_momentum (at 384x256x192: 0.240 -> 0.230 ms; the per-thread
// version served its 14 neighbour reads per point from L1 and L2), and
// loads divjac one plane ahead (-> 0.212 ms).  Same expressions on the same
// values.
__global__ __launch_bounds__(kThreads) void k_rhs_energy(const double* __restrict__ T,
                                                         const double* __restrict__ cond,
                                                         const double* __restrict__ divjac,
                                                         double* __restrict__ RHS, double* __restrict__ dvn,
                                                         const int* __restrict__ bot, const int* __restrict__ top,
                                                         Grid g, BandGrid bg, Geom G, Phys p) {
    __shared__ double sF[2][3][kRingP];
    int i0, j0, k0, k1;
    if (!ring_tile(bg, g.ny, i0, j0, k0, k1)) return;
    const int i = i0 + (int)threadIdx.x, j = j0 + (int)threadIdx.y;
    const bool active = i < g.nx && j < g.nz;
    const double* fields[2] = {T, cond};
    PlaneRing<2> ring(sF, fields, i0, j0, g.nx, g.nz);
    ring.fetch(k0 - 1);
    ring.put(k0 - 1);
    ring.fetch(k0);
    ring.put(k0);
    ring.fetch(k0 + 1);
    const int e0 = ring_point();
    // the column's rough-wall slab (plane-independent) and divjac one plane
    // ahead, so that no step waits for a load of its own
    int cb = -1, ct = g.ny;
    if (bot != nullptr && active) {
        const long col = (long)i + (long)g.nx * j;
        cb = bot[col];
        ct = top[col];
    }
    double djn = active ? divjac[g.at(i, j, k0)] : 0.0;
    for (int kk = k0; kk < k1; ++kk) {
        ring.put(kk + 1);  // the slot of plane kk - 2, free since the last barrier
        __syncthreads();
        if (kk + 2 <= k1) ring.fetch(kk + 2);  // in flight during this step's arithmetic
        const double djac = djn;
        if (active && kk + 1 < k1) djn = divjac[g.at(i, j, kk + 1)];
        if (active) {
            const int sm = (kk + 2) % 3, s0 = kk % 3, sp = (kk + 1) % 3;
            auto slot = [&](int dk) { return dk < 0 ? sm : (dk > 0 ? sp : s0); };
            auto t = [&](int di, int dj, int dk) { return sF[0][slot(dk)][e0 + di + dj * kRingX]; };
            auto q = [&](int di, int dj, int dk) { return sF[1][slot(dk)][e0 + di + dj * kRingX]; };
            const long c = g.at(i, j, kk);
            const long o = (long)i + (long)g.nx * ((long)j + (long)g.nz * kk);
            const double Ra = G.ruw[3 * kk], Rb = G.ruw[3 * kk + 1], Rc = G.ruw[3 * kk + 2];
            const double R2a = G.ruw2[3 * kk], R2b = G.ruw2[3 * kk + 1], R2c = G.ruw2[3 * kk + 2];
            const double d2tx = ((t(1, 0, 0) - 2 * t(0, 0, 0)) + t(-1, 0, 0)) * p.idx2;
            const double d2tz = ((t(0, 1, 0) - 2 * t(0, 0, 0)) + t(0, -1, 0)) * p.idz2;
            const double d2ty = (t(0, 0, -1) * R2a + t(0, 0, 0) * R2b) + t(0, 0, 1) * R2c;
            const double grad =
                (((q(1, 0, 0) - q(-1, 0, 0)) * p.idx__2) * ((t(1, 0, 0) - t(-1, 0, 0)) * p.idx__2) +
                 ((q(0, 1, 0) - q(0, -1, 0)) * p.idz__2) * ((t(0, 1, 0) - t(0, -1, 0)) * p.idz__2)) +
                ((q(0, 0, -1) * Ra + q(0, 0, 0) * Rb) + q(0, 0, 1) * Rc) *
                    ((t(0, 0, -1) * Ra + t(0, 0, 0) * Rb) + t(0, 0, 1) * Rc);
            const double rhs = (q(0, 0, 0) * ((d2tx + d2tz) + d2ty) + grad) + p.Sq;
            RHS[o] = rhs;
            double dn = djac * rhs;
            if (kk <= cb || kk >= ct) dn = 0.0;
            dvn[c] = dn;
        }
        __syncthreads();  // slot (kk + 2) % 3 is restaged next
    }
}

}  // namespace

// --- the object ------------------------------------------------------------------------

struct dnsp_predictor {
    dnsp_step* step = nullptr;  // borrowed
    Grid g{};
    Phys p{};
    int incompressible = 1;
    int kmin_Vhat = 1;
    int quady = 0;  // use_rough_surf_quady ordering of the rough-wall branch
    double walls[8] = {};  // U, V, W, T bottom; U, V, W, T top
    double extrap[4] = {};
    double* d_geom = nullptr;
    Geom G{};
    double* f[DNSP_PRED_NFIELDS] = {};
    double* d_rt_now = nullptr;
    PropSet ps{};
    std::vector<double> h_seg;  // host copy of every law's segments
    double* d_seg = nullptr;
    size_t d_seg_n = 0;
    std::vector<int> h_tab = std::vector<int>(Q_N * kBuckets, 0);  // bucket tables, kBuckets per quantity
    int* d_tab = nullptr;
    // merged tables when every law is cubic (ml.m >= 0): one device array of
    // the m knots (doubles), the 6 (m + 1) segment map and the kMergedBuckets
    // buckets (ints)
    MergedLaws ml{nullptr, nullptr, nullptr, -1, 0.0, 0.0};
    void* d_merged = nullptr;
    size_t d_merged_bytes = 0;
    int* d_bot[4] = {};  // rough-wall column slabs of U, V, W, T (nx*nz each), or NULL
    int* d_top[4] = {};
    dnsp_ibm* ibm[5] = {};  // U, V, W, T, div_now forcing (borrowed)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

bool halo_field(int which) {
    return !(which == DNSP_PRED_RU_OLD || which == DNSP_PRED_RV_OLD || which == DNSP_PRED_RW_OLD ||
             which == DNSP_PRED_RT_OLD || which == DNSP_PRED_RHS_ENERGY);
}

void release(dnsp_predictor* pr) {
    if (!pr) return;
    if (pr->exec) (void)hipGraphExecDestroy(pr->exec);
    if (pr->graph) (void)hipGraphDestroy(pr->graph);
    for (double*& q : pr->f) (void)hipFree(q);
    (void)hipFree(pr->d_rt_now);
    (void)hipFree(pr->d_geom);
    (void)hipFree(pr->d_seg);
    (void)hipFree(pr->d_tab);
    (void)hipFree(pr->d_merged);
    for (int c = 0; c < 4; ++c) {
        (void)hipFree(pr->d_bot[c]);
        (void)hipFree(pr->d_top[c]);
    }
    delete pr;
}

// d/dy and d2/dy2 at xc = 0 of the quadratic through (xs, fs), (0, fc),
// (xn, fn) (calc_bands.py:84-104: fd_coeffs_ref at xc = 0, its terms in
// order, the exact-zero xc terms dropped); bc 0 normal, 1 mirrored ghost at
// the s side, 2 at the n side.
void fd_coeffs(double xs, double xn, int bc, double* ddx, double* d2) {
    if (bc == 0) {
        const double D = xn * xn * xs - xn * (xs * xs);
        ddx[0] = (xn * xn) / D;
        ddx[1] = (-(xn * xn) + xs * xs) / D;
        ddx[2] = (-(xs * xs)) / D;
        d2[0] = 2 * (-xn) / D;
        d2[1] = 2 * (xn - xs) / D;
        d2[2] = 2 * (xs) / D;
    } else if (bc == 1) {
        const double D = xn * xn * xs - 2 * xn * (xs * xs);
        ddx[0] = (xn * xn) / D;
        ddx[1] = (-(xn * xn) + 2 * (xs * xs)) / D;
        ddx[2] = (-(2 * (xs * xs))) / D;
        d2[0] = 2 * (-xn) / D;
        d2[1] = 2 * (xn - xs) / D;
        d2[2] = 2 * (xs) / D;
    } else {
        const double D = -(2 * (xn * xn) * xs) + xn * (xs * xs);
        ddx[0] = (-(2 * (xn * xn))) / D;
        ddx[1] = (2 * (xn * xn) - xs * xs) / D;
        ddx[2] = (xs * xs) / D;
        d2[0] = 2 * (xn) / D;
        d2[1] = 2 * (-xn + xs) / D;
        d2[2] = 2 * (-xs) / D;
    }
}

}  // namespace

extern "C" {

size_t dnsp_predictor_geom_size(int ny) { return ny >= 2 ? (size_t)6 * ny + (size_t)9 * (ny + 1) + 4 : 0; }

int dnsp_predictor_geom(int ny, const double* yu, double* out) {
    clear_error();
    DNSP_ARG_CHECK(ny >= 2 && yu && out, "invalid arguments");
    std::vector<double> yp((size_t)ny), ext((size_t)ny + 2);
    for (int i = 0; i < ny; ++i) yp[(size_t)i] = (yu[i] + yu[i + 1]) / 2;
    ext[0] = yu[0];
    for (int i = 0; i < ny; ++i) ext[(size_t)i + 1] = yp[(size_t)i];
    ext[(size_t)ny + 1] = yu[ny];
    for (int i = 1; i <= ny; ++i)
        DNSP_ARG_CHECK(yu[i] > yu[i - 1], "faces must increase (at %d)", i);
    double* ruw = out;
    double* ruw2 = ruw + 3 * ny;
    double* rv = ruw2 + 3 * ny;
    double* rv2 = rv + 3 * (ny + 1);
    double* itp = rv2 + 3 * (ny + 1);
    double* idyu = itp + 2 * (ny + 1);
    double* ex = idyu + (ny + 1);
    for (int i = 0; i < ny; ++i) {
        const int bc = i == 0 ? 1 : (i == ny - 1 ? 2 : 0);
        fd_coeffs(ext[(size_t)i] - ext[(size_t)i + 1], ext[(size_t)i + 2] - ext[(size_t)i + 1], bc, ruw + 3 * i,
                  ruw2 + 3 * i);
    }
    for (int i = 0; i <= ny; ++i) {
        if (i == 0 || i == ny) {
            for (int c = 0; c < 3; ++c) rv[3 * i + c] = rv2[3 * i + c] = 0.0;
            itp[2 * i] = itp[2 * i + 1] = 0.0;
        } else {
            fd_coeffs(yu[i - 1] - yu[i], yu[i + 1] - yu[i], 0, rv + 3 * i, rv2 + 3 * i);
            const double dy_top = yp[(size_t)i] - yu[i];
            const double dy_bottom = yu[i] - yp[(size_t)i - 1];
            itp[2 * i] = 1 - (dy_bottom / (dy_top + dy_bottom));
            itp[2 * i + 1] = 1 - (dy_top / (dy_top + dy_bottom));
        }
        idyu[i] = 1. / (ext[(size_t)i + 1] - ext[(size_t)i]);
    }
    // linear extrapolation weights of the ghost planes (calc_bands.py:212-215)
    auto w = [](double xc, double xa, double xb, double* o) {
        o[0] = (xc - xa) / (xb - xa) * (-1) + 1;
        o[1] = (xc - xa) / (xb - xa);
    };
    w(ext[0], ext[1], ext[2], ex);
    w(ext[(size_t)ny + 1], ext[(size_t)ny], ext[(size_t)ny - 1], ex + 2);
    return DNSP_OK;
}

int dnsp_predictor_create(dnsp_step* s, const double* yu, const dnsp_pred_params* prm, dnsp_predictor** out) {
    clear_error();
    DNSP_ARG_CHECK(out, "out is NULL");
    *out = nullptr;
    DNSP_ARG_CHECK(s && yu && prm, "NULL argument");
    int nx, nz, ny, inc;
    double idx, idz, dt;
    int rc = dnsp_step_info(s, &nx, &nz, &ny, &idx, &idz, &dt, &inc);
    if (rc) return rc;
    DNSP_ARG_CHECK(ny >= 3, "the predictor needs ny >= 3 (wall extrapolation from two planes)");
    DNSP_ARG_CHECK(prm->kmin_Vhat >= 0 && prm->kmin_Vhat <= ny, "kmin_Vhat=%d", prm->kmin_Vhat);
    dnsp_predictor* pr = new (std::nothrow) dnsp_predictor();
    if (!pr) {
        set_error("host allocation failed");
        return DNSP_ERR_NOMEM;
    }
    pr->step = s;
    pr->g = Grid{nx, nz, ny};
    pr->incompressible = inc;
    pr->kmin_Vhat = prm->kmin_Vhat;
    pr->p = Phys{dt, idx, idz, idx * idx, idz * idz, 0.5 * idx, 0.5 * idz, prm->nu, prm->cond, prm->Sf_x,
                 prm->Sf_z, prm->Sq, prm->Gb_x, prm->Gb_y, prm->Gb_z};
    std::memcpy(pr->walls, prm->walls, sizeof(pr->walls));
    // default property laws: constants (rho = cp = 1, mu = nu, cond = cond), div_jacobian 0, no buoyancy
    const double defaults[Q_N] = {0.0, 1.0, prm->nu, prm->cond, 1.0, 0.0};
    for (int q = 0; q < Q_N; ++q) pr->ps.law[q] = Law{0, 0, 0, defaults[q], 0.0};
    do {
        std::vector<double> gh(dnsp_predictor_geom_size(ny));
        rc = dnsp_predictor_geom(ny, yu, gh.data());
        if (rc) break;
        std::memcpy(pr->extrap, gh.data() + gh.size() - 4, sizeof(pr->extrap));
        const size_t hc = (size_t)pr->g.total(), ic = (size_t)pr->g.interior();
        bool ok = hipMalloc(&pr->d_geom, sizeof(double) * gh.size()) == hipSuccess &&
                  hipMemcpy(pr->d_geom, gh.data(), sizeof(double) * gh.size(), hipMemcpyHostToDevice) == hipSuccess &&
                  hipMalloc(&pr->d_rt_now, sizeof(double) * ic) == hipSuccess &&
                  hipMemset(pr->d_rt_now, 0, sizeof(double) * ic) == hipSuccess &&
                  hipMalloc(&pr->d_tab, sizeof(int) * pr->h_tab.size()) == hipSuccess &&
                  hipMemcpy(pr->d_tab, pr->h_tab.data(), sizeof(int) * pr->h_tab.size(), hipMemcpyHostToDevice) ==
                      hipSuccess;
        for (int i = 0; i < DNSP_PRED_NFIELDS && ok; ++i) {
            // the property fields only for the variable-property form
            if (inc && i >= DNSP_PRED_MU) continue;
            const size_t n = halo_field(i) ? hc : ic;
            ok = hipMalloc(&pr->f[i], sizeof(double) * n) == hipSuccess &&
                 hipMemset(pr->f[i], 0, sizeof(double) * n) == hipSuccess;
        }
        if (!ok) {
            set_error("device allocation of the predictor fields failed");
            rc = DNSP_ERR_NOMEM;
            break;
        }
        const double* d = pr->d_geom;
        pr->G = Geom{d, d + 3 * ny, d + 6 * ny, d + 6 * ny + 3 * (ny + 1), d + 6 * ny + 6 * (ny + 1),
                     d + 6 * ny + 8 * (ny + 1)};
    } while (0);
    if (rc) {
        release(pr);
        return rc;
    }
    *out = pr;
    return DNSP_OK;
}

int dnsp_predictor_field(dnsp_predictor* pr, int which, double** ptr, size_t* count) {
    clear_error();
    DNSP_ARG_CHECK(pr && ptr, "NULL argument");
    DNSP_ARG_CHECK(which >= 0 && which < DNSP_PRED_NFIELDS, "field %d out of range", which);
    DNSP_ARG_CHECK(pr->f[which], "field %d is not used by the incompressible form", which);
    *ptr = pr->f[which];
    if (count) *count = (size_t)(halo_field(which) ? pr->g.total() : pr->g.interior());
    return DNSP_OK;
}

int dnsp_predictor_write(dnsp_predictor* pr, int which, const double* src, void* stream) {
    double* dst = nullptr;
    size_t n = 0;
    const int rc = dnsp_predictor_field(pr, which, &dst, &n);
    if (rc) return rc;
    DNSP_ARG_CHECK(src, "NULL source");
    DNSP_HIP_TRY(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, as_stream(stream)));
    return DNSP_OK;
}

int dnsp_predictor_read(dnsp_predictor* pr, int which, double* dst, void* stream) {
    double* src = nullptr;
    size_t n = 0;
    const int rc = dnsp_predictor_field(pr, which, &src, &n);
    if (rc) return rc;
    DNSP_ARG_CHECK(dst, "NULL destination");
    DNSP_HIP_TRY(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, as_stream(stream)));
    return DNSP_OK;
}

namespace {

// The merged-interval tables of the property kernel (MergedLaws) from the
// current laws; ml.m = -1 unless every law is cubic.
int build_merged(dnsp_predictor* pr) {
    pr->ml.m = -1;
    for (int q = 0; q < Q_N; ++q)
        if (pr->ps.law[q].kind != DNSP_LAW_CUBIC) return DNSP_OK;
    std::vector<double> K;
    for (int q = 0; q < Q_N; ++q) {
        const Law& L = pr->ps.law[q];
        for (int t = 0; t + 1 < L.nseg; ++t) K.push_back(pr->h_seg[6 * (size_t)(L.off + t)]);
    }
    std::sort(K.begin(), K.end());
    K.erase(std::unique(K.begin(), K.end()), K.end());
    const int m = (int)K.size();
    // the kernel stages the segments and these tables in LDS: above 64 KB it
    // keeps the per-law search (at most 54 KB)
    const size_t lds = sizeof(double) * pr->h_seg.size() + sizeof(double) * m +
                       sizeof(int) * ((size_t)6 * (m + 1) + kMergedBuckets);
    if (lds > 64 * 1024) return DNSP_OK;
    std::vector<int> ms((size_t)6 * (m + 1)), mt(kMergedBuckets);
    for (int q = 0; q < Q_N; ++q) {
        const Law& L = pr->ps.law[q];
        int t = 0;  // first segment whose bound is >= the interval's upper knot
        for (int i = 0; i <= m; ++i) {
            if (i == m) t = L.nseg - 1;
            else
                while (t < L.nseg - 1 && pr->h_seg[6 * (size_t)(L.off + t)] < K[(size_t)i]) ++t;
            ms[(size_t)6 * i + q] = L.off + t;
        }
    }
    double blo = 0.0, binv = 0.0;
    if (m >= 2) {
        blo = K[0];
        binv = kMergedBuckets / (K[(size_t)m - 1] - K[0]);
    }
    for (int b = 0, i = 0; b < kMergedBuckets; ++b) {
        const double start = binv > 0.0 ? blo + b / binv : -1e300;
        while (i < m && K[(size_t)i] < start) ++i;
        mt[(size_t)b] = i;
    }
    const size_t bytes = sizeof(double) * m + sizeof(int) * (ms.size() + mt.size());
    if (bytes > pr->d_merged_bytes) {
        (void)hipFree(pr->d_merged);
        pr->d_merged = nullptr;
        pr->d_merged_bytes = 0;
        DNSP_HIP_TRY(hipMalloc(&pr->d_merged, bytes));
        pr->d_merged_bytes = bytes;
    }
    double* dk = static_cast<double*>(pr->d_merged);
    int* dms = reinterpret_cast<int*>(dk + m);
    int* dmt = dms + ms.size();
    if (m > 0) DNSP_HIP_TRY(hipMemcpy(dk, K.data(), sizeof(double) * m, hipMemcpyHostToDevice));
    DNSP_HIP_TRY(hipMemcpy(dms, ms.data(), sizeof(int) * ms.size(), hipMemcpyHostToDevice));
    DNSP_HIP_TRY(hipMemcpy(dmt, mt.data(), sizeof(int) * mt.size(), hipMemcpyHostToDevice));
    pr->ml = MergedLaws{dk, dms, dmt, m, blo, binv};
    return DNSP_OK;
}

}  // namespace

int dnsp_predictor_set_law(dnsp_predictor* pr, int quantity, int kind, const double* data, int n) {
    clear_error();
    DNSP_ARG_CHECK(pr && data, "NULL argument");
    DNSP_ARG_CHECK(quantity >= 0 && quantity < Q_N, "quantity %d out of range", quantity);
    DNSP_ARG_CHECK(!pr->incompressible, "property laws belong to the variable-property form");
    Law L{kind, 0, 0, 0.0, 0.0, 0.0, 0.0};
    if (kind == DNSP_LAW_CONST) {
        L.c = data[0];
    } else if (kind == DNSP_LAW_POWER) {
        L.c = data[0];
        L.p = data[1];
    } else if (kind == DNSP_LAW_CUBIC) {
        DNSP_ARG_CHECK(n >= 1 && n <= 1 << 20, "cubic law needs 1..2^20 segments (got %d)", n);
        // bounds of segments 0..n-2 increase (the last segment has none)
        for (int s = 0; s + 2 < n; ++s)
            DNSP_ARG_CHECK(data[6 * s] < data[6 * (s + 1)], "segment bounds must increase (at %d)", s);
        // replace this quantity's segments in the host table, then upload the whole table
        std::vector<double> keep;
        int off = 0;
        for (int q = 0; q < Q_N; ++q) {
            Law& o = pr->ps.law[q];
            if (q == quantity || o.kind != DNSP_LAW_CUBIC) continue;
            keep.insert(keep.end(), pr->h_seg.begin() + 6 * o.off, pr->h_seg.begin() + 6 * (o.off + o.nseg));
            o.off = off;
            off += o.nseg;
        }
        DNSP_ARG_CHECK(off + n <= kMaxSegments, "at most %d cubic segments over all laws (got %d)", kMaxSegments,
                       off + n);
        L.nseg = n;
        L.off = off;
        // bucket table over [first bound, last bound]: entry b = first segment
        // whose bound is >= the bucket's start (the device refines it exactly)
        int* tb = pr->h_tab.data() + quantity * kBuckets;
        if (n >= 3 && data[6 * (n - 2)] > data[0]) {
            L.blo = data[0];
            L.binv = kBuckets / (data[6 * (n - 2)] - data[0]);
            for (int b = 0, sg = 0; b < kBuckets; ++b) {
                const double start = L.blo + b / L.binv;
                while (sg < n - 1 && data[6 * sg] < start) ++sg;
                tb[b] = sg;
            }
        } else {
            L.blo = 0.0;
            L.binv = 0.0;
            for (int b = 0; b < kBuckets; ++b) tb[b] = 0;
        }
        DNSP_HIP_TRY(hipMemcpy(pr->d_tab, pr->h_tab.data(), sizeof(int) * pr->h_tab.size(), hipMemcpyHostToDevice));
        keep.insert(keep.end(), data, data + 6 * (size_t)n);
        pr->h_seg.swap(keep);
        if (pr->h_seg.size() > pr->d_seg_n) {
            (void)hipFree(pr->d_seg);
            pr->d_seg = nullptr;
            pr->d_seg_n = 0;
            DNSP_HIP_TRY(hipMalloc(&pr->d_seg, sizeof(double) * pr->h_seg.size()));
            pr->d_seg_n = pr->h_seg.size();
        }
        DNSP_HIP_TRY(hipMemcpy(pr->d_seg, pr->h_seg.data(), sizeof(double) * pr->h_seg.size(),
                               hipMemcpyHostToDevice));
    } else {
        DNSP_ARG_CHECK(false, "unknown law kind %d", kind);
    }
    pr->ps.law[quantity] = L;
    return build_merged(pr);
}

int dnsp_predictor_set_derived(dnsp_predictor* pr, int divjac_formula, double divjac_const, int buoy_linear,
                               double b1, double b0) {
    clear_error();
    DNSP_ARG_CHECK(pr, "NULL predictor");
    pr->ps.divjac_formula = divjac_formula ? 1 : 0;
    pr->ps.divjac_c = divjac_const;
    pr->ps.buoy_linear = buoy_linear ? 1 : 0;
    pr->ps.b1 = b1;
    pr->ps.b0 = b0;
    return DNSP_OK;
}

int dnsp_predictor_set_component_columns(dnsp_predictor* pr, int component, const int* bot, const int* top) {
    clear_error();
    DNSP_ARG_CHECK(pr && component >= 0 && component < 4, "NULL predictor or component %d", component);
    return set_slab_pair(&pr->d_bot[component], &pr->d_top[component], bot, top, (size_t)pr->g.nx * pr->g.nz,
                         nullptr);
}

int dnsp_predictor_set_columns(dnsp_predictor* pr, const int* bot, const int* top) {
    clear_error();
    DNSP_ARG_CHECK(pr, "NULL predictor");
    for (int c = 0; c < 4; ++c) {
        const int rc = dnsp_predictor_set_component_columns(pr, c, bot, top);
        if (rc) return rc;
    }
    return DNSP_OK;
}

int dnsp_predictor_set_quady(dnsp_predictor* pr, int quady) {
    clear_error();
    DNSP_ARG_CHECK(pr, "NULL predictor");
    pr->quady = quady ? 1 : 0;
    return DNSP_OK;
}

int dnsp_predictor_set_ibm(dnsp_predictor* pr, int component, dnsp_ibm* h) {
    clear_error();
    DNSP_ARG_CHECK(pr, "NULL predictor");
    DNSP_ARG_CHECK(component >= 0 && component < 5, "component %d out of range (U, V, W, T, div)", component);
    pr->ibm[component] = h;
    return DNSP_OK;
}

int dnsp_predictor_enqueue(dnsp_predictor* pr, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(pr, "NULL predictor");
    hipStream_t st = as_stream(stream);
    const Grid g = pr->g;
    double *U, *V, *W, *Uh, *Vh, *Wh, *rho, *dvn;
    int rc = dnsp_step_field(pr->step, DNSP_STEP_U, &U);
    if (!rc) rc = dnsp_step_field(pr->step, DNSP_STEP_V, &V);
    if (!rc) rc = dnsp_step_field(pr->step, DNSP_STEP_W, &W);
    if (!rc) rc = dnsp_step_field(pr->step, DNSP_STEP_U_HAT, &Uh);
    if (!rc) rc = dnsp_step_field(pr->step, DNSP_STEP_V_HAT, &Vh);
    if (!rc) rc = dnsp_step_field(pr->step, DNSP_STEP_W_HAT, &Wh);
    if (!rc) rc = dnsp_step_field(pr->step, DNSP_STEP_RHO, &rho);
    if (!rc) rc = dnsp_step_field(pr->step, DNSP_STEP_DIV_NOW, &dvn);
    if (rc) return rc;
    double* T = pr->f[DNSP_PRED_T];
    const double* wl = pr->walls;
    const long per_plane = 2L * (g.nz + 2) + 2L * g.nx;
    const long nper = 9 * g.sy() + 4 * per_plane * g.ny;
    const BandGrid bg = band_grid(g.nx, g.nz, g.ny);
    const dim3 grid3((unsigned)bg.blocks());
    // tiles x runs of planes, for the kernels that stage planes in LDS
    const BandGrid br = ring_grid(g.nx, g.nz, g.ny);
    const dim3 gridr((unsigned)br.blocks());
    const dim3 blk3 = band_block();
    const bool vp = !pr->incompressible;
    const bool columns = pr->d_bot[0] || pr->d_bot[1] || pr->d_bot[2] || pr->d_bot[3];
    const bool rough = columns || pr->ibm[0] || pr->ibm[1] || pr->ibm[2] || pr->ibm[3] || pr->ibm[4];
    void* vs = stream;
    double* F4[4] = {U, V, W, T};
    double* H4[4] = {Uh, Vh, Wh, T};
    const double wb[4] = {wl[0], wl[1], wl[2], wl[3]}, wt[4] = {wl[4], wl[5], wl[6], wl[7]};
    // ibm_part_2 (mod_cfd_ibm.f90:47-141) on four fields, each from its own slabs
    auto fill_columns = [&](double* const* A4) {
        int r = DNSP_OK;
        for (int c = 0; c < 4 && !r; ++c)
            if (pr->d_bot[c])
                r = dnsp_ibm_part2(A4[c], pr->d_bot[c], pr->d_top[c], g.nx, g.nz, g.ny, c == 1 ? g.ny : g.ny - 1,
                                   wb[c], wt[c], vs);
        return r;
    };
    // ibm_part_1 (mod_cfd_ibm.f90:10-45): forcing of four fields, their wall
    // planes, and (force_div) the forcing of div_now
    auto part_1 = [&](double* const* A4, bool force_div) {
        int r = DNSP_OK;
        for (int c = 0; c < 4 && !r; ++c)
            if (pr->ibm[c]) r = dnsp_ibm_exec(pr->ibm[c], A4[c], vs);
        if (r) return r;
        k_walls4<<<blocks_for(8 * g.sy()), kThreads, 0, st>>>(A4[0], A4[1], A4[2], A4[3], g, wl[0], wl[1], wl[2],
                                                              wl[3], wl[4], wl[5], wl[6], wl[7]);
        if (force_div && pr->ibm[4]) r = dnsp_ibm_exec(pr->ibm[4], dvn, vs);
        return r;
    };
    if (!rough) {
        // 1. cfd_periodicity(U, V, W, T)
        k_periodicity4<<<blocks_for(nper), kThreads, 0, st>>>(U, V, W, T, g, wl[0], wl[1], wl[2], wl[3], wl[4],
                                                              wl[5], wl[6], wl[7]);
    } else {
        // 1. rough walls (mod_cfd_subroutines.f90:98-108): with quady the columns
        // of U, V, W, T and cfd_periodicity first; then ibm_part_1
        if (pr->quady) {
            rc = fill_columns(F4);
            if (rc) return rc;
            k_periodicity4<<<blocks_for(nper), kThreads, 0, st>>>(U, V, W, T, g, wl[0], wl[1], wl[2], wl[3], wl[4],
                                                                  wl[5], wl[6], wl[7]);
        }
        rc = part_1(F4, vp);
        if (rc) return rc;
    }
    // 2. div_now ghost planes
    if (vp)
        k_div_walls<<<blocks_for(2 * g.sy()), kThreads, 0, st>>>(dvn, g, pr->extrap[0], pr->extrap[1],
                                                                 pr->extrap[2], pr->extrap[3]);
    // 3. momentum predictor
    double* mu = pr->f[DNSP_PRED_MU];
    double* buoy = pr->f[DNSP_PRED_BUOY];
    if (vp)
        k_momentum<true><<<gridr, blk3, 0, st>>>(U, V, W, mu, rho, dvn, buoy, pr->f[DNSP_PRED_RU_OLD],
                                                     pr->f[DNSP_PRED_RV_OLD], pr->f[DNSP_PRED_RW_OLD], Uh, Vh, Wh, g,
                                                     br, pr->G, pr->p, pr->kmin_Vhat);
    else
        k_momentum<false><<<gridr, blk3, 0, st>>>(U, V, W, nullptr, nullptr, nullptr, nullptr,
                                                      pr->f[DNSP_PRED_RU_OLD], pr->f[DNSP_PRED_RV_OLD],
                                                      pr->f[DNSP_PRED_RW_OLD], Uh, Vh, Wh, g, br, pr->G, pr->p,
                                                      pr->kmin_Vhat);
    // 4. temperature
    if (vp)
        k_temp_rhs<true><<<grid3, blk3, 0, st>>>(T, U, V, W, rho, pr->f[DNSP_PRED_CP],
                                                     pr->f[DNSP_PRED_RHS_ENERGY], pr->d_rt_now, g, bg, pr->G, pr->p);
    else
        k_temp_rhs<false><<<grid3, blk3, 0, st>>>(T, U, V, W, nullptr, nullptr, nullptr, pr->d_rt_now, g, bg, pr->G,
                                                      pr->p);
    k_temp_update<<<grid3, blk3, 0, st>>>(T, pr->d_rt_now, pr->f[DNSP_PRED_RT_OLD], g, bg, pr->p.dt);
    // 5. rough walls: columns of the hats and T, each from its own slabs
    if (columns) {
        rc = fill_columns(H4);
        if (rc) return rc;
    }
    // periodicity of the hats and T with their wall planes (with rough walls
    // the reference sets only the halos the right-hand side reads, V_hat's
    // walls and T's; the extra halo and wall values are read by nothing)
    k_periodicity4<<<blocks_for(nper), kThreads, 0, st>>>(Uh, Vh, Wh, T, g, wl[0], wl[1], wl[2], wl[3], wl[4],
                                                          wl[5], wl[6], wl[7]);
    if (rough && pr->quady) {
        // quady (mod_cfd_subroutines.f90:417-427): ibm_part_1 on the hats and T;
        // its last argument (.false.) leaves the div_now forcing on; then T's halos
        rc = part_1(H4, true);
        if (!rc) rc = dnsp_halo_periodic_xz(T, g.nx, g.nz, g.ny, vs);
        if (rc) return rc;
    }
    // 6. properties at the new temperature, energy right-hand side, div_now
    if (vp) {
        const int nseg = (int)(pr->h_seg.size() / 6);
        const int nbp = std::min(blocks_for(g.total()), 2048);
        const bool all_cubic = pr->ml.m >= 0;
        const bool idx32 = g.total() < (1L << 31);
        auto kp = all_cubic ? (idx32 ? k_props<true, true> : k_props<true, false>)
                            : (idx32 ? k_props<false, true> : k_props<false, false>);
        const size_t lds = sizeof(double) * 6 * nseg +
                           (all_cubic ? sizeof(double) * pr->ml.m + sizeof(int) * (6 * (pr->ml.m + 1) + kMergedBuckets)
                                      : sizeof(int) * Q_N * kBuckets);
        kp<<<nbp, kThreads, lds, st>>>(T, pr->f[DNSP_PRED_T_EXT], pr->f[DNSP_PRED_ENTH], rho, mu,
                                       pr->f[DNSP_PRED_COND], pr->f[DNSP_PRED_CP], pr->f[DNSP_PRED_DIVJAC], buoy,
                                       pr->d_bot[3], pr->d_top[3], wl[3], wl[7], pr->d_seg, nseg, pr->d_tab, pr->ps,
                                       g, pr->ml);
        if (pr->ibm[3]) {
            // build_rhs_energy (:640-645): T forcing, then T's wall planes
            rc = dnsp_ibm_exec(pr->ibm[3], T, vs);
            if (!rc) rc = dnsp_set_val_slice(T, g.nx, g.nz, g.ny, 0, 0, wl[3], vs);
            if (!rc) rc = dnsp_set_val_slice(T, g.nx, g.nz, g.ny, -1, -1, wl[7], vs);
            if (rc) return rc;
        }
        k_rhs_energy<<<gridr, blk3, 0, st>>>(T, pr->f[DNSP_PRED_COND], pr->f[DNSP_PRED_DIVJAC],
                                                 pr->f[DNSP_PRED_RHS_ENERGY], dvn, pr->d_bot[3], pr->d_top[3], g, br,
                                                 pr->G, pr->p);
        k_div_finish<<<blocks_for(2 * g.sy() + per_plane * g.ny), kThreads, 0, st>>>(dvn, g);
    }
    DNSP_LAUNCH_CHECK();
    return DNSP_OK;
}

int dnsp_predictor_capture(dnsp_predictor* pr, int with_step, int refinements, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(pr, "NULL predictor");
    hipStream_t st = as_stream(stream);
    if (pr->exec) {
        (void)hipGraphExecDestroy(pr->exec);
        pr->exec = nullptr;
    }
    if (pr->graph) {
        (void)hipGraphDestroy(pr->graph);
        pr->graph = nullptr;
    }
    DNSP_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = dnsp_predictor_enqueue(pr, stream);
    if (!rc && with_step) rc = dnsp_step_enqueue(pr->step, refinements, stream);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) {
        set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
        return DNSP_ERR_HIP;
    }
    pr->graph = graph;
    DNSP_HIP_TRY(hipGraphInstantiate(&pr->exec, pr->graph, nullptr, nullptr, 0));
    return DNSP_OK;
}

int dnsp_predictor_launch(dnsp_predictor* pr, void* stream) {
    clear_error();
    DNSP_ARG_CHECK(pr && pr->exec, "no captured graph (call dnsp_predictor_capture first)");
    DNSP_HIP_TRY(hipGraphLaunch(pr->exec, as_stream(stream)));
    return DNSP_OK;
}

int dnsp_predictor_free(dnsp_predictor* pr) {
    clear_error();
    release(pr);
    return DNSP_OK;
}

}  // extern "C"
