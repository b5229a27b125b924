// Compile-time-sized kernels of the Poisson solve: x and z transforms for
// lengths 32..768 and the partitioned wall-normal tridiagonal solve.  Layout and halfcomplex conventions are those
// of the runtime-sized kernels in poisson.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

#include "fft_ct.h"

namespace dnsp {

// FFT kernel variants: (complex elements per workgroup tile, threads), one
// table per direction.  The tuner picks one per direction and handle.  x
// variant 0 is one complex line (two real lines) per workgroup at nx = 768:
// the smallest tiles keep the most workgroups in flight, which the
// contiguous x lines turn into bandwidth (x forward 0.374 -> 0.336 ms at
// 768x384x384 against two complex lines).
constexpr int kCtVariants = 4;
__host__ __device__ constexpr int ct_xvar_tile(int v) { return v == 0 ? 768 : v == 1 ? 3072 : v == 2 ? 1536 : 6144; }
__host__ __device__ constexpr int ct_xvar_threads(int v) { return v == 0 ? 256 : v == 1 ? 512 : v == 2 ? 256 : 512; }
__host__ __device__ constexpr int ct_zvar_tile(int v) { return v == 0 ? 3072 : v == 1 ? 3072 : v == 2 ? 1536 : 6144; }
__host__ __device__ constexpr int ct_zvar_threads(int v) { return v == 0 ? 256 : v == 1 ? 512 : v == 2 ? 256 : 512; }

template <int N, int TILE>
constexpr int ct_lines() {  // complex lines (= 2 real lines) per x workgroup
    return (TILE / N) > 0 ? (TILE / N) : 1;
}
template <int N, int TILE>
constexpr int ct_cols() {  // complex columns per z workgroup, at most 16
    return (TILE / N) > 16 ? 16 : ((TILE / N) > 0 ? (TILE / N) : 1);
}

// ---------------------------------------------------------------------------
// x forward: 2G real lines (one contiguous block of memory) per workgroup
// ---------------------------------------------------------------------------
// Bodies take the tile index and the LDS tile (G * ct_pitch(N) complex
// values); the __global__ wrappers pick the tile with xcd_tile.  All threads
// of the workgroup call a body.
// XB = true: S receives the x-blocked layout of the blocked solve,
// [k][a/16][j][a%16] (line k*nz + j, halfcomplex index a; nx a multiple of
// 16), instead of the lines in place.
// Plane k and index j of line l0 + r of a tile whose first line l0 is
// k0 * nz + j0 (0 <= j0 < nz), with mz = line_div_magic(nz): r + j0 divided
// by nz as a multiply-high by ceil(2^32 / nz), exact while (j0 + r) * nz <
// 2^32 (nz is an FFT length, far below 2^16).  Branch-free: control flow
// here would split an unrolled element loop into basic blocks, which keeps
// the scheduler from issuing the next elements' loads early; and a division
// per element is a dozen instructions.
__device__ __forceinline__ unsigned long long line_div_magic(int nz) {
    return 0xFFFFFFFFull / (unsigned)nz + 1ull;
}
__device__ __forceinline__ void tile_line_kj(int k0, int j0, int r, int nz, unsigned long long mz, int& k, int& j) {
    const unsigned jr = (unsigned)(j0 + r);
    const unsigned dk = (unsigned)(((unsigned long long)jr * mz) >> 32);
    k = k0 + (int)dk;
    j = (int)(jr - dk * (unsigned)nz);
}

template <int N, int G, int NT, bool XB = false>
__device__ __forceinline__ void xfwd_body(const double* P, double* S, long nlines, const double2* __restrict__ tw,
                                          long tile, double2* sm, int nz = 0) {
    constexpr int LP = ct_pitch(N);
    constexpr int NL = (G * N + NT - 1) / NT;  // double2 loads per thread
    double* smd = reinterpret_cast<double*>(sm);
    const long l0 = tile * 2 * G;
    const double2* src = reinterpret_cast<const double2*>(P + l0 * N);
    int kb = 0, jb = 0;
    unsigned long long mz = 0;
    if constexpr (XB) {
        kb = (int)(l0 / nz);
        jb = (int)(l0 - (long)kb * nz);
        mz = line_div_magic(nz);
    }
    double2 v[NL];
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const int r = (2 * e) / N;
        v[q] = ((G * N % NT == 0 || e < G * N) && l0 + r < nlines) ? src[e] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        if (G * N % NT == 0 || e < G * N) {
            const int r = (2 * e) / N;
            const int t = 2 * e - r * N;
            const int g = r >> 1, c = r & 1;
            smd[(g * LP + ct_pad(t)) * 2 + c] = v[q].x;
            smd[(g * LP + ct_pad(t + 1)) * 2 + c] = v[q].y;
        }
    }
    __syncthreads();
    fft_ct<N, G, NT, false>(sm, tw);
    double2* dst = reinterpret_cast<double2*>(S + l0 * N);
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const int r = (2 * e) / N;
        if ((G * N % NT == 0 || e < G * N) && l0 + r < nlines) {
            const int t0 = 2 * e - r * N;
            const int g = r >> 1;
            double out[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int t = t0 + h;
                const bool lo = 2 * t <= N;
                const int u = lo ? t : N - t;
                const double2 C = sm[g * LP + ct_pad(u)];
                const double2 Cm = sm[g * LP + ct_pad(u == 0 ? 0 : N - u)];
                if ((r & 1) == 0) out[h] = lo ? 0.5 * (C.x + Cm.x) : 0.5 * (C.y - Cm.y);
                else              out[h] = lo ? 0.5 * (C.y + Cm.y) : 0.5 * (Cm.x - C.x);
            }
            if constexpr (XB) {
                int k, j;
                tile_line_kj(kb, jb, r, nz, mz, k, j);
                *reinterpret_cast<double2*>(S + (((long)k * (N / 16) + (t0 >> 4)) * nz + j) * 16 + (t0 & 15)) =
                    make_double2(out[0], out[1]);
            } else {
                dst[e] = make_double2(out[0], out[1]);
            }
        }
    }
}

// Every transform wrapper stages the twiddles in LDS (ct_stage_twiddles); the
// body's barrier before its transform publishes them (at 768x384x384 the
// x-blocked forward kernel went from 0.370 to 0.336 ms, G = 2).
template <int N, int G, int NT>
__global__ __launch_bounds__(NT) void kt_xfwd(const double* P, double* S, long nlines,
                                              const double2* __restrict__ tw) {
    __shared__ double2 sm[G * ct_pitch(N)];
    __shared__ double2 stw[N];
    ct_stage_twiddles<N, NT>(stw, tw);
    xfwd_body<N, G, NT>(P, S, nlines, stw, xcd_tile(blockIdx.x, gridDim.x), sm);
}

// ---------------------------------------------------------------------------
// x backward (+ scale): halfcomplex pairs of lines -> real lines
// ---------------------------------------------------------------------------
template <int N, int G, int NT>
__device__ __forceinline__ void xbwd_body(const double* S, double* P, long nlines, const double2* __restrict__ tw,
                                          double scale, long tile, double2* sm) {
    constexpr int LP = ct_pitch(N);
    constexpr int H = N / 2 + 1;
    constexpr int NI = (G * H + NT - 1) / NT;
    constexpr int NL = (G * N + NT - 1) / NT;
    const long l0 = tile * 2 * G;
    const double* src = S + l0 * N;
    double xr[NI], xi[NI], yr[NI], yi[NI];
    if ((((size_t)src) & 15) == 0) {
        // 16-byte aligned tile: the 2G lines are staged in LDS (inside sm,
        // which they fit) with contiguous 16-byte loads from clamped lines,
        // and the t / N - t gather reads LDS instead of 8-byte global gathers
        constexpr int NV = G * N;  // double2 of the tile
        constexpr int NQ = (NV + NT - 1) / NT;
        double2 v[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            int e = (int)threadIdx.x + q * NT;
            e = (NV % NT == 0 || e < NV) ? e : 0;
            const int line = (2 * e) / N;
            const int off = 2 * e - line * N;
            const int ln = l0 + line < nlines ? line : 0;
            v[q] = *reinterpret_cast<const double2*>(src + (long)ln * N + off);
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = (int)threadIdx.x + q * NT;
            if (NV % NT == 0 || e < NV) sm[e] = v[q];
        }
        __syncthreads();
        const double* sd = reinterpret_cast<const double*>(sm);
#pragma unroll
        for (int q = 0; q < NI; ++q) {
            const int e = (int)threadIdx.x + q * NT;
            const bool ok = e < G * H;
            const int ee = ok ? e : 0;
            const int g = ee / H;
            const int t = ee - g * H;
            const bool single = (t == 0) || (2 * t == N);
            const bool va = ok && (l0 + 2 * g < nlines);
            const bool vb = ok && (l0 + 2 * g + 1 < nlines);
            const int tm = single ? t : N - t;
            const double* la = sd + (2 * g) * N;
            const double* lb = la + N;
            const double a0 = la[t], a1 = la[tm], b0 = lb[t], b1 = lb[tm];
            xr[q] = va ? a0 : 0.0;
            xi[q] = (va && !single) ? a1 : 0.0;
            yr[q] = vb ? b0 : 0.0;
            yi[q] = (vb && !single) ? b1 : 0.0;
        }
        __syncthreads();  // staged lines read; sm becomes the FFT lines
    } else {
#pragma unroll
        for (int q = 0; q < NI; ++q) {
            const int e = (int)threadIdx.x + q * NT;
            const int g = e / H;
            const int t = e - g * H;
            const bool ok = e < G * H;
            const bool single = (t == 0) || (2 * t == N);
            const bool va = ok && (l0 + 2 * g < nlines);
            const bool vb = ok && (l0 + 2 * g + 1 < nlines);
            const double* la = src + (2 * g) * N;
            const double* lb = la + N;
            xr[q] = va ? la[t] : 0.0;
            xi[q] = (va && !single) ? la[N - t] : 0.0;
            yr[q] = vb ? lb[t] : 0.0;
            yi[q] = (vb && !single) ? lb[N - t] : 0.0;
        }
    }
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        if (e < G * H) {
            const int g = e / H;
            const int t = e - g * H;
            const bool single = (t == 0) || (2 * t == N);
            sm[g * LP + ct_pad(t)] = make_double2(xr[q] - yi[q], xi[q] + yr[q]);
            if (!single) sm[g * LP + ct_pad(N - t)] = make_double2(xr[q] + yi[q], yr[q] - xi[q]);
        }
    }
    __syncthreads();
    fft_ct<N, G, NT, true>(sm, tw);
    const double* smd = reinterpret_cast<const double*>(sm);
    double2* dst = reinterpret_cast<double2*>(P + l0 * N);
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const int r = (2 * e) / N;
        if ((G * N % NT == 0 || e < G * N) && l0 + r < nlines) {
            const int t = 2 * e - r * N;
            const int g = r >> 1, c = r & 1;
            dst[e] = make_double2(smd[(g * LP + ct_pad(t)) * 2 + c] * scale,
                                  smd[(g * LP + ct_pad(t + 1)) * 2 + c] * scale);
        }
    }
}

template <int N, int G, int NT>
__global__ __launch_bounds__(NT) void kt_xbwd(const double* S, double* P, long nlines,
                                              const double2* __restrict__ tw, double scale) {
    __shared__ double2 sm[G * ct_pitch(N)];
    __shared__ double2 stw[N];
    ct_stage_twiddles<N, NT>(stw, tw);
    xbwd_body<N, G, NT>(S, P, nlines, stw, scale, xcd_tile(blockIdx.x, gridDim.x), sm);
}

// ---------------------------------------------------------------------------
// z: TA complex columns of one plane per workgroup.  Complex column c pairs
// the adjacent halfcomplex x rows 2c (real part) and 2c+1 (imaginary part),
// so a tile is 2*TA contiguous doubles of every z row; for odd nx the last
// row is paired with zero.  (Any two real rows can share one complex
// transform; the runtime-sized kernels pair rows c and nx-c instead.)
// ---------------------------------------------------------------------------
// T is the storage (and arithmetic) type, double or float.  When lsum is not
// NULL, the workgroup holding x row 0 of plane k also writes
// msum[k] = sum_j lsum[k*N + j] (FP64 line sums -> plane sum, fixed order).
template <int N, int TA, int NT, typename T>
__device__ __forceinline__ void zfwd_body(T* S, int nx, int ntc, const typename Cx<T>::V* __restrict__ tw,
                                          const double* __restrict__ lsum, double* __restrict__ msum, int tile,
                                          typename Cx<T>::V* sm) {
    using V = typename Cx<T>::V;
    constexpr int LP = ct_pitch(N);
    constexpr int W = 2 * TA;  // real rows per tile
    constexpr int TOT = W * N;
    constexpr int NL = (TOT + NT - 1) / NT;
    T* smd = reinterpret_cast<T*>(sm);
    const int k = tile / ntc;
    const int r0 = (tile - k * ntc) * W;  // first real row of the tile
    T* Sk = S + (size_t)k * nx * N;
    // even nx (and an aligned array): the two rows of a complex column are
    // adjacent in memory, so each lane moves one (re, im) pair (the tile's
    // start r0 is even)
    constexpr int TOT2 = TA * N;
    constexpr int NL2 = (TOT2 + NT - 1) / NT;
    const bool pairs = (nx & 1) == 0 && ((size_t)S & (sizeof(V) - 1)) == 0;
    if (pairs) {
        V v[NL2];
#pragma unroll
        for (int q = 0; q < NL2; ++q) {
            const int e = (int)threadIdx.x + q * NT;
            const int tt = e % TA;
            const int j = e / TA;
            const bool ok = (TOT2 % NT == 0 || e < TOT2) && r0 + 2 * tt < nx;
            v[q] = Cx<T>::mk(T(0), T(0));
            if (ok) v[q] = *reinterpret_cast<const V*>(Sk + (size_t)j * nx + r0 + 2 * tt);
        }
#pragma unroll
        for (int q = 0; q < NL2; ++q) {
            const int e = (int)threadIdx.x + q * NT;
            if (TOT2 % NT == 0 || e < TOT2) sm[(e % TA) * LP + ct_pad(e / TA)] = v[q];
        }
    } else {
        T v[NL];
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int e = (int)threadIdx.x + q * NT;
            const int r = e % W;
            const int j = e / W;
            const bool ok = (TOT % NT == 0 || e < TOT) && r0 + r < nx;
            v[q] = ok ? Sk[(size_t)j * nx + r0 + r] : T(0);
        }
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int e = (int)threadIdx.x + q * NT;
            if (TOT % NT == 0 || e < TOT) {
                const int r = e % W;
                const int j = e / W;
                smd[((r >> 1) * LP + ct_pad(j)) * 2 + (r & 1)] = v[q];
            }
        }
    }
    __syncthreads();
    fft_ct<N, TA, NT, false>(sm, tw);
    const T h = T(0.5);
    if (pairs) {
#pragma unroll
        for (int q = 0; q < NL2; ++q) {
            const int e = (int)threadIdx.x + q * NT;
            const int tt = e % TA;
            const int b = e / TA;
            if ((TOT2 % NT == 0 || e < TOT2) && r0 + 2 * tt < nx) {
                const bool lo = 2 * b <= N;
                const int qq = lo ? b : N - b;
                const V Z = sm[tt * LP + ct_pad(qq)];
                const V Zm = sm[tt * LP + ct_pad(qq == 0 ? 0 : N - qq)];
                const T ve = lo ? h * (Z.x + Zm.x) : h * (Z.y - Zm.y);
                const T vo = lo ? h * (Z.y + Zm.y) : h * (Zm.x - Z.x);
                *reinterpret_cast<V*>(Sk + (size_t)b * nx + r0 + 2 * tt) = Cx<T>::mk(ve, vo);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int e = (int)threadIdx.x + q * NT;
            const int r = e % W;
            const int b = e / W;
            if ((TOT % NT == 0 || e < TOT) && r0 + r < nx) {
                const int tt = r >> 1;
                const bool lo = 2 * b <= N;
                const int qq = lo ? b : N - b;
                const V Z = sm[tt * LP + ct_pad(qq)];
                const V Zm = sm[tt * LP + ct_pad(qq == 0 ? 0 : N - qq)];
                T val;
                if ((r & 1) == 0) val = lo ? h * (Z.x + Zm.x) : h * (Z.y - Zm.y);
                else              val = lo ? h * (Z.y + Zm.y) : h * (Zm.x - Z.x);
                Sk[(size_t)b * nx + r0 + r] = val;
            }
        }
    }
    if (lsum != nullptr && r0 == 0 && threadIdx.x < 64) {
        double acc = 0.0;
        for (int j = (int)threadIdx.x; j < N; j += 64) acc += lsum[(size_t)k * N + j];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (threadIdx.x == 0) msum[k] = acc;
    }
}

template <int N, int TA, int NT, typename T = double>
__global__ __launch_bounds__(NT) void kt_zfwd(T* S, int nx, int ncol, int ntc,
                                              const typename Cx<T>::V* __restrict__ tw,
                                              const double* __restrict__ lsum = nullptr,
                                              double* __restrict__ msum = nullptr) {
    __shared__ typename Cx<T>::V sm[TA * ct_pitch(N)];
    __shared__ typename Cx<T>::V stw[N];
    ct_stage_twiddles<N, NT>(stw, tw);
    zfwd_body<N, TA, NT, T>(S, nx, ntc, stw, lsum, msum, xcd_tile(blockIdx.x, gridDim.x), sm);
}

template <int N, int TA, int NT, typename T>
__device__ __forceinline__ void zbwd_body(T* S, int nx, int ntc, const typename Cx<T>::V* __restrict__ tw,
                                          int tile, typename Cx<T>::V* sm) {
    constexpr int LP = ct_pitch(N);
    constexpr int W = 2 * TA;
    constexpr int H = N / 2 + 1;
    constexpr int NI = (H * TA + NT - 1) / NT;
    constexpr int TOT = W * N;
    constexpr int NL = (TOT + NT - 1) / NT;
    const int k = tile / ntc;
    const int r0 = (tile - k * ntc) * W;
    T* Sk = S + (size_t)k * nx * N;
    // Even nx and an aligned array: one (u, v) pair per row (8 or 16 bytes)
    // from clamped addresses (unconditional loads issue together), masked by
    // exact 0/1 factors afterwards; else element-wise loads.  (Paired loads
    // under lane branches had measured slower than element-wise ones.)
    T ur[NI], ui[NI], vr[NI], vi[NI];
    const bool pairs = (nx & 1) == 0 && ((size_t)S & (2 * sizeof(T) - 1)) == 0;
    if (pairs) {
        using V = typename Cx<T>::V;
        V pu[NI], pi[NI];
#pragma unroll
        for (int q = 0; q < NI; ++q) {
            const int e = (int)threadIdx.x + q * NT;
            const bool ok = e < H * TA;
            const int ee = ok ? e : 0;
            const int tt = ee % TA;
            const int qq = ee / TA;
            const int ru = r0 + 2 * tt < nx ? r0 + 2 * tt : r0;
            const int qm = (qq == 0 || 2 * qq == N) ? qq : N - qq;
            pu[q] = *reinterpret_cast<const V*>(Sk + (size_t)qq * nx + ru);
            pi[q] = *reinterpret_cast<const V*>(Sk + (size_t)qm * nx + ru);
        }
#pragma unroll
        for (int q = 0; q < NI; ++q) {
            const int e = (int)threadIdx.x + q * NT;
            const int tt = e % TA;
            const int qq = e / TA;
            const bool single = (qq == 0) || (2 * qq == N);
            const T m = (e < H * TA && r0 + 2 * tt < nx) ? T(1) : T(0);
            const T mi = single ? T(0) : m;
            ur[q] = pu[q].x * m;
            vr[q] = pu[q].y * m;
            ui[q] = pi[q].x * mi;
            vi[q] = pi[q].y * mi;
        }
    } else {
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const int tt = e % TA;
        const int qq = e / TA;
        const bool ok = e < H * TA;
        const int ru = r0 + 2 * tt, rv = ru + 1;
        const bool single = (qq == 0) || (2 * qq == N);
        const bool hu = ok && ru < nx, hv = ok && rv < nx;
        ur[q] = hu ? Sk[(size_t)qq * nx + ru] : T(0);
        vr[q] = hv ? Sk[(size_t)qq * nx + rv] : T(0);
        ui[q] = (hu && !single) ? Sk[(size_t)(N - qq) * nx + ru] : T(0);
        vi[q] = (hv && !single) ? Sk[(size_t)(N - qq) * nx + rv] : T(0);
    }
    }
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        if (e < H * TA) {
            const int tt = e % TA;
            const int qq = e / TA;
            const bool single = (qq == 0) || (2 * qq == N);
            sm[tt * LP + ct_pad(qq)] = Cx<T>::mk(ur[q] - vi[q], ui[q] + vr[q]);
            if (!single) sm[tt * LP + ct_pad(N - qq)] = Cx<T>::mk(ur[q] + vi[q], vr[q] - ui[q]);
        }
    }
    __syncthreads();
    fft_ct<N, TA, NT, true>(sm, tw);
    // even nx (and an aligned array): the two rows of a complex column are
    // adjacent, stored as one (re, im) pair
    if ((nx & 1) == 0 && ((size_t)S & (2 * sizeof(T) - 1)) == 0) {
        using V = typename Cx<T>::V;
        constexpr int TOT2 = TA * N;
        constexpr int NL2 = (TOT2 + NT - 1) / NT;
#pragma unroll
        for (int q = 0; q < NL2; ++q) {
            const int e = (int)threadIdx.x + q * NT;
            const int tt = e % TA;
            const int j = e / TA;
            if ((TOT2 % NT == 0 || e < TOT2) && r0 + 2 * tt < nx)
                *reinterpret_cast<V*>(Sk + (size_t)j * nx + r0 + 2 * tt) = sm[tt * LP + ct_pad(j)];
        }
    } else {
        const T* smd = reinterpret_cast<const T*>(sm);
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int e = (int)threadIdx.x + q * NT;
            const int r = e % W;
            const int j = e / W;
            if ((TOT % NT == 0 || e < TOT) && r0 + r < nx)
                Sk[(size_t)j * nx + r0 + r] = smd[((r >> 1) * LP + ct_pad(j)) * 2 + (r & 1)];
        }
    }
}

template <int N, int TA, int NT, typename T = double>
__global__ __launch_bounds__(NT) void kt_zbwd(T* S, int nx, int ncol, int ntc,
                                              const typename Cx<T>::V* __restrict__ tw) {
    __shared__ typename Cx<T>::V sm[TA * ct_pitch(N)];
    __shared__ typename Cx<T>::V stw[N];
    ct_stage_twiddles<N, NT>(stw, tw);
    zbwd_body<N, TA, NT, T>(S, nx, ntc, stw, xcd_tile(blockIdx.x, gridDim.x), sm);
}

// ---------------------------------------------------------------------------
// Blocked solve (nx a multiple of 16), through a workspace W.  Every pass
// reads contiguous memory and, where it must transpose, scatters its writes
// (a scattered read waits for its round trip, a scattered write does not):
//   x forward   P lines -> W  x-blocked [k][a/16][j][a%16]      (kt_xfwd_xb)
//   z forward   W       -> P  mode-major [a/16][b][k][a%16]     (kt_zfwd_mm)
//   tridiagonal P       -> W  x-blocked, from whole ny x 16 blocks
//   z inverse   W       -> P  halfcomplex lines                 (kt_zbwd_xl)
//   x inverse   P lines, in place                               (kt_xbwd)
// A z tile (plane k, halfcomplex x rows 16c..16c+15) is one contiguous block
// of nz * 16 values in W, and a tridiagonal workgroup reads one contiguous
// block of ny * 16 values.  Same transforms and arithmetic as the in-place
// kernels above, only the addresses differ.
// ---------------------------------------------------------------------------
// Waves per SIMD that workgroups of nt threads holding lds bytes of LDS each
// leave room for (160 KB of LDS per CU, four SIMDs), between 1 and cap: the
// occupancy a launch bound asks the register allocator for.
__host__ __device__ constexpr int lds_waves(int lds, int nt, int cap = 8) {
    return (163840 / lds) * (nt / 64) / 4 < 1 ? 1
         : ((163840 / lds) * (nt / 64) / 4 > cap ? cap : (163840 / lds) * (nt / 64) / 4);
}
template <int N, int G, int NT>
__global__ __launch_bounds__(NT) void kt_xfwd_xb(const double* P, double* W, long nlines, int nz,
                                                 const double2* __restrict__ tw) {
    __shared__ double2 sm[G * ct_pitch(N)];
    __shared__ double2 stw[N];
    ct_stage_twiddles<N, NT>(stw, tw);
    xfwd_body<N, G, NT, true>(P, W, nlines, stw, xcd_tile(blockIdx.x, gridDim.x), sm, nz);
}

// z forward, XB -> MM.  Tile (k, c): the 16 x rows 16c..16c+15 of plane k as
// 8 complex columns (rows 2t and 2t + 1 paired as in kt_zfwd); N = nz.
template <int N, int NT>
__global__ __launch_bounds__(NT) void kt_zfwd_mm(const double* __restrict__ W, double* __restrict__ S, int nx,
                                                 int ny, const double2* __restrict__ tw) {
    constexpr int TA = 8;
    constexpr int LP = ct_pitch(N);
    constexpr int TOT2 = TA * N;
    constexpr int NL2 = (TOT2 + NT - 1) / NT;
    __shared__ double2 sm[TA * LP];
    __shared__ double2 stw[N];
    ct_stage_twiddles<N, NT>(stw, tw);
    const int nab = nx / 16;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int k = tile / nab;
    const int c = tile - k * nab;
    // the tile is one contiguous block: pair e = (j, t) sits at 2e
    const double2* src = reinterpret_cast<const double2*>(W + ((size_t)k * nab + c) * N * 16);
    double2 v[NL2];
#pragma unroll
    for (int q = 0; q < NL2; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        v[q] = (TOT2 % NT == 0 || e < TOT2) ? src[e] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < NL2; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        if (TOT2 % NT == 0 || e < TOT2) sm[(e % TA) * LP + ct_pad(e / TA)] = v[q];
    }
    __syncthreads();
    fft_ct<N, TA, NT, false>(sm, stw);
    double* dst = S + (size_t)c * N * ny * 16 + (size_t)k * 16;
#pragma unroll
    for (int q = 0; q < NL2; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const int tt = e % TA;
        const int b = e / TA;
        if (TOT2 % NT == 0 || e < TOT2) {
            const bool lo = 2 * b <= N;
            const int qq = lo ? b : N - b;
            const double2 Z = sm[tt * LP + ct_pad(qq)];
            const double2 Zm = sm[tt * LP + ct_pad(qq == 0 ? 0 : N - qq)];
            const double ve = lo ? 0.5 * (Z.x + Zm.x) : 0.5 * (Z.y - Zm.y);
            const double vo = lo ? 0.5 * (Z.y + Zm.y) : 0.5 * (Zm.x - Z.x);
            *reinterpret_cast<double2*>(dst + (size_t)b * ny * 16 + 2 * tt) = make_double2(ve, vo);
        }
    }
}

// z inverse, x-blocked -> halfcomplex lines: the gather of rows q and N - q of
// kt_zbwd from the tile's contiguous block of W; N = nz.
template <int N, int NT>
__global__ __launch_bounds__(NT) void kt_zbwd_xl(const double* __restrict__ W, double* __restrict__ P, int nx,
                                                 int ny, const double2* __restrict__ tw) {
    constexpr int TA = 8;
    constexpr int LP = ct_pitch(N);
    constexpr int H = N / 2 + 1;
    constexpr int NI = (H * TA + NT - 1) / NT;
    constexpr int TOT2 = TA * N;
    constexpr int NL2 = (TOT2 + NT - 1) / NT;
    __shared__ double2 sm[TA * LP];
    __shared__ double2 stw[N];
    ct_stage_twiddles<N, NT>(stw, tw);
    const int nab = nx / 16;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int k = tile / nab;
    const int c = tile - k * nab;
    const double* src = W + ((size_t)k * nab + c) * N * 16;
    double2 pu[NI], pi[NI];
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const int ee = e < H * TA ? e : 0;
        const int tt = ee % TA;
        const int qq = ee / TA;
        const int qm = (qq == 0 || 2 * qq == N) ? qq : N - qq;
        pu[q] = *reinterpret_cast<const double2*>(src + (size_t)qq * 16 + 2 * tt);
        pi[q] = *reinterpret_cast<const double2*>(src + (size_t)qm * 16 + 2 * tt);
    }
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        if (e < H * TA) {
            const int tt = e % TA;
            const int qq = e / TA;
            const bool single = (qq == 0) || (2 * qq == N);
            const double ur = pu[q].x, vr = pu[q].y;
            const double ui = single ? 0.0 : pi[q].x, vi = single ? 0.0 : pi[q].y;
            sm[tt * LP + ct_pad(qq)] = make_double2(ur - vi, ui + vr);
            if (!single) sm[tt * LP + ct_pad(N - qq)] = make_double2(ur + vi, vr - ui);
        }
    }
    __syncthreads();
    fft_ct<N, TA, NT, true>(sm, stw);
    double* dst = P + (size_t)k * nx * N + (size_t)c * 16;
#pragma unroll
    for (int q = 0; q < NL2; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const int tt = e % TA;
        const int j = e / TA;
        if (TOT2 % NT == 0 || e < TOT2)
            *reinterpret_cast<double2*>(dst + (size_t)j * nx + 2 * tt) = sm[tt * LP + ct_pad(j)];
    }
}

// ---------------------------------------------------------------------------
// y: partitioned tridiagonal solve, one workgroup = M consecutive modes x P
// chunks of MR rows.  Thread (mode mi, chunk p) keeps its MR right-hand sides
// in registers; rows past ny are identity rows (a = c = 0, b = 1, d = 0).
//
// Chunk rows s..e (e = s+MR-1).  Downward elimination from row s+1 keeps the
// fill-in on x_s:  A_k x_s + B_k x_k + c_k x_{k+1} = D_k  (k = s+1..e), i.e.
//   f = a_k / B_{k-1},  A_k = -f A_{k-1},  B_k = beta_k - f c_{k-1},  D_k = d_k - f D_{k-1}.
// Interior rows then read x_k = D^_k - A^_k x_s - C^_k x_{k+1} (hat = divided
// by B_k).  An upward pass folds them into x_{s+1} = P - Q x_s - R x_e, which
// turns row s into  a_s x_{s-1} + (beta_s - c_s Q) x_s - c_s R x_e = d_s - c_s P.
// The 2P tip rows (s and e of every chunk) form a tridiagonal system per mode.
// Row s couples x_{s-1}, x_s, x_e and row e couples x_s, x_e, x_{e+1}; every
// chunk eliminates its own x_s and the next chunk's x_{e+1} (= that chunk's
// x_s, whose row arrives through LDS) from its row e, which leaves P rows in
// the unknowns x_e, tridiagonal again (the stride-1 step of cyclic reduction
// applied to the e rows only).  A workgroup whose modes all have
// |lambda| >= pcr_min solves them by parallel cyclic reduction in LDS, one
// row per thread, log2(P) steps.  Cyclic reduction loses digits by
// cancellation on the nearly singular low-wavenumber modes, so a workgroup
// holding any mode with |lambda| < pcr_min (the (0,0) mode among them) solves
// them by a serial Thomas sweep per mode instead.  x_s then follows from row
// s, the interior rows by the backward recurrence above.  Reciprocals in the chunk sweeps are v_rcp_f64 refined by
// two Newton steps (full double precision for the nonzero pivots met here).
// ---------------------------------------------------------------------------

__device__ __forceinline__ double rcp64(double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    r = fma(r, fma(-b, r, 1.0), r);
    return r;
}

// FP32 arithmetic of the mixed form's well-conditioned modes: v_rcp_f32
// (1 ulp) refined by one Newton step.
__device__ __forceinline__ float rcp64(float b) {
    float r = __builtin_amdgcn_rcpf(b);
    r = fmaf(r, fmaf(-b, r, 1.0f), r);
    return r;
}

// Bitwise mask of a value: all bits kept (keep) or +0.
__device__ __forceinline__ double keep_or_zero(double v, bool keep) {
    return __longlong_as_double(__double_as_longlong(v) & (keep ? -1LL : 0LL));
}
__device__ __forceinline__ float keep_or_zero(float v, bool keep) {
    return __int_as_float(__float_as_int(v) & (keep ? -1 : 0));
}

// A workgroup solves one mode group of M modes.  OCC is the minimum number
// of resident waves per SIMD the register allocation must allow.  T is the
// storage type of S (double, or float for the mixed-precision form); the
// arithmetic type R is FP64, or FP32 in the mixed form's workgroups whose
// modes are all well conditioned.  When msum is not NULL the (0,0) mode is
// solved from the FP64 plane sums msum[k] instead of S, its solution times
// inv_nxnz goes to pbar[k] and its S entries are set to zero.
// Element offset of row 0 of mode md (= a + nx b) and the stride between its
// rows: the reference layout (md + row * plane), or MM = true, the
// mode-major blocks [a/16][b][row][a%16] of the blocked solve (nx a
// multiple of 16), where a workgroup's 16 consecutive modes are one
// contiguous block of ny * 16 values (read; the rows go out to So in the
// x-blocked layout, tri_addr_out).
template <bool MM>
__device__ __forceinline__ void tri_addr(long md, long plane, int nx, int nz, int ny, size_t& base, size_t& rs) {
    if constexpr (MM) {
        const unsigned um = (unsigned)md;
        const unsigned a = um % (unsigned)nx, b = um / (unsigned)nx;
        base = ((size_t)(a >> 4) * (size_t)nz + b) * (size_t)ny * 16 + (a & 15);
        rs = 16;
    } else {
        base = (size_t)md;
        rs = (size_t)plane;
    }
}

// Where a lane stores its rows: in place (reference layout), or MM = true,
// the x-blocked layout [k][a/16][b][a%16] of the blocked solve's z inverse.
template <bool MM>
__device__ __forceinline__ void tri_addr_out(long md, long plane, int nx, int nz, int ny, size_t& base, size_t& rs) {
    if constexpr (MM) {
        const unsigned um = (unsigned)md;
        const unsigned a = um % (unsigned)nx, b = um / (unsigned)nx;
        base = ((size_t)(a >> 4) * (size_t)nz + b) * 16 + (a & 15);
        rs = (size_t)plane;
    } else {
        tri_addr<false>(md, plane, nx, nz, ny, base, rs);
    }
}

template <int MR, int P, int M, typename T, typename R = double, bool MM = false>
__device__ __forceinline__ void tri_load_rows(R (&d)[MR], const T* S, long mode, long plane, int ny, int s,
                                              const double* __restrict__ msum, int nx = 0, int nz = 0) {
    const bool valid = mode < plane;
    const bool mean = msum != nullptr && mode == 0;
    // Branch-free: every row load goes to a valid address (row and mode
    // clamped) and the masking follows all of them, so the MR loads are in
    // flight together (a load under a branch waits for its own round trip
    // before the next one issues).  The mask is a bitwise AND, not a select
    // the compiler could turn back into a branch, nor a 0/1 factor it could
    // fuse into the elimination's first multiply-add (which would change the
    // rounding); masked rows are +0.
    const long md = valid ? mode : 0;
    size_t base, rs;
    tri_addr<MM>(md, plane, nx, nz, ny, base, rs);
    T raw[MR];
#pragma unroll
    for (int r = 0; r < MR; ++r) {
        const int row = s + r < ny ? s + r : ny - 1;
        raw[r] = S[base + (size_t)row * rs];
    }
#pragma unroll
    for (int r = 0; r < MR; ++r) d[r] = keep_or_zero((R)raw[r], valid && !mean && s + r < ny);
    if (mean) {
#pragma unroll
        for (int r = 0; r < MR; ++r) d[r] = s + r < ny ? (R)msum[s + r] : R(0);
    }
}

// LDS of the reduced system: two buffers of P rows per mode, mode fastest.
// b holds the diagonal for the serial sweep and its reciprocal for cyclic
// reduction.  Buffer 1 first carries the s rows between neighbouring chunks.
template <int P, int M, typename R = double>
struct TriRed {
    R a[2][P][M], b[2][P][M], c[2][P][M], d[2][P][M];
};

// The loads a workgroup needs before its solve, issued together so that
// their round trips overlap (one after the other they cost three): the
// mode's eigenvalue lambda = ax[i] + az[j] (returned; -1 for idle lanes,
// which then solve a regular system), the group's band set `set` and the
// lane's MR rows (into d[], masked as tri_load_rows does).  The band set
// (a, b, c of ny rows each) goes to smb as three arrays of MR * P rows;
// rows past ny get a = c = 0, b = -1: decoupled from the rows above and
// below, with the diagonal -1 + lambda nonzero for every mode (lambda <= 0),
// so that they solve to x = 0 from d = 0 and the band reads need no bounds
// test.  Workgroups of M * P threads; the caller places a workgroup barrier
// before smb is read.
template <int MR, int P, int M, typename T, typename Ar = double, bool MM = false>
__device__ __forceinline__ double tri_prologue(Ar (&d)[MR], Ar* smb, const double* __restrict__ set, const T* S,
                                               const double* __restrict__ ax, const double* __restrict__ az,
                                               long mode, long plane, int nx, int ny,
                                               const double* __restrict__ msum, int nz = 0) {
    constexpr int NP = MR * P, NT = M * P, KB = (3 * NP + NT - 1) / NT;
    const bool valid = mode < plane;
    // plane is far below 2^32 (the FFT lengths are bounded by LDS)
    const unsigned um = valid ? (unsigned)mode : 0u;
    const double axv = ax[um % (unsigned)nx], azv = az[um / (unsigned)nx];
    double bv[KB];
#pragma unroll
    for (int q = 0; q < KB; ++q) {
        const int e = min((int)threadIdx.x + q * NT, 3 * NP - 1);
        const int which = e / NP, row = e - which * NP;
        bv[q] = set[which * ny + min(row, ny - 1)];
    }
    tri_load_rows<MR, P, M, T, Ar, MM>(d, S, mode, plane, ny, ((int)threadIdx.x / M) * MR, msum, nx, nz);
#pragma unroll
    for (int q = 0; q < KB; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        if (e < 3 * NP) {
            const int which = e / NP, row = e - which * NP;
            smb[e] = row < ny ? (Ar)bv[q] : (which == 1 ? Ar(-1) : Ar(0));
        }
    }
    return valid ? axv + azv : -1.0;
}

// One mode group: thread (mi, p) holds rows s..s+MR-1 of mode `mode` in d[]
// and writes its solution rows (mode >= plane: an idle lane).  smb: the band
// set of every mode of the group, staged in LDS by tri_prologue; R: the
// LDS reduced system.  Contains workgroup barriers; all threads call it with
// the same group.  Ar: the arithmetic type (bands in smb, d[], R).
//
// TP = false: the elimination keeps the normalised rows (A^, C^, D^) of the
// chunk interior in registers and back-substitutes them once x_s and x_e are
// known.  TP = true: the elimination keeps only the right-hand side and
// accumulates x_{s+1} = Pp - Qq x_s - Rr x_e forwards; once x_s and x_e are
// known the interior rows are solved again as a tridiagonal system with those
// two values as boundary data.  About a third fewer values stay live across
// the reduced solve, for one more elimination of the chunk.
template <int MR, int P, int M, typename T, bool TP = false, typename Ar = double, bool MM = false>
__device__ __forceinline__ void tri_group(Ar (&d)[MR], T* S, double lam64, const Ar* smb, int nx, int ny,
                                          long plane, double pcr_min, long mode, const double* __restrict__ msum,
                                          double* __restrict__ pbar, double inv_nxnz, TriRed<P, M, Ar>& R,
                                          int nz = 0, T* So = nullptr) {
    const int mi = (int)threadIdx.x % M;
    const int p = (int)threadIdx.x / M;
    const int s = p * MR;
    const bool valid = mode < plane;
    const long md = valid ? mode : 0;
    const Ar lam = (Ar)lam64;
    const Ar* A = smb;
    const Ar* B = A + MR * P;
    const Ar* C = B + MR * P;
    auto band_a = [&](int row) { return A[row]; };
    auto band_c = [&](int row) { return C[row]; };
    auto beta = [&](int row) { return B[row] + lam; };

    Ar Ah[TP ? 1 : MR], Ch[TP ? 1 : MR];  // TP = false: d[] is overwritten by D^ for interior rows
    Ar Ae, Be, De;
    Ar Pp = 0, Qq = 0, Rr = -1;  // x_{s+1} = Pp - Qq x_s - Rr x_e
    if constexpr (TP) {
        Ar Ak = band_a(s + 1);
        Ar Bk = beta(s + 1);
        Ar Dk = d[1];
        Ar cprev = band_c(s + 1);
        Ar pi = 1;  // product of -C^ over the rows eliminated so far
#pragma unroll
        for (int l = 2; l < MR; ++l) {
            const Ar rinv = rcp64(Bk);
            Pp += pi * (Dk * rinv);
            Qq += pi * (Ak * rinv);
            pi = -pi * (cprev * rinv);
            const Ar f = band_a(s + l) * rinv;
            Ak = -f * Ak;
            Bk = beta(s + l) - f * cprev;
            Dk = d[l] - f * Dk;
            cprev = band_c(s + l);
        }
        Rr = -pi;
        Ae = Ak;
        Be = Bk;
        De = Dk;
    } else {
        Ar Ak = band_a(s + 1);
        Ar Bk = beta(s + 1);
        Ar Dk = d[1];
        Ar cprev = band_c(s + 1);
#pragma unroll
        for (int l = 2; l < MR; ++l) {
            const Ar rinv = rcp64(Bk);
            d[l - 1] = Dk * rinv;
            Ah[l - 1] = Ak * rinv;
            Ch[l - 1] = cprev * rinv;
            const Ar f = band_a(s + l) * rinv;
            Ak = -f * Ak;
            Bk = beta(s + l) - f * cprev;
            Dk = d[l] - f * Dk;
            cprev = band_c(s + l);
        }
        Ae = Ak;
        Be = Bk;
        De = Dk;
        // upward pass
#pragma unroll
        for (int l = MR - 2; l >= 1; --l) {
            Pp = d[l] - Ch[l] * Pp;
            Qq = Ah[l] - Ch[l] * Qq;
            Rr = -Ch[l] * Rr;
        }
    }
    // row s: sa x_{s-1} + sb x_s + sc x_e = sd (sib = 1/sb)
    const Ar cs = band_c(s);
    const Ar sa = band_a(s), sb = beta(s) - cs * Qq, sc = -cs * Rr, sd = d[0] - cs * Pp;
    const Ar sib = rcp64(sb);
    R.a[1][p][mi] = sa;
    R.b[1][p][mi] = sib;
    R.c[1][p][mi] = sc;
    R.d[1][p][mi] = sd;
    const bool serial = __syncthreads_or(valid && fabs(lam64) < pcr_min) != 0;
    // row e with x_s and x_{e+1} eliminated: ya x_{e-} + yb x_e + yc x_{e+} = yd
    Ar ya, yb, yc, yd;
    {
        const Ar al = -Ae * sib;
        ya = p > 0 ? al * sa : Ar(0);
        yb = Be + al * sc;
        yd = De + al * sd;
        yc = Ar(0);
        if (p + 1 < P) {
            const Ar g = -band_c(s + MR - 1) * R.b[1][p + 1][mi];
            yc = g * R.c[1][p + 1][mi];
            yb += g * R.a[1][p + 1][mi];
            yd += g * R.d[1][p + 1][mi];
        }
    }
    Ar yib = serial ? yb : rcp64(yb);
    R.a[0][p][mi] = ya;
    R.b[0][p][mi] = yib;
    R.c[0][p][mi] = yc;
    R.d[0][p][mi] = yd;
    __syncthreads();
    Ar xe, xem;  // x_e of this chunk and of the one below
    if (serial) {
        // one thread per mode: Thomas on the P rows; x overwrites d
        if (p == 0) {
#pragma unroll 1
            for (int q = 1; q < P; ++q) {
                const Ar m = R.a[0][q][mi] / R.b[0][q - 1][mi];
                R.b[0][q][mi] = R.b[0][q][mi] - m * R.c[0][q - 1][mi];
                R.d[0][q][mi] = R.d[0][q][mi] - m * R.d[0][q - 1][mi];
            }
            Ar xq = R.d[0][P - 1][mi] / R.b[0][P - 1][mi];
            R.d[0][P - 1][mi] = xq;
#pragma unroll 1
            for (int q = P - 2; q >= 0; --q) {
                xq = (R.d[0][q][mi] - R.c[0][q][mi] * xq) / R.b[0][q][mi];
                R.d[0][q][mi] = xq;
            }
        }
        __syncthreads();
        xe = R.d[0][p][mi];
        xem = p > 0 ? R.d[0][p - 1][mi] : Ar(0);
    } else {
        // parallel cyclic reduction; step st reads buffer cur, writes cur ^ 1
        int cur = 0;
        for (int st = 1; st < P; st *= 2) {
            Ar a2 = 0, b2 = yb, c2 = 0, d2 = yd;
            if (p - st >= 0) {
                const Ar f = -ya * R.b[cur][p - st][mi];
                a2 = f * R.a[cur][p - st][mi];
                b2 += f * R.c[cur][p - st][mi];
                d2 += f * R.d[cur][p - st][mi];
            }
            if (p + st < P) {
                const Ar g = -yc * R.b[cur][p + st][mi];
                c2 = g * R.c[cur][p + st][mi];
                b2 += g * R.a[cur][p + st][mi];
                d2 += g * R.d[cur][p + st][mi];
            }
            ya = a2;
            yb = b2;
            yc = c2;
            yd = d2;
            yib = rcp64(yb);
            cur ^= 1;
            R.a[cur][p][mi] = ya;
            R.b[cur][p][mi] = yib;
            R.c[cur][p][mi] = yc;
            R.d[cur][p][mi] = yd;
            __syncthreads();
        }
        xe = yd * yib;
        xem = p > 0 ? R.d[cur][p - 1][mi] * R.b[cur][p - 1][mi] : Ar(0);
    }
    const Ar xs = (sd - sa * xem - sc * xe) * sib;
    d[0] = xs;
    if constexpr (TP) {
        // rows s+1..e-1 with x_s and x_e as boundary data; d[] becomes the
        // forward-eliminated right-hand side, then the solution.  The bands
        // are read from LDS again rather than kept in registers across the
        // reduced solve.
        asm volatile("" ::: "memory");
        if constexpr (MR > 2) {
            Ar cp[MR];
#pragma unroll
            for (int l = 1; l < MR - 1; ++l) {
                const Ar al = band_a(s + l);
                const Ar cl = band_c(s + l);
                Ar piv = beta(s + l), r = d[l];
                if (l == 1) r -= al * xs;
                else {
                    piv -= al * cp[l - 1];
                    r -= al * d[l - 1];
                }
                if (l == MR - 2) r -= cl * xe;
                const Ar rinv = rcp64(piv);
                cp[l] = cl * rinv;
                d[l] = r * rinv;
            }
            Ar xn = d[MR - 2];
#pragma unroll
            for (int l = MR - 3; l >= 1; --l) {
                xn = d[l] - cp[l] * xn;
                d[l] = xn;
            }
        }
    } else {
        Ar xn = xe;
#pragma unroll
        for (int l = MR - 2; l >= 1; --l) {
            xn = d[l] - Ah[l] * xs - Ch[l] * xn;
            d[l] = xn;
        }
    }
    d[MR - 1] = xe;
    if (valid) {
        const bool mean = msum != nullptr && mode == 0;
        // addresses from one offset the compiler cannot relate to the loads'
        // (otherwise it keeps every row index live across the solve)
        size_t base, rs;
        tri_addr_out<MM>(md, plane, nx, nz, ny, base, rs);
        size_t off = base + (size_t)s * rs;
        int nrow = ny - s;
        asm volatile("" : "+v"(off), "+v"(nrow));
        T* q = (MM ? So : S) + off;
#pragma unroll
        for (int r = 0; r < MR; ++r) {
            if (r < nrow) {
                if (mean) pbar[s + r] = (double)d[r] * inv_nxnz;
                q[(size_t)r * rs] = mean ? T(0) : (T)d[r];
            }
        }
    }
}

// Mode groups: only mode 0 (kx = kz = 0) uses the (0,0) band set, so group 0
// holds mode 0 alone and stages that set, group `ngroups` (one past the
// plain ones) holds modes 1..M-1, and group g in between modes g*M..g*M+M-1;
// every group stages one band set.
__device__ __forceinline__ long tri_group_mode(long grp, long ngroups, int M, int mi) {
    if (grp == 0) return mi == 0 ? 0 : LONG_MAX;
    if (grp == ngroups) return mi == 0 ? LONG_MAX : mi;
    return grp * M + mi;
}

// The same groups for the mode-major blocks of the blocked solve (M = 16, nx
// a multiple of 16), numbered z index first: group g in between holds x
// block g / nz of z index g % nz, so that consecutive groups (which run side
// by side on one XCD) read adjacent blocks and store adjacent 128-byte pieces.
__device__ __forceinline__ long tri_group_mode_mm(long grp, long ngroups, int mi, int nx, int nz) {
    if (grp == 0 || grp == ngroups) return tri_group_mode(grp, ngroups, 16, mi);
    const unsigned g = (unsigned)grp;
    const unsigned ab = g / (unsigned)nz, b = g - ab * (unsigned)nz;
    return (long)ab * 16 + mi + (long)nx * b;
}

// Grid: tri_groups(plane, M) workgroups of M * P threads, 3 * MR * P values
// of type Ar of dynamic LDS.  With glist, workgroup b solves group
// glist[b] (gridDim.x entries): the mixed form launches its FP32-arithmetic
// groups and its FP64 ones from two such lists.
template <int MR, int P, int M, typename T = double, int OCC = 1, bool TP = false, typename Ar = double,
          bool MM = false>
__global__ __launch_bounds__(M * P, OCC) void kt_tri_part(T* S, const double* __restrict__ ax,
                                                          const double* __restrict__ az,
                                                          const double* __restrict__ bands, int nx, int ny,
                                                          long plane, double pcr_min,
                                                          const double* __restrict__ msum = nullptr,
                                                          double* __restrict__ pbar = nullptr,
                                                          double inv_nxnz = 0.0,
                                                          const int* __restrict__ glist = nullptr, int nz = 0,
                                                          T* So = nullptr) {
    __shared__ TriRed<P, M, Ar> red;
    extern __shared__ double smb_raw[];  // the group's band set, 3 * MR * P values
    Ar* smb = reinterpret_cast<Ar*>(smb_raw);
    const int mi = (int)threadIdx.x % M;
    // adjacent mode groups share cache lines: keep them on one XCD (each has
    // its own L2; workgroups reach XCDs round-robin)
    const int slot = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const long grp = glist ? (long)glist[slot] : (long)slot;
    const long mode = MM && M == 16 ? tri_group_mode_mm(grp, plane / 16, mi, nx, nz)
                                    : tri_group_mode(grp, (plane + M - 1) / M, M, mi);
    Ar d[MR];
    const double lam64 = tri_prologue<MR, P, M, T, Ar, MM>(d, smb, bands + (grp == 0 ? 3 * ny : 0), S, ax, az, mode,
                                                           plane, nx, ny, msum, nz);
    __syncthreads();  // bands staged
    tri_group<MR, P, M, T, TP, Ar, MM>(d, S, lam64, smb, nx, ny, plane, pcr_min, mode, msum, pbar, inv_nxnz, red,
                                       nz, So);
}

}  // namespace dnsp
