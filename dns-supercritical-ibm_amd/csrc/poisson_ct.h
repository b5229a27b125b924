// Compile-time-sized kernels of the Poisson solve: x and z transforms for
// lengths 32..768 and the partitioned wall-normal tridiagonal solve.  Layout and halfcomplex conventions are those
// of the runtime-sized kernels in poisson.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_ct.h"

namespace dnsp {

// FFT kernel variants: (complex elements per workgroup tile, threads).
// The tuner picks one per direction and handle.
constexpr int kCtVariants = 4;
__host__ __device__ constexpr int ct_var_tile(int v) { return v == 0 ? 3072 : v == 1 ? 3072 : v == 2 ? 1536 : 6144; }
__host__ __device__ constexpr int ct_var_threads(int v) { return v == 0 ? 256 : v == 1 ? 512 : v == 2 ? 256 : 512; }

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
template <int N, int G, int NT>
__device__ __forceinline__ void xfwd_body(const double* P, double* S, long nlines, const double2* __restrict__ tw,
                                          long tile, double2* sm) {
    constexpr int LP = ct_pitch(N);
    constexpr int NL = (G * N + NT - 1) / NT;  // double2 loads per thread
    double* smd = reinterpret_cast<double*>(sm);
    const long l0 = tile * 2 * G;
    const double2* src = reinterpret_cast<const double2*>(P + l0 * N);
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
            dst[e] = make_double2(out[0], out[1]);
        }
    }
}

template <int N, int G, int NT>
__global__ __launch_bounds__(NT) void kt_xfwd(const double* P, double* S, long nlines,
                                              const double2* __restrict__ tw) {
    __shared__ double2 sm[G * ct_pitch(N)];
    xfwd_body<N, G, NT>(P, S, nlines, tw, xcd_tile(blockIdx.x, gridDim.x), sm);
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
    xbwd_body<N, G, NT>(S, P, nlines, tw, scale, xcd_tile(blockIdx.x, gridDim.x), sm);
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
    __syncthreads();
    fft_ct<N, TA, NT, false>(sm, tw);
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
            const T h = T(0.5);
            T val;
            if ((r & 1) == 0) val = lo ? h * (Z.x + Zm.x) : h * (Z.y - Zm.y);
            else              val = lo ? h * (Z.y + Zm.y) : h * (Zm.x - Z.x);
            Sk[(size_t)b * nx + r0 + r] = val;
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
    zfwd_body<N, TA, NT, T>(S, nx, ntc, tw, lsum, msum, xcd_tile(blockIdx.x, gridDim.x), sm);
}

template <int N, int TA, int NT, typename T>
__device__ __forceinline__ void zbwd_body(T* S, int nx, int ntc, const typename Cx<T>::V* __restrict__ tw,
                                          int tile, typename Cx<T>::V* sm) {
    using V = typename Cx<T>::V;
    constexpr int LP = ct_pitch(N);
    constexpr int W = 2 * TA;
    constexpr int H = N / 2 + 1;
    constexpr int NI = (H * TA + NT - 1) / NT;
    constexpr int TOT = W * N;
    constexpr int NL = (TOT + NT - 1) / NT;
    const int k = tile / ntc;
    const int r0 = (tile - k * ntc) * W;
    T* Sk = S + (size_t)k * nx * N;
    T ur[NI], ui[NI], vr[NI], vi[NI];
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

template <int N, int TA, int NT, typename T = double>
__global__ __launch_bounds__(NT) void kt_zbwd(T* S, int nx, int ncol, int ntc,
                                              const typename Cx<T>::V* __restrict__ tw) {
    __shared__ typename Cx<T>::V sm[TA * ct_pitch(N)];
    zbwd_body<N, TA, NT, T>(S, nx, ntc, tw, xcd_tile(blockIdx.x, gridDim.x), sm);
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
// A workgroup whose modes all have |lambda| >= pcr_min solves it by parallel
// cyclic reduction in LDS.  Cyclic reduction loses digits by cancellation on
// the nearly singular low-wavenumber modes, so a workgroup holding any mode
// with |lambda| < pcr_min (the (0,0) mode among them) solves it by a serial
// Thomas sweep per mode instead.  The interior rows follow by the backward
// recurrence above.  Reciprocals in the chunk sweeps are v_rcp_f64 refined by
// two Newton steps (full double precision for the nonzero pivots met here).
// ---------------------------------------------------------------------------

__device__ __forceinline__ double rcp64(double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    r = fma(r, fma(-b, r, 1.0), r);
    return r;
}

// Without LOOP a workgroup solves one mode group of M modes.  With LOOP a
// workgroup loops over mode groups and stages the bands in LDS once; the loop
// costs registers (fewer resident waves).  OCC is the minimum number of
// resident waves per SIMD the register allocation must allow.  T is the storage type of S
// (double, or float for the mixed-precision form); the arithmetic is FP64.
// When msum is not NULL the (0,0) mode is solved from the FP64 plane sums
// msum[k] instead of S, its solution times inv_nxnz goes to pbar[k] and its
// S entries are set to zero.
template <int MR, int P, int M, typename T>
__device__ __forceinline__ void tri_load_rows(double (&d)[MR], const T* S, long mode, long plane, int ny, int s,
                                              const double* __restrict__ msum) {
    const bool valid = mode < plane;
    const bool mean = msum != nullptr && mode == 0;
#pragma unroll
    for (int r = 0; r < MR; ++r) {
        const int row = s + r;
        if (mean) d[r] = row < ny ? msum[row] : 0.0;
        else d[r] = (valid && row < ny) ? (double)S[(size_t)row * plane + mode] : 0.0;
    }
}

// One mode group: thread (mi, p) holds rows s..s+MR-1 of mode `mode` in d[]
// and writes its solution rows.  smb: the staged bands (plain set, then the
// (0,0)-mode set where needed); ra..rd: the LDS reduced system.  Contains
// workgroup barriers; all threads call it with the same group.
template <int MR, int P, int M, typename T>
__device__ __forceinline__ void tri_group(double (&d)[MR], T* S, const double* __restrict__ ax,
                                          const double* __restrict__ az, const int* __restrict__ is0x,
                                          const int* __restrict__ is0z, const double* smb, int nx, int ny,
                                          long plane, double pcr_min, long mode, const double* __restrict__ msum,
                                          double* __restrict__ pbar, double inv_nxnz, double (*ra)[M],
                                          double (*rb)[M], double (*rc)[M], double (*rd)[M]) {
    constexpr int NT = M * P;
    constexpr int NR = 2 * P;  // reduced rows per mode
    const int mi = (int)threadIdx.x % M;
    const int p = (int)threadIdx.x / M;
    const int s = p * MR;
    const bool valid = mode < plane;
    const long md = valid ? mode : 0;
    const int i = (int)(md % nx);
    const int j = (int)(md / nx);
    const double lam = ax[i] + az[j];
    const double* A = smb + ((is0x[i] && is0z[j]) ? 3 * ny : 0);
    const double* B = A + ny;
    const double* C = B + ny;
    auto band_a = [&](int row) { return row < ny ? A[row] : 0.0; };
    auto band_c = [&](int row) { return row < ny ? C[row] : 0.0; };
    auto beta = [&](int row) { return row < ny ? B[row] + lam : 1.0; };

    double Ah[MR], Ch[MR];  // d[] is overwritten by D^ for interior rows
    double Ae, Be, De;
    {
        double Ak = band_a(s + 1);
        double Bk = beta(s + 1);
        double Dk = d[1];
        double cprev = band_c(s + 1);
#pragma unroll
        for (int l = 2; l < MR; ++l) {
            const double rinv = rcp64(Bk);
            d[l - 1] = Dk * rinv;
            Ah[l - 1] = Ak * rinv;
            Ch[l - 1] = cprev * rinv;
            const double f = band_a(s + l) * rinv;
            Ak = -f * Ak;
            Bk = beta(s + l) - f * cprev;
            Dk = d[l] - f * Dk;
            cprev = band_c(s + l);
        }
        Ae = Ak;
        Be = Bk;
        De = Dk;
    }
    // upward pass: x_{s+1} = Pp - Qq x_s - Rr x_e
    double Pp = 0.0, Qq = 0.0, Rr = -1.0;
#pragma unroll
    for (int l = MR - 2; l >= 1; --l) {
        Pp = d[l] - Ch[l] * Pp;
        Qq = Ah[l] - Ch[l] * Qq;
        Rr = -Ch[l] * Rr;
    }
    {
        const double cs = band_c(s);
        ra[2 * p][mi] = band_a(s);
        rb[2 * p][mi] = beta(s) - cs * Qq;
        rc[2 * p][mi] = -cs * Rr;
        rd[2 * p][mi] = d[0] - cs * Pp;
        ra[2 * p + 1][mi] = Ae;
        rb[2 * p + 1][mi] = Be;
        rc[2 * p + 1][mi] = band_c(s + MR - 1);
        rd[2 * p + 1][mi] = De;
    }
    const bool serial = __syncthreads_or(valid && fabs(lam) < pcr_min) != 0;
    if (serial) {
        // one thread per mode: Thomas on the NR tip rows; x overwrites rd, rb = 1
        if (p == 0) {
            for (int q = 1; q < NR; ++q) {
                const double m = ra[q][mi] / rb[q - 1][mi];
                rb[q][mi] = rb[q][mi] - m * rc[q - 1][mi];
                rd[q][mi] = rd[q][mi] - m * rd[q - 1][mi];
            }
            double xq = rd[NR - 1][mi] / rb[NR - 1][mi];
            rd[NR - 1][mi] = xq;
            rb[NR - 1][mi] = 1.0;
            for (int q = NR - 2; q >= 0; --q) {
                xq = (rd[q][mi] - rc[q][mi] * xq) / rb[q][mi];
                rd[q][mi] = xq;
                rb[q][mi] = 1.0;
            }
        }
        __syncthreads();
    }
    // parallel cyclic reduction on the NR tip rows of each mode
    for (int st = 1; !serial && st < NR; st *= 2) {
        double na[2], nb[2], nc[2], nd[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = (int)threadIdx.x + h * NT;
            const int mm = e % M;
            const int q = e / M;
            const double aq = ra[q][mm], bq = rb[q][mm], cq = rc[q][mm], dq = rd[q][mm];
            double a2 = 0.0, b2 = bq, c2 = 0.0, d2 = dq;
            if (q - st >= 0) {
                const double al = -aq * rcp64(rb[q - st][mm]);
                a2 = al * ra[q - st][mm];
                b2 += al * rc[q - st][mm];
                d2 += al * rd[q - st][mm];
            }
            if (q + st < NR) {
                const double ga = -cq * rcp64(rb[q + st][mm]);
                c2 = ga * rc[q + st][mm];
                b2 += ga * ra[q + st][mm];
                d2 += ga * rd[q + st][mm];
            }
            na[h] = a2; nb[h] = b2; nc[h] = c2; nd[h] = d2;
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = (int)threadIdx.x + h * NT;
            const int mm = e % M;
            const int q = e / M;
            ra[q][mm] = na[h]; rb[q][mm] = nb[h]; rc[q][mm] = nc[h]; rd[q][mm] = nd[h];
        }
        __syncthreads();
    }
    const double xs = serial ? rd[2 * p][mi] : rd[2 * p][mi] * rcp64(rb[2 * p][mi]);
    const double xe = serial ? rd[2 * p + 1][mi] : rd[2 * p + 1][mi] * rcp64(rb[2 * p + 1][mi]);
    d[0] = xs;
    double xn = xe;
#pragma unroll
    for (int l = MR - 2; l >= 1; --l) {
        xn = d[l] - Ah[l] * xs - Ch[l] * xn;
        d[l] = xn;
    }
    d[MR - 1] = xe;
    if (valid) {
        const bool mean = msum != nullptr && mode == 0;
#pragma unroll
        for (int r = 0; r < MR; ++r) {
            const int row = s + r;
            if (row < ny) {
                if (mean) pbar[row] = d[r] * inv_nxnz;
                S[(size_t)row * plane + md] = mean ? T(0) : (T)d[r];
            }
        }
    }
}

template <int MR, int P, int M, typename T = double, bool LOOP = false, int OCC = 1>
__global__ __launch_bounds__(M * P, OCC) void kt_tri_part(T* S, const double* __restrict__ ax,
                                                          const double* __restrict__ az,
                                                          const int* __restrict__ is0x,
                                                          const int* __restrict__ is0z,
                                                          const double* __restrict__ bands, int nx, int ny,
                                                          long plane, double pcr_min, int ngroups,
                                                          const double* __restrict__ msum = nullptr,
                                                          double* __restrict__ pbar = nullptr,
                                                          double inv_nxnz = 0.0) {
    constexpr int NT = M * P;
    constexpr int NR = 2 * P;
    __shared__ double ra[NR][M], rb[NR][M], rc[NR][M], rd[NR][M];  // mode fastest
    extern __shared__ double smb[];                                  // 6*ny band values
    const int mi = (int)threadIdx.x % M;
    const int s = ((int)threadIdx.x / M) * MR;

    // Adjacent mode groups share cache lines, so they go to one XCD (each has
    // its own L2; workgroups reach XCDs round-robin): without LOOP through
    // xcd_tile, with LOOP as a contiguous range of groups per XCD, strided
    // over that XCD's workgroups (gridDim.x >= 8).
    int grp, g_end, g_step;
    if constexpr (LOOP) {
        const int xcd = (int)blockIdx.x % 8;
        g_step = (int)gridDim.x / 8 + (xcd < (int)gridDim.x % 8 ? 1 : 0);
        grp = (int)((long)ngroups * xcd / 8) + (int)blockIdx.x / 8;
        g_end = (int)((long)ngroups * (xcd + 1) / 8);
    } else {
        grp = xcd_tile((int)blockIdx.x, (int)gridDim.x);
        g_end = grp + 1;
        g_step = 1;
    }
    if (grp >= g_end) return;  // workgroup-uniform
    // the (0,0)-mode band set is only needed by the workgroup holding mode 0
    const int nband = (grp == 0 ? 6 : 3) * ny;
    for (int e = (int)threadIdx.x; e < nband; e += NT) smb[e] = bands[e];
    double d[MR];
    tri_load_rows<MR, P, M, T>(d, S, (long)grp * M + mi, plane, ny, s, msum);
    __syncthreads();  // bands staged.  Later groups need no barrier before
                      // their rows: each thread's last reads of the reduced
                      // rows are its own slots, which only it rewrites.
    for (; grp < g_end; grp += g_step) {
        tri_group<MR, P, M, T>(d, S, ax, az, is0x, is0z, smb, nx, ny, plane, pcr_min, (long)grp * M + mi, msum,
                               pbar, inv_nxnz, ra, rb, rc, rd);
        if constexpr (!LOOP) break;
        const int nxt = grp + g_step;
        if (nxt < g_end) tri_load_rows<MR, P, M, T>(d, S, (long)nxt * M + mi, plane, ny, s, msum);
    }
}

}  // namespace dnsp
