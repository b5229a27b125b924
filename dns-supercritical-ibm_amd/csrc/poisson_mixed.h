// Kernels of the mixed-precision Poisson solve (FP32 transforms and FP32
// spectra, FP64 tridiagonal arithmetic, FP64 residual refinement).
//
// Sweep 0 solves L p0 = f with p0 = pt + pbar0(k): pt an FP32 field whose
// (0,0) mode is zero in the spectral solve, pbar0 the FP64 plane-mean profile
// solved from the FP64 plane sums of f.  A refinement sweep forms the FP64
// residual r = f - L p on the fly inside the forward x transform, solves
// L d = r the same way and writes p + d in FP64.  The operator L is the one
// the FP64 solve inverts (periodic second differences in x and z, the y bands,
// and the (0,0)-mode band set on the plane means), so the iterates converge to
// the FP64 solution; the plane means enter through per-plane constants q(k).
//
// Every kernel that produces a field also writes the FP64 sums of its x lines
// (fixed summation order, so results are bit-reproducible run to run).
#pragma once

#include <hip/hip_runtime.h>

#include "poisson_ct.h"

namespace dnsp {

constexpr int cgcd(int a, int b) { return b == 0 ? a : cgcd(b, a % b); }

// Slots of the line-sum reduction of an x tile: pairs are reduced over groups
// of cgcd(N/2, 64) lanes, which always lie in one line and one wavefront.
template <int N, int G>
constexpr int line_slots() {
    return G * N / cgcd(N / 2, 64);
}

// FP64 sums of the 2G real lines of an x tile.  Thread pair e = tid + q*NT
// covers elements 2e, 2e+1 of the tile (line 2e/N) and contributes ps[q]
// (zero for pairs outside the tile or past nlines).  Contains a barrier.
template <int N, int G, int NT, int NL>
__device__ __forceinline__ void tile_line_sums(const double (&ps)[NL], double* slot, double* __restrict__ lsum,
                                               long l0, long nlines) {
    constexpr int S = N / 2;
    constexpr int GW = cgcd(S, 64);
    static_assert(N % 2 == 0, "even lengths only");
    static_assert(2 * G <= NT, "one thread per line for the final sums");
    static_assert(NT % 64 == 0, "whole wavefronts");
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        double s = ps[q];
#pragma unroll
        for (int o = GW / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (e % GW == 0 && e < G * N) slot[e / GW] = s;
    }
    __syncthreads();
    const int t = (int)threadIdx.x;
    if (t < 2 * G && l0 + t < nlines) {
        double acc = 0.0;
        for (int u = 0; u < S / GW; ++u) acc += slot[t * (S / GW) + u];
        lsum[l0 + t] = acc;
    }
}

// Halfcomplex output of a forward x tile (as kt_xfwd), FP32.
template <int N, int G, int NT>
__device__ __forceinline__ void xfwd_store_f32(const float2* sm, float* S, long l0, long nlines) {
    constexpr int LP = ct_pitch(N);
    constexpr int NL = (G * N + NT - 1) / NT;
    float2* dst = reinterpret_cast<float2*>(S + l0 * N);
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const int r = (2 * e) / N;
        if ((G * N % NT == 0 || e < G * N) && l0 + r < nlines) {
            const int t0 = 2 * e - r * N;
            const int g = r >> 1;
            float out[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int t = t0 + h;
                const bool lo = 2 * t <= N;
                const int u = lo ? t : N - t;
                const float2 C = sm[g * LP + ct_pad(u)];
                const float2 Cm = sm[g * LP + ct_pad(u == 0 ? 0 : N - u)];
                if ((r & 1) == 0) out[h] = lo ? 0.5f * (C.x + Cm.x) : 0.5f * (C.y - Cm.y);
                else              out[h] = lo ? 0.5f * (C.y + Cm.y) : 0.5f * (Cm.x - C.x);
            }
            dst[e] = make_float2(out[0], out[1]);
        }
    }
}

// ---------------------------------------------------------------------------
// Sweep 0, forward x: FP64 right-hand side -> FP32 halfcomplex lines, plus the
// FP64 sums of the input lines.
// ---------------------------------------------------------------------------
// Bodies take the tile index, the LDS tile (G * ct_pitch(N) float2) and the
// line-sum slots (line_slots<N, G>() doubles); the __global__ wrappers pick
// the tile with xcd_tile.
template <int N, int G, int NT>
__device__ __forceinline__ void xfwd_mf_body(const double* __restrict__ f, float* __restrict__ S,
                                             double* __restrict__ lsum, long nlines, const float2* __restrict__ tw,
                                             long tile, float2* sm, double* slot) {
    constexpr int LP = ct_pitch(N);
    constexpr int NL = (G * N + NT - 1) / NT;
    float* smf = reinterpret_cast<float*>(sm);
    const long l0 = tile * 2 * G;
    const double2* src = reinterpret_cast<const double2*>(f + l0 * N);
    double2 v[NL];
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const int r = (2 * e) / N;
        v[q] = ((G * N % NT == 0 || e < G * N) && l0 + r < nlines) ? src[e] : make_double2(0.0, 0.0);
    }
    double ps[NL];
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        ps[q] = v[q].x + v[q].y;
        if (G * N % NT == 0 || e < G * N) {
            const int r = (2 * e) / N;
            const int t = 2 * e - r * N;
            const int g = r >> 1, c = r & 1;
            smf[(g * LP + ct_pad(t)) * 2 + c] = (float)v[q].x;
            smf[(g * LP + ct_pad(t + 1)) * 2 + c] = (float)v[q].y;
        }
    }
    tile_line_sums<N, G, NT, NL>(ps, slot, lsum, l0, nlines);
    fft_ct<N, G, NT, false>(sm, tw);
    xfwd_store_f32<N, G, NT>(sm, S, l0, nlines);
}

template <int N, int G, int NT>
__global__ __launch_bounds__(NT) void kt_xfwd_mf(const double* __restrict__ f, float* __restrict__ S,
                                                 double* __restrict__ lsum, long nlines,
                                                 const float2* __restrict__ tw) {
    __shared__ float2 sm[G * ct_pitch(N)];
    __shared__ double slot[line_slots<N, G>()];
    __shared__ float2 stw[N];
    ct_stage_twiddles<N, NT>(stw, tw);
    xfwd_mf_body<N, G, NT>(f, S, lsum, nlines, stw, xcd_tile(blockIdx.x, gridDim.x), sm, slot);
}

// ---------------------------------------------------------------------------
// Refinement, forward x: r = f - (L_xz + L_y) p - q(k) in FP64 for the tile's
// lines (p of type TP: the FP32 sweep-0 field or an FP64 iterate), then the
// FP32 halfcomplex transform of r and the FP64 sums of r's lines.
// bands = a, b, c of the plain y band set (ny each).  Lines are (j, k) with
// line = j + nz*k; x neighbours wrap, z neighbours wrap, y neighbours stop at
// the walls (the operator of the FP64 solve).
// ---------------------------------------------------------------------------
// Plane-strip tile order for the residual's wall-normal stencil: JB
// consecutive tiles of plane k, then the same tiles of plane k+1, and so on
// through all planes before the next strip, so a plane's lines are read again
// (as the k+1 and k-1 neighbours) while they are still in L2; in plane order
// they come back ~2 * nz lines later, after more traffic than L2 holds.  A
// permutation of [0, ny * nz / (2G)) when 2G divides nz (identity otherwise).
template <int G>
__device__ __forceinline__ long strip_tile(long t, int nz, int ny) {
    constexpr int LG = 2 * G;
    constexpr int JB = 8;
    if (nz % LG != 0) return t;
    const int tpp = nz / LG;
    if (t >= (long)tpp * ny) return t;
    const int nfull = tpp / JB;
    const long full = (long)nfull * JB * ny;
    long k, jg;
    if (t < full) {
        const long sidx = t / ((long)JB * ny);
        const long r = t - sidx * JB * ny;
        k = r / JB;
        jg = sidx * JB + (r - k * JB);
    } else {
        const int w = tpp - nfull * JB;
        const long r = t - full;
        k = r / w;
        jg = (long)nfull * JB + (r - k * w);
    }
    return k * tpp + jg;
}

template <int N, int G, int NT, typename TP>
__device__ __forceinline__ void xfwd_mr_body(const TP* __restrict__ p, const double* __restrict__ f,
                                             const double* __restrict__ qk, const double* __restrict__ bands,
                                             float* __restrict__ S, double* __restrict__ lsum, int nz, int ny,
                                             double ix2, double iz2, long nlines, const float2* __restrict__ tw,
                                             long tile, float2* sm, double* slot) {
    using P2 = typename Cx<TP>::V;
    constexpr int LP = ct_pitch(N);
    constexpr int NL = (G * N + NT - 1) / NT;
    float* smf = reinterpret_cast<float*>(sm);
    const long l0 = strip_tile<G>(tile, nz, ny) * 2 * G;
    const int k0 = (int)(l0 / nz);
    const int j0 = (int)(l0 - (long)k0 * nz);
    const unsigned long long mz = line_div_magic(nz);
    // Branch-free: every address is clamped to a valid line and the wall
    // terms carry a zero coefficient, so all loads of an element issue
    // together (branches around loads serialise their round trips).
    double ps[NL];
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const int r = (2 * e) / N;
        const int t0 = 2 * e - r * N;
        const bool in_tile = (G * N % NT == 0 || e < G * N);
        const bool ok = in_tile && l0 + r < nlines;
        const int rr = ok ? r : 0;
        const long line = l0 + rr;
        int k, j;
        tile_line_kj(k0, j0, rr, nz, mz, k, j);
        const TP* pl = p + line * N;
        const P2 c = *reinterpret_cast<const P2*>(pl + t0);
        const TP xl = pl[t0 == 0 ? N - 1 : t0 - 1];
        const TP xr = pl[t0 + 2 == N ? 0 : t0 + 2];
        const long lzm = line + (j == 0 ? nz - 1 : -1);
        const long lzp = line + (j == nz - 1 ? 1 - nz : 1);
        const P2 zm = *reinterpret_cast<const P2*>(p + lzm * N + t0);
        const P2 zp = *reinterpret_cast<const P2*>(p + lzp * N + t0);
        const P2 yd = *reinterpret_cast<const P2*>(p + (k > 0 ? line - nz : line) * N + t0);
        const P2 yu = *reinterpret_cast<const P2*>(p + (k < ny - 1 ? line + nz : line) * N + t0);
        const double2 fv = *reinterpret_cast<const double2*>(f + line * N + t0);
        // loads unconditional (0 <= k < ny), the wall selection after them
        const double av = bands[k], bk = bands[ny + k], cv = bands[2 * ny + k], qv = qk[k];
        const double ak = k > 0 ? av : 0.0;
        const double ck = k < ny - 1 ? cv : 0.0;
        const double cx = (double)c.x, cy = (double)c.y;
        double v0 = ix2 * (cy - 2 * cx + (double)xl) + iz2 * ((double)zp.x - 2 * cx + (double)zm.x) + bk * cx;
        double v1 = ix2 * ((double)xr - 2 * cy + cx) + iz2 * ((double)zp.y - 2 * cy + (double)zm.y) + bk * cy;
        v0 += ak * (double)yd.x;
        v1 += ak * (double)yd.y;
        v0 += ck * (double)yu.x;
        v1 += ck * (double)yu.y;
        // a 0/1 factor rather than a select, which the compiler would turn
        // back into a branch holding half of the loads (clamped loads are
        // finite, so the product is exact)
        const double okf = ok ? 1.0 : 0.0;
        const double r0 = (fv.x - (v0 + qv)) * okf;
        const double r1 = (fv.y - (v1 + qv)) * okf;
        ps[q] = r0 + r1;
        if (in_tile) {
            const int g = r >> 1, cc = r & 1;
            smf[(g * LP + ct_pad(t0)) * 2 + cc] = (float)r0;
            smf[(g * LP + ct_pad(t0 + 1)) * 2 + cc] = (float)r1;
        }
    }
    tile_line_sums<N, G, NT, NL>(ps, slot, lsum, l0, nlines);
    fft_ct<N, G, NT, false>(sm, tw);
    xfwd_store_f32<N, G, NT>(sm, S, l0, nlines);
}

// Staged residual tile, for nz % 2G == 0: a tile is then 2G consecutive z
// lines of one plane k0 (strip_tile keeps tiles whole), and every tile is
// full.  The tile's lines and the z line on either side of it (2G + 2 lines,
// periodic in z) are staged in LDS, so the x and z neighbours are LDS reads
// and a p line is read from L2/HBM once per tile instead of three times; the
// y neighbours and f are read directly.  The band values of plane k0 are
// workgroup-uniform.  Term for term the arithmetic of xfwd_mr_body.  stg
// (which aliases sm) holds (2G + 2) N values of TP.
template <int N, int G, typename TP>
constexpr bool mr_staged() {
    return (2 * G + 2) * N * (int)sizeof(TP) <= 48 * 1024;
}

// float2 slots of the residual kernel's LDS tile (staging area or FFT lines)
template <int N, int G, typename TP>
constexpr int mr_lds_f2() {
    constexpr int a = G * ct_pitch(N);
    constexpr int b = (2 * G + 2) * N * (int)sizeof(TP) / 8;
    return (mr_staged<N, G, TP>() && b > a) ? b : a;
}

template <int N, int G, int NT, typename TP>
__device__ __forceinline__ void xfwd_mr_staged(const TP* __restrict__ p, const double* __restrict__ f,
                                               const double* __restrict__ qk, const double* __restrict__ bands,
                                               float* __restrict__ S, double* __restrict__ lsum, int nz, int ny,
                                               double ix2, double iz2, long nlines,
                                               const float2* __restrict__ tw, long tile, float2* sm,
                                               double* slot) {
    using P2 = typename Cx<TP>::V;
    constexpr int LP = ct_pitch(N);
    constexpr int NL = (G * N + NT - 1) / NT;
    constexpr int NPS = (2 * G + 2) * N / 2;  // staged pairs
    constexpr int NS = (NPS + NT - 1) / NT;
    TP* stg = reinterpret_cast<TP*>(sm);
    float* smf = reinterpret_cast<float*>(sm);
    const long l0 = strip_tile<G>(tile, nz, ny) * 2 * G;
    const int k0 = (int)(l0 / nz);
    const int j0 = (int)(l0 - (long)k0 * nz);
    const long pl = (long)k0 * nz;
    // Loads first (clamped, unconditional), LDS writes after them.  Every
    // address is a workgroup-uniform base plus a 32-bit offset (a plane of
    // nz lines of N values is far below 2^32 elements), so the loads take a
    // scalar base and no 64-bit address arithmetic per element.
    const TP* pk = p + pl * N;  // plane k0
    P2 sv[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        int e = (int)threadIdx.x + q * NT;
        e = (NPS % NT == 0 || e < NPS) ? e : 0;
        const int u = (2 * e) / N;
        const int t = 2 * e - u * N;
        int jj = j0 - 1 + u;
        jj = jj < 0 ? jj + nz : (jj >= nz ? jj - nz : jj);
        sv[q] = *reinterpret_cast<const P2*>(reinterpret_cast<const char*>(pk) +
                                             (unsigned)(jj * N + t) * (unsigned)sizeof(TP));
    }
    const long dn = k0 > 0 ? -(long)nz : 0;
    const long up = k0 < ny - 1 ? (long)nz : 0;
    // line l0 + r, value t0 is element 2 ee of the tile's lines (r N + t0 = 2 ee)
    const TP* pd = p + (l0 + dn) * N;
    const TP* pu = p + (l0 + up) * N;
    const double* pf = f + l0 * N;
    P2 yd[NL], yu[NL];
    double2 fv[NL];
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const bool in_tile = (G * N % NT == 0 || e < G * N);
        const unsigned o2 = 2u * (unsigned)(in_tile ? e : 0);  // byte offsets below
        yd[q] = *reinterpret_cast<const P2*>(reinterpret_cast<const char*>(pd) + o2 * (unsigned)sizeof(TP));
        yu[q] = *reinterpret_cast<const P2*>(reinterpret_cast<const char*>(pu) + o2 * (unsigned)sizeof(TP));
        fv[q] = *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(pf) + o2 * 8u);
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        if (NPS % NT == 0 || e < NPS) *reinterpret_cast<P2*>(stg + 2 * e) = sv[q];
    }
    const double av = bands[k0], bk = bands[ny + k0], cv = bands[2 * ny + k0], qv = qk[k0];
    const double ak = k0 > 0 ? av : 0.0;
    const double ck = k0 < ny - 1 ? cv : 0.0;
    __syncthreads();
    double r0s[NL], r1s[NL], ps[NL];
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const bool in_tile = (G * N % NT == 0 || e < G * N);
        const int ee = in_tile ? e : 0;
        const int r = (2 * ee) / N;
        const int t0 = 2 * ee - r * N;
        const TP* ln = stg + (r + 1) * N;
        const P2 c = *reinterpret_cast<const P2*>(ln + t0);
        const TP xl = ln[t0 == 0 ? N - 1 : t0 - 1];
        const TP xr = ln[t0 + 2 == N ? 0 : t0 + 2];
        const P2 zm = *reinterpret_cast<const P2*>(ln - N + t0);
        const P2 zp = *reinterpret_cast<const P2*>(ln + N + t0);
        const double cx = (double)c.x, cy = (double)c.y;
        double v0 = ix2 * (cy - 2 * cx + (double)xl) + iz2 * ((double)zp.x - 2 * cx + (double)zm.x) + bk * cx;
        double v1 = ix2 * ((double)xr - 2 * cy + cx) + iz2 * ((double)zp.y - 2 * cy + (double)zm.y) + bk * cy;
        v0 += ak * (double)yd[q].x;
        v1 += ak * (double)yd[q].y;
        v0 += ck * (double)yu[q].x;
        v1 += ck * (double)yu[q].y;
        const double okf = in_tile ? 1.0 : 0.0;
        r0s[q] = (fv[q].x - (v0 + qv)) * okf;
        r1s[q] = (fv[q].y - (v1 + qv)) * okf;
        ps[q] = r0s[q] + r1s[q];
    }
    __syncthreads();  // staged lines read; sm becomes the FFT tile
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        if (G * N % NT == 0 || e < G * N) {
            const int r = (2 * e) / N;
            const int t0 = 2 * e - r * N;
            const int g = r >> 1, cc = r & 1;
            smf[(g * LP + ct_pad(t0)) * 2 + cc] = (float)r0s[q];
            smf[(g * LP + ct_pad(t0 + 1)) * 2 + cc] = (float)r1s[q];
        }
    }
    tile_line_sums<N, G, NT, NL>(ps, slot, lsum, l0, nlines);
    fft_ct<N, G, NT, false>(sm, tw);
    xfwd_store_f32<N, G, NT>(sm, S, l0, nlines);
}

template <int N, int G, int NT, typename TP>
__global__ __launch_bounds__(NT, (lds_waves(mr_lds_f2<N, G, TP>() * 8 + line_slots<N, G>() * 8 + N * 8, NT, 5)))
void kt_xfwd_mr(const TP* __restrict__ p, const double* __restrict__ f,
                                                 const double* __restrict__ qk, const double* __restrict__ bands,
                                                 float* __restrict__ S, double* __restrict__ lsum, int nz,
                                                 int ny, double ix2, double iz2, long nlines,
                                                 const float2* __restrict__ tw) {
    __shared__ __align__(16) float2 sm[mr_lds_f2<N, G, TP>()];
    __shared__ double slot[line_slots<N, G>()];
    __shared__ float2 stw[N];
    ct_stage_twiddles<N, NT>(stw, tw);
    const long tile = xcd_tile(blockIdx.x, gridDim.x);
    if constexpr (mr_staged<N, G, TP>()) {
        if (nz % (2 * G) == 0) {
            xfwd_mr_staged<N, G, NT, TP>(p, f, qk, bands, S, lsum, nz, ny, ix2, iz2, nlines, stw, tile, sm, slot);
            return;
        }
    }
    xfwd_mr_body<N, G, NT, TP>(p, f, qk, bands, S, lsum, nz, ny, ix2, iz2, nlines, stw, tile, sm, slot);
}

// Halfcomplex pairs of lines -> LDS complex lines ready for the inverse FFT
// (as kt_xbwd), FP32.
template <int N, int G, int NT>
__device__ __forceinline__ void xbwd_load_f32(float2* sm, const float* S, long l0, long nlines) {
    constexpr int LP = ct_pitch(N);
    constexpr int H = N / 2 + 1;
    constexpr int NI = (G * H + NT - 1) / NT;
    const float* src = S + l0 * N;
    float xr[NI], xi[NI], yr[NI], yi[NI];
    // 16-byte aligned tile (every supported length is a multiple of 4): the
    // 2G lines are staged in LDS (inside sm, which they fit) with contiguous
    // 16-byte loads from clamped lines, and the t / N - t gather reads LDS
    // (a few percent faster than 4-byte global gathers at N = 768).
    bool staged = false;
    if constexpr (N % 4 == 0) {
        if ((((size_t)src) & 15) == 0) {
            staged = true;
            constexpr int NV = 2 * G * N / 4;
            constexpr int NQ = (NV + NT - 1) / NT;
            float4 v[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                int e = (int)threadIdx.x + q * NT;
                e = (NV % NT == 0 || e < NV) ? e : 0;
                const int line = (4 * e) / N;
                const int off = 4 * e - line * N;
                const int ln = l0 + line < nlines ? line : 0;
                v[q] = *reinterpret_cast<const float4*>(src + (long)ln * N + off);
            }
            float2* stg = sm;  // float index 4e -> float2 index 2e
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int e = (int)threadIdx.x + q * NT;
                if (NV % NT == 0 || e < NV) {
                    stg[2 * e] = make_float2(v[q].x, v[q].y);
                    stg[2 * e + 1] = make_float2(v[q].z, v[q].w);
                }
            }
            __syncthreads();
            const float* sf = reinterpret_cast<const float*>(sm);
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
                const float* la = sf + (2 * g) * N;
                const float* lb = la + N;
                const float a0 = la[t], a1 = la[tm], b0 = lb[t], b1 = lb[tm];
                xr[q] = va ? a0 : 0.f;
                xi[q] = (va && !single) ? a1 : 0.f;
                yr[q] = vb ? b0 : 0.f;
                yi[q] = (vb && !single) ? b1 : 0.f;
            }
            __syncthreads();  // staged lines read; sm becomes the FFT lines
        }
    }
    if (!staged) {
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const int g = e / H;
        const int t = e - g * H;
        const bool ok = e < G * H;
        const bool single = (t == 0) || (2 * t == N);
        const bool va = ok && (l0 + 2 * g < nlines);
        const bool vb = ok && (l0 + 2 * g + 1 < nlines);
        const float* la = src + (2 * g) * N;
        const float* lb = la + N;
        xr[q] = va ? la[t] : 0.f;
        xi[q] = (va && !single) ? la[N - t] : 0.f;
        yr[q] = vb ? lb[t] : 0.f;
        yi[q] = (vb && !single) ? lb[N - t] : 0.f;
    }
    }
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        if (e < G * H) {
            const int g = e / H;
            const int t = e - g * H;
            const bool single = (t == 0) || (2 * t == N);
            sm[g * LP + ct_pad(t)] = make_float2(xr[q] - yi[q], xi[q] + yr[q]);
            if (!single) sm[g * LP + ct_pad(N - t)] = make_float2(xr[q] + yi[q], yr[q] - xi[q]);
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Sweep 0, inverse x: FP32 halfcomplex lines -> FP32 field pt = scale * x-HC2R
// (in place allowed: S == pt), plus the FP64 sums of the stored pt lines.
// ---------------------------------------------------------------------------
template <int N, int G, int NT>
__device__ __forceinline__ void xbwd_mf_body(const float* S, float* pt, double* __restrict__ lsum, long nlines,
                                             const float2* __restrict__ tw, float scale, long tile, float2* sm,
                                             double* slot) {
    constexpr int LP = ct_pitch(N);
    constexpr int NL = (G * N + NT - 1) / NT;
    const long l0 = tile * 2 * G;
    xbwd_load_f32<N, G, NT>(sm, S, l0, nlines);
    fft_ct<N, G, NT, true>(sm, tw);
    const float* smf = reinterpret_cast<const float*>(sm);
    float2* dst = reinterpret_cast<float2*>(pt + l0 * N);
    double ps[NL];
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const int r = (2 * e) / N;
        ps[q] = 0.0;
        if ((G * N % NT == 0 || e < G * N) && l0 + r < nlines) {
            const int t = 2 * e - r * N;
            const int g = r >> 1, c = r & 1;
            const float2 o = make_float2(smf[(g * LP + ct_pad(t)) * 2 + c] * scale,
                                         smf[(g * LP + ct_pad(t + 1)) * 2 + c] * scale);
            dst[e] = o;
            ps[q] = (double)o.x + (double)o.y;
        }
    }
    tile_line_sums<N, G, NT, NL>(ps, slot, lsum, l0, nlines);
}

template <int N, int G, int NT>
__global__ __launch_bounds__(NT) void kt_xbwd_mf(const float* S, float* pt, double* __restrict__ lsum,
                                                 long nlines, const float2* __restrict__ tw, float scale) {
    __shared__ float2 sm[G * ct_pitch(N)];
    __shared__ double slot[line_slots<N, G>()];
    __shared__ float2 stw[N];
    ct_stage_twiddles<N, NT>(stw, tw);
    xbwd_mf_body<N, G, NT>(S, pt, lsum, nlines, stw, scale, xcd_tile(blockIdx.x, gridDim.x), sm, slot);
}

// ---------------------------------------------------------------------------
// Refinement, inverse x with accumulation into FP64:
//   P = base + (pbar0[k] + pbar1[k]) + scale * x-HC2R(S)
// base is the FP32 sweep-0 field (with its profile pbar0), the FP64 iterate
// itself (base == P, pbar0 == NULL) or NULL (zero); pbar0, pbar1 may be NULL.  lsum (optional) receives the FP64 sums
// of P's lines for a following sweep or residual check.
// ---------------------------------------------------------------------------
template <int N, int G, int NT, typename TP>
__device__ __forceinline__ void xbwd_ma_body(const float* __restrict__ S, const TP* base,
                                             const double* __restrict__ pbar0, const double* __restrict__ pbar1,
                                             double* P, double* __restrict__ lsum, int nz, long nlines,
                                             const float2* __restrict__ tw, double scale, long tile, float2* sm,
                                             double* slot) {
    using P2 = typename Cx<TP>::V;
    constexpr int LP = ct_pitch(N);
    constexpr int NL = (G * N + NT - 1) / NT;
    const long l0 = tile * 2 * G;
    const int k0 = (int)(l0 / nz);
    const int j0 = (int)(l0 - (long)k0 * nz);
    const unsigned long long mz = line_div_magic(nz);
    xbwd_load_f32<N, G, NT>(sm, S, l0, nlines);
    fft_ct<N, G, NT, true>(sm, tw);
    const float* smf = reinterpret_cast<const float*>(sm);
    double ps[NL];
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int e = (int)threadIdx.x + q * NT;
        const int r = (2 * e) / N;
        ps[q] = 0.0;
        if ((G * N % NT == 0 || e < G * N) && l0 + r < nlines) {
            const int t = 2 * e - r * N;
            const int g = r >> 1, c = r & 1;
            const long line = l0 + r;
            int k, j;
            tile_line_kj(k0, j0, r, nz, mz, k, j);
            const double add = (pbar0 ? pbar0[k] : 0.0) + (pbar1 ? pbar1[k] : 0.0);
            P2 b;
            b.x = TP(0);
            b.y = TP(0);
            if (base != nullptr) b = *reinterpret_cast<const P2*>(base + line * N + t);
            const double o0 = (double)b.x + (add + scale * (double)smf[(g * LP + ct_pad(t)) * 2 + c]);
            const double o1 = (double)b.y + (add + scale * (double)smf[(g * LP + ct_pad(t + 1)) * 2 + c]);
            *reinterpret_cast<double2*>(P + line * N + t) = make_double2(o0, o1);
            ps[q] = o0 + o1;
        }
    }
    if (lsum != nullptr) tile_line_sums<N, G, NT, NL>(ps, slot, lsum, l0, nlines);
}

template <int N, int G, int NT, typename TP>
__global__ __launch_bounds__(NT) void kt_xbwd_ma(const float* __restrict__ S, const TP* base,
                                                 const double* __restrict__ pbar0,
                                                 const double* __restrict__ pbar1, double* P,
                                                 double* __restrict__ lsum, int nz, long nlines,
                                                 const float2* __restrict__ tw, double scale) {
    __shared__ float2 sm[G * ct_pitch(N)];
    __shared__ double slot[line_slots<N, G>()];
    __shared__ float2 stw[N];
    ct_stage_twiddles<N, NT>(stw, tw);
    xbwd_ma_body<N, G, NT, TP>(S, base, pbar0, pbar1, P, lsum, nz, nlines, stw, scale,
                               xcd_tile(blockIdx.x, gridDim.x), sm, slot);
}

}  // namespace dnsp
