// Compile-time-sized in-LDS batched complex FFT (FP64 or FP32) for gfx950.
//
// Same Stockham formulation as fft_lds.h, with the length N, the radix
// sequence, the lines per workgroup G and the thread count NT known at
// compile time, so every index is a constant expression or a shift and each
// thread's butterflies are unrolled into registers.  LDS rows are padded with
// one complex slot per 8 (ct_pad) so the radix-8 scatter writes of a
// wavefront spread over all 64 banks.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_lds.h"

namespace dnsp {

struct CtPlan {
    int nf;
    int r[16];
};

// Radix order: 8s (16 -> 4*4), then 4, 2, 3, 5, 7.
constexpr CtPlan ct_plan(int n) {
    CtPlan p{};
    p.nf = 0;
    while (n % 8 == 0 && n != 16) { p.r[p.nf++] = 8; n /= 8; }
    while (n % 4 == 0) { p.r[p.nf++] = 4; n /= 4; }
    while (n % 2 == 0) { p.r[p.nf++] = 2; n /= 2; }
    while (n % 3 == 0) { p.r[p.nf++] = 3; n /= 3; }
    while (n % 5 == 0) { p.r[p.nf++] = 5; n /= 5; }
    while (n % 7 == 0) { p.r[p.nf++] = 7; n /= 7; }
    if (n != 1) p.nf = -1;
    return p;
}

constexpr int ct_ns(int n, int s) {
    CtPlan p = ct_plan(n);
    int v = 1;
    for (int i = 0; i < s; ++i) v *= p.r[i];
    return v;
}

__host__ __device__ constexpr int ct_pad(int i) { return i + (i >> 3); }
// LDS pitch (in complex elements) of one padded line of length n
__host__ __device__ constexpr int ct_pitch(int n) { return n + (n >> 3) + 1; }

template <int N, int R, int NS, int G, int NT, bool INV, typename V>
__device__ __forceinline__ void ct_stage(V* buf, const V* __restrict__ tw) {
    constexpr int m = N / R;
    constexpr int tstep = N / (NS * R);
    constexpr int total = G * m;
    constexpr int NB = (total + NT - 1) / NT;
    constexpr int LP = ct_pitch(N);
    V v[NB][R];
    int base[NB], dst[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int idx = (int)threadIdx.x + q * NT;
        if ((total % NT == 0) || idx < total) {
            const int g = idx / m;
            const int j = idx - g * m;
            const int k = j % NS;
            base[q] = g * LP;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = buf[base[q] + ct_pad(j + r * m)];
            // no branch on k != 0: tw[0] = (1, 0) exactly, so k = 0 multiplies
            // by one, and the twiddle loads of every q issue together (a load
            // under a lane branch waits for its round trip before the next)
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const V w = tw[r * k * tstep];
                    v[q][r] = INV ? cmulc(v[q][r], w) : cmul(v[q][r], w);
                }
            }
            dst[q] = (j / NS) * NS * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int idx = (int)threadIdx.x + q * NT;
        if ((total % NT == 0) || idx < total) {
            dft<R, INV>(v[q], tw, N / R);
#pragma unroll
            for (int r = 0; r < R; ++r) buf[base[q] + ct_pad(dst[q] + r * NS)] = v[q][r];
        }
    }
    __syncthreads();
}

template <int N, int S, int G, int NT, bool INV, typename V>
__device__ __forceinline__ void ct_stages(V* buf, const V* __restrict__ tw) {
    constexpr CtPlan P = ct_plan(N);
    static_assert(P.nf >= 0, "FFT length must factor into 2, 3, 5, 7");
    if constexpr (S < P.nf) {
        ct_stage<N, P.r[S], ct_ns(N, S), G, NT, INV, V>(buf, tw);
        ct_stages<N, S + 1, G, NT, INV, V>(buf, tw);
    }
}

// Copies the N twiddles exp(-2 pi i t/N) to LDS (stw, N entries); the
// caller's next __syncthreads publishes them.  A stage gathers tw[r k tstep]
// at up to 64 different k per wavefront: from global memory that touches
// several times as many cache lines as the tile's own data, from LDS it costs
// bank cycles only.  Same values, so the transform's results are unchanged.
template <int N, int NT, typename V>
__device__ __forceinline__ void ct_stage_twiddles(V* stw, const V* __restrict__ tw) {
    constexpr int NQ = (N + NT - 1) / NT;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int i = (int)threadIdx.x + q * NT;
        if (N % NT == 0 || i < N) stw[i] = tw[i];
    }
}

// Transforms G padded lines (pitch ct_pitch(N)) in LDS.  All NT threads call it.
// V is double2 or float2; tw holds exp(-2 pi i t/N) rounded to that type.
template <int N, int G, int NT, bool INV, typename V>
__device__ __forceinline__ void fft_ct(V* buf, const V* __restrict__ tw) {
    ct_stages<N, 0, G, NT, INV, V>(buf, tw);
}

// Maps a linear workgroup id to a logical tile id so that each of the 8 XCDs
// (which receive workgroups round-robin) works on one contiguous range of
// tiles: neighbouring tiles then share that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    constexpr int kX = 8;
    const int xcd = b % kX;
    const int per = nb / kX;
    const int rem = nb % kX;
    return xcd * per + (xcd < rem ? xcd : rem) + b / kX;
}

}  // namespace dnsp
