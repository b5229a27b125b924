// In-LDS batched complex FFT engine (FP64) for gfx950: radices 2, 3, 4, 5, 7,
// 8, 11, 13.
//
// A workgroup holds G complex lines of length n in LDS, line g at
// buf[g*n .. g*n+n).  The engine runs a mixed-radix Stockham autosort
// transform: every stage reads each thread's butterfly inputs into
// registers, waits at a barrier, then writes the butterfly outputs back
// into the same LDS buffer, so one buffer suffices.  Twiddles come from a
// per-length table tw[t] = exp(-2*pi*i*t/n) computed on the host in
// extended precision.
//
// Stage s with radix R and Ns = product of the earlier radices handles
// butterfly j (0 <= j < n/R, k = j mod Ns):
//   v_r = x[j + r*n/R] * w_{Ns*R}^{r*k},  V = DFT_R(v),
//   y[(j/Ns)*Ns*R + k + q*Ns] = V_q.
#pragma once

#include <hip/hip_runtime.h>

namespace dnsp {

constexpr int kFftMaxFactors = 24;
// Complex values one thread keeps in registers per stage.
constexpr int kFftRegBudget = 24;

struct FftPlan {
    int n;
    int nf;
    int fac[kFftMaxFactors];
    const double2* tw;  // device table, n entries
};

// Complex helpers and radix kernels for T = float or double.
template <typename T>
struct Cx;
template <>
struct Cx<double> {
    using V = double2;
    __device__ static __forceinline__ V mk(double a, double b) { return make_double2(a, b); }
};
template <>
struct Cx<float> {
    using V = float2;
    __device__ static __forceinline__ V mk(float a, float b) { return make_float2(a, b); }
};

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// Multiplication by -i (forward) or +i (inverse).
template <bool INV, typename V>
__device__ __forceinline__ V rot(V a) {
    V r;
    if (INV) { r.x = -a.y; r.y = a.x; }
    else     { r.x = a.y;  r.y = -a.x; }
    return r;
}

template <bool INV, typename V>
__device__ __forceinline__ void dft2(V* v) {
    V a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <bool INV, typename V>
__device__ __forceinline__ void dft3(V* v) {
    using T = decltype(v[0].x);
    const T c = T(-0.5);
    const T s = INV ? T(0.86602540378443864676) : T(-0.86602540378443864676);
    V t = cadd(v[1], v[2]);
    V d = csub(v[1], v[2]);
    V m, isd;
    m.x = v[0].x + c * t.x; m.y = v[0].y + c * t.y;
    isd.x = -s * d.y; isd.y = s * d.x;
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, isd);
    v[2] = csub(m, isd);
}

template <bool INV, typename V>
__device__ __forceinline__ void dft4(V* v) {
    V t0 = cadd(v[0], v[2]);
    V t1 = csub(v[0], v[2]);
    V t2 = cadd(v[1], v[3]);
    V t3 = rot<INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}

template <bool INV, typename V>
__device__ __forceinline__ void dft5(V* v) {
    using T = decltype(v[0].x);
    const T c1 = T(0.30901699437494742410);   // cos(2pi/5)
    const T c2 = T(-0.80901699437494742410);  // cos(4pi/5)
    const T s1 = T(0.95105651629515357212);   // sin(2pi/5)
    const T s2 = T(0.58778525229247312917);   // sin(4pi/5)
    V t1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
    V t2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
    V m1, m2, e1, e2;
    m1.x = v[0].x + c1 * t1.x + c2 * t2.x; m1.y = v[0].y + c1 * t1.y + c2 * t2.y;
    m2.x = v[0].x + c2 * t1.x + c1 * t2.x; m2.y = v[0].y + c2 * t1.y + c1 * t2.y;
    e1.x = s1 * d1.x + s2 * d2.x; e1.y = s1 * d1.y + s2 * d2.y;
    e2.x = s2 * d1.x - s1 * d2.x; e2.y = s2 * d1.y - s1 * d2.y;
    e1 = rot<INV>(e1);
    e2 = rot<INV>(e2);
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(m1, e1);
    v[4] = csub(m1, e1);
    v[2] = cadd(m2, e2);
    v[3] = csub(m2, e2);
}

template <bool INV, typename V>
__device__ __forceinline__ void dft8(V* v) {
    using T = decltype(v[0].x);
    const T h = T(0.70710678118654752440);  // sqrt(1/2)
    V e[4] = {v[0], v[2], v[4], v[6]};
    V o[4] = {v[1], v[3], v[5], v[7]};
    dft4<INV>(e);
    dft4<INV>(o);
    // o1 *= w8, o2 *= w8^2 (= -i fwd / +i inv), o3 *= w8^3
    V o1 = o[1], o3 = o[3];
    if (!INV) {
        o[1].x = h * (o1.x + o1.y); o[1].y = h * (o1.y - o1.x);
        o[3].x = h * (o3.y - o3.x); o[3].y = -h * (o3.x + o3.y);
    } else {
        o[1].x = h * (o1.x - o1.y); o[1].y = h * (o1.x + o1.y);
        o[3].x = -h * (o3.x + o3.y); o[3].y = h * (o3.x - o3.y);
    }
    o[2] = rot<INV>(o[2]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[q] = cadd(e[q], o[q]);
        v[q + 4] = csub(e[q], o[q]);
    }
}

// Any radix, twiddles read from the length-n table (rs = n/R).
template <int R, bool INV, typename V>
__device__ __forceinline__ void dft_generic(V* v, const V* __restrict__ tw, int rs) {
    V o[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        V s = v[0];
#pragma unroll
        for (int r = 1; r < R; ++r) {
            const V w = tw[((q * r) % R) * rs];
            s = cadd(s, INV ? cmulc(v[r], w) : cmul(v[r], w));
        }
        o[q] = s;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = o[q];
}

template <int R, bool INV, typename V>
__device__ __forceinline__ void dft(V* v, const V* __restrict__ tw, int rs) {
    if constexpr (R == 2) dft2<INV>(v);
    else if constexpr (R == 3) dft3<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v);
    else if constexpr (R == 5) dft5<INV>(v);
    else if constexpr (R == 8) dft8<INV>(v);
    else dft_generic<R, INV>(v, tw, rs);
}

template <int R>
__host__ __device__ constexpr int fft_nb() {
    return (kFftRegBudget / R) > 0 ? (kFftRegBudget / R) : 1;
}

template <int R, bool INV>
__device__ __forceinline__ void fft_stage(double2* buf, int G, int n, int Ns,
                                          const double2* __restrict__ tw) {
    constexpr int NB = fft_nb<R>();
    const int m = n / R;
    const int tstep = n / (Ns * R);
    const int total = G * m;
    const int nthr = blockDim.x;
    double2 v[NB][R];
    int dst[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int idx = threadIdx.x + q * nthr;
        dst[q] = -1;
        if (idx < total) {
            const int g = idx / m;
            const int j = idx - g * m;
            const int k = j % Ns;
            const double2* src = buf + g * n + j;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = src[r * m];
            if (k != 0) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const double2 w = tw[r * k * tstep];
                    v[q][r] = INV ? cmulc(v[q][r], w) : cmul(v[q][r], w);
                }
            }
            dst[q] = g * n + (j / Ns) * Ns * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        if (dst[q] >= 0) {
            dft<R, INV>(v[q], tw, n / R);
#pragma unroll
            for (int r = 0; r < R; ++r) buf[dst[q] + r * Ns] = v[q][r];
        }
    }
    __syncthreads();
}

// Transforms the G lines in buf (unnormalised; INV uses exp(+2 pi i ...)).
// Every thread of the block must call it.  The caller guarantees, for each
// radix R of the plan, G*n <= fft_nb<R>()*R*blockDim.x.
template <bool INV>
__device__ void fft_lds(double2* buf, int G, const FftPlan& pl) {
    int Ns = 1;
    for (int s = 0; s < pl.nf; ++s) {
        const int R = pl.fac[s];
        switch (R) {
            case 2: fft_stage<2, INV>(buf, G, pl.n, Ns, pl.tw); break;
            case 3: fft_stage<3, INV>(buf, G, pl.n, Ns, pl.tw); break;
            case 4: fft_stage<4, INV>(buf, G, pl.n, Ns, pl.tw); break;
            case 5: fft_stage<5, INV>(buf, G, pl.n, Ns, pl.tw); break;
            case 7: fft_stage<7, INV>(buf, G, pl.n, Ns, pl.tw); break;
            case 11: fft_stage<11, INV>(buf, G, pl.n, Ns, pl.tw); break;
            case 13: fft_stage<13, INV>(buf, G, pl.n, Ns, pl.tw); break;
            case 8: fft_stage<8, INV>(buf, G, pl.n, Ns, pl.tw); break;
            default: break;  // rejected by the host planner
        }
        Ns *= R;
    }
}

}  // namespace dnsp
