// Batched Poisson solves of many independent instances of uneven sizes.
//
// Every stage of the solve is one launch per distinct transform length (x, z)
// or rows-per-chunk (tridiagonal) covering the tiles of all instances of that
// size: a workgroup decodes its task through a slot table (instance, tile)
// and runs the same tile bodies as the single-instance kernels.  A batch of
// 48 instances over 5 x lengths, 5 z lengths and 4 row counts thus takes 24
// launches per FP64 solve instead of 240.  The FP64 form and the
// mixed-precision form (sweep 0 and one FP64 residual refinement) are
// provided; the launch sequence allocates nothing and can be captured in a
// graph.  The FP64 solve also runs as one persistent cooperative launch
// (kb_persist) over all stages and instances.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <new>
#include <type_traits>
#include <vector>

#include <hip/hip_cooperative_groups.h>

#include "common.h"
#include "poisson_ct.h"
#include "poisson_ct_launch.h"
#include "poisson_mixed.h"

namespace dnsp {

// Device description of one instance.
struct BInst {
    int nx, nz, ny, mr;
    long nlines, plane;
    double ix2, iz2, pcr_min, inv_nxnz;
    const double* ax;
    const double* az;
    const double* bands;
    const double2* twx;
    const double2* twz;
    const float2* twx32;
    const float2* twz32;
    const double* f;  // bound right-hand side
    double* P;        // bound solution
    float* A;         // mixed form: FP32 fields, line sums, msum | pbar0 | pbar1 | q
    float* B;
    double* lsA;
    double* lsB;
    double* small;
};

// Launch group: the instances of one size, their task offsets.
struct BGroup {
    int key = 0;  // transform length or rows per chunk
    int cnt = 0;
    long total = 0;
    int* ids = nullptr;     // device, cnt
    long* start = nullptr;  // device, cnt
    int* slot = nullptr;    // device, total: the instance slot u of every task (start[u] <= task < start[u + 1])
    // tridiagonal groups of the mixed form: (instance slot u, mode group)
    // pairs solved in FP64 ([0]) and in FP32 ([1]) arithmetic
    int2* tl[2] = {nullptr, nullptr};
    long ntl[2] = {0, 0};
};

enum BxKind { BX_F64_FWD = 0, BX_F64_BWD, BX_MX_FWD, BX_MX_RES, BX_MX_BWD, BX_MX_ACC, BX_MX_OUT0 };

template <int N, int V, int KIND>
__global__ __launch_bounds__(ct_xvar_threads(V)) void kb_x(const BInst* __restrict__ inst,
                                                          const int* __restrict__ ids,
                                                          const long* __restrict__ start, const int* __restrict__ tslot, int sums) {
    constexpr int G = ct_lines<N, ct_xvar_tile(V)>();
    constexpr int NT = ct_xvar_threads(V);
    constexpr bool F64 = KIND == BX_F64_FWD || KIND == BX_F64_BWD;
    using VT = typename std::conditional<F64, double2, float2>::type;
    // the residual kind may stage 2G + 2 lines of its instance (mr_lds_f2)
    __shared__ __align__(16) VT sm[KIND == BX_MX_RES ? mr_lds_f2<N, G, float>() : G * ct_pitch(N)];
    __shared__ double slot[F64 ? 1 : line_slots<N, G>()];
    const long task = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int u = tslot[task];
    const BInst& b = inst[ids[u]];
    const long tile = task - start[u];
    // the instance's twiddles in LDS (the bodies' barrier before the
    // transform publishes them), as the single-solve kernels
    __shared__ VT stw[N];
    if constexpr (F64) ct_stage_twiddles<N, NT>(stw, b.twx);
    else ct_stage_twiddles<N, NT>(stw, b.twx32);
    if constexpr (KIND == BX_F64_FWD) {
        xfwd_body<N, G, NT>(b.f, b.P, b.nlines, stw, tile, sm);
    } else if constexpr (KIND == BX_F64_BWD) {
        xbwd_body<N, G, NT>(b.P, b.P, b.nlines, stw, b.inv_nxnz, tile, sm);
    } else if constexpr (KIND == BX_MX_FWD) {
        xfwd_mf_body<N, G, NT>(b.f, b.A, b.lsA, b.nlines, stw, tile, sm, slot);
    } else if constexpr (KIND == BX_MX_RES) {
        // same dispatch as kt_xfwd_mr (the instance's nz is workgroup-uniform)
        bool done = false;
        if constexpr (mr_staged<N, G, float>()) {
            if (b.nz % (2 * G) == 0) {
                xfwd_mr_staged<N, G, NT, float>(b.A, b.f, b.small + 3 * b.ny, b.bands, b.B, b.lsB, b.nz, b.ny, b.ix2,
                                                b.iz2, b.nlines, stw, tile, sm, slot);
                done = true;
            }
        }
        if (!done)
            xfwd_mr_body<N, G, NT, float>(b.A, b.f, b.small + 3 * b.ny, b.bands, b.B, b.lsB, b.nz, b.ny, b.ix2,
                                          b.iz2, b.nlines, stw, tile, sm, slot);
    } else if constexpr (KIND == BX_MX_BWD) {
        xbwd_mf_body<N, G, NT>(b.A, b.A, b.lsA, b.nlines, stw, (float)b.inv_nxnz, tile, sm, slot);
    } else if constexpr (KIND == BX_MX_ACC) {
        xbwd_ma_body<N, G, NT, float>(b.B, b.A, b.small + b.ny, b.small + 2 * b.ny, b.P, sums ? b.lsA : nullptr,
                                      b.nz, b.nlines, stw, b.inv_nxnz, tile, sm, slot);
    } else {  // BX_MX_OUT0: sweep 0 written straight to P (no refinement)
        xbwd_ma_body<N, G, NT, float>(b.A, nullptr, b.small + b.ny, nullptr, b.P, nullptr, b.nz, b.nlines,
                                      stw, b.inv_nxnz, tile, sm, slot);
    }
}

enum BzKind { BZ_F64_FWD = 0, BZ_F64_BWD, BZ_A_FWD, BZ_A_BWD, BZ_B_FWD, BZ_B_BWD };

template <int N, int V, int KIND>
__global__ __launch_bounds__(ct_zvar_threads(V)) void kb_z(const BInst* __restrict__ inst,
                                                          const int* __restrict__ ids,
                                                          const long* __restrict__ start, const int* __restrict__ tslot) {
    constexpr int TA = ct_cols<N, ct_zvar_tile(V)>();
    constexpr int NT = ct_zvar_threads(V);
    constexpr bool F64 = KIND == BZ_F64_FWD || KIND == BZ_F64_BWD;
    using T = typename std::conditional<F64, double, float>::type;
    using VT = typename Cx<T>::V;
    __shared__ VT sm[TA * ct_pitch(N)];
    const long task = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int u = tslot[task];
    const BInst& b = inst[ids[u]];
    const int tile = (int)(task - start[u]);
    __shared__ VT stw[N];
    if constexpr (F64) ct_stage_twiddles<N, NT>(stw, b.twz);
    else ct_stage_twiddles<N, NT>(stw, b.twz32);
    const int ntc = ((b.nx + 1) / 2 + TA - 1) / TA;
    if constexpr (KIND == BZ_F64_FWD) zfwd_body<N, TA, NT, double>(b.P, b.nx, ntc, stw, nullptr, nullptr, tile, sm);
    else if constexpr (KIND == BZ_F64_BWD) zbwd_body<N, TA, NT, double>(b.P, b.nx, ntc, stw, tile, sm);
    else if constexpr (KIND == BZ_A_FWD) zfwd_body<N, TA, NT, float>(b.A, b.nx, ntc, stw, b.lsA, b.small, tile, sm);
    else if constexpr (KIND == BZ_A_BWD) zbwd_body<N, TA, NT, float>(b.A, b.nx, ntc, stw, tile, sm);
    else if constexpr (KIND == BZ_B_FWD) zfwd_body<N, TA, NT, float>(b.B, b.nx, ntc, stw, b.lsB, b.small, tile, sm);
    else zbwd_body<N, TA, NT, float>(b.B, b.nx, ntc, stw, tile, sm);
}

// Tridiagonal: one mode group of 8 modes per task.  KIND 0: FP64 on P;
// 1: FP32 on A with the (0,0) mode -> pbar0; 2: FP32 on B -> pbar1.
// Variants 0..3: P = kPartChunks chunks per mode, {registers uncapped,
// capped for 4 waves} x {one-pass, two-pass}; 4 and 5: 16 chunks of twice
// the rows (one reduction step fewer; the groups' rows per chunk at 32
// chunks, doubled, are again a listed row count), one-pass and two-pass.
// (8 chunks of four times the rows measured the same as 16 on the
// 48-instance batch: not kept.)
constexpr int kBatchM = 8;
constexpr int kBatchTriVariants = 6;

template <int MR, int KIND, int OCC, bool TP, int P = kPartChunks>
__global__ __launch_bounds__(kBatchM * P, OCC) void kb_tri(const BInst* __restrict__ inst,
                                                           const int* __restrict__ ids,
                                                           const long* __restrict__ start, const int* __restrict__ tslot) {
    constexpr int M = kBatchM;
    using T = typename std::conditional<KIND == 0, double, float>::type;
    __shared__ TriRed<P, M> red;
    extern __shared__ double smb[];
    const long task = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int u = tslot[task];
    const BInst& b = inst[ids[u]];
    const long grp = task - start[u];
    const int mi = (int)threadIdx.x % M;
    const long mode = tri_group_mode(grp, (b.plane + M - 1) / M, M, mi);
    T* S = KIND == 0 ? reinterpret_cast<T*>(b.P) : reinterpret_cast<T*>(KIND == 1 ? b.A : b.B);
    const double* msum = KIND == 0 ? nullptr : b.small;
    double* pbar = KIND == 0 ? nullptr : b.small + (KIND == 1 ? 1 : 2) * b.ny;
    double d[MR];
    const double lam64 = tri_prologue<MR, P, M, T>(d, smb, b.bands + (grp == 0 ? 3 * b.ny : 0), S, b.ax, b.az, mode,
                                                   b.plane, b.nx, b.ny, msum);
    __syncthreads();
    tri_group<MR, P, M, T, TP>(d, S, lam64, smb, b.nx, b.ny, b.plane, b.pcr_min, mode, msum, pbar, b.inv_nxnz, red);
}

// The mixed form's tridiagonal from a task list: workgroup b solves mode
// group tl[b].y of instance slot tl[b].x, in arithmetic Ar.  KIND 1 or 2.
template <int MR, int KIND, int OCC, bool TP, typename Ar, int P = kPartChunks>
__global__ __launch_bounds__(kBatchM * P, OCC) void kb_tri_list(const BInst* __restrict__ inst,
                                                                const int* __restrict__ ids,
                                                                const int2* __restrict__ tl) {
    constexpr int M = kBatchM;
    __shared__ TriRed<P, M, Ar> red;
    extern __shared__ double smb_raw[];
    Ar* smb = reinterpret_cast<Ar*>(smb_raw);
    const int2 t = tl[xcd_tile((int)blockIdx.x, (int)gridDim.x)];
    const BInst& b = inst[ids[t.x]];
    const long grp = t.y;
    const int mi = (int)threadIdx.x % M;
    const long mode = tri_group_mode(grp, (b.plane + M - 1) / M, M, mi);
    float* S = KIND == 1 ? b.A : b.B;
    const double* msum = b.small;
    double* pbar = b.small + (KIND == 1 ? 1 : 2) * b.ny;
    Ar d[MR];
    const double lam64 = tri_prologue<MR, P, M, float, Ar>(d, smb, b.bands + (grp == 0 ? 3 * b.ny : 0), S, b.ax, b.az,
                                                           mode, b.plane, b.nx, b.ny, msum);
    __syncthreads();
    tri_group<MR, P, M, float, TP, Ar>(d, S, lam64, smb, b.nx, b.ny, b.plane, b.pcr_min, mode, msum, pbar,
                                       b.inv_nxnz, red);
}

// q(k) of the first refinement for every plane of every instance (one
// wavefront per plane; see k_mixed_q).
__global__ __launch_bounds__(64) void kb_q(const BInst* __restrict__ inst, const int* __restrict__ ids,
                                           const long* __restrict__ start, const int* __restrict__ tslot) {
    const long task = blockIdx.x;
    const int u = tslot[task];
    const BInst& b = inst[ids[u]];
    const int k = (int)(task - start[u]);
    const int ny = b.ny, nz = b.nz;
    const double* pbar = b.small + ny;
    const int lane = (int)threadIdx.x;
    double m[3];
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        const int kk = k - 1 + w;
        const bool in = kk >= 0 && kk < ny;
        double acc = 0.0;
        if (in)
            for (int j = lane; j < nz; j += 64) acc += b.lsA[(size_t)kk * nz + j];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        m[w] = in ? acc * b.inv_nxnz + pbar[kk] : 0.0;
    }
    if (lane == 0) {
        const double* a = b.bands;
        const double* bb = a + ny;
        const double* c = bb + ny;
        const double* a0 = c + ny;
        const double* b0 = a0 + ny;
        const double* c0 = b0 + ny;
        double v = bb[k] * pbar[k];
        if (k > 0) v += a[k] * pbar[k - 1];
        if (k < ny - 1) v += c[k] * pbar[k + 1];
        v += (b0[k] - bb[k]) * m[1];
        if (k > 0) v += (a0[k] - a[k]) * m[0];
        if (k < ny - 1) v += (c0[k] - c[k]) * m[2];
        b.small[3 * ny + k] = v;
    }
}

#define DNSP_CT_SIZES(X) X(32) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768)
#define DNSP_PART_ROWS(X) X(2) X(3) X(4) X(6) X(8) X(12) X(16) X(24) X(32)

template <int N, int V>
static void bx_launch_nv(int kind, const BInst* inst, const BGroup& g, int sums, hipStream_t st) {
    const int nb = (int)g.total;
    constexpr int NT = ct_xvar_threads(V);
    switch (kind) {
        case BX_F64_FWD: kb_x<N, V, BX_F64_FWD><<<nb, NT, 0, st>>>(inst, g.ids, g.start, g.slot, sums); break;
        case BX_F64_BWD: kb_x<N, V, BX_F64_BWD><<<nb, NT, 0, st>>>(inst, g.ids, g.start, g.slot, sums); break;
        case BX_MX_FWD: kb_x<N, V, BX_MX_FWD><<<nb, NT, 0, st>>>(inst, g.ids, g.start, g.slot, sums); break;
        case BX_MX_RES: kb_x<N, V, BX_MX_RES><<<nb, NT, 0, st>>>(inst, g.ids, g.start, g.slot, sums); break;
        case BX_MX_BWD: kb_x<N, V, BX_MX_BWD><<<nb, NT, 0, st>>>(inst, g.ids, g.start, g.slot, sums); break;
        case BX_MX_ACC: kb_x<N, V, BX_MX_ACC><<<nb, NT, 0, st>>>(inst, g.ids, g.start, g.slot, sums); break;
        default: kb_x<N, V, BX_MX_OUT0><<<nb, NT, 0, st>>>(inst, g.ids, g.start, g.slot, sums); break;
    }
}

template <int N>
static void bx_launch_n(int kind, int v, const BInst* inst, const BGroup& g, int sums, hipStream_t st) {
    switch (v) {
        case 0: bx_launch_nv<N, 0>(kind, inst, g, sums, st); break;
        case 1: bx_launch_nv<N, 1>(kind, inst, g, sums, st); break;
        case 2: bx_launch_nv<N, 2>(kind, inst, g, sums, st); break;
        default: bx_launch_nv<N, 3>(kind, inst, g, sums, st); break;
    }
}

static void bx_launch(int kind, int v, const BInst* inst, const BGroup& g, int sums, hipStream_t st) {
    if (g.total <= 0) return;
    switch (g.key) {
#define DNSP_CASE(N) case N: bx_launch_n<N>(kind, v, inst, g, sums, st); break;
        DNSP_CT_SIZES(DNSP_CASE)
#undef DNSP_CASE
        default: break;
    }
}

template <int N, int V>
static void bz_launch_nv(int kind, const BInst* inst, const BGroup& g, hipStream_t st) {
    const int nb = (int)g.total;
    constexpr int NT = ct_zvar_threads(V);
    switch (kind) {
        case BZ_F64_FWD: kb_z<N, V, BZ_F64_FWD><<<nb, NT, 0, st>>>(inst, g.ids, g.start, g.slot); break;
        case BZ_F64_BWD: kb_z<N, V, BZ_F64_BWD><<<nb, NT, 0, st>>>(inst, g.ids, g.start, g.slot); break;
        case BZ_A_FWD: kb_z<N, V, BZ_A_FWD><<<nb, NT, 0, st>>>(inst, g.ids, g.start, g.slot); break;
        case BZ_A_BWD: kb_z<N, V, BZ_A_BWD><<<nb, NT, 0, st>>>(inst, g.ids, g.start, g.slot); break;
        case BZ_B_FWD: kb_z<N, V, BZ_B_FWD><<<nb, NT, 0, st>>>(inst, g.ids, g.start, g.slot); break;
        default: kb_z<N, V, BZ_B_BWD><<<nb, NT, 0, st>>>(inst, g.ids, g.start, g.slot); break;
    }
}

template <int N>
static void bz_launch_n(int kind, int v, const BInst* inst, const BGroup& g, hipStream_t st) {
    switch (v) {
        case 0: bz_launch_nv<N, 0>(kind, inst, g, st); break;
        case 1: bz_launch_nv<N, 1>(kind, inst, g, st); break;
        case 2: bz_launch_nv<N, 2>(kind, inst, g, st); break;
        default: bz_launch_nv<N, 3>(kind, inst, g, st); break;
    }
}

static void bz_launch(int kind, int v, const BInst* inst, const BGroup& g, hipStream_t st) {
    if (g.total <= 0) return;
    switch (g.key) {
#define DNSP_CASE(N) case N: bz_launch_n<N>(kind, v, inst, g, st); break;
        DNSP_CT_SIZES(DNSP_CASE)
#undef DNSP_CASE
        default: break;
    }
}

template <int MR, int OCC, bool TP, int P = kPartChunks>
static void bt_launch_m(int kind, const BInst* inst, const BGroup& g, hipStream_t st) {
    const int nb = (int)g.total;
    constexpr int NT = kBatchM * P;
    constexpr size_t lds = 3 * sizeof(double) * MR * P;
    if (kind == 0) {
        kb_tri<MR, 0, OCC, TP, P><<<nb, NT, lds, st>>>(inst, g.ids, g.start, g.slot);
        return;
    }
    if (g.tl[0]) {
        // the mixed form's split: FP64 groups, then FP32 ones
        constexpr size_t lds32 = 3 * sizeof(float) * MR * P;
        if (g.ntl[0] > 0) {
            if (kind == 1) kb_tri_list<MR, 1, OCC, TP, double, P><<<(int)g.ntl[0], NT, lds, st>>>(inst, g.ids, g.tl[0]);
            else kb_tri_list<MR, 2, OCC, TP, double, P><<<(int)g.ntl[0], NT, lds, st>>>(inst, g.ids, g.tl[0]);
        }
        if (g.ntl[1] > 0) {
            if (kind == 1) kb_tri_list<MR, 1, OCC, TP, float, P><<<(int)g.ntl[1], NT, lds32, st>>>(inst, g.ids, g.tl[1]);
            else kb_tri_list<MR, 2, OCC, TP, float, P><<<(int)g.ntl[1], NT, lds32, st>>>(inst, g.ids, g.tl[1]);
        }
        return;
    }
    if (kind == 1) kb_tri<MR, 1, OCC, TP, P><<<nb, NT, lds, st>>>(inst, g.ids, g.start, g.slot);
    else kb_tri<MR, 2, OCC, TP, P><<<nb, NT, lds, st>>>(inst, g.ids, g.start, g.slot);
}

// 16 chunks of 2 MR rows for groups keyed by MR rows at 32 chunks (MR <= 16)
template <int MR, bool TP>
static void bt_launch_16(int kind, const BInst* inst, const BGroup& g, hipStream_t st) {
    if constexpr (2 * MR <= 32) bt_launch_m<2 * MR, 1, TP, 16>(kind, inst, g, st);
    else bt_launch_m<MR, 1, TP>(kind, inst, g, st);
}

static void bt_launch(int kind, int v, const BInst* inst, const BGroup& g, hipStream_t st) {
    if (g.total <= 0) return;
    switch (g.key) {
#define DNSP_CASE(MR)                                            \
    case MR:                                                     \
        if (v == 0) bt_launch_m<MR, 1, false>(kind, inst, g, st);      \
        else if (v == 1) bt_launch_m<MR, 4, false>(kind, inst, g, st); \
        else if (v == 2) bt_launch_m<MR, 1, true>(kind, inst, g, st);  \
        else if (v == 3) bt_launch_m<MR, 4, true>(kind, inst, g, st);  \
        else if (v == 4) bt_launch_16<MR, false>(kind, inst, g, st);   \
        else bt_launch_16<MR, true>(kind, inst, g, st);                \
        break;
        DNSP_PART_ROWS(DNSP_CASE)
#undef DNSP_CASE
        default: break;
    }
}

// ---------------------------------------------------------------------------
// The FP64 batched solve in one persistent launch.  A cooperative grid of
// kPersistThreads-thread workgroups (all resident at once) runs the five
// stages; within a stage workgroup w takes tasks w, w + grid, w + 2 grid, ...
// (w numbered by XCD as xcd_tile numbers tiles, so that each XCD holds
// consecutive tasks at every step, and tasks of every size spread over all
// workgroups), finds
// each task's launch group (transform length or rows per chunk) and
// instance, and runs the same tile body as the per-length kernels of that
// group with x and z variant kPersistXV / kPersistZV and the two-pass
// tridiagonal; a grid barrier separates the stages.  One dynamic LDS buffer,
// sized for the largest body of the batch, serves every task.
// ---------------------------------------------------------------------------
constexpr int kPersistThreads = 256;
constexpr int kPersistXV = 2, kPersistZV = 2;  // (1536, 256) tiles
constexpr int kPersistMaxRows = 12;           // rows per chunk (ny <= 12 * kPartChunks)
static_assert(ct_xvar_threads(kPersistXV) == kPersistThreads && ct_zvar_threads(kPersistZV) == kPersistThreads &&
                  kBatchM * kPartChunks == kPersistThreads,
              "one workgroup size for every stage");

// One launch group of one stage: its tasks are [first, first + the group's
// total) of the stage.
struct PSeg {
    long first;
    int key, cnt;
    const int* ids;
    const long* start;
    const int* slot;
};
struct PPlan {
    int s0[6];      // stage s's groups: seg[s0[s] .. s0[s + 1])
    long total[5];  // tasks per stage
};

// One task of each kind and size, each a function of its own (not
// inlined), so that every body keeps the register allocation it has in its
// own kernel; one inlined switch over all of them would be allocated for
// all of them at once.
template <int N, bool FWD>
__device__ __attribute__((noinline)) void px_task_n(const BInst& b, long tile, double2* sm) {
    constexpr int G = ct_lines<N, ct_xvar_tile(kPersistXV)>();
    if constexpr (FWD) xfwd_body<N, G, kPersistThreads>(b.f, b.P, b.nlines, b.twx, tile, sm);
    else xbwd_body<N, G, kPersistThreads>(b.P, b.P, b.nlines, b.twx, b.inv_nxnz, tile, sm);
}

template <int N, bool FWD>
__device__ __attribute__((noinline)) void pz_task_n(const BInst& b, int tile, double2* sm) {
    constexpr int TA = ct_cols<N, ct_zvar_tile(kPersistZV)>();
    const int ntc = ((b.nx + 1) / 2 + TA - 1) / TA;
    if constexpr (FWD) zfwd_body<N, TA, kPersistThreads, double>(b.P, b.nx, ntc, b.twz, nullptr, nullptr, tile, sm);
    else zbwd_body<N, TA, kPersistThreads, double>(b.P, b.nx, ntc, b.twz, tile, sm);
}

template <int MR>
__device__ __attribute__((noinline)) void pt_task_m(const BInst& b, long grp, char* lds) {
    constexpr int M = kBatchM, P = kPartChunks;
    TriRed<P, M>& red = *reinterpret_cast<TriRed<P, M>*>(lds);
    double* smb = reinterpret_cast<double*>(lds + sizeof(TriRed<P, M>));
    const int mi = (int)threadIdx.x % M;
    const long mode = tri_group_mode(grp, (b.plane + M - 1) / M, M, mi);
    double d[MR];
    const double lam64 = tri_prologue<MR, P, M, double>(d, smb, b.bands + (grp == 0 ? 3 * b.ny : 0), b.P, b.ax, b.az,
                                                        mode, b.plane, b.nx, b.ny, nullptr);
    __syncthreads();
    tri_group<MR, P, M, double, true>(d, b.P, lam64, smb, b.nx, b.ny, b.plane, b.pcr_min, mode, nullptr, nullptr,
                                      b.inv_nxnz, red);
}

template <bool FWD>
__device__ __forceinline__ void px_task(int n, const BInst& b, long tile, double2* sm) {
    switch (n) {
#define DNSP_CASE(N) case N: px_task_n<N, FWD>(b, tile, sm); break;
        DNSP_CT_SIZES(DNSP_CASE)
#undef DNSP_CASE
        default: break;
    }
}

template <bool FWD>
__device__ __forceinline__ void pz_task(int n, const BInst& b, int tile, double2* sm) {
    switch (n) {
#define DNSP_CASE(N) case N: pz_task_n<N, FWD>(b, tile, sm); break;
        DNSP_CT_SIZES(DNSP_CASE)
#undef DNSP_CASE
        default: break;
    }
}

__device__ __forceinline__ void pt_task(int mr, const BInst& b, long grp, char* lds) {
    switch (mr) {
        case 2: pt_task_m<2>(b, grp, lds); break;
        case 3: pt_task_m<3>(b, grp, lds); break;
        case 4: pt_task_m<4>(b, grp, lds); break;
        case 6: pt_task_m<6>(b, grp, lds); break;
        case 8: pt_task_m<8>(b, grp, lds); break;
        case 12: pt_task_m<12>(b, grp, lds); break;
        default: break;
    }
}

// LDS bytes a task of stage kind s (0, 4: x; 1, 3: z; 2: tridiagonal) with
// key n takes (host side).
inline size_t persist_lds(int s, int n) {
    if (s == 2) return sizeof(TriRed<kPartChunks, kBatchM>) + 3 * sizeof(double) * (size_t)n * kPartChunks;
    const int tile = s == 0 || s == 4 ? ct_xvar_tile(kPersistXV) : ct_zvar_tile(kPersistZV);
    const int lines = s == 0 || s == 4 ? std::max(1, tile / n) : std::min(16, std::max(1, tile / n));
    return sizeof(double2) * (size_t)lines * (size_t)ct_pitch(n);
}

// Every workgroup runs every stage and reaches every grid barrier (no early
// return); the launch is cooperative, so all workgroups are resident.
__global__ __launch_bounds__(kPersistThreads) void kb_persist(const BInst* __restrict__ inst,
                                                              const PSeg* __restrict__ seg, PPlan plan) {
    extern __shared__ __align__(16) char plds[];
    double2* sm = reinterpret_cast<double2*>(plds);
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const long lid = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const long nwg = gridDim.x;
    for (int s = 0; s < 5; ++s) {
        const long tot = plan.total[s];
        int q = plan.s0[s];
        for (long t = lid; t < tot; t += nwg) {
            while (q + 1 < plan.s0[s + 1] && seg[q + 1].first <= t) ++q;
            const PSeg& g = seg[q];
            const long task = t - g.first;
            const int u = g.slot[task];
            const BInst& b = inst[g.ids[u]];
            const long tile = task - g.start[u];
            switch (s) {
                case 0: px_task<true>(g.key, b, tile, sm); break;
                case 1: pz_task<true>(g.key, b, (int)tile, sm); break;
                case 2: pt_task(g.key, b, tile, plds); break;
                case 3: pz_task<false>(g.key, b, (int)tile, sm); break;
                default: px_task<false>(g.key, b, tile, sm); break;
            }
            __syncthreads();  // LDS is reused by the next task
        }
        if (s < 4) grid.sync();
    }
}

}  // namespace dnsp

using namespace dnsp;

struct dnsp_batch {
    int n = 0;
    int precision = 0;
    int device = 0;
    std::vector<int> nx, nz, ny, mr;
    std::vector<double> pcr_min;
    // host copies for the mixed form's FP32/FP64 split of the tridiagonal
    std::vector<std::vector<double>> hax, haz;
    std::vector<double> f32_min;  // |lambda| from which a mode runs in FP32
    BInst* d_inst = nullptr;
    std::vector<BInst> h_inst;
    std::vector<void*> allocs;  // everything hipMalloc'd, freed at the end
    // launch groups: x and z per FFT variant (tile sizes differ), tridiagonal, planes
    std::vector<BGroup> gx[kCtVariants], gz[kCtVariants], gt;
    BGroup gq;
    int max_ny = 0;
    int xvar = 0, zvar = 0, tvar = 1;     // FP64 form
    int mxvar = 0, mzvar = 0, mtvar = 1;  // mixed form
    bool bound = false;
    // persistent launch (built at its first use): group table, plan, LDS
    // bytes, grid
    PSeg* d_pseg = nullptr;
    PPlan pplan{};
    size_t p_lds = 0;
    int p_grid = 0;
};

namespace {

int dev_alloc(dnsp_batch* b, void** p, size_t bytes) {
    DNSP_HIP_TRY(hipMalloc(p, bytes));
    b->allocs.push_back(*p);
    return DNSP_OK;
}

template <typename T>
int dev_copy(dnsp_batch* b, T** p, const T* src, size_t count) {
    int rc = dev_alloc(b, reinterpret_cast<void**>(p), sizeof(T) * (count ? count : 1));
    if (rc) return rc;
    if (count) DNSP_HIP_TRY(hipMemcpy(*p, src, sizeof(T) * count, hipMemcpyHostToDevice));
    return DNSP_OK;
}

void release(dnsp_batch* b) {
    if (!b) return;
    int cur = 0;
    if (hipGetDevice(&cur) == hipSuccess && cur != b->device) (void)hipSetDevice(b->device);
    for (void* p : b->allocs) (void)hipFree(p);
    if (cur != b->device) (void)hipSetDevice(cur);
    delete b;
}

int hc_wave(int i, int n) {
    if (i <= n / 2) return i;
    const int ii = (n + 1) / 2 - 1 - (i - (n / 2 + 1));
    return ii > 1 ? ii : 1;
}

// poisson_solver_multigpu.f90:139-152
void eig(int n, double inv_d, std::vector<double>& lam) {
    const double pi = 4.0 * std::atan(1.0);
    lam.resize((size_t)n);
    for (int i = 0; i < n; ++i) {
        const int ii = hc_wave(i, n);
        const double sv = std::sin(ii * pi / n);
        lam[(size_t)i] = -4 * (inv_d * inv_d) * (sv * sv);
    }
}

template <typename V>
int twiddles(dnsp_batch* b, int n, V** out) {
    std::vector<V> tw((size_t)n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int t = 0; t < n; ++t) {
        const long double ang = 2.0L * pi * (long double)t / (long double)n;
        tw[(size_t)t].x = (decltype(tw[0].x))cosl(ang);
        tw[(size_t)t].y = (decltype(tw[0].x))(-sinl(ang));
    }
    return dev_copy(b, out, tw.data(), (size_t)n);
}

// Groups instances by key; per instance `tasks` work items.
int make_groups(dnsp_batch* b, const std::vector<int>& key, const std::vector<long>& tasks, std::vector<BGroup>& out) {
    std::map<int, std::vector<int>> by;
    for (int i = 0; i < b->n; ++i) by[key[(size_t)i]].push_back(i);
    out.clear();
    for (auto& kv : by) {
        BGroup g;
        g.key = kv.first;
        g.cnt = (int)kv.second.size();
        std::vector<long> st((size_t)g.cnt);
        long tot = 0;
        for (int u = 0; u < g.cnt; ++u) {
            st[(size_t)u] = tot;
            tot += tasks[(size_t)kv.second[(size_t)u]];
        }
        g.total = tot;
        if (tot > 0x7fffffffL) {
            set_error("batch launch group too large (%ld tasks)", tot);
            return DNSP_ERR_ARG;
        }
        // the slot of every task, one load in the kernels (a search over
        // start[] costs a chain of dependent loads per workgroup)
        std::vector<int> sl((size_t)tot);
        for (int u = 0; u < g.cnt; ++u) {
            const long end = u + 1 < g.cnt ? st[(size_t)u + 1] : tot;
            for (long t = st[(size_t)u]; t < end; ++t) sl[(size_t)t] = u;
        }
        int rc = dev_copy(b, &g.ids, kv.second.data(), kv.second.size());
        if (!rc) rc = dev_copy(b, &g.start, st.data(), st.size());
        if (!rc && tot > 0) rc = dev_copy(b, &g.slot, sl.data(), sl.size());
        if (rc) return rc;
        out.push_back(g);
    }
    return DNSP_OK;
}

int variant_tasks(int n, int v, bool xdir) {
    const int tile = xdir ? ct_xvar_tile(v) : ct_zvar_tile(v);
    if (xdir) {
        const int G = std::max(1, tile / n);
        return G;
    }
    return std::min(16, std::max(1, tile / n));
}

int setup(dnsp_batch* b, const double* bands_cat) {
    DNSP_HIP_TRY(hipGetDevice(&b->device));
    const int n = b->n;
    std::map<int, double2*> tw64;
    std::map<int, float2*> tw32;
    b->h_inst.resize((size_t)n);
    size_t boff = 0;
    const bool mixed = b->precision == DNSP_PRECISION_MIXED;
    for (int i = 0; i < n; ++i) {
        BInst& I = b->h_inst[(size_t)i];
        std::memset(&I, 0, sizeof(I));
        const int nx = b->nx[(size_t)i], nz = b->nz[(size_t)i], ny = b->ny[(size_t)i];
        I.nx = nx;
        I.nz = nz;
        I.ny = ny;
        I.mr = b->mr[(size_t)i];
        I.nlines = (long)nz * ny;
        I.plane = (long)nx * nz;
        I.inv_nxnz = 1.0 / ((double)nx * (double)nz);
        const double* bi = bands_cat + boff;
        boff += 6 * (size_t)ny;
        double mn = 1e300;
        const int k0 = ny > 2 ? 1 : 0, k1 = ny > 2 ? ny - 1 : ny;
        for (int k = k0; k < k1; ++k) mn = std::min(mn, std::fabs(bi[k]) + std::fabs(bi[2 * ny + k]));
        I.pcr_min = mn;
        // the single-instance rule (poisson_mixed.hip): FP32 sweeps where
        // |lambda| >= max_k(|a_k| + |c_k|) / 128
        double cmax = 0.0;
        for (int k = 0; k < ny; ++k) cmax = std::max(cmax, std::fabs(bi[k]) + std::fabs(bi[2 * ny + k]));
        b->f32_min.push_back(cmax / 128.0);
        int rc = dev_copy(b, const_cast<double**>(&I.bands), bi, 6 * (size_t)ny);
        if (rc) return rc;
        for (int dir = 0; dir < 2; ++dir) {
            const int len = dir == 0 ? nx : nz;
            if (!tw64.count(len)) {
                double2* t = nullptr;
                rc = twiddles(b, len, &t);
                if (rc) return rc;
                tw64[len] = t;
            }
            if (mixed && !tw32.count(len)) {
                float2* t = nullptr;
                rc = twiddles(b, len, &t);
                if (rc) return rc;
                tw32[len] = t;
            }
        }
        I.twx = tw64[nx];
        I.twz = tw64[nz];
        I.twx32 = mixed ? tw32[nx] : nullptr;
        I.twz32 = mixed ? tw32[nz] : nullptr;
        b->max_ny = std::max(b->max_ny, ny);
        if (mixed) {
            const size_t N = (size_t)nx * nz * ny;
            rc = dev_alloc(b, reinterpret_cast<void**>(&I.A), sizeof(float) * N);
            if (!rc) rc = dev_alloc(b, reinterpret_cast<void**>(&I.B), sizeof(float) * N);
            if (!rc) rc = dev_alloc(b, reinterpret_cast<void**>(&I.lsA), sizeof(double) * I.nlines);
            if (!rc) rc = dev_alloc(b, reinterpret_cast<void**>(&I.lsB), sizeof(double) * I.nlines);
            if (!rc) rc = dev_alloc(b, reinterpret_cast<void**>(&I.small), sizeof(double) * 4 * ny);
            if (rc) return rc;
        }
    }
    return DNSP_OK;
}

int setup_eig(dnsp_batch* b, const double* inv_dx, const double* inv_dz) {
    for (int i = 0; i < b->n; ++i) {
        BInst& I = b->h_inst[(size_t)i];
        std::vector<double> ax, az;
        eig(I.nx, inv_dx[i], ax);
        eig(I.nz, inv_dz[i], az);
        b->hax.push_back(ax);
        b->haz.push_back(az);
        I.ix2 = inv_dx[i] * inv_dx[i];
        I.iz2 = inv_dz[i] * inv_dz[i];
        int rc = dev_copy(b, const_cast<double**>(&I.ax), ax.data(), ax.size());
        if (!rc) rc = dev_copy(b, const_cast<double**>(&I.az), az.data(), az.size());
        if (rc) return rc;
    }
    return DNSP_OK;
}

int setup_groups(dnsp_batch* b) {
    const int n = b->n;
    for (int v = 0; v < kCtVariants; ++v) {
        std::vector<long> tx((size_t)n), tz((size_t)n);
        for (int i = 0; i < n; ++i) {
            const BInst& I = b->h_inst[(size_t)i];
            const int G = variant_tasks(I.nx, v, true);
            tx[(size_t)i] = (I.nlines + 2 * G - 1) / (2 * G);
            const int TA = variant_tasks(I.nz, v, false);
            const long ntc = ((I.nx + 1) / 2 + TA - 1) / TA;
            tz[(size_t)i] = ntc * I.ny;
        }
        int rc = make_groups(b, b->nx, tx, b->gx[v]);
        if (!rc) rc = make_groups(b, b->nz, tz, b->gz[v]);
        if (rc) return rc;
    }
    std::vector<long> tt((size_t)n), tq((size_t)n);
    std::vector<int> zero((size_t)n, 0);
    for (int i = 0; i < n; ++i) {
        const BInst& I = b->h_inst[(size_t)i];
        tt[(size_t)i] = (I.plane + kBatchM - 1) / kBatchM + 1;  // plus the group of modes 1..M-1
        tq[(size_t)i] = I.ny;
    }
    int rc = make_groups(b, b->mr, tt, b->gt);
    if (rc) return rc;
    if (b->precision == DNSP_PRECISION_MIXED) {
        // per row-count launch group: (slot, mode group) pairs with a mode of
        // |lambda| < f32_min (FP64; mode 0 and modes 1..M-1 among them, so
        // tl[0] is never empty) and the rest (FP32); slots follow
        // make_groups' ascending instance order
        for (BGroup& g : b->gt) {
            std::vector<int2> tl[2];
            int u = 0;
            for (int i = 0; i < n; ++i) {
                if (b->mr[(size_t)i] != g.key) continue;
                const BInst& I = b->h_inst[(size_t)i];
                const long ng = (I.plane + kBatchM - 1) / kBatchM;
                tl[0].push_back(make_int2(u, 0));
                tl[0].push_back(make_int2(u, (int)ng));
                for (long q = 1; q < ng; ++q) {
                    bool f32 = true;
                    for (long m = q * kBatchM; m < std::min(I.plane, (q + 1) * kBatchM) && f32; ++m)
                        f32 = std::fabs(b->hax[(size_t)i][(size_t)(m % I.nx)] + b->haz[(size_t)i][(size_t)(m / I.nx)]) >=
                              b->f32_min[(size_t)i];
                    tl[f32 ? 1 : 0].push_back(make_int2(u, (int)q));
                }
                ++u;
            }
            for (int w = 0; w < 2; ++w) {
                g.ntl[w] = (long)tl[w].size();
                if (!tl[w].empty()) {
                    rc = dev_copy(b, &g.tl[w], tl[w].data(), tl[w].size());
                    if (rc) return rc;
                }
            }
        }
    }
    std::vector<BGroup> gq;
    rc = make_groups(b, zero, tq, gq);
    if (rc) return rc;
    b->gq = gq.empty() ? BGroup() : gq[0];
    return DNSP_OK;
}

int upload_inst(dnsp_batch* b) {
    if (!b->d_inst) {
        int rc = dev_alloc(b, reinterpret_cast<void**>(&b->d_inst), sizeof(BInst) * (size_t)b->n);
        if (rc) return rc;
    }
    DNSP_HIP_TRY(hipMemcpy(b->d_inst, b->h_inst.data(), sizeof(BInst) * (size_t)b->n, hipMemcpyHostToDevice));
    return DNSP_OK;
}


void run_fp64(dnsp_batch* b, int xv, int zv, int tv, hipStream_t st) {
    for (const BGroup& g : b->gx[xv]) bx_launch(BX_F64_FWD, xv, b->d_inst, g, 0, st);
    for (const BGroup& g : b->gz[zv]) bz_launch(BZ_F64_FWD, zv, b->d_inst, g, st);
    for (const BGroup& g : b->gt) bt_launch(0, tv, b->d_inst, g, st);
    for (const BGroup& g : b->gz[zv]) bz_launch(BZ_F64_BWD, zv, b->d_inst, g, st);
    for (const BGroup& g : b->gx[xv]) bx_launch(BX_F64_BWD, xv, b->d_inst, g, 0, st);
}

void run_mixed(dnsp_batch* b, int refinements, int xv, int zv, int tv, hipStream_t st) {
    for (const BGroup& g : b->gx[xv]) bx_launch(BX_MX_FWD, xv, b->d_inst, g, 0, st);
    for (const BGroup& g : b->gz[zv]) bz_launch(BZ_A_FWD, zv, b->d_inst, g, st);
    for (const BGroup& g : b->gt) bt_launch(1, tv, b->d_inst, g, st);
    for (const BGroup& g : b->gz[zv]) bz_launch(BZ_A_BWD, zv, b->d_inst, g, st);
    if (refinements == 0) {
        for (const BGroup& g : b->gx[xv]) bx_launch(BX_MX_OUT0, xv, b->d_inst, g, 0, st);
        return;
    }
    for (const BGroup& g : b->gx[xv]) bx_launch(BX_MX_BWD, xv, b->d_inst, g, 0, st);
    if (b->gq.total > 0) kb_q<<<(int)b->gq.total, 64, 0, st>>>(b->d_inst, b->gq.ids, b->gq.start, b->gq.slot);
    for (const BGroup& g : b->gx[xv]) bx_launch(BX_MX_RES, xv, b->d_inst, g, 0, st);
    for (const BGroup& g : b->gz[zv]) bz_launch(BZ_B_FWD, zv, b->d_inst, g, st);
    for (const BGroup& g : b->gt) bt_launch(2, tv, b->d_inst, g, st);
    for (const BGroup& g : b->gz[zv]) bz_launch(BZ_B_BWD, zv, b->d_inst, g, st);
    for (const BGroup& g : b->gx[xv]) bx_launch(BX_MX_ACC, xv, b->d_inst, g, 0, st);
}

// The persistent launch's group table and grid, at its first use.
int persist_setup(dnsp_batch* b) {
    if (b->d_pseg) return DNSP_OK;
    const std::vector<BGroup>* st[5] = {&b->gx[kPersistXV], &b->gz[kPersistZV], &b->gt, &b->gz[kPersistZV],
                                        &b->gx[kPersistXV]};
    std::vector<PSeg> segs;
    PPlan plan{};
    size_t lds = 0;
    for (int s = 0; s < 5; ++s) {
        plan.s0[s] = (int)segs.size();
        long first = 0;
        for (const BGroup& g : *st[s]) {
            if (g.total <= 0) continue;
            if (s == 2 && g.key > kPersistMaxRows) {
                set_error("persistent batched solve: %d rows per chunk (ny > %d) not supported", g.key,
                          kPersistMaxRows * kPartChunks);
                return DNSP_ERR_UNSUPPORTED;
            }
            segs.push_back(PSeg{first, g.key, g.cnt, g.ids, g.start, g.slot});
            first += g.total;
            lds = std::max(lds, persist_lds(s, g.key));
        }
        plan.total[s] = first;
    }
    plan.s0[5] = (int)segs.size();
    int coop = 0, ncu = 0, per_cu = 0;
    DNSP_HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, b->device));
    if (!coop) {
        set_error("device %d does not support cooperative launches", b->device);
        return DNSP_ERR_UNSUPPORTED;
    }
    DNSP_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device));
    DNSP_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kb_persist),
                                                              kPersistThreads, lds));
    if (per_cu < 1) {
        set_error("persistent batched solve: no workgroup fits a CU (%zu bytes of LDS)", lds);
        return DNSP_ERR_UNSUPPORTED;
    }
    int rc = dev_copy(b, &b->d_pseg, segs.data(), segs.size());
    if (rc) return rc;
    b->pplan = plan;
    b->p_lds = lds;
    b->p_grid = per_cu * ncu;
    return DNSP_OK;
}

int check_bound(const dnsp_batch* b) {
    DNSP_ARG_CHECK(b, "NULL batch");
    DNSP_ARG_CHECK(b->bound, "no arrays bound to the batch (dnsp_batch_bind)");
    return DNSP_OK;
}

}  // namespace

extern "C" {

int dnsp_batch_create(int n, const int* nx, const int* nz, const int* ny, const double* inv_dx, const double* inv_dz,
                      const double* bands, int precision, dnsp_batch** out) {
    clear_error();
    DNSP_ARG_CHECK(out, "out is NULL");
    *out = nullptr;
    DNSP_ARG_CHECK(n >= 1 && n <= (1 << 20), "n=%d instances", n);
    DNSP_ARG_CHECK(nx && nz && ny && inv_dx && inv_dz && bands, "NULL argument");
    DNSP_ARG_CHECK(precision == DNSP_PRECISION_FP64 || precision == DNSP_PRECISION_MIXED, "precision=%d",
                   precision);
    size_t nb = 0;
    for (int i = 0; i < n; ++i) {
        DNSP_ARG_CHECK(ct_fft_supported(nx[i]) && ct_fft_supported(nz[i]),
                       "instance %d: nx=%d, nz=%d must be in {32,64,96,128,192,256,384,512,768}", i, nx[i], nz[i]);
        DNSP_ARG_CHECK(ny[i] >= 2 && part_rows_for(ny[i]) > 0, "instance %d: ny=%d not in [2, %d]", i, ny[i],
                       32 * kPartChunks);
        DNSP_ARG_CHECK(std::isfinite(inv_dx[i]) && std::isfinite(inv_dz[i]), "instance %d: spacing not finite", i);
        nb += 6 * (size_t)ny[i];
    }
    for (size_t e = 0; e < nb; ++e) DNSP_ARG_CHECK(std::isfinite(bands[e]), "band value %zu is not finite", e);
    dnsp_batch* b = new (std::nothrow) dnsp_batch();
    if (!b) {
        set_error("host allocation failed");
        return DNSP_ERR_NOMEM;
    }
    b->n = n;
    b->precision = precision;
    b->nx.assign(nx, nx + n);
    b->nz.assign(nz, nz + n);
    b->ny.assign(ny, ny + n);
    b->mr.resize((size_t)n);
    for (int i = 0; i < n; ++i) b->mr[(size_t)i] = part_rows_for(ny[i]);
    int rc = setup(b, bands);
    if (!rc) rc = setup_eig(b, inv_dx, inv_dz);
    if (!rc) rc = setup_groups(b);
    if (!rc) rc = upload_inst(b);
    if (rc) {
        release(b);
        return rc;
    }
    *out = b;
    return DNSP_OK;
}

int dnsp_batch_bind(dnsp_batch* b, const double* const* rhs, double* const* P) {
    clear_error();
    DNSP_ARG_CHECK(b && rhs && P, "NULL argument");
    for (int i = 0; i < b->n; ++i) {
        DNSP_ARG_CHECK(rhs[i] && P[i], "instance %d: NULL array", i);
        DNSP_ARG_CHECK(rhs[i] != P[i], "instance %d: rhs and P must be distinct arrays", i);
        // the x kernels move (re, im) pairs as 16-byte accesses
        DNSP_ARG_CHECK(((reinterpret_cast<uintptr_t>(rhs[i]) | reinterpret_cast<uintptr_t>(P[i])) & 15) == 0,
                       "instance %d: rhs and P must be 16-byte aligned", i);
    }
    for (int i = 0; i < b->n; ++i) {
        b->h_inst[(size_t)i].f = rhs[i];
        b->h_inst[(size_t)i].P = P[i];
    }
    int rc = upload_inst(b);
    if (rc) return rc;
    b->bound = true;
    return DNSP_OK;
}

int dnsp_batch_run(dnsp_batch* b, void* stream) {
    clear_error();
    int rc = check_bound(b);
    if (rc) return rc;
    run_fp64(b, b->xvar, b->zvar, b->tvar, as_stream(stream));
    DNSP_LAUNCH_CHECK();
    return DNSP_OK;
}

int dnsp_batch_run_persistent(dnsp_batch* b, void* stream) {
    clear_error();
    int rc = check_bound(b);
    if (!rc) rc = persist_setup(b);  // on the caller's current device, as dnsp_batch_run
    if (rc) return rc;
    const BInst* inst = b->d_inst;
    const PSeg* seg = b->d_pseg;
    PPlan plan = b->pplan;
    void* args[] = {&inst, &seg, &plan};
    DNSP_HIP_TRY(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kb_persist), dim3(b->p_grid),
                                            dim3(kPersistThreads), args, (unsigned)b->p_lds, as_stream(stream)));
    return DNSP_OK;
}

int dnsp_batch_run_mixed(dnsp_batch* b, int refinements, void* stream) {
    clear_error();
    int rc = check_bound(b);
    if (rc) return rc;
    DNSP_ARG_CHECK(b->precision == DNSP_PRECISION_MIXED, "batch was created with precision %d", b->precision);
    DNSP_ARG_CHECK(refinements == 0 || refinements == 1, "refinements=%d: the batched form does 0 or 1",
                   refinements);
    run_mixed(b, refinements, b->mxvar, b->mzvar, b->mtvar, as_stream(stream));
    DNSP_LAUNCH_CHECK();
    return DNSP_OK;
}

int dnsp_batch_set_variants(dnsp_batch* b, int xvar, int zvar, int tvar, int mxvar, int mzvar, int mtvar) {
    clear_error();
    DNSP_ARG_CHECK(b, "NULL batch");
    for (int v : {xvar, zvar, mxvar, mzvar}) DNSP_ARG_CHECK(v >= 0 && v < kCtVariants, "FFT variant %d", v);
    for (int v : {tvar, mtvar}) DNSP_ARG_CHECK(v >= 0 && v < kBatchTriVariants, "tridiagonal variant %d", v);
    b->xvar = xvar;
    b->zvar = zvar;
    b->tvar = tvar;
    b->mxvar = mxvar;
    b->mzvar = mzvar;
    b->mtvar = mtvar;
    return DNSP_OK;
}

int dnsp_batch_get_variants(const dnsp_batch* b, int* v6) {
    clear_error();
    DNSP_ARG_CHECK(b && v6, "NULL argument");
    const int v[6] = {b->xvar, b->zvar, b->tvar, b->mxvar, b->mzvar, b->mtvar};
    std::memcpy(v6, v, sizeof(v));
    return DNSP_OK;
}

// Tunes the variants of the whole batched solve: each FP64 (and, for a mixed
// batch, mixed) solve is timed with one knob changed at a time, x then z then
// tridiagonal; overwrites the bound solution arrays.  times_ms (optional,
// DNSP_TUNE_SLOTS): ms per solve, FP64 x, z, tridiagonal variants (at the
// slots of dnsp_poisson_autotune), then the mixed ones; -1 where not
// applicable.
int dnsp_batch_autotune(dnsp_batch* b, int reps, void* stream, double* times_ms) {
    clear_error();
    int rc = check_bound(b);
    if (rc) return rc;
    DNSP_ARG_CHECK(reps >= 1 && reps <= 1000, "reps=%d", reps);
    hipStream_t st = as_stream(stream);
    if (times_ms)
        for (int i = 0; i < DNSP_TUNE_SLOTS; ++i) times_ms[i] = -1.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    DNSP_HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_error("event creation failed");
        return DNSP_ERR_HIP;
    }
    for (int form = 0; form < 2 && rc == DNSP_OK; ++form) {
        if (form == 1 && b->precision != DNSP_PRECISION_MIXED) break;
        int* var[3] = {form ? &b->mxvar : &b->xvar, form ? &b->mzvar : &b->zvar, form ? &b->mtvar : &b->tvar};
        for (int knob = 0; knob < 3 && rc == DNSP_OK; ++knob) {
            const int nv = knob == 2 ? kBatchTriVariants : kCtVariants;
            double best = 1e30;
            int pick = *var[knob];
            for (int v = 0; v < nv; ++v) {
                *var[knob] = v;
                auto once = [&]() {
                    if (form == 0) run_fp64(b, b->xvar, b->zvar, b->tvar, st);
                    else run_mixed(b, 1, b->mxvar, b->mzvar, b->mtvar, st);
                };
                once();
                (void)hipEventRecord(e0, st);
                for (int r = 0; r < reps; ++r) once();
                (void)hipEventRecord(e1, st);
                if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                    set_error("batch autotune: launch or synchronize failed");
                    rc = DNSP_ERR_HIP;
                    break;
                }
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                const double per = (double)ms / reps;
                if (times_ms) times_ms[form * kTuneFormSlots + kTuneSlot[knob] + v] = per;
                if (per < best) {
                    best = per;
                    pick = v;
                }
            }
            *var[knob] = pick;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int dnsp_batch_size(const dnsp_batch* b, int* n) {
    clear_error();
    DNSP_ARG_CHECK(b && n, "NULL argument");
    *n = b->n;
    return DNSP_OK;
}

int dnsp_batch_free(dnsp_batch* b) {
    clear_error();
    release(b);
    return DNSP_OK;
}

}  // extern "C"
