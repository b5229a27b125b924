// Block order of the point-wise and stencil kernels on (nx, nz, ny) grids.
//
// Blocks of 64 x 4 threads cover 64 x 4 (i, j) points of one plane k.  The
// tile rows of a plane are split into bands, one per XCD (each XCD has its
// own L2 and receives every 8th workgroup), and an XCD walks its band plane
// by plane: the blocks it holds at once cover a few consecutive planes of
// one band, so the k +- 1 planes and the halo rows a stencil reads are
// shared in its L2.  Dispatching plane by plane instead spreads each plane's
// tiles over all XCDs (the predictor's momentum kernel then read 177 B/pt
// from beyond L2 against 80 algorithmic).  The point of a block comes from
// a few workgroup-uniform 32-bit divisions, not per-point 64-bit ones.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_ct.h"

namespace dnsp {

constexpr int kTileX = 64, kTileZ = 4;

struct BandGrid {
    int nx, nz, nbx, hb, nbands, ny;
    __host__ __device__ int per_band() const { return nbx * hb * ny; }
    __host__ __device__ int blocks() const { return per_band() * nbands; }
};

// Grid of nx x nz x ny points (a padded grid passes its padded sizes).
inline BandGrid band_grid(int nx, int nz, int ny) {
    BandGrid b;
    b.nx = nx;
    b.nz = nz;
    b.ny = ny;
    b.nbx = (nx + kTileX - 1) / kTileX;
    const int nbz = (nz + kTileZ - 1) / kTileZ;
    b.hb = (nbz + 7) / 8;
    b.nbands = (nbz + b.hb - 1) / b.hb;
    return b;
}

inline dim3 band_block() { return dim3(kTileX, kTileZ, 1); }

// This workgroup's tile (bx, bz) and plane (or run of planes) k.  Launch:
// band_grid(..).blocks() workgroups of band_block().
__device__ __forceinline__ void band_tile(const BandGrid& b, int& bx, int& bz, int& k) {
    const int L = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int band = L / b.per_band();
    const int r = L - band * b.per_band();
    const int tb = b.nbx * b.hb;
    k = r / tb;
    const int t = r - k * tb;
    bz = band * b.hb + t / b.nbx;
    bx = t - (t / b.nbx) * b.nbx;
}

// This thread's point (i, j, k), 0-based; false outside the grid (rows of
// the last band past nz, columns past nx).
__device__ __forceinline__ bool band_point(const BandGrid& b, int& i, int& j, int& k) {
    int bx, bz;
    band_tile(b, bx, bz, k);
    i = bx * kTileX + (int)threadIdx.x;
    j = bz * kTileZ + (int)threadIdx.y;
    return i < b.nx && j < b.nz;
}

// Ring of three planes of NF halo fields in LDS, for a workgroup of
// kTileX x kTileZ threads that owns one (i, j) tile and marches through a run
// of planes in k.  A slot holds one plane of the tile plus one halo point on
// each side in x and z (kRingX x kRingZ values, point (i, j) of the tile at
// (j + 1) * kRingX + i + 1); plane kp (-1..ny) lives in slot (kp + 3) % 3.
// Each thread stages halo-tile points tid and tid + 256: fetch(kp) loads
// plane kp into registers, put(kp) stores them to the slot, so the loads of
// the next plane can be issued before a step's arithmetic.  Fields are the
// reference's halo arrays ((nx+2)(nz+2)(ny+2), x fastest).
constexpr int kRingX = kTileX + 2, kRingZ = kTileZ + 2, kRingP = kRingX * kRingZ;
constexpr int kRunK = 16;  // planes per workgroup

// Band order over tiles and runs of kRunK planes (an XCD walks the tiles of
// its band one run at a time, so neighbouring tiles' halo rows meet in its
// L2): launch ring_grid(nx, nz, ny).blocks() workgroups of band_block().
inline BandGrid ring_grid(int nx, int nz, int ny) { return band_grid(nx, nz, (ny + kRunK - 1) / kRunK); }

// This workgroup's tile origin (i0, j0) and planes [k0, k1); false for a
// workgroup past the last band's rows (uniform, before any barrier).
__device__ __forceinline__ bool ring_tile(const BandGrid& b, int ny, int& i0, int& j0, int& k0, int& k1) {
    int bx, bz, run;
    band_tile(b, bx, bz, run);
    i0 = bx * kTileX;
    j0 = bz * kTileZ;
    k0 = run * kRunK;
    k1 = k0 + kRunK < ny ? k0 + kRunK : ny;
    return j0 < b.nz;
}

template <int NF>
struct PlaneRing {
    static constexpr int kNT = kTileX * kTileZ;
    static_assert(kRingP <= 2 * kNT, "two staged points per thread");
    double (*lds)[3][kRingP];
    const double* f[NF];
    long go[2], sy;
    bool ok[2];
    double r[NF][2];
    int tid;

    // Tile at (i0, j0) of an nx x nz x ny grid; lds points to NF x 3 slots.
    __device__ __forceinline__ PlaneRing(double (*lds_)[3][kRingP], const double* const* fields, int i0, int j0,
                                         int nx, int nz)
        : lds(lds_) {
        for (int q = 0; q < NF; ++q) f[q] = fields[q];
        tid = (int)threadIdx.y * kTileX + (int)threadIdx.x;
        const long sz = nx + 2;
        sy = sz * (nz + 2);
        for (int u = 0; u < 2; ++u) {
            const int e = tid + u * kNT;
            const int rr = e / kRingX, qq = e - rr * kRingX;
            const int gi = i0 - 1 + qq, gj = j0 - 1 + rr;
            ok[u] = e < kRingP && gi <= nx && gj <= nz;
            go[u] = ok[u] ? (long)(gi + 1) + sz * (long)(gj + 1) + sy : 0;  // plane 0
        }
    }
    __device__ __forceinline__ void fetch(int kp) {
        for (int u = 0; u < 2; ++u)
            if (ok[u])
                for (int q = 0; q < NF; ++q) r[q][u] = f[q][go[u] + (long)kp * sy];
    }
    __device__ __forceinline__ void put(int kp) {
        const int slot = (kp + 3) % 3;
        for (int u = 0; u < 2; ++u)
            if (ok[u])
                for (int q = 0; q < NF; ++q) lds[q][slot][tid + u * kNT] = r[q][u];
    }
};

// This thread's point in a ring slot.
__device__ __forceinline__ int ring_point() { return ((int)threadIdx.y + 1) * kRingX + (int)threadIdx.x + 1; }

}  // namespace dnsp
