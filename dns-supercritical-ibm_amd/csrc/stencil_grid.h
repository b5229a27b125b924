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

// This thread's point (i, j, k), 0-based; false outside the grid (rows of
// the last band past nz, columns past nx).  Launch: band_grid(..).blocks()
// workgroups of band_block().
__device__ __forceinline__ bool band_point(const BandGrid& b, int& i, int& j, int& k) {
    const int L = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int band = L / b.per_band();
    const int r = L - band * b.per_band();
    const int tb = b.nbx * b.hb;
    k = r / tb;
    const int t = r - k * tb;
    const int bz = band * b.hb + t / b.nbx;
    const int bx = t - (t / b.nbx) * b.nbx;
    i = bx * kTileX + (int)threadIdx.x;
    j = bz * kTileZ + (int)threadIdx.y;
    return i < b.nx && j < b.nz;
}

}  // namespace dnsp
