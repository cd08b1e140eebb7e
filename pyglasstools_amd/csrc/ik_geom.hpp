// ik_geom.hpp -- geometry and launch-layout helpers shared by every kernel
// translation unit of the IK stress-field engine: the launch geometry
// (box, cell grid, search radii), minimum image, cell windows, the mapping
// of workgroups to the 8 XCDs, and wavefront sums.
#pragma once
#include <hip/hip_runtime.h>

#include "ik_device.hpp"

namespace pgt {
// Geometry shared by all kernels of one launch.
struct Geom {
    double L[3], invL[3], lo[3];
    double pinvL[3];  // invL on periodic axes, 0 on the others
    int periodic[3];
    int nc[3];     // cells per axis: 1 (non-periodic axes, short boxes) or at least 2 reach + 1
    double cs[3];  // cell edge, at least (search radius) / reach
    int ncell;
    int reach;     // cells on each side of a cell that a search covers
    double rc2;      // cg_rcut^2
    double inv_rc;   // 1/cg_rcut
    double inv_rc2;  // 1/cg_rcut^2
    double mr2;      // max_rcut^2
    // FP32 pre-tests of the 3D candidate scans: with |a|^2 from FP32 copies
    // of the positions, |a|^2 < in32 means inside the coarse-graining ball
    // and |a|^2 > out32 outside, whatever the FP64 rounding; in between the
    // FP64 test decides (in32 < 0 and out32 = inf send every candidate to
    // it).  bin32 and bout32 do the same for the bond search radius.
    float L32[3], invL32[3], pinvL32[3];
    float in32, out32;
    float bin32, bout32;
};

// Minimum image with the image index rint(d/L) taken as rint(d * (1/L)).
// The two agree except when d/L is within a few ulp of a half-integer, i.e.
// for separations of half a box, which every search here rejects (search
// radii are required to be below half a periodic box).  pinvL is 1/L on a
// periodic axis and 0 on another, where the expression is exactly d (the
// image index is rint(0) = 0 and fma(-L, 0, d) = d), so no select is needed.
__device__ __forceinline__ double minimg(double d, double L, double pinvL) { return fma(-L, rint(d * pinvL), d); }

__device__ __forceinline__ int cell_coord(const Geom &g, int a, double p)
{
    if (g.nc[a] == 1) return 0;
    double u = p - g.lo[a];
    u -= g.L[a] * floor(u * g.invL[a]);
    int k = (int)floor(u / g.cs[a]);
    return k < 0 ? 0 : (k >= g.nc[a] ? g.nc[a] - 1 : k);
}

constexpr int kMaxRows = 25;  // rows of cells in a search window, (2 reach + 1)^2 with reach <= 2

// Up to two contiguous particle ranges covering cells [cx - reach, cx + reach]
// of one row of the cell grid along x (wrapping periodically).
__device__ __forceinline__ int row_ranges(const Geom &g, const int *__restrict__ cstart, int rowbase, int cx,
                                          int *beg, int *end)
{
    const int nx = g.nc[0], R = g.reach;
    if (nx == 1) {
        beg[0] = cstart[rowbase];
        end[0] = cstart[rowbase + 1];
        return 1;
    }
    const int lo = cx - R, hi = cx + R;
    if (lo >= 0 && hi < nx) {
        beg[0] = cstart[rowbase + lo];
        end[0] = cstart[rowbase + hi + 1];
        return 1;
    }
    if (lo < 0) {  // cells [lo + nx, nx) and [0, hi]
        beg[0] = cstart[rowbase + lo + nx];
        end[0] = cstart[rowbase + nx];
        beg[1] = cstart[rowbase];
        end[1] = cstart[rowbase + hi + 1];
        return 2;
    }
    // hi >= nx: cells [lo, nx) and [0, hi - nx]
    beg[0] = cstart[rowbase + lo];
    end[0] = cstart[rowbase + nx];
    beg[1] = cstart[rowbase];
    end[1] = cstart[rowbase + hi - nx + 1];
    return 2;
}

// Base indices of the (up to kMaxRows) rows of cells within reach of cell
// (cy, cz), z outer, both ascending from -reach.
__device__ __forceinline__ int neighbour_rows(const Geom &g, int cy, int cz, int *rowbase)
{
    const int ny = g.nc[1], nz = g.nc[2], R = g.reach;
    int ys[5], zs[5], nyy = 0, nzz = 0;
    if (ny == 1) ys[nyy++] = 0;
    else
        for (int k = -R; k <= R; ++k) ys[nyy++] = (cy + k + ny) % ny;
    if (nz == 1) zs[nzz++] = 0;
    else
        for (int k = -R; k <= R; ++k) zs[nzz++] = (cz + k + nz) % nz;
    int n = 0;
    for (int a = 0; a < nzz; ++a)
        for (int b = 0; b < nyy; ++b) rowbase[n++] = (zs[a] * ny + ys[b]) * g.nc[0];
    return n;
}

// Workgroups are dealt to the 8 XCDs of the chip round-robin by linear
// workgroup id, and each XCD has its own L2.  With nx blocks per frame row
// (nx a multiple of 8, so the XCD of block x is x mod 8), block x is given
// the tile (x mod 8) * nx/8 + x / 8: each XCD then works through one
// contiguous eighth of the frame's tiles (grid points in Morton order,
// particles in cell order, so a compact piece of space) and its L2 holds
// only that piece's particles and bonds.
__device__ __forceinline__ int xcd_tile(int x, int nx)
{
    return (nx & 7) ? x : (x & 7) * (nx >> 3) + (x >> 3);
}

// Frames dealt to XCDs whole: the workgroups of frame f all run on XCD
// f mod 8 (workgroups go to the XCDs round-robin by linear id), so one
// frame's particle arrays, counters and partly written lines stay in one
// XCD's L2.  Launch with gridDim.y = frame_rows(nf); returns false for the
// padding workgroups (frames nf .. 8*ceil(nf/8)-1).
__host__ __device__ inline int frame_rows(int nf) { return (nf + 7) / 8 * 8; }

__device__ __forceinline__ bool frame_block(int nf, int &f, int &bx)
{
    const int L = blockIdx.y * gridDim.x + blockIdx.x;
    const int j = L >> 3;
    f = (L & 7) + 8 * (j / gridDim.x);
    bx = j % gridDim.x;
    return f < nf;
}

// Sum over the 64 lanes of a wavefront (all lanes must be active).
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Outputs of a field gather.
struct GatherOut {
    double *vir;
    double *kin;
    double *ela;
    // per-wavefront counts (balls, contributions), two per wavefront in
    // launch order; summed by k_reduce_stats.  One device-scope atomic per
    // wavefront on a shared counter would serialise every wavefront of the
    // launch on a single address.
    unsigned long long *wave_counts;
};

}  // namespace pgt
