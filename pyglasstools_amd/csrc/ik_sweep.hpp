// ik_sweep.hpp -- the sweep: the frames of several systems, all 2D or all 3D
// (uneven particle counts, each with its own box, grid, cell geometry and
// pair potential), computed with one launch per stage over all of them,
// instead of one launch per stage and system.  The kernels' bodies are those of the
// single-system launches; only the way a workgroup finds its system, frame
// and block differs.
#pragma once
#include <hip/hip_runtime.h>

#include "ik_device.hpp"
#include "ik_geom.hpp"

namespace pgt {

constexpr int kSweepMaxSystems = 16;

// The stages with per-system workgroup ranges: radix-sort histogram and
// scatter (one chunk of kSortChunk particles per workgroup), the sorted
// arrays (256 particles), the bond list (256 particles) and the field
// gather (kTilePoints grid points).
enum SweepStage { SW_SORT = 0, SW_FINISH = 1, SW_BONDS = 2, SW_GATHER = 3, SW_NSTAGE = 4 };

// One system of a sweep.  The work arrays stack the systems' frames: frame f
// of this system starts at particle p0 + f n (keys, values, sorted arrays,
// bond counts, bond entries at (p0 + f n) cap entries), at cell start
// c0 + f (ncell + 1) and at histogram entry h0 + f radix nchunk; its fields
// start at value f G 4 of the system's own output arrays.  In stage s the
// system's workgroups are blk0[s] .. blk0[s] + nblk[s] - 1, bpf[s] per frame;
// blk0[s] is a multiple of 8, so that a workgroup's XCD (its linear id mod 8)
// is its index inside the system's range mod 8.
struct SweepSys {
    Geom g;
    PotParams P;
    double cgnorm;
    int n, G, nf, nchunk;
    long long f0;  // first frame of this system in the sweep's frame numbering
    long long p0, c0, h0;
    int blk0[SW_NSTAGE], nblk[SW_NSTAGE], bpf[SW_NSTAGE];
    const double *grid;  // the sweep's copy of the system's grid (x, y, z per point)
    const int *perm;     // and of its Morton order
    const double *pos, *diam, *mass, *vel;  // caller's frame-stacked inputs
    double *vir, *kin;                      // caller's frame-stacked outputs
};

// The system whose stage-s range holds workgroup b (ranges in system order).
__device__ __forceinline__ int sweep_system(const SweepSys *__restrict__ T, int ns, int s, int b)
{
    int k = 0;
    while (k + 1 < ns && b >= T[k + 1].blk0[s]) ++k;
    return k;
}

// The system of frame q of the sweep's frame numbering.
__device__ __forceinline__ int sweep_system_of_frame(const SweepSys *__restrict__ T, int ns, long long q)
{
    int k = 0;
    while (k + 1 < ns && q >= T[k + 1].f0) ++k;
    return k;
}

// Frame f and block bx of workgroup b in system S's range of stage s, frames
// dealt to XCDs whole as frame_block deals them (the range is
// frame_rows(nf) * bpf workgroups); false for the padding workgroups.
__device__ __forceinline__ bool sweep_frame_block(const SweepSys &S, int s, int b, int &f, int &bx)
{
    const int L = b - S.blk0[s], bpf = S.bpf[s];
    const int j = L >> 3;
    f = (L & 7) + 8 * (j / bpf);
    bx = j % bpf;
    return f < S.nf;
}

// Frame f and block bx of workgroup b in system S's range of stage s, frame
// after frame, workgroup x of a frame taking block xcd_tile(x, bpf) (each XCD
// a contiguous eighth of the frame); false for the padding workgroups that
// round the range up to a multiple of 8.
__device__ __forceinline__ bool sweep_frame_tile(const SweepSys &S, int s, int b, int &f, int &bx)
{
    const int L = b - S.blk0[s], bpf = S.bpf[s];
    f = L / bpf;
    bx = xcd_tile(L - f * bpf, bpf);
    return f < S.nf;
}

}  // namespace pgt
