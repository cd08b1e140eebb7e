// ik_launch.hpp -- launch context shared by the engine's host code and the
// field-gather translation units (ik_gather_*.hip), which are compiled separately and in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "ik_device.hpp"
#include "ik_geom.hpp"
#include "ik_sweep.hpp"

namespace pgt {

// Everything a batch's bond-list and gather launches read.
struct LaunchCtx {
    Geom g;
    int n, G, nf;
    int cap;
    int block;
    int pl;     // grid points per gather lane
    bool pack;  // LDS list entries packed into one int (see k_gather)
    mutable int pl_launched;  // points per lane of the gather actually launched
    const double *grid;
    const int *perm;
    GLRule R;
    const double *gld;
    double cgnorm;
    const int *cstart;
    const double *spos, *skin;
    const int *bcount;
    const float4 *rec;  // FP32 position and bond count per sorted particle
    const void *bonds;
    GatherOut out;
    int *flags;  // job flags: 1 an adaptive integral ran out of workspace, 2 a tile item had more than 255 bonds
    int tile;    // 1: item-balanced tile gather (Octic 5-node rule), 0: lane gather
    hipStream_t stream;
};

// Gather points per lane in use: the tuned value where the gather launchers
// pick the unrolled Octic 5-node rule, 1 elsewhere.
inline int gather_pl(int tuned, int cg, int nq, int D, int OD)
{
    return (cg == 0 && nq == 5 && !(D == 3 && OD == 2)) ? tuned : 1;
}

// Whether the tile gather serves this form (the Octic 5-node rule).
inline bool tile_form(int cg, int nq, int D, int OD) { return cg == 0 && nq == 5 && !(D == 3 && OD == 2); }

// The tile gather's launch shape: kTilePoints grid points per workgroup of
// kTileThreads lanes (one wavefront, one point per lane), in every form.
constexpr int kTilePoints = 64, kTileThreads = 64;
// Particles per frame the tile gather can index (26 bits of an item word).
constexpr int64_t kTileMaxParticles = (int64_t)1 << 26;

// Field gather of one batch (fixed-order rule), one function per form and
// translation unit: 2D positions with FP64 or FP32 bond entries, 3D positions
// with 2 x 2 output tensors (a 2D system off the z = 0 plane), and 3D with
// FP64 or FP32 entries.  fields: PGT_FIELD_* bits; cg: kernel kind.
void launch_gather_2d64(const LaunchCtx &c, unsigned fields, int cg);  // ik_gather_2d64.hip
void launch_gather_2d32(const LaunchCtx &c, unsigned fields, int cg);  // ik_gather_2d32.hip
void launch_gather_3d2(const LaunchCtx &c, unsigned fields, int cg);   // ik_gather_3d2.hip
void launch_gather_3d64(const LaunchCtx &c, unsigned fields, int cg);  // ik_gather_3d64.hip
void launch_gather_3d32(const LaunchCtx &c, unsigned fields, int cg);  // ik_gather_3d32.hip
// Lanes per grid point of the adaptive gather: the point's bond integrals
// are dealt to them, so that a small grid still fills the GPU (about 64K
// lanes per frame; a power of two from 1 to 32).  Measured on a 1,024-point
// frame: 8 lanes 5.6 ms, 16 lanes 3.4 ms, 32 lanes 2.2 ms (Octic).
inline int adaptive_lanes(int64_t G)
{
    int l = 1;
    while (l < 32 && (int64_t)l * G < 65536) l <<= 1;
    return l;
}

// Field gather with the adaptive (QAG) bond integral.  ws null: each
// integral's interval table in the lane's private memory (kQagMax
// intervals), and capped gets 1 when an integral needs more; else the global
// workspace ws of adaptive_workspace_doubles(c, wcap) doubles holds wcap
// intervals per lane.
void launch_gather_adaptive(const LaunchCtx &c, unsigned fields, int cg, int D, int OD, AdaptQ Q, int *capped,
                            double *ws, int wcap);
size_t adaptive_workspace_doubles(const LaunchCtx &c, int wcap);  // ik_gather_adaptive.hip

// The field gather of a sweep (ik_sweep.hpp): one launch over the gather
// ranges of every system, `blocks` workgroups of kTileThreads lanes.
struct SweepGatherCtx {
    const SweepSys *tab;  // device copy of the sweep's table
    int ns;
    int blocks;
    GLRule R;
    const int *cstart;
    const double *spos, *skin;
    int cap;
    const int *bcount;
    const float4 *rec;  // 3D: FP32 position and bond count per sorted particle
    const void *bonds;
    unsigned long long *wave_counts;
    int *flags;
    hipStream_t stream;
};
// D: 2 (2 x 2 fields) or 3 (3 x 3)
void launch_sweep_gather(const SweepGatherCtx &c, int D, bool fp32, unsigned fields);  // ik_sweep_gather.hip

}  // namespace pgt
