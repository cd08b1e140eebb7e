// ik_gather_launch.hpp -- template dispatch from run-time settings (fields,
// kernel, gather kind, workgroup size, points per lane) to the k_gather and
// k_gather_tile instantiations.
// Included by the ik_gather_<form>.hip translation units only.
#pragma once
#include <hip/hip_runtime.h>

#include "ik_gather.hpp"
#include "ik_gather_tile.hpp"
#include "ik_launch.hpp"
#include "pgt_ik.h"

namespace pgt {

template <int D, int OD, int CG, int NQ, bool VIR, bool ELA, typename ER, int BLK, int PL>
void launch_gather_t(const LaunchCtx &c)
{
    // per-lane list length: ~2x the mean number of particles in a ball at
    // unit density (12.6 in 2D, 33.5 in 3D for r_cg = 2), more for the union
    // of several points' balls; longer lists are processed in pieces
    constexpr int MAXL = (D == 2 ? 24 : 48) * (PL == 1 ? 3 : PL == 2 ? 4 : 5) / 3;
    dim3 grid((unsigned)((c.G + BLK * PL - 1) / (BLK * PL)), (unsigned)c.nf);
    c.pl_launched = PL;
    if (c.pack)
        hipLaunchKernelGGL((k_gather<D, OD, CG, NQ, VIR, ELA, ER, BLK, MAXL, true, PL>), grid, dim3(BLK), 0, c.stream,
                           c.g, c.n, c.G, c.grid, c.perm, c.R, c.gld, c.cgnorm, c.cstart, c.spos, c.skin, c.cap,
                           c.bcount, c.rec, (const ER *)c.bonds, c.out);
    else
        hipLaunchKernelGGL((k_gather<D, OD, CG, NQ, VIR, ELA, ER, BLK, MAXL, false, PL>), grid, dim3(BLK), 0,
                           c.stream, c.g, c.n, c.G, c.grid, c.perm, c.R, c.gld, c.cgnorm, c.cstart, c.spos, c.skin,
                           c.cap, c.bcount, c.rec, (const ER *)c.bonds, c.out);
}

template <int D, int OD, int CG, int NQ, bool VIR, bool ELA, typename ER, int BLK>
void launch_gather_p(const LaunchCtx &c)
{
    if constexpr (CG == 0 && NQ == 5) {
        if (c.pl == 4) return launch_gather_t<D, OD, CG, NQ, VIR, ELA, ER, BLK, 4>(c);
        if (c.pl == 2) return launch_gather_t<D, OD, CG, NQ, VIR, ELA, ER, BLK, 2>(c);
    }
    launch_gather_t<D, OD, CG, NQ, VIR, ELA, ER, BLK, 1>(c);
}

template <int D, int OD, int CG, int NQ, bool VIR, bool ELA, typename ER>
void launch_gather_b(const LaunchCtx &c)
{
    if (c.block == 128) launch_gather_p<D, OD, CG, NQ, VIR, ELA, ER, 128>(c);
    else launch_gather_p<D, OD, CG, NQ, VIR, ELA, ER, 256>(c);
}

template <int D, int OD, int CG, int NQ, typename ER>
void launch_gather_f(const LaunchCtx &c, unsigned fields)
{
    const bool V = fields & PGT_FIELD_VIRIAL, E = fields & PGT_FIELD_ELASTIC;
    if (V && E) launch_gather_b<D, OD, CG, NQ, true, true, ER>(c);
    else if (V) launch_gather_b<D, OD, CG, NQ, true, false, ER>(c);
    else if (E) launch_gather_b<D, OD, CG, NQ, false, true, ER>(c);
    else launch_gather_b<D, OD, CG, NQ, false, false, ER>(c);
}

template <int D, int OD, bool VIR, bool ELA, bool KIN, typename ER>
void launch_tile_t(const LaunchCtx &c)
{
    using CP = TileCaps;
    const dim3 grid((unsigned)((c.G + CP::TP - 1) / CP::TP), (unsigned)c.nf);
    c.pl_launched = 1;
    hipLaunchKernelGGL((k_gather_tile<D, OD, 0, 5, VIR, ELA, KIN, ER>), grid, dim3(CP::BLK), 0, c.stream, c.g, c.n, c.G,
                       c.grid, c.perm, c.R, c.gld, c.cgnorm, c.cstart, c.spos, c.skin, c.cap, c.bcount, c.rec,
                       (const ER *)c.bonds, c.out, c.flags);
}

template <int D, int OD, bool VIR, bool ELA, typename ER>
void launch_tile_k(const LaunchCtx &c)
{
    if (c.skin) launch_tile_t<D, OD, VIR, ELA, true, ER>(c);
    else launch_tile_t<D, OD, VIR, ELA, false, ER>(c);
}

template <int D, int OD, typename ER>
void launch_tile_f(const LaunchCtx &c, unsigned fields)
{
    const bool V = fields & PGT_FIELD_VIRIAL, E = fields & PGT_FIELD_ELASTIC;
    if (V && E) launch_tile_k<D, OD, true, true, ER>(c);
    else if (V) launch_tile_k<D, OD, true, false, ER>(c);
    else if (E) launch_tile_k<D, OD, false, true, ER>(c);
    else launch_tile_k<D, OD, false, false, ER>(c);
}

// The Octic kernel gets the unrolled 5-node rule (the driver scripts' order)
// and the FP32 form, by the tile gather when c.tile is set; Rect and
// Mollifier run the generic FP64 rule.
template <int D, int OD, typename ER>
void launch_gather_c(const LaunchCtx &c, unsigned fields, int cg)
{
    if constexpr (!(D == 3 && OD == 2)) {
        if (c.tile && cg == 0 && c.R.n == 5) return launch_tile_f<D, OD, ER>(c, fields);
    }
    if (cg == 0) {
        if (c.R.n == 5 && !(D == 3 && OD == 2)) launch_gather_f<D, OD, 0, 5, ER>(c, fields);
        else launch_gather_f<D, OD, 0, 0, ER>(c, fields);
    } else if constexpr (std::is_same<ER, double>::value) {
        if (cg == 1) launch_gather_f<D, OD, 1, 0, double>(c, fields);
        else launch_gather_f<D, OD, 2, 0, double>(c, fields);
    }
}

}  // namespace pgt
