// ik_sweep_gather.hip -- the sweep's field-gather instantiations (2D or 3D,
// Octic 5-node rule, FP64 or FP32 bond entries; virial, kinetic or both).
#include "ik_gather_tile.hpp"
#include "ik_launch.hpp"
#include "pgt_ik.h"

namespace pgt {

namespace {

template <int D, bool VIR, bool KIN, typename ER>
void launch_t(const SweepGatherCtx &c)
{
    hipLaunchKernelGGL((k_sweep_gather<D, D, VIR, KIN, ER>), dim3((unsigned)c.blocks), dim3(kTileThreads), 0,
                       c.stream, c.tab, c.ns, c.R, c.cstart, c.spos, c.skin, c.cap, c.bcount, c.rec,
                       (const ER *)c.bonds, c.wave_counts, c.flags);
}

template <int D, typename ER>
void launch_f(const SweepGatherCtx &c, unsigned fields)
{
    const bool V = fields & PGT_FIELD_VIRIAL, K = fields & PGT_FIELD_KINETIC;
    if (V && K) launch_t<D, true, true, ER>(c);
    else if (V) launch_t<D, true, false, ER>(c);
    else launch_t<D, false, true, ER>(c);
}

}  // namespace

void launch_sweep_gather(const SweepGatherCtx &c, int D, bool fp32, unsigned fields)
{
    if (D == 2) {
        if (fp32) launch_f<2, float>(c, fields);
        else launch_f<2, double>(c, fields);
    } else {
        if (fp32) launch_f<3, float>(c, fields);
        else launch_f<3, double>(c, fields);
    }
}

}  // namespace pgt
