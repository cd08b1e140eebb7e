// ik_gather_adaptive.hip -- instantiations of the field gather with the
// adaptive (QAG) bond integral.
#include "ik_gather_adaptive.hpp"
#include "ik_launch.hpp"
#include "pgt_ik.h"

namespace pgt {

template <int D, int OD, bool GWS>
static void launch_adaptive_w(const LaunchCtx &c, unsigned fields, int cg, AdaptQ Q, int *capped, double *ws,
                              int wcap)
{
    const int lpp = adaptive_lanes(c.G);
    const dim3 grid((unsigned)(((size_t)c.G * lpp + 127) / 128), (unsigned)c.nf);
    const bool V = fields & PGT_FIELD_VIRIAL, E = fields & PGT_FIELD_ELASTIC;
    const double *bonds = (const double *)c.bonds;
    if (V && E)
        hipLaunchKernelGGL((k_gather_adaptive<D, OD, true, true, GWS>), grid, dim3(128), 0, c.stream, c.g, c.n, c.G,
                           c.grid, c.perm, cg, Q, c.cgnorm, c.cstart, c.spos, c.skin, c.cap, c.bcount, bonds, c.out,
                           capped, ws, wcap, lpp);
    else if (E)
        hipLaunchKernelGGL((k_gather_adaptive<D, OD, false, true, GWS>), grid, dim3(128), 0, c.stream, c.g, c.n, c.G,
                           c.grid, c.perm, cg, Q, c.cgnorm, c.cstart, c.spos, c.skin, c.cap, c.bcount, bonds, c.out,
                           capped, ws, wcap, lpp);
    else
        hipLaunchKernelGGL((k_gather_adaptive<D, OD, true, false, GWS>), grid, dim3(128), 0, c.stream, c.g, c.n, c.G,
                           c.grid, c.perm, cg, Q, c.cgnorm, c.cstart, c.spos, c.skin, c.cap, c.bcount, bonds, c.out,
                           capped, ws, wcap, lpp);
}

template <int D, int OD>
static void launch_adaptive_d(const LaunchCtx &c, unsigned fields, int cg, AdaptQ Q, int *capped, double *ws,
                              int wcap)
{
    if (ws) launch_adaptive_w<D, OD, true>(c, fields, cg, Q, capped, ws, wcap);
    else launch_adaptive_w<D, OD, false>(c, fields, cg, Q, capped, nullptr, 0);
}

size_t adaptive_workspace_doubles(const LaunchCtx &c, int wcap)
{
    return (((size_t)c.G * adaptive_lanes(c.G) + 127) / 128) * 128 * (size_t)c.nf * 4 * (size_t)wcap;
}

void launch_gather_adaptive(const LaunchCtx &c, unsigned fields, int cg, int D, int OD, AdaptQ Q, int *capped,
                            double *ws, int wcap)
{
    if (D == 2) launch_adaptive_d<2, 2>(c, fields, cg, Q, capped, ws, wcap);
    else if (OD == 2) launch_adaptive_d<3, 2>(c, fields, cg, Q, capped, ws, wcap);
    else launch_adaptive_d<3, 3>(c, fields, cg, Q, capped, ws, wcap);
}

}  // namespace pgt
