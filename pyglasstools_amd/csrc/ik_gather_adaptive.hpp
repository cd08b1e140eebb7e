// ik_gather_adaptive.hpp -- field gather with the adaptive bond integral.
// Included by ik_gather_adaptive.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ik_bonds.hpp"
#include "ik_device.hpp"
#include "ik_geom.hpp"
#include "ik_launch.hpp"

namespace pgt {

// lpp lanes per grid point (Morton order, as k_gather), so that even
// a grid of a thousand points fills the GPU: every lane of a point walks the
// cells around it, and the point's (particle, bond) integrals are dealt to
// its lanes round-robin in that walk's order.  Each particle inside the ball
// adds its kinetic term, and each of its bonds the pair term with the bond
// integral
//   int_0^1 Delta(|x - (r_i + s r_ij)|^2) ds
// by QAG (qag_dev), the integrand normalised as the reference integrates it
// (CoarseGrainFunction.h:42-45), so the tolerance tests see the same values.
// The kernel shape is chosen at run time (cg), the bond list is the FP64
// layout of BondLayout.  Each lane sums its share in FP64 in walk order, and
// the point's lanes are added by a fixed butterfly.  GWS false: the
// interval table of each integral in private memory (kQagMax intervals), an
// integral that needs more sets *capped; GWS true: in the global workspace
// ws, wcap intervals per lane (the caller's limit), lane-interleaved.
template <int D, int OD, bool VIR, bool ELA, bool GWS>
__global__ __launch_bounds__(128) void k_gather_adaptive(Geom g, int n, int G, const double *__restrict__ grid,
                                                         const int *__restrict__ perm, int cg, AdaptQ Q,
                                                         double cgnorm, const int *__restrict__ cstart,
                                                         const double *__restrict__ spos,
                                                         const double *__restrict__ skin, int cap,
                                                         const int *__restrict__ bcount,
                                                         const double *__restrict__ bonds, GatherOut out,
                                                         int *__restrict__ capped, double *__restrict__ ws, int wcap,
                                                         int lpp)
{
    constexpr int BLK = 128;
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int KN = OD * (OD + 1) / 2;
    constexpr bool PAIRS = VIR || ELA;
    using LY = BondLayout<D, OD, ELA, double>;
    constexpr int ES = LY::ES;
    const int LPP = lpp;  // a power of two from 1 to 32 (adaptive_lanes)
    const int f = blockIdx.y;
    const int lane = threadIdx.x;
    const int sub = lane & (LPP - 1);  // this lane's share of the point's integrals
    const int t = (xcd_tile(blockIdx.x, gridDim.x) * BLK + lane) / LPP;
    const bool live = t < G;
    const bool KIN = skin != nullptr;
    const double *sp = spos + (size_t)f * n * PS;
    const int *st = cstart + (size_t)f * (g.ncell + 1);
    const int *bc = bcount + (size_t)f * n;
    const double *bd = bonds + (size_t)f * n * (size_t)cap * ES;
    const double *sk = KIN ? skin + (size_t)f * n * KN : nullptr;

    double av[KN], ae[KN], ak[KN];
#pragma unroll
    for (int k = 0; k < KN; ++k) { av[k] = 0.0; ae[k] = 0.0; ak[k] = 0.0; }
    unsigned long long nball = 0, ncontrib = 0;
    bool hit_cap = false;
    const int gp = live ? perm[t] : 0;
    using WS = typename std::conditional<GWS, QagGlobal, QagPrivate>::type;
    WS wsp;
    if constexpr (GWS) {
        const size_t stride = (size_t)gridDim.x * gridDim.y * BLK;
        wsp.p = ws + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * BLK + lane;
        wsp.stride = stride;
        wsp.n = wcap;
    }

    int pcount = 0, bcount_walk = 0;  // particles and bonds met in the walk, for dealing them to lanes
    if (live) {
        const double x0 = grid[(size_t)gp * 3], x1 = grid[(size_t)gp * 3 + 1];
        const double x2 = (D == 3) ? grid[(size_t)gp * 3 + 2] : 0.0;
        const int cx = cell_coord(g, 0, x0), cy = cell_coord(g, 1, x1), cz = cell_coord(g, 2, x2);
        int rows[kMaxRows];
        const int nrows = neighbour_rows(g, cy, cz, rows);
        for (int r = 0; r < nrows; ++r) {
            int rb[2], re[2];
            const int nr = row_ranges(g, st, rows[r], cx, rb, re);
            for (int s = 0; s < nr; ++s)
                for (int p = rb[s]; p < re[s]; ++p) {
                    double a[3];
                    a[0] = minimg(sp[(size_t)p * PS] - x0, g.L[0], g.pinvL[0]);
                    a[1] = minimg(sp[(size_t)p * PS + 1] - x1, g.L[1], g.pinvL[1]);
                    a[2] = (D == 3) ? minimg(sp[(size_t)p * PS + 2] - x2, g.L[2], g.pinvL[2]) : 0.0;
                    const double asq = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
                    if (!(asq < g.rc2)) continue;
                    if (sub == 0) ++nball;
                    if (KIN && (pcount++ & (LPP - 1)) == sub) {
                        const double w = shape_rt(cg, asq * g.inv_rc2);
#pragma unroll
                        for (int k = 0; k < KN; ++k) ak[k] = fma(w, sk[(size_t)p * KN + k], ak[k]);
                    }
                    if constexpr (PAIRS) {
                        const int nb = min(bc[p], cap);
                        double as[3];  // r_i - x in units of the kernel width
#pragma unroll
                        for (int k = 0; k < 3; ++k) as[k] = a[k] * g.inv_rc;
                        for (int j = 0; j < nb; ++j) {
                            if ((bcount_walk++ & (LPP - 1)) != sub) continue;
                            ++ncontrib;
                            const double *e = bd + ((size_t)p * cap + j) * ES;
                            double b[3] = {e[0], e[1], D == 3 ? e[2] : 0.0};
                            auto obj = [&](double sv) {
                                double tt = 0.0;
#pragma unroll
                                for (int k = 0; k < D; ++k) {
                                    const double d = as[k] + sv * b[k];
                                    tt += d * d;
                                }
                                return cgnorm * shape_rt(cg, tt);
                            };
                            const double w = qag_dev(obj, Q, hit_cap, wsp) / cgnorm;
                            if (VIR) {
                                const double c = w * e[LY::OC];
                                av[0] = fma(c * b[0], b[0], av[0]);
                                av[1] = fma(c * b[1], b[1], av[1]);
                                if (OD == 2) {
                                    av[2] = fma(c * b[0], b[1], av[2]);
                                } else {
                                    av[2] = fma(c * b[2], b[2], av[2]);
                                    av[3] = fma(c * b[0], b[1], av[3]);
                                    av[4] = fma(c * b[0], b[2], av[4]);
                                    av[5] = fma(c * b[1], b[2], av[5]);
                                }
                            }
                            if (ELA)
#pragma unroll
                                for (int m = 0; m < KN; ++m) ae[m] = fma(w, e[LY::OE + m], ae[m]);
                        }
                    }
                }
        }
    }
    // the point's lanes: a fixed butterfly over the LPP lanes of the group
    // (every lane of a group is live or none is)
    for (int m = 1; m < LPP; m <<= 1)
#pragma unroll
        for (int k = 0; k < KN; ++k) {
            av[k] += __shfl_xor(av[k], m, 64);
            ae[k] += __shfl_xor(ae[k], m, 64);
            ak[k] += __shfl_xor(ak[k], m, 64);
        }
    if (live && sub == 0) {
        const size_t ob = ((size_t)f * G + gp) * OD * OD;
        auto store = [&](double *dst, const double *acc, double scale) {
            for (int aa = 0; aa < OD; ++aa) dst[ob + aa + OD * aa] = acc[aa] * scale;
            if (OD == 2) {
                dst[ob + 1] = acc[2] * scale;
                dst[ob + 2] = acc[2] * scale;
            } else {
                dst[ob + 1] = dst[ob + 3] = acc[3] * scale;
                dst[ob + 2] = dst[ob + 6] = acc[4] * scale;
                dst[ob + 5] = dst[ob + 7] = acc[5] * scale;
            }
        };
        if (VIR) store(out.vir, av, 1.0);
        if (ELA) store(out.ela, ae, 1.0);
        if (KIN) store(out.kin, ak, cgnorm);
    }
    if (hit_cap) atomicOr(capped, 1);
    nball = wave_sum(nball);
    ncontrib = wave_sum(ncontrib);
    if ((lane & 63) == 0) {
        const size_t w = ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * (BLK / 64) + (lane >> 6);
        out.wave_counts[2 * w] = nball;
        out.wave_counts[2 * w + 1] = ncontrib;
    }
}

}  // namespace pgt
