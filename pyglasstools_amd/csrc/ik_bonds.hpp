// ik_bonds.hpp -- bond-list kernel of the IK stress-field engine.
// Included by ik_kernels.hpp.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ik_device.hpp"
#include "ik_geom.hpp"

namespace pgt {

// Entries of particle p (cell-sorted index) sit at bonds[(f*n + p)*cap*ES ...],
// one per neighbour j within max_rcut with a nonzero pair force (or, with the
// elastic field, a nonzero force at r_ij + u_ij).  N_cg is the kernel
// normalisation and w the quadrature sum of the unnormalised kernel shape.
//
// FP64 (ER = double), ES = D + 1 (+ KN) values, no padding:
//   [0, D)     b' = r_ij / rc                   (scaled bond vector)
//   [D]        c' = 0.5 * N_cg * (F/r) * rc^2   (virial: T += c' w b' (x) b')
//   [D+1, ..)  elastic: 0.5 * N_cg * (F1_a (u+r)_b - F_a r_b), packed symmetric
// The gather reads entries two at a time (a group of 2*ES doubles, a whole
// number of 16-byte pieces, 16-byte aligned since cap is even) and
// recomputes |b'|^2 itself: the memory path of a CU returns about one
// scattered 16-byte piece per lane per cycle, so bytes per bond set the
// gather's speed.
//
// FP32 (ER = float), the same fields unpadded (a 2D virial entry is 12
// bytes, a 3D one 16), read in groups of 4, 2 or 1 entries (the fewest that
// make a whole number of 16-byte pieces).
template <int D, int OD, bool ELA, typename ER>
struct BondLayout {
    static constexpr int KN = OD * (OD + 1) / 2;
    static constexpr bool F64 = std::is_same<ER, double>::value;
    static constexpr bool HAS_C2 = false;  // readers recompute |b'|^2
    static constexpr int OC = D;           // offset of c'
    static constexpr int OE = OC + 1;      // offset of the elastic terms
    static constexpr int RAW = OE + (ELA ? KN : 0);
    static constexpr int PER16 = 16 / (int)sizeof(ER);
    static constexpr int ES = RAW;
    // entries per read group (FP64 always pairs)
    static constexpr int GRP = F64 ? 2 : (RAW % PER16 == 0 ? 1 : ((2 * RAW) % PER16 == 0 ? 2 : 4));
    static constexpr int GV = GRP * ES;  // values per group, a multiple of PER16
    static_assert(GV % PER16 == 0, "a read group is a whole number of 16-byte pieces");
};

// One lane per particle.  Pass 1 walks the cells within reach of the particle's
// cell (candidate positions loaded a batch at a time) and keeps, in a per-lane
// LDS list, the candidates within max_rcut (about a third of them).  Pass 2
// evaluates the pair force only for listed candidates, so the force code runs
// with the lanes of a wavefront in step instead of behind a branch that some
// lane takes at almost every candidate.  A full list is processed early.
// Entries keep the candidate order (cells in a fixed order, input order
// inside a cell).  In 3D pass 1 reads the FP32 records and tests distances in
// FP32 against bounds of the rounding (Geom::bin32, bout32), falling back to
// the FP64 test in the band between them, so the list is that of the FP64
// test; the kernel then stores each particle's bond count in its record.
//
// The body works on one frame's arrays (cell starts st, sorted positions sp,
// diameters sd, displacements sdp, counts bc, entries bd, FP32 records rf)
// for the particles of block pb (BLK consecutive cell-sorted particles), and
// stores the wavefronts' largest counts at wmax[0 .. BLK/64).
template <int D, int OD, bool VIR, bool ELA, typename ER, int BLK, int MAXC>
__device__ __forceinline__ void bonds_body(const Geom &g, int n, const PotParams &P, double cgnorm,
                                           const int *__restrict__ st, const double *__restrict__ sp,
                                           const double *__restrict__ sd, const double *__restrict__ sdp, int cap,
                                           int *__restrict__ bc, ER *__restrict__ bd, int *__restrict__ wmax,
                                           float4 *__restrict__ rf, int pb)
{
    constexpr int PS = (D == 2) ? 2 : 4;
    using LY = BondLayout<D, OD, ELA, ER>;
    constexpr int ES = LY::ES;
    __shared__ int lq[MAXC][BLK];
    const int lane = threadIdx.x;
    const int p_raw = pb * BLK + lane;
    // lanes past the end shadow the last particle without writing, so the
    // whole wavefront takes part in the max-reduction at the end
    const bool live = p_raw < n;
    const int p = live ? p_raw : n - 1;
    const double xi = sp[(size_t)p * PS], yi = sp[(size_t)p * PS + 1], zi = (D == 3) ? sp[(size_t)p * PS + 2] : 0.0;
    const double di = sd[p];
    const float xi32 = (float)xi, yi32 = (float)yi, zi32 = (float)zi;
    ER *out = bd + (size_t)p * (size_t)cap * ES;
    const double h = 0.5 * cgnorm;
    int cnt = 0;
    constexpr int NP = LY::GRP > 1 ? LY::GRP - 1 : 1;
    ER pend[NP][LY::RAW];  // entries awaiting the rest of their read group
#pragma unroll
    for (int s = 0; s < NP; ++s)
#pragma unroll
        for (int k = 0; k < LY::RAW; ++k) pend[s][k] = ER(0);
    auto store_group = [](ER *dst, const ER (&gv)[LY::GV]) {
        using V = typename std::conditional<std::is_same<ER, double>::value, double2, float4>::type;
        V *q = reinterpret_cast<V *>(dst);
#pragma unroll
        for (int k = 0; k < LY::GV / LY::PER16; ++k) {
            V t;
            if constexpr (LY::PER16 == 2) {
                t.x = gv[2 * k];
                t.y = gv[2 * k + 1];
            } else {
                t.x = gv[4 * k];
                t.y = gv[4 * k + 1];
                t.z = gv[4 * k + 2];
                t.w = gv[4 * k + 3];
            }
            q[k] = t;
        }
    };

    // pass 2 over the first m listed candidates, NF at a time with their
    // positions and diameters loaded together
    constexpr int NF = 4;
    auto forces = [&](int m) {
        for (int k0 = 0; k0 < m; k0 += NF) {
            int qq[NF];
            double cqs[NF][PS], djs[NF];
#pragma unroll
            for (int u = 0; u < NF; ++u) {
                qq[u] = lq[min(k0 + u, m - 1)][lane];
#pragma unroll
                for (int a = 0; a < PS; ++a) cqs[u][a] = sp[(size_t)qq[u] * PS + a];
                djs[u] = sd[qq[u]];
            }
#pragma unroll
            for (int u = 0; u < NF; ++u) {
                if (k0 + u >= m) break;
                const int q = qq[u];
                const double dj = djs[u];
                double b[3];
                b[0] = minimg(cqs[u][0] - xi, g.L[0], g.pinvL[0]);
                b[1] = minimg(cqs[u][1] - yi, g.L[1], g.pinvL[1]);
                b[2] = (D == 3) ? minimg(cqs[u][2] - zi, g.L[2], g.pinvL[2]) : 0.0;
                const double rsq = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
                double f0, f1 = 0.0;
                double ur[3] = {0, 0, 0};
                if constexpr (std::is_same<ER, double>::value) f0 = force_divr<double>(P, rsq, di, dj);
                else f0 = (double)force_divr<float>(P, (float)rsq, (float)di, (float)dj);
                if (ELA) {
                    for (int a = 0; a < 3; ++a) ur[a] = sdp[(size_t)q * 3 + a] - sdp[(size_t)p * 3 + a] + b[a];
                    const double ursq = ur[0] * ur[0] + ur[1] * ur[1] + ur[2] * ur[2];
                    if constexpr (std::is_same<ER, double>::value) f1 = force_divr<double>(P, ursq, di, dj);
                    else f1 = (double)force_divr<float>(P, (float)ursq, (float)di, (float)dj);
                    if (f0 == 0.0 && f1 == 0.0) continue;
                } else if (f0 == 0.0) {
                    continue;
                }
                if (live && cnt < cap) {
                    ER ev[LY::RAW];
                    for (int a = 0; a < D; ++a) ev[a] = (ER)(b[a] * g.inv_rc);
                    if constexpr (LY::HAS_C2) ev[D] = (ER)(rsq * g.inv_rc2);
                    ev[LY::OC] = VIR ? (ER)(h * f0 / g.inv_rc2) : (ER)0;
                    if constexpr (ELA) {
                        ER *t = ev + LY::OE;
                        auto T = [&](int a, int c) { return h * (f1 * ur[a] * ur[c] - f0 * b[a] * b[c]); };
                        for (int a = 0; a < OD; ++a) t[a] = (ER)T(a, a);
                        if (OD == 2) t[2] = (ER)T(0, 1);
                        else { t[3] = (ER)T(0, 1); t[4] = (ER)T(0, 2); t[5] = (ER)T(1, 2); }
                    }
                    if constexpr (LY::GRP > 1) {
                        // the first GRP - 1 entries of a group wait in
                        // registers; the last sends the group as 16-byte stores
                        const int slot = cnt % LY::GRP;
                        if (slot < LY::GRP - 1) {
#pragma unroll
                            for (int s = 0; s < NP; ++s)
                                if (s == slot)
#pragma unroll
                                    for (int k = 0; k < LY::RAW; ++k) pend[s][k] = ev[k];
                        } else {
                            ER gv[LY::GV];
#pragma unroll
                            for (int s = 0; s < LY::GRP - 1; ++s)
#pragma unroll
                                for (int k = 0; k < LY::RAW; ++k) gv[s * LY::ES + k] = pend[s][k];
#pragma unroll
                            for (int k = 0; k < LY::RAW; ++k) gv[(LY::GRP - 1) * LY::ES + k] = ev[k];
                            store_group(out + (size_t)(cnt - (LY::GRP - 1)) * ES, gv);
                        }
                    } else {
                        ER *e = out + (size_t)cnt * ES;
#pragma unroll
                        for (int k = 0; k < LY::RAW; ++k) e[k] = ev[k];
                    }
                }
                ++cnt;
            }
        }
    };

    // pass 1: candidates in batches of NB, the batch's loads issued together
    constexpr int NB = 8;
    static_assert(MAXC >= NB, "list shorter than a batch");
    const int cx = cell_coord(g, 0, xi), cy = cell_coord(g, 1, yi), cz = cell_coord(g, 2, zi);
    int rows[kMaxRows];
    const int nrows = neighbour_rows(g, cy, cz, rows);
    int m = 0;
    for (int r = 0; r < nrows; ++r) {
        int rb[2], re[2];
        const int nr = row_ranges(g, st, rows[r], cx, rb, re);
        for (int s = 0; s < nr; ++s) {
            for (int q0 = rb[s]; q0 < re[s]; q0 += NB) {
                if (m > MAXC - NB) {  // room for a whole batch
                    forces(m);
                    m = 0;
                }
                // exact test: FP64 minimum image and |r_ij|^2 against max_rcut^2
                auto within = [&](const double *cq) {
                    const double b0 = minimg(cq[0] - xi, g.L[0], g.pinvL[0]);
                    const double b1 = minimg(cq[1] - yi, g.L[1], g.pinvL[1]);
                    const double b2 = (D == 3) ? minimg(cq[2] - zi, g.L[2], g.pinvL[2]) : 0.0;
                    return b0 * b0 + b1 * b1 + b2 * b2 < g.mr2;
                };
                if constexpr (D == 3) {
                    // FP32 positions from the records, tested against
                    // bounds of the FP64 rounding; the band between them
                    // goes to the exact test (see Geom::bin32)
                    float c32[NB][3];
#pragma unroll
                    for (int u = 0; u < NB; ++u) {
                        const float *r = reinterpret_cast<const float *>(rf + min(q0 + u, re[s] - 1));
#pragma unroll
                        for (int a = 0; a < 3; ++a) c32[u][a] = r[a];
                    }
#pragma unroll
                    for (int u = 0; u < NB; ++u) {
                        const int q = q0 + u;
                        if (q >= re[s]) break;
                        float d0 = c32[u][0] - xi32, d1 = c32[u][1] - yi32, d2 = c32[u][2] - zi32;
                        d0 = fmaf(-g.L32[0], rintf(d0 * g.pinvL32[0]), d0);
                        d1 = fmaf(-g.L32[1], rintf(d1 * g.pinvL32[1]), d1);
                        d2 = fmaf(-g.L32[2], rintf(d2 * g.pinvL32[2]), d2);
                        const float a2f = d0 * d0 + d1 * d1 + d2 * d2;
                        bool in = a2f < g.bin32;
                        if (!in && !(a2f > g.bout32)) {
                            double cq[3];
#pragma unroll
                            for (int a = 0; a < 3; ++a) cq[a] = sp[(size_t)q * PS + a];
                            in = within(cq);
                        }
                        if (q == p || !in) continue;
                        lq[m++][lane] = q;
                    }
                } else {
                    double cq[NB][PS];
#pragma unroll
                    for (int u = 0; u < NB; ++u) {
                        const int q = min(q0 + u, re[s] - 1);
#pragma unroll
                        for (int a = 0; a < PS; ++a) cq[u][a] = sp[(size_t)q * PS + a];
                    }
#pragma unroll
                    for (int u = 0; u < NB; ++u) {
                        const int q = q0 + u;
                        if (q >= re[s]) break;
                        if (q == p || !within(cq[u])) continue;
                        lq[m++][lane] = q;
                    }
                }
            }
        }
    }
    if (m > 0) forces(m);
    if constexpr (LY::GRP > 1) {
        const int r = cnt % LY::GRP;  // entries of an unfinished last group
        if (live && r && cnt <= cap)
#pragma unroll
            for (int s = 0; s < NP; ++s)
                if (s < r) {
                    ER *e = out + (size_t)(cnt - r + s) * ES;
#pragma unroll
                    for (int k = 0; k < LY::RAW; ++k) e[k] = pend[s][k];
                }
    }

    if (live) {
        bc[p] = cnt;
        // 3D: the bond count completes the particle's FP32 record (the
        // candidate scans read only its position words)
        if (D == 3) reinterpret_cast<int *>(rf + p)[3] = cnt;
    }
    int mc = live ? cnt : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mc = max(mc, __shfl_xor(mc, off, 64));
    // the largest count of the wavefront, stored per wavefront in launch
    // order (k_reduce_stats takes the maximum)
    if ((lane & 63) == 0) wmax[lane >> 6] = mc;
}

// One frame per y row; block x takes the particles of block
// xcd_tile(x, gridDim.x), so each XCD works through a contiguous eighth of
// the frame.
template <int D, int OD, bool VIR, bool ELA, typename ER, int BLK, int MAXC>
__global__ __launch_bounds__(BLK) void k_bonds(Geom g, int n, PotParams P, double cgnorm,
                                               const int *__restrict__ cstart, const double *__restrict__ spos,
                                               const double *__restrict__ sdiam, const double *__restrict__ sdisp,
                                               int cap, int *__restrict__ bcount, ER *__restrict__ bonds,
                                               int *__restrict__ wave_max, float4 *__restrict__ rec)
{
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int ES = BondLayout<D, OD, ELA, ER>::ES;
    const int f = blockIdx.y;
    const size_t fo = (size_t)f * n;
    bonds_body<D, OD, VIR, ELA, ER, BLK, MAXC>(
        g, n, P, cgnorm, cstart + (size_t)f * (g.ncell + 1), spos + fo * PS, sdiam + fo,
        ELA ? sdisp + fo * 3 : nullptr, cap, bcount + fo, bonds + fo * (size_t)cap * ES,
        wave_max + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * (BLK / 64), D == 3 ? rec + fo : nullptr,
        xcd_tile(blockIdx.x, gridDim.x));
}

}  // namespace pgt
