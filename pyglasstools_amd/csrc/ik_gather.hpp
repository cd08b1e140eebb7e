// ik_gather.hpp -- the field-gather kernel of the IK stress-field engine.
// Included by ik_kernels.hpp.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ik_device.hpp"

namespace pgt {

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

// sum_k w_k shape(t(s_k)), t(s) = A + s B + s^2 C: the bond integral of the
// unnormalised kernel along the segment, in units of the kernel width.
template <int CG, int NQ, typename ER>
__device__ __forceinline__ ER bond_quad(const GLRule &R, const double *__restrict__ gld, int nq, ER A, ER B, ER C)
{
    ER acc = ER(0);
    if constexpr (NQ > 0) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const ER s = (ER)R.s[k];
            const ER t = fma(s, fma(s, C, B), A);
            acc = fma((ER)R.w[k], shape<CG, ER>(t), acc);
        }
    } else {
        for (int k = 0; k < nq; ++k) {
            const ER s = (ER)gld[k];
            const ER t = fma(s, fma(s, C, B), A);
            acc = fma((ER)gld[nq + k], shape<CG, ER>(t), acc);
        }
    }
    return acc;
}

// N values (a whole number of 16-byte pieces, 16-byte aligned) in 16-byte loads.
template <typename ER, int ES>
__device__ __forceinline__ void load_entry(const ER *__restrict__ p, ER (&v)[ES])
{
    using V = typename std::conditional<std::is_same<ER, double>::value, double2, float4>::type;
    constexpr int PER = 16 / (int)sizeof(ER);
    const V *q = reinterpret_cast<const V *>(p);
#pragma unroll
    for (int k = 0; k < ES / PER; ++k) {
        const V t = q[k];
        if constexpr (PER == 2) {
            v[2 * k] = t.x;
            v[2 * k + 1] = t.y;
        } else {
            v[4 * k] = t.x;
            v[4 * k + 1] = t.y;
            v[4 * k + 2] = t.z;
            v[4 * k + 3] = t.w;
        }
    }
}

// Field gather, one lane per grid point (grid points taken in the order of
// `perm`, which puts nearby points in the same wavefront), reading particles
// and bond entries straight from global memory.
//
// Pass A scans the cells around the grid point and compacts the particles
// inside the coarse-graining ball, with their bond counts, into a per-lane list in LDS
// (one int per entry, index | count << 26, when PACK is set: fewer than 2^26
// particles and at most 63 bonds each; otherwise a second array holds the
// counts).  The list is then sorted by bond count (largest first), so that at
// every step of pass B the lanes of a wavefront hold particles with similar
// bond counts and the bond loop wastes few lanes.  Pass B adds each
// particle's kinetic term and integrates its bonds; the next bond entry and
// the next particle's data are loaded before the current ones are used, so
// memory latency overlaps the arithmetic.  A full list is processed early.
// All sums are FP64 in a fixed order.
template <int D, int OD, int CG, int NQ, bool VIR, bool ELA, typename ER, int BLK, int MAXL, bool PACK>
__global__ __launch_bounds__(BLK) void k_gather(Geom g, int n, int G, const double *__restrict__ grid,
                                                const int *__restrict__ perm, GLRule R,
                                                const double *__restrict__ gld, double cgnorm,
                                                const int *__restrict__ cstart, const double *__restrict__ spos,
                                                const double *__restrict__ skin, int cap,
                                                const int *__restrict__ bcount, const ER *__restrict__ bonds,
                                                GatherOut out)
{
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int KN = OD * (OD + 1) / 2;
    constexpr bool PAIRS = VIR || ELA;
    using LY = BondLayout<D, OD, ELA, ER>;
    constexpr int ES = LY::ES;    // values per entry
    constexpr int GRP = LY::GRP;  // entries per read group
    constexpr int GV = LY::GV;    // values per read group
    __shared__ int lp[MAXL][BLK];
    __shared__ unsigned short lnb[(PAIRS && !PACK) ? MAXL : 1][BLK];

    const int f = blockIdx.y;
    const int lane = threadIdx.x;
    // list entry k of this lane: particle index and bond count
    auto getp = [&](int k) -> int { return PACK ? (lp[k][lane] & 0x3FFFFFF) : lp[k][lane]; };
    auto getnb = [&](int k) -> int {
        if constexpr (!PAIRS) return 0;
        else if constexpr (PACK) return (int)((unsigned)lp[k][lane] >> 26);
        else return lnb[k][lane];
    };
    const int t = xcd_tile(blockIdx.x, gridDim.x) * BLK + lane;
    const bool live = t < G;
    const bool KIN = skin != nullptr;
    const double *sp = spos + (size_t)f * n * PS;
    const int *st = cstart + (size_t)f * (g.ncell + 1);
    const int *bc = bcount + (size_t)f * n;
    const ER *bd = bonds + (size_t)f * n * (size_t)cap * ES;
    const double *sk = KIN ? skin + (size_t)f * n * KN : nullptr;

    double av[KN], ae[KN], ak[KN];
#pragma unroll
    for (int k = 0; k < KN; ++k) { av[k] = 0.0; ae[k] = 0.0; ak[k] = 0.0; }
    unsigned long long nball = 0, ncontrib = 0;
    const int gp = live ? perm[t] : 0;

    if (live) {
        const double x0 = grid[(size_t)gp * 3], x1 = grid[(size_t)gp * 3 + 1];
        const double x2 = (D == 3) ? grid[(size_t)gp * 3 + 2] : 0.0;
        const double s2 = 2.0 * g.inv_rc;

        // integrate the entry at offset O (a std::integral_constant) of a
        // read group
        auto integrate = [&](const ER (&G)[GV], auto O, ER ta0, ER ta1, ER ta2, ER A) {
            constexpr int o = decltype(O)::value;
            const ER *e = G + o;
            ER B = ta0 * e[0] + ta1 * e[1];
            if (D == 3) B = fma(ta2, e[2], B);
            ER C;
            if constexpr (LY::HAS_C2) {
                C = e[D];
            } else {
                C = e[0] * e[0];
                C = fma(e[1], e[1], C);
                if (D == 3) C = fma(e[2], e[2], C);
            }
            const ER w = bond_quad<CG, NQ, ER>(R, gld, R.n, A, B, C);
            if (VIR) {
                const ER c = w * e[LY::OC];
                const ER cb0 = c * e[0], cb1 = c * e[1];
                if constexpr (std::is_same<ER, double>::value) {
                    av[0] = fma(cb0, e[0], av[0]);
                    av[1] = fma(cb1, e[1], av[1]);
                    if (OD == 2) {
                        av[2] = fma(cb0, e[1], av[2]);
                    } else {
                        const ER cb2 = c * e[2];
                        av[2] = fma(cb2, e[2], av[2]);
                        av[3] = fma(cb0, e[1], av[3]);
                        av[4] = fma(cb0, e[2], av[4]);
                        av[5] = fma(cb1, e[2], av[5]);
                    }
                } else {
                    av[0] += (double)(cb0 * e[0]);
                    av[1] += (double)(cb1 * e[1]);
                    if (OD == 2) {
                        av[2] += (double)(cb0 * e[1]);
                    } else {
                        const ER cb2 = c * e[2];
                        av[2] += (double)(cb2 * e[2]);
                        av[3] += (double)(cb0 * e[1]);
                        av[4] += (double)(cb0 * e[2]);
                        av[5] += (double)(cb1 * e[2]);
                    }
                }
            }
            if (ELA) {
#pragma unroll
                for (int m = 0; m < KN; ++m) {
                    if constexpr (std::is_same<ER, double>::value) ae[m] = fma(w, e[LY::OE + m], ae[m]);
                    else ae[m] += (double)(w * e[LY::OE + m]);
                }
            }
        };
        // the entries of one group that exist (k-th group of a particle with nb)
        auto integrate_group = [&](const ER (&G)[GV], int k, int nb, ER ta0, ER ta1, ER ta2, ER A) {
            integrate(G, std::integral_constant<int, 0>(), ta0, ta1, ta2, A);
            if constexpr (GRP == 2)
                if (2 * k + 1 < nb) integrate(G, std::integral_constant<int, ES>(), ta0, ta1, ta2, A);
        };

        // pass B over the first cnt list entries
        auto process = [&](int cnt) {
            if constexpr (PAIRS) {
                // stable insertion sort by bond count, largest first
                for (int k = 1; k < cnt; ++k) {
                    const int ve = lp[k][lane];
                    const int vn = getnb(k);
                    const unsigned short vs = PACK ? 0 : lnb[k][lane];
                    int m = k - 1;
                    while (m >= 0 && getnb(m) < vn) {
                        lp[m + 1][lane] = lp[m][lane];
                        if constexpr (!PACK) lnb[m + 1][lane] = lnb[m][lane];
                        --m;
                    }
                    lp[m + 1][lane] = ve;
                    if constexpr (!PACK) lnb[m + 1][lane] = vs;
                }
            }
            // particle data and first bond entry of list entry q
            auto fetch = [&](int q, double (&ps)[PS], double (&kv)[KN], ER (&e0)[GV]) {
                const int pp = getp(q);
#pragma unroll
                for (int a = 0; a < PS; ++a) ps[a] = sp[(size_t)pp * PS + a];
                if (KIN)
#pragma unroll
                    for (int k = 0; k < KN; ++k) kv[k] = sk[(size_t)pp * KN + k];
                if constexpr (PAIRS) load_entry<ER, GV>(bd + (size_t)pp * cap * ES, e0);
            };
            double pos[PS], kk[KN];
            ER E0[GV];
            fetch(0, pos, kk, E0);
            for (int q = 0; q < cnt; ++q) {
                const int p = getp(q);
                // the next particle's data, loaded ahead
                double npos[PS], nkk[KN];
                ER nE0[GV];
                fetch(q + 1 < cnt ? q + 1 : q, npos, nkk, nE0);

                const double a0 = minimg(pos[0] - x0, g.L[0], g.invL[0], g.periodic[0]);
                const double a1 = minimg(pos[1] - x1, g.L[1], g.invL[1], g.periodic[1]);
                const double a2 = (D == 3) ? minimg(pos[2] - x2, g.L[2], g.invL[2], g.periodic[2]) : 0.0;
                const double A = (a0 * a0 + a1 * a1 + a2 * a2) * g.inv_rc2;
                if (KIN) {
                    const double w = shape<CG, double>(A);
#pragma unroll
                    for (int k = 0; k < KN; ++k) ak[k] = fma(w, kk[k], ak[k]);
                }
                if constexpr (PAIRS) {
                    const int nb = getnb(q);
                    ncontrib += nb;
                    const ER ta0 = (ER)(a0 * s2), ta1 = (ER)(a1 * s2), ta2 = (ER)(a2 * s2), Ae = (ER)A;
                    const ER *ptr = bd + (size_t)p * cap * ES;
                    // read groups k and k+1 alternate between E0 and E1, each
                    // loaded one step before it is used
                    const int ng = (nb + GRP - 1) / GRP;
                    ER E1[GV];
                    int k = 0;
                    for (; k + 1 < ng; k += 2) {
                        load_entry<ER, GV>(ptr + (size_t)(k + 1) * GV, E1);
                        integrate_group(E0, k, nb, ta0, ta1, ta2, Ae);
                        load_entry<ER, GV>(ptr + (size_t)(k + 2 < ng ? k + 2 : k + 1) * GV, E0);
                        integrate_group(E1, k + 1, nb, ta0, ta1, ta2, Ae);
                    }
                    if (k < ng) integrate_group(E0, k, nb, ta0, ta1, ta2, Ae);
                }
#pragma unroll
                for (int a = 0; a < PS; ++a) pos[a] = npos[a];
                if (KIN)
#pragma unroll
                    for (int k = 0; k < KN; ++k) kk[k] = nkk[k];
                if constexpr (PAIRS)
#pragma unroll
                    for (int k = 0; k < GV; ++k) E0[k] = nE0[k];
            }
        };

        // pass A: candidates in batches of NB, the batch's loads issued
        // together (a load whose result the next step needs would hold the
        // lane for a full memory latency per candidate)
        constexpr int NB = 8;
        static_assert(MAXL >= NB, "list shorter than a batch");
        const int cx = cell_coord(g, 0, x0), cy = cell_coord(g, 1, x1), cz = cell_coord(g, 2, x2);
        int rows[9];
        const int nrows = neighbour_rows(g, cy, cz, rows);
        int cnt = 0;
        for (int r = 0; r < nrows; ++r) {
            int rb[2], re[2];
            const int nr = row_ranges(g, st, rows[r], cx, rb, re);
            for (int s = 0; s < nr; ++s) {
                for (int p0 = rb[s]; p0 < re[s]; p0 += NB) {
                    if (cnt > MAXL - NB) {  // room for a whole batch
                        process(cnt);
                        cnt = 0;
                    }
                    double q[NB][PS];
                    int qb[NB];
#pragma unroll
                    for (int u = 0; u < NB; ++u) {
                        const int p = min(p0 + u, re[s] - 1);
#pragma unroll
                        for (int a = 0; a < PS; ++a) q[u][a] = sp[(size_t)p * PS + a];
                        if constexpr (PAIRS) qb[u] = bc[p];
                    }
#pragma unroll
                    for (int u = 0; u < NB; ++u) {
                        const int p = p0 + u;
                        if (p >= re[s]) break;
                        const double a0 = minimg(q[u][0] - x0, g.L[0], g.invL[0], g.periodic[0]);
                        const double a1 = minimg(q[u][1] - x1, g.L[1], g.invL[1], g.periodic[1]);
                        const double a2 = (D == 3) ? minimg(q[u][2] - x2, g.L[2], g.invL[2], g.periodic[2]) : 0.0;
                        const double asq = a0 * a0 + a1 * a1 + a2 * a2;
                        if (!(asq < g.rc2)) continue;
                        ++nball;
                        if constexpr (PAIRS) {
                            const int nb = min(qb[u], cap);
                            if constexpr (PACK) {
                                lp[cnt][lane] = p | (nb << 26);
                            } else {
                                lp[cnt][lane] = p;
                                lnb[cnt][lane] = (unsigned short)nb;
                            }
                        } else {
                            lp[cnt][lane] = p;
                        }
                        ++cnt;
                    }
                }
            }
        }
        if (cnt > 0) process(cnt);
    }

    if (live) {
        const size_t ob = ((size_t)f * G + gp) * OD * OD;
        auto store = [&](double *dst, const double *acc, double scale) {
            for (int a = 0; a < OD; ++a) dst[ob + a + OD * a] = acc[a] * scale;
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
    // counts: one plain store per wavefront
    nball = wave_sum(nball);
    ncontrib = wave_sum(ncontrib);
    if ((lane & 63) == 0) {
        const size_t w = ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * (BLK / 64) + (lane >> 6);
        out.wave_counts[2 * w] = nball;
        out.wave_counts[2 * w + 1] = ncontrib;
    }
}

}  // namespace pgt
