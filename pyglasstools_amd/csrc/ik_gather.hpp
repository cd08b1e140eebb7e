// ik_gather.hpp -- the field-gather kernel of the IK stress-field engine.
// Included by the gather translation units (ik_gather_*.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ik_device.hpp"
#include "ik_bonds.hpp"
#include "ik_geom.hpp"

namespace pgt {

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

// Cells along one axis of the cell grid that cover cells c_j - R .. c_j + R
// of PL points (point 0 always valid, the others where v[j]): `lo` (in
// [0, nc)) and `cnt`, taken periodically; the whole axis when cnt would reach
// nc (so no cell is listed twice).  An axis of one cell is non-periodic or
// too short for a window.
template <int PL>
__device__ __forceinline__ void axis_window(int nc, int R, const int (&c)[PL], const bool (&v)[PL], int &lo,
                                            int &cnt)
{
    if (nc == 1) {
        lo = 0;
        cnt = 1;
        return;
    }
    int dmin = 0, dmax = 0;
#pragma unroll
    for (int j = 1; j < PL; ++j) {
        if (!v[j]) continue;
        int d = c[j] - c[0];
        if (2 * d > nc) d -= nc;
        else if (2 * d <= -nc) d += nc;
        dmin = min(dmin, d);
        dmax = max(dmax, d);
    }
    cnt = dmax - dmin + 2 * R + 1;
    if (cnt >= nc) {
        lo = 0;
        cnt = nc;
        return;
    }
    lo = c[0] + dmin - R;
    if (lo < 0) lo += nc;
    else if (lo >= nc) lo -= nc;
}

// Field gather.  Each lane takes PL consecutive grid points in the order of
// `perm` (Morton order, so 2 points are x-neighbours and 4 a 2x2 block on a
// regular grid) and reads particles and bond entries straight from global
// memory.  A particle inside the coarse-graining ball of several of the
// lane's points is read once, and each of its bond entries once, for all of
// them: on a regular 2D grid with spacing r_cg/2 a lane of 4 points reads a
// particle for 2.3 points on average, which divides the scattered loads that
// bound this kernel.
//
// Pass A scans the cells around the lane's points and compacts the particles
// inside at least one ball, with their bond counts and the mask of the points
// whose ball holds them, into a per-lane list in LDS.  In 3D with bonds it
// reads one 16-byte record per candidate (FP32 position and bond count,
// written by the bond kernel) instead of 32 bytes of FP64 position and a
// separate count, and tests |a|^2 in FP32 against thresholds that bound the
// FP64 rounding (Geom::in32, out32); a candidate between them is decided by
// the FP64 test, so the particles listed are exactly those of the FP64 scan.  With PACK one int per
// entry holds index | count << IB | mask << (IB + 6) (IB = 26 for one point
// per lane, 22 otherwise: fewer than 2^IB particles and at most 63 bonds
// each); otherwise separate arrays hold the counts and masks.  The list is
// then sorted by bond count (largest first), so that at every step of pass B
// the lanes of a wavefront hold particles with similar bond counts and the
// bond loop wastes few lanes.  Pass B adds each particle's kinetic term and
// integrates its bonds for the points of its mask; the next bond entries and
// the next particle's data are loaded before the current ones are used, so
// memory latency overlaps the arithmetic.  A full list is processed early.
// All sums are FP64, per grid point in list order.
template <int D, int OD, int CG, int NQ, bool VIR, bool ELA, typename ER, int BLK, int MAXL, bool PACK, int PL>
__global__ __launch_bounds__(BLK) void k_gather(Geom g, int n, int G, const double *__restrict__ grid,
                                                const int *__restrict__ perm, GLRule R,
                                                const double *__restrict__ gld, double cgnorm,
                                                const int *__restrict__ cstart, const double *__restrict__ spos,
                                                const double *__restrict__ skin, int cap,
                                                const int *__restrict__ bcount, const float4 *__restrict__ rec,
                                                const ER *__restrict__ bonds, GatherOut out)
{
    static_assert(PL == 1 || PL == 2 || PL == 4, "1, 2 or 4 grid points per lane");
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int KN = OD * (OD + 1) / 2;
    constexpr bool PAIRS = VIR || ELA;
    using LY = BondLayout<D, OD, ELA, ER>;
    constexpr int ES = LY::ES;    // values per entry
    constexpr int GRP = LY::GRP;  // entries per read group
    constexpr int GV = LY::GV;    // values per read group
    constexpr int IB = PL > 1 ? 22 : 26;  // index bits of a packed list entry
    __shared__ int lp[MAXL][BLK];
    __shared__ unsigned short lnb[(PAIRS && !PACK) ? MAXL : 1][BLK];
    __shared__ unsigned char lmk[(PL > 1 && !PACK) ? MAXL : 1][BLK];

    const int f = blockIdx.y;
    const int lane = threadIdx.x;
    // list entry k of this lane: particle index, bond count, point mask
    auto getp = [&](int k) -> int { return PACK ? (lp[k][lane] & ((1 << IB) - 1)) : lp[k][lane]; };
    auto getnb = [&](int k) -> int {
        if constexpr (!PAIRS) return 0;
        else if constexpr (PACK) return (int)(((unsigned)lp[k][lane] >> IB) & 63u);
        else return lnb[k][lane];
    };
    auto getm = [&](int k) -> int {
        if constexpr (PL == 1) return 1;
        else if constexpr (PACK) return (int)((unsigned)lp[k][lane] >> (IB + 6));
        else return lmk[k][lane];
    };
    const int t0 = (xcd_tile(blockIdx.x, gridDim.x) * BLK + lane) * PL;  // first point slot of the lane
    const bool live = t0 < G;
    const bool KIN = skin != nullptr;
    const double *sp = spos + (size_t)f * n * PS;
    const int *st = cstart + (size_t)f * (g.ncell + 1);
    // 3D with bonds: the candidate scan reads the bond kernel's 16-byte
    // records (FP32 position, bond count) instead of the FP64 positions and
    // the counts, and tests in FP32 first
    constexpr bool RECS = PAIRS && D == 3;
    const int *bc = bcount + (size_t)f * n;
    const float4 *rc4 = RECS ? rec + (size_t)f * n : nullptr;
    const ER *bd = bonds + (size_t)f * n * (size_t)cap * ES;
    const double *sk = KIN ? skin + (size_t)f * n * KN : nullptr;

    double av[PL][KN], ae[PL][KN], ak[PL][KN];
#pragma unroll
    for (int j = 0; j < PL; ++j)
#pragma unroll
        for (int k = 0; k < KN; ++k) { av[j][k] = 0.0; ae[j][k] = 0.0; ak[j][k] = 0.0; }
    unsigned long long nball = 0, ncontrib = 0;
    int gp[PL];
    bool pv[PL];
#pragma unroll
    for (int j = 0; j < PL; ++j) {
        pv[j] = t0 + j < G;
        gp[j] = pv[j] ? perm[t0 + j] : 0;
    }

    if (live) {
        // coordinates of the lane's points; a missing point takes point 0's
        // (its mask bit is never set)
        double xg[PL][3];
#pragma unroll
        for (int j = 0; j < PL; ++j) {
            const int q = pv[j] ? gp[j] : gp[0];
            xg[j][0] = grid[(size_t)q * 3];
            xg[j][1] = grid[(size_t)q * 3 + 1];
            xg[j][2] = (D == 3) ? grid[(size_t)q * 3 + 2] : 0.0;
        }
        const double s2 = 2.0 * g.inv_rc;
        float xg32[PL][3];
#pragma unroll
        for (int j = 0; j < PL; ++j)
#pragma unroll
            for (int a = 0; a < 3; ++a) xg32[j][a] = (float)xg[j][a];

        // integrate the entry at offset O (a std::integral_constant) of a
        // read group for the points of mask m
        auto integrate = [&](const ER (&Gv)[GV], auto O, int m, const ER (&ta)[PL][3], const ER (&Aj)[PL]) {
            constexpr int o = decltype(O)::value;
            const ER *e = Gv + o;
            ER C;
            if constexpr (LY::HAS_C2) {
                C = e[D];
            } else {
                C = e[0] * e[0];
                C = fma(e[1], e[1], C);
                if (D == 3) C = fma(e[2], e[2], C);
            }
#pragma unroll
            for (int j = 0; j < PL; ++j) {
                if (PL > 1 && !((m >> j) & 1)) continue;
                ER B = ta[j][0] * e[0] + ta[j][1] * e[1];
                if (D == 3) B = fma(ta[j][2], e[2], B);
                const ER w = bond_quad<CG, NQ, ER>(R, gld, R.n, Aj[j], B, C);
                if (VIR) {
                    const ER c = w * e[LY::OC];
                    const ER cb0 = c * e[0], cb1 = c * e[1];
                    double *v = av[j];
                    if constexpr (std::is_same<ER, double>::value) {
                        v[0] = fma(cb0, e[0], v[0]);
                        v[1] = fma(cb1, e[1], v[1]);
                        if (OD == 2) {
                            v[2] = fma(cb0, e[1], v[2]);
                        } else {
                            const ER cb2 = c * e[2];
                            v[2] = fma(cb2, e[2], v[2]);
                            v[3] = fma(cb0, e[1], v[3]);
                            v[4] = fma(cb0, e[2], v[4]);
                            v[5] = fma(cb1, e[2], v[5]);
                        }
                    } else if constexpr (OD == 2) {
                        // FP32 entries, 2D: products in FP32, sums in FP64
                        // (as the tile gather)
                        v[0] += (double)(cb0 * e[0]);
                        v[1] += (double)(cb1 * e[1]);
                        v[2] += (double)(cb0 * e[1]);
                    } else {
                        // FP32 entries, 3D: c' w and b' widened, the products
                        // and sums in FP64
                        const double cd = (double)c, d0 = (double)e[0], d1 = (double)e[1];
                        const double d2 = (double)e[2];
                        const double q0 = cd * d0, q1 = cd * d1, q2 = cd * d2;
                        v[0] = fma(q0, d0, v[0]);
                        v[1] = fma(q1, d1, v[1]);
                        v[2] = fma(q2, d2, v[2]);
                        v[3] = fma(q0, d1, v[3]);
                        v[4] = fma(q0, d2, v[4]);
                        v[5] = fma(q1, d2, v[5]);
                    }
                }
                if (ELA) {
#pragma unroll
                    for (int q = 0; q < KN; ++q) {
                        if constexpr (std::is_same<ER, double>::value) ae[j][q] = fma(w, e[LY::OE + q], ae[j][q]);
                        else ae[j][q] += (double)(w * e[LY::OE + q]);
                    }
                }
            }
        };
        // the entries of one group that exist (k-th group of a particle with nb)
        auto integrate_group = [&](const ER (&Gv)[GV], int k, int nb, int m, const ER (&ta)[PL][3],
                                   const ER (&Aj)[PL]) {
            integrate(Gv, std::integral_constant<int, 0>(), m, ta, Aj);
            if constexpr (GRP >= 2)
                if (GRP * k + 1 < nb) integrate(Gv, std::integral_constant<int, ES>(), m, ta, Aj);
            if constexpr (GRP >= 3)
                if (GRP * k + 2 < nb) integrate(Gv, std::integral_constant<int, 2 * ES>(), m, ta, Aj);
            if constexpr (GRP >= 4)
                if (GRP * k + 3 < nb) integrate(Gv, std::integral_constant<int, 3 * ES>(), m, ta, Aj);
        };

        // pass B over the first cnt list entries
        auto process = [&](int cnt) {
            if constexpr (PAIRS) {
                // stable insertion sort by bond count, largest first
                for (int k = 1; k < cnt; ++k) {
                    const int ve = lp[k][lane];
                    const int vn = getnb(k);
                    const unsigned short vs = PACK ? 0 : lnb[k][lane];
                    const unsigned char vm = (PL > 1 && !PACK) ? lmk[k][lane] : 0;
                    int m = k - 1;
                    while (m >= 0 && getnb(m) < vn) {
                        lp[m + 1][lane] = lp[m][lane];
                        if constexpr (!PACK) lnb[m + 1][lane] = lnb[m][lane];
                        if constexpr (PL > 1 && !PACK) lmk[m + 1][lane] = lmk[m][lane];
                        --m;
                    }
                    lp[m + 1][lane] = ve;
                    if constexpr (!PACK) lnb[m + 1][lane] = vs;
                    if constexpr (PL > 1 && !PACK) lmk[m + 1][lane] = vm;
                }
            }
            // particle data and first bond entries of list entry q
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
                const int m = getm(q);
                // the next particle's data, loaded ahead
                double npos[PS], nkk[KN];
                ER nE0[GV];
                fetch(q + 1 < cnt ? q + 1 : q, npos, nkk, nE0);

                ER ta[PL][3], Aj[PL];
#pragma unroll
                for (int j = 0; j < PL; ++j) {
                    const double a0 = minimg(pos[0] - xg[j][0], g.L[0], g.pinvL[0]);
                    const double a1 = minimg(pos[1] - xg[j][1], g.L[1], g.pinvL[1]);
                    const double a2 = (D == 3) ? minimg(pos[2] - xg[j][2], g.L[2], g.pinvL[2]) : 0.0;
                    const double A = (a0 * a0 + a1 * a1 + a2 * a2) * g.inv_rc2;
                    if (KIN && (PL == 1 || ((m >> j) & 1))) {
                        const double w = shape<CG, double>(A);
#pragma unroll
                        for (int k = 0; k < KN; ++k) ak[j][k] = fma(w, kk[k], ak[j][k]);
                    }
                    ta[j][0] = (ER)(a0 * s2);
                    ta[j][1] = (ER)(a1 * s2);
                    ta[j][2] = (ER)(a2 * s2);
                    Aj[j] = (ER)A;
                }
                if constexpr (PAIRS) {
                    const int nb = getnb(q);
                    ncontrib += (unsigned long long)nb * (PL == 1 ? 1 : __popc(m));
                    const ER *ptr = bd + (size_t)p * cap * ES;
                    // read groups k and k+1 alternate between E0 and E1, each
                    // loaded one step before it is used
                    const int ng = (nb + GRP - 1) / GRP;
                    ER E1[GV];
                    int k = 0;
                    for (; k + 1 < ng; k += 2) {
                        load_entry<ER, GV>(ptr + (size_t)(k + 1) * GV, E1);
                        integrate_group(E0, k, nb, m, ta, Aj);
                        load_entry<ER, GV>(ptr + (size_t)(k + 2 < ng ? k + 2 : k + 1) * GV, E0);
                        integrate_group(E1, k + 1, nb, m, ta, Aj);
                    }
                    if (k < ng) integrate_group(E0, k, nb, m, ta, Aj);
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

        // pass A: the cells around the lane's points, candidates in batches
        // of NB with the batch's loads issued together (a load whose result
        // the next step needs would hold the lane for a full memory latency
        // per candidate)
        constexpr int NB = 8;
        static_assert(MAXL >= NB, "list shorter than a batch");
        int wlo[3], wn[3];
        {
            int cc[3][PL];
#pragma unroll
            for (int j = 0; j < PL; ++j) {
                cc[0][j] = cell_coord(g, 0, xg[j][0]);
                cc[1][j] = cell_coord(g, 1, xg[j][1]);
                cc[2][j] = cell_coord(g, 2, xg[j][2]);
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) axis_window<PL>(g.nc[a], g.reach, cc[a], pv, wlo[a], wn[a]);
        }
        const int nx = g.nc[0], ny = g.nc[1], nz = g.nc[2];
        int cnt = 0;
        for (int iz = 0; iz < wn[2]; ++iz) {
            const int z = wlo[2] + iz < nz ? wlo[2] + iz : wlo[2] + iz - nz;
            for (int iy = 0; iy < wn[1]; ++iy) {
                const int y = wlo[1] + iy < ny ? wlo[1] + iy : wlo[1] + iy - ny;
                const int rowbase = (z * ny + y) * nx;
                // cells wlo[0] .. wlo[0] + wn[0] - 1 of the row, wrapping
                int rb[2], re[2], nr = 1;
                const int xe = wlo[0] + wn[0];
                rb[0] = st[rowbase + wlo[0]];
                if (xe <= nx) {
                    re[0] = st[rowbase + xe];
                } else {
                    re[0] = st[rowbase + nx];
                    rb[1] = st[rowbase];
                    re[1] = st[rowbase + xe - nx];
                    nr = 2;
                }
                for (int s = 0; s < nr; ++s) {
                    for (int p0 = rb[s]; p0 < re[s]; p0 += NB) {
                        if (cnt > MAXL - NB) {  // room for a whole batch
                            process(cnt);
                            cnt = 0;
                        }
                        // exact test: FP64 minimum image and |a|^2 against r_cg^2
                        auto exact_mask = [&](const double *qq) {
                            int mm = 0;
#pragma unroll
                            for (int j = 0; j < PL; ++j) {
                                const double a0 = minimg(qq[0] - xg[j][0], g.L[0], g.pinvL[0]);
                                const double a1 = minimg(qq[1] - xg[j][1], g.L[1], g.pinvL[1]);
                                const double a2 =
                                    (D == 3) ? minimg(qq[2] - xg[j][2], g.L[2], g.pinvL[2]) : 0.0;
                                const double asq = a0 * a0 + a1 * a1 + a2 * a2;
                                if (pv[j] && asq < g.rc2) mm |= 1 << j;
                            }
                            return mm;
                        };
                        double q[RECS ? 1 : NB][PS];
                        float4 r4[RECS ? NB : 1];
                        int qb[NB];
#pragma unroll
                        for (int u = 0; u < NB; ++u) {
                            const int p = min(p0 + u, re[s] - 1);
                            if constexpr (RECS) {
                                // one 16-byte record: FP32 position and bond count
                                r4[u] = rc4[p];
                                qb[u] = __float_as_int(r4[u].w);
                            } else {
#pragma unroll
                                for (int a = 0; a < PS; ++a) q[u][a] = sp[(size_t)p * PS + a];
                                if constexpr (PAIRS) qb[u] = bc[p];
                            }
                        }
#pragma unroll
                        for (int u = 0; u < NB; ++u) {
                            const int p = p0 + u;
                            if (p >= re[s]) break;
                            int m = 0;
                            if constexpr (RECS) {
                                // FP32 pre-test; a candidate in the band
                                // around r_cg (or any, when the geometry gives
                                // no bound) is decided by the exact test
                                bool unsure = false;
#pragma unroll
                                for (int j = 0; j < PL; ++j) {
                                    float d0 = r4[u].x - xg32[j][0], d1 = r4[u].y - xg32[j][1], d2 = 0.0f;
                                    d0 = fmaf(-g.L32[0], rintf(d0 * g.pinvL32[0]), d0);
                                    d1 = fmaf(-g.L32[1], rintf(d1 * g.pinvL32[1]), d1);
                                    if (D == 3) {
                                        d2 = r4[u].z - xg32[j][2];
                                        d2 = fmaf(-g.L32[2], rintf(d2 * g.pinvL32[2]), d2);
                                    }
                                    const float a2f = d0 * d0 + d1 * d1 + d2 * d2;
                                    if (pv[j]) {
                                        if (a2f < g.in32) m |= 1 << j;
                                        else if (!(a2f > g.out32)) unsure = true;
                                    }
                                }
                                if (unsure) {
                                    double qq[PS];
#pragma unroll
                                    for (int a = 0; a < PS; ++a) qq[a] = sp[(size_t)p * PS + a];
                                    m = exact_mask(qq);
                                }
                            } else {
                                m = exact_mask(q[u]);
                            }
                            if (!m) continue;
                            nball += PL == 1 ? 1 : __popc(m);
                            const int nb = PAIRS ? min(qb[u], cap) : 0;
                            if constexpr (PACK) {
                                lp[cnt][lane] = (int)((unsigned)p | ((unsigned)nb << IB) |
                                                      (PL > 1 ? ((unsigned)m << (IB + 6)) : 0u));
                            } else {
                                lp[cnt][lane] = p;
                                if constexpr (PAIRS) lnb[cnt][lane] = (unsigned short)nb;
                                if constexpr (PL > 1) lmk[cnt][lane] = (unsigned char)m;
                            }
                            ++cnt;
                        }
                    }
                }
            }
        }
        if (cnt > 0) process(cnt);
    }

    if (live) {
        auto store = [&](double *dst, size_t ob, const double *acc, double scale) {
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
#pragma unroll
        for (int j = 0; j < PL; ++j) {
            if (!pv[j]) continue;
            const size_t ob = ((size_t)f * G + gp[j]) * OD * OD;
            if (VIR) store(out.vir, ob, av[j], 1.0);
            if (ELA) store(out.ela, ob, ae[j], 1.0);
            if (KIN) store(out.kin, ob, ak[j], cgnorm);
        }
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
