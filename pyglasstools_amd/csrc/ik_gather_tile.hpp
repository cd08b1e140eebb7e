// ik_gather_tile.hpp -- the item-balanced field gather ("tile" gather) of
// the IK stress-field engine.  Included by ik_gather_launch.hpp.
//
// A workgroup is one wavefront and takes 64 consecutive grid points in Morton
// order (a compact tile of the grid), one per lane.  Work items are the
// (grid point, particle) pairs with the particle inside the point's ball;
// an item integrates the particle's bonds for the point.
//
// 1. Scan: each lane walks the cells around its point, as the lane gather
//    does, in steps of NB candidates that every lane takes together, and
//    tests them exactly (FP64 minimum image, |a|^2 < r_cg^2; in 3D with
//    bonds an FP32 test on the bond kernel's records first, with the FP64
//    test for the band where FP32 rounding could decide).  The items found
//    go to a list shared by the wavefront (ballot, prefix count).  A 2D
//    scan reads candidate positions only; the bond counts of the items found
//    are loaded before the sort, a batch per lane (in 3D the record holds
//    the count).
// 2. When the list could overflow, and at the end: the list is sorted by
//    the particle's bond count, largest first (stable counting sort), and
//    lane l takes items l, l + 64, ...: the 64 items of a step have nearly
//    equal bond counts, so the bond loop wastes few lanes whatever the
//    spread of the counts between points.  The next step's particle data and
//    first bond entries are loaded while the current step integrates.  A lane
//    sums its item in registers and adds it to the point's FP64 accumulators
//    in LDS.
//
// The wavefront runs alone in its workgroup and every step is in a fixed
// order, so the sums are the same from run to run.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ik_bonds.hpp"
#include "ik_device.hpp"
#include "ik_gather.hpp"
#include "ik_geom.hpp"
#include "ik_launch.hpp"
#include "ik_sweep.hpp"

namespace pgt {

// Capacities of the tile gather.
struct TileCaps {
    static constexpr int BLK = kTileThreads;  // one wavefront
    static constexpr int TP = kTilePoints;    // grid points per tile, one per lane
    static constexpr int MAXI = 768;          // items in the list
    static constexpr int NB = 4;              // candidates per lane per scan step
    static constexpr int PBITS = 26;          // item word: particle index | point << PBITS
    static constexpr int MAXNB = 255;         // bond count of an item (one byte); more: overflow
    static_assert(TP == 64 && BLK == 64, "one wavefront of 64 points");
    static_assert(MAXI >= 2 * 64 * NB, "the list holds two scan steps");
};

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// Wavefronts per SIMD the compiler sizes registers for: four for the 2D forms
// without elastic terms (a few registers spilled at most) and for the 3D FP32
// virial form (126 registers, no spill; 131 and three waves unbounded, with
// the gather 7% slower on the 3D N = 131,072 job), which hides more memory
// latency; the others are left to the compiler (two or three: the 3D FP64
// virial form would spill 88 bytes a lane at four).
template <int D, bool ELA, bool KIN, typename ER>
struct TileWaves {
    static constexpr bool F32 = std::is_same<ER, float>::value;
    static constexpr int value = ((D == 2 && !ELA) || (D == 3 && F32 && !ELA && !KIN)) ? 4 : 1;
};

// The body works on one frame's arrays (cell starts st, sorted positions
// sp, kinetic terms sk, bond counts bc, FP32 records rf, entries bd; out's
// field pointers at the frame's first value) for the tile-th 64 points of
// the Morton order perm, and stores its counts at out.wave_counts[2 wi ..].
template <int D, int OD, int CG, int NQ, bool VIR, bool ELA, bool KIN, typename ER>
__device__ __forceinline__ void gather_tile_body(const Geom &g, int n, int G, const double *__restrict__ grid,
                                                 const int *__restrict__ perm, const GLRule &R,
                                                 const double *__restrict__ gld, double cgnorm,
                                                 const int *__restrict__ st, const double *__restrict__ sp,
                                                 const double *__restrict__ sk, int cap,
                                                 const int *__restrict__ bc, const float4 *__restrict__ rf,
                                                 const ER *__restrict__ bd, const GatherOut &out, size_t wi, int tile,
                                                 int *__restrict__ overflow)
{
    using CP = TileCaps;
    using LY = BondLayout<D, OD, ELA, ER>;
    constexpr int MAXI = CP::MAXI, NB = CP::NB, PBITS = CP::PBITS;
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int KN = OD * (OD + 1) / 2;
    constexpr bool PAIRS = VIR || ELA;
    constexpr int KV = (VIR ? KN : 0) + (ELA ? KN : 0);  // pair accumulators per point
    constexpr int KT = KV + (KIN ? KN : 0);             // all accumulators per point
    constexpr int KTA = KT | 1;  // accumulator stride: odd, so 64 points fall on distinct LDS banks
    constexpr int KNA = KIN ? KN : 1;
    constexpr int ES = LY::ES;    // values per entry
    constexpr int GRP = LY::GRP;  // entries per read group
    constexpr int GV = LY::GV;    // values per read group
    constexpr unsigned PMASK = (1u << PBITS) - 1u;
    static_assert(((int64_t)1 << PBITS) == kTileMaxParticles, "item word: particle bits");

    __shared__ double s_pt[64][D];            // the tile's points
    __shared__ double s_acc[64][KTA];         // their accumulators
    __shared__ unsigned s_item[MAXI];         // items in scan order
    __shared__ unsigned char s_inb[MAXI];     // their bond counts
    __shared__ unsigned short s_perm[MAXI];   // the items in bond-count order
    __shared__ int s_bucket[64];

    const int l = threadIdx.x;
    const int t0 = tile * 64;
    const int t = t0 + l;
    const bool live = t < G;
    // 3D with bonds: the scan reads the bond kernel's 16-byte records (FP32
    // position, bond count) and tests in FP32 first, as the lane gather does
    constexpr bool RECS = PAIRS && D == 3;

    const int gp = perm[live ? t : t0];
    double xg[3];
    xg[0] = grid[(size_t)gp * 3];
    xg[1] = grid[(size_t)gp * 3 + 1];
    xg[2] = (D == 3) ? grid[(size_t)gp * 3 + 2] : 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) s_pt[l][a] = xg[a];
#pragma unroll
    for (int k = 0; k < KTA; ++k) s_acc[l][k] = 0.0;

    unsigned long long nball = 0;  // items listed (the same in every lane)
    unsigned long long ncontrib = 0;  // bond counts of the items, per lane
    bool over = false;  // an item's bond count exceeds MAXNB
    const double s2 = 2.0 * g.inv_rc;

    // ---- sort the first ni items of the list by bond count and integrate --
    auto flush = [&](int ni) {
        // counting sort by bond count, largest first: bucket sizes, their
        // offsets, then each item takes the next slot of its bucket (an LDS
        // atomic; the lanes of one instruction are served in a fixed order)
        s_bucket[l] = 0;
        __syncthreads();
        nball += (unsigned long long)ni;
        if constexpr (PAIRS && !RECS) {
            // the items' bond counts, loaded here U at a time rather than
            // with every scanned candidate (2D scans read positions only)
            constexpr int U = 8;
            for (int k0 = l; k0 < ni; k0 += 64 * U) {
                int nbv[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int k = k0 + 64 * u;
                    nbv[u] = k < ni ? bc[s_item[k] & PMASK] : 0;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int k = k0 + 64 * u;
                    if (k < ni) {
                        const int nbc = min(nbv[u], cap);
                        if (nbc > CP::MAXNB) over = true;
                        const int nbk = min(nbc, CP::MAXNB);
                        s_inb[k] = (unsigned char)nbk;
                        ncontrib += (unsigned long long)nbk;
                        atomicAdd(&s_bucket[63 - min(nbk, 63)], 1);
                    }
                }
            }
        } else {
            for (int k = l; k < ni; k += 64) {
                const int nbk = s_inb[k];
                ncontrib += (unsigned long long)nbk;
                atomicAdd(&s_bucket[63 - min(nbk, 63)], 1);
            }
        }
        __syncthreads();
        {
            const int c = s_bucket[l];
            int v = c;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int u = __shfl_up(v, off, 64);
                if (l >= off) v += u;
            }
            __syncthreads();
            s_bucket[l] = v - c;  // first slot of the bucket
        }
        __syncthreads();
        for (int k = l; k < ni; k += 64) {
            const int slot = atomicAdd(&s_bucket[63 - min((int)s_inb[k], 63)], 1);
            s_perm[slot] = (unsigned short)k;
        }
        __syncthreads();

        // lane l takes sorted items l, l + 64, ...; the data of the next one
        // is loaded before the current one is integrated
        auto fetch = [&](int k, int &p, int &ip, int &nb, double (&pos)[3], double (&kk)[KNA], ER (&E)[GV]) {
            const bool valid = k < ni;
            const int ix = s_perm[valid ? k : 0];
            const unsigned it = s_item[ix];
            p = (int)(it & PMASK);
            ip = (int)(it >> PBITS);
            nb = (PAIRS && valid) ? (int)s_inb[ix] : 0;
            pos[0] = sp[(size_t)p * PS];
            pos[1] = sp[(size_t)p * PS + 1];
            pos[2] = (D == 3) ? sp[(size_t)p * PS + 2] : 0.0;
            if constexpr (KIN)
#pragma unroll
                for (int q = 0; q < KN; ++q) kk[q] = sk[(size_t)p * KN + q];
            if constexpr (PAIRS)
                if (nb > 0) load_entry<ER, GV>(bd + (size_t)p * cap * ES, E);
        };
        int p, ip, nb;
        double pos[3], kk[KNA];
        ER E0[GV];
        fetch(l, p, ip, nb, pos, kk, E0);
        for (int k0 = 0; k0 < ni; k0 += 64) {
            const int k = k0 + l;
            int np_, nip, nnb;
            double npos[3], nkk[KNA];
            ER nE0[GV];
            fetch(k + 64, np_, nip, nnb, npos, nkk, nE0);
            if (k < ni) {
                double a[3];
                a[0] = minimg(pos[0] - s_pt[ip][0], g.L[0], g.pinvL[0]);
                a[1] = minimg(pos[1] - s_pt[ip][1], g.L[1], g.pinvL[1]);
                a[2] = 0.0;
                if constexpr (D == 3) a[2] = minimg(pos[2] - s_pt[ip][D - 1], g.L[2], g.pinvL[2]);
                const double A = (a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) * g.inv_rc2;
                double acc[KTA];
#pragma unroll
                for (int q = 0; q < KTA; ++q) acc[q] = 0.0;
                if constexpr (KIN) {
                    const double wk = shape<CG, double>(A);
#pragma unroll
                    for (int q = 0; q < KN; ++q) acc[KV + q] = wk * kk[q];
                }
                if constexpr (PAIRS) {
                    ER ta[3];
                    ta[0] = (ER)(a[0] * s2);
                    ta[1] = (ER)(a[1] * s2);
                    ta[2] = (ER)(a[2] * s2);
                    const ER Aj = (ER)A;
                    auto one = [&](const ER (&Gv)[GV], auto O) {
                        constexpr int o = decltype(O)::value;
                        const ER *e = Gv + o;
                        ER C = e[0] * e[0];
                        C = fma(e[1], e[1], C);
                        if (D == 3) C = fma(e[2], e[2], C);
                        ER B = ta[0] * e[0] + ta[1] * e[1];
                        if (D == 3) B = fma(ta[2], e[2], B);
                        const ER wq = bond_quad<CG, NQ, ER>(R, gld, R.n, Aj, B, C);
                        if constexpr (VIR) {
                            const ER c = wq * e[LY::OC];
                            const ER cb0 = c * e[0], cb1 = c * e[1];
                            if constexpr (std::is_same<ER, double>::value) {
                                acc[0] = fma(cb0, e[0], acc[0]);
                                acc[1] = fma(cb1, e[1], acc[1]);
                                if constexpr (OD == 2) {
                                    acc[2] = fma(cb0, e[1], acc[2]);
                                } else {
                                    const ER cb2 = c * e[2];
                                    acc[2] = fma(cb2, e[2], acc[2]);
                                    acc[3] = fma(cb0, e[1], acc[3]);
                                    acc[4] = fma(cb0, e[2], acc[4]);
                                    acc[5] = fma(cb1, e[2], acc[5]);
                                }
                            } else if constexpr (OD == 2) {
                                // FP32 entries, 2D: products in FP32, sums
                                // in FP64 (the registers of the 3D form
                                // below would spill here)
                                acc[0] += (double)(cb0 * e[0]);
                                acc[1] += (double)(cb1 * e[1]);
                                acc[2] += (double)(cb0 * e[1]);
                            } else {
                                // FP32 entries, 3D: c' w and b' widened, the
                                // products and sums in FP64
                                const double cd = (double)c, d0 = (double)e[0], d1 = (double)e[1];
                                const double d2 = (double)e[2];
                                const double q0 = cd * d0, q1 = cd * d1, q2 = cd * d2;
                                acc[0] = fma(q0, d0, acc[0]);
                                acc[1] = fma(q1, d1, acc[1]);
                                acc[2] = fma(q2, d2, acc[2]);
                                acc[3] = fma(q0, d1, acc[3]);
                                acc[4] = fma(q0, d2, acc[4]);
                                acc[5] = fma(q1, d2, acc[5]);
                            }
                        }
                        if constexpr (ELA) {
                            constexpr int EO = VIR ? KN : 0;
#pragma unroll
                            for (int q = 0; q < KN; ++q) {
                                if constexpr (std::is_same<ER, double>::value)
                                    acc[EO + q] = fma(wq, e[LY::OE + q], acc[EO + q]);
                                else acc[EO + q] += (double)(wq * e[LY::OE + q]);
                            }
                        }
                    };
                    auto group = [&](const ER (&Gv)[GV], int kk2) {
                        one(Gv, std::integral_constant<int, 0>());
                        if constexpr (GRP >= 2)
                            if (GRP * kk2 + 1 < nb) one(Gv, std::integral_constant<int, ES>());
                        if constexpr (GRP >= 3)
                            if (GRP * kk2 + 2 < nb) one(Gv, std::integral_constant<int, 2 * ES>());
                        if constexpr (GRP >= 4)
                            if (GRP * kk2 + 3 < nb) one(Gv, std::integral_constant<int, 3 * ES>());
                    };
                    const ER *ptr = bd + (size_t)p * cap * ES;
                    const int ng = (nb + GRP - 1) / GRP;
                    ER E1[GV];
                    int kk2 = 0;
                    for (; kk2 + 1 < ng; kk2 += 2) {
                        load_entry<ER, GV>(ptr + (size_t)(kk2 + 1) * GV, E1);
                        group(E0, kk2);
                        if (kk2 + 2 < ng) load_entry<ER, GV>(ptr + (size_t)(kk2 + 2) * GV, E0);
                        group(E1, kk2 + 1);
                    }
                    if (kk2 < ng) group(E0, kk2);
                }
#pragma unroll
                for (int q = 0; q < KT; ++q)
                    __hip_atomic_fetch_add(&s_acc[ip][q], acc[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            }
            p = np_;
            ip = nip;
            nb = nnb;
#pragma unroll
            for (int a = 0; a < 3; ++a) pos[a] = npos[a];
            if constexpr (KIN)
#pragma unroll
                for (int q = 0; q < KN; ++q) kk[q] = nkk[q];
            if constexpr (PAIRS)
#pragma unroll
                for (int q = 0; q < GV; ++q) E0[q] = nE0[q];
        }
        __syncthreads();
    };

    // ---- scan: the cells around each lane's point ---------------------------
    // per axis the window is cells c - reach .. c + reach, or the whole axis
    // when that would reach around it (the same count for every lane)
    int wlo[3], wn[3];
    for (int a = 0; a < 3; ++a) {
        const int nc = g.nc[a];
        const int c = cell_coord(g, a, xg[a]);
        if (nc == 1 || 2 * g.reach + 1 >= nc) {
            wlo[a] = 0;
            wn[a] = nc;
        } else {
            wn[a] = 2 * g.reach + 1;
            int lo = c - g.reach;
            if (lo < 0) lo += nc;
            wlo[a] = lo;
        }
    }
    const int nx = g.nc[0], ny = g.nc[1], nz = g.nc[2];
    const int nrows = wn[1] * wn[2];
    // row r of the window: cells wlo[0] .. wlo[0] + wn[0] - 1, as one or two
    // particle ranges read as one run [b0, b0 + len0) + [b1, b1 + len - len0)
    auto row = [&](int r, int &b0, int &len0, int &b1, int &len) {
        const int iz = r / wn[1], iy = r - iz * wn[1];
        const int z = wlo[2] + iz < nz ? wlo[2] + iz : wlo[2] + iz - nz;
        const int y = wlo[1] + iy < ny ? wlo[1] + iy : wlo[1] + iy - ny;
        const int rowbase = (z * ny + y) * nx;
        const int xe = wlo[0] + wn[0];
        b0 = st[rowbase + wlo[0]];
        int e0, e1 = 0;
        b1 = 0;
        if (xe <= nx) {
            e0 = st[rowbase + xe];
        } else {
            e0 = st[rowbase + nx];
            b1 = st[rowbase];
            e1 = st[rowbase + xe - nx];
        }
        len0 = e0 - b0;
        len = live ? len0 + (e1 - b1) : 0;
    };
    // candidates c0 .. c0 + NB - 1 of a row (clamped into it; a lane with an
    // empty row reads particle 0, and every lane has one when any reads)
    struct Cand {
        int p[NB], nb[RECS ? NB : 1];
        double q[RECS ? 1 : NB][3];
        float r[RECS ? NB : 1][3];
    };
    auto load = [&](int c0, int b0, int len0, int b1, int len, Cand &C) {
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int c = min(c0 + u, len - 1);
            C.p[u] = len == 0 ? 0 : (c < len0 ? b0 + c : b1 + (c - len0));
            if constexpr (RECS) {
                const float4 r4 = rf[C.p[u]];
                C.r[u][0] = r4.x;
                C.r[u][1] = r4.y;
                C.r[u][2] = r4.z;
                C.nb[u] = __float_as_int(r4.w);
            } else {
                C.q[u][0] = sp[(size_t)C.p[u] * PS];
                C.q[u][1] = sp[(size_t)C.p[u] * PS + 1];
                C.q[u][2] = (D == 3) ? sp[(size_t)C.p[u] * PS + 2] : 0.0;
            }
        }
    };
    float xg32[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) xg32[a] = (float)xg[a];
    // the exact test: FP64 minimum image and |a|^2 against r_cg^2
    auto exact_in = [&](const double *q) {
        const double a0 = minimg(q[0] - xg[0], g.L[0], g.pinvL[0]);
        const double a1 = minimg(q[1] - xg[1], g.L[1], g.pinvL[1]);
        const double a2 = (D == 3) ? minimg(q[2] - xg[2], g.L[2], g.pinvL[2]) : 0.0;
        return a0 * a0 + a1 * a1 + a2 * a2 < g.rc2;
    };
    // the scan visits rows r = 0 .. nrows - 1, each in steps of NB candidates
    // per lane (one step for a row that is empty in every lane)
    auto test_append = [&](const Cand &C, int c0, int len, int &ni) {
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            bool in;
            if constexpr (RECS) {
                // FP32 pre-test against bounds of the FP64 rounding; a
                // candidate in the band between them (or any, when the
                // geometry gives no bound) is decided by the exact test
                float d0 = C.r[u][0] - xg32[0], d1 = C.r[u][1] - xg32[1], d2 = C.r[u][2] - xg32[2];
                d0 = fmaf(-g.L32[0], rintf(d0 * g.pinvL32[0]), d0);
                d1 = fmaf(-g.L32[1], rintf(d1 * g.pinvL32[1]), d1);
                d2 = fmaf(-g.L32[2], rintf(d2 * g.pinvL32[2]), d2);
                const float a2f = d0 * d0 + d1 * d1 + d2 * d2;
                in = a2f < g.in32;
                if (!in && !(a2f > g.out32) && c0 + u < len) {
                    double q[3];
#pragma unroll
                    for (int a = 0; a < 3; ++a) q[a] = sp[(size_t)C.p[u] * PS + a];
                    in = exact_in(q);
                }
                in = in && c0 + u < len;
            } else {
                in = c0 + u < len && exact_in(C.q[u]);
            }
            const unsigned long long bm = __builtin_amdgcn_ballot_w64(in);
            if (in) {
                const int k = ni + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bm >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((unsigned)bm, 0u));
                s_item[k] = (unsigned)C.p[u] | ((unsigned)l << PBITS);
                if constexpr (RECS) {
                    const int nbc = min(C.nb[u], cap);
                    if (nbc > CP::MAXNB) over = true;
                    s_inb[k] = (unsigned char)min(nbc, CP::MAXNB);
                } else if constexpr (!PAIRS) {
                    s_inb[k] = 0;
                }
            }
            ni += __popcll(bm);
        }
    };
    int ni = 0;  // items in the list (the same in every lane)
    int r = 0, c0 = 0;
    int b0, len0, b1, len, nb0 = 0, nlen0 = 0, nb1 = 0, nlen = 0;
    row(0, b0, len0, b1, len);
    int maxlen = wave_max(len);
    if (nrows > 1) row(1, nb0, nlen0, nb1, nlen);  // the next row's ranges load ahead
    auto advance = [&](bool same) {
        if (same) {
            c0 += NB;
            return;
        }
        ++r;
        c0 = 0;
        b0 = nb0;
        len0 = nlen0;
        b1 = nb1;
        len = nlen;
        maxlen = wave_max(len);
        if (r + 1 < nrows) row(r + 1, nb0, nlen0, nb1, nlen);
    };
    Cand C0;
    if (n > 0) load(0, b0, len0, b1, len, C0);
    for (; n > 0;) {
        const bool same = c0 + NB < maxlen;
        const bool last = !same && r + 1 >= nrows;
        // the next step's candidates load while these are tested (loaded
        // again after a flush, which needs the registers)
        Cand C1;
        if (!last) {
            if (same) load(c0 + NB, b0, len0, b1, len, C1);
            else load(0, nb0, nlen0, nb1, nlen, C1);
        }
        test_append(C0, c0, len, ni);
        if (last) break;
        advance(same);
        if (ni > MAXI - 64 * NB) {
            flush(ni);
            ni = 0;
            load(c0, b0, len0, b1, len, C1);
        }
        C0 = C1;
    }
    if (ni > 0) flush(ni);
    __syncthreads();

    if (live) {
        const size_t ob = (size_t)gp * OD * OD;
        double av[KN], ae[KN], ak[KN];
#pragma unroll
        for (int k = 0; k < KN; ++k) {
            av[k] = VIR ? s_acc[l][k] : 0.0;
            ae[k] = ELA ? s_acc[l][(VIR ? KN : 0) + k] : 0.0;
            ak[k] = KIN ? s_acc[l][KV + k] : 0.0;
        }
        auto store = [&](double *dst, const double *acc, double scale) {
            for (int a = 0; a < OD; ++a) dst[ob + a + OD * a] = acc[a] * scale;
            if constexpr (OD == 2) {
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
    ncontrib = wave_sum(ncontrib);
    if (l == 0) {
        out.wave_counts[2 * wi] = nball;  // uniform: not summed over lanes
        out.wave_counts[2 * wi + 1] = ncontrib;
    }
    if (over) atomicOr(overflow, 2);
}

// One frame per y row; block x takes the tile xcd_tile(x, gridDim.x), so
// each XCD works through a contiguous eighth of the frame's points.
template <int D, int OD, int CG, int NQ, bool VIR, bool ELA, bool KIN, typename ER>
__global__ __launch_bounds__(64, (TileWaves<D, ELA, KIN, ER>::value)) void k_gather_tile(
    Geom g, int n, int G, const double *__restrict__ grid, const int *__restrict__ perm, GLRule R,
    const double *__restrict__ gld, double cgnorm, const int *__restrict__ cstart, const double *__restrict__ spos,
    const double *__restrict__ skin, int cap, const int *__restrict__ bcount, const float4 *__restrict__ rec,
    const ER *__restrict__ bonds, GatherOut out, int *__restrict__ overflow)
{
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int KN = OD * (OD + 1) / 2;
    constexpr int ES = BondLayout<D, OD, ELA, ER>::ES;
    constexpr bool RECS = (VIR || ELA) && D == 3;
    const int f = blockIdx.y;
    const size_t fo = (size_t)f * n, oo = (size_t)f * G * OD * OD;
    GatherOut of = out;
    if (VIR) of.vir = out.vir + oo;
    if (ELA) of.ela = out.ela + oo;
    if (KIN) of.kin = out.kin + oo;
    gather_tile_body<D, OD, CG, NQ, VIR, ELA, KIN, ER>(
        g, n, G, grid, perm, R, gld, cgnorm, cstart + (size_t)f * (g.ncell + 1), spos + fo * PS,
        KIN ? skin + fo * KN : nullptr, cap, bcount + fo, RECS ? rec + fo : nullptr, bonds + fo * (size_t)cap * ES, of,
        (size_t)blockIdx.y * gridDim.x + blockIdx.x, xcd_tile(blockIdx.x, gridDim.x), overflow);
}

// The sweep's field gather (2D or 3D, virial and/or kinetic): each
// workgroup finds its system, frame and tile in the sweep's table
// (ik_sweep.hpp).  Padding workgroups store zero counts for their slot, so
// k_reduce_stats can read every slot.  Three waves per SIMD: the system's
// geometry and pointers come from the table and stay in registers (a kernel
// argument would be reloaded instead), which at the single-system kernel's
// four waves spills 24 bytes a lane in the 2D FP64 form; with the room of
// three there is no spill, and the eight-system 2D sweep's gather takes
// 5.87 ms against 5.89 at four.
constexpr int kSweepGatherWaves = 3;

template <int D, int OD, bool VIR, bool KIN, typename ER>
__global__ __launch_bounds__(64, kSweepGatherWaves) void k_sweep_gather(
    const SweepSys *__restrict__ T, int ns, GLRule R, const int *__restrict__ cstart,
    const double *__restrict__ spos, const double *__restrict__ skin, int cap, const int *__restrict__ bcount,
    const float4 *__restrict__ rec, const ER *__restrict__ bonds, unsigned long long *__restrict__ wave_counts,
    int *__restrict__ overflow)
{
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int KN = OD * (OD + 1) / 2;
    constexpr int ES = BondLayout<D, OD, false, ER>::ES;
    constexpr bool RECS = VIR && D == 3;
    const SweepSys &S = T[sweep_system(T, ns, SW_GATHER, blockIdx.x)];
    int f, tile;
    if (!sweep_frame_tile(S, SW_GATHER, blockIdx.x, f, tile)) {
        if (threadIdx.x == 0) {
            wave_counts[2 * (size_t)blockIdx.x] = 0;
            wave_counts[2 * (size_t)blockIdx.x + 1] = 0;
        }
        return;
    }
    const Geom &g = S.g;
    const int n = S.n, G = S.G;
    const size_t fo = (size_t)S.p0 + (size_t)f * n, oo = (size_t)f * G * OD * OD;
    GatherOut of{};
    of.vir = VIR ? S.vir + oo : nullptr;
    of.kin = KIN ? S.kin + oo : nullptr;
    of.wave_counts = wave_counts;
    gather_tile_body<D, OD, 0, 5, VIR, false, KIN, ER>(
        g, n, G, S.grid, S.perm, R, nullptr, S.cgnorm, cstart + (size_t)S.c0 + (size_t)f * (g.ncell + 1),
        spos + fo * PS, KIN ? skin + fo * KN : nullptr, cap, VIR ? bcount + fo : nullptr, RECS ? rec + fo : nullptr,
        VIR ? bonds + fo * (size_t)cap * ES : nullptr, of, blockIdx.x, tile, overflow);
}

}  // namespace pgt
