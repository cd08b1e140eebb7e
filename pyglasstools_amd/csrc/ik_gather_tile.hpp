// ik_gather_tile.hpp -- the LDS-staged field gather ("tile" gather) of the
// IK stress-field engine.  Included by ik_gather_launch.hpp.
//
// One workgroup of BLK lanes takes BLK consecutive grid points in Morton
// order (a compact tile of the grid), one point per lane.  The workgroup
// first finds the window of cells that holds every particle within reach of
// its points, then copies those particles (position, kinetic tensor, bond
// count) and their bond entries from global memory into LDS: particle data
// with one lane per particle, bond entries with four lanes per particle
// reading consecutive 16-byte pieces, so a load instruction touches a few
// cache lines instead of one line per lane.  Each particle's entries are read
// from global memory once per tile instead of once per grid point whose ball
// holds it (12 points on the 2D benchmark grids, 60 in 3D), and the
// per-lane candidate scan and bond loop read LDS only.
//
// A window larger than the LDS arrays is processed in pieces: blocks of at
// most MAXWC window cells (row-major), and inside a block chunks of whole
// cells whose particles and entries fit MAXP and MAXE; a lane then takes the
// candidates of its own cells inside the piece.  Typical 2D windows are one
// piece.  Sums are FP64 per grid point in list order (pieces in order, then
// the lane's list sorted by bond count, largest first), so the result does
// not depend on scheduling.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ik_bonds.hpp"
#include "ik_device.hpp"
#include "ik_geom.hpp"
#include "ik_launch.hpp"

namespace pgt {

// Exclusive scan of a[0 .. n) in LDS by the whole workgroup (every thread must
// call it); returns the total.  tmp holds BLK / 64 + 1 ints.
template <int BLK>
__device__ int block_exscan(int *a, int n, int *tmp)
{
    constexpr int NW = BLK / 64;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (n + BLK - 1) / BLK;
    const int b0 = min(n, t * per), b1 = min(n, b0 + per);
    int s = 0;
    for (int k = b0; k < b1; ++k) s += a[k];
    // inclusive scan of the thread sums inside the wavefront
    int v = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    if (lane == 63) tmp[w] = v;
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int k = 0; k < NW; ++k) {
            const int c = tmp[k];
            tmp[k] = acc;
            acc += c;
        }
        tmp[NW] = acc;
    }
    __syncthreads();
    int run = tmp[w] + v - s;
    for (int k = b0; k < b1; ++k) {
        const int c = a[k];
        a[k] = run;
        run += c;
    }
    const int total = tmp[NW];
    __syncthreads();
    return total;
}

// LDS capacities of the tile gather for one form.
template <int D, int OD, bool ELA, typename ER>
struct TileCaps {
    using LY = BondLayout<D, OD, ELA, ER>;
    static constexpr int BLK = kTileThreads;          // lanes per workgroup: two per grid point
    static constexpr int TP = kTilePoints;            // grid points per tile
    static constexpr int MAXP = D == 2 ? 320 : 384;   // staged particles per chunk
    static constexpr int EBUDGET = D == 2 ? 46 * 1024 : 44 * 1024;
    static constexpr int MAXE = (EBUDGET / (LY::ES * (int)sizeof(ER))) / 4 * 4;  // staged entries per chunk
    static constexpr int MAXWC = 256;                 // window cells per block
    static constexpr int MAXL = D == 2 ? 16 : 24;     // per-lane list entries before an early flush
    static_assert(MAXP <= 1024, "list entries hold a 10-bit staged index");
};

// Tile gather.  Lanes t and t + 128 share grid point t of the tile: each
// takes every other candidate of the point's cells (so the two lists split
// the point's particles) and the two partial sums are added at the end in a
// fixed order.  Two lanes per point give the workgroup four wavefronts for
// the LDS that 128 points need, so two workgroups per CU keep two
// wavefronts on every SIMD while one of them stages.
template <int D, int OD, int CG, int NQ, bool VIR, bool ELA, bool KIN, typename ER>
__global__ __launch_bounds__(256) void k_gather_tile(Geom g, int n, int G, const double *__restrict__ grid,
                                                     const int *__restrict__ perm, GLRule R,
                                                     const double *__restrict__ gld, double cgnorm,
                                                     const int *__restrict__ cstart, const double *__restrict__ spos,
                                                     const double *__restrict__ skin, int cap,
                                                     const int *__restrict__ bcount, const ER *__restrict__ bonds,
                                                     GatherOut out, int *__restrict__ overflow)
{
    using CP = TileCaps<D, OD, ELA, ER>;
    using LY = BondLayout<D, OD, ELA, ER>;
    constexpr int BLK = CP::BLK, TP = CP::TP, MAXP = CP::MAXP, MAXE = CP::MAXE, MAXWC = CP::MAXWC, MAXL = CP::MAXL;
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int KN = OD * (OD + 1) / 2;
    constexpr bool PAIRS = VIR || ELA;
    constexpr int ES = LY::ES;    // values per entry
    constexpr int GRP = LY::GRP;  // entries per 16-byte-aligned read group
    constexpr int GV = LY::GV;    // values per read group
    constexpr int PIECES = GV * (int)sizeof(ER) / 16;  // 16-byte pieces per read group
    constexpr int NWV = BLK / 64;
    using V = typename std::conditional<std::is_same<ER, double>::value, double2, float4>::type;

    __shared__ __align__(16) double s_pos[MAXP][PS];
    __shared__ __align__(16) double s_kin[KIN ? MAXP : 1][KN];
    __shared__ __align__(16) ER s_bonds[PAIRS ? MAXE * ES : 2 * TP * 3 * KN];
    __shared__ int s_nb[MAXP];
    __shared__ int s_off[MAXP + 1];
    __shared__ int s_gidx[MAXP];
    __shared__ int s_cell[MAXWC + 1];  // staged start of each cell of the block (then the end)
    __shared__ int s_cg[MAXWC];        // first sorted particle index of each cell of the block
    __shared__ unsigned short s_list[MAXL][BLK];
    __shared__ int s_tmp[NWV + 1];
    __shared__ int s_win[12];  // ref[3], dmin[3], dmax[3], chunk end, flags
    static_assert(sizeof(s_bonds) >= (size_t)TP * 3 * KN * sizeof(double), "room to combine the two halves");

    const int f = blockIdx.y;
    const int tid = threadIdx.x;
    const int half = tid / TP;  // which of the point's two lanes
    const int t0 = xcd_tile(blockIdx.x, gridDim.x) * TP;  // first point slot of the tile
    const int t = t0 + (tid % TP);
    const bool live = t < G;
    const double *sp = spos + (size_t)f * n * PS;
    const int *st = cstart + (size_t)f * (g.ncell + 1);
    const int *bc = bcount + (size_t)f * n;
    const ER *bd = bonds + (size_t)f * n * (size_t)cap * ES;
    const double *sk = KIN ? skin + (size_t)f * n * KN : nullptr;

    const int gp = perm[live ? t : t0];
    double xg[3];
    xg[0] = grid[(size_t)gp * 3];
    xg[1] = grid[(size_t)gp * 3 + 1];
    xg[2] = (D == 3) ? grid[(size_t)gp * 3 + 2] : 0.0;
    int cc[3];
    for (int a = 0; a < 3; ++a) cc[a] = cell_coord(g, a, xg[a]);

    // ---- window: cells within reach of every point of the tile ------------
    // offsets from the first point's cell, wrapped into (-nc/2, nc/2], reduced
    // to their range over the tile
    if (tid == 0)
        for (int a = 0; a < 3; ++a) {
            s_win[a] = cc[a];
            s_win[3 + a] = 0;
            s_win[6 + a] = 0;
        }
    __syncthreads();
    if (live && half == 0)
        for (int a = 0; a < 3; ++a) {
            const int nc = g.nc[a];
            if (nc == 1) continue;
            int d = cc[a] - s_win[a];
            if (2 * d > nc) d -= nc;
            else if (2 * d <= -nc) d += nc;
            atomicMin(&s_win[3 + a], d);
            atomicMax(&s_win[6 + a], d);
        }
    __syncthreads();
    int wlo[3], wn[3];
    for (int a = 0; a < 3; ++a) {
        const int nc = g.nc[a];
        if (nc == 1) {
            wlo[a] = 0;
            wn[a] = 1;
            continue;
        }
        const int cnt = s_win[6 + a] - s_win[3 + a] + 2 * g.reach + 1;
        if (cnt >= nc) {
            wlo[a] = 0;
            wn[a] = nc;
        } else {
            int lo = s_win[a] + s_win[3 + a] - g.reach;
            lo %= nc;
            if (lo < 0) lo += nc;
            wlo[a] = lo;
            wn[a] = cnt;
        }
    }
    const int wcells = wn[0] * wn[1] * wn[2];
    // window-relative coordinate of cell c along axis a
    auto rel = [&](int a, int c) {
        int k = c - wlo[a];
        if (k < 0) k += g.nc[a];
        return k;
    };

    double av[KN], ae[KN], ak[KN];
#pragma unroll
    for (int k = 0; k < KN; ++k) { av[k] = 0.0; ae[k] = 0.0; ak[k] = 0.0; }
    unsigned long long nball = 0, ncontrib = 0;
    const double s2 = 2.0 * g.inv_rc;

    // integrate the entry at offset O (a std::integral_constant) of a read group
    auto integrate = [&](const ER (&Gv)[GV], auto O, const ER (&ta)[3], ER Aj) {
        constexpr int o = decltype(O)::value;
        const ER *e = Gv + o;
        ER C = e[0] * e[0];
        C = fma(e[1], e[1], C);
        if (D == 3) C = fma(e[2], e[2], C);
        ER B = ta[0] * e[0] + ta[1] * e[1];
        if (D == 3) B = fma(ta[2], e[2], B);
        const ER w = bond_quad<CG, NQ, ER>(R, gld, R.n, Aj, B, C);
        if (VIR) {
            const ER c = w * e[LY::OC];
            const ER cb0 = c * e[0], cb1 = c * e[1];
            if constexpr (std::is_same<ER, double>::value) {
                av[0] = fma(cb0, e[0], av[0]);
                av[1] = fma(cb1, e[1], av[1]);
                if constexpr (OD == 2) {
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
                if constexpr (OD == 2) {
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
            for (int q = 0; q < KN; ++q) {
                if constexpr (std::is_same<ER, double>::value) ae[q] = fma(w, e[LY::OE + q], ae[q]);
                else ae[q] += (double)(w * e[LY::OE + q]);
            }
        }
    };
    auto integrate_group = [&](const ER (&Gv)[GV], int k, int nb, const ER (&ta)[3], ER Aj) {
        integrate(Gv, std::integral_constant<int, 0>(), ta, Aj);
        if constexpr (GRP >= 2)
            if (GRP * k + 1 < nb) integrate(Gv, std::integral_constant<int, ES>(), ta, Aj);
        if constexpr (GRP >= 3)
            if (GRP * k + 2 < nb) integrate(Gv, std::integral_constant<int, 2 * ES>(), ta, Aj);
        if constexpr (GRP >= 4)
            if (GRP * k + 3 < nb) integrate(Gv, std::integral_constant<int, 3 * ES>(), ta, Aj);
    };
    // read group k of staged particle s (entries at s_off[s])
    auto load_group = [&](int off, int k, ER (&v)[GV]) {
        const V *q = reinterpret_cast<const V *>(s_bonds + (size_t)(off + k * GRP) * ES);
#pragma unroll
        for (int j = 0; j < PIECES; ++j) {
            const V u = q[j];
            if constexpr (std::is_same<ER, double>::value) {
                v[2 * j] = u.x;
                v[2 * j + 1] = u.y;
            } else {
                v[4 * j] = u.x;
                v[4 * j + 1] = u.y;
                v[4 * j + 2] = u.z;
                v[4 * j + 3] = u.w;
            }
        }
    };

    // sort the first cnt list entries by bond count (largest first), then
    // integrate them; the next particle's position, count and first read
    // group are loaded while the current one is integrated
    auto process = [&](int cnt) {
        if constexpr (PAIRS) {
            for (int k = 1; k < cnt; ++k) {
                const unsigned short v = s_list[k][tid];
                int m = k - 1;
                while (m >= 0 && s_list[m][tid] < v) {
                    s_list[m + 1][tid] = s_list[m][tid];
                    --m;
                }
                s_list[m + 1][tid] = v;
            }
        }
        auto fetch = [&](int q, double (&ps)[3], int &nb, int &off, ER (&e0)[GV]) {
            const int s = s_list[q][tid] & 1023;
            ps[0] = s_pos[s][0];
            ps[1] = s_pos[s][1];
            ps[2] = (D == 3) ? s_pos[s][2] : 0.0;
            nb = PAIRS ? s_nb[s] : 0;
            off = PAIRS ? s_off[s] : 0;
            if constexpr (PAIRS) load_group(off, 0, e0);
            return s;
        };
        double pos[3];
        int nb, off;
        ER E0[GV];
        int s = fetch(0, pos, nb, off, E0);
        for (int q = 0; q < cnt; ++q) {
            double npos[3];
            int nnb, noff;
            ER nE0[GV];
            const int ns = fetch(q + 1 < cnt ? q + 1 : q, npos, nnb, noff, nE0);
            double a[3];
            a[0] = minimg(pos[0] - xg[0], g.L[0], g.invL[0], g.periodic[0]);
            a[1] = minimg(pos[1] - xg[1], g.L[1], g.invL[1], g.periodic[1]);
            a[2] = (D == 3) ? minimg(pos[2] - xg[2], g.L[2], g.invL[2], g.periodic[2]) : 0.0;
            const double A = (a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) * g.inv_rc2;
            if (KIN) {
                const double w = shape<CG, double>(A);
#pragma unroll
                for (int k = 0; k < KN; ++k) ak[k] = fma(w, s_kin[s][k], ak[k]);
            }
            if constexpr (PAIRS) {
                ncontrib += (unsigned long long)nb;
                ER ta[3];
                ta[0] = (ER)(a[0] * s2);
                ta[1] = (ER)(a[1] * s2);
                ta[2] = (ER)(a[2] * s2);
                const ER Aj = (ER)A;
                const int ng = (nb + GRP - 1) / GRP;
                ER E1[GV];
                int k = 0;
                for (; k + 1 < ng; k += 2) {
                    load_group(off, k + 1, E1);
                    integrate_group(E0, k, nb, ta, Aj);
                    if (k + 2 < ng) load_group(off, k + 2, E0);
                    integrate_group(E1, k + 1, nb, ta, Aj);
                }
                if (k < ng) integrate_group(E0, k, nb, ta, Aj);
            }
            s = ns;
#pragma unroll
            for (int a2 = 0; a2 < 3; ++a2) pos[a2] = npos[a2];
            nb = nnb;
            off = noff;
#pragma unroll
            for (int k = 0; k < GV; ++k) E0[k] = nE0[k];
        }
    };

    // ---- window cells in blocks of at most MAXWC --------------------------
    for (int K0 = 0; K0 < wcells; K0 += MAXWC) {
        const int nwc = min(MAXWC, wcells - K0);
        // particle counts of the block's cells (window row-major: x fastest)
        for (int k = tid; k < nwc; k += BLK) {
            const int kk = K0 + k;
            const int kx = kk % wn[0], ky = (kk / wn[0]) % wn[1], kz = kk / (wn[0] * wn[1]);
            int x = wlo[0] + kx, y = wlo[1] + ky, z = wlo[2] + kz;
            if (x >= g.nc[0]) x -= g.nc[0];
            if (y >= g.nc[1]) y -= g.nc[1];
            if (z >= g.nc[2]) z -= g.nc[2];
            const int cell = (z * g.nc[1] + y) * g.nc[0] + x;
            const int b = st[cell];
            s_cg[k] = b;
            s_cell[k] = st[cell + 1] - b;
        }
        __syncthreads();
        const int ptot = block_exscan<BLK>(s_cell, nwc, s_tmp);
        if (tid == 0) s_cell[nwc] = ptot;
        __syncthreads();

        // chunks of whole cells of the block
        int k0 = 0;
        while (k0 < nwc) {
            // largest k1 with the particles of cells [k0, k1) within MAXP
            if (tid == 0) {
                int lo = k0 + 1, hi = nwc;  // s_cell is non-decreasing
                if (s_cell[lo] - s_cell[k0] > MAXP) {
                    s_win[10] = 1;  // one cell holds more particles than the chunk
                } else {
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (s_cell[mid] - s_cell[k0] <= MAXP) lo = mid;
                        else hi = mid - 1;
                    }
                    s_win[10] = 0;
                }
                s_win[9] = lo;
            }
            __syncthreads();
            int k1 = s_win[9];
            if (s_win[10] != 0) {
                if (tid == 0) atomicOr(overflow, 2);
                break;
            }
            const int S0 = s_cell[k0];
            int np = s_cell[k1] - S0;
            // staged index -> sorted particle index, one thread per cell
            for (int k = k0 + tid; k < k1; k += BLK) {
                const int b = s_cell[k] - S0, e = s_cell[k + 1] - S0, gb = s_cg[k];
                for (int i = b; i < e; ++i) s_gidx[i] = gb + (i - b);
            }
            __syncthreads();
            // particle data, one lane per particle, every load of the lane
            // issued before the first store
            if (np > 0) {
                constexpr int PU = (MAXP + BLK - 1) / BLK;
                double pv[PU][PS], kv[PU][KN];
                int nbv[PU];
#pragma unroll
                for (int u = 0; u < PU; ++u) {
                    const int i = min(tid + u * BLK, np - 1);
                    const int p = s_gidx[i];
#pragma unroll
                    for (int a = 0; a < PS; ++a) pv[u][a] = sp[(size_t)p * PS + a];
                    if (KIN)
#pragma unroll
                        for (int k = 0; k < KN; ++k) kv[u][k] = sk[(size_t)p * KN + k];
                    nbv[u] = PAIRS ? min(bc[p], cap) : 0;
                }
#pragma unroll
                for (int u = 0; u < PU; ++u) {
                    const int i = tid + u * BLK;
                    if (i >= np) continue;
#pragma unroll
                    for (int a = 0; a < PS; ++a) s_pos[i][a] = pv[u][a];
                    if (KIN)
#pragma unroll
                        for (int k = 0; k < KN; ++k) s_kin[i][k] = kv[u][k];
                    s_nb[i] = nbv[u];
                    s_off[i] = (nbv[u] + GRP - 1) / GRP * GRP;
                }
            }
            __syncthreads();
            if constexpr (PAIRS) {
                const int etot = block_exscan<BLK>(s_off, np, s_tmp);
                if (etot > MAXE) {
                    // fewer cells: the largest k1 whose entries fit
                    if (tid == 0) {
                        s_off[np] = etot;  // the end, for a last cell without particles
                        int lo = k0, hi = k1 - 1;
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (s_off[s_cell[mid] - S0] <= MAXE) lo = mid;
                            else hi = mid - 1;
                        }
                        // cells k0 .. lo - 1 fit (lo == k0: not even one)
                        s_win[9] = lo;
                    }
                    __syncthreads();
                    k1 = s_win[9];
                    if (k1 == k0) {
                        if (tid == 0) atomicOr(overflow, 4);
                        break;
                    }
                    np = s_cell[k1] - S0;
                }
                // bond entries: the read groups of all staged particles as one
                // run of 16-byte pieces; lane l copies pieces l, l + BLK, ...,
                // UP at a time with their loads issued together.  The owner of
                // a piece is found by a fixed-step binary search of s_off.
                const int etot_np = (np > 0) ? s_off[np - 1] + (s_nb[np - 1] + GRP - 1) / GRP * GRP : 0;
                const int totp = etot_np / GRP * PIECES;
                V *dst = reinterpret_cast<V *>(s_bonds);
                const V *src = reinterpret_cast<const V *>(bd);
                const size_t slot = (size_t)cap * ES * sizeof(ER) / 16;  // pieces per particle slot
                constexpr int UP = 4;
                for (int q0 = 0; q0 < totp; q0 += BLK * UP) {
                    V v[UP];
#pragma unroll
                    for (int u = 0; u < UP; ++u) {
                        const int q = min(q0 + u * BLK + tid, totp - 1);
                        const int e = q / PIECES * GRP;  // first entry of the piece's read group
                        int lo = 0;
#pragma unroll
                        for (int step = 512; step >= 1; step >>= 1)
                            if (lo + step < np && s_off[lo + step] <= e) lo += step;
                        v[u] = src[(size_t)s_gidx[lo] * slot + (q - s_off[lo] / GRP * PIECES)];
                    }
#pragma unroll
                    for (int u = 0; u < UP; ++u) {
                        const int q = q0 + u * BLK + tid;
                        if (q < totp) dst[q] = v[u];
                    }
                }
            }
            __syncthreads();

            // the lane's candidates inside cells [k0, k1) of this block: every
            // other one (this lane's half), NB at a time with their positions
            // read together
            if (live) {
                int cnt = 0;
                const int KA = K0 + k0, KB = K0 + k1;  // window cell range of the chunk
                const int Rr = g.reach;
                const int ndz = g.nc[2] == 1 ? 0 : Rr, ndy = g.nc[1] == 1 ? 0 : Rr;
                for (int dz = -ndz; dz <= ndz; ++dz) {
                    int z = cc[2] + dz;
                    if (z < 0) z += g.nc[2];
                    else if (z >= g.nc[2]) z -= g.nc[2];
                    const int kz = rel(2, z);
                    for (int dy = -ndy; dy <= ndy; ++dy) {
                        int y = cc[1] + dy;
                        if (y < 0) y += g.nc[1];
                        else if (y >= g.nc[1]) y -= g.nc[1];
                        const int ky = rel(1, y);
                        const int rowc = (kz * wn[1] + ky) * wn[0];
                        // x cells [cx - R, cx + R] as window columns: one or two ranges
                        int xa[2], xb[2], nr = 1;
                        if (g.nc[0] == 1) {
                            xa[0] = 0;
                            xb[0] = 0;
                        } else {
                            int x0 = cc[0] - Rr;
                            if (x0 < 0) x0 += g.nc[0];
                            const int ka = rel(0, x0);
                            if (ka + 2 * Rr < wn[0]) {
                                xa[0] = ka;
                                xb[0] = ka + 2 * Rr;
                            } else {  // whole axis, the range wraps
                                xa[0] = ka;
                                xb[0] = wn[0] - 1;
                                xa[1] = 0;
                                xb[1] = ka + 2 * Rr - wn[0];
                                nr = 2;
                            }
                        }
                        for (int r = 0; r < nr; ++r) {
                            const int ca = max(rowc + xa[r], KA), cb = min(rowc + xb[r] + 1, KB);
                            if (ca >= cb) continue;
                            const int sa = s_cell[ca - K0] - S0, sb = s_cell[cb - K0] - S0;
                            constexpr int NB = 4;
                            for (int s0 = sa + half; s0 < sb; s0 += 2 * NB) {
                                double cq[NB][3];
#pragma unroll
                                for (int u = 0; u < NB; ++u) {
                                    const int sx = min(s0 + 2 * u, sb - 1);
                                    cq[u][0] = s_pos[sx][0];
                                    cq[u][1] = s_pos[sx][1];
                                    cq[u][2] = (D == 3) ? s_pos[sx][2] : 0.0;
                                }
#pragma unroll
                                for (int u = 0; u < NB; ++u) {
                                    const int sx = s0 + 2 * u;
                                    if (sx >= sb) break;
                                    const double a0 = minimg(cq[u][0] - xg[0], g.L[0], g.invL[0], g.periodic[0]);
                                    const double a1 = minimg(cq[u][1] - xg[1], g.L[1], g.invL[1], g.periodic[1]);
                                    const double a2 =
                                        (D == 3) ? minimg(cq[u][2] - xg[2], g.L[2], g.invL[2], g.periodic[2]) : 0.0;
                                    if (!(a0 * a0 + a1 * a1 + a2 * a2 < g.rc2)) continue;
                                    ++nball;
                                    if (cnt == MAXL) {  // a full list is processed early
                                        process(cnt);
                                        cnt = 0;
                                    }
                                    s_list[cnt][tid] = (unsigned short)(sx | (min(s_nb[sx], 63) << 10));
                                    ++cnt;
                                }
                            }
                        }
                    }
                }
                if (cnt > 0) process(cnt);
            }
            __syncthreads();  // the next chunk overwrites the staged data
            k0 = k1;
        }
        __syncthreads();
    }

    // the two halves of each point: half 1 leaves its sums in LDS, half 0
    // adds them (in this fixed order) and stores the point's tensors
    double *s_acc = reinterpret_cast<double *>(s_bonds);
    const int pt = tid % TP;
    if (half == 1)
#pragma unroll
        for (int k = 0; k < KN; ++k) {
            s_acc[(0 * KN + k) * TP + pt] = av[k];
            s_acc[(1 * KN + k) * TP + pt] = ae[k];
            s_acc[(2 * KN + k) * TP + pt] = ak[k];
        }
    __syncthreads();
    if (half == 0 && live) {
#pragma unroll
        for (int k = 0; k < KN; ++k) {
            av[k] += s_acc[(0 * KN + k) * TP + pt];
            ae[k] += s_acc[(1 * KN + k) * TP + pt];
            ak[k] += s_acc[(2 * KN + k) * TP + pt];
        }
        const size_t ob = ((size_t)f * G + gp) * OD * OD;
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
    nball = wave_sum(nball);
    ncontrib = wave_sum(ncontrib);
    if ((tid & 63) == 0) {
        const size_t w = ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * NWV + (tid >> 6);
        out.wave_counts[2 * w] = nball;
        out.wave_counts[2 * w + 1] = ncontrib;
    }
}

}  // namespace pgt
