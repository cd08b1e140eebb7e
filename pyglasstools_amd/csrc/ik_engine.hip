// ik_engine.hip -- MI355X (gfx950) Irving-Kirkwood stress-field engine behind
// the C ABI declared in include/pgt_ik.h.
//
// Per frame the engine runs:
//   1. cell list   k_cell_count -> k_cell_scan -> k_cell_fill -> k_cell_sort
//                  (counting sort of particles into cells of edge >= the
//                  largest search radius; inside a cell particles are ordered
//                  by input index, so every later sum has a fixed order)
//   2. bond list   k_bonds: for every particle i, the neighbours j within
//                  max_rcut with a nonzero pair force; each entry holds the
//                  scaled bond vector, |b|^2/rc^2 and the pair tensors
//                  (0.5 * F (x) r_ij, times the kernel normalisation), so the
//                  force is evaluated once per ordered pair instead of once per
//                  (grid point, i, j) as in the reference
//   3. field gather k_gather: lanes per grid point walk the 3x3(x3) cells
//                  around the grid point, keep the particles i inside the
//                  coarse-graining ball and, for each bond of i, evaluate the
//                  Gauss-Legendre bond integral and accumulate the tensors in
//                  FP64; the lanes of a grid point are combined by a
//                  wavefront shuffle reduction and one lane stores the result.
// Frames are batched along blockIdx.y so one launch covers many frames.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "ik_device.hpp"
#include "pgt_ik.h"

using namespace pgt;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                               \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return fail(PGT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));           \
    } while (0)

// ------------------------------------------------------------------------
// Geometry passed to every kernel
// ------------------------------------------------------------------------
struct Geom {
    double L[3], invL[3], lo[3];
    int periodic[3];
    int nc[3];     // cells per axis
    double cs[3];  // cell edge
    int ncell;
    double rc2;       // cg_rcut^2
    double inv_rc;    // 1/cg_rcut
    double inv_rc2;   // 1/cg_rcut^2
    double mr2;       // max_rcut^2
};

__device__ __forceinline__ double minimg_div(double d, double L, int periodic)
{
    return periodic ? d - L * rint(d / L) : d;
}

__device__ __forceinline__ double minimg_mul(double d, double L, double invL, int periodic)
{
    return periodic ? d - L * rint(d * invL) : d;
}

__device__ __forceinline__ int cell_coord(const Geom &g, int a, double p)
{
    if (g.nc[a] == 1) return 0;
    double u = p - g.lo[a];
    u -= g.L[a] * floor(u * g.invL[a]);
    int k = (int)floor(u / g.cs[a]);
    return k < 0 ? 0 : (k >= g.nc[a] ? g.nc[a] - 1 : k);
}

// Up to two contiguous particle ranges covering cells [c-1, c+1] of one row
// of the cell grid along x (wrapping periodically).  Returns the number of
// ranges written.
__device__ __forceinline__ int row_ranges(const Geom &g, const int *__restrict__ cstart, int rowbase, int cx,
                                          int *beg, int *end)
{
    const int nx = g.nc[0];
    if (nx < 3) {
        beg[0] = cstart[rowbase];
        end[0] = cstart[rowbase + nx];
        return 1;
    }
    int lo = cx - 1, hi = cx + 1;
    if (lo >= 0 && hi < nx) {
        beg[0] = cstart[rowbase + lo];
        end[0] = cstart[rowbase + hi + 1];
        return 1;
    }
    if (lo < 0) {  // cells nx-1 and [0, 1]
        beg[0] = cstart[rowbase + nx - 1];
        end[0] = cstart[rowbase + nx];
        beg[1] = cstart[rowbase];
        end[1] = cstart[rowbase + 2];
        return 2;
    }
    // hi == nx: cells [nx-2, nx-1] and 0
    beg[0] = cstart[rowbase + nx - 2];
    end[0] = cstart[rowbase + nx];
    beg[1] = cstart[rowbase];
    end[1] = cstart[rowbase + 1];
    return 2;
}

// Rows of neighbouring cells around cell (cy, cz): writes row base indices.
__device__ __forceinline__ int neighbour_rows(const Geom &g, int cy, int cz, int *rowbase)
{
    int ny = g.nc[1], nz = g.nc[2];
    int ys[3], zs[3], nyy, nzz;
    if (ny >= 3) { ys[0] = (cy + ny - 1) % ny; ys[1] = cy; ys[2] = (cy + 1) % ny; nyy = 3; }
    else { for (int k = 0; k < ny; ++k) ys[k] = k; nyy = ny; }
    if (nz >= 3) { zs[0] = (cz + nz - 1) % nz; zs[1] = cz; zs[2] = (cz + 1) % nz; nzz = 3; }
    else { for (int k = 0; k < nz; ++k) zs[k] = k; nzz = nz; }
    int n = 0;
    for (int a = 0; a < nzz; ++a)
        for (int b = 0; b < nyy; ++b) rowbase[n++] = (zs[a] * ny + ys[b]) * g.nc[0];
    return n;
}

// ------------------------------------------------------------------------
// 1. cell list
// ------------------------------------------------------------------------
__global__ void k_cell_count(Geom g, int n, const double *__restrict__ pos, int *__restrict__ pcell,
                             int *__restrict__ pslot, int *__restrict__ ccount)
{
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *p = pos + ((size_t)f * n + i) * 3;
    int cx = cell_coord(g, 0, p[0]), cy = cell_coord(g, 1, p[1]), cz = cell_coord(g, 2, p[2]);
    int c = (cz * g.nc[1] + cy) * g.nc[0] + cx;
    pcell[(size_t)f * n + i] = c;
    pslot[(size_t)f * n + i] = atomicAdd(&ccount[(size_t)f * g.ncell + c], 1);
}

// exclusive scan of ccount into cstart (ncell+1 entries per frame), one
// workgroup of 1024 threads per frame; also clears ccount for the next use.
__global__ __launch_bounds__(1024) void k_cell_scan(int ncell, int *__restrict__ ccount, int *__restrict__ cstart)
{
    const int f = blockIdx.x;
    int *cnt = ccount + (size_t)f * ncell;
    int *st = cstart + (size_t)f * (ncell + 1);
    const int T = blockDim.x, t = threadIdx.x;
    const int chunk = (ncell + T - 1) / T;
    const int b0 = min(ncell, t * chunk), b1 = min(ncell, b0 + chunk);
    int s = 0;
    for (int k = b0; k < b1; ++k) s += cnt[k];
    __shared__ int sh[1024];
    sh[t] = s;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        int v = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    int run = sh[t] - s;  // exclusive prefix of this thread's chunk
    for (int k = b0; k < b1; ++k) {
        st[k] = run;
        run += cnt[k];
        cnt[k] = 0;
    }
    if (t == T - 1) st[ncell] = sh[T - 1];
}

__global__ void k_cell_fill(int n, int ncell, const int *__restrict__ pcell, const int *__restrict__ pslot,
                            const int *__restrict__ cstart, int *__restrict__ order)
{
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t fi = (size_t)f * n + i;
    order[(size_t)f * n + cstart[(size_t)f * (ncell + 1) + pcell[fi]] + pslot[fi]] = i;
}

// Per cell: insertion-sort its particle indices (input order), then write the
// cell-sorted particle arrays.
template <int D, int OD>
__global__ void k_cell_sort(int n, int ncell, const int *__restrict__ cstart, int *__restrict__ order,
                            const double *__restrict__ pos, const double *__restrict__ diam,
                            const double *__restrict__ mass, const double *__restrict__ vel,
                            const double *__restrict__ disp, double *__restrict__ spos,
                            double *__restrict__ sdiam, double *__restrict__ skin, double *__restrict__ sdisp)
{
    const int f = blockIdx.y;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int *st = cstart + (size_t)f * (ncell + 1);
    int *ord = order + (size_t)f * n;
    const int b = st[c], e = st[c + 1];
    for (int k = b + 1; k < e; ++k) {
        int v = ord[k], m = k - 1;
        while (m >= b && ord[m] > v) {
            ord[m + 1] = ord[m];
            --m;
        }
        ord[m + 1] = v;
    }
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int KN = OD * (OD + 1) / 2;
    for (int k = b; k < e; ++k) {
        const int i = ord[k];
        const size_t fi = (size_t)f * n + i, fk = (size_t)f * n + k;
        const double *p = pos + fi * 3;
        double *sp = spos + fk * PS;
        sp[0] = p[0];
        sp[1] = p[1];
        if (D == 3) { sp[2] = p[2]; sp[3] = 0.0; }
        sdiam[fk] = diam[fi];
        if (skin) {
            const double m = mass ? mass[fi] : 1.0;
            double v[3] = {0, 0, 0};
            if (vel) { v[0] = vel[fi * 3]; v[1] = vel[fi * 3 + 1]; v[2] = vel[fi * 3 + 2]; }
            double *kk = skin + fk * KN;
            // symmetric packing: diagonal first, then (0,1), (0,2), (1,2)
            for (int a = 0; a < OD; ++a) kk[a] = m * v[a] * v[a];
            if (OD == 2) kk[2] = m * v[0] * v[1];
            else { kk[3] = m * v[0] * v[1]; kk[4] = m * v[0] * v[2]; kk[5] = m * v[1] * v[2]; }
        }
        if (sdisp) {
            for (int a = 0; a < 3; ++a) sdisp[fk * 3 + a] = disp ? disp[fi * 3 + a] : 0.0;
        }
    }
}

// ------------------------------------------------------------------------
// 2. bond list
// ------------------------------------------------------------------------
// Entry layout (ER = double or float), stride ES = D + 1 + NT*KN:
//   [0..D)   b / rc               (scaled bond vector r_ij)
//   [D]      |r_ij|^2 / rc^2
//   [D+1 ..) NT packed symmetric tensors (virial first, then elastic), each
//            pre-multiplied by 0.5 * (kernel normalisation)
template <int D, int OD, bool VIR, bool ELA, typename ER>
__global__ void k_bonds(Geom g, int n, PotParams P, double cgnorm, const int *__restrict__ cstart,
                        const double *__restrict__ spos, const double *__restrict__ sdiam,
                        const double *__restrict__ sdisp, int cap, int *__restrict__ bcount, ER *__restrict__ bonds,
                        int *__restrict__ maxcount)
{
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int KN = OD * (OD + 1) / 2;
    constexpr int NT = (VIR ? 1 : 0) + (ELA ? 1 : 0);
    constexpr int ES = D + 1 + NT * KN;
    const int f = blockIdx.y;
    const int p_raw = blockIdx.x * blockDim.x + threadIdx.x;
    // lanes past the end shadow the last particle without writing, so the
    // whole wavefront takes part in the max-reduction below
    const bool live = p_raw < n;
    const int p = live ? p_raw : n - 1;
    const double *sp = spos + (size_t)f * n * PS;
    const double *sd = sdiam + (size_t)f * n;
    const int *st = cstart + (size_t)f * (g.ncell + 1);
    const double xi = sp[(size_t)p * PS], yi = sp[(size_t)p * PS + 1], zi = (D == 3) ? sp[(size_t)p * PS + 2] : 0.0;
    const double di = sd[p];
    int cx = cell_coord(g, 0, xi), cy = cell_coord(g, 1, yi), cz = cell_coord(g, 2, zi);
    int rows[9];
    const int nrows = neighbour_rows(g, cy, cz, rows);
    ER *out = bonds + ((size_t)f * n + p) * (size_t)cap * ES;
    int cnt = 0;
    for (int r = 0; r < nrows; ++r) {
        int rb[2], re[2];
        const int nr = row_ranges(g, st, rows[r], cx, rb, re);
        for (int s = 0; s < nr; ++s) {
            for (int q = rb[s]; q < re[s]; ++q) {
                if (q == p) continue;
                double b[3];
                b[0] = minimg_div(sp[(size_t)q * PS] - xi, g.L[0], g.periodic[0]);
                b[1] = minimg_div(sp[(size_t)q * PS + 1] - yi, g.L[1], g.periodic[1]);
                b[2] = (D == 3) ? minimg_div(sp[(size_t)q * PS + 2] - zi, g.L[2], g.periodic[2]) : 0.0;
                const double rsq = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
                if (!(rsq < g.mr2)) continue;
                const double dj = sd[q];
                double f0, f1 = 0.0;
                double ur[3] = {0, 0, 0};
                if constexpr (std::is_same<ER, double>::value) f0 = force_divr<double>(P, rsq, di, dj);
                else f0 = (double)force_divr<float>(P, (float)rsq, (float)di, (float)dj);
                if (ELA) {
                    const double *sdp = sdisp + (size_t)f * n * 3;
                    for (int a = 0; a < 3; ++a) ur[a] = sdp[(size_t)q * 3 + a] - sdp[(size_t)p * 3 + a] + b[a];
                    const double ursq = ur[0] * ur[0] + ur[1] * ur[1] + ur[2] * ur[2];
                    if constexpr (std::is_same<ER, double>::value) f1 = force_divr<double>(P, ursq, di, dj);
                    else f1 = (double)force_divr<float>(P, (float)ursq, (float)di, (float)dj);
                    if (f0 == 0.0 && f1 == 0.0) continue;
                } else {
                    if (f0 == 0.0) continue;
                }
                if (live && cnt < cap) {
                    ER *e = out + (size_t)cnt * ES;
                    for (int a = 0; a < D; ++a) e[a] = (ER)(b[a] * g.inv_rc);
                    e[D] = (ER)(rsq * g.inv_rc2);
                    int o = D + 1;
                    const double h = 0.5 * cgnorm;
                    if (VIR) {
                        const double c = h * f0;
                        for (int a = 0; a < OD; ++a) e[o + a] = (ER)(c * b[a] * b[a]);
                        if (OD == 2) e[o + 2] = (ER)(c * b[0] * b[1]);
                        else {
                            e[o + 3] = (ER)(c * b[0] * b[1]);
                            e[o + 4] = (ER)(c * b[0] * b[2]);
                            e[o + 5] = (ER)(c * b[1] * b[2]);
                        }
                        o += KN;
                    }
                    if (ELA) {
                        // 0.5 * (F1_a (u+r)_b - F_a r_b), ListofCommonObservables.h:113-120
                        auto T = [&](int a, int bb) { return f1 * ur[a] * ur[bb] - f0 * b[a] * b[bb]; };
                        for (int a = 0; a < OD; ++a) e[o + a] = (ER)(h * T(a, a));
                        if (OD == 2) e[o + 2] = (ER)(h * T(0, 1));
                        else {
                            e[o + 3] = (ER)(h * T(0, 1));
                            e[o + 4] = (ER)(h * T(0, 2));
                            e[o + 5] = (ER)(h * T(1, 2));
                        }
                    }
                }
                ++cnt;
            }
        }
    }
    if (live) bcount[(size_t)f * n + p] = cnt;
    // largest bond count of the job: one atomic per wavefront
    int m = live ? cnt : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(maxcount, m);
}

// ------------------------------------------------------------------------
// 3. field gather
// ------------------------------------------------------------------------
struct GatherOut {
    double *vir;
    double *kin;
    double *ela;
    unsigned long long *n_ball;
    unsigned long long *n_contrib;
};

template <int CG, int NQ, typename ER>
__device__ __forceinline__ ER bond_gl(const GLRule &R, const double *__restrict__ gld, int nq, ER A, ER B, ER C)
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

template <int D, int OD, int CG, int NQ, bool VIR, bool ELA, int LPG, typename ER>
__global__ __launch_bounds__(256) void k_gather(Geom g, int n, int G, const double *__restrict__ grid, GLRule R,
                                                const double *__restrict__ gld, double cgnorm,
                                                const int *__restrict__ cstart, const double *__restrict__ spos,
                                                const double *__restrict__ skin, int cap,
                                                const int *__restrict__ bcount, const ER *__restrict__ bonds,
                                                GatherOut out, int count_stats)
{
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int KN = OD * (OD + 1) / 2;
    constexpr int NT = (VIR ? 1 : 0) + (ELA ? 1 : 0);
    constexpr int ES = D + 1 + NT * KN;
    const int f = blockIdx.y;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int gp = tid / LPG;
    const int sub = tid % LPG;
    const bool live = gp < G;
    const double *sp = spos + (size_t)f * n * PS;
    const int *st = cstart + (size_t)f * (g.ncell + 1);
    const int *bc = bcount + (size_t)f * n;
    const ER *bd = bonds + (size_t)f * n * (size_t)cap * ES;
    // the kinetic field is a per-launch (uniform) choice
    const bool KIN = skin != nullptr;
    const double *sk = KIN ? skin + (size_t)f * n * KN : nullptr;

    double av[KN], ae[KN], ak[KN];
#pragma unroll
    for (int k = 0; k < KN; ++k) { av[k] = 0.0; ae[k] = 0.0; ak[k] = 0.0; }
    unsigned long long nball = 0, ncontrib = 0;

    if (live) {
        const double x0 = grid[(size_t)gp * 3], x1 = grid[(size_t)gp * 3 + 1];
        const double x2 = (D == 3) ? grid[(size_t)gp * 3 + 2] : 0.0;
        const int cx = cell_coord(g, 0, x0), cy = cell_coord(g, 1, x1), cz = cell_coord(g, 2, x2);
        int rows[9];
        const int nrows = neighbour_rows(g, cy, cz, rows);
        for (int r = 0; r < nrows; ++r) {
            int rb[2], re[2];
            const int nr = row_ranges(g, st, rows[r], cx, rb, re);
            for (int s = 0; s < nr; ++s) {
                for (int p = rb[s] + sub; p < re[s]; p += LPG) {
                    const double a0 = minimg_mul(sp[(size_t)p * PS] - x0, g.L[0], g.invL[0], g.periodic[0]);
                    const double a1 = minimg_mul(sp[(size_t)p * PS + 1] - x1, g.L[1], g.invL[1], g.periodic[1]);
                    const double a2v = (D == 3) ? minimg_mul(sp[(size_t)p * PS + 2] - x2, g.L[2], g.invL[2], g.periodic[2]) : 0.0;
                    const double asq = a0 * a0 + a1 * a1 + a2v * a2v;
                    if (!(asq < g.rc2)) continue;
                    ++nball;
                    const double A = asq * g.inv_rc2;
                    if (KIN) {
                        const double w = shape<CG, double>(A);
#pragma unroll
                        for (int k = 0; k < KN; ++k) ak[k] = fma(w, sk[(size_t)p * KN + k], ak[k]);
                    }
                    if (VIR || ELA) {
                        const double s2 = 2.0 * g.inv_rc;
                        const ER ta0 = (ER)(a0 * s2), ta1 = (ER)(a1 * s2), ta2 = (ER)(a2v * s2);
                        const ER Ae = (ER)A;
                        const int nb = min(bc[p], cap);
                        ncontrib += nb;
                        const ER *e = bd + (size_t)p * cap * ES;
                        for (int k = 0; k < nb; ++k, e += ES) {
                            ER B = ta0 * e[0] + ta1 * e[1];
                            if (D == 3) B = fma(ta2, e[2], B);
                            const ER wgl = bond_gl<CG, NQ, ER>(R, gld, R.n, Ae, B, e[D]);
                            int o = D + 1;
                            if (VIR) {
#pragma unroll
                                for (int m = 0; m < KN; ++m) {
                                    if constexpr (std::is_same<ER, double>::value) av[m] = fma(wgl, e[o + m], av[m]);
                                    else av[m] += (double)(wgl * e[o + m]);
                                }
                                o += KN;
                            }
                            if (ELA) {
#pragma unroll
                                for (int m = 0; m < KN; ++m) {
                                    if constexpr (std::is_same<ER, double>::value) ae[m] = fma(wgl, e[o + m], ae[m]);
                                    else ae[m] += (double)(wgl * e[o + m]);
                                }
                            }
                        }
                    }
                }
            }
        }
    }
    // combine the LPG lanes of one grid point (lanes gp*LPG .. gp*LPG+LPG-1
    // are adjacent in the wavefront)
#pragma unroll
    for (int off = LPG / 2; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < KN; ++k) {
            if (VIR) av[k] += __shfl_xor(av[k], off, LPG);
            if (ELA) ae[k] += __shfl_xor(ae[k], off, LPG);
            if (KIN) ak[k] += __shfl_xor(ak[k], off, LPG);
        }
        if (count_stats) {
            nball += __shfl_xor(nball, off, LPG);
            ncontrib += __shfl_xor(ncontrib, off, LPG);
        }
    }
    if (!live || sub != 0) return;
    const size_t ob = ((size_t)f * G + gp) * OD * OD;
    auto store = [&](double *dst, const double *acc, double scale) {
        // unpack the symmetric accumulator into the full OD x OD tensor
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
    if (count_stats) {
        atomicAdd(out.n_ball, nball);
        atomicAdd(out.n_contrib, ncontrib);
    }
}

// ------------------------------------------------------------------------
// scalar / batched evaluators
// ------------------------------------------------------------------------
__global__ void k_eval_force(PotParams P, int n, const double *rij, const double *di, const double *dj,
                             double *fo, double *eo, double *ro)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double rsq = rij[3 * i] * rij[3 * i] + rij[3 * i + 1] * rij[3 * i + 1] + rij[3 * i + 2] * rij[3 * i + 2];
    if (fo) fo[i] = force_divr<double>(P, rsq, di[i], dj[i]);
    if (eo) eo[i] = pair_energy<double>(P, rsq, di[i], dj[i]);
    if (ro) {
        const double sigma = 0.5 * (di[i] + dj[i]) * (1 - P.eps * fabs(di[i] - dj[i]));
        ro[i] = P.rcutsq * sigma * sigma;
    }
}

// Bond integral exactly in the reference's arithmetic form: the integrand is
// Delta(|x - (ri + s rij)|^2) (CoarseGrainFunction.h:42-45) summed over the
// unfolded rule.
__global__ void k_eval_cg(int cg, double rc, int mode, int n, const double *s, const double *x, const double *ri,
                          const double *rij, const double *gld, int nq, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *X = x + 3 * i, *P = ri + 3 * i;
    auto delta_at = [&](double sv) {
        double dsq = 0.0;
        for (int a = 0; a < 3; ++a) {
            const double d = X[a] - (rij ? P[a] + sv * rij[3 * i + a] : P[a]);
            dsq += d * d;
        }
        return cg_delta_ref(cg, rc, dsq);
    };
    if (mode == 0) out[i] = delta_at(0.0);
    else if (mode == 1) out[i] = delta_at(s[i]);
    else {
        double acc = 0.0;
        for (int k = 0; k < nq; ++k) acc += gld[nq + k] * delta_at(gld[k]);
        out[i] = acc;
    }
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------
// Gauss-Legendre nodes on [-1,1] by Newton iteration on P_n in long double,
// unfolded onto [0,1] with the half-length in the weights.
int gl_rule(int n, std::vector<double> &s, std::vector<double> &w)
{
    if (n < 1 || n > 1024) return -1;
    s.assign(n, 0.0);
    w.assign(n, 0.0);
    const int m = (n + 1) / 2;
    for (int k = 0; k < m; ++k) {
        long double z = cosl(3.14159265358979323846264338327950288L * (k + 0.75L) / (n + 0.5L));
        long double dp = 1.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = z;
            for (int j = 2; j <= n; ++j) {
                long double p2 = ((2.0L * j - 1.0L) * z * p1 - (j - 1.0L) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p0 = 1.0L; p1 = z; }
            dp = n * (z * p1 - p0) / (z * z - 1.0L);
            long double dz = p1 / dp;
            z -= dz;
            if (fabsl(dz) < 1e-19L) {
                p0 = 1.0L; p1 = z;
                for (int j = 2; j <= n; ++j) {
                    long double p2 = ((2.0L * j - 1.0L) * z * p1 - (j - 1.0L) * p0) / j;
                    p0 = p1;
                    p1 = p2;
                }
                if (n == 1) { p0 = 1.0L; p1 = z; }
                dp = n * (z * p1 - p0) / (z * z - 1.0L);
                break;
            }
        }
        if ((n & 1) && k == m - 1) z = 0.0L;
        const long double wt = 2.0L / ((1.0L - z * z) * dp * dp);
        // node pair +-z -> s = (1 -+ z)/2, weight wt/2
        s[k] = (double)(0.5L - 0.5L * z);
        w[k] = (double)(0.5L * wt);
        s[n - 1 - k] = (double)(0.5L + 0.5L * z);
        w[n - 1 - k] = (double)(0.5L * wt);
    }
    return 0;
}

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    int ensure(size_t want)
    {
        if (want <= n && p) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (want == 0) want = 1;
        hipError_t e = hipMalloc(&p, want * sizeof(T));
        if (e != hipSuccess) { p = nullptr; return fail(PGT_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e)); }
        n = want;
        return 0;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace

struct pgt_ik_engine {
    pgt_ik_params prm{};
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> ev;
    bool box_set = false;
    double L[3] = {0, 0, 0}, lo[3] = {0, 0, 0};
    int periodic[3] = {1, 1, 1};
    int64_t n = 0, G = 0;
    bool have_particles = false, have_grid = false, have_disp = false, have_vel = false, have_mass = false;
    bool z_flat = true;  // all particle and grid z equal to zero (2D fast path)
    bool grid_z_flat = true;
    // host-uploaded single-frame particle data
    DBuf<double> d_pos, d_diam, d_mass, d_vel, d_disp;
    DBuf<double> d_grid;
    // per-batch work arrays
    DBuf<int> d_pcell, d_pslot, d_ccount, d_cstart, d_order, d_bcount, d_maxcount;
    DBuf<double> d_spos, d_sdiam, d_skin, d_sdisp;
    DBuf<unsigned char> d_bonds;
    DBuf<double> d_out;
    DBuf<double> d_gl;
    DBuf<unsigned long long> d_stats;
    std::vector<double> gl_s, gl_w;
    int cap = 0;
    int lpg = 4;
    int frames_per_launch = 8;
    pgt_ik_stats stats{};
};

namespace {

int check_engine(pgt_ik_engine *e)
{
    if (!e) return fail(PGT_ERR_ARG, "null engine");
    return 0;
}

int ensure_events(pgt_ik_engine *e, size_t want)
{
    while (e->ev.size() < want) {
        hipEvent_t v = nullptr;
        HIPCHK(hipEventCreate(&v));
        e->ev.push_back(v);
    }
    return 0;
}

PotParams make_pot(const pgt_ik_params &p)
{
    PotParams P{};
    P.kind = p.potential_kind;
    P.v0 = p.pot_params[0];
    P.eps = p.pot_params[1];
    P.c0 = p.pot_params[2];
    P.c1 = p.pot_params[3];
    P.c2 = p.pot_params[4];
    P.c3 = p.n_pot_params > 5 ? p.pot_params[5] : 0.0;
    P.rcutsq = p.scaled_rcut * p.scaled_rcut;
    return P;
}

int validate_params(const pgt_ik_params *p)
{
    if (!p) return fail(PGT_ERR_ARG, "null params");
    if (p->dim != 2 && p->dim != 3) return fail(PGT_ERR_ARG, "dim must be 2 or 3");
    if (p->potential_kind < 0 || p->potential_kind > 4) return fail(PGT_ERR_ARG, "unknown potential kind");
    const int need = p->potential_kind == PGT_POT_POLY10 ? 6 : 5;
    if (p->n_pot_params < need)
        return fail(PGT_ERR_ARG, "potential needs " + std::to_string(need) + " parameters");
    if (p->cg_kind < 0 || p->cg_kind > 2) return fail(PGT_ERR_ARG, "unknown coarse-graining kernel");
    if (!(p->cg_rcut > 0)) return fail(PGT_ERR_ARG, "cg_rcut must be positive");
    if (!(p->max_rcut >= 0)) return fail(PGT_ERR_ARG, "max_rcut must be non-negative");
    if (p->quad_mode == PGT_QUAD_FIXED) {
        if (p->quad_order < 1 || p->quad_order > 1024) return fail(PGT_ERR_ARG, "quadrature order must be in 1..1024");
    } else if (p->quad_mode == PGT_QUAD_ADAPTIVE) {
        return fail(PGT_ERR_UNSUPPORTED, "adaptive quadrature is not available in the field engine");
    } else {
        return fail(PGT_ERR_ARG, "unknown quadrature mode");
    }
    if (p->precision != PGT_PREC_FP64 && p->precision != PGT_PREC_FP32)
        return fail(PGT_ERR_ARG, "unknown precision");
    if (p->precision == PGT_PREC_FP32 && p->cg_kind != PGT_CG_OCTIC)
        return fail(PGT_ERR_UNSUPPORTED, "the FP32 form is built for the Octic kernel only");
    return 0;
}

Geom make_geom(const pgt_ik_engine *e, int D)
{
    Geom g{};
    const double rmax = std::max(e->prm.cg_rcut, e->prm.max_rcut);
    g.ncell = 1;
    for (int a = 0; a < 3; ++a) {
        g.L[a] = e->L[a];
        g.invL[a] = 1.0 / e->L[a];
        g.lo[a] = e->lo[a];
        g.periodic[a] = e->periodic[a];
        int nc = (int)std::floor(e->L[a] / rmax);
        if (!e->periodic[a] || nc < 3 || (a == 2 && D == 2)) nc = 1;
        g.nc[a] = nc;
        g.cs[a] = e->L[a] / nc;
        g.ncell *= nc;
    }
    g.rc2 = e->prm.cg_rcut * e->prm.cg_rcut;
    g.inv_rc = 1.0 / e->prm.cg_rcut;
    g.inv_rc2 = 1.0 / g.rc2;
    g.mr2 = e->prm.max_rcut * e->prm.max_rcut;
    return g;
}

// Dispatch tables ----------------------------------------------------------
struct LaunchCtx {
    Geom g;
    int n, G, nf;
    int cap;
    int lpg;
    const double *grid;
    GLRule R;
    const double *gld;
    double cgnorm;
    const int *cstart;
    const double *spos, *skin;
    const int *bcount;
    const void *bonds;
    GatherOut out;
    int count_stats;
    hipStream_t stream;
};

template <int D, int OD, int CG, int NQ, bool VIR, bool ELA, int LPG, typename ER>
void launch_gather_t(const LaunchCtx &c)
{
    const int threads = 256;
    const long long total = (long long)c.G * LPG;
    dim3 grid((unsigned)((total + threads - 1) / threads), (unsigned)c.nf);
    hipLaunchKernelGGL((k_gather<D, OD, CG, NQ, VIR, ELA, LPG, ER>), grid, dim3(threads), 0, c.stream, c.g, c.n,
                       c.G, c.grid, c.R, c.gld, c.cgnorm, c.cstart, c.spos, c.skin, c.cap, c.bcount,
                       (const ER *)c.bonds, c.out, c.count_stats);
}

template <int D, int OD, int CG, int NQ, bool VIR, bool ELA, typename ER>
void launch_gather_l(const LaunchCtx &c)
{
    switch (c.lpg) {
    case 1: launch_gather_t<D, OD, CG, NQ, VIR, ELA, 1, ER>(c); break;
    case 8: launch_gather_t<D, OD, CG, NQ, VIR, ELA, 8, ER>(c); break;
    default: launch_gather_t<D, OD, CG, NQ, VIR, ELA, 4, ER>(c); break;
    }
}

template <int D, int OD, int CG, int NQ, typename ER>
void launch_gather_f(const LaunchCtx &c, unsigned fields)
{
    const bool V = fields & PGT_FIELD_VIRIAL, E = fields & PGT_FIELD_ELASTIC;
    if (V && E) launch_gather_l<D, OD, CG, NQ, true, true, ER>(c);
    else if (V) launch_gather_l<D, OD, CG, NQ, true, false, ER>(c);
    else if (E) launch_gather_l<D, OD, CG, NQ, false, true, ER>(c);
    else launch_gather_l<D, OD, CG, NQ, false, false, ER>(c);
}

// Octic gets the unrolled 5-node rule (the driver scripts' order) and the
// FP32 form; the other kernels run the generic FP64 rule.
template <int D, int OD, typename ER>
void launch_gather_c(const LaunchCtx &c, unsigned fields, int cg)
{
    if (cg == 0) {
        if (c.R.n == 5 && !(D == 3 && OD == 2)) launch_gather_f<D, OD, 0, 5, ER>(c, fields);
        else launch_gather_f<D, OD, 0, 0, ER>(c, fields);
    } else if constexpr (std::is_same<ER, double>::value) {
        if (cg == 1) launch_gather_f<D, OD, 1, 0, double>(c, fields);
        else launch_gather_f<D, OD, 2, 0, double>(c, fields);
    }
}

template <int D, int OD, typename ER>
void launch_bonds_t(const LaunchCtx &c, unsigned fields, const PotParams &P, const double *sdiam, const double *sdisp,
                    int *bcount, void *bonds, int *maxcount)
{
    const bool V = fields & PGT_FIELD_VIRIAL, E = fields & PGT_FIELD_ELASTIC;
    dim3 grid((unsigned)((c.n + 255) / 256), (unsigned)c.nf);
    if (V && E)
        hipLaunchKernelGGL((k_bonds<D, OD, true, true, ER>), grid, dim3(256), 0, c.stream, c.g, c.n, P, c.cgnorm,
                           c.cstart, c.spos, sdiam, sdisp, c.cap, bcount, (ER *)bonds, maxcount);
    else if (E)
        hipLaunchKernelGGL((k_bonds<D, OD, false, true, ER>), grid, dim3(256), 0, c.stream, c.g, c.n, P, c.cgnorm,
                           c.cstart, c.spos, sdiam, sdisp, c.cap, bcount, (ER *)bonds, maxcount);
    else
        hipLaunchKernelGGL((k_bonds<D, OD, true, false, ER>), grid, dim3(256), 0, c.stream, c.g, c.n, P, c.cgnorm,
                           c.cstart, c.spos, sdiam, sdisp, c.cap, bcount, (ER *)bonds, maxcount);
}

int entry_stride(int D, int OD, unsigned fields)
{
    const int KN = OD * (OD + 1) / 2;
    const int NT = ((fields & PGT_FIELD_VIRIAL) ? 1 : 0) + ((fields & PGT_FIELD_ELASTIC) ? 1 : 0);
    return D + 1 + NT * KN;
}

// Run nf frames starting at the given device arrays.
int run_frames(pgt_ik_engine *e, int nf, int64_t n_particles, const double *pos, const double *diam, const double *mass,
               const double *vel, const double *disp, unsigned fields, double *ovir, double *okin, double *oela,
               bool z_flat, hipEvent_t ev_begin, hipEvent_t ev_gather, hipEvent_t ev_end)
{
    const int dim = e->prm.dim;
    const int D = (dim == 3 || !z_flat) ? 3 : 2;
    const int OD = dim;
    const int n = (int)n_particles;
    const int G = (int)e->G;
    Geom g = make_geom(e, D);
    const int KN = OD * (OD + 1) / 2;
    const int PS = (D == 2) ? 2 : 4;
    const bool pairs = fields & (PGT_FIELD_VIRIAL | PGT_FIELD_ELASTIC);
    const bool kin = fields & PGT_FIELD_KINETIC;
    const bool ela = fields & PGT_FIELD_ELASTIC;
    // FP32 form: Octic kernel on the 2D (z == 0) and 3D paths; a 2D system
    // with nonzero z runs in FP64.
    const bool fp32 = e->prm.precision == PGT_PREC_FP32 && !(D == 3 && OD == 2);
    const size_t esz = fp32 ? sizeof(float) : sizeof(double);
    const int ES = entry_stride(D, OD, fields);
    int rc;
    if ((rc = e->d_pcell.ensure((size_t)nf * n)) || (rc = e->d_pslot.ensure((size_t)nf * n)) ||
        (rc = e->d_order.ensure((size_t)nf * n)) || (rc = e->d_ccount.ensure((size_t)nf * g.ncell)) ||
        (rc = e->d_cstart.ensure((size_t)nf * (g.ncell + 1))) || (rc = e->d_spos.ensure((size_t)nf * n * PS)) ||
        (rc = e->d_sdiam.ensure((size_t)nf * n)) || (rc = e->d_bcount.ensure((size_t)nf * n)))
        return rc;
    if (kin && (rc = e->d_skin.ensure((size_t)nf * n * KN))) return rc;
    if (ela && (rc = e->d_sdisp.ensure((size_t)nf * n * 3))) return rc;
    hipStream_t s = e->stream;
    if (ev_begin) HIPCHK(hipEventRecord(ev_begin, s));
    HIPCHK(hipMemsetAsync(e->d_ccount.p, 0, sizeof(int) * (size_t)nf * g.ncell, s));

    dim3 pgrid((unsigned)((n + 255) / 256), (unsigned)nf);
    hipLaunchKernelGGL(k_cell_count, pgrid, dim3(256), 0, s, g, n, pos, e->d_pcell.p, e->d_pslot.p, e->d_ccount.p);
    hipLaunchKernelGGL(k_cell_scan, dim3(nf), dim3(1024), 0, s, g.ncell, e->d_ccount.p, e->d_cstart.p);
    hipLaunchKernelGGL(k_cell_fill, pgrid, dim3(256), 0, s, n, g.ncell, e->d_pcell.p, e->d_pslot.p, e->d_cstart.p,
                       e->d_order.p);
    dim3 cgrid((unsigned)((g.ncell + 255) / 256), (unsigned)nf);
    double *skin = kin ? e->d_skin.p : nullptr;
    double *sdisp = ela ? e->d_sdisp.p : nullptr;
    if (D == 2)
        hipLaunchKernelGGL((k_cell_sort<2, 2>), cgrid, dim3(256), 0, s, n, g.ncell, e->d_cstart.p, e->d_order.p, pos,
                           diam, mass, vel, disp, e->d_spos.p, e->d_sdiam.p, skin, sdisp);
    else if (OD == 2)
        hipLaunchKernelGGL((k_cell_sort<3, 2>), cgrid, dim3(256), 0, s, n, g.ncell, e->d_cstart.p, e->d_order.p, pos,
                           diam, mass, vel, disp, e->d_spos.p, e->d_sdiam.p, skin, sdisp);
    else
        hipLaunchKernelGGL((k_cell_sort<3, 3>), cgrid, dim3(256), 0, s, n, g.ncell, e->d_cstart.p, e->d_order.p, pos,
                           diam, mass, vel, disp, e->d_spos.p, e->d_sdiam.p, skin, sdisp);
    HIPCHK(hipGetLastError());

    LaunchCtx c{};
    c.g = g;
    c.n = n;
    c.G = G;
    c.nf = nf;
    c.lpg = e->lpg;
    c.grid = e->d_grid.p;
    c.R.n = (int)e->gl_s.size();
    for (int k = 0; k < c.R.n && k < kMaxQArg; ++k) {
        c.R.s[k] = e->gl_s[k];
        c.R.w[k] = e->gl_w[k];
    }
    c.gld = e->d_gl.p;
    c.cgnorm = cg_norm(e->prm.cg_kind, e->prm.cg_rcut);
    c.cstart = e->d_cstart.p;
    c.spos = e->d_spos.p;
    c.skin = skin;
    c.out.vir = ovir;
    c.out.kin = okin;
    c.out.ela = oela;
    c.out.n_ball = e->d_stats.p;
    c.out.n_contrib = e->d_stats.p + 1;
    c.count_stats = 1;
    c.stream = s;

    if (pairs) {
        // the bond list may need to grow: retry once with the observed maximum
        // entries beyond the capacity are counted, not written; the caller
        // checks d_maxcount after the job and re-runs with a larger capacity
        if (e->cap < 1) e->cap = D == 3 ? 32 : 16;
        if ((rc = e->d_bonds.ensure((size_t)nf * n * (size_t)e->cap * ES * esz))) return rc;
        c.cap = e->cap;
        const PotParams P = make_pot(e->prm);
        if (D == 2) {
            if (fp32) launch_bonds_t<2, 2, float>(c, fields, P, e->d_sdiam.p, sdisp, e->d_bcount.p, e->d_bonds.p, e->d_maxcount.p);
            else launch_bonds_t<2, 2, double>(c, fields, P, e->d_sdiam.p, sdisp, e->d_bcount.p, e->d_bonds.p, e->d_maxcount.p);
        } else if (OD == 2) {
            launch_bonds_t<3, 2, double>(c, fields, P, e->d_sdiam.p, sdisp, e->d_bcount.p, e->d_bonds.p, e->d_maxcount.p);
        } else {
            if (fp32) launch_bonds_t<3, 3, float>(c, fields, P, e->d_sdiam.p, sdisp, e->d_bcount.p, e->d_bonds.p, e->d_maxcount.p);
            else launch_bonds_t<3, 3, double>(c, fields, P, e->d_sdiam.p, sdisp, e->d_bcount.p, e->d_bonds.p, e->d_maxcount.p);
        }
        HIPCHK(hipGetLastError());
        e->stats.bond_capacity = e->cap;
    } else {
        c.cap = 1;
    }
    c.bcount = e->d_bcount.p;
    c.bonds = e->d_bonds.p;
    if (ev_gather) HIPCHK(hipEventRecord(ev_gather, s));
    const int cg = e->prm.cg_kind;
    if (D == 2) {
        if (fp32) launch_gather_c<2, 2, float>(c, fields, cg);
        else launch_gather_c<2, 2, double>(c, fields, cg);
    } else if (OD == 2) {
        launch_gather_c<3, 2, double>(c, fields, cg);
    } else {
        if (fp32) launch_gather_c<3, 3, float>(c, fields, cg);
        else launch_gather_c<3, 3, double>(c, fields, cg);
    }
    HIPCHK(hipGetLastError());
    if (ev_end) HIPCHK(hipEventRecord(ev_end, s));
    return 0;
}

struct FrameSet {
    int nframes;
    int64_t n;
    const double *pos, *diam, *mass, *vel, *disp;
    double *vir, *kin, *ela;  // device outputs, nframes * G * D * D each
    bool z_flat;
};

int ensure_events(pgt_ik_engine *e, size_t want);

// Run every batch of a job on the engine's stream, then check the bond-list
// capacity once; on overflow grow it and run the whole job again.
int run_job(pgt_ik_engine *e, const FrameSet &fs, unsigned fields)
{
    const int OD = e->prm.dim;
    const size_t fsz = (size_t)e->G * OD * OD;
    const int fpl = std::max(1, e->frames_per_launch);
    const int nb = (fs.nframes + fpl - 1) / fpl;
    int rc;
    if ((rc = ensure_events(e, (size_t)nb * 3))) return rc;
    if ((rc = e->d_maxcount.ensure(1)) || (rc = e->d_stats.ensure(2))) return rc;
    int mx = 0;
    for (int attempt = 0; attempt < 3; ++attempt) {
        HIPCHK(hipMemsetAsync(e->d_maxcount.p, 0, sizeof(int), e->stream));
        HIPCHK(hipMemsetAsync(e->d_stats.p, 0, sizeof(unsigned long long) * 2, e->stream));
        for (int b = 0; b < nb; ++b) {
            const int f0 = b * fpl;
            const int nf = std::min(fpl, fs.nframes - f0);
            const size_t po = (size_t)f0 * fs.n;
            rc = run_frames(e, nf, fs.n, fs.pos + po * 3, fs.diam + po, fs.mass ? fs.mass + po : nullptr,
                            fs.vel ? fs.vel + po * 3 : nullptr, fs.disp ? fs.disp + po * 3 : nullptr, fields,
                            fs.vir ? fs.vir + (size_t)f0 * fsz : nullptr, fs.kin ? fs.kin + (size_t)f0 * fsz : nullptr,
                            fs.ela ? fs.ela + (size_t)f0 * fsz : nullptr, fs.z_flat, e->ev[3 * b],
                            e->ev[3 * b + 1], e->ev[3 * b + 2]);
            if (rc) return rc;
        }
        mx = 0;
        HIPCHK(hipMemcpyAsync(&mx, e->d_maxcount.p, sizeof(int), hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        if (mx <= e->cap) break;
        e->cap = ((mx + 7) / 8) * 8;
        if (attempt == 2) return fail(PGT_ERR_HIP, "bond list capacity did not settle");
    }
    unsigned long long h[2] = {0, 0};
    HIPCHK(hipMemcpy(h, e->d_stats.p, sizeof(h), hipMemcpyDeviceToHost));
    double g_ms = 0.0;
    for (int b = 0; b < nb; ++b) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, e->ev[3 * b + 1], e->ev[3 * b + 2]) == hipSuccess) g_ms += t;
    }
    float tot = 0.0f;
    (void)hipEventElapsedTime(&tot, e->ev[0], e->ev[3 * (nb - 1) + 2]);
    e->stats.n_particles = fs.n;
    e->stats.n_grid = e->G;
    e->stats.n_ball = (int64_t)h[0];
    e->stats.n_contrib = (int64_t)h[1];
    e->stats.gather_ms = g_ms;
    e->stats.total_ms = tot;
    e->stats.bond_capacity = e->cap;
    e->stats.max_bonds = mx;
    return 0;
}

}  // namespace

// ==========================================================================
// C ABI
// ==========================================================================
extern "C" {

const char *pgt_last_error(void) { return g_err.c_str(); }

int pgt_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int pgt_ik_create(const pgt_ik_params *params, pgt_ik_engine **out)
{
    if (!out) return fail(PGT_ERR_ARG, "null output pointer");
    *out = nullptr;
    int rc = validate_params(params);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(PGT_ERR_HIP, "no HIP device available: the IK engine runs on the GPU only");
    pgt_ik_engine *e = new pgt_ik_engine();
    e->prm = *params;
    (void)hipGetDevice(&e->device);
    hipError_t he = hipStreamCreateWithFlags(&e->own_stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete e;
        return fail(PGT_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
    }
    e->stream = e->own_stream;
    if (gl_rule(params->quad_order, e->gl_s, e->gl_w) != 0) {
        pgt_ik_destroy(e);
        return fail(PGT_ERR_ARG, "bad quadrature order");
    }
    std::vector<double> packed(e->gl_s);
    packed.insert(packed.end(), e->gl_w.begin(), e->gl_w.end());
    if ((rc = e->d_gl.ensure(packed.size()))) {
        pgt_ik_destroy(e);
        return rc;
    }
    he = hipMemcpy(e->d_gl.p, packed.data(), packed.size() * sizeof(double), hipMemcpyHostToDevice);
    if (he != hipSuccess) {
        pgt_ik_destroy(e);
        return fail(PGT_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(he));
    }
    *out = e;
    return PGT_OK;
}

void pgt_ik_destroy(pgt_ik_engine *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    DBuf<double> *dbufs[] = {&e->d_pos, &e->d_diam, &e->d_mass, &e->d_vel, &e->d_disp, &e->d_grid,
                             &e->d_spos, &e->d_sdiam, &e->d_skin, &e->d_sdisp, &e->d_out, &e->d_gl};
    for (auto *b : dbufs) b->release();
    DBuf<int> *ibufs[] = {&e->d_pcell, &e->d_pslot, &e->d_ccount, &e->d_cstart, &e->d_order, &e->d_bcount, &e->d_maxcount};
    for (auto *b : ibufs) b->release();
    e->d_bonds.release();
    e->d_stats.release();
    for (hipEvent_t v : e->ev) (void)hipEventDestroy(v);
    if (e->own_stream) (void)hipStreamDestroy(e->own_stream);
    delete e;
}

int pgt_ik_set_box(pgt_ik_engine *e, const double boxsize[3], const double lowerbound[3], const int periodic[3])
{
    if (int rc = check_engine(e)) return rc;
    if (!boxsize || !periodic) return fail(PGT_ERR_ARG, "null box argument");
    const double rmax = std::max(e->prm.cg_rcut, e->prm.max_rcut);
    for (int a = 0; a < 3; ++a) {
        if (!(boxsize[a] > 0)) return fail(PGT_ERR_ARG, "box lengths must be positive");
        const bool per = periodic[a] != 0;
        if (per && !(rmax < 0.5 * boxsize[a]))
            return fail(PGT_ERR_RANGE, "search radius max(cg rcut, max_rcut) must be below half the box length");
        e->L[a] = boxsize[a];
        e->lo[a] = lowerbound ? lowerbound[a] : -0.5 * boxsize[a];
        e->periodic[a] = per ? 1 : 0;
    }
    e->box_set = true;
    return PGT_OK;
}

int pgt_ik_set_particles(pgt_ik_engine *e, int64_t n, const double *pos, const double *diam, const double *mass,
                         const double *vel, const double *disp)
{
    if (int rc = check_engine(e)) return rc;
    if (n < 0 || n > (int64_t)1 << 30) return fail(PGT_ERR_ARG, "particle count out of range");
    if (n > 0 && (!pos || !diam)) return fail(PGT_ERR_ARG, "positions and diameters are required");
    int rc;
    if ((rc = e->d_pos.ensure((size_t)n * 3)) || (rc = e->d_diam.ensure((size_t)n))) return rc;
    hipStream_t s = e->stream;
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(e->d_pos.p, pos, sizeof(double) * n * 3, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(e->d_diam.p, diam, sizeof(double) * n, hipMemcpyHostToDevice, s));
    }
    e->have_mass = mass != nullptr;
    e->have_vel = vel != nullptr;
    e->have_disp = disp != nullptr;
    if (mass && n > 0) {
        if ((rc = e->d_mass.ensure((size_t)n))) return rc;
        HIPCHK(hipMemcpyAsync(e->d_mass.p, mass, sizeof(double) * n, hipMemcpyHostToDevice, s));
    }
    if (vel && n > 0) {
        if ((rc = e->d_vel.ensure((size_t)n * 3))) return rc;
        HIPCHK(hipMemcpyAsync(e->d_vel.p, vel, sizeof(double) * n * 3, hipMemcpyHostToDevice, s));
    }
    if (disp && n > 0) {
        if ((rc = e->d_disp.ensure((size_t)n * 3))) return rc;
        HIPCHK(hipMemcpyAsync(e->d_disp.p, disp, sizeof(double) * n * 3, hipMemcpyHostToDevice, s));
    }
    bool flat = true;
    for (int64_t i = 0; i < n && flat; ++i) flat = pos[3 * i + 2] == 0.0;
    e->z_flat = flat;
    HIPCHK(hipStreamSynchronize(s));
    e->n = n;
    e->have_particles = true;
    return PGT_OK;
}

int pgt_ik_set_grid(pgt_ik_engine *e, int64_t g, const double *grid)
{
    if (int rc = check_engine(e)) return rc;
    if (g < 0 || g > (int64_t)1 << 30) return fail(PGT_ERR_ARG, "grid size out of range");
    if (g > 0 && !grid) return fail(PGT_ERR_ARG, "null grid");
    int rc;
    if ((rc = e->d_grid.ensure((size_t)g * 3))) return rc;
    if (g > 0) HIPCHK(hipMemcpyAsync(e->d_grid.p, grid, sizeof(double) * g * 3, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    bool flat = true;
    for (int64_t i = 0; i < g && flat; ++i) flat = grid[3 * i + 2] == 0.0;
    e->grid_z_flat = flat;
    e->G = g;
    e->have_grid = true;
    return PGT_OK;
}

int pgt_ik_set_stream(pgt_ik_engine *e, void *hip_stream)
{
    if (int rc = check_engine(e)) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    e->stream = hip_stream ? (hipStream_t)hip_stream : e->own_stream;
    return PGT_OK;
}

int pgt_ik_set_tuning(pgt_ik_engine *e, int lanes_per_point, int frames_per_launch)
{
    if (int rc = check_engine(e)) return rc;
    if (lanes_per_point) {
        if (lanes_per_point != 1 && lanes_per_point != 2 && lanes_per_point != 4 && lanes_per_point != 8 &&
            lanes_per_point != 16)
            return fail(PGT_ERR_ARG, "lanes per point must be 1, 2, 4, 8 or 16");
        e->lpg = lanes_per_point;
    }
    if (frames_per_launch) {
        if (frames_per_launch < 1 || frames_per_launch > 65535) return fail(PGT_ERR_ARG, "bad frames per launch");
        e->frames_per_launch = frames_per_launch;
    }
    return PGT_OK;
}

int pgt_ik_get_stats(pgt_ik_engine *e, pgt_ik_stats *out)
{
    if (int rc = check_engine(e)) return rc;
    if (!out) return fail(PGT_ERR_ARG, "null stats");
    *out = e->stats;
    return PGT_OK;
}

int pgt_ik_compute(pgt_ik_engine *e, unsigned fields, double *virial, double *kinetic, double *elastic)
{
    if (int rc = check_engine(e)) return rc;
    if (!e->box_set || !e->have_particles || !e->have_grid)
        return fail(PGT_ERR_STATE, "set_box, set_particles and set_grid must precede compute");
    if (fields == 0 || (fields & ~7u)) return fail(PGT_ERR_ARG, "bad field mask");
    if (((fields & PGT_FIELD_VIRIAL) && !virial) || ((fields & PGT_FIELD_KINETIC) && !kinetic) ||
        ((fields & PGT_FIELD_ELASTIC) && !elastic))
        return fail(PGT_ERR_ARG, "null output for a selected field");
    const int OD = e->prm.dim;
    const size_t fsz = (size_t)e->G * OD * OD;
    e->stats = pgt_ik_stats{};
    if (e->G == 0) return PGT_OK;
    if (e->n == 0) {
        if (virial) std::memset(virial, 0, fsz * sizeof(double));
        if (kinetic) std::memset(kinetic, 0, fsz * sizeof(double));
        if (elastic) std::memset(elastic, 0, fsz * sizeof(double));
        return PGT_OK;
    }
    int rc;
    if ((rc = e->d_out.ensure(fsz * 3))) return rc;
    FrameSet fs{};
    fs.nframes = 1;
    fs.n = e->n;
    fs.pos = e->d_pos.p;
    fs.diam = e->d_diam.p;
    fs.mass = e->have_mass ? e->d_mass.p : nullptr;
    fs.vel = e->have_vel ? e->d_vel.p : nullptr;
    fs.disp = e->have_disp ? e->d_disp.p : nullptr;
    fs.vir = e->d_out.p;
    fs.kin = e->d_out.p + fsz;
    fs.ela = e->d_out.p + 2 * fsz;
    fs.z_flat = e->z_flat && e->grid_z_flat;
    if ((rc = run_job(e, fs, fields))) return rc;
    if (fields & PGT_FIELD_VIRIAL) HIPCHK(hipMemcpy(virial, fs.vir, fsz * sizeof(double), hipMemcpyDeviceToHost));
    if (fields & PGT_FIELD_KINETIC) HIPCHK(hipMemcpy(kinetic, fs.kin, fsz * sizeof(double), hipMemcpyDeviceToHost));
    if (fields & PGT_FIELD_ELASTIC) HIPCHK(hipMemcpy(elastic, fs.ela, fsz * sizeof(double), hipMemcpyDeviceToHost));
    return PGT_OK;
}

int pgt_ik_compute_frames(pgt_ik_engine *e, int nframes, int64_t n, const double *d_pos, const double *d_diam,
                          const double *d_mass, const double *d_vel, const double *d_disp, unsigned fields,
                          double *d_out_virial, double *d_out_kinetic, double *d_out_elastic)
{
    if (int rc = check_engine(e)) return rc;
    if (!e->box_set || !e->have_grid) return fail(PGT_ERR_STATE, "set_box and set_grid must precede compute_frames");
    if (nframes < 0 || n < 0 || n > (int64_t)1 << 30) return fail(PGT_ERR_ARG, "bad frame or particle count");
    if (fields == 0 || (fields & ~7u)) return fail(PGT_ERR_ARG, "bad field mask");
    if (((fields & PGT_FIELD_VIRIAL) && !d_out_virial) || ((fields & PGT_FIELD_KINETIC) && !d_out_kinetic) ||
        ((fields & PGT_FIELD_ELASTIC) && !d_out_elastic))
        return fail(PGT_ERR_ARG, "null output for a selected field");
    if (n > 0 && (!d_pos || !d_diam)) return fail(PGT_ERR_ARG, "positions and diameters are required");
    e->stats = pgt_ik_stats{};
    if (nframes == 0 || e->G == 0) return PGT_OK;
    const int OD = e->prm.dim;
    const size_t fsz = (size_t)e->G * OD * OD;
    if (n == 0) {
        for (double *o : {d_out_virial, d_out_kinetic, d_out_elastic})
            if (o) HIPCHK(hipMemsetAsync(o, 0, fsz * nframes * sizeof(double), e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        return PGT_OK;
    }
    FrameSet fs{};
    fs.nframes = nframes;
    fs.n = n;
    fs.pos = d_pos;
    fs.diam = d_diam;
    fs.mass = d_mass;
    fs.vel = d_vel;
    fs.disp = d_disp;
    fs.vir = (fields & PGT_FIELD_VIRIAL) ? d_out_virial : nullptr;
    fs.kin = (fields & PGT_FIELD_KINETIC) ? d_out_kinetic : nullptr;
    fs.ela = (fields & PGT_FIELD_ELASTIC) ? d_out_elastic : nullptr;
    // Device-resident 2D frames take the z == 0 fast path when the grid is
    // flat: positions' z components are not read for such frames.
    fs.z_flat = e->grid_z_flat;
    return run_job(e, fs, fields);
}

int pgt_eval_pair_force(int potential_kind, const double *pot_params, int n_pot_params, double scaled_rcut, int64_t n,
                        const double *rij, const double *di, const double *dj, double *out_force_divr,
                        double *out_energy, double *out_rcutsq)
{
    pgt_ik_params p{};
    p.dim = 3;
    p.potential_kind = potential_kind;
    p.scaled_rcut = scaled_rcut;
    p.n_pot_params = n_pot_params;
    for (int k = 0; k < 6 && k < n_pot_params; ++k) p.pot_params[k] = pot_params[k];
    p.cg_kind = 0;
    p.cg_rcut = 1.0;
    p.quad_order = 1;
    if (int rc = validate_params(&p)) return rc;
    if (n < 0 || n > (int64_t)1 << 28) return fail(PGT_ERR_ARG, "bad count");
    if (n == 0) return PGT_OK;
    if (!rij || !di || !dj) return fail(PGT_ERR_ARG, "null input");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(PGT_ERR_HIP, "no HIP device available: pair forces are evaluated on the GPU");
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(double) * n * 8));
    double *drij = d, *ddi = d + 3 * n, *ddj = d + 4 * n, *fo = d + 5 * n, *eo = d + 6 * n, *ro = d + 7 * n;
    hipError_t he = hipMemcpy(drij, rij, sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(ddi, di, sizeof(double) * n, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(ddj, dj, sizeof(double) * n, hipMemcpyHostToDevice);
    if (he == hipSuccess) {
        hipLaunchKernelGGL(k_eval_force, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, make_pot(p), (int)n,
                           drij, ddi, ddj, fo, eo, ro);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipDeviceSynchronize();
    if (he == hipSuccess && out_force_divr) he = hipMemcpy(out_force_divr, fo, sizeof(double) * n, hipMemcpyDeviceToHost);
    if (he == hipSuccess && out_energy) he = hipMemcpy(out_energy, eo, sizeof(double) * n, hipMemcpyDeviceToHost);
    if (he == hipSuccess && out_rcutsq) he = hipMemcpy(out_rcutsq, ro, sizeof(double) * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (he != hipSuccess) return fail(PGT_ERR_HIP, std::string("pair force evaluation: ") + hipGetErrorString(he));
    return PGT_OK;
}

int pgt_eval_cgfunc(int cg_kind, double cg_rcut, int quad_mode, int quad_order, int adapt_limit, double adapt_epsrel,
                    double adapt_epsabs, int mode, int64_t n, const double *s, const double *x, const double *ri,
                    const double *rij, double *out)
{
    if (cg_kind < 0 || cg_kind > 2) return fail(PGT_ERR_ARG, "unknown coarse-graining kernel");
    if (mode < 0 || mode > 2) return fail(PGT_ERR_ARG, "bad mode");
    if (n < 0 || n > (int64_t)1 << 28) return fail(PGT_ERR_ARG, "bad count");
    if (n == 0) return PGT_OK;
    if (!x || !ri || !out || (mode >= 1 && !rij) || (mode == 1 && !s)) return fail(PGT_ERR_ARG, "null input");
    std::vector<double> packed;
    int nq = 0;
    if (mode == 2) {
        if (quad_mode != PGT_QUAD_FIXED)
            return fail(PGT_ERR_UNSUPPORTED, "adaptive quadrature is not available in the evaluator");
        std::vector<double> gs, gw;
        if (gl_rule(quad_order, gs, gw) != 0) return fail(PGT_ERR_ARG, "quadrature order must be in 1..1024");
        nq = quad_order;
        packed = gs;
        packed.insert(packed.end(), gw.begin(), gw.end());
    }
    (void)adapt_limit;
    (void)adapt_epsrel;
    (void)adapt_epsabs;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(PGT_ERR_HIP, "no HIP device available: kernel functions are evaluated on the GPU");
    double *d = nullptr;
    const size_t tot = (size_t)n * 11 + packed.size();
    HIPCHK(hipMalloc(&d, sizeof(double) * tot));
    double *dx = d, *dri = d + 3 * n, *drij = d + 6 * n, *ds = d + 9 * n, *dout = d + 10 * n, *dgl = d + 11 * n;
    hipError_t he = hipMemcpy(dx, x, sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(dri, ri, sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    if (he == hipSuccess && rij) he = hipMemcpy(drij, rij, sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    if (he == hipSuccess && s) he = hipMemcpy(ds, s, sizeof(double) * n, hipMemcpyHostToDevice);
    if (he == hipSuccess && !packed.empty())
        he = hipMemcpy(dgl, packed.data(), sizeof(double) * packed.size(), hipMemcpyHostToDevice);
    if (he == hipSuccess) {
        hipLaunchKernelGGL(k_eval_cg, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, cg_kind, cg_rcut, mode,
                           (int)n, s ? ds : nullptr, dx, dri, rij ? drij : nullptr, dgl, nq, dout);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipDeviceSynchronize();
    if (he == hipSuccess) he = hipMemcpy(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (he != hipSuccess) return fail(PGT_ERR_HIP, std::string("kernel evaluation: ") + hipGetErrorString(he));
    return PGT_OK;
}

}  // extern "C"
