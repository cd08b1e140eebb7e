// ik_engine.hip -- MI355X (gfx950) Irving-Kirkwood stress-field engine behind
// the C ABI declared in include/pgt_ik.h.
//
// Per batch of frames the engine runs (kernels in ik_kernels.hpp):
//   1. cell list    stable radix sort of the particles by cell (cells of edge
//                   >= the largest search radius / reach); inside a cell
//                   particles keep their input order, so every later sum has
//                   a fixed order
//   2. bond list    for every particle i, the neighbours j within max_rcut
//                   with a nonzero pair force, each stored once with the
//                   scaled bond vector and the force coefficient: the force is
//                   evaluated once per ordered pair instead of once per
//                   (grid point, i, j) triple as in the reference loop
//   3. field gather one lane per 1, 2 or 4 grid points: the particles inside
//                   the coarse-graining balls are compacted into a per-lane
//                   LDS list, then the lane integrates every bond of every
//                   listed particle (5-node Gauss-Legendre on the Octic kernel
//                   by default) and accumulates the tensors in FP64
// Frames are stacked along blockIdx.y so one launch covers many frames (the
// cell-list kernels deal whole frames to XCDs, see frame_block).  The learned
// bond-integral surrogate (ik_surrogate.hpp) has entry points of its own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "ik_band.hpp"
#include "ik_kernels.hpp"
#include "ik_launch.hpp"
#include "ik_sweep_kernels.hpp"
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

template <typename T>
int fill_async(T *p, T v, size_t n, hipStream_t s)
{
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
    HIPCHK(hipGetLastError());
    return 0;
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

// Device memory the adaptive rule may take for interval tables beyond the
// kQagMax private ones (tables of the caller's limit, 32 bytes an interval,
// one per lane): compute jobs fit their frames per launch to it.
constexpr size_t kQagWorkspaceBytes = (size_t)4 << 30;

// Kernel and bond integral in the reference's arithmetic form: the integrand
// is Delta(|x - (ri + s rij)|^2) (CoarseGrainFunction.h:42-45), summed over
// the unfolded Gauss-Legendre rule.
__global__ void k_eval_cg(int cg, double rc, int mode, int n, const double *s, const double *x, const double *ri,
                          const double *rij, const double *gld, int nq, int adapt, AdaptQ Q, int *capped,
                          double *ws, double *out)
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
    else if (adapt) {
        // interval tables of the caller's limit in ws when it exceeds kQagMax
        bool cp = false;
        if (ws) {
            QagGlobal w{ws + i, (size_t)n, Q.limit};
            out[i] = qag_dev(delta_at, Q, cp, w);
        } else {
            QagPrivate w;
            out[i] = qag_dev(delta_at, Q, cp, w);
        }
        if (cp) *capped = 1;
    } else {
        double acc = 0.0;
        for (int k = 0; k < nq; ++k) acc += gld[nq + k] * delta_at(gld[k]);
        out[i] = acc;
    }
}

// ------------------------------------------------------------------------
// host helpers
// ------------------------------------------------------------------------
// Gauss-Legendre nodes on [-1,1] by Newton iteration on P_n in long double,
// unfolded onto [0,1] with the half-length folded into the weights.
int gl_rule(int n, std::vector<double> &s, std::vector<double> &w)
{
    if (n < 1 || n > 1024) return -1;
    s.assign(n, 0.0);
    w.assign(n, 0.0);
    const int m = (n + 1) / 2;
    auto legendre = [n](long double z, long double &p0, long double &p1) {
        p0 = 1.0L;
        p1 = z;
        for (int j = 2; j <= n; ++j) {
            const long double p2 = ((2.0L * j - 1.0L) * z * p1 - (j - 1.0L) * p0) / j;
            p0 = p1;
            p1 = p2;
        }
        if (n == 1) { p0 = 1.0L; p1 = z; }
    };
    for (int k = 0; k < m; ++k) {
        long double z = cosl(3.14159265358979323846264338327950288L * (k + 0.75L) / (n + 0.5L));
        long double p0, p1;
        for (int it = 0; it < 100; ++it) {
            legendre(z, p0, p1);
            const long double dp = n * (z * p1 - p0) / (z * z - 1.0L);
            const long double dz = p1 / dp;
            z -= dz;
            if (fabsl(dz) < 1e-19L) break;
        }
        if ((n & 1) && k == m - 1) z = 0.0L;
        legendre(z, p0, p1);
        const long double dp = n * (z * p1 - p0) / (z * z - 1.0L);
        const long double wt = 2.0L / ((1.0L - z * z) * dp * dp);
        s[k] = (double)(0.5L - 0.5L * z);
        w[k] = (double)(0.5L * wt);
        s[n - 1 - k] = (double)(0.5L + 0.5L * z);
        w[n - 1 - k] = (double)(0.5L * wt);
    }
    return 0;
}

// Morton (Z-order) permutation of the grid points, so that a wavefront or a
// workgroup takes a compact block of grid points and shares the particles it
// reads.  Points are rounded onto a lattice whose spacing along each axis is
// extent / (n - 1), n = G^(1/d): a regular n^d grid whose end points sit on
// the box faces (as numpy.linspace makes them) maps onto the indices 0..n-1
// exactly, so for n a power of two every run of 4^k (2D) or 8^k (3D) points
// is a square (cube) block.
std::vector<int> morton_order(int64_t G, const double *grid, bool use_z)
{
    std::vector<int> perm((size_t)G);
    std::iota(perm.begin(), perm.end(), 0);
    if (G < 2) return perm;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < G; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], grid[3 * i + a]);
            hi[a] = std::max(hi[a], grid[3 * i + a]);
        }
    const int d = use_z ? 3 : 2;
    const double nside = std::max(2.0, std::round(std::pow((double)G, 1.0 / d)));
    double h[3];
    for (int a = 0; a < 3; ++a) h[a] = std::max(hi[a] - lo[a], 1e-300) / (nside - 1.0);
    const int bits = use_z ? 10 : 16;
    const double qmax = (double)((1u << bits) - 1);
    std::vector<uint64_t> key((size_t)G);
    for (int64_t i = 0; i < G; ++i) {
        uint64_t k = 0;
        uint32_t q[3] = {0, 0, 0};
        for (int a = 0; a < d; ++a)
            q[a] = (uint32_t)std::min(qmax, std::max(0.0, std::floor((grid[3 * i + a] - lo[a]) / h[a] + 0.5)));
        for (int b = bits - 1; b >= 0; --b)
            for (int a = d - 1; a >= 0; --a) k = (k << 1) | ((q[a] >> b) & 1u);
        key[(size_t)i] = k;
    }
    std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return key[(size_t)x] < key[(size_t)y]; });
    return perm;
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
        if (e != hipSuccess) {
            p = nullptr;
            return fail(PGT_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
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
    bool z_flat = true;       // all particle z are zero
    bool grid_z_flat = true;  // all grid z are zero
    double grid_absmax = 0.0;  // largest |grid coordinate|
    DBuf<double> d_pos, d_diam, d_mass, d_vel, d_disp;  // single-frame host uploads
    DBuf<double> d_grid;
    DBuf<int> d_perm;
    // work arrays of a batch (see WorkArray), as raw bytes
    DBuf<unsigned char> d_keys0, d_vals0, d_keys1, d_vals1, d_hist, d_cstart, d_bcount;
    DBuf<unsigned char> d_spos, d_sdiam, d_skin, d_sdisp;
    DBuf<unsigned char> d_bonds;
    DBuf<int> d_maxcount;  // [0] largest bond count, [1] adaptive workspace ran out
    DBuf<double> d_out;
    std::vector<double> h_grid;  // the current grid (set_grid skips an unchanged one)
    DBuf<double> d_gl;
    DBuf<double> d_qag;  // adaptive rule: interval tables beyond kQagMax intervals
    DBuf<unsigned long long> d_stats;              // job totals (balls, contributions)
    DBuf<unsigned char> d_wcounts, d_wmax, d_rec;  // per-wavefront counts and maxima; 3D FP32 records
    // surrogate field form: pair counts and offsets per grid point, pair
    // records, network inputs and outputs, final bond integrals, counter
    DBuf<int> d_pcount, d_poff;
    DBuf<PairRec> d_prec;
    DBuf<float> d_pxin, d_pymlp;
    DBuf<double> d_pw;
    DBuf<unsigned long long> d_pn;
    std::vector<double> gl_s, gl_w;
    int cap = 0;
    int block = 256;
    int frames_per_launch = 8;
    int points_per_lane = 2;  // grid points per gather lane (Octic 5-node rule; 1 elsewhere)
    int cell_reach = 1;       // cells searched on each side (cell edge >= search radius / reach)
    int gather_mode = 0;      // 1: item-balanced tile gather where the form has one, 0: one lane per point
    // set by pgt_ik_enqueue_frames: launches that a captured graph may replay
    // hold the current device buffers, so nothing may free or move them (or
    // change the bond capacity the launches were built with) until
    // pgt_ik_release_enqueued
    bool held = false;
    // the no-allocation entry point: this device's copy of the static scratch
    // array (address from hipGetSymbolAddress at create) and the bond
    // capacity that entry point uses
    unsigned char *scratch = nullptr;
    int scratch_cap = 0;
    // pinned host copy of a job's counters (flags, then the two totals), read
    // back with one synchronisation per job
    unsigned long long *h_counts = nullptr;
    pgt_ik_stats stats{};
};

// Scratch memory of pgt_ik_compute_frames_noalloc: every work array of its
// batches (cell list, bond list, statistics) is carved out of this array, so
// that entry point never allocates.  Each GPU the library runs on holds its
// own copy for the life of the process.  In host code `scratch` names a
// host-side stand-in, not device memory: the device address comes from
// hipGetSymbolAddress (scratch_on_device), made once per device with that
// device current, and only that address is passed to kernels.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<unsigned char *> g_scratch_addr;  // per device id

// Device address of the scratch array on device `dev`, which must be the
// current device.
int scratch_on_device(int dev, unsigned char **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0) return fail(PGT_ERR_ARG, "bad device id");
    if ((int)g_scratch_addr.size() <= dev) g_scratch_addr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_addr[(size_t)dev]) {
        void *ptr = nullptr;
        const hipError_t er = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (er != hipSuccess || !ptr)
            return fail(PGT_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(er));
        g_scratch_addr[(size_t)dev] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch_addr[(size_t)dev];
    return 0;
}

int check_engine(pgt_ik_engine *e)
{
    if (!e) return fail(PGT_ERR_ARG, "null engine");
    return 0;
}

int fail_held()
{
    return fail(PGT_ERR_STATE, "device buffers are held by enqueued work that a captured graph may replay: call "
                               "pgt_ik_release_enqueued before a call that would regrow or replace them, and capture "
                               "again afterwards");
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
    if (p->n_pot_params < need || p->n_pot_params > 6)
        return fail(PGT_ERR_ARG, "potential needs " + std::to_string(need) + " parameters");
    if (p->cg_kind < 0 || p->cg_kind > 2) return fail(PGT_ERR_ARG, "unknown coarse-graining kernel");
    if (!(p->cg_rcut > 0)) return fail(PGT_ERR_ARG, "cg_rcut must be positive");
    if (!(p->max_rcut >= 0)) return fail(PGT_ERR_ARG, "max_rcut must be non-negative");
    if (p->quad_mode == PGT_QUAD_FIXED) {
        if (p->quad_order < 1 || p->quad_order > 1024) return fail(PGT_ERR_ARG, "quadrature order must be in 1..1024");
    } else if (p->quad_mode == PGT_QUAD_ADAPTIVE) {
        if (p->adapt_limit < 1) return fail(PGT_ERR_ARG, "adaptive limit (maxiter) must be at least 1");
        if (!(p->adapt_epsabs >= 0) || !(p->adapt_epsrel >= 0))
            return fail(PGT_ERR_ARG, "adaptive tolerances must be non-negative");
        if (p->precision == PGT_PREC_FP32)
            return fail(PGT_ERR_UNSUPPORTED, "the adaptive rule runs in FP64 only");
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
    // cells of at least search radius / R, windows of R cells each side
    // (R = 2 scans 30% fewer candidates in 2D, 42% in 3D, for 4x / 8x the
    // cells; which is faster depends on the workload, so it is tuned)
    const int R = e->cell_reach;
    g.reach = R;
    g.ncell = 1;
    for (int a = 0; a < 3; ++a) {
        g.L[a] = e->L[a];
        g.invL[a] = 1.0 / e->L[a];
        g.lo[a] = e->lo[a];
        g.periodic[a] = e->periodic[a];
        g.pinvL[a] = e->periodic[a] ? g.invL[a] : 0.0;
        int nc = (int)std::floor(e->L[a] * R / rmax);
        if (!e->periodic[a] || nc < 2 * R + 1 || (a == 2 && D == 2)) nc = 1;
        g.nc[a] = nc;
        g.cs[a] = e->L[a] / nc;
        g.ncell *= nc;
    }
    g.rc2 = e->prm.cg_rcut * e->prm.cg_rcut;
    g.inv_rc = 1.0 / e->prm.cg_rcut;
    g.inv_rc2 = 1.0 / g.rc2;
    g.mr2 = e->prm.max_rcut * e->prm.max_rcut;
    // FP32 pre-test bounds for the two search radii (r_cg for the gather,
    // max_rcut for the bond list): fp32_band, with M the largest coordinate
    // magnitude of box and grid; every searched axis must be periodic, so
    // that particle coordinates are wrapped into the box
    for (int a = 0; a < 3; ++a) {
        g.L32[a] = (float)g.L[a];
        g.invL32[a] = (float)g.invL[a];
        g.pinvL32[a] = g.periodic[a] ? g.invL32[a] : 0.0f;
    }
    bool ok = std::isfinite(e->grid_absmax);
    double M = e->grid_absmax, Lm = 0.0, Lmin = INFINITY;
    for (int a = 0; a < D; ++a) {
        ok = ok && e->periodic[a];
        M = std::max(M, std::max(std::fabs(e->lo[a]), std::fabs(e->lo[a] + e->L[a])));
        Lm = std::max(Lm, e->L[a]);
        Lmin = std::min(Lmin, e->L[a]);
    }
    fp32_band(e->prm.cg_rcut, M, Lm, Lmin, D, ok, g.in32, g.out32);
    fp32_band(e->prm.max_rcut, M, Lm, Lmin, D, ok, g.bin32, g.bout32);
    return g;
}

// ------------------------------------------------------------------------
// launch dispatch
// ------------------------------------------------------------------------
// k_bonds threads per workgroup and LDS candidate-list length per lane (a
// full list is processed early: the length must stay well above the ~11.5
// (2D) or ~33 (3D) neighbours within max_rcut at unit density, or most
// particles process their list twice).  Measured per step: 2D, 256 lanes and
// 32 entries (32 KB of LDS) 1% faster than 128 lanes on the N = 65,536 job,
// 24 entries equal, 16 entries 17% slower; 3D, 64 lanes and 48 entries
// (12 KB) 9% faster on the N = 131,072 FP32 job than 128 lanes and 64
// entries, 40 entries 12% slower.
constexpr int bond_block(int D) { return D == 2 ? 256 : 64; }
constexpr int bond_list(int D) { return D == 2 ? 32 : 48; }

template <int D, int OD, typename ER>
void launch_bonds_t(const LaunchCtx &c, unsigned fields, const PotParams &P, const double *sdiam, const double *sdisp,
                    int *bcount, void *bonds, int *wave_max, float4 *rec)
{
    const bool V = fields & PGT_FIELD_VIRIAL, E = fields & PGT_FIELD_ELASTIC;
    constexpr int BLK = bond_block(D), MAXC = bond_list(D);
    dim3 grid((unsigned)((c.n + BLK - 1) / BLK), (unsigned)c.nf);
    if (V && E)
        hipLaunchKernelGGL((k_bonds<D, OD, true, true, ER, BLK, MAXC>), grid, dim3(BLK), 0, c.stream, c.g, c.n, P,
                           c.cgnorm, c.cstart, c.spos, sdiam, sdisp, c.cap, bcount, (ER *)bonds, wave_max, rec);
    else if (E)
        hipLaunchKernelGGL((k_bonds<D, OD, false, true, ER, BLK, MAXC>), grid, dim3(BLK), 0, c.stream, c.g, c.n, P,
                           c.cgnorm, c.cstart, c.spos, sdiam, sdisp, c.cap, bcount, (ER *)bonds, wave_max, rec);
    else
        hipLaunchKernelGGL((k_bonds<D, OD, true, false, ER, BLK, MAXC>), grid, dim3(BLK), 0, c.stream, c.g, c.n, P,
                           c.cgnorm, c.cstart, c.spos, sdiam, sdisp, c.cap, bcount, (ER *)bonds, wave_max, rec);
}

template <int D, int OD, typename ER>
size_t bond_entry_bytes(unsigned fields)
{
    if (fields & PGT_FIELD_ELASTIC) return sizeof(ER) * BondLayout<D, OD, true, ER>::ES;
    return sizeof(ER) * BondLayout<D, OD, false, ER>::ES;
}

// What a job's shape fixes about its batches: dimensions, fields, the cell
// geometry and radix-sort passes, the bond-entry size and the gather shape.
struct JobShape {
    int D, OD, KN, PS;
    bool pairs, kin, ela, fp32, adaptive;
    Geom g;
    int npass, nbits, radix;
    size_t ebytes;
    int block, pl;
    bool tile;       // the item-balanced tile gather (kTilePoints points per workgroup of kTileThreads lanes)
    int64_t gwaves;  // gather wavefronts per frame
    bool qag_global; // adaptive rule: interval tables of the caller's limit in e->d_qag
};

// n: particles per frame.  allow_tile false: the lane gather even where the
// tile gather is selected (the fallback when a tile reports an overflow)
JobShape job_shape(const pgt_ik_engine *e, unsigned fields, bool z_flat, int64_t n, bool allow_tile = true)
{
    JobShape j{};
    j.OD = e->prm.dim;
    j.D = (j.OD == 3 || !z_flat) ? 3 : 2;
    j.KN = j.OD * (j.OD + 1) / 2;
    j.PS = (j.D == 2) ? 2 : 4;
    j.pairs = fields & (PGT_FIELD_VIRIAL | PGT_FIELD_ELASTIC);
    j.kin = fields & PGT_FIELD_KINETIC;
    j.ela = fields & PGT_FIELD_ELASTIC;
    // FP32 form: Octic kernel on the 2D (z == 0) and 3D paths; a 2D system
    // with nonzero z runs in FP64.
    j.fp32 = e->prm.precision == PGT_PREC_FP32 && !(j.D == 3 && j.OD == 2);
    j.adaptive = e->prm.quad_mode == PGT_QUAD_ADAPTIVE && j.pairs;
    j.g = make_geom(e, j.D);
    // radix sort of the cell keys: digits of nbits bits, npass passes
    int kbits = 0;
    while ((1 << kbits) < j.g.ncell) ++kbits;
    j.npass = std::max(1, (kbits + 7) / 8);
    j.nbits = std::max(1, (kbits + j.npass - 1) / j.npass);
    j.radix = 1 << j.nbits;
    if (j.D == 2) j.ebytes = j.fp32 ? bond_entry_bytes<2, 2, float>(fields) : bond_entry_bytes<2, 2, double>(fields);
    else if (j.OD == 2) j.ebytes = bond_entry_bytes<3, 2, double>(fields);
    else j.ebytes = j.fp32 ? bond_entry_bytes<3, 3, float>(fields) : bond_entry_bytes<3, 3, double>(fields);
    j.block = j.adaptive ? 128 : e->block;  // k_gather_adaptive runs 128-thread workgroups
    j.pl = j.adaptive ? 1 : gather_pl(e->points_per_lane, e->prm.cg_kind, (int)e->gl_s.size(), j.D, j.OD);
    j.tile = allow_tile && !j.adaptive && e->gather_mode == 1 && n <= kTileMaxParticles &&
             tile_form(e->prm.cg_kind, (int)e->gl_s.size(), j.D, j.OD);
    if (j.tile) {
        j.block = kTileThreads;
        j.pl = 1;
        j.gwaves = (e->G + kTilePoints - 1) / kTilePoints * (kTileThreads / 64);
    } else {
        // the adaptive gather deals each point's integrals to several lanes
        const int64_t lanes = j.adaptive ? e->G * adaptive_lanes(e->G) : (e->G + j.pl - 1) / j.pl;
        j.gwaves = (lanes + j.block - 1) / j.block * (j.block / 64);
    }
    return j;
}

// Device work arrays of one batch of frames, and the job's counters.
enum WorkArray {
    W_KEYS0, W_VALS0, W_KEYS1, W_VALS1, W_HIST, W_CSTART, W_SPOS, W_SDIAM, W_BCOUNT, W_SKIN, W_SDISP, W_REC,
    W_BONDS, W_WCOUNTS, W_WMAX, W_NARRAYS
};

struct Work {
    void *p[W_NARRAYS];
    unsigned long long *stats;  // job totals: balls, contributions
    int *maxcount;              // [0] largest bond count, [1] an adaptive integral ran out of workspace
    template <typename T>
    T *get(WorkArray a) const { return static_cast<T *>(p[a]); }
};

// Bytes of each work array for nf frames of n particles at bond capacity cap
// (0 for arrays the job does not use).
void work_bytes(const JobShape &j, int nf, int64_t n, int64_t G, int cap, size_t out[W_NARRAYS])
{
    const size_t fn = (size_t)nf * (size_t)n;
    const size_t nchunk = (size_t)((n + kSortChunk - 1) / kSortChunk);
    const size_t nw_gather_max = (size_t)std::max<int64_t>(j.gwaves, (G + j.block - 1) / j.block * (j.block / 64)) * nf;
    const int bb = bond_block(j.D);
    const size_t nw_bonds = j.pairs ? (size_t)((n + bb - 1) / bb) * nf * (bb / 64) : 0;
    for (int k = 0; k < W_NARRAYS; ++k) out[k] = 0;
    out[W_KEYS0] = out[W_VALS0] = out[W_KEYS1] = out[W_VALS1] = fn * sizeof(int);
    out[W_HIST] = (size_t)nf * j.radix * std::max<size_t>(nchunk, 1) * sizeof(int);
    out[W_CSTART] = (size_t)nf * (j.g.ncell + 1) * sizeof(int);
    out[W_SPOS] = fn * j.PS * sizeof(double);
    out[W_SDIAM] = fn * sizeof(double);
    out[W_BCOUNT] = fn * sizeof(int);
    if (j.kin) out[W_SKIN] = fn * j.KN * sizeof(double);
    if (j.ela) out[W_SDISP] = fn * 3 * sizeof(double);
    if (j.pairs && j.D == 3) out[W_REC] = fn * sizeof(float4);
    if (j.pairs) out[W_BONDS] = fn * (size_t)cap * j.ebytes;
    // per-wavefront statistics, sized for one point per gather lane (the most
    // wavefronts any shape launches)
    out[W_WCOUNTS] = 2 * std::max<size_t>(nw_gather_max, 1) * sizeof(unsigned long long);
    out[W_WMAX] = std::max<size_t>(nw_bonds, 1) * sizeof(int);
}

constexpr size_t kWorkAlign = 256;

size_t align_up(size_t b) { return (b + kWorkAlign - 1) / kWorkAlign * kWorkAlign; }

// Work arrays from the engine's own buffers, grown as needed unless no_alloc
// (graph capture: the buffers must already be large enough).
int work_from_engine(pgt_ik_engine *e, const JobShape &j, int nf, int64_t n, bool no_alloc, Work &w)
{
    size_t b[W_NARRAYS];
    work_bytes(j, nf, n, e->G, e->cap, b);
    DBuf<unsigned char> *bufs[W_NARRAYS] = {&e->d_keys0, &e->d_vals0, &e->d_keys1, &e->d_vals1, &e->d_hist,
                                            &e->d_cstart, &e->d_spos, &e->d_sdiam, &e->d_bcount, &e->d_skin,
                                            &e->d_sdisp, &e->d_rec, &e->d_bonds, &e->d_wcounts, &e->d_wmax};
    for (int k = 0; k < W_NARRAYS; ++k) {
        w.p[k] = nullptr;
        if (!b[k]) continue;
        DBuf<unsigned char> &buf = *bufs[k];
        if (!(buf.p && buf.n >= b[k])) {
            if (no_alloc)
                return fail(PGT_ERR_STATE, "engine buffers are smaller than this job: run pgt_ik_compute_frames with "
                                           "the same frame count and fields first");
            if (e->held) return fail_held();
            if (int rc = buf.ensure(b[k])) return rc;
        }
        w.p[k] = buf.p;
    }
    w.stats = e->d_stats.p;
    w.maxcount = e->d_maxcount.p;
    return 0;
}

// Work arrays carved out of `base` (the no-allocation entry point's scratch
// array): the job counters first, then each array, all 256-byte aligned.
// Returns the bytes used.
size_t work_carve(const JobShape &j, int nf, int64_t n, int64_t G, int cap, unsigned char *base, Work *w)
{
    size_t b[W_NARRAYS];
    work_bytes(j, nf, n, G, cap, b);
    size_t off = 0;
    if (w) {
        w->stats = reinterpret_cast<unsigned long long *>(base);
        w->maxcount = reinterpret_cast<int *>(base + 2 * sizeof(unsigned long long));
    }
    off = align_up(2 * sizeof(unsigned long long) + 2 * sizeof(int));
    for (int k = 0; k < W_NARRAYS; ++k) {
        if (w) w->p[k] = b[k] ? base + off : nullptr;
        off += align_up(b[k]);
    }
    return off;
}

// Launch the cell list, bond list and gather for nf frames (no host sync).
int run_frames(pgt_ik_engine *e, const JobShape &j, const Work &w, int cap, int nf, int64_t n_particles,
               const double *pos, const double *diam, const double *mass, const double *vel, const double *disp,
               unsigned fields, double *ovir, double *okin, double *oela, hipEvent_t ev_begin, hipEvent_t ev_gather,
               hipEvent_t ev_end, bool bonds_only)
{
    const int D = j.D, OD = j.OD;
    const int n = (int)n_particles;
    const int G = (int)e->G;
    const Geom &g = j.g;
    const int nchunk = (n + kSortChunk - 1) / kSortChunk;
    float4 *rec = w.get<float4>(W_REC);
    double *skin = w.get<double>(W_SKIN);
    double *sdisp = w.get<double>(W_SDISP);
    int *hist = w.get<int>(W_HIST);
    int *cstart = w.get<int>(W_CSTART);
    double *spos = w.get<double>(W_SPOS);
    double *sdiam = w.get<double>(W_SDIAM);
    int *bcount = w.get<int>(W_BCOUNT);
    hipStream_t s = e->stream;
    if (ev_begin) HIPCHK(hipEventRecord(ev_begin, s));
    // cell list: per pass a histogram, a scan and a scatter of one digit of
    // the cell keys; frames are dealt to XCDs whole (frame_block)
    {
        const dim3 cgrid((unsigned)nchunk, (unsigned)frame_rows(nf));
        int *kb[2] = {w.get<int>(W_KEYS0), w.get<int>(W_KEYS1)}, *vb[2] = {w.get<int>(W_VALS0), w.get<int>(W_VALS1)};
        for (int p = 0; p < j.npass; ++p) {
            const int shift = p * j.nbits, radix = j.radix;
            int *kin = kb[p & 1], *vin = vb[p & 1], *kout = kb[(p + 1) & 1], *vout = vb[(p + 1) & 1];
            if (p == 0) {
                hipLaunchKernelGGL(k_sort_hist<true>, cgrid, dim3(kSortBlock), 0, s, g, n, nf, shift, radix, pos, kin,
                                   hist);
                hipLaunchKernelGGL(k_sort_scan, dim3(nf), dim3(1024), 0, s, radix * nchunk, hist);
                hipLaunchKernelGGL(k_sort_scatter<true>, cgrid, dim3(kSortBlock), 0, s, n, nf, shift, radix, kin, vin,
                                   hist, kout, vout);
            } else {
                hipLaunchKernelGGL(k_sort_hist<false>, cgrid, dim3(kSortBlock), 0, s, g, n, nf, shift, radix, pos,
                                   kin, hist);
                hipLaunchKernelGGL(k_sort_scan, dim3(nf), dim3(1024), 0, s, radix * nchunk, hist);
                hipLaunchKernelGGL(k_sort_scatter<false>, cgrid, dim3(kSortBlock), 0, s, n, nf, shift, radix, kin,
                                   vin, hist, kout, vout);
            }
        }
        const int *keys = kb[j.npass & 1], *vals = vb[j.npass & 1];
        const dim3 pgrid((unsigned)((n + 255) / 256), (unsigned)frame_rows(nf));
        if (D == 2)
            hipLaunchKernelGGL((k_sort_finish<2, 2>), pgrid, dim3(256), 0, s, g, n, nf, g.ncell, keys, vals, pos, diam,
                               mass, vel, disp, cstart, spos, sdiam, skin, sdisp, rec);
        else if (OD == 2)
            hipLaunchKernelGGL((k_sort_finish<3, 2>), pgrid, dim3(256), 0, s, g, n, nf, g.ncell, keys, vals, pos, diam,
                               mass, vel, disp, cstart, spos, sdiam, skin, sdisp, rec);
        else
            hipLaunchKernelGGL((k_sort_finish<3, 3>), pgrid, dim3(256), 0, s, g, n, nf, g.ncell, keys, vals, pos, diam,
                               mass, vel, disp, cstart, spos, sdiam, skin, sdisp, rec);
    }
    HIPCHK(hipGetLastError());

    LaunchCtx c{};
    c.g = g;
    c.n = n;
    c.G = G;
    c.nf = nf;
    c.block = j.block;
    c.grid = e->d_grid.p;
    c.perm = e->d_perm.p;
    c.R.n = (int)e->gl_s.size();
    for (int k = 0; k < c.R.n && k < kMaxQArg; ++k) {
        c.R.s[k] = e->gl_s[k];
        c.R.w[k] = e->gl_w[k];
    }
    c.gld = e->d_gl.p;
    c.cgnorm = cg_norm(e->prm.cg_kind, e->prm.cg_rcut);
    c.cstart = cstart;
    c.spos = spos;
    c.skin = skin;
    c.out.vir = ovir;
    c.out.kin = okin;
    c.out.ela = oela;
    c.flags = w.maxcount + 1;
    c.tile = j.tile ? 1 : 0;
    c.stream = s;

    // per-wavefront statistics of this batch (see k_reduce_stats)
    c.pl = j.pl;
    const int nw_gather = (int)(j.gwaves * nf);
    const int bb = bond_block(D);
    const int nw_bonds = j.pairs ? (int)(((size_t)n + bb - 1) / bb) * nf * (bb / 64) : 0;
    c.out.wave_counts = w.get<unsigned long long>(W_WCOUNTS);
    int *wmax = w.get<int>(W_WMAX);
    void *bonds = w.p[W_BONDS];
    if (j.pairs) {
        // entries beyond the capacity are counted, not written; the caller
        // checks maxcount after the job and re-runs with a larger capacity
        c.cap = cap;
        const PotParams P = make_pot(e->prm);
        if (D == 2) {
            if (j.fp32) launch_bonds_t<2, 2, float>(c, fields, P, sdiam, sdisp, bcount, bonds, wmax, rec);
            else launch_bonds_t<2, 2, double>(c, fields, P, sdiam, sdisp, bcount, bonds, wmax, rec);
        } else if (OD == 2) {
            launch_bonds_t<3, 2, double>(c, fields, P, sdiam, sdisp, bcount, bonds, wmax, rec);
        } else {
            if (j.fp32) launch_bonds_t<3, 3, float>(c, fields, P, sdiam, sdisp, bcount, bonds, wmax, rec);
            else launch_bonds_t<3, 3, double>(c, fields, P, sdiam, sdisp, bcount, bonds, wmax, rec);
        }
        HIPCHK(hipGetLastError());
    } else {
        c.cap = 1;
    }
    c.bcount = bcount;
    c.rec = rec;
    c.bonds = bonds;
    c.pack = n < (1 << (c.pl > 1 ? 22 : 26)) && c.cap <= 63;
    if (ev_gather) HIPCHK(hipEventRecord(ev_gather, s));
    const int cg = e->prm.cg_kind;
    if (bonds_only) {
        // cell list and bond list only (the surrogate field form reads them)
    } else if (j.adaptive) {
        const AdaptQ Q{e->prm.adapt_limit, e->prm.adapt_epsabs, e->prm.adapt_epsrel};
        if (j.qag_global) {
            // sized by run_job for this batch's frames
            if (e->d_qag.n < adaptive_workspace_doubles(c, Q.limit))
                return fail(PGT_ERR_STATE, "adaptive workspace smaller than the batch needs");
            launch_gather_adaptive(c, fields, cg, D, OD, Q, c.flags, e->d_qag.p, Q.limit);
        } else {
            launch_gather_adaptive(c, fields, cg, D, OD, Q, c.flags, nullptr, 0);
        }
    } else if (D == 2) {
        if (j.fp32) launch_gather_2d32(c, fields, cg);
        else launch_gather_2d64(c, fields, cg);
    } else if (OD == 2) {
        launch_gather_3d2(c, fields, cg);
    } else {
        if (j.fp32) launch_gather_3d32(c, fields, cg);
        else launch_gather_3d64(c, fields, cg);
    }
    HIPCHK(hipGetLastError());
    if (!bonds_only && !j.adaptive && c.pl_launched != c.pl)
        return fail(PGT_ERR_STATE, "gather launched with an unexpected number of points per lane");
    if (ev_end) HIPCHK(hipEventRecord(ev_end, s));
    const int nwg = bonds_only ? 0 : nw_gather;
    const int nrb = std::max(1, std::min(256, (std::max(nwg, nw_bonds) + 4095) / 4096));
    hipLaunchKernelGGL(k_reduce_stats, dim3(nrb), dim3(1024), 0, s, nwg, c.out.wave_counts, nw_bonds, wmax, w.stats,
                       w.maxcount);
    HIPCHK(hipGetLastError());
    return 0;
}

struct FrameSet {
    int nframes;
    int64_t n;
    const double *pos, *diam, *mass, *vel, *disp;
    double *vir, *kin, *ela;  // device outputs, nframes * G * D * D each
    bool z_flat;
    bool bonds_only;  // build the cell and bond lists only (no field gather)
};

// Launch batch f0 .. f0 + nf - 1 of a job with work arrays w.
int run_batch(pgt_ik_engine *e, const JobShape &j, const Work &w, int cap, const FrameSet &fs, unsigned fields, int f0,
              int nf, hipEvent_t e0, hipEvent_t e1, hipEvent_t e2)
{
    const size_t fsz = (size_t)e->G * j.OD * j.OD;
    const size_t po = (size_t)f0 * fs.n;
    return run_frames(e, j, w, cap, nf, fs.n, fs.pos + po * 3, fs.diam + po, fs.mass ? fs.mass + po : nullptr,
                      fs.vel ? fs.vel + po * 3 : nullptr, fs.disp ? fs.disp + po * 3 : nullptr, fields,
                      fs.vir ? fs.vir + (size_t)f0 * fsz : nullptr, fs.kin ? fs.kin + (size_t)f0 * fsz : nullptr,
                      fs.ela ? fs.ela + (size_t)f0 * fsz : nullptr, e0, e1, e2, fs.bonds_only);
}

// Run every batch of a job on the engine's stream, then check the bond-list
// capacity once; on overflow grow it and run the whole job again.
int run_job(pgt_ik_engine *e, const FrameSet &fs, unsigned fields)
{
    int fpl = std::max(1, e->frames_per_launch);
    int nb = (fs.nframes + fpl - 1) / fpl;
    JobShape j = job_shape(e, fields, fs.z_flat, fs.n);
    int rc;
    if ((rc = ensure_events(e, (size_t)nb * 3))) return rc;
    if ((rc = e->d_maxcount.ensure(2)) || (rc = e->d_stats.ensure(2))) return rc;
    if (j.pairs && e->cap < 1) e->cap = j.D == 3 ? 32 : 16;
    int mx = 0;
    for (int attempt = 0; attempt < 4; ++attempt) {
        Work w{};
        if ((rc = work_from_engine(e, j, std::min(fpl, fs.nframes), fs.n, false, w))) return rc;
        if ((rc = fill_async(w.maxcount, 0, 2, e->stream)) || (rc = fill_async(w.stats, 0ull, 2, e->stream))) return rc;
        for (int b = 0; b < nb; ++b) {
            const int f0 = b * fpl;
            rc = run_batch(e, j, w, e->cap, fs, fields, f0, std::min(fpl, fs.nframes - f0), e->ev[3 * b],
                           e->ev[3 * b + 1], e->ev[3 * b + 2]);
            if (rc) return rc;
        }
        // flags and totals into pinned memory, one synchronisation
        int *hf = reinterpret_cast<int *>(e->h_counts);
        HIPCHK(hipMemcpyAsync(hf, e->d_maxcount.p, 2 * sizeof(int), hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipMemcpyAsync(e->h_counts + 1, e->d_stats.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        const int flags[2] = {hf[0], hf[1]};
        mx = flags[0];
        const bool grow = j.pairs && mx > e->cap;
        if (grow && mx > 65535) return fail(PGT_ERR_RANGE, "a particle has more than 65535 interacting neighbours");
        if (grow && e->held) return fail_held();  // the graph's launches keep the old capacity
        if (flags[1] & 1) {
            // an adaptive integral needed more than the kQagMax intervals of
            // the private tables while the caller's limit allows more: run
            // the job again with tables of the limit in global memory, as
            // many frames per launch as kQagWorkspaceBytes holds
            if (j.qag_global) return fail(PGT_ERR_STATE, "adaptive workspace overflow with the caller's limit");
            const size_t per_frame =
                (((size_t)e->G * adaptive_lanes(e->G) + 127) / 128) * 128 * 4 * (size_t)e->prm.adapt_limit;
            const size_t fit = kQagWorkspaceBytes / sizeof(double) / per_frame;
            if (fit < 1)
                return fail(PGT_ERR_RANGE, "an adaptive bond integral needs more than " + std::to_string(kQagMax) +
                                               " intervals, and the tables of the caller's limit for one frame's "
                                               "grid exceed the " + std::to_string(kQagWorkspaceBytes >> 20) +
                                               " MiB workspace");
            fpl = (int)std::min<size_t>((size_t)fpl, fit);
            nb = (fs.nframes + fpl - 1) / fpl;
            if ((rc = ensure_events(e, (size_t)nb * 3))) return rc;
            if ((rc = e->d_qag.ensure(per_frame * (size_t)std::min(fpl, fs.nframes)))) return rc;
            j.qag_global = true;
            if (grow) e->cap = ((mx + 7) / 8) * 8;
            continue;
        }
        if (flags[1] & 6) {
            // a tile item had more bonds than its one-byte count holds:
            // run the job again with the lane gather
            if (!j.tile) return fail(PGT_ERR_STATE, "tile overflow reported by a lane gather");
            j = job_shape(e, fields, fs.z_flat, fs.n, false);
        } else if (!grow) {
            break;
        }
        if (grow) e->cap = ((mx + 7) / 8) * 8;
        if (attempt == 3) return fail(PGT_ERR_HIP, "bond list capacity did not settle");
    }
    const unsigned long long h[2] = {e->h_counts[1], e->h_counts[2]};
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

// Enqueue every batch of a job on the engine's stream and return: no host
// synchronisation, no allocation, no events, no memset, so the launches can
// be captured into a HIP graph.  Needs buffers and a bond capacity sized by an
// earlier run_job of the same shape; an overflow of that capacity during the
// enqueued work shows up in pgt_ik_finish.
int enqueue_job(pgt_ik_engine *e, const FrameSet &fs, unsigned fields)
{
    const int fpl = std::max(1, e->frames_per_launch);
    const int nb = (fs.nframes + fpl - 1) / fpl;
    if (e->cap < 1 || !e->d_maxcount.p || e->d_maxcount.n < 2 || !e->d_stats.p)
        return fail(PGT_ERR_STATE, "run pgt_ik_compute_frames of the same shape before enqueueing");
    const JobShape j = job_shape(e, fields, fs.z_flat, fs.n);
    Work w{};
    if (int rc = work_from_engine(e, j, std::min(fpl, fs.nframes), fs.n, true, w)) return rc;
    for (int b = 0; b < nb; ++b) {
        const int f0 = b * fpl;
        const int rc = run_batch(e, j, w, e->cap, fs, fields, f0, std::min(fpl, fs.nframes - f0), nullptr, nullptr,
                                 nullptr);
        if (rc) return rc;
    }
    return 0;
}

}  // namespace

// ==========================================================================
// C ABI
// ==========================================================================
namespace {

GLRule engine_rule(const pgt_ik_engine *e)
{
    GLRule R{};
    R.n = (int)e->gl_s.size();
    for (int k = 0; k < R.n && k < kMaxQArg; ++k) {
        R.s[k] = e->gl_s[k];
        R.w[k] = e->gl_w[k];
    }
    return R;
}

// The surrogate stands in for the Octic 5-node rule of a 2D (z = 0) system
// with FP64 bond entries.
int surrogate_check(pgt_ik_engine *e)
{
    if (int rc = check_engine(e)) return rc;
    if (!e->box_set || !e->have_particles || !e->have_grid)
        return fail(PGT_ERR_STATE, "set_box, set_particles and set_grid must precede the surrogate form");
    if (e->prm.dim != 2 || !(e->z_flat && e->grid_z_flat))
        return fail(PGT_ERR_UNSUPPORTED, "the bond-integral surrogate is built for 2D systems (z = 0)");
    if (e->prm.cg_kind != PGT_CG_OCTIC || e->prm.quad_mode != PGT_QUAD_FIXED || e->gl_s.size() != 5 ||
        e->prm.precision != PGT_PREC_FP64)
        return fail(PGT_ERR_UNSUPPORTED,
                    "the bond-integral surrogate stands in for the Octic kernel's 5-node rule with FP64 bond entries");
    return 0;
}

// Cell and bond lists of the engine's frame, then the (grid point, bond)
// pairs: counts, offsets and *np.  With write, the pair records and network
// input rows (into xin, or the engine's buffer when xin is null).
int prepare_pairs(pgt_ik_engine *e, int64_t *np, bool write, float *xin)
{
    *np = 0;
    if (e->n == 0 || e->G == 0) return 0;
    FrameSet fs{};
    fs.nframes = 1;
    fs.n = e->n;
    fs.pos = e->d_pos.p;
    fs.diam = e->d_diam.p;
    fs.z_flat = true;
    fs.bonds_only = true;
    int rc;
    if ((rc = run_job(e, fs, PGT_FIELD_VIRIAL))) return rc;
    const int G = (int)e->G;
    const Geom g = make_geom(e, 2);
    hipStream_t s = e->stream;
    // d_pn: [0] pairs taken from the network, [1] the 64-bit pair total
    if ((rc = e->d_pcount.ensure(G)) || (rc = e->d_poff.ensure(G)) || (rc = e->d_pn.ensure(2))) return rc;
    if ((rc = fill_async(e->d_pn.p + 1, 0ull, 1, s))) return rc;
    const dim3 gg((unsigned)((G + 127) / 128));
    const int *cstart = reinterpret_cast<const int *>(e->d_cstart.p);
    const double *spos = reinterpret_cast<const double *>(e->d_spos.p);
    const int *bcount = reinterpret_cast<const int *>(e->d_bcount.p);
    hipLaunchKernelGGL(k_pair_count, gg, dim3(128), 0, s, g, G, e->d_grid.p, cstart, spos, e->cap, bcount,
                       e->d_pcount.p, e->d_pn.p + 1);
    unsigned long long total = 0;
    HIPCHK(hipMemcpyAsync(&total, e->d_pn.p + 1, sizeof(total), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (total > (unsigned long long)INT_MAX)
        return fail(PGT_ERR_RANGE, "the frame has more than 2^31 - 1 (grid point, bond) pairs; split the grid");
    *np = (int64_t)total;
    if (!write || *np == 0) return 0;
    HIPCHK(hipMemcpyAsync(e->d_poff.p, e->d_pcount.p, (size_t)G * sizeof(int), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(1024), 0, s, G, e->d_poff.p);
    if ((rc = e->d_prec.ensure((size_t)*np)) || (rc = e->d_pw.ensure((size_t)*np))) return rc;
    if (!xin) {
        if ((rc = e->d_pxin.ensure((size_t)*np * kMlpInPad))) return rc;
        xin = e->d_pxin.p;
    }
    hipLaunchKernelGGL(k_pair_write, gg, dim3(128), 0, s, g, G, e->d_grid.p, cstart, spos, e->cap, bcount,
                       (const double *)e->d_bonds.p, e->d_poff.p, e->prm.cg_rcut, e->d_prec.p, xin);
    HIPCHK(hipGetLastError());
    return 0;
}

// H1 H2 H3 dZa dZb (64 n each), Y, dY (n each), loss partials (doubles),
// gradient partials (kGradSplit rows of 64 * 64 + 64)
size_t mlp_part_offset(int64_t n) { return (size_t)n * 322 + 2 * (size_t)((n + 255) / 256) + 2; }
size_t mlp_work_floats(int64_t n) { return mlp_part_offset(n) + (size_t)kGradSplit * (kMlpW * kMlpW + kMlpW); }

unsigned blocks_for(int64_t threads) { return (unsigned)((threads + 255) / 256); }

}  // namespace

extern "C" {

const char *pgt_last_error(void) { return g_err.c_str(); }

int pgt_pretest_band(double r, double max_coord, double max_box, double min_box, int dim, int periodic,
                     float *in32, float *out32)
{
    if (!in32 || !out32) return fail(PGT_ERR_ARG, "null output");
    if (dim != 2 && dim != 3) return fail(PGT_ERR_ARG, "dim must be 2 or 3");
    fp32_band(r, max_coord, max_box, min_box, dim, periodic != 0, *in32, *out32);
    return PGT_OK;
}

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
    if ((rc = scratch_on_device(e->device, &e->scratch))) {
        delete e;
        return rc;
    }
    hipError_t he = hipStreamCreateWithFlags(&e->own_stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete e;
        return fail(PGT_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
    }
    e->stream = e->own_stream;
    he = hipHostMalloc((void **)&e->h_counts, 4 * sizeof(unsigned long long), hipHostMallocDefault);
    if (he != hipSuccess) {
        e->h_counts = nullptr;
        pgt_ik_destroy(e);
        return fail(PGT_ERR_HIP, std::string("hipHostMalloc: ") + hipGetErrorString(he));
    }
    if (gl_rule(params->quad_mode == PGT_QUAD_FIXED ? params->quad_order : 5, e->gl_s, e->gl_w) != 0) {
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
                             &e->d_out, &e->d_gl};
    for (auto *b : dbufs) b->release();
    DBuf<int> *ibufs[] = {&e->d_perm, &e->d_maxcount};
    for (auto *b : ibufs) b->release();
    DBuf<unsigned char> *wbufs[] = {&e->d_keys0, &e->d_vals0, &e->d_keys1, &e->d_vals1, &e->d_hist, &e->d_cstart,
                                    &e->d_bcount, &e->d_spos, &e->d_sdiam, &e->d_skin, &e->d_sdisp, &e->d_bonds,
                                    &e->d_wcounts, &e->d_wmax, &e->d_rec};
    for (auto *b : wbufs) b->release();
    e->d_pcount.release();
    e->d_poff.release();
    e->d_prec.release();
    e->d_pxin.release();
    e->d_pymlp.release();
    e->d_pw.release();
    e->d_pn.release();
    e->d_stats.release();
    for (hipEvent_t v : e->ev) (void)hipEventDestroy(v);
    if (e->h_counts) (void)hipHostFree(e->h_counts);
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
    }
    for (int a = 0; a < 3; ++a) {
        e->L[a] = boxsize[a];
        e->lo[a] = lowerbound ? lowerbound[a] : -0.5 * boxsize[a];
        e->periodic[a] = periodic[a] != 0 ? 1 : 0;
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
    if (mass && n > 0 && (rc = e->d_mass.ensure((size_t)n))) return rc;
    if (vel && n > 0 && (rc = e->d_vel.ensure((size_t)n * 3))) return rc;
    if (disp && n > 0 && (rc = e->d_disp.ensure((size_t)n * 3))) return rc;
    hipStream_t s = e->stream;
    e->have_mass = mass != nullptr;
    e->have_vel = vel != nullptr;
    e->have_disp = disp != nullptr;
    if (n > 0) {
        // straight from the caller's (pageable) arrays: the runtime stages
        // them through its own pinned buffers with the copies pipelined,
        // which measured faster than one host copy into pinned memory of
        // ours followed by the DMA
        const double *src[5] = {pos, diam, mass, vel, disp};
        double *dst[5] = {e->d_pos.p, e->d_diam.p, e->d_mass.p, e->d_vel.p, e->d_disp.p};
        const size_t cnt[5] = {3, 1, 1, 3, 3};
        for (int k = 0; k < 5; ++k)
            if (src[k])
                HIPCHK(hipMemcpyAsync(dst[k], src[k], sizeof(double) * cnt[k] * (size_t)n, hipMemcpyHostToDevice, s));
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
    // the grid already set (a driver passes the same grid on every frame):
    // no Morton sort, upload or synchronisation
    if (e->have_grid && g == e->G && (g == 0 || std::memcmp(e->h_grid.data(), grid, sizeof(double) * 3 * g) == 0))
        return PGT_OK;
    if (e->held) return fail_held();  // enqueued launches read the grid with its size built in
    bool flat = true;
    double amax = 0.0;
    for (int64_t i = 0; i < g; ++i) {
        flat = flat && grid[3 * i + 2] == 0.0;
        for (int a = 0; a < 3; ++a) amax = std::max(amax, std::fabs(grid[3 * i + a]));
    }
    if (!(amax < INFINITY)) amax = INFINITY;  // NaN coordinates: no FP32 pre-test
    const std::vector<int> perm = morton_order(g, grid, !flat || e->prm.dim == 3);
    int rc;
    if ((rc = e->d_grid.ensure((size_t)g * 3)) || (rc = e->d_perm.ensure((size_t)g))) return rc;
    if (g > 0) {
        HIPCHK(hipMemcpyAsync(e->d_grid.p, grid, sizeof(double) * g * 3, hipMemcpyHostToDevice, e->stream));
        HIPCHK(hipMemcpyAsync(e->d_perm.p, perm.data(), sizeof(int) * g, hipMemcpyHostToDevice, e->stream));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    e->h_grid.assign(grid, grid + 3 * g);
    e->grid_z_flat = flat;
    e->grid_absmax = amax;
    e->G = g;
    e->have_grid = true;
    return PGT_OK;
}

int pgt_ik_enqueue_frames(pgt_ik_engine *e, int nframes, int64_t n, const double *d_pos, const double *d_diam,
                          const double *d_mass, const double *d_vel, const double *d_disp, unsigned fields,
                          double *d_out_virial, double *d_out_kinetic, double *d_out_elastic)
{
    if (int rc = check_engine(e)) return rc;
    if (!e->box_set || !e->have_grid) return fail(PGT_ERR_STATE, "set_box and set_grid must precede enqueue_frames");
    if (nframes < 0 || n < 0 || n > (int64_t)1 << 30) return fail(PGT_ERR_ARG, "bad frame or particle count");
    if (fields == 0 || (fields & ~7u)) return fail(PGT_ERR_ARG, "bad field mask");
    if (((fields & PGT_FIELD_VIRIAL) && !d_out_virial) || ((fields & PGT_FIELD_KINETIC) && !d_out_kinetic) ||
        ((fields & PGT_FIELD_ELASTIC) && !d_out_elastic))
        return fail(PGT_ERR_ARG, "null output for a selected field");
    if (n > 0 && (!d_pos || !d_diam)) return fail(PGT_ERR_ARG, "positions and diameters are required");
    if (nframes == 0 || e->G == 0) return PGT_OK;
    const int OD = e->prm.dim;
    const size_t fsz = (size_t)e->G * OD * OD;
    if (n == 0) {
        const unsigned bits[3] = {PGT_FIELD_VIRIAL, PGT_FIELD_KINETIC, PGT_FIELD_ELASTIC};
        double *outs[3] = {d_out_virial, d_out_kinetic, d_out_elastic};
        for (int k = 0; k < 3; ++k)
            if (fields & bits[k])
                if (int rc = fill_async(outs[k], 0.0, fsz * nframes, e->stream)) return rc;
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
    fs.z_flat = e->grid_z_flat;
    const int rc = enqueue_job(e, fs, fields);
    if (rc == PGT_OK) e->held = true;
    return rc;
}

int pgt_ik_release_enqueued(pgt_ik_engine *e)
{
    if (int rc = check_engine(e)) return rc;
    e->held = false;
    return PGT_OK;
}

int pgt_ik_finish(pgt_ik_engine *e)
{
    if (int rc = check_engine(e)) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    if (!e->d_maxcount.p || !e->d_stats.p) return PGT_OK;
    int flags[2] = {0, 0};
    unsigned long long h[2] = {0, 0};
    HIPCHK(hipMemcpy(flags, e->d_maxcount.p, sizeof(flags), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(h, e->d_stats.p, sizeof(h), hipMemcpyDeviceToHost));
    int rc = fill_async(e->d_maxcount.p, 0, 2, e->stream);
    if (!rc) rc = fill_async(e->d_stats.p, 0ull, 2, e->stream);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    e->stats.n_ball = (int64_t)h[0];
    e->stats.n_contrib = (int64_t)h[1];
    e->stats.max_bonds = flags[0];
    e->stats.bond_capacity = e->cap;
    if (flags[1] & 1)
        return fail(PGT_ERR_RANGE, "an adaptive bond integral needed more than " + std::to_string(kQagMax) +
                                       " intervals: the enqueued results are incomplete; pgt_ik_compute_frames "
                                       "keeps tables of the caller's limit");
    if (flags[1] & 6)
        return fail(PGT_ERR_RANGE, "a particle had more than 255 bonds, more than the tile gather holds: "
                                   "the enqueued results are incomplete; select the lane gather "
                                   "(pgt_ik_set_gather_mode) and capture again");
    if (flags[0] > e->cap)
        return fail(PGT_ERR_RANGE, "a particle had more bonds than the capacity sized by the last "
                                   "pgt_ik_compute_frames: the enqueued results are incomplete; run "
                                   "pgt_ik_compute_frames on these frames to grow it");
    return PGT_OK;
}

int pgt_ik_set_stream(pgt_ik_engine *e, void *hip_stream)
{
    if (int rc = check_engine(e)) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    e->stream = hip_stream ? (hipStream_t)hip_stream : e->own_stream;
    return PGT_OK;
}

int pgt_ik_set_tuning(pgt_ik_engine *e, int threads_per_block, int frames_per_launch, int points_per_lane,
                      int cell_reach)
{
    if (int rc = check_engine(e)) return rc;
    if (threads_per_block) {
        if (threads_per_block != 128 && threads_per_block != 256)
            return fail(PGT_ERR_ARG, "threads per block must be 128 or 256");
        e->block = threads_per_block;
    }
    if (frames_per_launch) {
        if (frames_per_launch < 1 || frames_per_launch > 65535) return fail(PGT_ERR_ARG, "bad frames per launch");
        e->frames_per_launch = frames_per_launch;
    }
    if (points_per_lane) {
        if (points_per_lane != 1 && points_per_lane != 2 && points_per_lane != 4)
            return fail(PGT_ERR_ARG, "grid points per lane must be 1, 2 or 4");
        e->points_per_lane = points_per_lane;
    }
    if (cell_reach) {
        if (cell_reach != 1 && cell_reach != 2) return fail(PGT_ERR_ARG, "cell reach must be 1 or 2");
        e->cell_reach = cell_reach;
    }
    return PGT_OK;
}

int pgt_ik_set_gather_mode(pgt_ik_engine *e, int mode)
{
    if (int rc = check_engine(e)) return rc;
    if (mode != 0 && mode != 1) return fail(PGT_ERR_ARG, "gather mode must be 0 (lane) or 1 (tile)");
    e->gather_mode = mode;
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
        if (fields & PGT_FIELD_VIRIAL) std::memset(virial, 0, fsz * sizeof(double));
        if (fields & PGT_FIELD_KINETIC) std::memset(kinetic, 0, fsz * sizeof(double));
        if (fields & PGT_FIELD_ELASTIC) std::memset(elastic, 0, fsz * sizeof(double));
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
    // straight into the caller's arrays (see pgt_ik_set_particles)
    const unsigned bits[3] = {PGT_FIELD_VIRIAL, PGT_FIELD_KINETIC, PGT_FIELD_ELASTIC};
    const double *dsrc[3] = {fs.vir, fs.kin, fs.ela};
    double *hdst[3] = {virial, kinetic, elastic};
    for (int k = 0; k < 3; ++k)
        if (fields & bits[k])
            HIPCHK(hipMemcpyAsync(hdst[k], dsrc[k], fsz * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
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
        const unsigned bits[3] = {PGT_FIELD_VIRIAL, PGT_FIELD_KINETIC, PGT_FIELD_ELASTIC};
        double *outs[3] = {d_out_virial, d_out_kinetic, d_out_elastic};
        for (int k = 0; k < 3; ++k)
            if (fields & bits[k]) HIPCHK(hipMemsetAsync(outs[k], 0, fsz * nframes * sizeof(double), e->stream));
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

int pgt_ik_compute_frames_noalloc(pgt_ik_engine *e, int nframes, int64_t n, const double *d_pos,
                                  const double *d_diam, const double *d_mass, const double *d_vel,
                                  const double *d_disp, unsigned fields, double *d_out_virial,
                                  double *d_out_kinetic, double *d_out_elastic)
{
    if (int rc = check_engine(e)) return rc;
    if (!e->box_set || !e->have_grid)
        return fail(PGT_ERR_STATE, "set_box and set_grid must precede compute_frames_noalloc");
    if (nframes < 0 || n < 0 || n > (int64_t)1 << 30) return fail(PGT_ERR_ARG, "bad frame or particle count");
    if (fields == 0 || (fields & ~7u)) return fail(PGT_ERR_ARG, "bad field mask");
    if (((fields & PGT_FIELD_VIRIAL) && !d_out_virial) || ((fields & PGT_FIELD_KINETIC) && !d_out_kinetic) ||
        ((fields & PGT_FIELD_ELASTIC) && !d_out_elastic))
        return fail(PGT_ERR_ARG, "null output for a selected field");
    if (n > 0 && (!d_pos || !d_diam)) return fail(PGT_ERR_ARG, "positions and diameters are required");
    if (!e->scratch) return fail(PGT_ERR_STATE, "no scratch array on the engine's device");
    e->stats = pgt_ik_stats{};
    if (nframes == 0 || e->G == 0) return PGT_OK;
    const int OD = e->prm.dim;
    const size_t fsz = (size_t)e->G * OD * OD;
    hipStream_t s = e->stream;
    if (n == 0) {  // fill kernels, no memset
        const unsigned bits[3] = {PGT_FIELD_VIRIAL, PGT_FIELD_KINETIC, PGT_FIELD_ELASTIC};
        double *outs[3] = {d_out_virial, d_out_kinetic, d_out_elastic};
        for (int k = 0; k < 3; ++k)
            if (fields & bits[k])
                if (int rc = fill_async(outs[k], 0.0, fsz * nframes, s)) return rc;
        HIPCHK(hipStreamSynchronize(s));
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
    fs.z_flat = e->grid_z_flat;
    JobShape j = job_shape(e, fields, fs.z_flat, fs.n);
    int cap = e->scratch_cap;
    if (j.pairs && cap < 1) cap = j.D == 3 ? 32 : 16;
    if (!j.pairs) cap = 1;
    int mx = 0;
    unsigned long long h[2] = {0, 0};
    for (int attempt = 0; attempt < 3; ++attempt) {
        // the most frames per launch whose work arrays fit the scratch array
        int nf = std::min(std::max(1, e->frames_per_launch), nframes);
        while (nf > 0 && work_carve(j, nf, n, e->G, cap, nullptr, nullptr) > kScratchBytes) nf = nf / 2;
        if (nf < 1)
            return fail(PGT_ERR_RANGE, "one frame of this job needs " +
                                           std::to_string(work_carve(j, 1, n, e->G, cap, nullptr, nullptr)) +
                                           " bytes of work memory; the no-allocation entry point has " +
                                           std::to_string(kScratchBytes) + " (use pgt_ik_compute_frames)");
        Work w{};
        work_carve(j, nf, n, e->G, cap, e->scratch, &w);
        // counters start at zero on every call (the array keeps the last call's contents)
        int rc;
        if ((rc = fill_async(w.maxcount, 0, 2, s)) || (rc = fill_async(w.stats, 0ull, 2, s))) return rc;
        for (int f0 = 0; f0 < nframes; f0 += nf)
            if ((rc = run_batch(e, j, w, cap, fs, fields, f0, std::min(nf, nframes - f0), nullptr, nullptr, nullptr)))
                return rc;
        int flags[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(flags, w.maxcount, sizeof(flags), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(h, w.stats, sizeof(h), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        mx = flags[0];
        if (flags[1] & 1)
            return fail(PGT_ERR_RANGE, "an adaptive bond integral needed more than " + std::to_string(kQagMax) +
                                           " intervals, which this entry point has no memory for; "
                                           "pgt_ik_compute_frames keeps tables of the caller's limit");
        const bool grow = j.pairs && mx > cap;
        if (grow && mx > 65535) return fail(PGT_ERR_RANGE, "a particle has more than 65535 interacting neighbours");
        if (flags[1] & 6) {
            if (!j.tile) return fail(PGT_ERR_STATE, "tile overflow reported by a lane gather");
            j = job_shape(e, fields, fs.z_flat, fs.n, false);
        } else if (!grow) {
            break;
        }
        if (grow) cap = ((mx + 7) / 8) * 8;
        if (attempt == 2) return fail(PGT_ERR_HIP, "bond list capacity did not settle");
    }
    if (j.pairs) e->scratch_cap = cap;
    e->stats.n_particles = n;
    e->stats.n_grid = e->G;
    e->stats.n_ball = (int64_t)h[0];
    e->stats.n_contrib = (int64_t)h[1];
    e->stats.bond_capacity = cap;
    e->stats.max_bonds = mx;
    return PGT_OK;
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
    if (!pot_params && n_pot_params > 0) return fail(PGT_ERR_ARG, "null parameters");
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
    const int adapt = mode == 2 && quad_mode == PGT_QUAD_ADAPTIVE;
    AdaptQ Q{adapt_limit, adapt_epsabs, adapt_epsrel};
    if (mode == 2 && quad_mode != PGT_QUAD_FIXED && quad_mode != PGT_QUAD_ADAPTIVE)
        return fail(PGT_ERR_ARG, "unknown quadrature mode");
    if (adapt && (adapt_limit < 1 || !(adapt_epsabs >= 0) || !(adapt_epsrel >= 0)))
        return fail(PGT_ERR_ARG, "adaptive limit must be >= 1 and tolerances non-negative");
    if (mode == 2 && !adapt) {
        std::vector<double> gs, gw;
        if (gl_rule(quad_order, gs, gw) != 0) return fail(PGT_ERR_ARG, "quadrature order must be in 1..1024");
        nq = quad_order;
        packed = gs;
        packed.insert(packed.end(), gw.begin(), gw.end());
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(PGT_ERR_HIP, "no HIP device available: kernel functions are evaluated on the GPU");
    double *d = nullptr;
    // interval tables of the caller's limit when it exceeds kQagMax
    const size_t wsn = (adapt && adapt_limit > kQagMax) ? (size_t)n * 4 * (size_t)adapt_limit : 0;
    if (wsn * sizeof(double) > kQagWorkspaceBytes)
        return fail(PGT_ERR_RANGE, "the adaptive interval tables of " + std::to_string(n) + " integrals at limit " +
                                       std::to_string(adapt_limit) + " exceed the " +
                                       std::to_string(kQagWorkspaceBytes >> 20) + " MiB workspace");
    const size_t tot = (size_t)n * 11 + packed.size() + 1 + wsn;  // + the capped flag + the tables
    HIPCHK(hipMalloc(&d, sizeof(double) * tot));
    double *dx = d, *dri = d + 3 * n, *drij = d + 6 * n, *ds = d + 9 * n, *dout = d + 10 * n, *dgl = d + 11 * n;
    int *dcap = reinterpret_cast<int *>(d + 11 * n + packed.size());
    double *dws = wsn ? d + 11 * n + packed.size() + 1 : nullptr;
    hipError_t he = hipMemset(dcap, 0, sizeof(int));
    if (he == hipSuccess) he = hipMemcpy(dx, x, sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(dri, ri, sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    if (he == hipSuccess && rij) he = hipMemcpy(drij, rij, sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    if (he == hipSuccess && s) he = hipMemcpy(ds, s, sizeof(double) * n, hipMemcpyHostToDevice);
    if (he == hipSuccess && !packed.empty())
        he = hipMemcpy(dgl, packed.data(), sizeof(double) * packed.size(), hipMemcpyHostToDevice);
    if (he == hipSuccess) {
        hipLaunchKernelGGL(k_eval_cg, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, cg_kind, cg_rcut, mode,
                           (int)n, s ? ds : nullptr, dx, dri, rij ? drij : nullptr, dgl, nq, adapt, Q, dcap, dws,
                           dout);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipDeviceSynchronize();
    if (he == hipSuccess) he = hipMemcpy(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost);
    int capped = 0;
    if (he == hipSuccess) he = hipMemcpy(&capped, dcap, sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (he != hipSuccess) return fail(PGT_ERR_HIP, std::string("kernel evaluation: ") + hipGetErrorString(he));
    if (capped) return fail(PGT_ERR_STATE, "adaptive interval tables smaller than the caller's limit");
    return PGT_OK;
}

// ------------------------------------------------------------------------
// learned bond-integral surrogate
// ------------------------------------------------------------------------
int pgt_mlp_nparam(void) { return kMlpNParam; }

size_t pgt_mlp_work_floats(int64_t n) { return n > 0 ? mlp_work_floats(n) : 0; }

int pgt_mlp_forward(const float *d_params, int64_t n, const float *d_x, float *d_y, float *d_work, void *stream)
{
    if (n < 0 || (n > 0 && (!d_params || !d_x || !d_y || !d_work))) return fail(PGT_ERR_ARG, "bad network arguments");
    if (n == 0) return PGT_OK;
    hipStream_t s = (hipStream_t)stream;
    const float *P = d_params;
    float *H1 = d_work, *H2 = d_work + 64 * n, *H3 = d_work + 128 * n;
    hipLaunchKernelGGL((k_dense_fwd<kMlpIn, kMlpW>), dim3(blocks_for(n * 64)), dim3(256), 0, s, n, d_x, kMlpInPad,
                       P + kOffW1, P + kOffB1, true, H1);
    hipLaunchKernelGGL((k_dense_fwd<kMlpW, kMlpW>), dim3(blocks_for(n * 64)), dim3(256), 0, s, n, H1, kMlpW,
                       P + kOffW2, P + kOffB2, true, H2);
    hipLaunchKernelGGL((k_dense_fwd<kMlpW, kMlpW>), dim3(blocks_for(n * 64)), dim3(256), 0, s, n, H2, kMlpW,
                       P + kOffW3, P + kOffB3, true, H3);
    hipLaunchKernelGGL((k_dense_fwd<kMlpW, 1>), dim3(blocks_for(n)), dim3(256), 0, s, n, H3, kMlpW, P + kOffW4,
                       P + kOffB4, false, d_y);
    HIPCHK(hipGetLastError());
    return PGT_OK;
}

int pgt_mlp_grad(const float *d_params, int64_t n, const float *d_x, const float *d_t, float *d_grad, float *d_loss,
                 float *d_work, void *stream)
{
    if (n <= 0 || !d_params || !d_x || !d_t || !d_grad || !d_loss || !d_work)
        return fail(PGT_ERR_ARG, "bad network arguments");
    hipStream_t s = (hipStream_t)stream;
    const float *P = d_params;
    float *H1 = d_work, *H2 = d_work + 64 * n, *H3 = d_work + 128 * n, *dZa = d_work + 192 * n,
          *dZb = d_work + 256 * n, *Y = d_work + 320 * n, *dY = d_work + 321 * n;
    double *part = reinterpret_cast<double *>(d_work + 322 * n);
    int rc;
    if ((rc = pgt_mlp_forward(d_params, n, d_x, Y, d_work, stream))) return rc;
    const unsigned nb = blocks_for(n);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, s, n, Y, d_t, dY, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, s, (int)nb, part, n, d_loss);
    float *g = d_grad;
    float *gpart = d_work + mlp_part_offset(n);
    // weight and bias gradient of a layer with K inputs X (row stride ldx), N outputs, error dZ
    auto wgrad = [&](int K, int N, const float *X, int ldx, const float *dZ, float *gW, float *gb) {
        const int M = K * N + N;
        hipLaunchKernelGGL(k_dense_bwd_weight, dim3(kGradSplit), dim3(256), 0, s, n, K, N, X, ldx, dZ, gpart);
        hipLaunchKernelGGL(k_sum_split, dim3(blocks_for(M)), dim3(256), 0, s, K, N, gpart, gW, gb);
    };
    wgrad(kMlpW, 1, H3, kMlpW, dY, g + kOffW4, g + kOffB4);
    hipLaunchKernelGGL(k_dense_bwd_data<1>, dim3(blocks_for(n * 64)), dim3(256), 0, s, n, dY, P + kOffW4, H3, dZa);
    wgrad(kMlpW, kMlpW, H2, kMlpW, dZa, g + kOffW3, g + kOffB3);
    hipLaunchKernelGGL(k_dense_bwd_data<kMlpW>, dim3(blocks_for(n * 64)), dim3(256), 0, s, n, dZa, P + kOffW3, H2,
                       dZb);
    wgrad(kMlpW, kMlpW, H1, kMlpW, dZb, g + kOffW2, g + kOffB2);
    hipLaunchKernelGGL(k_dense_bwd_data<kMlpW>, dim3(blocks_for(n * 64)), dim3(256), 0, s, n, dZb, P + kOffW2, H1,
                       dZa);
    wgrad(kMlpIn, kMlpW, d_x, kMlpInPad, dZa, g + kOffW1, g + kOffB1);
    HIPCHK(hipGetLastError());
    return PGT_OK;
}

int pgt_mlp_adam(float *d_params, float *d_m, float *d_v, const float *d_grad, int step, float lr, void *stream)
{
    if (!d_params || !d_m || !d_v || !d_grad || step < 1 || !(lr > 0.0f))
        return fail(PGT_ERR_ARG, "bad optimiser arguments");
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(kMlpNParam)), dim3(256), 0, (hipStream_t)stream, kMlpNParam, d_params,
                       d_m, d_v, d_grad, lr, b1, b2, 1e-8f, c1, c2);
    HIPCHK(hipGetLastError());
    return PGT_OK;
}

int pgt_mlp_infer(const float *d_params, int64_t n, const float *d_x, float *d_y, void *stream)
{
    if (n < 0 || (n > 0 && (!d_params || !d_x || !d_y))) return fail(PGT_ERR_ARG, "bad network arguments");
    if (n == 0) return PGT_OK;
    const int64_t groups = (n + 15) / 16;  // 16 samples per wavefront step
    const unsigned nblk = (unsigned)std::min<int64_t>((groups + 3) / 4, 2048);
    hipLaunchKernelGGL(k_mlp_infer, dim3(nblk), dim3(256), 0, (hipStream_t)stream, n, d_params, d_x, d_y);
    HIPCHK(hipGetLastError());
    return PGT_OK;
}

int pgt_ik_bond_samples(pgt_ik_engine *e, int64_t max_pairs, float *d_x, float *d_t, int64_t *n_pairs)
{
    if (int rc = surrogate_check(e)) return rc;
    if (!n_pairs || max_pairs < 0 || (d_x && !d_t)) return fail(PGT_ERR_ARG, "bad sample arguments");
    int rc;
    int64_t np = 0;
    if ((rc = prepare_pairs(e, &np, false, nullptr))) return rc;
    *n_pairs = np;
    if (!d_x || np == 0) return PGT_OK;
    if (np > max_pairs) return fail(PGT_ERR_RANGE, "the sample buffers hold fewer rows than the frame has pairs");
    if ((rc = prepare_pairs(e, &np, true, d_x))) return rc;
    hipLaunchKernelGGL(k_pair_finalize, dim3(blocks_for(np)), dim3(256), 0, e->stream, np, e->d_prec.p,
                       (const double *)e->d_bonds.p, engine_rule(e), (const float *)nullptr, 0.0, 0.0, -1.0,
                       e->d_pw.p, d_t, (unsigned long long *)nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    return PGT_OK;
}

int pgt_ik_compute_surrogate(pgt_ik_engine *e, const float *d_params, double bound_inside, double bound_crossing,
                             double tol, double *virial)
{
    if (int rc = surrogate_check(e)) return rc;
    if (!d_params || !virial) return fail(PGT_ERR_ARG, "null network parameters or output");
    if (!(bound_inside >= 0.0) || !(bound_crossing >= 0.0)) return fail(PGT_ERR_ARG, "error bounds must be >= 0");
    const int G = (int)e->G;
    e->stats = pgt_ik_stats{};
    if (G == 0) return PGT_OK;
    if (e->n == 0) {
        std::memset(virial, 0, (size_t)G * 4 * sizeof(double));
        return PGT_OK;
    }
    int rc;
    // no pair class qualifies for the network: the field is the plain form's,
    // so run the plain gather (the pair list would only cost time)
    if (!(tol >= 0.0 && std::min(bound_inside, bound_crossing) <= tol)) {
        if ((rc = pgt_ik_compute(e, PGT_FIELD_VIRIAL, virial, nullptr, nullptr))) return rc;
        e->stats.n_pairs = e->stats.n_contrib;
        e->stats.n_network = 0;
        e->stats.surrogate_ms = 0.0;
        return PGT_OK;
    }
    // events 0..2 belong to the bond-list job inside prepare_pairs
    if ((rc = ensure_events(e, 6))) return rc;
    if ((rc = e->d_out.ensure((size_t)G * 4)) || (rc = e->d_pn.ensure(2))) return rc;
    hipStream_t s = e->stream;
    HIPCHK(hipEventRecord(e->ev[3], s));
    int64_t np = 0;
    if ((rc = prepare_pairs(e, &np, true, nullptr))) return rc;
    const int mx = e->stats.max_bonds, cap = e->stats.bond_capacity;
    HIPCHK(hipEventRecord(e->ev[4], s));
    if ((rc = fill_async(e->d_pn.p, 0ull, 1, s))) return rc;
    const bool any = tol >= 0.0 && std::min(bound_inside, bound_crossing) <= tol;
    if (np > 0) {
        if (any) {
            if ((rc = e->d_pymlp.ensure((size_t)np))) return rc;
            if ((rc = pgt_mlp_infer(d_params, np, e->d_pxin.p, e->d_pymlp.p, s))) return rc;
        }
        hipLaunchKernelGGL(k_pair_finalize, dim3(blocks_for(np)), dim3(256), 0, s, np, e->d_prec.p,
                           (const double *)e->d_bonds.p, engine_rule(e), any ? (const float *)e->d_pymlp.p : nullptr,
                           bound_inside, bound_crossing, any ? tol : -1.0, e->d_pw.p, (float *)nullptr, e->d_pn.p);
    }
    hipLaunchKernelGGL(k_pair_reduce, dim3((unsigned)((G + 127) / 128)), dim3(128), 0, s, G, e->d_pcount.p,
                       e->d_poff.p, e->d_prec.p, (const double *)e->d_bonds.p, e->d_pw.p, e->d_out.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e->ev[5], s));
    unsigned long long used = 0;
    HIPCHK(hipMemcpyAsync(&used, e->d_pn.p, sizeof(used), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(virial, e->d_out.p, (size_t)G * 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    float tot = 0.0f, pairs_ms = 0.0f;
    (void)hipEventElapsedTime(&tot, e->ev[3], e->ev[5]);
    (void)hipEventElapsedTime(&pairs_ms, e->ev[4], e->ev[5]);
    e->stats.n_particles = e->n;
    e->stats.n_grid = G;
    e->stats.n_contrib = np;
    e->stats.bond_capacity = cap;
    e->stats.max_bonds = mx;
    e->stats.total_ms = tot;
    e->stats.n_pairs = np;
    e->stats.n_network = (int64_t)used;
    e->stats.surrogate_ms = pairs_ms;
    return PGT_OK;
}

}  // extern "C"

// ==========================================================================
// sweep: the frames of several systems, all 2D or all 3D, one launch per stage
// ==========================================================================
struct pgt_ik_sweep {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    unsigned fields = 0;
    bool vir = false, kin = false, fp32 = false;
    int D = 2;             // 2: 2D systems (z = 0), 2 x 2 fields; 3: 3D, 3 x 3
    int ns = 0;
    std::vector<SweepSys> tab;
    GLRule R{};
    int npass = 1, nbits = 1, radix = 2;
    int64_t nframes = 0;   // frames of every system
    int64_t np = 0;        // particles of every frame
    int64_t npoints = 0;   // grid points of every frame
    int64_t ncs = 0;       // cell starts of every frame
    int64_t nh = 0;        // histogram entries of every frame
    int blocks[SW_NSTAGE] = {0, 0, 0, 0};
    size_t ebytes = 0;     // bytes of one bond entry
    int cap = 0;
    bool ready = false;    // buffers sized by a run
    bool held = false;     // enqueued launches a captured graph may replay
    DBuf<SweepSys> d_tab;
    DBuf<double> d_grid;   // the systems' grids, back to back
    DBuf<int> d_perm;      // and their Morton orders
    DBuf<unsigned char> d_keys0, d_vals0, d_keys1, d_vals1, d_hist, d_cstart, d_spos, d_sdiam, d_bcount, d_skin;
    DBuf<unsigned char> d_bonds, d_wcounts, d_wmax, d_rec;
    DBuf<int> d_maxcount;                // [0] largest bond count, [1] tile overflow bits
    DBuf<unsigned long long> d_stats;    // balls, contributions
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    pgt_ik_stats stats{};
};

namespace {

int check_sweep(pgt_ik_sweep *s)
{
    if (!s) return fail(PGT_ERR_ARG, "null sweep");
    return 0;
}

int fail_sweep_held()
{
    return fail(PGT_ERR_STATE, "the sweep's buffers are held by enqueued work that a captured graph may replay: "
                               "call pgt_ik_sweep_release first, and capture again afterwards");
}

// Work arrays of the sweep at its current bond capacity; grown unless held
// (then PGT_ERR_STATE when too small).
int sweep_buffers(pgt_ik_sweep *s)
{
    const size_t np = (size_t)s->np;
    const size_t PS = s->D == 2 ? 2 : 4, KN = s->D == 2 ? 3 : 6;
    const size_t want[14] = {np * sizeof(int), np * sizeof(int), np * sizeof(int), np * sizeof(int),
                             (size_t)s->nh * sizeof(int), (size_t)s->ncs * sizeof(int), np * PS * sizeof(double),
                             np * sizeof(double), s->vir ? np * sizeof(int) : 0, s->kin ? np * KN * sizeof(double) : 0,
                             s->vir ? np * (size_t)s->cap * s->ebytes : 0,
                             (size_t)s->blocks[SW_GATHER] * 2 * sizeof(unsigned long long),
                             s->vir ? (size_t)s->blocks[SW_BONDS] * (bond_block(s->D) / 64) * sizeof(int) : 0,
                             (s->vir && s->D == 3) ? np * sizeof(float4) : 0};
    DBuf<unsigned char> *bufs[14] = {&s->d_keys0, &s->d_vals0, &s->d_keys1, &s->d_vals1, &s->d_hist,
                                     &s->d_cstart, &s->d_spos, &s->d_sdiam, &s->d_bcount, &s->d_skin,
                                     &s->d_bonds, &s->d_wcounts, &s->d_wmax, &s->d_rec};
    for (int k = 0; k < 14; ++k) {
        if (!want[k] || (bufs[k]->p && bufs[k]->n >= want[k])) continue;
        if (s->held) return fail_sweep_held();
        if (int rc = bufs[k]->ensure(want[k])) return rc;
    }
    if (int rc = s->d_maxcount.ensure(2)) return rc;
    return s->d_stats.ensure(2);
}

// Every launch of the sweep on its stream, no host synchronisation.  With
// timed, events around the whole and around the gather.
int sweep_launch(pgt_ik_sweep *s, bool timed)
{
    hipStream_t st = s->stream;
    const SweepSys *T = s->d_tab.p;
    const int ns = s->ns;
    int *kb[2] = {(int *)s->d_keys0.p, (int *)s->d_keys1.p}, *vb[2] = {(int *)s->d_vals0.p, (int *)s->d_vals1.p};
    int *hist = (int *)s->d_hist.p, *cstart = (int *)s->d_cstart.p, *bcount = (int *)s->d_bcount.p;
    double *spos = (double *)s->d_spos.p, *sdiam = (double *)s->d_sdiam.p, *skin = (double *)s->d_skin.p;
    if (timed) HIPCHK(hipEventRecord(s->ev[0], st));
    const dim3 sg((unsigned)s->blocks[SW_SORT]);
    for (int p = 0; p < s->npass; ++p) {
        const int shift = p * s->nbits, radix = s->radix;
        int *kin = kb[p & 1], *vin = vb[p & 1], *kout = kb[(p + 1) & 1], *vout = vb[(p + 1) & 1];
        if (p == 0) {
            hipLaunchKernelGGL(k_sweep_hist<true>, sg, dim3(kSortBlock), 0, st, T, ns, shift, radix, kin, hist);
            hipLaunchKernelGGL(k_sweep_scan, dim3((unsigned)s->nframes), dim3(1024), 0, st, T, ns, radix, hist);
            hipLaunchKernelGGL(k_sweep_scatter<true>, sg, dim3(kSortBlock), 0, st, T, ns, shift, radix, kin, vin,
                               hist, kout, vout);
        } else {
            hipLaunchKernelGGL(k_sweep_hist<false>, sg, dim3(kSortBlock), 0, st, T, ns, shift, radix, kin, hist);
            hipLaunchKernelGGL(k_sweep_scan, dim3((unsigned)s->nframes), dim3(1024), 0, st, T, ns, radix, hist);
            hipLaunchKernelGGL(k_sweep_scatter<false>, sg, dim3(kSortBlock), 0, st, T, ns, shift, radix, kin, vin,
                               hist, kout, vout);
        }
    }
    float4 *rec = (s->vir && s->D == 3) ? (float4 *)s->d_rec.p : nullptr;
    const dim3 fg((unsigned)s->blocks[SW_FINISH]);
    const int *keys = kb[s->npass & 1], *vals = vb[s->npass & 1];
    if (s->D == 2)
        hipLaunchKernelGGL((k_sweep_finish<2, 2>), fg, dim3(256), 0, st, T, ns, keys, vals, cstart, spos, sdiam,
                           s->kin ? skin : nullptr, (float4 *)nullptr);
    else
        hipLaunchKernelGGL((k_sweep_finish<3, 3>), fg, dim3(256), 0, st, T, ns, keys, vals, cstart, spos, sdiam,
                           s->kin ? skin : nullptr, rec);
    HIPCHK(hipGetLastError());
    int *wmax = (int *)s->d_wmax.p;
    if (s->vir) {
        const dim3 bg((unsigned)s->blocks[SW_BONDS]);
        constexpr int B2 = bond_block(2), M2 = bond_list(2), B3 = bond_block(3), M3 = bond_list(3);
        if (s->D == 2 && s->fp32)
            hipLaunchKernelGGL((k_sweep_bonds<2, 2, float, B2, M2>), bg, dim3(B2), 0, st, T, ns, s->cap, cstart, spos,
                               sdiam, bcount, (float *)s->d_bonds.p, wmax, rec);
        else if (s->D == 2)
            hipLaunchKernelGGL((k_sweep_bonds<2, 2, double, B2, M2>), bg, dim3(B2), 0, st, T, ns, s->cap, cstart, spos,
                               sdiam, bcount, (double *)s->d_bonds.p, wmax, rec);
        else if (s->fp32)
            hipLaunchKernelGGL((k_sweep_bonds<3, 3, float, B3, M3>), bg, dim3(B3), 0, st, T, ns, s->cap, cstart, spos,
                               sdiam, bcount, (float *)s->d_bonds.p, wmax, rec);
        else
            hipLaunchKernelGGL((k_sweep_bonds<3, 3, double, B3, M3>), bg, dim3(B3), 0, st, T, ns, s->cap, cstart, spos,
                               sdiam, bcount, (double *)s->d_bonds.p, wmax, rec);
        HIPCHK(hipGetLastError());
    }
    if (timed) HIPCHK(hipEventRecord(s->ev[1], st));
    SweepGatherCtx c{};
    c.tab = T;
    c.ns = ns;
    c.blocks = s->blocks[SW_GATHER];
    c.R = s->R;
    c.cstart = cstart;
    c.spos = spos;
    c.skin = s->kin ? skin : nullptr;
    c.cap = s->cap;
    c.bcount = s->vir ? bcount : nullptr;
    c.rec = rec;
    c.bonds = s->vir ? (const void *)s->d_bonds.p : nullptr;
    c.wave_counts = (unsigned long long *)s->d_wcounts.p;
    c.flags = s->d_maxcount.p + 1;
    c.stream = st;
    launch_sweep_gather(c, s->D, s->fp32, s->fields);
    HIPCHK(hipGetLastError());
    if (timed) HIPCHK(hipEventRecord(s->ev[2], st));
    const int nwg = s->blocks[SW_GATHER], nwb = s->vir ? s->blocks[SW_BONDS] * (bond_block(s->D) / 64) : 0;
    const int nrb = std::max(1, std::min(256, (std::max(nwg, nwb) + 4095) / 4096));
    hipLaunchKernelGGL(k_reduce_stats, dim3(nrb), dim3(1024), 0, st, nwg, c.wave_counts, nwb, wmax, s->d_stats.p,
                       s->d_maxcount.p);
    HIPCHK(hipGetLastError());
    return 0;
}

int sweep_flags_error(int flags1)
{
    if (flags1 & 6)
        return fail(PGT_ERR_RANGE, "a particle had more than 255 bonds, more than the tile gather holds: compute "
                                   "these systems with their engines and the lane gather (pgt_ik_set_gather_mode)");
    return 0;
}

}  // namespace

extern "C" {

int pgt_ik_sweep_create(int nsys, const pgt_ik_sweep_system *systems, unsigned fields, void *hip_stream,
                        pgt_ik_sweep **out)
{
    if (!out) return fail(PGT_ERR_ARG, "null output handle");
    *out = nullptr;
    if (nsys < 1 || nsys > kSweepMaxSystems)
        return fail(PGT_ERR_ARG, "a sweep holds 1 to " + std::to_string(kSweepMaxSystems) + " systems");
    if (!systems) return fail(PGT_ERR_ARG, "null systems");
    if (fields == 0 || (fields & ~(unsigned)(PGT_FIELD_VIRIAL | PGT_FIELD_KINETIC)))
        return fail(fields & PGT_FIELD_ELASTIC ? PGT_ERR_UNSUPPORTED : PGT_ERR_ARG,
                    "a sweep computes the virial and/or kinetic fields");
    int ndev = 0, dev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(PGT_ERR_HIP, "no HIP device available: the sweep runs on the GPU only");
    HIPCHK(hipGetDevice(&dev));
    const bool vir = fields & PGT_FIELD_VIRIAL, kin = fields & PGT_FIELD_KINETIC;
    bool fp32 = false;
    int D = 2;
    for (int k = 0; k < nsys; ++k) {
        const pgt_ik_sweep_system &y = systems[k];
        const pgt_ik_engine *e = y.engine;
        const std::string at = "system " + std::to_string(k) + ": ";
        if (!e) return fail(PGT_ERR_ARG, at + "null engine");
        if (!e->box_set || !e->have_grid || e->G < 1)
            return fail(PGT_ERR_STATE, at + "set_box and a nonempty set_grid must precede the sweep");
        if (e->device != dev) return fail(PGT_ERR_ARG, at + "engine on another device");
        if (e->prm.dim == 2 && !e->grid_z_flat)
            return fail(PGT_ERR_UNSUPPORTED, at + "a sweep runs 2D systems with grid and positions at z = 0, or 3D ones");
        if (k == 0) D = e->prm.dim;
        else if (e->prm.dim != D) return fail(PGT_ERR_ARG, at + "every system of a sweep has the same dimension");
        if (e->prm.cg_kind != PGT_CG_OCTIC || e->prm.quad_mode != PGT_QUAD_FIXED || e->gl_s.size() != 5)
            return fail(PGT_ERR_UNSUPPORTED, at + "a sweep runs the Octic kernel with the fixed 5-node rule");
        const bool f32 = e->prm.precision == PGT_PREC_FP32;
        if (k == 0) fp32 = f32;
        else if (f32 != fp32) return fail(PGT_ERR_ARG, at + "every system of a sweep has the same precision");
        if (y.nframes < 1) return fail(PGT_ERR_ARG, at + "at least one frame");
        if (y.n < 1 || y.n > kTileMaxParticles)
            return fail(PGT_ERR_RANGE, at + "particles per frame must be 1 .. 2^26");
        if (!y.d_pos || !y.d_diam) return fail(PGT_ERR_ARG, at + "positions and diameters are required");
        if ((vir && !y.d_out_virial) || (kin && !y.d_out_kinetic))
            return fail(PGT_ERR_ARG, at + "null output for a selected field");
    }
    auto *s = new pgt_ik_sweep();
    auto bail = [&](int rc) {
        pgt_ik_sweep_destroy(s);
        return rc;
    };
    s->device = dev;
    s->fields = fields;
    s->vir = vir;
    s->kin = kin;
    s->fp32 = fp32;
    s->D = D;
    s->ns = nsys;
    s->R = engine_rule(systems[0].engine);
    if (D == 2)
        s->ebytes = fp32 ? sizeof(float) * BondLayout<2, 2, false, float>::ES
                         : sizeof(double) * BondLayout<2, 2, false, double>::ES;
    else
        s->ebytes = fp32 ? sizeof(float) * BondLayout<3, 3, false, float>::ES
                         : sizeof(double) * BondLayout<3, 3, false, double>::ES;
    s->tab.resize((size_t)nsys);
    int maxcell = 1, cap = 0;
    int64_t gtotal = 0;
    for (int k = 0; k < nsys; ++k) {
        const pgt_ik_engine *e = systems[k].engine;
        SweepSys &S = s->tab[(size_t)k];
        S = SweepSys{};
        S.g = make_geom(e, D);
        S.P = make_pot(e->prm);
        S.cgnorm = cg_norm(e->prm.cg_kind, e->prm.cg_rcut);
        S.n = (int)systems[k].n;
        S.G = (int)e->G;
        S.nf = systems[k].nframes;
        S.nchunk = (S.n + kSortChunk - 1) / kSortChunk;
        maxcell = std::max(maxcell, S.g.ncell);
        cap = std::max(cap, e->cap);
        gtotal += e->G;
    }
    int kbits = 0;
    while ((1 << kbits) < maxcell) ++kbits;
    s->npass = std::max(1, (kbits + 7) / 8);
    s->nbits = std::max(1, (kbits + s->npass - 1) / s->npass);
    s->radix = 1 << s->nbits;
    s->cap = cap > 0 ? cap : 16;
    // offsets and workgroup ranges, systems in order
    int64_t blk[SW_NSTAGE] = {0, 0, 0, 0};
    for (int k = 0; k < nsys; ++k) {
        SweepSys &S = s->tab[(size_t)k];
        S.f0 = s->nframes;
        S.p0 = s->np;
        S.c0 = s->ncs;
        S.h0 = s->nh;
        s->nframes += S.nf;
        s->np += (int64_t)S.nf * S.n;
        s->npoints += (int64_t)S.nf * S.G;
        s->ncs += (int64_t)S.nf * (S.g.ncell + 1);
        s->nh += (int64_t)S.nf * s->radix * S.nchunk;
        S.bpf[SW_SORT] = S.nchunk;
        S.bpf[SW_FINISH] = (S.n + 255) / 256;
        S.bpf[SW_BONDS] = (S.n + bond_block(D) - 1) / bond_block(D);
        S.bpf[SW_GATHER] = (S.G + kTilePoints - 1) / kTilePoints;
        for (int st = 0; st < SW_NSTAGE; ++st) {
            const int64_t nb = (st == SW_SORT || st == SW_FINISH)
                                   ? (int64_t)frame_rows(S.nf) * S.bpf[st]
                                   : ((int64_t)S.nf * S.bpf[st] + 7) / 8 * 8;
            S.blk0[st] = (int)blk[st];
            S.nblk[st] = (int)nb;
            blk[st] += nb;
            if (blk[st] > (int64_t)1 << 24)
                return bail(fail(PGT_ERR_RANGE, "the sweep's launches exceed 2^24 workgroups: split the sweep"));
        }
        S.pos = systems[k].d_pos;
        S.diam = systems[k].d_diam;
        S.mass = systems[k].d_mass;
        S.vel = systems[k].d_vel;
        S.vir = vir ? systems[k].d_out_virial : nullptr;
        S.kin = kin ? systems[k].d_out_kinetic : nullptr;
    }
    for (int st = 0; st < SW_NSTAGE; ++st) s->blocks[st] = (int)blk[st];
    if (s->np > ((int64_t)1 << 31) - 1)
        return bail(fail(PGT_ERR_RANGE, "the sweep holds more than 2^31 - 1 particles over all frames: split it"));
    int rc;
    if ((rc = s->d_grid.ensure((size_t)gtotal * 3)) || (rc = s->d_perm.ensure((size_t)gtotal)) ||
        (rc = s->d_tab.ensure((size_t)nsys)))
        return bail(rc);
    int64_t goff = 0;
    for (int k = 0; k < nsys; ++k) {
        const pgt_ik_engine *e = systems[k].engine;
        SweepSys &S = s->tab[(size_t)k];
        S.grid = s->d_grid.p + goff * 3;
        S.perm = s->d_perm.p + goff;
        if (hipMemcpy((void *)S.grid, e->d_grid.p, sizeof(double) * 3 * (size_t)e->G, hipMemcpyDeviceToDevice) !=
                hipSuccess ||
            hipMemcpy((void *)S.perm, e->d_perm.p, sizeof(int) * (size_t)e->G, hipMemcpyDeviceToDevice) != hipSuccess)
            return bail(fail(PGT_ERR_HIP, "copying a system's grid"));
        goff += e->G;
    }
    if (hipMemcpy(s->d_tab.p, s->tab.data(), sizeof(SweepSys) * (size_t)nsys, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(PGT_ERR_HIP, "uploading the sweep's table"));
    if (hipStreamCreate(&s->own_stream) != hipSuccess) return bail(fail(PGT_ERR_HIP, "hipStreamCreate"));
    s->stream = hip_stream ? (hipStream_t)hip_stream : s->own_stream;
    for (auto &v : s->ev)
        if (hipEventCreate(&v) != hipSuccess) return bail(fail(PGT_ERR_HIP, "hipEventCreate"));
    *out = s;
    return PGT_OK;
}

void pgt_ik_sweep_destroy(pgt_ik_sweep *s)
{
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    DBuf<unsigned char> *bufs[] = {&s->d_keys0, &s->d_vals0, &s->d_keys1, &s->d_vals1, &s->d_hist,
                                   &s->d_cstart, &s->d_spos, &s->d_sdiam, &s->d_bcount, &s->d_skin,
                                   &s->d_bonds, &s->d_wcounts, &s->d_wmax, &s->d_rec};
    for (auto *b : bufs) b->release();
    s->d_tab.release();
    s->d_grid.release();
    s->d_perm.release();
    s->d_maxcount.release();
    s->d_stats.release();
    for (hipEvent_t v : s->ev)
        if (v) (void)hipEventDestroy(v);
    if (s->own_stream) (void)hipStreamDestroy(s->own_stream);
    delete s;
}

int pgt_ik_sweep_run(pgt_ik_sweep *s)
{
    if (int rc = check_sweep(s)) return rc;
    int mx = 0;
    for (int attempt = 0; attempt < 4; ++attempt) {
        int rc;
        if ((rc = sweep_buffers(s))) return rc;
        if ((rc = fill_async(s->d_maxcount.p, 0, 2, s->stream)) || (rc = fill_async(s->d_stats.p, 0ull, 2, s->stream)))
            return rc;
        if ((rc = sweep_launch(s, true))) return rc;
        int flags[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(flags, s->d_maxcount.p, sizeof(flags), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        if ((rc = sweep_flags_error(flags[1]))) return rc;
        mx = flags[0];
        if (!(s->vir && mx > s->cap)) break;
        if (mx > 65535) return fail(PGT_ERR_RANGE, "a particle has more than 65535 interacting neighbours");
        if (s->held) return fail_sweep_held();  // the graph's launches keep the old capacity
        s->cap = ((mx + 7) / 8) * 8;
        if (attempt == 3) return fail(PGT_ERR_HIP, "bond list capacity did not settle");
    }
    unsigned long long h[2] = {0, 0};
    HIPCHK(hipMemcpy(h, s->d_stats.p, sizeof(h), hipMemcpyDeviceToHost));
    float g_ms = 0.0f, tot = 0.0f;
    (void)hipEventElapsedTime(&g_ms, s->ev[1], s->ev[2]);
    (void)hipEventElapsedTime(&tot, s->ev[0], s->ev[2]);
    s->stats = pgt_ik_stats{};
    s->stats.n_particles = s->np;
    s->stats.n_grid = s->npoints;
    s->stats.n_ball = (int64_t)h[0];
    s->stats.n_contrib = (int64_t)h[1];
    s->stats.bond_capacity = s->cap;
    s->stats.max_bonds = mx;
    s->stats.gather_ms = g_ms;
    s->stats.total_ms = tot;
    s->ready = true;
    return PGT_OK;
}

int pgt_ik_sweep_enqueue(pgt_ik_sweep *s)
{
    if (int rc = check_sweep(s)) return rc;
    if (!s->ready) return fail(PGT_ERR_STATE, "run pgt_ik_sweep_run before enqueueing the sweep");
    if (int rc = sweep_launch(s, false)) return rc;
    s->held = true;
    return PGT_OK;
}

int pgt_ik_sweep_release(pgt_ik_sweep *s)
{
    if (int rc = check_sweep(s)) return rc;
    s->held = false;
    return PGT_OK;
}

int pgt_ik_sweep_finish(pgt_ik_sweep *s)
{
    if (int rc = check_sweep(s)) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    if (!s->ready) return PGT_OK;
    int flags[2] = {0, 0};
    unsigned long long h[2] = {0, 0};
    HIPCHK(hipMemcpy(flags, s->d_maxcount.p, sizeof(flags), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(h, s->d_stats.p, sizeof(h), hipMemcpyDeviceToHost));
    int rc = fill_async(s->d_maxcount.p, 0, 2, s->stream);
    if (!rc) rc = fill_async(s->d_stats.p, 0ull, 2, s->stream);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    s->stats.n_ball = (int64_t)h[0];
    s->stats.n_contrib = (int64_t)h[1];
    s->stats.max_bonds = flags[0];
    s->stats.bond_capacity = s->cap;
    if ((rc = sweep_flags_error(flags[1]))) return rc;
    if (s->vir && flags[0] > s->cap)
        return fail(PGT_ERR_RANGE, "a particle had more bonds than the capacity sized by the last "
                                   "pgt_ik_sweep_run: the enqueued results are incomplete; release, run and "
                                   "capture again");
    return PGT_OK;
}

int pgt_ik_sweep_set_stream(pgt_ik_sweep *s, void *hip_stream)
{
    if (int rc = check_sweep(s)) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    s->stream = hip_stream ? (hipStream_t)hip_stream : s->own_stream;
    return PGT_OK;
}

int pgt_ik_sweep_get_stats(pgt_ik_sweep *s, pgt_ik_stats *out)
{
    if (int rc = check_sweep(s)) return rc;
    if (!out) return fail(PGT_ERR_ARG, "null stats");
    *out = s->stats;
    return PGT_OK;
}

}  // extern "C"
