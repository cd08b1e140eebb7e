/*
 * ik_oracle.c -- FP64 CPU restatement of the pyglasstools Irving-Kirkwood
 * stress-field path.  TEST INFRASTRUCTURE ONLY: this file is the checker that
 * the HIP engine is compared against, and the CPU baseline that bench.py
 * times beside it.  Nothing in the product (pyglasstools_amd/) may link, load
 * or call it.
 *
 * What it restates (reference files, relative to the reference checkout):
 *   - grid loop, ball rule, j search, 1/2 factor:
 *       pyglasstools/irvingkirkwood/IrvingKirkwood.h:74-110
 *   - virial accumulation:  pyglasstools/ListofCommonObservables.h:49-82
 *   - elastic virial:       pyglasstools/ListofCommonObservables.h:86-123
 *   - kinetic (intended formula; the reference's LocalField never sizes its
 *     tensors, CoarseGrainedField.h:58-64, so it computes nothing):
 *       pyglasstools/ListofCommonObservables.h:125-149
 *   - bond integral:        irvingkirkwood/cgfunc/CoarseGrainFunction.h:42-45,78-88
 *     with GSL's fixed Gauss-Legendre rule (FixedPointQuadrature.h:39-55),
 *     restated as the textbook n-point rule on [0,1] summed in symmetric
 *     node pairs (A&S 25.4.29).
 *   - adaptive bond integral: CoarseGrainFunction.h:124-127 ->
 *     Quadrature.h:39-51 (GSL QAG, 15-point Gauss-Kronrod, key=1), restated
 *     from the published QUADPACK QAG algorithm (Piessens et al. 1983).
 *   - coarse-graining kernels: cgfunc/CGFuncOctic.h:16-29, CGFuncRect.h:17-28,
 *     CGFuncMollifier.h:11-31
 *   - pair forces:  potential/PairPotential.h:98-104 with
 *     Polydisperse12.h:38-54, Polydisperse18.h:37-53, Polydisperse10.h:37-52,
 *     Polydisperse106.h:37-53, PolydisperseLJ.h:37-53
 *   - max_rcut:     pyglasstools/Calculator.h:41-45 (passed in by the caller)
 *
 * The neighbour search of the reference (Aboria's periodic cell list) is
 * replaced by a plain CPU linked-cell list, or by brute force when
 * `brute` is set; both enumerate the same (x,i,j) triples.  Aboria's dx()
 * sign convention does not matter: flipping every dx() reflects each bond
 * segment through the grid point, and both the radial kernel and F (x) r are
 * invariant under that reflection.
 *
 * Two loop flavours are provided:
 *   faithful (mode 0): the reference's per-triple costs are kept -- a heap
 *       quadrature table per bond integral (FixedPointQuadrature.h:35), a heap
 *       temporary tensor per triple (ListofCommonObservables.h:78), calls
 *       through function pointers, and the bond integral evaluated even when
 *       the pair force is zero.
 *   clean (mode 1): the same arithmetic without the allocations and with
 *       zero-force triples skipped (exact: they add 0.5*bond*0).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

enum { POT_POLY12 = 0, POT_POLY18 = 1, POT_POLY10 = 2, POT_POLY106 = 3, POT_POLYLJ = 4 };
enum { CG_OCTIC = 0, CG_RECT = 1, CG_MOLLIFIER = 2 };
enum { QUAD_FIXED = 0, QUAD_ADAPTIVE = 1 };

/* ------------------------------------------------------------------ */
/* Gauss-Legendre nodes on [-1,1]: Newton iteration on P_n in long double. */
/* Output: the m = (n+1)/2 non-negative abscissae in ascending order       */
/* (x[0] = 0 when n is odd) and their weights.                             */
/* ------------------------------------------------------------------ */
int oracle_gl_nodes(int n, double *x, double *w)
{
    if (n < 1 || n > 4096) return -1;
    int m = (n + 1) / 2;
    for (int k = 0; k < m; ++k) {
        /* k-th root from the top: cos(pi (k+0.75)/(n+0.5)) initial guess */
        long double z = cosl((long double)M_PI * ((long double)k + 0.75L) / ((long double)n + 0.5L));
        long double pp = 0.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = z;
            for (int j = 2; j <= n; ++j) {
                long double p2 = ((2.0L * j - 1.0L) * z * p1 - (j - 1.0L) * p0) / (long double)j;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = z; p0 = 1.0L; }
            /* P_n = p1, P_{n-1} = p0 */
            pp = (long double)n * (z * p1 - p0) / (z * z - 1.0L);
            long double dz = p1 / pp;
            z -= dz;
            if (fabsl(dz) < 1e-19L) break;
        }
        /* recompute derivative at the converged root */
        {
            long double p0 = 1.0L, p1 = z;
            for (int j = 2; j <= n; ++j) {
                long double p2 = ((2.0L * j - 1.0L) * z * p1 - (j - 1.0L) * p0) / (long double)j;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = z; p0 = 1.0L; }
            pp = (long double)n * (z * p1 - p0) / (z * z - 1.0L);
        }
        long double wt = 2.0L / ((1.0L - z * z) * pp * pp);
        /* roots come out descending; store ascending non-negative */
        int slot = m - 1 - k;
        if ((n & 1) && k == m - 1) z = 0.0L;
        x[slot] = (double)z;
        w[slot] = (double)wt;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Coarse-graining kernels (CGFuncOctic.h:16-29, CGFuncRect.h:17-28,    */
/* CGFuncMollifier.h:11-31).                                           */
/* ------------------------------------------------------------------ */
typedef struct {
    int kind;
    double rcut;
} cgfunc_t;

double oracle_cg_delta(int kind, double rcut, double dr_sq)
{
    double rcutsq = rcut * rcut;
    if (!(dr_sq < rcutsq)) return 0.0;
    if (kind == CG_OCTIC) {
        double r4 = dr_sq * dr_sq / (rcutsq * rcutsq);
        double r8 = r4 * r4;
        return 15.0 / (8 * M_PI * rcutsq) * (1 - 2 * r4 + r8);
    } else if (kind == CG_RECT) {
        return 1 / (M_PI * rcutsq);
    } else {
        double norm = 13.46842098743079272180855772032086697297031125109170011102 / (2.0 * M_PI * rcut * rcut);
        double r2 = dr_sq / rcutsq;
        return norm * exp(-1.0 / (1 - r2));
    }
}

/* integrand of the bond integral: Delta(|x - (ri + s rij)|^2)
 * (CoarseGrainFunction.h:42-45 -> :78-83) */
static double objfunc(const cgfunc_t *cg, double s, const double *x, const double *ri, const double *rij)
{
    double dr[3], drsq = 0.0;
    for (int a = 0; a < 3; ++a) {
        dr[a] = x[a] - (ri[a] + s * rij[a]);
        drsq += dr[a] * dr[a];
    }
    return oracle_cg_delta(cg->kind, cg->rcut, drsq);
}

/* Fixed-order Gauss-Legendre on [0,1] in symmetric node pairs.  The table
 * (x,w,n) is what the reference allocates per call (FixedPointQuadrature.h:35). */
typedef struct {
    int n;
    double *x;
    double *w;
} gltable_t;

static double gl_fixed(const cgfunc_t *cg, const gltable_t *t, const double *x, const double *ri, const double *rij)
{
    const double A = 0.5, B = 0.5;
    int m = (t->n + 1) >> 1;
    double sum;
    if (t->n & 1) {
        sum = t->w[0] * objfunc(cg, B, x, ri, rij);
        for (int i = 1; i < m; ++i) {
            double Ax = A * t->x[i];
            sum += t->w[i] * (objfunc(cg, B + Ax, x, ri, rij) + objfunc(cg, B - Ax, x, ri, rij));
        }
    } else {
        sum = 0.0;
        for (int i = 0; i < m; ++i) {
            double Ax = A * t->x[i];
            sum += t->w[i] * (objfunc(cg, B + Ax, x, ri, rij) + objfunc(cg, B - Ax, x, ri, rij));
        }
    }
    return A * sum;
}

/* ------------------------------------------------------------------ */
/* Adaptive QAG with the 15-point Gauss-Kronrod rule (QUADPACK QAG,      */
/* key=1), restated from the published algorithm: bisect the interval    */
/* with the largest error estimate until the total error estimate drops   */
/* below max(epsabs, epsrel*|result|) or `limit` intervals are in use.    */
/* ------------------------------------------------------------------ */
static const double gk15_xgk[8] = {
    0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
    0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
    0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
    0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
static const double gk15_wgk[8] = {
    0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
    0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
    0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
    0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double gk15_wg[4] = {
    0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
    0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

static void gk15(const cgfunc_t *cg, const double *x, const double *ri, const double *rij,
                 double a, double b, double *result, double *abserr, double *resabs, double *resasc)
{
    double center = 0.5 * (a + b), half = 0.5 * (b - a), abs_half = fabs(half);
    double fc = objfunc(cg, center, x, ri, rij);
    double resg = fc * gk15_wg[3];
    double resk = fc * gk15_wgk[7];
    double ra = fabs(resk);
    double fv1[7], fv2[7];
    for (int j = 0; j < 3; ++j) {
        int jtw = 2 * j + 1;
        double absc = half * gk15_xgk[jtw];
        double f1 = objfunc(cg, center - absc, x, ri, rij), f2 = objfunc(cg, center + absc, x, ri, rij);
        fv1[jtw] = f1; fv2[jtw] = f2;
        resg += gk15_wg[j] * (f1 + f2);
        resk += gk15_wgk[jtw] * (f1 + f2);
        ra += gk15_wgk[jtw] * (fabs(f1) + fabs(f2));
    }
    for (int j = 0; j < 4; ++j) {
        int jtwm1 = 2 * j;
        double absc = half * gk15_xgk[jtwm1];
        double f1 = objfunc(cg, center - absc, x, ri, rij), f2 = objfunc(cg, center + absc, x, ri, rij);
        fv1[jtwm1] = f1; fv2[jtwm1] = f2;
        resk += gk15_wgk[jtwm1] * (f1 + f2);
        ra += gk15_wgk[jtwm1] * (fabs(f1) + fabs(f2));
    }
    double mean = resk * 0.5;
    double rasc = gk15_wgk[7] * fabs(fc - mean);
    for (int j = 0; j < 7; ++j) rasc += gk15_wgk[j] * (fabs(fv1[j] - mean) + fabs(fv2[j] - mean));
    double err = fabs((resk - resg) * half);
    rasc *= abs_half;
    ra *= abs_half;
    if (rasc != 0.0 && err != 0.0) {
        double scale = pow(200.0 * err / rasc, 1.5);
        err = (scale < 1.0) ? rasc * scale : rasc;
    }
    if (ra > 2.2250738585072014e-308 / (50 * 2.2204460492503131e-16)) {
        double min_err = 50 * 2.2204460492503131e-16 * ra;
        if (min_err > err) err = min_err;
    }
    *result = resk * half;
    *abserr = err;
    *resabs = ra;
    *resasc = rasc;
}

double oracle_qag(int kind, double rcut, const double *x, const double *ri, const double *rij,
                  double epsabs, double epsrel, int limit)
{
    cgfunc_t cg = {kind, rcut};
    if (limit < 1) limit = 1;
    double *alist = (double *)malloc(sizeof(double) * 4 * (size_t)limit);
    double *blist = alist + limit, *rlist = blist + limit, *elist = rlist + limit;
    double result, abserr, resabs, resasc;
    gk15(&cg, x, ri, rij, 0.0, 1.0, &result, &abserr, &resabs, &resasc);
    alist[0] = 0.0; blist[0] = 1.0; rlist[0] = result; elist[0] = abserr;
    int n = 1;
    double tol = fmax(epsabs, epsrel * fabs(result));
    double round_off = 50 * 2.2204460492503131e-16 * resabs;
    if ((abserr <= round_off && abserr > tol) || abserr <= tol || limit == 1) {
        free(alist);
        return result;
    }
    double area = result, errsum = abserr;
    while (n < limit) {
        /* interval with the largest error */
        int imax = 0;
        for (int k = 1; k < n; ++k)
            if (elist[k] > elist[imax]) imax = k;
        double a1 = alist[imax], b2 = blist[imax];
        double a2 = 0.5 * (a1 + b2), b1 = a2;
        double r1, e1, ra1, rs1, r2, e2, ra2, rs2;
        gk15(&cg, x, ri, rij, a1, b1, &r1, &e1, &ra1, &rs1);
        gk15(&cg, x, ri, rij, a2, b2, &r2, &e2, &ra2, &rs2);
        area += r1 + r2 - rlist[imax];
        errsum += e1 + e2 - elist[imax];
        alist[imax] = a1; blist[imax] = b1; rlist[imax] = r1; elist[imax] = e1;
        alist[n] = a2; blist[n] = b2; rlist[n] = r2; elist[n] = e2;
        ++n;
        tol = fmax(epsabs, epsrel * fabs(area));
        if (errsum <= tol) break;
    }
    /* final result: sum of the interval results */
    double sum = 0.0;
    for (int k = 0; k < n; ++k) sum += rlist[k];
    free(alist);
    return sum;
}

/* ------------------------------------------------------------------ */
/* Pair forces F/r of the polydisperse m-n family (PairPotential.h:98-104)*/
/* params: {v0, eps, c0, c1, c2[, c3]} as built in potential/__init__.py */
/* ------------------------------------------------------------------ */
double oracle_force_divr(int kind, const double *p, double scaled_rcut, double rsq, double di, double dj)
{
    double rcutsq = scaled_rcut * scaled_rcut;
    double v0 = p[0], eps = p[1], c1 = p[3], c2 = p[4];
    double sigma = 0.5 * (di + dj) * (1 - eps * fabs(di - dj));
    double actualrcutsq = rcutsq * sigma * sigma;
    if (!(rsq < actualrcutsq)) return 0.0;
    double r2inv = 1.0 / rsq;
    r2inv *= sigma * sigma;
    double _rsq = rsq / (sigma * sigma);
    switch (kind) {
    case POT_POLY12: { /* Polydisperse12.h:38-54 */
        double r6inv = r2inv * r2inv * r2inv;
        return (12.0 * v0 * r2inv * r6inv * r6inv - 2.0 * c1 - 4.0 * c2 * _rsq) / (sigma * sigma);
    }
    case POT_POLY18: { /* Polydisperse18.h:37-53 */
        double r6inv = r2inv * r2inv * r2inv;
        return (18.0 * v0 * r2inv * r6inv * r6inv * r6inv - 2.0 * c1 - 4.0 * c2 * _rsq) / (sigma * sigma);
    }
    case POT_POLY10: { /* Polydisperse10.h:37-52 (six parameters) */
        double c3 = p[5];
        double r10inv = r2inv * r2inv * r2inv * r2inv * r2inv;
        return (10.0 * v0 * r2inv * r10inv - 2.0 * c1 - 4.0 * c2 * _rsq - 6.0 * c3 * _rsq * _rsq) / (sigma * sigma);
    }
    case POT_POLY106: { /* Polydisperse106.h:37-53 */
        double r10inv = r2inv * r2inv * r2inv * r2inv * r2inv;
        return (v0 * (10.0 * r10inv * r2inv - 6.0 * r2inv * r2inv * r2inv * r2inv) - 2.0 * c1 - 4.0 * c2 * _rsq) /
               (sigma * sigma);
    }
    case POT_POLYLJ: { /* PolydisperseLJ.h:37-53 */
        double r6inv = r2inv * r2inv * r2inv;
        return (12.0 * v0 * r2inv * r6inv * r6inv - 6.0 * v0 * r2inv * r6inv - 2.0 * c1 - 4.0 * c2 * _rsq) /
               (sigma * sigma);
    }
    }
    return 0.0;
}

/* pair energies (used by the tests to pin F/r = -phi'/r) */
double oracle_energy(int kind, const double *p, double scaled_rcut, double rsq, double di, double dj)
{
    double rcutsq = scaled_rcut * scaled_rcut;
    double v0 = p[0], eps = p[1], c0 = p[2], c1 = p[3], c2 = p[4];
    double sigma = 0.5 * (di + dj) * (1 - eps * fabs(di - dj));
    if (!(rsq < rcutsq * sigma * sigma)) return 0.0;
    double r2inv = sigma * sigma / rsq;
    double _rsq = rsq / (sigma * sigma);
    switch (kind) {
    case POT_POLY12: { double r6 = r2inv * r2inv * r2inv; return v0 * r6 * r6 + c0 + c1 * _rsq + c2 * _rsq * _rsq; }
    case POT_POLY18: { double r6 = r2inv * r2inv * r2inv; return v0 * r6 * r6 * r6 + c0 + c1 * _rsq + c2 * _rsq * _rsq; }
    case POT_POLY10: {
        double r10 = r2inv * r2inv * r2inv * r2inv * r2inv;
        return v0 * r10 + c0 + c1 * _rsq + c2 * _rsq * _rsq + p[5] * _rsq * _rsq * _rsq;
    }
    case POT_POLY106: {
        double r10 = r2inv * r2inv * r2inv * r2inv * r2inv;
        return v0 * (r10 - r2inv * r2inv * r2inv) + c0 + c1 * _rsq + c2 * _rsq * _rsq;
    }
    case POT_POLYLJ: { double r6 = r2inv * r2inv * r2inv; return v0 * r6 * r6 - v0 * r6 + c0 + c1 * _rsq + c2 * _rsq * _rsq; }
    }
    return 0.0;
}

/* ------------------------------------------------------------------ */
/* Scalar entry points used by the tests                                */
/* ------------------------------------------------------------------ */
double oracle_bond_fixed(int kind, double rcut, int order, const double *x, const double *ri, const double *rij)
{
    double xs[2048], ws[2048];
    if (oracle_gl_nodes(order, xs, ws) != 0) return NAN;
    cgfunc_t cg = {kind, rcut};
    gltable_t t = {order, xs, ws};
    return gl_fixed(&cg, &t, x, ri, rij);
}

/* ------------------------------------------------------------------ */
/* Periodic geometry                                                    */
/* ------------------------------------------------------------------ */
typedef struct {
    double L[3];
    double lo[3];
    int periodic[3];
} box_t;

static inline double minimg1(double d, double L, int periodic)
{
    if (periodic) d -= L * nearbyint(d / L);
    return d;
}

/* linked-cell list over wrapped positions */
typedef struct {
    int nc[3];
    double cs[3];
    int *head;
    int *next;
} cells_t;

static int cell_coord(const box_t *b, const cells_t *c, int a, double p)
{
    if (c->nc[a] == 1) return 0;
    double u = p - b->lo[a];
    if (b->periodic[a]) u -= b->L[a] * floor(u / b->L[a]);
    int k = (int)floor(u / c->cs[a]);
    if (k < 0) k = 0;
    if (k >= c->nc[a]) k = c->nc[a] - 1;
    return k;
}

static void cells_build(const box_t *b, cells_t *c, double rmax, int N, const double *pos)
{
    for (int a = 0; a < 3; ++a) {
        int n = (int)floor(b->L[a] / rmax);
        if (!b->periodic[a] || n < 3) n = 1;
        c->nc[a] = n;
        c->cs[a] = b->L[a] / n;
    }
    int ncell = c->nc[0] * c->nc[1] * c->nc[2];
    c->head = (int *)malloc(sizeof(int) * (size_t)ncell);
    c->next = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    for (int k = 0; k < ncell; ++k) c->head[k] = -1;
    for (int i = N - 1; i >= 0; --i) {
        int cx = cell_coord(b, c, 0, pos[3 * i]), cy = cell_coord(b, c, 1, pos[3 * i + 1]),
            cz = cell_coord(b, c, 2, pos[3 * i + 2]);
        int k = (cz * c->nc[1] + cy) * c->nc[0] + cx;
        c->next[i] = c->head[k];
        c->head[k] = i;
    }
}

/* enumerate candidate particles near point p: calls visit(j) for each particle
 * in the (up to) 3x3x3 block of cells around p */
#define MAXCAND_CELLS 27
static int cells_near(const box_t *b, const cells_t *c, const double *p, int *cell_ids)
{
    int base[3], cnt[3];
    for (int a = 0; a < 3; ++a) {
        int k = cell_coord(b, c, a, p[a]);
        if (c->nc[a] >= 3) { base[a] = k - 1; cnt[a] = 3; }
        else { base[a] = 0; cnt[a] = c->nc[a]; }
    }
    int n = 0;
    for (int dz = 0; dz < cnt[2]; ++dz)
        for (int dy = 0; dy < cnt[1]; ++dy)
            for (int dx = 0; dx < cnt[0]; ++dx) {
                int kx = (base[0] + dx + c->nc[0]) % c->nc[0];
                int ky = (base[1] + dy + c->nc[1]) % c->nc[1];
                int kz = (base[2] + dz + c->nc[2]) % c->nc[2];
                cell_ids[n++] = (kz * c->nc[1] + ky) * c->nc[0] + kx;
            }
    return n;
}

/* ------------------------------------------------------------------ */
/* The field computation                                                */
/* ------------------------------------------------------------------ */
typedef struct {
    int dim;
    double L[3];
    int periodic[3];
    int N;
    const double *pos;   /* N x 3 */
    const double *diam;  /* N */
    const double *mass;  /* N */
    const double *vel;   /* N x 3 */
    const double *disp;  /* N x 3 or NULL */
    int G;
    const double *grid;  /* G x 3 */
    int pot_kind;
    double params[6];
    double scaled_rcut;
    double max_rcut;
    int cg_kind;
    double cg_rcut;
    int quad_mode;
    int quad_order;
    int adapt_limit;
    double adapt_epsabs, adapt_epsrel;
    int mode;   /* 0 faithful, 1 clean */
    int brute;  /* 1: O(G N) / O(N) brute-force enumeration */
    int nthreads;
} oracle_args_t;

typedef struct {
    double *virial;   /* G x D x D (flat a + D b), may be NULL */
    double *kinetic;  /* G x D x D, may be NULL */
    double *elastic;  /* G x D x D, may be NULL */
    double *virial_abs; /* G x D x D: sum of |contribution|, may be NULL */
    int64_t n_ball;    /* (x,i) pairs */
    int64_t n_triples; /* (x,i,j) triples enumerated */
    int64_t n_nonzero; /* triples with nonzero pair force (U) */
} oracle_out_t;

typedef void (*pair_obs_fn)(const oracle_args_t *A, int i, int j, const double *rij, double bondval, double *val, int D);

/* VirialStress::compute_cg (ListofCommonObservables.h:72-82): a temporary
 * tensor per triple, then val += 0.5*bondval*temp */
static void virial_obs(const oracle_args_t *A, int i, int j, const double *rij, double bondval, double *val, int D)
{
    double rsq = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
    double forceval = oracle_force_divr(A->pot_kind, A->params, A->scaled_rcut, rsq, A->diam[i], A->diam[j]);
    double *tmp = (double *)calloc((size_t)(D * D), sizeof(double));
    double F[3] = {forceval * rij[0], forceval * rij[1], forceval * rij[2]};
    for (int b = 0; b < D; ++b)
        for (int a = 0; a < D; ++a) tmp[a + D * b] += F[a] * rij[b];
    for (int k = 0; k < D * D; ++k) val[k] += 0.5 * bondval * tmp[k];
    free(tmp);
}

/* ElasticVirialStress::compute (ListofCommonObservables.h:86-123) under the
 * VirialStress::compute_cg wrapper */
static void elastic_obs(const oracle_args_t *A, int i, int j, const double *rij, double bondval, double *val, int D)
{
    double u[3] = {0, 0, 0};
    if (A->disp)
        for (int a = 0; a < 3; ++a) u[a] = A->disp[3 * j + a] - A->disp[3 * i + a];
    double ur[3] = {u[0] + rij[0], u[1] + rij[1], u[2] + rij[2]};
    double rsq = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
    double ursq = ur[0] * ur[0] + ur[1] * ur[1] + ur[2] * ur[2];
    double f0 = oracle_force_divr(A->pot_kind, A->params, A->scaled_rcut, rsq, A->diam[i], A->diam[j]);
    double f1 = oracle_force_divr(A->pot_kind, A->params, A->scaled_rcut, ursq, A->diam[i], A->diam[j]);
    double *tmp = (double *)calloc((size_t)(D * D), sizeof(double));
    for (int b = 0; b < D; ++b)
        for (int a = 0; a < D; ++a) tmp[a + D * b] += f1 * ur[a] * ur[b] - f0 * rij[a] * rij[b];
    for (int k = 0; k < D * D; ++k) val[k] += 0.5 * bondval * tmp[k];
    free(tmp);
}

static double bond_value(const oracle_args_t *A, const double *x, const double *ri, const double *rij,
                         const double *glx, const double *glw)
{
    cgfunc_t cg = {A->cg_kind, A->cg_rcut};
    if (A->quad_mode == QUAD_ADAPTIVE)
        return oracle_qag(A->cg_kind, A->cg_rcut, x, ri, rij, A->adapt_epsabs, A->adapt_epsrel, A->adapt_limit);
    if (A->mode == 0) {
        /* per-call table, as gsl_integration_glfixed_table_alloc does */
        int m = (A->quad_order + 1) / 2;
        gltable_t *t = (gltable_t *)malloc(sizeof(gltable_t));
        t->n = A->quad_order;
        t->x = (double *)malloc(sizeof(double) * (size_t)m);
        t->w = (double *)malloc(sizeof(double) * (size_t)m);
        memcpy(t->x, glx, sizeof(double) * (size_t)m);
        memcpy(t->w, glw, sizeof(double) * (size_t)m);
        double v = gl_fixed(&cg, t, x, ri, rij);
        free(t->x);
        free(t->w);
        free(t);
        return v;
    }
    gltable_t t = {A->quad_order, (double *)glx, (double *)glw};
    return gl_fixed(&cg, &t, x, ri, rij);
}

int oracle_ik_compute(const oracle_args_t *A, oracle_out_t *O)
{
    const int D = A->dim;
    box_t box;
    for (int a = 0; a < 3; ++a) {
        box.L[a] = A->L[a];
        box.lo[a] = -0.5 * A->L[a];
        box.periodic[a] = A->periodic[a];
    }
    double rmax = fmax(A->cg_rcut, A->max_rcut);
    for (int a = 0; a < 3; ++a)
        if (box.periodic[a] && !(rmax < 0.5 * box.L[a])) return -2; /* images would not be unique */

    double glx[2048], glw[2048];
    if (A->quad_mode == QUAD_FIXED && oracle_gl_nodes(A->quad_order, glx, glw) != 0) return -3;

    cells_t cells = {{1, 1, 1}, {0, 0, 0}, NULL, NULL};
    if (!A->brute) cells_build(&box, &cells, rmax, A->N, A->pos);

    /* observables in the order the reference's std::map iterates them
     * (alphabetical by name: elasticvirialstress, kineticstress, virialstress) */
    pair_obs_fn pair_fns[2];
    double *pair_out[2];
    int npair = 0;
    if (O->elastic) { pair_fns[npair] = elastic_obs; pair_out[npair++] = O->elastic; }
    if (O->virial) { pair_fns[npair] = virial_obs; pair_out[npair++] = O->virial; }

    const double rc2 = A->cg_rcut * A->cg_rcut;
    const double mr2 = A->max_rcut * A->max_rcut;
    int64_t n_ball = 0, n_triples = 0, n_nonzero = 0;

#ifdef _OPENMP
    if (A->nthreads > 0) omp_set_num_threads(A->nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : n_ball, n_triples, n_nonzero)
    for (int g = 0; g < A->G; ++g) {
        const double *x = A->grid + 3 * (size_t)g;
        /* clearState (IrvingKirkwood.h:36-40) */
        for (int k = 0; k < npair; ++k) memset(pair_out[k] + (size_t)g * D * D, 0, sizeof(double) * D * D);
        if (O->kinetic) memset(O->kinetic + (size_t)g * D * D, 0, sizeof(double) * D * D);
        if (O->virial_abs) memset(O->virial_abs + (size_t)g * D * D, 0, sizeof(double) * D * D);

        int cids_x[MAXCAND_CELLS];
        int ncx = A->brute ? 1 : cells_near(&box, &cells, x, cids_x);
        for (int cxi = 0; cxi < ncx; ++cxi) {
            int i = A->brute ? 0 : cells.head[cids_x[cxi]];
            for (; i >= 0 && i < A->N; i = A->brute ? i + 1 : cells.next[i]) {
                /* particle i's image nearest x: dr = x - ri */
                double dr[3], ri[3], drsq = 0.0;
                for (int a = 0; a < 3; ++a) {
                    dr[a] = minimg1(x[a] - A->pos[3 * i + a], box.L[a], box.periodic[a]);
                    drsq += dr[a] * dr[a];
                }
                if (!(drsq < rc2)) continue;
                for (int a = 0; a < 3; ++a) ri[a] = x[a] - dr[a];
                ++n_ball;
                /* local observable: kinetic, intended formula m v (x) v Delta */
                if (O->kinetic) {
                    double cgval = oracle_cg_delta(A->cg_kind, A->cg_rcut, drsq);
                    double *kv = O->kinetic + (size_t)g * D * D;
                    for (int b = 0; b < D; ++b)
                        for (int a = 0; a < D; ++a)
                            kv[a + D * b] += A->mass[i] * A->vel[3 * i + a] * A->vel[3 * i + b] * cgval;
                }
                if (npair == 0 && !O->virial_abs) continue;
                /* j within max_rcut of particle i (IrvingKirkwood.h:93-98) */
                int cids_i[MAXCAND_CELLS];
                int nci = A->brute ? 1 : cells_near(&box, &cells, A->pos + 3 * (size_t)i, cids_i);
                for (int cii = 0; cii < nci; ++cii) {
                    int j = A->brute ? 0 : cells.head[cids_i[cii]];
                    for (; j >= 0 && j < A->N; j = A->brute ? j + 1 : cells.next[j]) {
                        if (j == i) continue;
                        double rij[3], rsq = 0.0;
                        for (int a = 0; a < 3; ++a) {
                            rij[a] = minimg1(A->pos[3 * j + a] - A->pos[3 * i + a], box.L[a], box.periodic[a]);
                            rsq += rij[a] * rij[a];
                        }
                        if (!(rsq < mr2)) continue;
                        ++n_triples;
                        double f = oracle_force_divr(A->pot_kind, A->params, A->scaled_rcut, rsq, A->diam[i], A->diam[j]);
                        if (f != 0.0) ++n_nonzero;
                        int has_elastic = (O->elastic != NULL);
                        if (A->mode == 1 && f == 0.0 && !has_elastic) continue;
                        double bondval = bond_value(A, x, ri, rij, glx, glw);
                        for (int k = 0; k < npair; ++k)
                            pair_fns[k](A, i, j, rij, bondval, pair_out[k] + (size_t)g * D * D, D);
                        if (O->virial_abs) {
                            double *sv = O->virial_abs + (size_t)g * D * D;
                            for (int b = 0; b < D; ++b)
                                for (int a = 0; a < D; ++a) sv[a + D * b] += fabs(0.5 * bondval * f * rij[a] * rij[b]);
                        }
                    }
                }
            }
        }
    }
    O->n_ball = n_ball;
    O->n_triples = n_triples;
    O->n_nonzero = n_nonzero;
    free(cells.head);
    free(cells.next);
    return 0;
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
