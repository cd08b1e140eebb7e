// ik_device.hpp -- device-side physics of the IK stress-field engine:
// polydisperse pair forces, coarse-graining kernels and the fixed-order
// bond line integral.  Included by ik_engine.hip only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgt {

constexpr double kPi = 3.14159265358979323846;

// Pair-potential parameters, as PairPotential stores them (params vector +
// scaled_rcut), flattened for kernel arguments.
struct PotParams {
    int kind;
    double v0, eps, c0, c1, c2, c3;
    double rcutsq;  // scaled_rcut^2
};

// F/r of the polydisperse m-n family with its even-polynomial smoothing.
// Arithmetic in the order of pyglasstools/potential/Polydisperse12.h:38-54
// (and Polydisperse18.h:37-53, Polydisperse10.h:37-52,
// Polydisperse106.h:37-53, PolydisperseLJ.h:37-53) so the FP64 result
// rounds like the reference's.  R is double (drop-in) or float (FP32 form).
template <typename R>
__host__ __device__ inline R force_divr(const PotParams &P, R rsq, R di, R dj)
{
    const R sigma = R(0.5) * (di + dj) * (R(1) - R(P.eps) * fabs(di - dj));
    const R sig2 = sigma * sigma;
    if (!(rsq < R(P.rcutsq) * sig2)) return R(0);
    R r2inv = R(1) / rsq;
    r2inv *= sig2;
    const R xsq = rsq / sig2;
    const R v0 = R(P.v0), c1 = R(P.c1), c2 = R(P.c2);
    switch (P.kind) {
    case 0: {
        const R r6 = r2inv * r2inv * r2inv;
        return (R(12) * v0 * r2inv * r6 * r6 - R(2) * c1 - R(4) * c2 * xsq) / sig2;
    }
    case 1: {
        const R r6 = r2inv * r2inv * r2inv;
        return (R(18) * v0 * r2inv * r6 * r6 * r6 - R(2) * c1 - R(4) * c2 * xsq) / sig2;
    }
    case 2: {
        const R r10 = r2inv * r2inv * r2inv * r2inv * r2inv;
        return (R(10) * v0 * r2inv * r10 - R(2) * c1 - R(4) * c2 * xsq - R(6) * R(P.c3) * xsq * xsq) / sig2;
    }
    case 3: {
        const R r10 = r2inv * r2inv * r2inv * r2inv * r2inv;
        return (v0 * (R(10) * r10 * r2inv - R(6) * r2inv * r2inv * r2inv * r2inv) - R(2) * c1 - R(4) * c2 * xsq) / sig2;
    }
    default: {
        const R r6 = r2inv * r2inv * r2inv;
        return (R(12) * v0 * r2inv * r6 * r6 - R(6) * v0 * r2inv * r6 - R(2) * c1 - R(4) * c2 * xsq) / sig2;
    }
    }
}

template <typename R>
__host__ __device__ inline R pair_energy(const PotParams &P, R rsq, R di, R dj)
{
    const R sigma = R(0.5) * (di + dj) * (R(1) - R(P.eps) * fabs(di - dj));
    const R sig2 = sigma * sigma;
    if (!(rsq < R(P.rcutsq) * sig2)) return R(0);
    R r2inv = R(1) / rsq;
    r2inv *= sig2;
    const R xsq = rsq / sig2;
    const R v0 = R(P.v0), c0 = R(P.c0), c1 = R(P.c1), c2 = R(P.c2);
    switch (P.kind) {
    case 0: { const R r6 = r2inv * r2inv * r2inv; return v0 * r6 * r6 + c0 + c1 * xsq + c2 * xsq * xsq; }
    case 1: { const R r6 = r2inv * r2inv * r2inv; return v0 * r6 * r6 * r6 + c0 + c1 * xsq + c2 * xsq * xsq; }
    case 2: {
        const R r10 = r2inv * r2inv * r2inv * r2inv * r2inv;
        return v0 * r10 + c0 + c1 * xsq + c2 * xsq * xsq + R(P.c3) * xsq * xsq * xsq;
    }
    case 3: {
        const R r10 = r2inv * r2inv * r2inv * r2inv * r2inv;
        return v0 * (r10 - r2inv * r2inv * r2inv) + c0 + c1 * xsq + c2 * xsq * xsq;
    }
    default: { const R r6 = r2inv * r2inv * r2inv; return v0 * r6 * r6 - v0 * r6 + c0 + c1 * xsq + c2 * xsq * xsq; }
    }
}

// Radial coarse-graining kernel as a function of the scaled squared distance
// t = |dr|^2 / rc^2, without its normalisation constant (which the callers
// fold into the tensors they accumulate):
//   Octic      (1 - t^2)^2              CGFuncOctic.h:16-29 (norm 15/(8 pi rc^2))
//   Rect       1                        CGFuncRect.h:17-28  (norm 1/(pi rc^2))
//   Mollifier  exp(-1/(1 - t))          CGFuncMollifier.h:11-31
// and zero for t >= 1.  For the Octic the clamp max(1 - t^2, 0) gives the
// support test for free: (1 - t^2) <= 0 exactly when t >= 1.
template <int CG, typename R>
__device__ __forceinline__ R shape(R t)
{
    if constexpr (CG == 0) {
        R u = fmax(fma(-t, t, R(1)), R(0));
        return u * u;
    } else if constexpr (CG == 1) {
        return t < R(1) ? R(1) : R(0);
    } else {
        return t < R(1) ? exp(R(-1) / (R(1) - t)) : R(0);
    }
}

__host__ __device__ inline double cg_norm(int cg, double rc)
{
    const double rcsq = rc * rc;
    if (cg == 0) return 15.0 / (8 * kPi * rcsq);
    if (cg == 1) return 1 / (kPi * rcsq);
    return 13.46842098743079272180855772032086697297031125109170011102 / (2.0 * kPi * rc * rc);
}

// The kernel exactly as the reference writes it (unscaled dr^2), used by the
// scalar evaluators that back getDeltaFunc/getObjFunc.
__host__ __device__ inline double cg_delta_ref(int cg, double rc, double dr_sq)
{
    const double rcutsq = rc * rc;
    if (!(dr_sq < rcutsq)) return 0.0;
    if (cg == 0) {
        const double r4 = dr_sq * dr_sq / (rcutsq * rcutsq);
        const double r8 = r4 * r4;
        return 15.0 / (8 * kPi * rcutsq) * (1 - 2 * r4 + r8);
    }
    if (cg == 1) return 1 / (kPi * rcutsq);
    const double r2 = dr_sq / rcutsq;
    return cg_norm(2, rc) * exp(-1.0 / (1 - r2));
}

// Gauss-Legendre rule on [0,1], unfolded: node k at s[k] with weight w[k]
// (the [0,1] half-length 1/2 is already in w).  MAXQ bounds the kernel-arg
// copy; longer rules are read from global memory.
constexpr int kMaxQArg = 64;
struct GLRule {
    int n;
    double s[kMaxQArg];
    double w[kMaxQArg];
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

// Adaptive quadrature (AdaptiveCGFunc::getBondFunc, CoarseGrainFunction.h:
// 124-127 -> gsl_quad::integrate, Quadrature.h:39-51: gsl_integration_qag
// with key 1, the 15-point Gauss-Kronrod rule).  The rule's abscissae and
// weights are the published QUADPACK constants.  QAG: apply the rule on
// [0, 1]; stop if the error estimate meets max(epsabs, epsrel |result|) (or
// is at the rounding floor); otherwise bisect the interval with the largest
// error estimate until the summed error meets the tolerance or `limit`
// intervals exist, and return the sum of the interval results in storage
// order.  The interval table lives in a workspace W: QagPrivate keeps up to
// kQagMax intervals in the lane's private memory, QagGlobal the caller's
// limit in global memory; a call that would need more than its workspace
// holds (limit larger and not converged) sets `capped`.
constexpr int kQagMax = 128;

// Interval table of kQagMax entries in private memory.
struct QagPrivate {
    double v[4][kQagMax];
    __device__ int size() const { return kQagMax; }
    __device__ double &a(int k) { return v[0][k]; }
    __device__ double &b(int k) { return v[1][k]; }
    __device__ double &r(int k) { return v[2][k]; }
    __device__ double &e(int k) { return v[3][k]; }
};

// Interval table of n entries in global memory, lane-interleaved: entry k's
// four values at p[(4k + q) * stride], p the lane's first slot.
struct QagGlobal {
    double *p;
    size_t stride;
    int n;
    __device__ int size() const { return n; }
    __device__ double &a(int k) { return p[(size_t)(4 * k) * stride]; }
    __device__ double &b(int k) { return p[(size_t)(4 * k + 1) * stride]; }
    __device__ double &r(int k) { return p[(size_t)(4 * k + 2) * stride]; }
    __device__ double &e(int k) { return p[(size_t)(4 * k + 3) * stride]; }
};

struct AdaptQ {
    int limit;
    double epsabs, epsrel;
};

__device__ __forceinline__ double gk15_x(int k)
{
    constexpr double v[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                             0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                             0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                             0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
    return v[k];
}
__device__ __forceinline__ double gk15_wk(int k)
{
    constexpr double v[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                             0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                             0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                             0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
    return v[k];
}
__device__ __forceinline__ double gk15_wg(int k)
{
    constexpr double v[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                             0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
    return v[k];
}

// The 15-point Gauss-Kronrod rule on [a, b] with QUADPACK's error estimate.
template <typename F>
__device__ void gk15_dev(const F &f, double a, double b, double &result, double &abserr, double &resabs,
                         double &resasc)
{
    const double center = 0.5 * (a + b), half = 0.5 * (b - a), abs_half = fabs(half);
    const double fc = f(center);
    double resg = fc * gk15_wg(3);
    double resk = fc * gk15_wk(7);
    double ra = fabs(resk);
    double fv1[7], fv2[7];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int jtw = 2 * j + 1;
        const double absc = half * gk15_x(jtw);
        const double f1 = f(center - absc), f2 = f(center + absc);
        fv1[jtw] = f1;
        fv2[jtw] = f2;
        resg += gk15_wg(j) * (f1 + f2);
        resk += gk15_wk(jtw) * (f1 + f2);
        ra += gk15_wk(jtw) * (fabs(f1) + fabs(f2));
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int jtwm1 = 2 * j;
        const double absc = half * gk15_x(jtwm1);
        const double f1 = f(center - absc), f2 = f(center + absc);
        fv1[jtwm1] = f1;
        fv2[jtwm1] = f2;
        resk += gk15_wk(jtwm1) * (f1 + f2);
        ra += gk15_wk(jtwm1) * (fabs(f1) + fabs(f2));
    }
    const double mean = resk * 0.5;
    double rasc = gk15_wk(7) * fabs(fc - mean);
#pragma unroll
    for (int j = 0; j < 7; ++j) rasc += gk15_wk(j) * (fabs(fv1[j] - mean) + fabs(fv2[j] - mean));
    double err = fabs((resk - resg) * half);
    rasc *= abs_half;
    ra *= abs_half;
    if (rasc != 0.0 && err != 0.0) {
        const double scale = pow(200.0 * err / rasc, 1.5);
        err = (scale < 1.0) ? rasc * scale : rasc;
    }
    if (ra > 2.2250738585072014e-308 / (50 * 2.2204460492503131e-16)) {
        const double min_err = 50 * 2.2204460492503131e-16 * ra;
        if (min_err > err) err = min_err;
    }
    result = resk * half;
    abserr = err;
    resabs = ra;
    resasc = rasc;
}

template <typename F, typename W>
__device__ double qag_dev(const F &f, const AdaptQ &Q, bool &capped, W &w)
{
    const int limit = Q.limit < 1 ? 1 : Q.limit;
    const int lim = limit < w.size() ? limit : w.size();
    double result, abserr, resabs, resasc;
    gk15_dev(f, 0.0, 1.0, result, abserr, resabs, resasc);
    w.a(0) = 0.0;
    w.b(0) = 1.0;
    w.r(0) = result;
    w.e(0) = abserr;
    int n = 1;
    double tol = fmax(Q.epsabs, Q.epsrel * fabs(result));
    const double round_off = 50 * 2.2204460492503131e-16 * resabs;
    if ((abserr <= round_off && abserr > tol) || abserr <= tol || limit == 1) return result;
    double area = result, errsum = abserr;
    while (n < lim) {
        int imax = 0;
        double emax = w.e(0);
        for (int k = 1; k < n; ++k) {
            const double ek = w.e(k);
            if (ek > emax) {
                emax = ek;
                imax = k;
            }
        }
        const double a1 = w.a(imax), b2 = w.b(imax);
        const double a2 = 0.5 * (a1 + b2), b1 = a2;
        double r1, e1, ra1, rs1, r2, e2, ra2, rs2;
        gk15_dev(f, a1, b1, r1, e1, ra1, rs1);
        gk15_dev(f, a2, b2, r2, e2, ra2, rs2);
        area += r1 + r2 - w.r(imax);
        errsum += e1 + e2 - w.e(imax);
        w.a(imax) = a1;
        w.b(imax) = b1;
        w.r(imax) = r1;
        w.e(imax) = e1;
        w.a(n) = a2;
        w.b(n) = b2;
        w.r(n) = r2;
        w.e(n) = e2;
        ++n;
        tol = fmax(Q.epsabs, Q.epsrel * fabs(area));
        if (errsum <= tol) break;
    }
    if (n == lim && lim < limit && errsum > tol) capped = true;
    double sum = 0.0;
    for (int k = 0; k < n; ++k) sum += w.r(k);
    return sum;
}

// The kernel shape (t = |dr|^2 / rc^2, no normalisation) chosen at run time.
__device__ __forceinline__ double shape_rt(int cg, double t)
{
    if (cg == 0) return shape<0, double>(t);
    if (cg == 1) return shape<1, double>(t);
    return shape<2, double>(t);
}

}  // namespace pgt
