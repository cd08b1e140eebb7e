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

}  // namespace pgt
