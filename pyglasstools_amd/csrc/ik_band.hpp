// ik_band.hpp -- thresholds of the FP32 distance pre-test of the 3D candidate
// scans (host code, no HIP).  Shared by the engine's geometry set-up and the
// C entry point pgt_fp32_band that lets the host tests check them.
#pragma once
#include <cmath>

namespace pgt {

// FP32 pre-test bounds for a search radius r.  Coordinates (box and grid) are
// at most M in magnitude and every searched axis is periodic (ok), so an FP32
// component of a separation (difference, minimum image) is within
// E = (4M + 2 Lmax + 4r) 2^-23 of the FP64 one (twice the rounding bound);
// near |a| = r the FP32 |a|^2 is then within D E (2.02 r + 2E) +
// 3 2^-24 |a|^2 of the FP64 value.  The band is twice that, and the
// thresholds are rounded outwards: |a|^2 < in32 in FP32 means inside in
// FP64, |a|^2 > out32 outside.  Without a bound (not ok, a band reaching
// zero, or a radius not well below half of the shortest box length Lmin, so
// that the FP32 and FP64 minimum images could differ near r) in32 = -1 and
// out32 = inf send every candidate to the FP64 test.
inline void fp32_band(double r, double M, double Lmax, double Lmin, int D, bool ok, float &in32, float &out32)
{
    in32 = -1.0f;
    out32 = INFINITY;
    const double r2 = r * r;
    const double E = (4 * M + 2 * Lmax + 4 * r) * std::ldexp(1.0, -23);
    const double B = 2 * (D * E * (2.02 * r + 2 * E) + 3 * std::ldexp(1.0, -24) * 1.03 * r2);
    if (!(ok && r2 - B > 0 && 2 * r + 8 * E < Lmin)) return;
    float lo32 = (float)(r2 - B), hi32 = (float)(r2 + B);
    if ((double)lo32 > r2 - B) lo32 = std::nextafter(lo32, -INFINITY);
    if ((double)hi32 < r2 + B) hi32 = std::nextafter(hi32, INFINITY);
    in32 = lo32;
    out32 = hi32;
}

}  // namespace pgt
