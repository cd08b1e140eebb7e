// ik_sweep_kernels.hpp -- the cell-list and bond-list launches of a sweep
// (ik_sweep.hpp): each workgroup finds its system, frame and block in the
// sweep's table and runs the single-system kernel's body on that frame's
// arrays; 2D systems (positions with z = 0) or 3D ones.  Included once, by
// ik_engine.hip, after ik_kernels.hpp.
#pragma once
#include <hip/hip_runtime.h>

#include "ik_kernels.hpp"
#include "ik_sweep.hpp"

namespace pgt {

template <bool FIRST>
__global__ __launch_bounds__(kSortBlock) void k_sweep_hist(const SweepSys *__restrict__ T, int ns, int shift,
                                                           int radix, int *__restrict__ keys, int *__restrict__ hist)
{
    const SweepSys &S = T[sweep_system(T, ns, SW_SORT, blockIdx.x)];
    int f, ch;
    if (!sweep_frame_block(S, SW_SORT, blockIdx.x, f, ch)) return;
    const Geom g = S.g;
    const int n = S.n, nchunk = S.nchunk;
    const size_t fo = (size_t)S.p0 + (size_t)f * n;
    sort_hist_body<FIRST>(g, n, ch, nchunk, shift, radix, S.pos + (size_t)f * n * 3, keys + fo,
                          hist + (size_t)S.h0 + (size_t)f * radix * nchunk);
}

// One workgroup per frame of the sweep.
__global__ __launch_bounds__(1024) void k_sweep_scan(const SweepSys *__restrict__ T, int ns, int radix,
                                                     int *__restrict__ hist)
{
    const SweepSys &S = T[sweep_system_of_frame(T, ns, blockIdx.x)];
    const int f = (int)(blockIdx.x - S.f0);
    const int len = radix * S.nchunk;
    scan_body(len, hist + (size_t)S.h0 + (size_t)f * len);
}

template <bool FIRST>
__global__ __launch_bounds__(kSortBlock) void k_sweep_scatter(const SweepSys *__restrict__ T, int ns, int shift,
                                                              int radix, const int *__restrict__ kin,
                                                              const int *__restrict__ vin,
                                                              const int *__restrict__ hist, int *__restrict__ kout,
                                                              int *__restrict__ vout)
{
    const SweepSys &S = T[sweep_system(T, ns, SW_SORT, blockIdx.x)];
    int f, ch;
    if (!sweep_frame_block(S, SW_SORT, blockIdx.x, f, ch)) return;
    const int n = S.n, nchunk = S.nchunk;
    const size_t fo = (size_t)S.p0 + (size_t)f * n;
    sort_scatter_body<FIRST>(n, ch, nchunk, shift, radix, kin + fo, vin + fo,
                             hist + (size_t)S.h0 + (size_t)f * radix * nchunk, kout + fo, vout + fo);
}

// skin null: no kinetic field; rec non-null (3D with bonds): the FP32
// candidate records of the bond kernel and the 3D gather.
template <int D, int OD>
__global__ __launch_bounds__(256) void k_sweep_finish(const SweepSys *__restrict__ T, int ns,
                                                      const int *__restrict__ keys, const int *__restrict__ vals,
                                                      int *__restrict__ cstart, double *__restrict__ spos,
                                                      double *__restrict__ sdiam, double *__restrict__ skin,
                                                      float4 *__restrict__ rec)
{
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int KN = OD * (OD + 1) / 2;
    const SweepSys &S = T[sweep_system(T, ns, SW_FINISH, blockIdx.x)];
    int f, bx;
    if (!sweep_frame_block(S, SW_FINISH, blockIdx.x, f, bx)) return;
    const Geom g = S.g;
    const int n = S.n;
    const size_t fi = (size_t)f * n;              // frame offset in the caller's inputs
    const size_t fo = (size_t)S.p0 + fi;          // and in the work arrays
    sort_finish_body<D, OD>(g, n, bx * 256 + threadIdx.x, g.ncell, keys + fo, vals + fo, S.pos + fi * 3,
                            S.diam + fi, S.mass ? S.mass + fi : nullptr, S.vel ? S.vel + fi * 3 : nullptr, nullptr,
                            cstart + (size_t)S.c0 + (size_t)f * (g.ncell + 1), spos + fo * PS, sdiam + fo,
                            skin ? skin + fo * KN : nullptr, nullptr, rec ? rec + fo : nullptr);
}

// Virial bond entries.  Padding workgroups store zero maxima for their
// wavefronts, so k_reduce_stats can read every slot.
template <int D, int OD, typename ER, int BLK, int MAXC>
__global__ __launch_bounds__(BLK) void k_sweep_bonds(const SweepSys *__restrict__ T, int ns, int cap,
                                                     const int *__restrict__ cstart, const double *__restrict__ spos,
                                                     const double *__restrict__ sdiam, int *__restrict__ bcount,
                                                     ER *__restrict__ bonds, int *__restrict__ wave_max,
                                                     float4 *__restrict__ rec)
{
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int ES = BondLayout<D, OD, false, ER>::ES;
    const SweepSys &S = T[sweep_system(T, ns, SW_BONDS, blockIdx.x)];
    int *wm = wave_max + (size_t)blockIdx.x * (BLK / 64);
    int f, pb;
    if (!sweep_frame_tile(S, SW_BONDS, blockIdx.x, f, pb)) {
        if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = 0;
        return;
    }
    const Geom g = S.g;
    const PotParams P = S.P;
    const int n = S.n;
    const size_t fo = (size_t)S.p0 + (size_t)f * n;
    bonds_body<D, OD, true, false, ER, BLK, MAXC>(g, n, P, S.cgnorm, cstart + (size_t)S.c0 + (size_t)f * (g.ncell + 1),
                                                  spos + fo * PS, sdiam + fo, nullptr, cap, bcount + fo,
                                                  bonds + fo * (size_t)cap * ES, wm, D == 3 ? rec + fo : nullptr, pb);
}

}  // namespace pgt
