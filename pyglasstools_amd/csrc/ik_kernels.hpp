// ik_kernels.hpp -- cell-list, bond-list and surrogate kernels of the IK
// stress-field engine (gfx950).  Included once, by ik_engine.hip; the field
// gathers live in their own translation units (ik_gather_*.hip).
//
//   cell list   k_sort_hist -> k_sort_scan -> k_sort_scatter (per digit), k_sort_finish
//   bond list   k_bonds   (ik_bonds.hpp)
//   fields      k_gather  (ik_gather.hpp); with the adaptive bond integral
//               k_gather_adaptive (ik_gather_adaptive.hpp)
//
// All per-frame arrays are stacked along blockIdx.y (one frame per y index).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ik_device.hpp"
#include "ik_geom.hpp"

namespace pgt {

// ------------------------------------------------------------------------
// cell list: a stable LSD radix sort of (cell, input index) per frame.
// Every pass keeps the input order among equal digits, so particles end up
// in cell order and, inside a cell, in input order: the cell-sorted arrays
// do not depend on scheduling.  No global atomics are used.
// ------------------------------------------------------------------------
constexpr int kSortBlock = 256;                     // threads per workgroup
constexpr int kSortRows = 4;                        // rows of kSortBlock items per chunk
constexpr int kSortChunk = kSortBlock * kSortRows;  // items per chunk (one workgroup)
constexpr int kSortMaxRadix = 256;                  // digits of at most 8 bits

__device__ __forceinline__ int cell_of(const Geom &g, const double *p)
{
    const int cx = cell_coord(g, 0, p[0]), cy = cell_coord(g, 1, p[1]), cz = cell_coord(g, 2, p[2]);
    return (cz * g.nc[1] + cy) * g.nc[0] + cx;
}

// Count the digit (key >> shift) & (radix - 1) over chunk ch of one frame
// into hf[digit * nchunk + ch] (hf: the frame's histogram).  With FIRST the
// keys are the particles' cells, computed here from the frame's input
// positions and stored in kf.
template <bool FIRST>
__device__ __forceinline__ void sort_hist_body(const Geom &g, int n, int ch, int nchunk, int shift, int radix,
                                               const double *__restrict__ pos, int *__restrict__ kf,
                                               int *__restrict__ hf)
{
    __shared__ int h[kSortMaxRadix];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const int base = ch * kSortChunk;
    for (int r = 0; r < kSortRows; ++r) {
        const int i = base + r * kSortBlock + t;
        if (i >= n) break;
        int k;
        if constexpr (FIRST) {
            k = cell_of(g, pos + (size_t)i * 3);
            kf[i] = k;
        } else {
            k = kf[i];
        }
        atomicAdd(&h[(k >> shift) & (radix - 1)], 1);
    }
    __syncthreads();
    if (t < radix) hf[(size_t)t * nchunk + ch] = h[t];
}

// One frame per y row (frames dealt to XCDs whole), one chunk per workgroup:
// hist[f][digit][chunk].
template <bool FIRST>
__global__ __launch_bounds__(kSortBlock) void k_sort_hist(Geom g, int n, int nf, int shift, int radix,
                                                          const double *__restrict__ pos, int *__restrict__ keys,
                                                          int *__restrict__ hist)
{
    int f, ch;
    if (!frame_block(nf, f, ch)) return;
    sort_hist_body<FIRST>(g, n, ch, gridDim.x, shift, radix, pos + (size_t)f * n * 3, keys + (size_t)f * n,
                          hist + (size_t)f * radix * gridDim.x);
}

// In-place exclusive scan of the len ints at a, by one 1024-thread
// workgroup: a frame's hist[digit][chunk] becomes the first output slot of
// that digit's items of that chunk.
__device__ __forceinline__ void scan_body(int len, int *__restrict__ a)
{
    const int T = blockDim.x, t = threadIdx.x;
    const int chunk = (len + T - 1) / T;
    const int b0 = min(len, t * chunk), b1 = min(len, b0 + chunk);
    int s = 0;
    for (int k = b0; k < b1; ++k) s += a[k];
    __shared__ int sh[1024];
    sh[t] = s;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        const int v = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    int run = sh[t] - s;
    for (int k = b0; k < b1; ++k) {
        const int v = a[k];
        a[k] = run;
        run += v;
    }
}

// One workgroup per frame (blockIdx.x), len ints each.
__global__ __launch_bounds__(1024) void k_sort_scan(int len, int *__restrict__ hist)
{
    scan_body(len, hist + (size_t)blockIdx.x * len);
}

// Move the items of one chunk to their slots for this digit.  A chunk is
// taken row by row (kSortBlock consecutive items); inside a row an item's
// slot is the chunk's offset for its digit, plus the items of that digit in
// earlier rows, in earlier wavefronts of the row, and in lower lanes of its
// wavefront (found with one ballot per digit bit).  With FIRST the value is
// the input index.
template <bool FIRST>
__device__ __forceinline__ void sort_scatter_body(int n, int ch, int nchunk, int shift, int radix,
                                                  const int *__restrict__ kin, const int *__restrict__ vin,
                                                  const int *__restrict__ hf, int *__restrict__ kout,
                                                  int *__restrict__ vout)
{
    constexpr int NW = kSortBlock / 64;
    __shared__ int run[kSortMaxRadix];     // next slot of each digit
    __shared__ int wc[NW][kSortMaxRadix];  // per-wavefront digit counts of a row, then their slots
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t < radix) run[t] = hf[(size_t)t * nchunk + ch];
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;  // lanes under this one
    int nbits = 0;
    while ((1 << nbits) < radix) ++nbits;
    const int base = ch * kSortChunk;
    for (int r = 0; r < kSortRows; ++r) {
        if (base + r * kSortBlock >= n) break;  // uniform over the workgroup
        for (int k = t; k < NW * kSortMaxRadix; k += kSortBlock) wc[k / kSortMaxRadix][k % kSortMaxRadix] = 0;
        __syncthreads();
        const int i = base + r * kSortBlock + t;
        const bool valid = i < n;
        int key = 0, val = 0, d = 0;
        if (valid) {
            key = kin[i];
            val = FIRST ? i : vin[i];
            d = (key >> shift) & (radix - 1);
        }
        unsigned long long peers = __ballot(valid);
        for (int b = 0; b < nbits; ++b) {
            const bool bit = (d >> b) & 1;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & below);
        if (valid && rank == 0) wc[w][d] = __popcll(peers);
        __syncthreads();
        if (t < radix) {
            int acc = run[t];
            for (int q = 0; q < NW; ++q) {
                const int c = wc[q][t];
                wc[q][t] = acc;
                acc += c;
            }
            run[t] = acc;
        }
        __syncthreads();
        if (valid) {
            const int slot = wc[w][d] + rank;
            kout[slot] = key;
            vout[slot] = val;
        }
        __syncthreads();
    }
}

template <bool FIRST>
__global__ __launch_bounds__(kSortBlock) void k_sort_scatter(int n, int nf, int shift, int radix,
                                                             const int *__restrict__ kin, const int *__restrict__ vin,
                                                             const int *__restrict__ hist, int *__restrict__ kout,
                                                             int *__restrict__ vout)
{
    int f, ch;
    if (!frame_block(nf, f, ch)) return;
    const size_t fo = (size_t)f * n;
    sort_scatter_body<FIRST>(n, ch, gridDim.x, shift, radix, kin + fo, vin + fo, hist + (size_t)f * radix * gridDim.x,
                             kout + fo, vout + fo);
}

// Write the cell-sorted particle arrays (slot k takes input particle
// vals[k]) and the cell starts: cstart[c] = first slot of cell c, and
// cstart[ncell] = n (keys are sorted, so slot k writes the starts of the
// cells after its predecessor's up to its own).  Positions are wrapped into
// the box along periodic axes (x - L floor((x - lo) / L): unchanged, bit for
// bit, inside the box), so a particle sits where its cell says.  In 3D with
// bonds it also writes each particle's FP32 record (x, y, z; the bond kernel
// adds the count) for the candidate scans' FP32 pre-tests.
template <int D, int OD>
__device__ __forceinline__ void sort_finish_body(const Geom &g, int n, int k, int ncell, const int *__restrict__ keys,
                                                 const int *__restrict__ vals, const double *__restrict__ pos,
                                                 const double *__restrict__ diam, const double *__restrict__ mass,
                                                 const double *__restrict__ vel, const double *__restrict__ disp,
                                                 int *__restrict__ st, double *__restrict__ spos,
                                                 double *__restrict__ sdiam, double *__restrict__ skin,
                                                 double *__restrict__ sdisp, float4 *__restrict__ rec)
{
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int KN = OD * (OD + 1) / 2;
    if (k >= n) return;
    const int c = keys[k];
    const int cp = k == 0 ? -1 : keys[k - 1];
    for (int q = cp + 1; q <= c; ++q) st[q] = k;
    if (k == n - 1)
        for (int q = c + 1; q <= ncell; ++q) st[q] = n;
    const size_t fi = (size_t)vals[k];
    const double *p = pos + fi * 3;
    double *sp = spos + (size_t)k * PS;
    auto wrap = [&](int a) {
        return g.periodic[a] ? p[a] - g.L[a] * floor((p[a] - g.lo[a]) * g.invL[a]) : p[a];
    };
    sp[0] = wrap(0);
    sp[1] = wrap(1);
    if (D == 3) { sp[2] = wrap(2); sp[3] = 0.0; }
    // 3D with bonds: the FP32 record of the candidate scans (its last word,
    // the bond count, is filled in by the bond kernel)
    if (D == 3 && rec) rec[k] = make_float4((float)sp[0], (float)sp[1], (float)sp[2], 0.0f);
    sdiam[k] = diam[fi];
    if (skin) {
        const double m = mass ? mass[fi] : 1.0;
        double v[3] = {0, 0, 0};
        if (vel) { v[0] = vel[fi * 3]; v[1] = vel[fi * 3 + 1]; v[2] = vel[fi * 3 + 2]; }
        double *kk = skin + (size_t)k * KN;
        // symmetric packing: diagonal, then (0,1) [, (0,2), (1,2)]
        for (int a = 0; a < OD; ++a) kk[a] = m * v[a] * v[a];
        if (OD == 2) kk[2] = m * v[0] * v[1];
        else { kk[3] = m * v[0] * v[1]; kk[4] = m * v[0] * v[2]; kk[5] = m * v[1] * v[2]; }
    }
    if (sdisp)
        for (int a = 0; a < 3; ++a) sdisp[(size_t)k * 3 + a] = disp ? disp[fi * 3 + a] : 0.0;
}

template <int D, int OD>
__global__ __launch_bounds__(256) void k_sort_finish(Geom g, int n, int nf, int ncell, const int *__restrict__ keys,
                                                     const int *__restrict__ vals, const double *__restrict__ pos,
                                                     const double *__restrict__ diam, const double *__restrict__ mass,
                                                     const double *__restrict__ vel, const double *__restrict__ disp,
                                                     int *__restrict__ cstart, double *__restrict__ spos,
                                                     double *__restrict__ sdiam, double *__restrict__ skin,
                                                     double *__restrict__ sdisp, float4 *__restrict__ rec)
{
    constexpr int PS = (D == 2) ? 2 : 4;
    constexpr int KN = OD * (OD + 1) / 2;
    int f, bx;
    if (!frame_block(nf, f, bx)) return;
    const size_t fo = (size_t)f * n;
    sort_finish_body<D, OD>(g, n, bx * blockDim.x + threadIdx.x, ncell, keys + fo, vals + fo, pos + fo * 3,
                            diam + fo, mass ? mass + fo : nullptr, vel ? vel + fo * 3 : nullptr,
                            disp ? disp + fo * 3 : nullptr, cstart + (size_t)f * (ncell + 1), spos + fo * PS,
                            sdiam + fo, skin ? skin + fo * KN : nullptr, sdisp ? sdisp + fo * 3 : nullptr,
                            rec ? rec + fo : nullptr);
}

// Clears that a captured HIP graph can hold (a memset node is avoided there).
template <typename T>
__global__ void k_fill(T *__restrict__ p, T v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Fold the per-wavefront statistics of one batch into the job totals: sums
// of (balls, contributions) into stats[0..1], the largest bond count into
// maxcount[0].  Each workgroup reduces a stretch of the per-wavefront arrays
// and adds its result with one atomic per total (a few hundred atomics per
// batch, against one per wavefront of the batch's kernels before).
__global__ __launch_bounds__(1024) void k_reduce_stats(int nw, const unsigned long long *__restrict__ wave_counts,
                                                       int nm, const int *__restrict__ wave_max,
                                                       unsigned long long *__restrict__ stats, int *__restrict__ maxcount)
{
    __shared__ unsigned long long s0[16], s1[16];
    __shared__ int sm[16];
    unsigned long long a = 0, b = 0;
    int m = 0;
    const int stride = gridDim.x * blockDim.x;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nw; k += stride) {
        a += wave_counts[2 * k];
        b += wave_counts[2 * k + 1];
    }
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nm; k += stride) m = max(m, wave_max[k]);
    a = wave_sum(a);
    b = wave_sum(b);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s0[w] = a;
        s1[w] = b;
        sm[w] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nwv = blockDim.x >> 6;
        for (int k = 1; k < nwv; ++k) {
            a += s0[k];
            b += s1[k];
            m = max(m, sm[k]);
        }
        if (a) atomicAdd(&stats[0], a);
        if (b) atomicAdd(&stats[1], b);
        if (m) atomicMax(maxcount, m);
    }
}

}  // namespace pgt

#include "ik_bonds.hpp"
#include "ik_surrogate.hpp"
