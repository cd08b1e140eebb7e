// ik_surrogate.hpp -- the learned bond-integral surrogate (gfx950).
// Included by ik_kernels.hpp.
//
// A tanh MLP 5 -> 64 -> 64 -> 64 -> 1 maps a bond seen from a grid point,
//   x = (a_x, a_y, b_x, b_y, r_cg),  a = (r_i - x) / r_cg,  b = (r_j - x) / r_cg,
// to the unnormalised Octic line integral w = sum_k w_k shape(|a + s_k (b - a)|^2)
// of the 5-node rule (2D).  Training runs in FP32 on the kernels below
// (dense forward, backward, mean-squared loss, Adam); inference runs in BF16
// on the matrix cores (v_mfma_f32_16x16x16_bf16) with FP32 accumulation.
//
// The field form with the surrogate enumerates the (grid point, bond) pairs
// of a frame into a pair buffer (k_pair_count, a scan, k_pair_write), infers
// w for every pair (k_mlp_infer), keeps it where the calibrated error bound
// of the pair's class is within the tolerance and recomputes it by the
// quadrature elsewhere (k_pair_finalize), and sums c' w b' (x) b' per grid
// point in pair order in FP64 (k_pair_reduce).
#pragma once
#include <hip/hip_runtime.h>

#include "ik_device.hpp"

namespace pgt {

// parameter layout (FP32): W1[5][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4
constexpr int kMlpIn = 5;
constexpr int kMlpInPad = 8;  // floats per input row in memory
constexpr int kMlpW = 64;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpIn * kMlpW, kOffW2 = kOffB1 + kMlpW, kOffB2 = kOffW2 + kMlpW * kMlpW,
              kOffW3 = kOffB2 + kMlpW, kOffB3 = kOffW3 + kMlpW * kMlpW, kOffW4 = kOffB3 + kMlpW,
              kOffB4 = kOffW4 + kMlpW, kMlpNParam = kOffB4 + 1;

// ------------------------------------------------------------------------
// FP32 training kernels (row-major activations, one sample per row)
// ------------------------------------------------------------------------

// Y[s][o] = act(b[o] + sum_k X[s][k] W[k][o]); X rows have stride ldx.  The
// layer sizes are template arguments, so a thread's K loads of X and of W
// issue together instead of one per step of a loop of unknown length.
template <int K, int N>
__global__ __launch_bounds__(256) void k_dense_fwd(int64_t n, const float *__restrict__ X, int ldx,
                                                   const float *__restrict__ W, const float *__restrict__ b,
                                                   bool tanh_act, float *__restrict__ Y)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * N) return;
    const int64_t s = idx / N;
    const int o = (int)(idx % N);
    float x[K];
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] = X[s * ldx + k];
    float acc = b[o];
#pragma unroll
    for (int k = 0; k < K; ++k) acc = fmaf(x[k], W[k * N + o], acc);
    Y[idx] = tanh_act ? tanhf(acc) : acc;
}

// Mean-squared loss over n outputs: dY = 2 (Y - T) / n and per-block sums of
// (Y - T)^2 into part[blockIdx.x] (folded by k_sum_parts).
__global__ __launch_bounds__(256) void k_mse(int64_t n, const float *__restrict__ Y, const float *__restrict__ T,
                                             float *__restrict__ dY, double *__restrict__ part)
{
    __shared__ double sh[256];
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double e2 = 0.0;
    if (s < n) {
        const float d = Y[s] - T[s];
        dY[s] = 2.0f * d / (float)n;
        e2 = (double)d * d;
    }
    sh[threadIdx.x] = e2;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// out[0] = sum(part[0..np)) / n, in a fixed order (one workgroup).
__global__ __launch_bounds__(256) void k_sum_parts(int np, const double *__restrict__ part, int64_t n,
                                                   float *__restrict__ out)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (int k = threadIdx.x; k < np; k += blockDim.x) s += part[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = (float)(sh[0] / (double)n);
}

// dZp[s][k] = (sum_o dZ[s][o] W[k][o]) * (1 - H[s][k]^2): the error at the
// pre-activation of the previous (tanh) layer of K = kMlpW units, H being
// that layer's output; N outputs (a template argument, for the same reason
// as in k_dense_fwd).
template <int N>
__global__ __launch_bounds__(256) void k_dense_bwd_data(int64_t n, const float *__restrict__ dZ,
                                                        const float *__restrict__ W, const float *__restrict__ H,
                                                        float *__restrict__ dZp)
{
    constexpr int K = kMlpW;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * K) return;
    const int64_t s = idx / K;
    const int k = (int)(idx % K);
    float z[N];
#pragma unroll
    for (int o = 0; o < N; ++o) z[o] = dZ[s * N + o];
    float acc = 0.0f;
#pragma unroll
    for (int o = 0; o < N; ++o) acc = fmaf(z[o], W[k * N + o], acc);
    const float h = H[idx];
    dZp[idx] = acc * (1.0f - h * h);
}

// Weight and bias gradients in two deterministic stages.  Stage 1: workgroup
// c (of S = gridDim.x) sums samples [n c / S, n (c+1) / S) into
// part[c][k*N + o] = sum X[s][k] dZ[s][o] and part[c][K*N + o] = sum dZ[s][o]
// (K, N <= 64).  The slice is staged through LDS kWChunk samples at a time
// (the X and dZ rows of the chunk), and thread t keeps the 4 x 4 outputs
// k = 4 (t / 16) + i, o = 4 (t % 16) + j in registers, adding the chunk's
// samples in order: 16 FMAs per pair of 16-byte LDS reads, against a
// dependent chain of global loads per output before.  Stage 2 (k_sum_split)
// adds the S rows in order.
constexpr int kGradSplit = 64;
constexpr int kWChunk = 32;
__global__ __launch_bounds__(256) void k_dense_bwd_weight(int64_t n, int K, int N, const float *__restrict__ X,
                                                          int ldx, const float *__restrict__ dZ,
                                                          float *__restrict__ part)
{
    __shared__ float4 sx[kWChunk][kMlpW / 4];  // X rows of the chunk, zero past K
    __shared__ float4 sz[kWChunk][kMlpW / 4];  // dZ rows, zero past N
    const int t = threadIdx.x;
    const int ty = t >> 4, tx = t & 15;
    const int c = blockIdx.x, S = gridDim.x;
    const int64_t s0 = n * c / S, s1 = n * (c + 1) / S;
    float acc[4][4], bacc = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
    float *fx = reinterpret_cast<float *>(sx), *fz = reinterpret_cast<float *>(sz);
    for (int64_t b = s0; b < s1; b += kWChunk) {
        __syncthreads();  // the previous chunk's reads are done
        for (int q = t; q < kWChunk * kMlpW; q += 256) {
            const int r = q / kMlpW, k = q % kMlpW;
            const int64_t smp = b + r;
            const bool ok = smp < s1;
            fx[q] = (ok && k < K) ? X[smp * ldx + k] : 0.0f;
            fz[q] = (ok && k < N) ? dZ[smp * N + k] : 0.0f;
        }
        __syncthreads();
#pragma unroll 4
        for (int r = 0; r < kWChunk; ++r) {
            const float4 a = sx[r][ty], z = sz[r][tx];
            const float av[4] = {a.x, a.y, a.z, a.w}, zv[4] = {z.x, z.y, z.z, z.w};
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(av[i], zv[j], acc[i][j]);
        }
        if (t < N)
            for (int r = 0; r < kWChunk; ++r) bacc += fz[r * kMlpW + t];
    }
    const int M = K * N + N;
    float *out = part + (size_t)c * M;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = 4 * ty + i, o = 4 * tx + j;
            if (k < K && o < N) out[k * N + o] = acc[i][j];
        }
    if (t < N) out[K * N + t] = bacc;
}

// gW, gb = the sum of the kGradSplit rows of part, in row order (the rows'
// loads issue together).
__global__ __launch_bounds__(256) void k_sum_split(int K, int N, const float *__restrict__ part,
                                                   float *__restrict__ gW, float *__restrict__ gb)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int M = K * N + N;
    if (t >= M) return;
    float v[kGradSplit];
#pragma unroll
    for (int c = 0; c < kGradSplit; ++c) v[c] = part[(size_t)c * M + t];
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < kGradSplit; ++c) acc += v[c];
    if (t < K * N) gW[t] = acc;
    else gb[t - K * N] = acc;
}

// Adam (Kingma & Ba 2015), bias-corrected, on np parameters.
__global__ void k_adam(int np, float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ------------------------------------------------------------------------
// BF16 inference on the matrix cores
// ------------------------------------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ short to_bf16(float x)
{
    const unsigned u = __float_as_uint(x);
    return (short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// One wavefront evaluates 16 samples per step with v_mfma_f32_16x16x16_bf16.
// Every layer is computed transposed, D = W^T (out x in) * H^T (in x 16):
// lane l holds A[l % 16][4 (l / 16) + r] (weights), B[4 (l / 16) + r][l % 16]
// (activations) and D[4 (l / 16) + r][l % 16] (r = 0..3), so the D tile T of
// one layer, after bias and tanh, is the B operand of k-tile T of the next
// layer with no data movement.  The weights of all layers stay in registers.
__global__ __launch_bounds__(256) void k_mlp_infer(int64_t n, const float *__restrict__ P, const float *__restrict__ X,
                                                   float *__restrict__ Y)
{
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, grp = lane >> 4;
    // weights as A operands: layer 1 (4 out tiles x 1 k tile), 2 and 3 (4 x 4), 4 (1 x 4)
    bf16x4 a1[4], a2[4][4], a3[4][4], a4[4];
    float bias1[4][4], bias2[4][4], bias3[4][4];
#pragma unroll
    for (int T = 0; T < 4; ++T) {
        const int out = 16 * T + col;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int in1 = 4 * grp + r;
            a1[T][r] = to_bf16(in1 < kMlpIn ? P[kOffW1 + in1 * kMlpW + out] : 0.0f);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int in = 16 * t + 4 * grp + r;
                a2[T][t][r] = to_bf16(P[kOffW2 + in * kMlpW + out]);
                a3[T][t][r] = to_bf16(P[kOffW3 + in * kMlpW + out]);
            }
            bias1[T][r] = P[kOffB1 + 16 * T + 4 * grp + r];
            bias2[T][r] = P[kOffB2 + 16 * T + 4 * grp + r];
            bias3[T][r] = P[kOffB3 + 16 * T + 4 * grp + r];
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) a4[t][r] = to_bf16(col == 0 ? P[kOffW4 + 16 * t + 4 * grp + r] : 0.0f);
    const float bias4 = P[kOffB4];

    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwave = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t g0 = wave * 16; g0 < n; g0 += nwave * 16) {
        const int64_t s = g0 + col;
        bf16x4 h[4];
        // layer 1: input features 4 grp + r of sample s
        bf16x4 x;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int k = 4 * grp + r;
            x[r] = to_bf16(s < n && k < kMlpIn ? X[s * kMlpInPad + k] : 0.0f);
        }
#pragma unroll
        for (int T = 0; T < 4; ++T) {
            f32x4 d = {0.0f, 0.0f, 0.0f, 0.0f};
            d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1[T], x, d, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) h[T][r] = to_bf16(tanhf(d[r] + bias1[T][r]));
        }
        // layers 2 and 3
        bf16x4 h2[4];
#pragma unroll
        for (int T = 0; T < 4; ++T) {
            f32x4 d = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int t = 0; t < 4; ++t) d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[T][t], h[t], d, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) h2[T][r] = to_bf16(tanhf(d[r] + bias2[T][r]));
        }
#pragma unroll
        for (int T = 0; T < 4; ++T) {
            f32x4 d = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int t = 0; t < 4; ++t) d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3[T][t], h2[t], d, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) h[T][r] = to_bf16(tanhf(d[r] + bias3[T][r]));
        }
        // layer 4: output row 0 of the single out tile sits in lanes 0..15, item 0
        f32x4 d = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < 4; ++t) d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a4[t], h[t], d, 0, 0, 0);
        if (grp == 0 && s < n) Y[s] = d[0] + bias4;
    }
}

// ------------------------------------------------------------------------
// the field form with the surrogate (2D)
// ------------------------------------------------------------------------
// One enumerated (grid point, bond) pair: the quadratic form of its bond
// integral (A, B; C comes with the bond entry) and the bond entry's index.
struct PairRec {
    double A, B;
    int entry;  // particle * cap + slot
    int pad;
};

// Walk the particles inside the ball of grid point t (cells around it, in
// cell order) and call visit(p, a0, a1, A) for each.
template <typename F>
__device__ __forceinline__ void for_ball(const Geom &g, const int *__restrict__ st, const double *__restrict__ sp,
                                         double x0, double x1, F &&visit)
{
    const int cx = cell_coord(g, 0, x0), cy = cell_coord(g, 1, x1);
    int rows[kMaxRows];
    const int nrows = neighbour_rows(g, cy, 0, rows);
    for (int r = 0; r < nrows; ++r) {
        int rb[2], re[2];
        const int nr = row_ranges(g, st, rows[r], cx, rb, re);
        for (int s = 0; s < nr; ++s)
            for (int p = rb[s]; p < re[s]; ++p) {
                const double a0 = minimg(sp[(size_t)p * 2] - x0, g.L[0], g.pinvL[0]);
                const double a1 = minimg(sp[(size_t)p * 2 + 1] - x1, g.L[1], g.pinvL[1]);
                const double asq = a0 * a0 + a1 * a1;
                if (asq < g.rc2) visit(p, a0, a1, asq * g.inv_rc2);
            }
    }
}

// Pairs of each grid point of one frame into counts[gp].
// The 64-bit sum of all counts goes to *total (one atomic per wavefront), so
// the host can refuse a frame whose pairs overflow the int offsets.
__global__ __launch_bounds__(128) void k_pair_count(Geom g, int G, const double *__restrict__ grid,
                                                    const int *__restrict__ st, const double *__restrict__ sp, int cap,
                                                    const int *__restrict__ bc, int *__restrict__ counts,
                                                    unsigned long long *__restrict__ total)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    if (t < G) {
        for_ball(g, st, sp, grid[(size_t)t * 3], grid[(size_t)t * 3 + 1],
                 [&](int p, double, double, double) { c += min(bc[p], cap); });
        counts[t] = c;
    }
    const unsigned long long w = wave_sum((unsigned long long)c);
    if ((threadIdx.x & 63) == 0 && w) atomicAdd(total, w);
}

// Write the pairs of grid point t at its offset (exclusive scan of the
// counts): quadratic form, entry index and the network input row.
__global__ __launch_bounds__(128) void k_pair_write(Geom g, int G, const double *__restrict__ grid,
                                                    const int *__restrict__ st, const double *__restrict__ sp, int cap,
                                                    const int *__restrict__ bc, const double *__restrict__ bd,
                                                    const int *__restrict__ offs, double rcg, PairRec *__restrict__ rec,
                                                    float *__restrict__ xin)
{
    constexpr int ES = BondLayout<2, 2, false, double>::ES;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= G) return;
    const double s2 = 2.0 * g.inv_rc;
    size_t o = (size_t)offs[t];
    for_ball(g, st, sp, grid[(size_t)t * 3], grid[(size_t)t * 3 + 1], [&](int p, double a0, double a1, double A) {
        const int nb = min(bc[p], cap);
        const double s0 = a0 * g.inv_rc, s1 = a1 * g.inv_rc;
        const double ta0 = a0 * s2, ta1 = a1 * s2;
        for (int j = 0; j < nb; ++j) {
            const double *e = bd + ((size_t)p * cap + j) * ES;
            PairRec r;
            r.A = A;
            r.B = ta0 * e[0] + ta1 * e[1];
            r.entry = p * cap + j;
            r.pad = 0;
            rec[o] = r;
            float *x = xin + o * kMlpInPad;
            x[0] = (float)s0;
            x[1] = (float)s1;
            x[2] = (float)(s0 + e[0]);
            x[3] = (float)(s1 + e[1]);
            x[4] = (float)rcg;
            x[5] = x[6] = x[7] = 0.0f;
            ++o;
        }
    });
}

// Final bond integral of each pair: the network's value where the bound of
// the pair's class (0: both ends inside the ball, so the integrand is a
// polynomial along the whole segment; 1: the segment leaves the ball) is
// within tol, the 5-node quadrature otherwise.  With tol < 0 every pair takes
// the quadrature (training targets).  Counts network pairs per workgroup.
__global__ __launch_bounds__(256) void k_pair_finalize(int64_t np, const PairRec *__restrict__ rec,
                                                       const double *__restrict__ bonds, GLRule R,
                                                       const float *__restrict__ ymlp, double bound_in,
                                                       double bound_out, double tol, double *__restrict__ w,
                                                       float *__restrict__ wf, unsigned long long *__restrict__ nmlp)
{
    constexpr int ES = BondLayout<2, 2, false, double>::ES;
    __shared__ unsigned long long sh[4];
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long used = 0;
    if (k < np) {
        const PairRec r = rec[k];
        const double *e = bonds + (size_t)r.entry * ES;
        double C = e[0] * e[0];
        C = fma(e[1], e[1], C);
        const bool inside = r.A < 1.0 && r.A + r.B + C < 1.0;
        double v;
        if (tol >= 0.0 && (inside ? bound_in : bound_out) <= tol) {
            v = (double)ymlp[k];
            used = 1;
        } else {
            v = bond_quad<0, 5, double>(R, nullptr, 5, r.A, r.B, C);
        }
        w[k] = v;
        if (wf) wf[k] = (float)v;
    }
    used = wave_sum(used);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = used;
    __syncthreads();
    if (threadIdx.x == 0 && nmlp) {
        const unsigned long long tot = sh[0] + sh[1] + sh[2] + sh[3];
        if (tot) atomicAdd(nmlp, tot);
    }
}

// Virial field of grid point t: sum over its pairs, in order, of c' w b' (x) b'.
__global__ __launch_bounds__(128) void k_pair_reduce(int G, const int *__restrict__ counts,
                                                     const int *__restrict__ offs, const PairRec *__restrict__ rec,
                                                     const double *__restrict__ bonds, const double *__restrict__ w,
                                                     double *__restrict__ out)
{
    constexpr int ES = BondLayout<2, 2, false, double>::ES;
    constexpr int OC = BondLayout<2, 2, false, double>::OC;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= G) return;
    const int c = counts[t];
    const size_t o = (size_t)offs[t];
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    for (int k = 0; k < c; ++k) {
        const double *e = bonds + (size_t)rec[o + k].entry * ES;
        const double cw = w[o + k] * e[OC];
        const double cb0 = cw * e[0], cb1 = cw * e[1];
        v0 = fma(cb0, e[0], v0);
        v1 = fma(cb1, e[1], v1);
        v2 = fma(cb0, e[1], v2);
    }
    double *d = out + (size_t)t * 4;
    d[0] = v0;
    d[1] = v2;
    d[2] = v2;
    d[3] = v1;
}

}  // namespace pgt
