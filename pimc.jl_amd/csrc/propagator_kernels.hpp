// Integral terms of the pair-product propagator on the device
// (prop_rel_interpolate_terms, src/propagator.jl:36-70):
//
//   terms(r1, r2) = -term1 - term2 + term3
//                 = sum_q [ -a_q J0(k_q r1) J0(k_q r2)
//                           - b_q (J0(k_q r1) Y0(k_q r2) + J0(k_q r2) Y0(k_q r1))
//                           + a_q Y0(k_q r1) Y0(k_q r2) ]
//
// over the nodes k_q of a fixed Gauss-Legendre rule on [0, K] (in place of
// quadgk), with a_q = w_q k_q e^{-tau k_q^2} / (2 pi) tk^2 / (1 + tk^2) and
// b_q = ... tk / (1 + tk^2), tk = 1 / ((2/pi)(eulergamma + log(k/2)) - 4/g0).
// The node list and a_q, b_q are built on the host (pimc_prop_int_terms).
//
// Step 1 (k_prop_bessel) tabulates J0 and Y0 at every (grid radius, node);
// step 2 (k_prop_terms) is the contraction over the nodes, a product of an
// (n x 2nk) and a (2nk x n) matrix with the 2 x 2 node weights
// [[-a, -b], [-b, a]] folded into one operand: LDS tiles of 64 x 64 outputs,
// 4 x 4 per thread, nodes staged 16 at a time.
#pragma once
#include <hip/hip_runtime.h>

namespace pimc {

// J[i * nk + q] = J0(k_q r_i), Y[i * nk + q] = Y0(k_q r_i), r_i = r0 + i h.
__global__ void k_prop_bessel(const double *__restrict__ k, int nk, double r0, double h, int n, double *__restrict__ J,
                              double *__restrict__ Y)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * nk) return;
    const int i = (int)(t / nk), q = (int)(t - (long long)i * nk);
    const double r = r0 + i * h;
    J[t] = j0(k[q] * r);
    Y[t] = y0(k[q] * r);
}

constexpr int kPropTile = 64, kPropChunk = 16, kPropThreads = 256;

// out[i * n + l] = terms(r_i, r_l); grid (ceil(n/64), ceil(n/64)), 256 threads.
__global__ void __launch_bounds__(kPropThreads)
    k_prop_terms(const double *__restrict__ J, const double *__restrict__ Y, const double *__restrict__ ca,
                 const double *__restrict__ cb, int nk, int n, double *__restrict__ out)
{
    __shared__ double sJi[kPropChunk][kPropTile], sYi[kPropChunk][kPropTile];
    __shared__ double sJl[kPropChunk][kPropTile], sYl[kPropChunk][kPropTile];
    __shared__ double sa[kPropChunk], sb[kPropChunk];
    const int tid = threadIdx.x, ti = tid & 15, tl = tid >> 4;
    const int i0 = blockIdx.y * kPropTile, l0 = blockIdx.x * kPropTile;
    double acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = 0.0;

    for (int q0 = 0; q0 < nk; q0 += kPropChunk) {
        // 64 rows x 16 nodes of each operand: thread t loads node t % 16 of rows t / 16 + 16 m
        const int qq = tid & (kPropChunk - 1);
        const int q = q0 + qq;
#pragma unroll
        for (int m = 0; m < kPropTile * kPropChunk / kPropThreads; ++m) {
            const int row = (tid >> 4) + 16 * m;
            const int gi = i0 + row, gl = l0 + row;
            const bool qi = q < nk;
            sJi[qq][row] = (qi && gi < n) ? J[(size_t)gi * nk + q] : 0.0;
            sYi[qq][row] = (qi && gi < n) ? Y[(size_t)gi * nk + q] : 0.0;
            sJl[qq][row] = (qi && gl < n) ? J[(size_t)gl * nk + q] : 0.0;
            sYl[qq][row] = (qi && gl < n) ? Y[(size_t)gl * nk + q] : 0.0;
        }
        if (tid < kPropChunk) {
            sa[tid] = (q0 + tid < nk) ? ca[q0 + tid] : 0.0;
            sb[tid] = (q0 + tid < nk) ? cb[q0 + tid] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int c = 0; c < kPropChunk; ++c) {
            const double a = sa[c], b = sb[c];
            double u[4], v[4], jl[4], yl[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                const double ji = sJi[c][ti + 16 * x], yi = sYi[c][ti + 16 * x];
                u[x] = -a * ji - b * yi;
                v[x] = -b * ji + a * yi;
            }
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                jl[y] = sJl[c][tl + 16 * y];
                yl[y] = sYl[c][tl + 16 * y];
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) acc[x][y] += u[x] * jl[y] + v[x] * yl[y];
        }
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int gi = i0 + ti + 16 * x, gl = l0 + tl + 16 * y;
            if (gi < n && gl < n) out[(size_t)gi * n + gl] = acc[x][y];
        }
}

} // namespace pimc
