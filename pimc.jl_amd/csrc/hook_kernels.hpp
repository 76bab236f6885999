// Single-move and single-estimate kernels behind the C ABI: one reference
// move (ReshapeLinear, SingleCenterOfMass), one interaction_action! value, one
// Energy or Density evaluation, and the adjoint of the potential action.
// They reuse the device physics of the sweep kernels on one chain's HBM
// paths, so checking them against the oracle checks that arithmetic.
#pragma once
#include "sweep_kernels.hpp"

namespace pimc {

struct HookArgs {
    const ChainDev *chains;
    void *pos;
    int *cell_cnt;
    unsigned short *cell_mem;
    int *bins;
    int *status;
    int chain, n, j0, m, op;  // op: 0 propose, 1 propose + Metropolis + commit
    const double *in;         // normals ((m-1) x dim) or displacement (dim)
    double u;
    double *out;              // r' then the four weights, then the accept flag
    double *work;             // scratch, (M + 1) * dim doubles
    int use_cells;
    const int *next;          // the chain's permutation (next[p] follows p across slice M - 1)
};

template <typename T, int DIM> __device__ inline Cells hook_cells(const HookArgs &a, const ChainDev &ch)
{
    return Cells{a.cell_cnt + ch.cell_off, a.cell_mem + ch.mem_off, a.bins + ch.bin_off, ch.M,
                 ch.ncell, ch.cap, ch.nbins, ch.swaps > 0 ? a.next : nullptr};
}

template <typename T, int DIM, bool CELLS, typename PathT>
__device__ inline T hook_link(const PathT &path, const Cells &C, const ChainDev &ch, const Prm<T> &P, int n,
                              int j, const T *pj, const T *pj1)
{
    if (ch.pair == kPairNone) return T(0);
    return link_sum<T, DIM, CELLS>(path, C, ch.N, ch.nbins, P, n, j, pj, pj1);
}

// ReshapeLinear on particle n, start slice j0, m links (src/updates/
// reshape.jl:31-91): the bridge of levy! with the injected normals, then the
// link sums w_initial / w_updated of the potential and pair actions.
template <typename T, int DIM, bool CELLS, bool SM> __global__ void k_hook_reshape(HookArgs a)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const ChainDev ch = a.chains[a.chain];
    const Prm<T> P = Prm<T>::make(ch);
    const int M = ch.M, m = a.m, n = a.n, j0 = a.j0, nrows = m + 1;
    const Path<T, DIM, SM> path{reinterpret_cast<T *>(a.pos) + ch.pos_off, ch.M, ch.N};
    const Cells C = hook_cells<T, DIM>(a, ch);
    T *rp = reinterpret_cast<T *>(a.work); // (m+1) x DIM, row-major per bead
    const int jm = (j0 + m) % M;
    T e0[DIM], em[DIM];
    path.get(n, j0, e0);
    path.get(n, jm, em);
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
        em[e] = shift_end(e0[e], em[e], P);
        rp[0 * DIM + e] = e0[e];
        rp[m * DIM + e] = em[e];
    }
    T cur[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) cur[e] = e0[e];
    for (int i = 1; i < m; ++i) {
        const T alpha = bridge_alpha<T>(m, i);
        const T sd = bridge_sd(alpha, P);
#pragma unroll
        for (int e = 0; e < DIM; ++e) {
            cur[e] = bridge_step(cur[e], em[e], alpha, (T)a.in[(i - 1) * DIM + e], sd);
            rp[i * DIM + e] = cur[e];
        }
    }
    for (int i = 0; i <= m; ++i)
#pragma unroll
        for (int e = 0; e < DIM; ++e) rp[i * DIM + e] = teleport(rp[i * DIM + e], P);
    double wVi = 0.0, wVu = 0.0, wUi = 0.0, wUu = 0.0;
    for (int t = 0; t < m; ++t) {
        const int j = (j0 + t) % M, jn = (j + 1) % M;
        T r0[DIM], r1[DIM];
        path.get(n, j, r0);
        path.get(n, jn, r1);
        wVi += (double)((T(-0.5) * P.tau) * (potV<T, DIM>(r0, P) + potV<T, DIM>(r1, P)));
        wVu += (double)((T(-0.5) * P.tau) * (potV<T, DIM>(rp + t * DIM, P) + potV<T, DIM>(rp + (t + 1) * DIM, P)));
        wUi += (double)hook_link<T, DIM, CELLS>(path, C, ch, P, n, j, r0, r1);
        wUu += (double)hook_link<T, DIM, CELLS>(path, C, ch, P, n, j, rp + t * DIM, rp + (t + 1) * DIM);
    }
    for (int i = 0; i <= m; ++i)
        for (int e = 0; e < DIM; ++e) a.out[e * nrows + i] = (double)rp[i * DIM + e];
    double *w = a.out + nrows * DIM;
    w[0] = wVi;
    w[1] = wVu;
    w[2] = wUi;
    w[3] = wUu;
    w[4] = 0.0;
    if (a.op == 1) {
        T lw = (T)(wVu - wVi);
        if (ch.pair != kPairNone && ch.pair_in_accept) lw += (T)(wUu - wUi);
        const bool acc = metropolis(ex(lw), (T)a.u);
        w[4] = acc ? 1.0 : 0.0;
        if (acc) {
            // rows 0..m-1 are written back, the start bead as teleported (src/updates/reshape.jl:82-86)
            bool ok = true;
            for (int t = 0; t < m; ++t) {
                const int j = (j0 + t) % M;
                path.put(n, j, rp + t * DIM);
                if (CELLS) ok = cells_move<T, DIM>(C, P, n, j, rp + t * DIM) && ok;
            }
            if (!ok) atomicOr(a.status + a.chain, 1);
        }
    }
}

// SingleCenterOfMass with displacement in[0..dim) (src/updates/com.jl:136-224).
template <typename T, int DIM, bool CELLS, bool SM> __global__ void k_hook_com(HookArgs a)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const ChainDev ch = a.chains[a.chain];
    const Prm<T> P = Prm<T>::make(ch);
    const int M = ch.M, n = a.n;
    const Path<T, DIM, SM> path{reinterpret_cast<T *>(a.pos) + ch.pos_off, ch.M, ch.N};
    const Cells C = hook_cells<T, DIM>(a, ch);
    T *rp = reinterpret_cast<T *>(a.work); // M x DIM row-major
    for (int j = 0; j < M; ++j)
#pragma unroll
        for (int e = 0; e < DIM; ++e) rp[j * DIM + e] = teleport(path.at(n, e, j) + (T)a.in[e], P);
    double wVi = 0.0, wVu = 0.0, wUi = 0.0, wUu = 0.0;
    for (int j = 0; j < M; ++j) {
        const int jn = (j + 1) % M;
        T r0[DIM], r1[DIM];
        path.get(n, j, r0);
        path.get(n, jn, r1);
        wVi += (double)((T(-0.5) * P.tau) * (potV<T, DIM>(r0, P) + potV<T, DIM>(r1, P)));
        wVu += (double)((T(-0.5) * P.tau) * (potV<T, DIM>(rp + j * DIM, P) + potV<T, DIM>(rp + jn * DIM, P)));
        wUi += (double)hook_link<T, DIM, CELLS>(path, C, ch, P, n, j, r0, r1);
        wUu += (double)hook_link<T, DIM, CELLS>(path, C, ch, P, n, j, rp + j * DIM, rp + jn * DIM);
    }
    for (int j = 0; j < M; ++j)
        for (int e = 0; e < DIM; ++e) a.out[e * M + j] = (double)rp[j * DIM + e];
    double *w = a.out + M * DIM;
    w[0] = wVi;
    w[1] = wVu;
    w[2] = wUi;
    w[3] = wUu;
    w[4] = 0.0;
    if (a.op == 1) {
        T lw = (T)(wVu - wVi);
        if (ch.pair != kPairNone && ch.pair_in_accept) lw += (T)(wUu - wUi);
        const bool acc = metropolis(ex(lw), (T)a.u);
        w[4] = acc ? 1.0 : 0.0;
        if (acc) {
            bool ok = true;
            for (int j = 0; j < M; ++j) {
                path.put(n, j, rp + j * DIM);
                if (CELLS) ok = cells_move<T, DIM>(C, P, n, j, rp + j * DIM) && ok;
            }
            if (!ok) atomicOr(a.status + a.chain, 1);
        }
    }
}

// interaction_action! value for particle n at slice j (src/updates/helper.jl:306-324)
template <typename T, int DIM, bool CELLS, bool SM> __global__ void k_hook_interaction(HookArgs a)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const ChainDev ch = a.chains[a.chain];
    const Prm<T> P = Prm<T>::make(ch);
    const Path<T, DIM, SM> path{reinterpret_cast<T *>(a.pos) + ch.pos_off, ch.M, ch.N};
    const Cells C = hook_cells<T, DIM>(a, ch);
    T r0[DIM], r1[DIM];
    path.get(a.n, a.j0, r0);
    path.get(a.n, (a.j0 + 1) % ch.M, r1);
    a.out[0] = (double)hook_link<T, DIM, CELLS>(path, C, ch, P, a.n, a.j0, r0, r1);
}

// Energy (src/measurement.jl:92-122) plus the pair energy, one workgroup.
template <typename T, int DIM, bool CELLS, bool SM> __global__ void k_hook_energy(HookArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    double *buf = reinterpret_cast<double *>(smem);
    const ChainDev ch = a.chains[a.chain];
    const Prm<T> P = Prm<T>::make(ch);
    const int N = ch.N, M = ch.M;
    const Path<T, DIM, SM> path{reinterpret_cast<T *>(a.pos) + ch.pos_off, ch.M, ch.N};
    const Cells C = hook_cells<T, DIM>(a, ch);
    const double inv_s2 = (ch.pair == kPairGaussian) ? 1.0 / (ch.sigma * ch.sigma) : 0.0;
    double link = 0.0, pot = 0.0, vkin = 0.0, su = 0.0, rgp = 0.0;
    for (int b = threadIdx.x; b < N * M; b += blockDim.x) {
        const int i = b / M, j = b - i * M;
        const int jn = (j + 1 == M) ? 0 : j + 1;
        T r0[DIM], r1[DIM];
        path.get(i, j, r0);
        path.get(jn == 0 ? a.next[i] : i, jn, r1); // inext of src/measurement.jl:98
        T dd = T(0);
#pragma unroll
        for (int e = 0; e < DIM; ++e) {
            const T dx = mindist(r0[e], r1[e], P);
            dd += dx * dx;
        }
        link += (double)dd;
        pot += (double)(potV<T, DIM>(r0, P) + potV<T, DIM>(r1, P));
        vkin += (double)rgradV<T, DIM>(r0, P);
        if (ch.pair == kPairGaussian) { // no pair energy for the tabulated propagator
            double x, y;
            pair_energy_bead<T, DIM, CELLS>(path, C, N, ch.nbins, P, inv_s2, i, j, x, y);
            su += x;
            rgp += y;
        }
    }
    link = block_sum(link, buf);
    pot = block_sum(pot, buf);
    vkin = block_sum(vkin, buf);
    su = block_sum(su, buf);
    rgp = block_sum(rgp, buf);
    if (threadIdx.x == 0) {
        const double tau = ch.tau;
        a.out[0] = ((double)(ch.dim * N) / (2.0 * tau) - (1.0 / (((4.0 * ch.lambda) * (tau * tau)) * M)) * link) +
                   (1.0 / (2.0 * M)) * pot;
        a.out[1] = (1.0 / (2.0 * M)) * vkin + (1.0 / (2.0 * M)) * pot;
        a.out[2] = ch.g * su / ch.beta;
        a.out[3] = ch.g * (su / ch.beta + rgp / (2.0 * ch.beta));
    }
}

// Density (src/measurement.jl:45-55): counts into a uint64 histogram.
template <typename T, int DIM, bool SM> __global__ void k_hook_density(HookArgs a, int nbins, unsigned long long *hist)
{
    const ChainDev ch = a.chains[a.chain];
    const Path<T, DIM, SM> path{reinterpret_cast<T *>(a.pos) + ch.pos_off, ch.M, ch.N};
    const double bw = (2.0 * ch.L) / nbins;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < ch.N * ch.M; b += gridDim.x * blockDim.x) {
        const int i = b / ch.M, j = b - i * ch.M;
        long long idx = 0, stride = 1;
        bool ok = true;
        for (int e = 0; e < DIM; ++e) {
            const long long ib = (long long)floor(((double)path.at(i, e, j) + ch.L) / bw);
            ok = ok && ib > 0 && ib < nbins + 1;
            idx += (ib - 1) * stride;
            stride *= nbins;
        }
        if (ok) atomicAdd(&hist[idx], 1ull);
    }
}

// Adjoint of S_pot = tau sum_beads V + S_U with
// S_U = sum_links sum_{pairs within h at slice j} tau/2 (v(r_j) + v(r_j+1)).
// One thread per bead: dS/dr of bead (n, j) collects tau grad V and the pair
// terms of links j (neighbour set at slice j) and j-1 (set at slice j-1);
// distances are recomputed, nothing is stored.  Block sums give S_pot,
// dS/domega and dS/dg.
template <typename T, int DIM, bool CELLS, bool SM> __global__ void k_hook_adjoint(HookArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    double *buf = reinterpret_cast<double *>(smem);
    const ChainDev ch = a.chains[a.chain];
    const Prm<T> P = Prm<T>::make(ch);
    const int N = ch.N, M = ch.M;
    const Path<T, DIM, SM> path{reinterpret_cast<T *>(a.pos) + ch.pos_off, ch.M, ch.N};
    const Cells C = hook_cells<T, DIM>(a, ch);
    const T inv_s2 = (ch.pair != kPairNone) ? T(1.0 / (ch.sigma * ch.sigma)) : T(0);
    const T ht = T(0.5) * P.tau;
    double S = 0.0, Sw = 0.0, SU = 0.0;
    for (int b = threadIdx.x; b < N * M; b += blockDim.x) {
        const int n = b / M, j = b - n * M;
        const int jp = (j == 0) ? M - 1 : j - 1;
        T r[DIM], g[DIM];
        path.get(n, j, r);
        gradV<T, DIM>(r, P, g);
#pragma unroll
        for (int e = 0; e < DIM; ++e) g[e] *= P.tau;
        S += (double)(P.tau * potV<T, DIM>(r, P));
        if (ch.pot == kPotHarmonic) Sw += (double)(P.tau * dVdomega<T, DIM>(r, ch.omega));
        if (ch.pair != kPairNone) {
            // link j: the neighbour set at slice j around bead (n, j); link j-1: the set at slice j-1
            for (int side = 0; side < 2; ++side) {
                const int js = side == 0 ? j : jp;
                T rs[DIM];
                path.get(n, js, rs);
                auto visit = [&](int k, int mult) {
                    T rk[DIM];
                    path.get(k, js, rk);
                    T d0 = mindist(rk[0], rs[0], P);
                    T r2 = d0 * d0;
                    if (DIM > 1) {
                        T d1 = mindist(rk[1], rs[1], P);
                        r2 += d1 * d1;
                    }
                    if (r2 > P.h2) return;
                    // gradient at slice j of v(|r_n^j - r_k^j|), signed minimum image
                    T rkj[DIM], x[DIM];
                    path.get(k, j, rkj);
                    T x2 = T(0);
#pragma unroll
                    for (int e = 0; e < DIM; ++e) {
                        T dx = r[e] - rkj[e];
                        dx -= P.twoL * fl(dx * P.inv2L + T(0.5));
                        x[e] = dx;
                        x2 += dx * dx;
                    }
                    const T v = P.g * ex(-x2 * P.inv2s2);
#pragma unroll
                    for (int e = 0; e < DIM; ++e) g[e] += T(mult) * ht * (-(x[e] * inv_s2) * v);
                    if (side == 0) {
                        // S_U: each unordered pair of link j counted once (half from each end)
                        const int jn = (j + 1 == M) ? 0 : j + 1;
                        T rkn[DIM], rnn[DIM];
                        path.get(k, jn, rkn);
                        path.get(n, jn, rnn);
                        T e0 = mindist(rkn[0], rnn[0], P);
                        T s2 = e0 * e0;
                        if (DIM > 1) {
                            T e1 = mindist(rkn[1], rnn[1], P);
                            s2 += e1 * e1;
                        }
                        // per unit coupling: S_U = g * SU, dS/dg = SU
                        SU += 0.5 * (double)(T(mult) * ht * (ex(-r2 * P.inv2s2) + ex(-s2 * P.inv2s2)));
                    }
                };
                if (CELLS) {
                    const int c0 = cell_of<T, DIM>(rs, P, C.nbins);
                    const int ns = (DIM == 1) ? 3 : 9;
                    for (int s = 0; s < ns; ++s) {
                        const int cc = stencil_cell(c0, s, C.nbins, DIM);
                        const int cnt = C.count(js, cc);
                        const unsigned short *mem = C.members(js, cc);
                        for (int q = 0; q < cnt; ++q)
                            if (mem[q] != n) visit(mem[q], 1);
                    }
                } else {
                    const int ns = (DIM == 1) ? 3 : 9;
                    const int c0 = (ch.nbins < 3) ? cell_of<T, DIM>(rs, P, ch.nbins) : 0;
                    for (int k = 0; k < N; ++k) {
                        if (k == n) continue;
                        int mult = 1;
                        if (ch.nbins < 3) {
                            T rk[DIM];
                            path.get(k, js, rk);
                            const int ck = cell_of<T, DIM>(rk, P, ch.nbins);
                            mult = 0;
                            for (int s = 0; s < ns; ++s) mult += (stencil_cell(c0, s, ch.nbins, DIM) == ck);
                            if (!mult) continue;
                        }
                        visit(k, mult);
                    }
                }
            }
        }
        for (int e = 0; e < DIM; ++e) a.out[((size_t)n * DIM + e) * M + j] = (double)g[e];
    }
    S = block_sum(S, buf);
    Sw = block_sum(Sw, buf);
    SU = block_sum(SU, buf);
    if (threadIdx.x == 0) {
        double *tail = a.out + (size_t)N * DIM * M;
        tail[0] = S + ch.g * SU; // S_pot
        tail[1] = Sw;            // dS/d omega
        tail[2] = SU;            // dS/d g
    }
}

// init_world (src/system.jl:36-78) with a = 0: each particle's first bead
// uniform in [-L, L)^dim, the last bead equal to it, and levy! over the M - 1
// links in between (the bridge arithmetic of levy_segment, teleport of every
// row).  One thread per particle, FP64 arithmetic, stored in the chain's
// layout; Philox purposes kInitStart and kInitLevy, sweep word 0.
template <typename T, int DIM>
__global__ void k_init_paths(const ChainDev *chains, const int *part_chain, const int *part_idx, int q0, int nq,
                             void *pos, uint64_t seed)
{
    const int gq = q0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (gq >= q0 + nq) return;
    const int c = part_chain[gq], p = part_idx[gq];
    const ChainDev ch = chains[c];
    const Prm<double> P = Prm<double>::make(ch);
    const int M = ch.M, N = ch.N;
    T *base = reinterpret_cast<T *>(pos) + ch.pos_off;
    auto put = [&](int j, int d, double v) {
        T x = (T)v;
        if (x >= (T)ch.L) x = nextafter((T)ch.L, (T)(-ch.L)); // keep [-L, L) after rounding to T
        if (ch.sm) base[((size_t)j * N + p) * DIM + d] = x;
        else base[((size_t)p * DIM + d) * M + j] = x;
    };
    U4 x = draw(seed, (uint32_t)p, 0u, (uint32_t)c, kInitStart);
    const uint32_t w[4] = {x.x, x.y, x.z, x.w};
    double r0[DIM], cur[DIM], em[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        r0[d] = 2.0 * ch.L * (u01d(w[d]) - 0.5);
        cur[d] = r0[d];
        em[d] = shift_end(r0[d], r0[d], P); // the ends coincide: no shift
    }
    const int K = M - 1; // links of the bridge
    for (int i = 1; i < K; ++i) {
        if (i & 1) x = draw(seed, (uint32_t)(p * M + i), 0u, (uint32_t)c, kInitLevy);
        double n[2];
        if (i & 1) box_muller(x.x, x.y, n[0], n[1]);
        else box_muller(x.z, x.w, n[0], n[1]);
        const double alpha = bridge_alpha<double>(K, i);
        const double sd = bridge_sd(alpha, P);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            cur[d] = bridge_step(cur[d], em[d], alpha, n[d], sd);
            put(i, d, teleport(cur[d], P));
        }
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        put(0, d, teleport(r0[d], P));
        put(M - 1, d, teleport(em[d], P));
    }
}

// init_world (src/system.jl:36-78) with hard spheres of radius a = hs_a > 0:
// one workgroup per listed chain, particles in order.  Try t of particle p is
// k_init_paths's path with sweep word t (thread 0 runs the bridge into LDS in
// FP64); the workgroup then tests every slice against the earlier particles'
// beads (FP64 copies in hs64, N x M x DIM per chain at its position offset):
// a bead closer than a (norm of the per-component |minimum-image| offsets)
// rejects the try.  At most 10 000 tries per particle, else status bit 4 and
// the chain's remaining particles keep their old paths.
template <typename T, int DIM>
__global__ void __launch_bounds__(256) k_init_paths_hs(const ChainDev *chains, const int *ids, int n, void *pos,
                                                       double *hs64, uint64_t seed, int *status)
{
    extern __shared__ __align__(16) unsigned char smem[];
    if ((int)blockIdx.x >= n) return;
    const int c = ids[blockIdx.x], tid = threadIdx.x, nt = blockDim.x;
    const ChainDev ch = chains[c];
    const Prm<double> P = Prm<double>::make(ch);
    const int M = ch.M, N = ch.N;
    const double a = ch.hs_a;
    double *cand = reinterpret_cast<double *>(smem); // [M][DIM]
    int *flag = reinterpret_cast<int *>(cand + (size_t)M * DIM);
    double *done = hs64 + ch.pos_off;                // [p][j][d] of the placed particles
    T *base = reinterpret_cast<T *>(pos) + ch.pos_off;
    for (int p = 0; p < N; ++p) {
        uint32_t t = 0;
        for (;; ++t) {
            if (t >= 10000u) {
                if (tid == 0) atomicOr(status + c, 4);
                return; // uniform: every thread reads the same flags
            }
            if (tid == 0) {
                U4 x = draw(seed, (uint32_t)p, t, (uint32_t)c, kInitStart);
                const uint32_t w[4] = {x.x, x.y, x.z, x.w};
                double r0[DIM], cur[DIM], em[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    r0[d] = 2.0 * ch.L * (u01d(w[d]) - 0.5);
                    cur[d] = r0[d];
                    em[d] = shift_end(r0[d], r0[d], P);
                }
                const int K = M - 1;
                for (int i = 1; i < K; ++i) {
                    if (i & 1) x = draw(seed, (uint32_t)(p * M + i), t, (uint32_t)c, kInitLevy);
                    double nn[2];
                    if (i & 1) box_muller(x.x, x.y, nn[0], nn[1]);
                    else box_muller(x.z, x.w, nn[0], nn[1]);
                    const double alpha = bridge_alpha<double>(K, i);
                    const double sd = bridge_sd(alpha, P);
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        cur[d] = bridge_step(cur[d], em[d], alpha, nn[d], sd);
                        cand[i * DIM + d] = teleport(cur[d], P);
                    }
                }
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    cand[d] = teleport(r0[d], P);
                    cand[(M - 1) * DIM + d] = teleport(em[d], P);
                }
                *flag = 0;
            }
            __syncthreads();
            for (int it = tid; it < p * M; it += nt) {
                const int i = it / M, j = it - i * M;
                double d2 = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    const double dd = mindist(done[((size_t)i * M + j) * DIM + d], cand[j * DIM + d], P);
                    d2 += dd * dd;
                }
                if (sqrt(d2) < a) *flag = 1;
            }
            __syncthreads();
            const bool clash = *flag != 0;
            __syncthreads();
            if (!clash) break;
        }
        for (int it = tid; it < M * DIM; it += nt) {
            const int j = it / DIM, d = it - j * DIM;
            const double v = cand[it];
            done[((size_t)p * M + j) * DIM + d] = v;
            T x = (T)v;
            if (x >= (T)ch.L) x = nextafter((T)ch.L, (T)(-ch.L)); // keep [-L, L) after rounding to T
            if (ch.sm) base[((size_t)j * N + p) * DIM + d] = x;
            else base[((size_t)p * DIM + d) * M + j] = x;
        }
        __syncthreads();
    }
}

// The batched kernel's FP32 fast paths on one particle of a tile chain whose
// staging length ell = MAXL divides M (every segment and centre-of-mass chunk
// is one skew block of the LDS rows): the particle's ring is staged into LDS
// in the tile kernel's row layout, and one wave runs the same device code as
// k_sweep_tiles with the caller's draws.  Staging (m = MAXL links from j0):
// levy_full_propose with the injected normals; centre of mass: com_chunk_full
// per chunk, summed by the lane-group butterfly when the chunks form an
// aligned lane group (nseg a power of two <= 64), else in chunk order.  The
// outputs follow k_hook_reshape / k_hook_com, with the log-weight change in
// w[1] (w[0] = 0: the fast paths form the difference directly).
template <int DIM, int MAXL, int POT> __global__ void __launch_bounds__(64) k_hook_tile_fast(HookArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const ChainDev ch = a.chains[a.chain];
    const Prm<float> P = Prm<float>::make(ch);
    const int M = ch.M, n = a.n, lane = threadIdx.x;
    int sh = 0;
    while ((2 << sh) <= MAXL) ++sh;
    const LdsRow R{tile_row_len(M, sh), sh};
    float *sp = reinterpret_cast<float *>(smem);
    float *s_red = sp + DIM * R.Mp;
    float *gp = reinterpret_cast<float *>(a.pos) + ch.pos_off + (size_t)n * DIM * M;
    for (int i = lane; i < DIM * M; i += 64) {
        const int d = i / M, j = i - d * M;
        sp[R.at(d, j)] = gp[i];
    }
    __syncthreads();
    double *w = nullptr;
    if (a.m > 0) { // staging segment
        const int j0 = a.j0, nrows = MAXL + 1;
        w = a.out + nrows * DIM;
        if (lane == 0) {
            float nb[MAXL - 1][DIM];
            GivenNormals<DIM> ns{a.in};
            const float lw = levy_full_propose<DIM, MAXL, POT>(sp, R, M, j0, P, ns, nb);
            const int jm = (j0 + MAXL) % M;
            for (int e = 0; e < DIM; ++e) {
                a.out[e * nrows] = (double)sp[R.at(e, j0)];
                a.out[e * nrows + MAXL] = (double)sp[R.at(e, jm)];
                for (int i = 1; i < MAXL; ++i) a.out[e * nrows + i] = (double)nb[i - 1][e];
            }
            w[0] = 0.0;
            w[1] = (double)lw;
            w[2] = w[3] = w[4] = 0.0;
            if (a.op == 1 && metropolis(ex(lw), (float)a.u)) {
                w[4] = 1.0;
                for (int i = 1; i < MAXL; ++i) {
                    const int j = (j0 + i) % M;
                    for (int e = 0; e < DIM; ++e) gp[e * M + j] = nb[i - 1][e];
                }
            }
        }
        return;
    }
    // centre of mass: chunk k = lane (every lane takes part in the butterfly)
    const int nseg = M / MAXL;
    float d[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) d[e] = (float)a.in[e];
    w = a.out + M * DIM;
    const bool grouped = nseg <= 64 && (nseg & (nseg - 1)) == 0;
    for (int k0 = 0; k0 < nseg; k0 += 64) {
        const int k = k0 + lane;
        bool wr = false;
        float v = 0.0f;
        if (k < nseg) {
            v = com_chunk_full<DIM, MAXL, POT>(sp + k * (MAXL + 1), R, d, P, wr);
            for (int jj = 0; jj < MAXL; ++jj) {
                const int j = k * MAXL + jj;
                for (int e = 0; e < DIM; ++e) {
                    const float x = sp[R.at(e, j)] + d[e];
                    a.out[e * M + j] = (double)(wr ? teleport(sp[R.at(e, j)] + d[e], P) : x);
                }
            }
            s_red[k] = v;
        }
        if (grouped) {
            const float tot = group_sum_dpp(v, nseg);
            if (lane == 0) s_red[0] = tot;
        }
    }
    __syncthreads();
    if (lane == 0) {
        float tot = s_red[0];
        if (!grouped) {
            tot = 0.0f;
            for (int k = 0; k < nseg; ++k) tot += s_red[k];
        }
        const float lw = -P.tau * tot;
        w[0] = 0.0;
        w[1] = (double)lw;
        w[2] = w[3] = w[4] = 0.0;
        if (a.op == 1 && metropolis(ex(lw), (float)a.u)) {
            w[4] = 1.0;
            for (int j = 0; j < M; ++j)
                for (int e = 0; e < DIM; ++e) gp[e * M + j] = (float)a.out[e * M + j];
        }
    }
}

} // namespace pimc
