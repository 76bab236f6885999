// Batched sweep kernels.
//
// A sweep of one chain is, per particle, staging segments of length `ell`
// tiling the ring at a Philox-drawn offset (ReshapeLinear, src/updates/
// reshape.jl:31-91, with levy!, src/updates/helper.jl:118-139), followed by
// one centre-of-mass try per particle (SingleCenterOfMass, src/updates/
// com.jl:136-224).  Segments that share only end beads touch disjoint
// interior beads, so the segments of a ring are independent given their ends.
//
//  * k_sweep_tiles: chains without pair action or swap moves.  Particles are
//    then independent, so a workgroup owns a tile of particles (of one or
//    several chains with equal M), stages the tile in LDS, runs K sweeps there
//    and writes it back once.  One thread per (particle, segment).
//  * k_sweep_whole: chains with a pair action, swap moves (permutation
//    cycles) or estimators that need whole-chain products.  A workgroup owns a
//    chain in HBM; particles are updated one after another, the segments (and
//    links) of the current particle in parallel, pair sums through per-slice
//    cell lists; swap tries follow the centre-of-mass pass.
//  * the window kernels (window_kernels.hpp) take the Gaussian-pair chains with
//    cell lists instead, many particles at once.
#pragma once
#include "physics.hpp"

namespace pimc {

constexpr int kItemsPerThread = 1; // register slots of the estimator sums per thread
constexpr int kTileMaxThreads = 256; // workgroup size bound of the tile kernel
// cumulative counters per chain: staging accepted, COM accepted, COM tries, swaps accepted,
// swap tries, staging tries, staging bead moves
constexpr int kCnt = 7;
constexpr size_t kTileLdsBudget = 40 * 1024; // LDS of one tile: 4 workgroups share a CU's 160 KiB
constexpr int kNQ = 12;            // per-chain per-block estimator quantities

// block accumulator layout (doubles), per (block, chain)
enum BlockQ {
    qN = 0,      // measurements
    qE = 1,      // sum E (thermodynamic + pair)
    qEv = 2,     // sum E_virial (+ pair virial)
    qEpair = 3,  // sum pair energy
    qSw = 4,     // sum dS/d omega
    qSg = 5,     // sum dS/d g
    qESw = 6,    // sum E dS/d omega
    qESg = 7,    // sum E dS/d g
    qdEw = 8,    // sum dE/d omega
    qdEg = 9,    // sum dE/d g
    qE2 = 10,    // sum E^2
    qWhole = 11, // 1 if the products above were measured on the whole chain
};

// Sweep and block counters of a context, kept on the device and advanced by
// k_advance after each launch, so the launch arguments never change and a
// captured launch sequence can be replayed.
struct RunState {
    unsigned sweep;
    int blk;
    int pad0, pad1;
};

__global__ void k_advance(RunState *rs, int nsweeps, int measure)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        rs->sweep += (unsigned)nsweeps;
        if (measure) rs->blk += 1;
    }
}

struct TileDev {
    long long pos_off; // element offset of the tile's first particle
    int q0, np, M, nseg;
    int ell, items, nslot, sh; // nslot: distinct chains in the tile; sh: LDS skew shift
    int pst, grouped;          // pst: LDS elements per particle (dim rows plus padding);
                               // grouped: FP32 tile whose particles' segments are aligned
                               // lane groups (ell = MAXL a power of two dividing M, nseg a
                               // power of two <= 64)
    int every_st, every_com;   // the staging and COM passes run on the sweeps s with
                               // s % every == 0 (the same for every chain of the tile)
};

// Padded LDS row length of M slices skewed by j >> sh.
__host__ __device__ inline int tile_row_len(int M, int sh) { return M + (M >> sh) + 1; }

// LDS elements per particle: its dim rows, padded so that the 64 lanes of a
// wave, which start on the segments k = 0..nseg-1 of consecutive particles,
// hit distinct banks (checked for 32 and for 64 four-byte banks).
inline int tile_particle_stride(int M, int ell, int sh, int dim)
{
    const int nseg = (M + ell - 1) / ell, base = dim * tile_row_len(M, sh);
    auto conflicts = [&](int ps) {
        int bad = 0;
        for (int banks = 32; banks <= 64; banks *= 2)
            for (int h = 0; h < 64; h += banks) {
                unsigned long long seen = 0ull;
                for (int lane = h; lane < h + banks; ++lane) {
                    const int q = lane / nseg, k = lane % nseg, j = k * ell;
                    const int b = (q * ps + j + (j >> sh)) % banks;
                    if (seen >> b & 1ull) ++bad;
                    seen |= 1ull << b;
                }
            }
        return bad;
    };
    int best = base, least = conflicts(base);
    for (int pad = 1; pad < 64 && least > 0; ++pad) {
        const int n = conflicts(base + pad);
        if (n < least) least = n, best = base + pad;
    }
    return best;
}

struct SweepArgs {
    const ChainDev *chains;
    const void *prm;       // Prm<T> per chain
    const TileDev *tiles;
    const int *part_chain; // global particle -> chain
    const int *part_idx;   // global particle -> index within its chain
    void *pos;
    double *part_acc;      // per particle, per launch: sum link^2, sum V, sum r.gradV, measurements
    int *part_lacc;        // per particle, per launch: accepted staging segments
    int *part_cacc;        // per particle, per launch: accepted COM moves
    const double *com_step;
    unsigned long long *radial;
    unsigned long long *cart;
    double radial_winv;
    int radial_bins, cart_bins;
    uint64_t seed;
    const struct RunState *rs; // first sweep of this launch (device side, so a captured graph replays)
    int nsweeps, measure;
    int measure_every; // estimators after the sweeps s with (s + 1) % measure_every == 0
    int sweep_off;     // whole-chain kernel: this launch starts sweep_off sweeps after the run state's sweep
    int split;         // whole-chain kernel: estimators left to k_meas_beads / k_meas_fold (one sweep per launch)
    int *next;         // per particle: the particle whose slice 0 follows its slice M - 1
    int tile0;             // first tile of this launch (tiles are grouped by potential kind)
    int link_group;        // lanes per link in the whole-chain kernel (power of two <= 64)
    // whole-chain kernel
    const int *whole;
    int *cell_cnt;
    unsigned short *cell_mem;
    int *bins;
    double *chain_meas;    // per chain, per launch: kNQ sums
    int *status;           // per chain: cell-list overflow flag
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// LDS layout of k_sweep_tiles; the host sizes the launch with the same function.
__host__ __device__ inline size_t tile_lds_bytes(int np, int dim, int pst, int items, int tsize, int nslot, int rbins)
{
    const size_t posb = (size_t)np * pst * tsize;
    const size_t redb = (size_t)items * 3 * sizeof(double); // end-of-launch estimator sums, over the positions
    size_t o = align16(posb > redb ? posb : redb);
    o += 5 * align16((size_t)np * 4);        // off, flag, lacc, cacc, chain slot
    o += align16((size_t)np * dim * tsize);  // COM displacement
    o += align16((size_t)np * tsize);        // COM uniform
    o += align16((size_t)items * tsize);     // COM partial sums
    o += align16((size_t)nslot * rbins * 4); // radial histogram per chain of the tile
    return o;
}

// Slice j of coordinate row d of one particle in LDS.  Rows are padded to
// Mp = M + (M >> sh) + 1 and slice j sits at j + (j >> sh): the threads of a
// wave walk the segments of a ring, which start `ell` slices apart, so with
// sh = floor(log2(ell)) neighbouring lanes hit addresses ell + 1 apart (an
// odd stride for even ell) instead of ell apart, which spreads them over the
// LDS banks.
struct LdsRow {
    int Mp, sh;
    __device__ __forceinline__ int at(int d, int j) const { return d * Mp + j + (j >> sh); }
};

// One staging segment of a particle held in LDS, start slice `start`, m
// links (levy!, src/updates/helper.jl:118-139, then the Metropolis test of
// src/updates/reshape.jl:66-80 on the potential link actions).  Returns true
// if accepted (interior beads written back).
template <typename T, int DIM, int MAXL, int POT>
__device__ __forceinline__ bool levy_segment(T *xp, const LdsRow &R, int M, int start, int m, const Prm<T> &P,
                                             uint64_t seed, uint32_t item0, uint32_t sweep, uint32_t chain)
{
    int jm = start + m;
    if (jm >= M) jm -= M;
    T cur[DIM], em[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        cur[d] = xp[R.at(d, start)];
        em[d] = shift_end(cur[d], xp[R.at(d, jm)], P);
    }
    T nb[MAXL - 1][DIM];
    T dv = T(0);
    U4 xr{0u, 0u, 0u, 0u};
#pragma unroll
    for (int i = 1; i < MAXL; ++i) {
        if (i < m) {
            int b = start + i;
            if (b >= M) b -= M;
            // one Philox block per interior bead pair (i, i+1), i odd
            if (i & 1) xr = draw(seed, item0 + (uint32_t)b, sweep, chain, kLevy);
            T n0, n1;
            if (i & 1) box_muller(xr.x, xr.y, n0, n1);
            else box_muller(xr.z, xr.w, n0, n1);
            const T alpha = bridge_alpha<T>(m, i);
            const T sd = bridge_sd(alpha, P);
            cur[0] = bridge_step(cur[0], em[0], alpha, n0, sd);
            nb[i - 1][0] = teleport(cur[0], P);
            if (DIM > 1) {
                cur[1] = bridge_step(cur[1], em[1], alpha, n1, sd);
                nb[i - 1][1] = teleport(cur[1], P);
            }
            T old[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) old[d] = xp[R.at(d, b)];
            dv += potVk<T, DIM, POT>(nb[i - 1], P) - potVk<T, DIM, POT>(old, P);
        }
    }
    const T lw = -P.tau * dv;
    const U4 xa = draw(seed, item0 + (uint32_t)start, sweep, chain, kLevyAccept);
    const bool acc = metropolis(ex(lw), Uniform<T>::get(xa.x));
    if (acc) {
#pragma unroll
        for (int i = 1; i < MAXL; ++i) {
            if (i < m) {
                int b = start + i;
                if (b >= M) b -= M;
#pragma unroll
                for (int d = 0; d < DIM; ++d) xp[R.at(d, b)] = nb[i - 1][d];
            }
        }
    }
    return acc;
}

// Sum over the lanes of an aligned group of G lanes (G a power of two <= 64).
template <typename T> __device__ __forceinline__ T group_sum(T v, int G)
{
    for (int off = G >> 1; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The value of another lane of the same row of 16, chosen by a DPP control.
template <int CTRL> __device__ __forceinline__ float dpp_lane(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL> __device__ __forceinline__ double dpp_lane(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// group_sum for groups of G <= 64 lanes (G uniform): butterfly partners by
// DPP within rows of 16 (lane ^ 1, lane ^ 2, mirror in 8, mirror in 16), then
// shuffles across rows.  Each step adds the same two numbers in both partner
// lanes, so every lane of a group ends with the same bits.
template <typename T> __device__ __forceinline__ T group_sum_dpp(T v, int G)
{
    if (G >= 2) v += dpp_lane<0xB1>(v);  // quad_perm [1,0,3,2]
    if (G >= 4) v += dpp_lane<0x4E>(v);  // quad_perm [2,3,0,1]
    if (G >= 8) v += dpp_lane<0x141>(v); // row_half_mirror
    if (G >= 16) v += dpp_lane<0x140>(v); // row_mirror
    if (G >= 32) v += __shfl_xor(v, 16, 64);
    if (G >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

// Normals of the interior beads of a staging segment, as the batched sweep
// draws them: one Philox block per bead pair (i, i+1), i odd, keyed by the
// slice b of bead i, words x, y for bead i and z, w for bead i + 1.
struct PhiloxNormals {
    uint64_t seed;
    uint32_t item0, sweep, chain;
    U4 xr;
    __device__ __forceinline__ void operator()(int i, int b, float (&n)[2])
    {
        if (i & 1) {
            xr = draw(seed, item0 + (uint32_t)b, sweep, chain, kLevy);
            box_muller(xr.x, xr.y, n[0], n[1]);
        } else {
            box_muller(xr.z, xr.w, n[0], n[1]);
        }
    }
};

// Normals given by the caller ((MAXL - 1) x DIM, row-major per bead): the
// single-move hooks run the batched kernel's staging code with them.
template <int DIM> struct GivenNormals {
    const double *nrm;
    __device__ __forceinline__ void operator()(int i, int, float (&n)[2])
    {
        n[0] = (float)nrm[(i - 1) * DIM];
        n[1] = DIM > 1 ? (float)nrm[(i - 1) * DIM + 1] : 0.0f;
    }
};

// FP32 staging segment of exactly MAXL links: the bridge of levy_segment and
// the log-weight change of the potential, with two changes that only touch
// rounding.  The bridge coefficients are compile-time constants (kBridge),
// and the periodic wrap is applied only when a new bead of the segment left
// [-L, L): inside the box teleport is the identity, and its FP32 evaluation
// ((x + L) - ...) - L would only round x to the spacing of L.  Fills the
// interior beads nb and returns -tau (sum V' - sum V).
template <int DIM, int MAXL, int POT, typename NS>
__device__ __forceinline__ float levy_full_propose(const float *xp, const LdsRow &R, int M, int start,
                                                   const Prm<float> &P, NS &ns, float (&nb)[MAXL - 1][DIM])
{
    int jm = start + MAXL;
    if (jm >= M) jm -= M;
    float cur[DIM], em[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        cur[d] = xp[R.at(d, start)];
        em[d] = shift_end(cur[d], xp[R.at(d, jm)], P);
    }
    float vn = 0.0f, vo = 0.0f, mx = 0.0f;
#pragma unroll
    for (int i = 1; i < MAXL; ++i) {
        int b = start + i;
        if (b >= M) b -= M;
        float n[2];
        ns(i, b, n);
        const float al = kBridge<MAXL>.alpha[i], sd = P.sdl * kBridge<MAXL>.root[i];
        float old[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            cur[d] = bridge_step(cur[d], em[d], al, n[d], sd);
            nb[i - 1][d] = cur[d];
            mx = fmaxf(mx, fabsf(cur[d]));
            old[d] = xp[R.at(d, b)];
        }
        if constexpr (POT == kPotHarmonic) {
            // sum of |new|^2 - |old|^2, one FMA per coordinate; times w^2/2 below
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                vn = fmaf(nb[i - 1][d], nb[i - 1][d], vn);
                vn = fmaf(-old[d], old[d], vn);
            }
        } else {
            vn += potVk<float, DIM, POT>(nb[i - 1], P);
            vo += potVk<float, DIM, POT>(old, P);
        }
    }
    if constexpr (POT == kPotHarmonic) vn *= P.halfw2;
    if (mx >= P.L) {
        if constexpr (POT == kPotHarmonic) {
            vo = 0.0f;
#pragma unroll
            for (int i = 1; i < MAXL; ++i) {
                int b = start + i;
                if (b >= M) b -= M;
                float old[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) old[d] = xp[R.at(d, b)];
                vo += potVk<float, DIM, POT>(old, P);
            }
        }
        vn = 0.0f;
#pragma unroll
        for (int i = 1; i < MAXL; ++i) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) nb[i - 1][d] = teleport(nb[i - 1][d], P);
            vn += potVk<float, DIM, POT>(nb[i - 1], P);
        }
    }
    return -P.tau * (vn - vo);
}

// The batched sweep's FP32 staging segment of MAXL links: levy_full_propose
// with the Philox draws, the Metropolis test, and the write-back on accept.
template <int DIM, int MAXL, int POT>
__device__ __forceinline__ bool levy_full(float *xp, const LdsRow &R, int M, int start, const Prm<float> &P,
                                          uint64_t seed, uint32_t item0, uint32_t sweep, uint32_t chain)
{
    float nb[MAXL - 1][DIM];
    PhiloxNormals ns{seed, item0, sweep, chain, U4{0u, 0u, 0u, 0u}};
    const float lw = levy_full_propose<DIM, MAXL, POT>(xp, R, M, start, P, ns, nb);
    const U4 xa = draw(seed, item0 + (uint32_t)start, sweep, chain, kLevyAccept);
    const bool acc = metropolis(ex(lw), Uniform<float>::get(xa.x));
    if (acc) {
#pragma unroll
        for (int i = 1; i < MAXL; ++i) {
            int b = start + i;
            if (b >= M) b -= M;
#pragma unroll
            for (int d = 0; d < DIM; ++d) xp[R.at(d, b)] = nb[i - 1][d];
        }
    }
    return acc;
}

// FP32 centre-of-mass chunk that is one skew block of the LDS rows (slice
// k MAXL + jj of the chunk at x0[jj]): the potential change sum_j V(r_j + d)
// - V(r_j) of its MAXL beads.  In the harmonic trap (w^2/2) (2 d.S + MAXL
// |d|^2) with S the sum of the beads, two adds per bead instead of both
// potentials.  wrap: a moved bead left [-L, L), so the moved beads are
// teleported (and the sum taken over them).
template <int DIM, int MAXL, int POT>
__device__ __forceinline__ float com_chunk_full(const float *x0, const LdsRow &R, const float (&d)[DIM],
                                                const Prm<float> &P, bool &wrap)
{
    float vn = 0.0f, vo = 0.0f, mx = 0.0f;
    if constexpr (POT == kPotHarmonic) {
        float sr[DIM];
#pragma unroll
        for (int e = 0; e < DIM; ++e) sr[e] = 0.0f;
#pragma unroll
        for (int jj = 0; jj < MAXL; ++jj)
#pragma unroll
            for (int e = 0; e < DIM; ++e) {
                const float ro = x0[e * R.Mp + jj];
                sr[e] += ro;
                mx = fmaxf(mx, fabsf(ro + d[e]));
            }
        float ds = d[0] * sr[0], dd = d[0] * d[0];
        if (DIM > 1) {
            ds += d[DIM - 1] * sr[DIM - 1];
            dd += d[DIM - 1] * d[DIM - 1];
        }
        vn = P.halfw2 * (2.0f * ds + float(MAXL) * dd);
        if (mx >= P.L) {
            vn = 0.0f;
#pragma unroll
            for (int jj = 0; jj < MAXL; ++jj) {
                float ro[DIM];
#pragma unroll
                for (int e = 0; e < DIM; ++e) ro[e] = x0[e * R.Mp + jj];
                vo += potVk<float, DIM, POT>(ro, P);
            }
        }
    } else {
#pragma unroll
        for (int jj = 0; jj < MAXL; ++jj) {
            float ro[DIM], rn[DIM];
#pragma unroll
            for (int e = 0; e < DIM; ++e) {
                ro[e] = x0[e * R.Mp + jj];
                rn[e] = ro[e] + d[e];
                mx = fmaxf(mx, fabsf(rn[e]));
            }
            vn += potVk<float, DIM, POT>(rn, P);
            vo += potVk<float, DIM, POT>(ro, P);
        }
    }
    wrap = mx >= P.L;
    if (wrap) {
        vn = 0.0f;
#pragma unroll
        for (int jj = 0; jj < MAXL; ++jj) {
            float rn[DIM];
#pragma unroll
            for (int e = 0; e < DIM; ++e) rn[e] = teleport(x0[e * R.Mp + jj] + d[e], P);
            vn += potVk<float, DIM, POT>(rn, P);
        }
    }
    return vn - vo;
}

// At most 256 threads; FP32 at 4 waves per SIMD (128 VGPRs: with tiles of
// at most 40 KiB of LDS, four workgroups share a CU), FP64 at 2 (its values
// take register pairs).
template <typename T, int DIM, int MAXL, int POT>
__global__ void __launch_bounds__(kTileMaxThreads) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2))) k_sweep_tiles(SweepArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const TileDev tl = a.tiles[a.tile0 + blockIdx.x];
    const Prm<T> *prm = reinterpret_cast<const Prm<T> *>(a.prm);
    const int tid = threadIdx.x, nt = blockDim.x;
    const int np = tl.np, M = tl.M, nseg = tl.nseg, ell = tl.ell, items = tl.items;
    const LdsRow R{tile_row_len(M, tl.sh), tl.sh};
    const uint32_t sweep0 = a.rs->sweep;
    const int pstride = tl.pst; // LDS elements per particle
    const int rbins = a.radial_bins;
    constexpr bool kF32 = std::is_same<T, float>::value;
    // FP32 fast paths when every segment is one skew block of the LDS rows
    // (ell == MAXL, both powers of two, and ell divides M)
    const bool aligned = kF32 && ell == MAXL && (M & (MAXL - 1)) == 0;
    // ... and a particle's segments form aligned lane groups of a wave (set by the host)
    const bool grouped = kF32 && tl.grouped;
    // item it = (particle q, segment k) with k fastest; thread tid visits it = tid + r nt,
    // so (q, k) advance by (nt / nseg, nt % nseg) per step (no divide in the loops)
    const int dq = nt / nseg, dk = nt - (nt / nseg) * nseg;
    const int q_first = tid / nseg, k_first = tid - (tid / nseg) * nseg;
    int slot_q[kItemsPerThread], slot_k[kItemsPerThread];
#pragma unroll
    for (int r = 0; r < kItemsPerThread; ++r) {
        const int it = tid + r * nt;
        slot_q[r] = it / nseg;
        slot_k[r] = it - slot_q[r] * nseg;
    }

    T *sp = reinterpret_cast<T *>(smem);
    const size_t posb = (size_t)np * pstride * sizeof(T), redb = (size_t)items * 3 * sizeof(double);
    size_t o = align16(posb > redb ? posb : redb);
    int *s_off = reinterpret_cast<int *>(smem + o);
    o += align16((size_t)np * 4);
    int *s_flag = reinterpret_cast<int *>(smem + o);
    o += align16((size_t)np * 4);
    int *s_lacc = reinterpret_cast<int *>(smem + o);
    o += align16((size_t)np * 4);
    int *s_cacc = reinterpret_cast<int *>(smem + o);
    o += align16((size_t)np * 4);
    int *s_slot = reinterpret_cast<int *>(smem + o);
    o += align16((size_t)np * 4);
    T *s_d = reinterpret_cast<T *>(smem + o);
    o += align16((size_t)np * DIM * sizeof(T));
    T *s_u = reinterpret_cast<T *>(smem + o);
    o += align16((size_t)np * sizeof(T));
    T *s_red = reinterpret_cast<T *>(smem + o);
    o += align16((size_t)items * sizeof(T));
    unsigned *s_hist = reinterpret_cast<unsigned *>(smem + o);

    // HBM rows (particle, coordinate) of M slices -> padded, skewed LDS rows
    T *gpos = reinterpret_cast<T *>(a.pos) + tl.pos_off;
    const int nrow = np * DIM;
    for (int i = tid; i < nrow * M; i += nt) {
        const int row = i / M, j = i - row * M, q = row / DIM, d = row - q * DIM;
        sp[q * pstride + d * R.Mp + j + (j >> R.sh)] = gpos[i];
    }
    for (int q = tid; q < np; q += nt) {
        s_lacc[q] = 0;
        s_cacc[q] = 0;
    }
    for (int i = tid; i < tl.nslot * rbins; i += nt) s_hist[i] = 0u;
    if (tid == 0) {
        // chains are contiguous in the particle order: slot = chain boundaries before q
        int slot = 0;
        for (int q = 0; q < np; ++q) {
            if (q > 0 && a.part_chain[tl.q0 + q] != a.part_chain[tl.q0 + q - 1]) ++slot;
            s_slot[q] = slot;
        }
    }
    double mL[kItemsPerThread], mV[kItemsPerThread], mR[kItemsPerThread];
#pragma unroll
    for (int r = 0; r < kItemsPerThread; ++r) mL[r] = mV[r] = mR[r] = 0.0;
    int nmeas = 0;
    // sweeps until the next measured one: (sw + 1) % measure_every == 0 without a divide per sweep
    int meas_in = (int)((uint32_t)a.measure_every - 1u - sweep0 % (uint32_t)a.measure_every);
    __syncthreads();

    for (int s = 0; s < a.nsweeps; ++s) {
        const uint32_t sw = sweep0 + (uint32_t)s;
        const bool meas_now = a.measure && meas_in == 0;
        meas_in = (meas_in == 0) ? a.measure_every - 1 : meas_in - 1;
        // move weights: uniform over the workgroup (one value per tile)
        const bool st_on = pass_on(tl.every_st, sw), com_on = pass_on(tl.every_com, sw);
        // per-particle draws of this sweep: staging offset, COM displacement and uniform
        for (int q = tid; q < np; q += nt) {
            const int gq = tl.q0 + q;
            const int c = a.part_chain[gq], p = a.part_idx[gq];
            const U4 x = draw(a.seed, 0u, sw, (uint32_t)c, kOffset); // one offset per chain and sweep
            s_off[q] = (int)(((uint64_t)x.x * (uint64_t)ell) >> 32);
            const U4 y = draw(a.seed, (uint32_t)p, sw, (uint32_t)c, kCom);
            const T step2 = (T)a.com_step[c] * T(2);
            s_d[q * DIM + 0] = step2 * (Uniform<T>::get(y.x) - T(0.5));
            if (DIM > 1) s_d[q * DIM + 1] = step2 * (Uniform<T>::get(y.y) - T(0.5));
            s_u[q] = Uniform<T>::get(y.z);
        }
        __syncthreads();

        // staging segments
        for (int it = st_on ? tid : items, q = q_first, k = k_first; it < items;
             it += nt, q += dq + (k + dk >= nseg ? 1 : 0), k = (k + dk >= nseg) ? k + dk - nseg : k + dk) {
            const int m = min(ell, M - k * ell);
            if (m < 2) continue;
            const int gq = tl.q0 + q;
            const int c = a.part_chain[gq], p = a.part_idx[gq];
            const Prm<T> P = prm[c];
            int start = s_off[q] + k * ell;
            if (start >= M) start -= M;
            bool acc;
            if constexpr (kF32) {
                if (m == MAXL)
                    acc = levy_full<DIM, MAXL, POT>(sp + (size_t)q * pstride, R, M, start, P, a.seed,
                                                    (uint32_t)(p * M), sw, (uint32_t)c);
                else
                    acc = levy_segment<T, DIM, MAXL, POT>(sp + (size_t)q * pstride, R, M, start, m, P, a.seed,
                                                          (uint32_t)(p * M), sw, (uint32_t)c);
            } else {
                acc = levy_segment<T, DIM, MAXL, POT>(sp + (size_t)q * pstride, R, M, start, m, P, a.seed,
                                                      (uint32_t)(p * M), sw, (uint32_t)c);
            }
            if (acc) atomicAdd(&s_lacc[q], 1);
        }
        __syncthreads();

        // centre of mass: partial Delta V per chunk of `ell` beads
        if (com_on) {
        unsigned wrap = 0u; // bit r: chunk r of this thread has a moved bead outside the box
        for (int it = tid, q = q_first, k = k_first, r = 0; it < items;
             it += nt, q += dq + (k + dk >= nseg ? 1 : 0), k = (k + dk >= nseg) ? k + dk - nseg : k + dk, ++r) {
            const int j0 = k * ell, j1 = min(M, j0 + ell);
            const Prm<T> P = prm[a.part_chain[tl.q0 + q]];
            const T *xp = sp + (size_t)q * pstride;
            T d[DIM];
#pragma unroll
            for (int e = 0; e < DIM; ++e) d[e] = s_d[q * DIM + e];
            if constexpr (kF32) {
                if (aligned) {
                    // the chunk is one skew block: slice k MAXL + jj sits at k (MAXL + 1) + jj
                    const T *x0 = xp + k * (MAXL + 1);
                    bool wr = false;
                    const T dvc = com_chunk_full<DIM, MAXL, POT>(x0, R, d, P, wr);
                    if (wr) wrap |= 1u << r;
                    if (!grouped) {
                        s_red[it] = dvc;
                        continue;
                    }
                    // the nseg chunks of particle q are an aligned lane group of this
                    // wave: every lane gets the same total (each butterfly step adds
                    // the same two numbers in both partner lanes), decides, and moves
                    // its own chunk
                    const T tot = group_sum_dpp(dvc, nseg);
                    const bool acc = metropolis(ex(-P.tau * tot), s_u[q]);
                    if (acc && k == 0) s_cacc[q] += 1;
                    if (acc) {
                        T *xw = sp + (size_t)q * pstride + k * (MAXL + 1);
                        if (!((wrap >> r) & 1u)) {
#pragma unroll
                            for (int jj = 0; jj < MAXL; ++jj)
#pragma unroll
                                for (int e = 0; e < DIM; ++e) xw[e * R.Mp + jj] += d[e];
                        } else {
#pragma unroll
                            for (int jj = 0; jj < MAXL; ++jj)
#pragma unroll
                                for (int e = 0; e < DIM; ++e) xw[e * R.Mp + jj] = teleport(xw[e * R.Mp + jj] + d[e], P);
                        }
                    }
                    continue;
                }
            }
            T dv = T(0);
            for (int j = j0; j < j1; ++j) {
                T ro[DIM], rn[DIM];
#pragma unroll
                for (int e = 0; e < DIM; ++e) {
                    ro[e] = xp[R.at(e, j)];
                    rn[e] = teleport(ro[e] + d[e], P);
                }
                dv += potVk<T, DIM, POT>(rn, P) - potVk<T, DIM, POT>(ro, P);
            }
            s_red[it] = dv;
        }
        if (!grouped) {
        __syncthreads();
        for (int q = tid; q < np; q += nt) {
            T sum = T(0);
            for (int k = 0; k < nseg; ++k) sum += s_red[q * nseg + k];
            const T lw = -prm[a.part_chain[tl.q0 + q]].tau * sum;
            const int acc = metropolis(ex(lw), s_u[q]) ? 1 : 0;
            s_flag[q] = acc;
            s_cacc[q] += acc;
        }
        __syncthreads();

        for (int it = tid, q = q_first, k = k_first, r = 0; it < items;
             it += nt, q += dq + (k + dk >= nseg ? 1 : 0), k = (k + dk >= nseg) ? k + dk - nseg : k + dk, ++r) {
            if (!s_flag[q]) continue;
            const int j0 = k * ell, j1 = min(M, j0 + ell);
            const Prm<T> P = prm[a.part_chain[tl.q0 + q]];
            T *xp = sp + (size_t)q * pstride;
            if constexpr (kF32) {
                if (aligned) {
                    T *x0 = xp + k * (MAXL + 1);
                    T d[DIM];
#pragma unroll
                    for (int e = 0; e < DIM; ++e) d[e] = s_d[q * DIM + e];
                    if (!((wrap >> r) & 1u)) {
#pragma unroll
                        for (int jj = 0; jj < MAXL; ++jj)
#pragma unroll
                            for (int e = 0; e < DIM; ++e) x0[e * R.Mp + jj] += d[e];
                    } else {
#pragma unroll
                        for (int jj = 0; jj < MAXL; ++jj)
#pragma unroll
                            for (int e = 0; e < DIM; ++e) x0[e * R.Mp + jj] = teleport(x0[e * R.Mp + jj] + d[e], P);
                    }
                    continue;
                }
            }
            for (int j = j0; j < j1; ++j) {
#pragma unroll
                for (int e = 0; e < DIM; ++e) xp[R.at(e, j)] = teleport(xp[R.at(e, j)] + s_d[q * DIM + e], P);
            }
        }
        } // !grouped
        } // com_on
        __syncthreads();

        // estimators (src/measurement.jl:45-55, 92-122): T sums over a chunk
        // of `ell` beads, FP64 accumulation across chunks, sweeps and particles
        if (meas_now) {
#pragma unroll
            for (int r = 0; r < kItemsPerThread; ++r) {
                const int it = tid + r * nt;
                if (it < items) {
                    const int q = slot_q[r], k = slot_k[r];
                    const int j0 = k * ell, j1 = min(M, j0 + ell);
                    const int c = a.part_chain[tl.q0 + q];
                    const Prm<T> P = prm[c];
                    const T *xp = sp + (size_t)q * pstride;
                    unsigned *hist = s_hist + s_slot[q] * rbins;
                    T l = T(0), v = T(0), g = T(0);
                    // link (r0 -> r1) of the kinetic sum, and the potential, virial and
                    // histogram terms of bead r0
                    // (FP32 in the harmonic trap: v sums |r|^2, scaled to V and r.grad V after
                    // the chunk)
                    constexpr bool kSumR2 = kF32 && POT == kPotHarmonic;
                    auto bead = [&](const T *r0, const T *r1) {
                        if constexpr (kF32) {
#pragma unroll
                            for (int e = 0; e < DIM; ++e) {
                                const T dx = mindist(r0[e], r1[e], P);
                                l = fma(dx, dx, l);
                            }
                        } else {
                            T dd = T(0);
#pragma unroll
                            for (int e = 0; e < DIM; ++e) {
                                const T dx = mindist(r0[e], r1[e], P);
                                dd += dx * dx;
                            }
                            l += dd;
                        }
                        // |r|^2 of the bead: the radial histogram's radius and, in the FP32
                        // harmonic trap, the potential sum's term
                        T rr = r0[0] * r0[0];
                        if (DIM > 1) rr += r0[1] * r0[1];
                        if constexpr (kSumR2) {
                            v += rr;
                        } else {
                            v += potVk<T, DIM, POT>(r0, P);
                            g += rgradVk<T, DIM, POT>(r0, P);
                        }
                        if (rbins > 0) {
                            const T qb = sq(rr) * (T)a.radial_winv;
                            if (qb < (T)rbins) atomicAdd(&hist[(int)qb], 1u);
                        }
                        if (a.cart_bins > 0) {
                            // the reference's Density: floor((r + L)/bin) used as a 1-based index
                            const double bw = (2.0 * a.chains[c].L) / a.cart_bins;
                            long long idx = 0, stride = 1;
                            bool ok = true;
#pragma unroll
                            for (int e = 0; e < DIM; ++e) {
                                const long long ib = (long long)floor(((double)r0[e] + a.chains[c].L) / bw);
                                ok = ok && ib > 0 && ib < a.cart_bins + 1;
                                idx += (ib - 1) * stride;
                                stride *= a.cart_bins;
                            }
                            if (ok) atomicAdd(&a.cart[(size_t)c * stride + idx], 1ull);
                        }
                    };
                    bool done = false;
                    if constexpr (kF32) {
                        if (aligned && a.cart_bins == 0) {
                            // one skew block; the link out of its last bead goes to slice
                            // (k + 1) MAXL, or to slice 0 from the last block
                            const T *x0 = xp + k * (MAXL + 1);
                            const T *xn = (k + 1 == nseg) ? xp : x0 + (MAXL + 1);
                            T r0[DIM];
#pragma unroll
                            for (int e = 0; e < DIM; ++e) r0[e] = x0[e * R.Mp];
#pragma unroll
                            for (int jj = 0; jj < MAXL; ++jj) {
                                T r1[DIM];
#pragma unroll
                                for (int e = 0; e < DIM; ++e) r1[e] = (jj + 1 < MAXL) ? x0[e * R.Mp + jj + 1] : xn[e * R.Mp];
                                bead(r0, r1);
#pragma unroll
                                for (int e = 0; e < DIM; ++e) r0[e] = r1[e];
                            }
                            done = true;
                        }
                    }
                    if (!done) {
                        T r0[DIM];
#pragma unroll
                        for (int e = 0; e < DIM; ++e) r0[e] = xp[R.at(e, j0)];
                        for (int j = j0; j < j1; ++j) {
                            const int jn = (j + 1 == M) ? 0 : j + 1;
                            T r1[DIM];
#pragma unroll
                            for (int e = 0; e < DIM; ++e) r1[e] = xp[R.at(e, jn)];
                            bead(r0, r1);
#pragma unroll
                            for (int e = 0; e < DIM; ++e) r0[e] = r1[e];
                        }
                    }
                    if constexpr (kSumR2) {
                        g = P.w2 * v;
                        v = P.halfw2 * v;
                    }
                    mL[r] += (double)l;
                    mV[r] += (double)v;
                    mR[r] += (double)g;
                }
            }
            ++nmeas;
        }
    }
    __syncthreads();

    for (int i = tid; i < nrow * M; i += nt) {
        const int row = i / M, j = i - row * M, q = row / DIM, d = row - q * DIM;
        gpos[i] = sp[q * pstride + d * R.Mp + j + (j >> R.sh)];
    }
    // radial histogram of the launch -> global counts (one atomic per non-empty bin)
    if (a.measure && rbins > 0) {
        for (int i = tid; i < tl.nslot * rbins; i += nt) {
            const unsigned cnt = s_hist[i];
            if (cnt == 0u) continue;
            const int slot = i / rbins, b = i - slot * rbins;
            int q = 0;
            while (q + 1 < np && s_slot[q] != slot) ++q;
            const int c = a.part_chain[tl.q0 + q];
            atomicAdd(&a.radial[(size_t)c * rbins + b], (unsigned long long)cnt);
        }
    }
    __syncthreads();
    if (a.measure) {
        double *red3 = reinterpret_cast<double *>(smem);
#pragma unroll
        for (int r = 0; r < kItemsPerThread; ++r) {
            const int it = tid + r * nt;
            if (it < items) {
                red3[it * 3 + 0] = mL[r];
                red3[it * 3 + 1] = mV[r];
                red3[it * 3 + 2] = mR[r];
            }
        }
        __syncthreads();
        for (int q = tid; q < np; q += nt) {
            double l = 0.0, v = 0.0, g = 0.0;
            for (int k = 0; k < nseg; ++k) {
                l += red3[(q * nseg + k) * 3 + 0];
                v += red3[(q * nseg + k) * 3 + 1];
                g += red3[(q * nseg + k) * 3 + 2];
            }
            double *pa = a.part_acc + (size_t)(tl.q0 + q) * 4;
            pa[0] = l;
            pa[1] = v;
            pa[2] = g;
            pa[3] = (double)nmeas;
        }
    }
    for (int q = tid; q < np; q += nt) {
        a.part_lacc[tl.q0 + q] = s_lacc[q];
        a.part_cacc[tl.q0 + q] = s_cacc[q];
    }
}
// Deterministic workgroup sum: butterfly within each wave (lane 0's order is
// fixed), wave totals through LDS, then every thread adds the wave totals in
// wave order.  `wbuf` holds 16 doubles.  Every thread gets the result.
__device__ inline double block_sum(double v, double *wbuf)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) wbuf[wave] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += wbuf[w];
    __syncthreads();
    return s;
}


// Squared periodic distance of two beads.
template <typename T, int DIM> __device__ __forceinline__ T dist2(const T *x, const T *y, const Prm<T> &P)
{
    const T d0 = mindist(x[0], y[0], P);
    T r2 = d0 * d0;
    if (DIM > 1) {
        const T d1 = mindist(x[1], y[1], P);
        r2 += d1 * d1;
    }
    return r2;
}

// Neighbour scan of the pair sums (whole-chain kernel, hooks, estimators).
// The candidates of query point q at slice j are the particles of the 9
// stencil cells around q, in the stencil order of bin_neighbors and the
// member order of each cell (CELLS), or all particles in index order, each
// counted as often as its cell appears in the stencil when nbins < 3; a
// candidate is a neighbour when its periodic distance to q is <= h (the
// neighbour set of find_nns, src/nearest_neighbours.jl:106-154).  Candidates
// are split over the G lanes of a group: lane g takes g, g + G, ...  For NQ
// query points at once (the new and the old bead of a move) a lane loads
// the ids, then the beads at slice j, of R candidates before it tests any,
// then the beads at slice jn of the neighbours only.  A query whose cell is
// query 0's shares query 0's list and loads.  Each query sees its candidates
// in the order of a one-at-a-time scan.  term(u, rk, rkn, r2, mult) is called
// for every neighbour k != n of query u, with rk = r_kj, r2 = |r_kj - q_u|^2
// and rkn = r_k,jn.
template <typename T, int DIM, bool CELLS, int NQ, int R, typename PathT, typename F>
__device__ __forceinline__ void scan_pairs(const PathT &path, const Cells &C, int N, int nbins, const Prm<T> &P,
                                           int n, int j, int jn, const T (&q)[NQ][DIM], int g, int G, F &&term)
{
    constexpr int NS = (DIM == 1) ? 3 : 9;
    if constexpr (CELLS) {
        int c0[NQ], lst[NQ], bx[NQ], by[NQ]; // lst[u]: the list query u reads (its own, or query 0's)
#pragma unroll
        for (int u = 0; u < NQ; ++u) {
            bin_xy<T, DIM>(q[u], P, C.nbins, bx[u], by[u]);
            c0[u] = bx[u] + C.nbins * by[u];
            lst[u] = (u > 0 && c0[u] == c0[0]) ? 0 : u;
        }
        // per stencil slot s: prefix pre[s] of the candidate counts and the offset moff[s]
        // with C.mem[moff[s] + idx] = member idx - pre[s] of cell s, so that a candidate
        // index picks its member with one select per slot
        int moff[NQ][NS], pre[NQ][NS + 1], total = 0;
#pragma unroll
        for (int u = 0; u < NQ; ++u) {
            pre[u][0] = 0;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int cell = stencil_xy(bx[u], by[u], s, C.nbins, DIM);
                moff[u][s] = (j * C.ncell + cell) * C.cap;
                pre[u][s + 1] = (lst[u] == u) ? C.count(j, cell) : 0;
            }
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                moff[u][s] -= pre[u][s];
                pre[u][s + 1] += pre[u][s];
            }
            total = max(total, pre[u][NS]);
        }
        for (int i0 = g; i0 < total; i0 += G * R) {
            int k[NQ][R];
            T rk[NQ][R][DIM];
#pragma unroll
            for (int u = 0; u < NQ; ++u) {
#pragma unroll
                for (int r = 0; r < R; ++r) k[u][r] = -1;
                if (lst[u] != u) continue; // a query sharing query 0's list reads query 0's candidates
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int idx = i0 + r * G;
                    int off = moff[u][0];
#pragma unroll
                    for (int s = 1; s < NS; ++s) off = (idx >= pre[u][s]) ? moff[u][s] : off;
                    k[u][r] = (idx < pre[u][NS]) ? (int)C.mem[off + idx] : -1;
                }
            }
#pragma unroll
            for (int u = 0; u < NQ; ++u)
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (lst[u] == u) path.get(k[u][r] < 0 ? n : k[u][r], j, rk[u][r]);
            T r2[NQ][R];
            bool near[NQ][R];
#pragma unroll
            for (int u = 0; u < NQ; ++u) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int l = lst[u];
                    const int kk = l == 0 ? k[0][r] : k[u][r];
                    r2[u][r] = dist2<T, DIM>(l == 0 ? rk[0][r] : rk[u][r], q[u], P);
                    near[u][r] = kk >= 0 && kk != n && r2[u][r] <= P.h2;
                }
            }
            T rkn[NQ][R][DIM];
#pragma unroll
            for (int u = 0; u < NQ; ++u) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    bool need = false;
#pragma unroll
                    for (int v = 0; v < NQ; ++v) need = need || (lst[v] == u && near[v][r]);
                    if (need) path.get(C.after(k[u][r], jn), jn, rkn[u][r]);
                }
            }
#pragma unroll
            for (int u = 0; u < NQ; ++u) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (near[u][r])
                        term(u, lst[u] == 0 ? k[0][r] : k[u][r], lst[u] == 0 ? rk[0][r] : rk[u][r],
                             lst[u] == 0 ? rkn[0][r] : rkn[u][r], r2[u][r], T(1));
            }
        }
    } else {
        int c0[NQ];
#pragma unroll
        for (int u = 0; u < NQ; ++u) c0[u] = (nbins < 3) ? cell_of<T, DIM>(q[u], P, nbins) : 0;
        for (int k0 = g; k0 < N; k0 += G * R) {
            T rk[R][DIM], rkn[R][DIM], r2[NQ][R], mult[NQ][R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int kk = k0 + r * G;
                path.get(kk < N ? kk : n, j, rk[r]);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int kk = k0 + r * G;
                const bool cand = kk < N && kk != n;
                const int ck = (nbins < 3) ? cell_of<T, DIM>(rk[r], P, nbins) : 0;
                bool need = false;
#pragma unroll
                for (int u = 0; u < NQ; ++u) {
                    int mm = 1;
                    if (nbins < 3) {
                        mm = 0;
                        for (int s = 0; s < NS; ++s) mm += (stencil_cell(c0[u], s, nbins, DIM) == ck);
                    }
                    r2[u][r] = dist2<T, DIM>(rk[r], q[u], P);
                    mult[u][r] = (cand && mm > 0 && r2[u][r] <= P.h2) ? T(mm) : T(0);
                    need = need || mult[u][r] > T(0);
                }
                if (need) path.get(C.after(kk, jn), jn, rkn[r]);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
#pragma unroll
                for (int u = 0; u < NQ; ++u)
                    if (mult[u][r] > T(0)) term(u, k0 + r * G, rk[r], rkn[r], r2[u][r], mult[u][r]);
            }
        }
    }
}

// This lane's share of the pair link sum of particle n over link j -> j+1,
// with its beads there replaced by pj, pj1: pair_lnU for
// every neighbour found at slice j.
template <typename T, int DIM, bool CELLS, int TAB = -1, typename PathT>
__device__ T link_part(const PathT &path, const Cells &C, int N, int nbins, const Prm<T> &P, int n, int j,
                       const T *pj, const T *pj1, int g, int G)
{
    const int jn = (j + 1 == path.M) ? 0 : j + 1;
    T q[1][DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) q[0][e] = pj[e];
    T acc = T(0);
    scan_pairs<T, DIM, CELLS, 1, 4>(path, C, N, nbins, P, n, j, jn, q, g, G,
                                    [&](int, int, const T *rk, const T *rkn, T r2, T mult) {
                                        const T s2 = dist2<T, DIM>(rkn, pj1, P);
                                        acc += mult * pair_lnU<T, DIM, TAB>(rk, pj, rkn, pj1, r2, s2, P);
                                    });
    return acc;
}

// This lane's share of (link sum with beads n0, n1) - (link sum with o0, o1)
// for particle n over link j: both scans in one batched pass.
template <typename T, int DIM, bool CELLS, int R, int TAB = -1, typename PathT>
__device__ __forceinline__ T link_delta(const PathT &path, const Cells &C, int N, int nbins, const Prm<T> &P, int n,
                                        int j, int jn, const T *n0, const T *n1, const T *o0, const T *o1, int g,
                                        int G)
{
    if (CELLS && TAB == 1 && G >= 4) {
        // (tabulated propagator, whose 512-thread kernel has the registers for it) half of
        // the group scans for the new beads, half for the old ones, each with twice the
        // candidates per lane; the group's sum of the returned shares is the difference
        const int h = G >> 1, u = g >= h ? 1 : 0;
        const T *a0 = u ? o0 : n0, *a1 = u ? o1 : n1;
        T q[1][DIM];
#pragma unroll
        for (int e = 0; e < DIM; ++e) q[0][e] = a0[e];
        T acc = T(0);
        scan_pairs<T, DIM, CELLS, 1, 2 * R>(path, C, N, nbins, P, n, j, jn, q, g - u * h, h,
                                            [&](int, int, const T *rk, const T *rkn, T r2, T mult) {
                                                const T s2 = dist2<T, DIM>(rkn, a1, P);
                                                acc += mult * pair_lnU<T, DIM, TAB>(rk, a0, rkn, a1, r2, s2, P);
                                            });
        return u ? -acc : acc;
    }
    T q[2][DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
        q[0][e] = n0[e];
        q[1][e] = o0[e];
    }
    T acc[2] = {T(0), T(0)};
    scan_pairs<T, DIM, CELLS, 2, R>(path, C, N, nbins, P, n, j, jn, q, g, G,
                                    [&](int u, int, const T *rk, const T *rkn, T r2, T mult) {
                                        const T *qn = u == 0 ? n1 : o1;
                                        const T s2 = dist2<T, DIM>(rkn, qn, P);
                                        acc[u] += mult * pair_lnU<T, DIM, TAB>(rk, u == 0 ? n0 : o0, rkn, qn, r2, s2, P);
                                    });
    return acc[0] - acc[1];
}

// How many times the neighbour scan of a bead at a lists a bead at b: the
// stencil slots of a's cell holding b's cell (several when nbins < 3), if b
// lies within the cut-off; scan_pairs' multiplicity for one pair.
template <typename T, int DIM> __device__ __forceinline__ T pair_mult(const T *a, const T *b, const Prm<T> &P, int nbins)
{
    if (!(dist2<T, DIM>(a, b, P) <= P.h2)) return T(0);
    if (nbins >= 3) return T(1);
    const int ca = cell_of<T, DIM>(a, P, nbins), cb = cell_of<T, DIM>(b, P, nbins);
    int mm = 0;
    for (int s = 0; s < ((DIM == 1) ? 3 : 9); ++s) mm += (stencil_cell(ca, s, nbins, DIM) == cb);
    return T(mm);
}

template <typename T, int DIM, bool CELLS, typename PathT>
__device__ __forceinline__ T link_sum(const PathT &path, const Cells &C, int N, int nbins, const Prm<T> &P, int n,
                                      int j, const T *pj, const T *pj1)
{
    return link_part<T, DIM, CELLS>(path, C, N, nbins, P, n, j, pj, pj1, 0, 1);
}

// The hard-sphere test of hardspherelevy! and move_polymer! (src/updates/
// helper.jl:166-170, 385-389): whether the nearest bead at slice j among the
// stencil candidates of q (find_nn), other than particle n's, is closer than
// a.  With a <= h every bead closer than a is a stencil candidate, so this is
// whether any such bead exists.  The 64 lanes of the calling wave split the
// candidates; every lane of the wave must call it, and all get the same answer.
template <typename T, int DIM, bool CELLS, typename PathT>
__device__ bool hs_near_wave(const PathT &path, const Cells &C, int N, const Prm<T> &P, int n, int j, const T *q,
                             int lane)
{
    bool near = false;
    if constexpr (CELLS) {
        // all stencil counts first (independent loads), then the flattened candidate list
        constexpr int NS = (DIM == 1) ? 3 : 9;
        int bx, by;
        bin_xy<T, DIM>(q, P, C.nbins, bx, by);
        int cs[NS], pre[NS + 1];
        pre[0] = 0;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            cs[s] = stencil_xy(bx, by, s, C.nbins, DIM);
            pre[s + 1] = C.count(j, cs[s]);
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) pre[s + 1] += pre[s];
        for (int idx = lane; idx < pre[NS]; idx += 64) {
            int cell = cs[0], base = 0;
#pragma unroll
            for (int s = 1; s < NS; ++s) {
                if (idx >= pre[s]) {
                    cell = cs[s];
                    base = pre[s];
                }
            }
            const int k = C.members(j, cell)[idx - base];
            if (k == n) continue;
            T rk[DIM];
            path.get(k, j, rk);
            near = near || dist2<T, DIM>(rk, q, P) < P.hs2;
        }
    } else {
        for (int k = lane; k < N; k += 64) {
            if (k == n) continue;
            T rk[DIM];
            path.get(k, j, rk);
            near = near || dist2<T, DIM>(rk, q, P) < P.hs2;
        }
    }
    return __ballot(near) != 0ull;
}

// Lane g's share of the same test when kHsLanes lanes test one bead together: the
// stencil cells s = g, g + G, ... (with cell lists) or the particles
// k = g, g + G, ... (without); the bead is too close when any lane's share
// is.  With cell lists a lane loads the counts and first 8 member ids (one
// 16-byte load; lists are 16-byte aligned) of its cells first, then the
// beads of their first 4 members together, then those of members 4..7 of
// the cells that have them; members past the 8th, rare at the densities
// where cell lists are used, one at a time.
constexpr int kHsLanes = 4;
template <typename T, int DIM, bool CELLS, typename PathT>
__device__ bool hs_near_part(const PathT &path, const Cells &C, int N, const Prm<T> &P, int n, int j, const T *q,
                             int g)
{
    constexpr int G = kHsLanes;
    if constexpr (CELLS) {
        constexpr int NS = (DIM == 1) ? 3 : 9, NL = (NS + G - 1) / G; // cells per lane
        // stencil offsets of bin_neighbors' order, dx + 1 and dy + 1 in 2 bits per cell
        constexpr unsigned kDx = (DIM == 1) ? (1u | (0u << 2) | (2u << 4))
                                            : (1u | (0u << 2) | (1u << 4) | (2u << 6) | (0u << 8) | (2u << 10) |
                                               (0u << 12) | (1u << 14) | (2u << 16));
        constexpr unsigned kDy = (1u | (2u << 2) | (2u << 4) | (2u << 6) | (1u << 8) | (1u << 10) | (0u << 12) |
                                  (0u << 14) | (0u << 16));
        int bx, by;
        bin_xy<T, DIM>(q, P, C.nbins, bx, by);
        int cnt[NL], cell[NL];
        uint4 ids[NL];
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const int s = g + i * G;
            cnt[i] = 0;
            cell[i] = 0;
            ids[i] = make_uint4(0u, 0u, 0u, 0u);
            if (s < NS) {
                const int cx = wrap_bin(bx + (int)((kDx >> (2 * s)) & 3u) - 1, C.nbins);
                const int cy = DIM == 1 ? 0 : wrap_bin(by + (int)((kDy >> (2 * s)) & 3u) - 1, C.nbins);
                cell[i] = cx + C.nbins * cy;
                cnt[i] = C.count(j, cell[i]);
                ids[i] = *reinterpret_cast<const uint4 *>(C.members(j, cell[i]));
            }
        }
        // members 4h .. 4h + 3 of this lane's cells
        auto batch = [&](int h) {
            int k[NL][4];
            T rk[NL][4][DIM];
#pragma unroll
            for (int i = 0; i < NL; ++i) {
                const unsigned w0 = h == 0 ? ids[i].x : ids[i].z, w1 = h == 0 ? ids[i].y : ids[i].w;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const unsigned w = t < 2 ? w0 : w1;
                    k[i][t] = (4 * h + t < cnt[i]) ? (int)((w >> (16 * (t & 1))) & 0xffffu) : n;
                    if (k[i][t] != n) path.get(k[i][t], j, rk[i][t]);
                }
            }
            bool near = false;
#pragma unroll
            for (int i = 0; i < NL; ++i)
#pragma unroll
                for (int t = 0; t < 4; ++t) near = near || (k[i][t] != n && dist2<T, DIM>(rk[i][t], q, P) < P.hs2);
            return near;
        };
        if (batch(0)) return true;
        bool more = false;
#pragma unroll
        for (int i = 0; i < NL; ++i) more = more || cnt[i] > 4;
        if (more && batch(1)) return true;
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            if (cnt[i] > 8) {
                const unsigned short *mem = C.members(j, cell[i]);
                for (int t = 8; t < cnt[i]; ++t) {
                    const int kk = mem[t];
                    if (kk == n) continue;
                    T r[DIM];
                    path.get(kk, j, r);
                    if (dist2<T, DIM>(r, q, P) < P.hs2) return true;
                }
            }
        }
    } else {
        for (int k = g; k < N; k += G) {
            if (k == n) continue;
            T rk[DIM];
            path.get(k, j, rk);
            if (dist2<T, DIM>(rk, q, P) < P.hs2) return true;
        }
    }
    return false;
}

// Pair energy of bead (i, j) per unit coupling g: its share of S_U / g
// (each unordered pair counted from both ends with weight 1/2) and of
// (r . grad S_U) / g.  Kept free of g so that dS/dg is defined at g = 0.
template <typename T, int DIM, bool CELLS, typename PathT>
__device__ inline void pair_energy_bead(const PathT &path, const Cells &C, int N, int nbins, const Prm<T> &P,
                                        double inv_s2, int i, int j, double &su, double &rg, int g = 0, int G = 1)
{
    const int M = path.M;
    const int jn = (j + 1 == M) ? 0 : j + 1;
    T ri[DIM], rin[DIM];
    path.get(i, j, ri);
    path.get(C.after(i, jn), jn, rin);
    su = 0.0;
    rg = 0.0;
    T q[1][DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) q[0][e] = ri[e];
    scan_pairs<T, DIM, CELLS, 1, 4>(path, C, N, nbins, P, i, j, jn, q, g, G,
                                    [&](int, int, const T *, const T *rkn, T r2, T mult) {
                                        const T s2 = dist2<T, DIM>(rkn, rin, P);
                                        const double ea = (double)ex(-r2 * P.inv2s2), eb = (double)ex(-s2 * P.inv2s2);
                                        const double w = 0.5 * (0.5 * (double)P.tau) * (double)mult;
                                        su += w * (ea + eb);
                                        rg += w * (-((double)r2 * inv_s2) * ea - ((double)s2 * inv_s2) * eb);
                                    });
}

// Per-measurement whole-chain sums -> the kNQ quantities.
struct ChainMeas {
    double link, V, rg, su, rgp, r2w;
};

template <typename T, int DIM, bool CELLS, typename PathT>
__device__ inline ChainMeas measure_chain(const PathT &path, const Cells &C, const ChainDev &ch, const Prm<T> &P,
                                          double *wbuf, unsigned long long *radial, int radial_bins,
                                          double radial_winv, unsigned long long *cart, int cart_bins, int c,
                                          int G = 1, const int *nxt = nullptr, int b0 = 0, int b1 = -1)
{
    const int N = ch.N, M = ch.M;
    if (b1 < 0) b1 = N * M;
    const double inv_s2 = (ch.pair == kPairGaussian) ? 1.0 / (ch.sigma * ch.sigma) : 0.0;
    double l = 0.0, v = 0.0, g = 0.0, su = 0.0, rgp = 0.0, r2w = 0.0;
    // bead b = (slice j, particle i) per group of G lanes; the pair scan is split over
    // the lanes, the single-bead terms are taken by lane 0 of the group
    const int lg = threadIdx.x & (G - 1), ngroups = blockDim.x / G;
    for (int b = b0 + (int)threadIdx.x / G; b < b1; b += ngroups) {
        const int j = b / N, i = b - j * N;
        if (ch.pair == kPairGaussian) { // no pair energy for the tabulated propagator
            double a, bb;
            pair_energy_bead<T, DIM, CELLS>(path, C, N, ch.nbins, P, inv_s2, i, j, a, bb, lg, G);
            su += a;
            rgp += bb;
        }
        if (lg != 0) continue;
        const int jn = (j + 1 == M) ? 0 : j + 1;
        T r0[DIM], r1[DIM];
        path.get(i, j, r0);
        path.get((jn == 0 && nxt) ? nxt[i] : i, jn, r1); // the next bead along the permutation cycle
        T dd = T(0);
#pragma unroll
        for (int e = 0; e < DIM; ++e) {
            const T dx = mindist(r0[e], r1[e], P);
            dd += dx * dx;
        }
        l += (double)dd;
        v += (double)potV<T, DIM>(r0, P);
        g += (double)rgradV<T, DIM>(r0, P);
        if (ch.pot == kPotHarmonic) r2w += (double)dVdomega<T, DIM>(r0, ch.omega);
        if (radial_bins > 0) {
            T rr = r0[0] * r0[0];
            if (DIM > 1) rr += r0[1] * r0[1];
            const double qb = (double)sq(rr) * radial_winv;
            if (qb < (double)radial_bins) atomicAdd(&radial[(size_t)c * radial_bins + (int)qb], 1ull);
        }
        if (cart_bins > 0) {
            const double bw = (2.0 * ch.L) / cart_bins;
            long long idx = 0, stride = 1;
            bool ok = true;
#pragma unroll
            for (int e = 0; e < DIM; ++e) {
                const long long ib = (long long)floor(((double)r0[e] + ch.L) / bw);
                ok = ok && ib > 0 && ib < cart_bins + 1;
                idx += (ib - 1) * stride;
                stride *= cart_bins;
            }
            if (ok) atomicAdd(&cart[(size_t)c * stride + idx], 1ull);
        }
    }
    ChainMeas r;
    r.link = block_sum(l, wbuf);
    r.V = block_sum(v, wbuf);
    r.rg = block_sum(g, wbuf);
    r.su = block_sum(su, wbuf);
    r.rgp = block_sum(rgp, wbuf);
    r.r2w = block_sum(r2w, wbuf);
    return r;
}

// Fold one measurement of a whole chain into the kNQ sums.
__device__ inline void fold_meas(const ChainMeas &m, const ChainDev &ch, double *q)
{
    const double tau = ch.tau, M = ch.M;
    const double Epair = ch.g * m.su / ch.beta;
    const double E = ((double)(ch.dim * ch.N) / (2.0 * tau) - m.link / (((4.0 * ch.lambda) * (tau * tau)) * M)) +
                     m.V / M + Epair;
    const double Ev = m.rg / (2.0 * M) + m.V / M + Epair + ch.g * m.rgp / (2.0 * ch.beta);
    const double Sw = tau * m.r2w;
    const double Sg = m.su;
    const double dEw = m.r2w / M;
    const double dEg = m.su / ch.beta;
    q[qN] += 1.0;
    q[qE] += E;
    q[qEv] += Ev;
    q[qEpair] += Epair;
    q[qSw] += Sw;
    q[qSg] += Sg;
    q[qESw] += E * Sw;
    q[qESg] += E * Sg;
    q[qdEw] += dEw;
    q[qdEg] += dEg;
    q[qE2] += E * E;
    q[qWhole] = 1.0;
}

// LDS of k_sweep_whole: two buffers of new beads, one of bridge normals and
// one of a staging move's old beads (M x DIM T each), Delta V per bead and Delta lnU per link (M T each), accept
// flags (M ints), two sets of 16 wave partials (double), 16 doubles for block
// sums, kNQ doubles of estimator sums.
__host__ __device__ inline size_t whole_lds_bytes(int M, int dim, int tsize, int nperm = 0)
{
    size_t o = 4 * align16((size_t)M * dim * tsize);
    o += 2 * align16((size_t)M * tsize);
    o += align16((size_t)M * 4);
    o += align16(32 * sizeof(double));
    o += align16(16 * sizeof(double));
    o += align16(kNQ * sizeof(double));
    // chains with swap moves: next[] and the starts and ends of a swap try (nperm = N)
    o += align16((size_t)nperm * 4) + 2 * align16((size_t)nperm * dim * tsize) + align16(8 * 4);
    return o;
}
__host__ __device__ inline size_t whole_lds_bytes(const ChainDev &h, int tsize)
{
    return whole_lds_bytes(h.M, h.dim, tsize, h.swaps > 0 ? h.N : 0);
}

// Free-particle propagator over m links up to its normalisation,
// exp(lnK) = exp(-|x - y|^2 / (4 lambda m tau)) with the minimum image
// (src/propagator.jl:16-19): the weights of the swap partner choice.
template <typename T, int DIM> __device__ __forceinline__ T rho0(const T *x, const T *y, const Prm<T> &P, int m)
{
    T dd = T(0);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const T r = mindist(x[d], y[d], P);
        dd += r * r;
    }
    return ex(-dd / ((T(4) * P.lambda) * ((T)m * P.tau)));
}

// LDS that holds a whole chain's positions (chains whose path stays in LDS).
__host__ __device__ inline size_t whole_path_bytes(const ChainDev &h, int tsize)
{
    return align16((size_t)h.N * h.M * h.dim * tsize);
}

// One workgroup per chain with a pair action; positions slice-major in HBM
// (LDSR: copied into LDS for the launch when they fit, for all-pairs chains).
// Staging, for each particle in order: (A1) one thread per interior bead pair
// draws the bridge normals, (A2) one thread per segment runs the bridge
// recursion into LDS, (B) a group of G lanes per link loads the old beads,
// takes Delta V of the link's first bead and the link's new-minus-old pair
// sum (batched neighbour scan), (C) one thread per segment decides, summing
// in the bead and link order of the oracle, (D) one thread per slice commits
// beads and cell lists.  New beads are double-buffered by particle parity.
// Centre of mass, for each particle in order: the link groups form the moved
// beads on the fly and add Delta V and the pair difference, one workgroup
// reduction gives a uniform decision, and the commit: two barriers.
// Workgroup size: 1024 threads, or 512 with the tabulated propagator, whose
// FP64 table lookups need more than the 128 VGPRs of a 1024-thread launch.
template <bool TAB> constexpr int whole_threads() { return TAB ? 512 : 1024; }
template <typename T, int DIM, bool CELLS, bool LDSR, bool TAB, bool TPERM = false>
__global__ void __launch_bounds__(whole_threads<TAB>()) k_sweep_whole(SweepArgs a)
{
    static_assert(!(CELLS && LDSR), "LDS-resident paths are for all-pairs chains");
    extern __shared__ __align__(16) unsigned char smem[];
    const int c = a.whole[blockIdx.x];
    const ChainDev ch = a.chains[c];
    const Prm<T> P = reinterpret_cast<const Prm<T> *>(a.prm)[c];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int N = ch.N, M = ch.M, ell = ch.ell, nseg = ch.nseg;
    T *const gpath = reinterpret_cast<T *>(a.pos) + ch.pos_off;
    T *const lpath = reinterpret_cast<T *>(smem + whole_lds_bytes(ch, sizeof(T)));
    if constexpr (LDSR) {
        for (int i = tid; i < N * M * DIM; i += nt) lpath[i] = gpath[i];
        __syncthreads();
    }
    const Path<T, DIM, true> path{LDSR ? lpath : gpath, M, N};
    Cells cells{a.cell_cnt + ch.cell_off, a.cell_mem + ch.mem_off, a.bins + ch.bin_off, M, ch.ncell, ch.cap,
                ch.nbins};
    const bool use_pair = (ch.pair != kPairNone) && ch.pair_in_accept;
    // hard spheres (tabulated-propagator chains only): s.ctr of run!, src/simulation.jl:31-32
    const bool hs = TAB && ch.hs_a > 0.0;
    const int hs_ctr = a.measure ? 1000 : 10000;
    const uint32_t sweep0 = a.rs->sweep + (uint32_t)a.sweep_off;
    const int G = a.link_group, ngroups = nt / G, g = tid & (G - 1), grp = tid / G;
    const int lane = tid & 63, wave = tid >> 6, nw = (nt + 63) >> 6;
    // candidates per lane and query loaded together (3 suits the cell-list scans of the
    // analytic pairs; more spills registers)
    constexpr int R = (CELLS && !TAB) ? 3 : 2;

    size_t o = 0;
    T *s_new0 = reinterpret_cast<T *>(smem + o);
    o += align16((size_t)M * DIM * sizeof(T));
    T *s_new1 = reinterpret_cast<T *>(smem + o);
    o += align16((size_t)M * DIM * sizeof(T));
    T *s_nrm = reinterpret_cast<T *>(smem + o);
    o += align16((size_t)M * DIM * sizeof(T));
    T *s_old = reinterpret_cast<T *>(smem + o); // old beads of the staging move, for the cell lists
    o += align16((size_t)M * DIM * sizeof(T));
    T *s_dvb = reinterpret_cast<T *>(smem + o);
    o += align16((size_t)M * sizeof(T));
    T *s_du = reinterpret_cast<T *>(smem + o);
    o += align16((size_t)M * sizeof(T));
    int *s_flag = reinterpret_cast<int *>(smem + o);
    o += align16((size_t)M * 4);
    double *s_wsum = reinterpret_cast<double *>(smem + o);
    o += align16(32 * sizeof(double));
    double *s_buf = reinterpret_cast<double *>(smem + o);
    o += align16(16 * sizeof(double));
    double *s_q = reinterpret_cast<double *>(smem + o); // estimator sums of the launch
    o += align16(kNQ * sizeof(double));
    // permutation cycles (chains with swap moves): next[] in LDS for the launch.  Chains with the
    // tabulated propagator and swaps run in their own instantiation (TPERM), so that the others
    // leave the permutation code out
    constexpr bool kPerm = !TAB || TPERM;
    const bool perm = kPerm && ch.swaps > 0;
    const int np_ = perm ? N : 0;
    int *s_next = reinterpret_cast<int *>(smem + o);
    o += align16((size_t)np_ * 4);
    T *s_E = reinterpret_cast<T *>(smem + o); // a swap try's end beads E_i
    o += align16((size_t)np_ * DIM * sizeof(T));
    T *s_B = reinterpret_cast<T *>(smem + o); // ... and start beads r_i(j0)
    o += align16((size_t)np_ * DIM * sizeof(T));
    int *s_sw = reinterpret_cast<int *>(smem + o);
    int *const gnext = a.next + ch.part_base;
    for (int i = tid; i < np_; i += nt) s_next[i] = gnext[i];
    // the particle holding bead u of the window that starts on particle pp:
    // u < M on pp, u >= M on next[pp] (pp itself without permutations)
    auto owner = [&](int pp, int u) { return (perm && u >= M) ? s_next[pp] : pp; };
    if (perm) cells.next = s_next; // pair links across the ring's end follow the cycles

    if (tid < kNQ) s_q[tid] = 0.0;
    int lacc = 0, cacc = 0, ctry = 0, sacc = 0, stry = 0, overflow = 0;
    __syncthreads();
    const int hp = ell >> 1; // Philox blocks per segment: interior bead pairs (i, i + 1), i odd
    // sweeps until the next measured one: (sw + 1) % measure_every == 0 without a divide per sweep
    int meas_in = (int)((uint32_t)a.measure_every - 1u - sweep0 % (uint32_t)a.measure_every);

    for (int s = 0; s < a.nsweeps; ++s) {
        const uint32_t sw = sweep0 + (uint32_t)s;
        const bool meas_now = a.measure && !a.split && meas_in == 0;
        meas_in = (meas_in == 0) ? a.measure_every - 1 : meas_in - 1;
        // move weights: the passes run on the sweeps s with s % every == 0 (uniform per chain)
        const bool st_on = pass_on(ch.every_st, sw), com_on = pass_on(ch.every_com, sw);
        const bool sw_on = pass_on(ch.every_sw, sw);
        // ---- staging: particles in order, segments (and links) of one particle in parallel
        const U4 xo = draw(a.seed, 0u, sw, (uint32_t)c, kOffset); // one offset per chain and sweep
        const int off = (int)(((uint64_t)xo.x * (uint64_t)ell) >> 32);
        for (int p = 0; p < (st_on ? N : 0); ++p) {
            T *buf = (p & 1) ? s_new1 : s_new0;
            // (A1) normals of the interior beads, by slice
            for (int w = tid; w < nseg * hp; w += nt) {
                const int k = w / hp, i = 2 * (w - k * hp) + 1;
                const int m = min(ell, M - k * ell);
                if (i < m) {
                    int b = off + k * ell + i;
                    if (b >= M) b -= M;
                    const U4 xr = draw(a.seed, (uint32_t)(p * M + b), sw, (uint32_t)c, kLevy);
                    T n0, n1;
                    box_muller(xr.x, xr.y, n0, n1);
                    s_nrm[b * DIM] = n0;
                    if (DIM > 1) s_nrm[b * DIM + 1] = n1;
                    if (i + 1 < m) {
                        const int b1 = (b + 1 == M) ? 0 : b + 1;
                        box_muller(xr.z, xr.w, n0, n1);
                        s_nrm[b1 * DIM] = n0;
                        if (DIM > 1) s_nrm[b1 * DIM + 1] = n1;
                    }
                }
            }
            __syncthreads();
            // (A2) bridges with the first draws; with hard spheres s_flag[k] = m: no bead of
            // segment k found too close yet
            for (int k = tid; k < nseg; k += nt) {
                const int m = min(ell, M - k * ell);
                if (TAB && hs) s_flag[k] = m;
                if (m >= 2) {
                    const int us = off + k * ell; // window position of the start bead (< 2M)
                    const int start = us >= M ? us - M : us;
                    int jm = start + m;
                    if (jm >= M) jm -= M;
                    T cur[DIM], em[DIM];
                    path.get(owner(p, us), start, cur);
                    path.get(owner(p, us + m), jm, em);
#pragma unroll
                    for (int e = 0; e < DIM; ++e) em[e] = shift_end(cur[e], em[e], P);
                    for (int i = 1; i < m; ++i) {
                        int b = start + i;
                        if (b >= M) b -= M;
                        const T alpha = bridge_alpha<T>(m, i);
                        const T sd = bridge_sd(alpha, P);
#pragma unroll
                        for (int e = 0; e < DIM; ++e) {
                            cur[e] = bridge_step(cur[e], em[e], alpha, s_nrm[b * DIM + e], sd);
                            buf[b * DIM + e] = teleport(cur[e], P);
                        }
                    }
                }
            }
            if (TAB && hs) {
                // hardspherelevy! (src/updates/helper.jl:141-181).  The first draws of every
                // interior bead are tested in parallel, kHsLanes threads per bead, keeping in
                // s_flag the first bead of each segment that lies closer than a to another
                // particle's bead
                __syncthreads();
                for (int w = tid; w < M * kHsLanes; w += nt) {
                    const int j = w / kHsLanes, hg = w - j * kHsLanes;
                    int rel = j - off;
                    if (rel < 0) rel += M;
                    const int k = rel / ell, t = rel - k * ell;
                    const int m = min(ell, M - k * ell);
                    if (m >= 2 && t >= 1) {
                        T q[DIM];
#pragma unroll
                        for (int e = 0; e < DIM; ++e) q[e] = buf[j * DIM + e];
                        if (hs_near_part<T, DIM, CELLS>(path, cells, N, P, p, j, q, hg)) atomicMin(&s_flag[k], t);
                    }
                }
                __syncthreads();
                // segments with such a bead: one wave walks the bridge on from it.  The failing bead
                // is redrawn (draw t >= 2: kHardSphere | t << 8; its first draw is known to fail)
                // until it lies no closer than a; then the rest of the segment, built from its first
                // draws, is tested together (kHsLanes lanes per bead), and the walk goes on from the
                // first bead that fails, as a bead-by-bead walk would.  A segment that runs out of
                // draws is a rejected try (s_flag = -1) and keeps its old beads; the others get
                // s_flag = 0
                for (int k = wave; k < nseg; k += nw) {
                    const int m = min(ell, M - k * ell);
                    if (m < 2) continue;
                    int i0 = s_flag[k];
                    if (i0 >= m) {
                        if (lane == 0) s_flag[k] = 0;
                        continue;
                    }
                    const int us = off + k * ell; // window position of the start bead (< 2M)
                    const int start = us >= M ? us - M : us;
                    int jm = start + m;
                    if (jm >= M) jm -= M;
                    T cur[DIM], em[DIM];
                    path.get(owner(p, us), start, cur); // the segment's ends along the cycle
                    path.get(owner(p, us + m), jm, em);
#pragma unroll
                    for (int e = 0; e < DIM; ++e) em[e] = shift_end(cur[e], em[e], P);
                    // cur: bead i - 1 of the segment; advance(i): bead i from its first draw
                    auto advance = [&](T *x, int i) {
                        int bb = start + i;
                        if (bb >= M) bb -= M;
                        const T alpha = bridge_alpha<T>(m, i);
                        const T sd = bridge_sd(alpha, P);
#pragma unroll
                        for (int e = 0; e < DIM; ++e) x[e] = bridge_step(x[e], em[e], alpha, s_nrm[bb * DIM + e], sd);
                    };
                    for (int i = 1; i < i0; ++i) advance(cur, i); // the bridge up to the failing bead, as built above
                    bool failed = false;
                    while (i0 < m) {
                        // redraws of bead i0
                        int b = start + i0;
                        if (b >= M) b -= M;
                        {
                            const T alpha = bridge_alpha<T>(m, i0);
                            const T sd = bridge_sd(alpha, P);
                            bool placed = false;
                            for (int t = 2; t <= hs_ctr; ++t) {
                                const U4 xr = draw(a.seed, (uint32_t)(p * M + b), sw, (uint32_t)c,
                                                   (uint32_t)kHardSphere | ((uint32_t)t << 8));
                                T nr[2];
                                box_muller(xr.x, xr.y, nr[0], nr[1]);
                                T nx[DIM], tb[DIM];
#pragma unroll
                                for (int e = 0; e < DIM; ++e) {
                                    nx[e] = bridge_step(cur[e], em[e], alpha, nr[e], sd);
                                    tb[e] = teleport(nx[e], P);
                                }
                                if (!hs_near_wave<T, DIM, CELLS>(path, cells, N, P, p, b, tb, lane)) {
#pragma unroll
                                    for (int e = 0; e < DIM; ++e) {
                                        cur[e] = nx[e];
                                        if (lane == 0) buf[b * DIM + e] = tb[e];
                                    }
                                    placed = true;
                                    break;
                                }
                            }
                            if (!placed) {
                                failed = true;
                                break;
                            }
                        }
                        // beads i0 + 1 .. from their first draws, 64 / kHsLanes at a time
                        int next = m; // the first of them that fails
                        for (int c0 = i0 + 1; c0 < m && next == m; c0 += 64 / kHsLanes) {
                            const int mine = c0 + lane / kHsLanes;
                            T x[DIM], xm[DIM];
#pragma unroll
                            for (int e = 0; e < DIM; ++e) x[e] = xm[e] = cur[e];
                            for (int i = i0 + 1; i < min(m, c0 + 64 / kHsLanes); ++i) {
                                advance(x, i);
                                if (i == mine) {
#pragma unroll
                                    for (int e = 0; e < DIM; ++e) xm[e] = teleport(x[e], P);
                                }
                            }
                            int bm = start + mine;
                            if (bm >= M) bm -= M;
                            const bool near = mine < m && hs_near_part<T, DIM, CELLS>(path, cells, N, P, p, bm, xm, lane % kHsLanes);
                            const unsigned long long bal = __ballot(near);
                            if (bal != 0ull) next = c0 + (__ffsll((long long)bal) - 1) / kHsLanes;
                            // the beads before the first failure keep their first draws
                            if (mine < next && mine < m && lane % kHsLanes == 0) {
#pragma unroll
                                for (int e = 0; e < DIM; ++e) buf[bm * DIM + e] = xm[e];
                            }
                        }
                        if (next >= m) break;
                        for (int i = i0 + 1; i < next; ++i) advance(cur, i);
                        i0 = next;
                    }
                    if (failed) { // the old beads: the link pass then sees no change
                        for (int i = 1 + lane; i < m; i += 64) {
                            int b = start + i;
                            if (b >= M) b -= M;
                            T ro[DIM];
                            path.get(owner(p, us + i), b, ro);
#pragma unroll
                            for (int e = 0; e < DIM; ++e) buf[b * DIM + e] = ro[e];
                        }
                    }
                    if (lane == 0) s_flag[k] = failed ? -1 : 0;
                }
            }
            __syncthreads();
            // (B) per link j: Delta V of bead j (interior beads only) and the pair difference
            for (int j0 = 0; j0 < M; j0 += ngroups) {
                const int j = j0 + grp;
                T part = T(0);
                if (j < M) {
                    int rel = j - off;
                    if (rel < 0) rel += M;
                    const int k = rel / ell, t = rel - k * ell;
                    const int m = min(ell, M - k * ell);
                    if (m >= 2) {
                        const int jn = (j + 1 == M) ? 0 : j + 1;
                        const int u = j >= off ? j : j + M; // window position of slice j
                        T o0[DIM], o1[DIM], n0[DIM], n1[DIM];
                        path.get(owner(p, u), j, o0);
                        path.get(owner(p, u + 1), jn, o1);
#pragma unroll
                        for (int e = 0; e < DIM; ++e) {
                            n0[e] = (t == 0) ? o0[e] : buf[j * DIM + e];
                            n1[e] = (t + 1 == m) ? o1[e] : buf[jn * DIM + e];
                        }
                        if (g == 0) {
                            s_dvb[j] = (t >= 1) ? potV<T, DIM>(n0, P) - potV<T, DIM>(o0, P) : T(0);
                            if (CELLS) {
#pragma unroll
                                for (int e = 0; e < DIM; ++e) s_old[j * DIM + e] = o0[e];
                            }
                        }
                        if (use_pair) // the neighbours of the bead's holder at slice j
                            part = link_delta<T, DIM, CELLS, R, TAB ? 1 : 0>(path, cells, N, ch.nbins, P, owner(p, u), j, jn,
                                                                             n0, n1, o0, o1, g, G);
                    }
                }
                if (use_pair) {
                    part = group_sum_dpp(part, G);
                    if (j < M && g == 0) s_du[j] = part;
                }
            }
            __syncthreads();
            // (C) decisions: Delta V over the interior beads, then the links, in order
            for (int k = tid; k < nseg; k += nt) {
                const int m = min(ell, M - k * ell);
                int acc = 0;
                if (m >= 2) {
                    int start = off + k * ell;
                    if (start >= M) start -= M;
                    T dv = T(0);
                    for (int i = 1; i < m; ++i) {
                        int b = start + i;
                        if (b >= M) b -= M;
                        dv += s_dvb[b];
                    }
                    T lw = -P.tau * dv;
                    if (use_pair) {
                        for (int t = 0; t < m; ++t) {
                            int j = start + t;
                            if (j >= M) j -= M;
                            lw += s_du[j];
                        }
                    }
                    const U4 xa = draw(a.seed, (uint32_t)(p * M + start), sw, (uint32_t)c, kLevyAccept);
                    acc = metropolis(ex(lw), Uniform<T>::get(xa.x)) ? 1 : 0;
                    if (TAB && hs && s_flag[k] < 0) acc = 0; // no bridge outside the hard spheres
                }
                s_flag[k] = acc;
                lacc += acc;
            }
            __syncthreads();
            // (D) commit interior beads of accepted segments
            for (int j = tid; j < M; j += nt) {
                int rel = j - off;
                if (rel < 0) rel += M;
                const int k = rel / ell, t = rel - k * ell;
                const int m = min(ell, M - k * ell);
                if (m >= 2 && t >= 1 && s_flag[k]) {
                    T rn[DIM];
#pragma unroll
                    for (int e = 0; e < DIM; ++e) rn[e] = buf[j * DIM + e];
                    const int q = owner(p, j >= off ? j : j + M);
                    path.put(q, j, rn);
                    if (CELLS) overflow |= !cells_move<T, DIM>(cells, P, q, j, rn, s_old + j * DIM);
                }
            }
        }
        __syncthreads();
        // ---- centre of mass: one particle after another, 1-cycles only
        // (SingleCenterOfMass picks among next[i] == i, src/updates/com.jl:141-160)
        for (int p = 0; p < (com_on ? N : 0); ++p) {
            if (perm && s_next[p] != p) continue; // uniform: LDS value
            ctry += (tid == 0) ? 1 : 0;
            const U4 y = draw(a.seed, (uint32_t)p, sw, (uint32_t)c, kCom);
            const T step2 = (T)a.com_step[c] * T(2);
            T d[DIM];
            d[0] = step2 * (Uniform<T>::get(y.x) - T(0.5));
            if (DIM > 1) d[1] = step2 * (Uniform<T>::get(y.y) - T(0.5));
            if (TAB && hs) {
                // move_polymer! (src/updates/helper.jl:368-395): redraw the shift (draw t >= 2:
                // kHardSphereCom | t << 8) while a moved bead lies closer than a to another
                // particle's bead at its slice; out of draws, the try is rejected
                bool placed = false;
                for (int t = 1; t <= hs_ctr; ++t) {
                    if (t > 1) {
                        const U4 yy = draw(a.seed, (uint32_t)p, sw, (uint32_t)c, (uint32_t)kHardSphereCom | ((uint32_t)t << 8));
                        d[0] = step2 * (Uniform<T>::get(yy.x) - T(0.5));
                        if (DIM > 1) d[1] = step2 * (Uniform<T>::get(yy.y) - T(0.5));
                    }
                    if (tid == 0) s_sw[1] = 0;
                    __syncthreads();
                    bool nr = false;
                    for (int w = tid; w < M * kHsLanes; w += nt) { // kHsLanes threads per moved bead
                        const int j = w / kHsLanes;
                        T q[DIM];
                        path.get(p, j, q);
#pragma unroll
                        for (int e = 0; e < DIM; ++e) q[e] = teleport(q[e] + d[e], P);
                        nr = nr || hs_near_part<T, DIM, CELLS>(path, cells, N, P, p, j, q, w - j * kHsLanes);
                    }
                    if (nr) s_sw[1] = 1;
                    __syncthreads();
                    placed = s_sw[1] == 0;
                    __syncthreads();
                    if (placed) break;
                }
                if (!placed) continue; // uniform: LDS value; counted as a try
            }
            T part = T(0);
            for (int j0 = 0; j0 < M; j0 += ngroups) {
                const int j = j0 + grp;
                if (j < M) {
                    const int jn = (j + 1 == M) ? 0 : j + 1;
                    T o0[DIM], o1[DIM], n0[DIM], n1[DIM];
                    path.get(p, j, o0);
                    path.get(p, jn, o1);
#pragma unroll
                    for (int e = 0; e < DIM; ++e) {
                        n0[e] = teleport(o0[e] + d[e], P);
                        n1[e] = teleport(o1[e] + d[e], P);
                    }
                    if (g == 0) part += -P.tau * (potV<T, DIM>(n0, P) - potV<T, DIM>(o0, P));
                    if (use_pair)
                        part += link_delta<T, DIM, CELLS, R, TAB ? 1 : 0>(path, cells, N, ch.nbins, P, p, j, jn, n0, n1, o0, o1, g, G);
                }
            }
            // workgroup sum of the log-weight change: waves, then wave totals in order
            double x = (double)part;
#pragma unroll
            for (int offs = 32; offs >= 1; offs >>= 1) x += __shfl_xor(x, offs, 64);
            double *ws = s_wsum + (p & 1) * 16;
            if (lane == 0) ws[wave] = x;
            __syncthreads();
            double lw = 0.0;
            for (int w = 0; w < nw; ++w) lw += ws[w];
            const bool acc = metropolis(ex((T)lw), Uniform<T>::get(y.z));
            cacc += (acc && tid == 0) ? 1 : 0;
            if (acc) {
                for (int j = tid; j < M; j += nt) {
                    T ro[DIM], rn[DIM];
                    path.get(p, j, ro);
#pragma unroll
                    for (int e = 0; e < DIM; ++e) rn[e] = teleport(ro[e] + d[e], P);
                    path.put(p, j, rn);
                    if (CELLS) overflow |= !cells_move<T, DIM>(cells, P, p, j, rn, ro);
                }
            }
            __syncthreads();
        }
        // ---- swap tries: the partner choice and the acceptance are serial in particle order
        // on one thread, as in the oracle; the pair sums are spread over the lane groups
        for (int aa = 0; aa < ((perm && sw_on) ? ch.swaps : 0); ++aa) {
            const U4 xs = draw(a.seed, (uint32_t)aa, sw, (uint32_t)c, kSwap);
            const int m = ch.swap_m;
            const int j0 = (int)(((uint64_t)xs.x * (uint64_t)M) >> 32);
            const int n1 = (int)(((uint64_t)xs.y * (uint64_t)N) >> 32);
            const int jm = j0 + m, sm = jm >= M ? jm - M : jm;
            stry += (tid == 0) ? 1 : 0;
            if (N < 2) continue;
            for (int i = tid; i < N; i += nt) {
                T e[DIM], b[DIM];
                path.get(jm < M ? i : s_next[i], sm, e);
                path.get(i, j0, b);
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    s_E[i * DIM + d] = e[d];
                    s_B[i * DIM + d] = b[d];
                }
            }
            __syncthreads();
            if (tid == 0) {
                const T *A = s_B + n1 * DIM, *C = s_E + n1 * DIM;
                T W1 = T(0);
                for (int i = 0; i < N; ++i) W1 += rho0<T, DIM>(A, s_E + i * DIM, P, m) * rho0<T, DIM>(s_B + i * DIM, C, P, m);
                const T target = Uniform<T>::get(xs.z) * W1;
                T cum = T(0);
                int n2 = N - 1;
                for (int i = 0; i < N; ++i) {
                    cum += rho0<T, DIM>(A, s_E + i * DIM, P, m) * rho0<T, DIM>(s_B + i * DIM, C, P, m);
                    if (cum > target) {
                        n2 = i;
                        break;
                    }
                }
                s_sw[0] = n2;
            }
            __syncthreads();
            const int n2 = s_sw[0];
            if (n2 == n1) continue; // uniform: LDS value
            // the two bridges: n1 from r_n1(j0) to E_n2, n2 from r_n2(j0) to E_n1.  With hard spheres
            // hardspherelevy!(r1, s, j0, fpcycle2) and (r2, s, j0, fpcycle1) (src/updates/reshape.jl:157-
            // 162): path 1's beads keep clear of every particle but n2, path 2's of every particle but n1;
            // bead i is redrawn (draw t >= 2: kHardSphereSwap | t << 8, item (2a + w) M + i) while too
            // close, and a bead out of draws rejects the try (s_sw[2 + w] = 0)
            if (tid < 2) {
                const int w = tid;
                T *bw = w == 0 ? s_new0 : s_new1;
                const T *st0 = s_B + (w == 0 ? n1 : n2) * DIM, *en = s_E + (w == 0 ? n2 : n1) * DIM;
                T cur[DIM], em[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    cur[d] = st0[d];
                    em[d] = shift_end(cur[d], en[d], P);
                }
                U4 xr{0u, 0u, 0u, 0u};
                int placed_all = 1;
                for (int i = 1; i < m && placed_all; ++i) {
                    if (i & 1) xr = draw(a.seed, (uint32_t)((2 * aa + w) * M + i), sw, (uint32_t)c, kSwapLevy);
                    T nn[2];
                    if (i & 1) box_muller(xr.x, xr.y, nn[0], nn[1]);
                    else box_muller(xr.z, xr.w, nn[0], nn[1]);
                    const T alpha = bridge_alpha<T>(m, i);
                    const T sd = bridge_sd(alpha, P);
                    T nx[DIM], tb[DIM];
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        nx[d] = bridge_step(cur[d], em[d], alpha, nn[d], sd);
                        tb[d] = teleport(nx[d], P);
                    }
                    if (TAB && hs) {
                        const int excl = w == 0 ? n2 : n1, slice = (j0 + i) % M;
                        auto near = [&](const T *q) {
                            bool r = false;
                            for (int hg = 0; hg < kHsLanes; ++hg)
                                r = r || hs_near_part<T, DIM, CELLS>(path, cells, N, P, excl, slice, q, hg);
                            return r;
                        };
                        for (int t = 2; near(tb); ++t) {
                            if (t > hs_ctr) {
                                placed_all = 0;
                                break;
                            }
                            const U4 xh = draw(a.seed, (uint32_t)((2 * aa + w) * M + i), sw, (uint32_t)c,
                                               (uint32_t)kHardSphereSwap | ((uint32_t)t << 8));
                            box_muller(xh.x, xh.y, nn[0], nn[1]);
#pragma unroll
                            for (int d = 0; d < DIM; ++d) {
                                nx[d] = bridge_step(cur[d], em[d], alpha, nn[d], sd);
                                tb[d] = teleport(nx[d], P);
                            }
                        }
                    }
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        cur[d] = nx[d];
                        bw[i * DIM + d] = tb[d];
                    }
                }
                s_sw[2 + w] = placed_all;
            }
            __syncthreads();
            if (!(s_sw[2] && s_sw[3])) continue; // uniform: LDS values (a rejected try)
            // pair sums (chains with a pair action, src/updates/reshape.jl:178-245): per path w and
            // link t of the window, the neighbours of the old bead (holder x_w at slice j) and of the
            // new bead, other than the two old holders x1, x2; the x1-x2 pair of the old beads comes
            // with the old sums, twice as the reference lists it (ref: each holder's list excludes
            // only itself) or once.  ref also adds the new sums to w_initial (the ratio takes minus
            // both); otherwise the log ratio takes new minus old, plus the pair of the two new paths
            const bool pairs = kPerm && ch.pair != kPairNone, ref = !ch.pair_in_accept;
            auto swap_bead = [&](int w, int i, T *out) { // bead i = 0..m of new path w
                const T *src = i == 0 ? s_B + (w == 0 ? n1 : n2) * DIM
                                      : (i == m ? s_E + (w == 0 ? n2 : n1) * DIM : (w == 0 ? s_new0 : s_new1) + i * DIM);
#pragma unroll
                for (int d = 0; d < DIM; ++d) out[d] = src[d];
            };
            if constexpr (kPerm) if (pairs) {
                for (int it = grp; it < 2 * m; it += ngroups) {
                    const int w = it >= m ? 1 : 0, t = it - w * m;
                    const int u = j0 + t, j = u >= M ? u - M : u, jn = (j + 1 == M) ? 0 : j + 1;
                    const int x1 = owner(n1, u), x2 = owner(n2, u);
                    const int xw = w == 0 ? x1 : x2, xo = w == 0 ? x2 : x1;
                    // one scan per query (a rare move: the plain one-query scan keeps the registers low)
                    T acc[2] = {T(0), T(0)};
#pragma unroll
                    for (int uq = 0; uq < 2; ++uq) { // 0: the new bead, 1: the old one
                        T q[1][DIM], qn[DIM];
                        if (uq == 0) {
                            swap_bead(w, t, q[0]);
                            swap_bead(w, t + 1, qn);
                        } else {
                            path.get(xw, j, q[0]);
                            path.get(owner(w == 0 ? n1 : n2, u + 1), jn, qn);
                        }
                        const bool keep_xo = uq == 1 && (ref || w == 0);
                        scan_pairs<T, DIM, CELLS, 1, 2>(path, cells, N, ch.nbins, P, xw, j, jn, q, g, G,
                                                        [&](int, int k, const T *rk, const T *rkn, T r2, T mult) {
                                                            if (k == xo && !keep_xo) return;
                                                            acc[uq] += mult * pair_lnU<T, DIM, TAB ? 1 : 0>(
                                                                                  rk, q[0], rkn, qn, r2, dist2<T, DIM>(rkn, qn, P), P);
                                                        });
                    }
                    T part = ref ? -(acc[0] + acc[1]) : acc[0] - acc[1];
                    part = group_sum_dpp(part, G);
                    if (g == 0) (w == 0 ? s_du : s_dvb)[t] = part;
                }
                if (!ref) {
                    for (int t = tid; t < m; t += nt) { // the pair of the two new paths at link t
                        T a0[DIM], a1[DIM], b0[DIM], b1[DIM];
                        swap_bead(0, t, a0);
                        swap_bead(0, t + 1, a1);
                        swap_bead(1, t, b0);
                        swap_bead(1, t + 1, b1);
                        const T mult = pair_mult<T, DIM>(a0, b0, P, ch.nbins);
                        s_old[t] = mult > T(0) ? mult * pair_lnU<T, DIM, TAB ? 1 : 0>(b0, a0, b1, a1, dist2<T, DIM>(b0, a0, P),
                                                                                      dist2<T, DIM>(b1, a1, P), P)
                                               : T(0);
                    }
                }
                __syncthreads();
            }
            if (tid == 0) {
                // Delta lnW over the interior beads of both paths (old beads along the old cycles)
                T dv = T(0);
                for (int w = 0; w < 2; ++w) {
                    const int nw = w == 0 ? n1 : n2;
                    const T *bw = w == 0 ? s_new0 : s_new1;
                    for (int i = 1; i < m; ++i) {
                        const int u = j0 + i;
                        T ob[DIM];
                        path.get(owner(nw, u), u >= M ? u - M : u, ob);
                        dv += potV<T, DIM>(bw + i * DIM, P) - potV<T, DIM>(ob, P);
                    }
                }
                T lw = -P.tau * dv;
                if (pairs)
                    for (int t = 0; t < m; ++t) lw += s_du[t] + s_dvb[t] + (ref ? T(0) : s_old[t]);
                T ratio = ex(lw);
                if (ch.swap_rule == 0) {
                    // reverse over forward partner-choice probability, either particle drawn first
                    const T *A = s_B + n1 * DIM, *C = s_E + n1 * DIM, *Bp = s_B + n2 * DIM, *D = s_E + n2 * DIM;
                    T W1 = T(0), W2 = T(0), W1p = T(0), W2p = T(0);
                    for (int i = 0; i < N; ++i) {
                        const T *Ei = s_E + i * DIM, *Bi = s_B + i * DIM;
                        const T *Epi = i == n1 ? D : (i == n2 ? C : Ei);
                        W1 += rho0<T, DIM>(A, Ei, P, m) * rho0<T, DIM>(Bi, C, P, m);
                        W2 += rho0<T, DIM>(Bp, Ei, P, m) * rho0<T, DIM>(Bi, D, P, m);
                        W1p += rho0<T, DIM>(A, Epi, P, m) * rho0<T, DIM>(Bi, D, P, m);
                        W2p += rho0<T, DIM>(Bp, Epi, P, m) * rho0<T, DIM>(Bi, C, P, m);
                    }
                    ratio *= (T(1) / W1p + T(1) / W2p) / (T(1) / W1 + T(1) / W2);
                }
                s_sw[1] = metropolis(ratio, Uniform<T>::get(xs.w)) ? 1 : 0;
            }
            __syncthreads();
            if (s_sw[1]) { // uniform: LDS value
                const int q1 = s_next[n1], q2 = s_next[n2];
                // one thread per slice: the interior slices of the window, then the tails
                for (int i = 1 + tid; i < m; i += nt) {
                    const int u = j0 + i, j = u >= M ? u - M : u;
                    const int qa = u < M ? n1 : q2, qb = u < M ? n2 : q1; // along the new cycles
                    T oa[DIM], ob[DIM];
                    if (CELLS) {
                        path.get(qa, j, oa);
                        path.get(qb, j, ob);
                    }
                    path.put(qa, j, s_new0 + i * DIM);
                    path.put(qb, j, s_new1 + i * DIM);
                    if (CELLS) {
                        overflow |= !cells_move<T, DIM>(cells, P, qa, j, s_new0 + i * DIM, oa);
                        overflow |= !cells_move<T, DIM>(cells, P, qb, j, s_new1 + i * DIM, ob);
                    }
                }
                if (jm < M) {
                    for (int j = jm + tid; j < M; j += nt) {
                        T r1[DIM], r2[DIM];
                        path.get(n1, j, r1);
                        path.get(n2, j, r2);
                        path.put(n1, j, r2);
                        path.put(n2, j, r1);
                        if (CELLS) cells_exchange(cells, n1, n2, j);
                    }
                }
                __syncthreads();
                if (tid == 0) {
                    s_next[n1] = q2;
                    s_next[n2] = q1;
                    sacc += 1;
                }
            }
            __syncthreads();
        }
        if (meas_now) {
            const ChainMeas m = measure_chain<T, DIM, CELLS>(path, cells, ch, P, s_buf, a.radial, a.radial_bins,
                                                              a.radial_winv, a.cart, a.cart_bins, c, 8,
                                                              perm ? s_next : nullptr);
            if (tid == 0) fold_meas(m, ch, s_q);
        }
    }
    __syncthreads();
    if constexpr (LDSR) {
        for (int i = tid; i < N * M * DIM; i += nt) gpath[i] = lpath[i];
    }
    for (int i = tid; i < np_; i += nt) gnext[i] = s_next[i];
    // per-launch outputs (accept counts reduced deterministically)
    const double la = block_sum((double)lacc, s_buf);
    const double ca = block_sum((double)cacc, s_buf);
    if (tid == 0) {
        double *out = a.chain_meas + (size_t)c * (kNQ + kCnt);
        if (a.split && a.sweep_off > 0) { // later one-sweep launches of a split step add their counts
            out[kNQ] += la;
            out[kNQ + 1] += ca;
            out[kNQ + 2] += ctry;
            out[kNQ + 3] += sacc;
            out[kNQ + 4] += stry;
        } else {
            for (int k = 0; k < kNQ; ++k) out[k] = s_q[k]; // zero when split: k_meas_fold adds
            out[kNQ] = la;
            out[kNQ + 1] = ca;
            out[kNQ + 2] = ctry;
            out[kNQ + 3] = sacc;
            out[kNQ + 4] = stry;
        }
    }
    if (overflow) atomicOr(a.status + c, 1);
}

// Estimators of whole chains as a wide kernel (split measuring steps): the
// sweep kernel runs one sweep per launch and leaves the estimators out; after
// each such launch k_meas_beads spreads the chains' beads over nblk workgroups
// per chain (the per-bead sums of measure_chain with 2 lanes per bead, FP64
// partial sums per workgroup) and k_meas_fold adds the partials in workgroup order and folds the
// measurement into the chain's kNQ sums.  Both skip sweeps that are not
// measured ((sweep + 1) % measure_every != 0).
struct MeasArgs {
    const ChainDev *chains;
    const void *prm;
    const int *ids; // whole chains of this launch
    int n_ids, nblk;
    const void *pos;
    int *cell_cnt;
    unsigned short *cell_mem;
    int *bins;
    const int *next;
    unsigned long long *radial, *cart;
    double radial_winv;
    int radial_bins, cart_bins;
    double *part; // [id][blk][6]
    double *chain_meas;
    const RunState *rs;
    int sweep_off, measure_every;
};

__device__ __forceinline__ bool split_measured(const MeasArgs &a)
{
    const uint32_t sw = a.rs->sweep + (uint32_t)a.sweep_off;
    return (sw + 1u) % (uint32_t)a.measure_every == 0u;
}

template <typename T, int DIM, bool CELLS> __global__ void __launch_bounds__(256) k_meas_beads(MeasArgs a)
{
    __shared__ double wbuf[32];
    if (!split_measured(a)) return;
    const int y = blockIdx.y, c = a.ids[y];
    const ChainDev ch = a.chains[c];
    const Prm<T> P = reinterpret_cast<const Prm<T> *>(a.prm)[c];
    const int nb = ch.N * ch.M, per = (nb + a.nblk - 1) / a.nblk;
    const int b0 = min(nb, (int)blockIdx.x * per), b1 = min(nb, b0 + per);
    const Path<T, DIM, true> path{reinterpret_cast<T *>(const_cast<void *>(a.pos)) + ch.pos_off, ch.M, ch.N};
    const int *nxt = ch.swaps > 0 ? a.next + ch.part_base : nullptr;
    const Cells cells{a.cell_cnt + ch.cell_off, a.cell_mem + ch.mem_off, a.bins + ch.bin_off, ch.M, ch.ncell, ch.cap,
                      ch.nbins, nxt};
    const ChainMeas m = measure_chain<T, DIM, CELLS>(path, cells, ch, P, wbuf, a.radial, a.radial_bins, a.radial_winv,
                                                      a.cart, a.cart_bins, c, 2, nxt, b0, b1);
    if (threadIdx.x == 0) {
        double *o = a.part + ((size_t)y * a.nblk + blockIdx.x) * 6;
        o[0] = m.link;
        o[1] = m.V;
        o[2] = m.rg;
        o[3] = m.su;
        o[4] = m.rgp;
        o[5] = m.r2w;
    }
}

__global__ void k_meas_fold(MeasArgs a)
{
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= a.n_ids || !split_measured(a)) return;
    const int c = a.ids[y];
    ChainMeas m{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < a.nblk; ++b) {
        const double *o = a.part + ((size_t)y * a.nblk + b) * 6;
        m.link += o[0];
        m.V += o[1];
        m.rg += o[2];
        m.su += o[3];
        m.rgp += o[4];
        m.r2w += o[5];
    }
    fold_meas(m, a.chains[c], a.chain_meas + (size_t)c * (kNQ + kCnt));
}

// After each launch: fold per-particle (tile kernel) or per-chain (whole
// kernel) launch outputs into the cumulative counters and, when measuring,
// into block `blk`; while thermalising, adapt the COM step per chain with the
// rule of adjust!(::Step) (src/updates/helper.jl:22-32) on this launch's
// acceptance.
struct PostArgs {
    ChainDev *chains;       // ell and nseg of whole chains adapt while thermalising (adapt_ell)
    int n_chains;
    const int *is_whole;
    const double *part_acc;
    const int *part_lacc, *part_cacc;
    const double *chain_meas;
    double *com_step;
    long long *counters; // per chain: levy accepted, COM accepted
    double *blocks;      // [blk][chain][kNQ]
    const RunState *rs; // block index of this launch
    int measure, nsweeps;
    int adapt_ell;      // NumbOfSlices rule on every chain while thermalising
};

// Staging tries and proposed beads of one sweep of a chain (segments of ell
// tiling the ring; a segment of m links proposes m - 1 beads).
__host__ __device__ inline void staging_counts(const ChainDev &h, long long &tries, long long &beads)
{
    tries = 0;
    beads = 0;
    for (int k = 0; k < h.nseg; ++k) {
        const int m = h.ell < h.M - k * h.ell ? h.ell : h.M - k * h.ell;
        if (m >= 2) {
            tries += 1;
            beads += m - 1;
        }
    }
    tries *= h.N;
    beads *= h.N;
}

__global__ void k_post(PostArgs a)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n_chains) return;
    ChainDev &ch = a.chains[c];
    long long lt, lb; // staging tries and proposed beads of this launch (the launch's ell)
    staging_counts(ch, lt, lb);
    // the sweeps of this launch that ran each pass (move weights); rs->sweep is the launch's first
    const long long nst = every_count(a.rs->sweep, a.nsweeps, ch.every_st);
    lt *= nst;
    lb *= nst;
    long long la = 0, ca = 0;
    double q[kNQ];
    for (int k = 0; k < kNQ; ++k) q[k] = 0.0;
    long long ct = (long long)ch.N * every_count(a.rs->sweep, a.nsweeps, ch.every_com), sa = 0, st = 0;
    if (a.is_whole[c]) {
        const double *src = a.chain_meas + (size_t)c * (kNQ + kCnt);
        for (int k = 0; k < kNQ; ++k) q[k] = src[k];
        la = (long long)src[kNQ];
        ca = (long long)src[kNQ + 1];
        ct = (long long)src[kNQ + 2];
        sa = (long long)src[kNQ + 3];
        st = (long long)src[kNQ + 4];
    } else {
        double l = 0.0, v = 0.0, g = 0.0, nm = 0.0;
        for (int p = 0; p < ch.N; ++p) {
            const int gq = ch.part_base + p;
            la += a.part_lacc[gq];
            ca += a.part_cacc[gq];
            if (a.measure) {
                l += a.part_acc[(size_t)gq * 4 + 0];
                v += a.part_acc[(size_t)gq * 4 + 1];
                g += a.part_acc[(size_t)gq * 4 + 2];
                nm = a.part_acc[(size_t)gq * 4 + 3];
            }
        }
        if (a.measure) {
            const double tau = ch.tau, M = ch.M;
            q[qN] = nm;
            q[qE] = nm * ((double)(ch.dim * ch.N) / (2.0 * tau)) - l / (((4.0 * ch.lambda) * (tau * tau)) * M) + v / M;
            q[qEv] = g / (2.0 * M) + v / M;
            if (ch.pot == kPotHarmonic && ch.omega != 0.0) {
                // sum omega r^2 = 2 sum V / omega for V = omega^2 r^2 / 2
                q[qSw] = tau * 2.0 * v / ch.omega;
                q[qdEw] = 2.0 * v / (ch.omega * M);
            }
        }
    }
    long long *cnt = a.counters + (size_t)c * kCnt;
    cnt[0] += la;
    cnt[1] += ca;
    cnt[2] += ct;
    cnt[3] += sa;
    cnt[4] += st;
    cnt[5] += lt;
    cnt[6] += lb;
    // adjust!(::NumbOfSlices) (src/updates/helper.jl:42-52) on this launch's staging
    // acceptance: one slice less below 0.6, one more above 0.8, within
    // [minslices, maxslices] = [2, M - 2] (ReshapeLinear's defaults,
    // src/updates/reshape.jl:12-21), capped at 32 and never below 2
    if (!a.measure && a.adapt_ell && lt > 0) {
        const double acc = (double)la / (double)lt;
        int ell = ch.ell;
        if (acc < 0.6) ell -= 1;
        else if (acc > 0.8) ell += 1;
        ell = max(2, min(min(32, ch.M - 2), ell));
        ch.ell = ell;
        ch.nseg = (ch.M + ell - 1) / ell;
    }
    if (a.measure) {
        double *b = a.blocks + ((size_t)a.rs->blk * a.n_chains + c) * kNQ;
        for (int k = 0; k < kNQ; ++k) b[k] = q[k];
    } else {
        const long long tries = ct;
        double step = a.com_step[c];
        if (tries > 0) {
            const double acc = (double)ca / (double)tries;
            if (acc < 0.4) step *= 0.9;
            else if (acc > 0.6) step *= 1.1;
            step = fmax(0.1, fmin(ch.L / 2.0, step));
        }
        a.com_step[c] = step;
    }
}

// Rebuild the cell lists of the whole chains from their positions (init_nn,
// src/system.jl:80-91; psort!, src/nearest_neighbours.jl:40-52): one thread
// per (chain, slice), particles in index order.
template <typename T, int DIM>
__global__ void k_build_cells(const ChainDev *chains, const int *whole, int n_whole, const void *pos, int *cell_cnt,
                              unsigned short *cell_mem, int *bins, int *status, int only)
{
    const int w = blockIdx.y;
    if (w >= n_whole) return;
    const int c = whole[w];
    if (only >= 0 && c != only) return;
    const ChainDev ch = chains[c];
    if (ch.ncell <= 0) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ch.M) return;
    const Prm<T> P = Prm<T>::make(ch);
    const Path<T, DIM, true> path{const_cast<T *>(reinterpret_cast<const T *>(pos)) + ch.pos_off, ch.M, ch.N};
    const Cells C{cell_cnt + ch.cell_off, cell_mem + ch.mem_off, bins + ch.bin_off, ch.M,
                  ch.ncell, ch.cap, ch.nbins};
    for (int cc = 0; cc < ch.ncell; ++cc) C.count(j, cc) = 0;
    for (int n = 0; n < ch.N; ++n) {
        T r[DIM];
        path.get(n, j, r);
        const int cell = cell_of<T, DIM>(r, P, ch.nbins);
        int &cnt = C.count(j, cell);
        if (cnt < ch.cap) {
            C.members(j, cell)[cnt] = (unsigned short)n;
            cnt += 1;
            C.bin[(long long)n * ch.M + j] = cell;
        } else {
            C.bin[(long long)n * ch.M + j] = -1;
            atomicOr(status + c, 1);
        }
    }
}

} // namespace pimc
