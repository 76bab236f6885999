// Sweep kernels of chains with a Gaussian pair action (the interacting
// workloads): staging moves by slice windows, centre-of-mass moves by
// speculation and dependency rounds.
//
// Staging (k_stage_win).  All particles of a chain share one staging offset
// per sweep, so their segments occupy the same slice windows [s, s + ell].
// A move of particle p in window w reads and writes only beads at the
// window's slices; the end slices s and s + ell are written by no staging
// move.  Moves in different windows therefore touch disjoint beads and
// commute: the result is the particle-by-particle schedule whatever order the
// windows run in.  One workgroup owns one window.
//
// Within a window every proposal depends only on the particle's own fixed end
// beads and its own draws, so all of them are formed first.  The log-weight
// change of particle p's try against the configuration the schedule presents
// to it is
//   Delta V_p + sum_q A(p, q) + sum_{q < p accepted} C(p, q),
// with, per link, T(x, y) the pair term of p's beads x (o: old, n: new) with
// q's beads y:  A = T(n, o) - T(o, o)  and  C = T(n, n) - T(o, n) - A.
// Both are computed for every pair whose window bounds (centre and radius of
// the old and new beads) come within the cut-off; other pairs have no term
// within the cut-off, so A = C = 0 exactly.  The decisions then need only the
// acceptances of each try's earlier partners: a chunk of 64 tries is decided
// in rounds, each round taking every try whose earlier partners in the chunk
// are decided.  Every sum runs in a fixed order (partners by index), so the
// result does not depend on thread timing.
//
// Centre of mass.  k_com_spec computes, by windows of the ring, every try's
// Delta V + sum_q A against the configuration at the start of the phase and
// the window bounds of the old rings; k_com_dec (one workgroup per chain)
// adds, for each try, the C terms of its accepted earlier partners, over the
// windows where their bounds (old and shifted) come within the cut-off, and
// decides in rounds as above; k_com_apply moves the accepted rings.
//
// The cell lists are not used by these kernels; k_cells_win rebuilds them
// (members by particle index) where the estimators or other callers need them.
#pragma once
#include "sweep_kernels.hpp"

namespace pimc {

constexpr int kChunk = 64; // tries decided per block of rounds (one wave)

// Slices per centre-of-mass window: at most 32 windows per ring.
__host__ __device__ inline int com_win_len(int M) { return M <= 256 ? 8 : (M + 31) / 32; }
__host__ __device__ inline int com_nwin(int M) { return (M + com_win_len(M) - 1) / com_win_len(M); }

// Signed minimum-image offset x1 - x2.
template <typename T> __device__ __forceinline__ T signed_mindist(T x1, T x2, const Prm<T> &p)
{
    T d = x1 - x2;
    if (d > p.L) d -= p.twoL;
    else if (d < -p.L) d += p.twoL;
    return d;
}

// Multiplicity of particle k (bead rk) among the stencil candidates of a
// query at q: 1 within the cut-off for nbins >= 3; when nbins < 3 the number
// of stencil slots of q's cell that hold rk's cell (find_nns repeats cells
// then).  Symmetric in q and rk.
template <typename T, int DIM>
__device__ __forceinline__ T stencil_mult(const T *q, const T *rk, const Prm<T> &P, int nbins)
{
    if (nbins >= 3) return T(1);
    constexpr int NS = (DIM == 1) ? 3 : 9;
    const int c0 = cell_of<T, DIM>(q, P, nbins), ck = cell_of<T, DIM>(rk, P, nbins);
    int mm = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) mm += (stencil_cell(c0, s, nbins, DIM) == ck);
    return T(mm);
}

// Exclusive prefix of one value per lane over a wave; returns the lane's
// prefix and sets the wave total.
__device__ __forceinline__ int wave_prefix(int v, int lane, int &total)
{
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

// --- window bounds (FP32; the overlap test's margin covers the rounding) ---

__device__ __forceinline__ float mi_f(float d, float twoL, float inv2L) { return d - twoL * rintf(d * inv2L); }

// Whether beads within ar of (ax, ay) and beads within br of (bx, by) can
// come within the cut-off h of each other (periodic box of side twoL).
template <int DIM>
__device__ __forceinline__ bool bnd_near(float ax, float ay, float ar, float bx, float by, float br, float h,
                                         float twoL, float inv2L)
{
    const float dx = mi_f(ax - bx, twoL, inv2L);
    float d2 = dx * dx;
    if (DIM > 1) {
        const float dy = mi_f(ay - by, twoL, inv2L);
        d2 += dy * dy;
    }
    const float reach = (ar + br + h) * 1.00001f + 1e-5f * twoL;
    return d2 <= reach * reach;
}

// The four pairings of two bounds shifted by their centre-of-mass shifts
// (x: particle a old/new, y: particle b old/new).
template <int DIM>
__device__ __forceinline__ bool bnd_near4(float4 A, const float *da, float4 B, const float *db, float h, float twoL,
                                          float inv2L)
{
    const float ax1 = A.x + da[0], ay1 = DIM > 1 ? A.y + da[1] : 0.f;
    const float bx1 = B.x + db[0], by1 = DIM > 1 ? B.y + db[1] : 0.f;
    return bnd_near<DIM>(A.x, A.y, A.z, B.x, B.y, B.z, h, twoL, inv2L) ||
           bnd_near<DIM>(ax1, ay1, A.z, B.x, B.y, B.z, h, twoL, inv2L) ||
           bnd_near<DIM>(A.x, A.y, A.z, bx1, by1, B.z, h, twoL, inv2L) ||
           bnd_near<DIM>(ax1, ay1, A.z, bx1, by1, B.z, h, twoL, inv2L);
}

// --- pair terms of one window ---

// Per slice of a link pairing: squared distance test, pair potential and
// stencil multiplicity of the query bead x (particle p) against y (particle q).
template <typename T, int DIM> struct SliceTerm {
    T e, mult;
    bool in;
    __device__ __forceinline__ void eval(const T *x, const T *y, const Prm<T> &P, int nbins)
    {
        const T d2 = dist2<T, DIM>(x, y, P);
        in = d2 <= P.h2;
        e = vpair_r2(d2, P);
        mult = (nbins >= 3) ? T(1) : stencil_mult<T, DIM>(x, y, P, nbins);
    }
    // lnU of the link from this slice (the query slice) to the next: the term a
    // neighbour adds in interaction_action! (src/updates/helper.jl:326-346), 0 for a
    // bead outside the cut-off (not among find_nns's neighbours)
    __device__ __forceinline__ T link(const SliceTerm &next, T ctau) const
    {
        return in ? mult * (ctau * (e + next.e)) : T(0);
    }
};

// A(p, q) and, for q < p (back), C(p, q) over the links 0..m-1 of one window.
// bead(t, who, ro, rn): old and new bead of particle p (who 0) or q (who 1) at
// window slice t.  Slices are fetched B at a time before any is used (B > 1
// lets loads from HBM overlap).  ENDS: the old and new beads coincide at slices
// 0 and m (staging), so one pairing is evaluated there.  hit: some pairing has
// a term within the cut-off.
template <typename T, int DIM, int B, bool ENDS = false, typename BeadFn>
__device__ __forceinline__ void pair_window(BeadFn bead, int m, bool back, const Prm<T> &P, int nbins, T &A, T &C,
                                            bool &hit)
{
    const T ctau = T(-0.5) * P.tau;
    SliceTerm<T, DIM> oo, no, on, nn;
    A = T(0);
    C = T(0);
    bool any = false;
    for (int t0 = 0; t0 <= m; t0 += B) {
        T po[B][DIM], pn[B][DIM], qo[B][DIM], qn[B][DIM];
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (t0 + k <= m) {
                bead(t0 + k, 0, po[k], pn[k]);
                bead(t0 + k, 1, qo[k], qn[k]);
            }
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int t = t0 + k;
            if (t > m) break;
            SliceTerm<T, DIM> oo1, no1, on1, nn1;
            oo1.eval(po[k], qo[k], P, nbins);
            if (ENDS && (t == 0 || t == m)) {
                no1 = oo1;
                on1 = oo1;
                nn1 = oo1;
            } else {
                no1.eval(pn[k], qo[k], P, nbins);
                if (back) {
                    on1.eval(po[k], qn[k], P, nbins);
                    nn1.eval(pn[k], qn[k], P, nbins);
                }
            }
            if (t > 0) {
                const T too = oo.link(oo1, ctau), tno = no.link(no1, ctau);
                const T a = tno - too;
                A += a;
                if (back) C += (nn.link(nn1, ctau) - on.link(on1, ctau)) - a;
            }
            if (t < m) {
                any = any || oo1.in || no1.in || (back && (on1.in || nn1.in));
                oo = oo1;
                no = no1;
                on = on1;
                nn = nn1;
            }
        }
    }
    hit = any;
}

// --- ordered pair lists of a chunk of particles ---

// LDS layout of the chunk machinery: candidate masks (kChunk x nq words),
// counts, offsets, ends of the earlier partners, per-try sums, then the slot
// pool (partner index with the hit flag in bit 15, owner, nval values: with
// one value, pa and pc name the same array).
struct ChunkLds {
    size_t mask, cnt, off, nb, spec, misc, pq, po, pa, pc, end;
};
__host__ __device__ inline ChunkLds chunk_lds(size_t o, int N, int cap, int tsize, int nval)
{
    ChunkLds L;
    const int nq = (N + 63) / 64;
    L.mask = o;
    o = align16(o + (size_t)kChunk * nq * 8);
    L.cnt = o;
    o = align16(o + kChunk * 4);
    L.off = o;
    o = align16(o + (kChunk + 1) * 4);
    L.nb = o;
    o = align16(o + kChunk * 4);
    L.spec = o;
    o = align16(o + kChunk * 8);
    L.misc = o;
    o = align16(o + 16);
    L.pq = o;
    o = align16(o + (size_t)cap * 2);
    L.po = o;
    o = align16(o + (size_t)cap);
    L.pa = o;
    o = align16(o + (size_t)cap * tsize);
    L.pc = L.pa; // one value per slot: pa and pc are the same array
    if (nval > 1) {
        L.pc = o;
        o = align16(o + (size_t)cap * tsize);
    }
    L.end = o;
    return L;
}

template <typename T> struct ChunkPtr {
    unsigned long long *mask;
    int *cnt, *off, *nb, *misc;
    double *spec;
    unsigned short *pq;
    unsigned char *po;
    T *pa, *pc;
    __device__ __forceinline__ ChunkPtr(unsigned char *s, const ChunkLds &L)
        : mask(reinterpret_cast<unsigned long long *>(s + L.mask)), cnt(reinterpret_cast<int *>(s + L.cnt)),
          off(reinterpret_cast<int *>(s + L.off)), nb(reinterpret_cast<int *>(s + L.nb)),
          misc(reinterpret_cast<int *>(s + L.misc)), spec(reinterpret_cast<double *>(s + L.spec)),
          pq(reinterpret_cast<unsigned short *>(s + L.pq)), po(s + L.po), pa(reinterpret_cast<T *>(s + L.pa)),
          pc(reinterpret_cast<T *>(s + L.pc))
    {
    }
};

// --- grid of window centres ---

// A uniform periodic grid over the particles' bound centres, rebuilt per
// window: cells of side >= 2 rmax + h (rmax the largest bound radius), so two
// bounds within reach sit in neighbouring cells and a particle's candidates
// are the members of its 3^DIM stencil.  ng cells per dimension, at most
// kGridMax; no grid (ng = 0) when the stencil would cover more than a quarter
// (2D: ng < 6) or three eighths (1D: ng < 8) of the box, where the full scan
// is cheaper.
constexpr int kGridMax = 16;

struct GridLds {
    size_t start, mem, pcell, misc, end;
};
__host__ __device__ inline GridLds grid_lds(size_t o, int N, int dim)
{
    GridLds L;
    const int nc = dim == 1 ? kGridMax : kGridMax * kGridMax;
    L.start = o;
    o = align16(o + (size_t)(nc + 1) * 4);
    L.mem = o;
    o = align16(o + (size_t)N * 2);
    L.pcell = o;
    o = align16(o + (size_t)N * 2);
    L.misc = o;
    o = align16(o + 16);
    L.end = o;
    return L;
}

struct WinGrid {
    int *start;            // after the build: cell c's members are mem[c ? start[c - 1] : 0 .. start[c])
    unsigned short *mem;   // particles by cell
    unsigned short *pcell; // each particle's cell
    int *misc;             // [0] largest radius (float bits)
    int ng;
    __device__ __forceinline__ WinGrid(unsigned char *s, const GridLds &L)
        : start(reinterpret_cast<int *>(s + L.start)), mem(reinterpret_cast<unsigned short *>(s + L.mem)),
          pcell(reinterpret_cast<unsigned short *>(s + L.pcell)), misc(reinterpret_cast<int *>(s + L.misc)), ng(0)
    {
    }
};

// Build the grid from centre(p, x, y, r) (all threads of the workgroup); sets G.ng.
template <int DIM, typename Centre>
__device__ inline void grid_build(WinGrid &G, int N, float h, float twoL, float inv2L, Centre centre)
{
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) G.misc[0] = 0;
    __syncthreads();
    for (int p = tid; p < N; p += nt) {
        float x, y, r;
        centre(p, x, y, r);
        atomicMax(G.misc, __float_as_int(fmaxf(r, 0.f)));
    }
    __syncthreads();
    const float rmax = __int_as_float(G.misc[0]);
    const float side = (2.f * rmax + h) * 1.0001f + 2e-5f * twoL;
    int ng = min((int)floorf(twoL / side), kGridMax);
    G.ng = ng >= (DIM == 1 ? 8 : 6) ? ng : 0;
    if (G.ng == 0) return; // uniform: every thread read the same rmax
    const int ncell = DIM == 1 ? ng : ng * ng;
    for (int c = tid; c <= ncell; c += nt) G.start[c] = 0;
    __syncthreads();
    const float scale = (float)ng * inv2L;
    auto axis = [&](float v) {
        const float w = v - twoL * floorf(v * inv2L + 0.5f); // into [-L, L]
        return min(max((int)((w + 0.5f * twoL) * scale), 0), ng - 1);
    };
    for (int p = tid; p < N; p += nt) {
        float x, y, r;
        centre(p, x, y, r);
        const int c = DIM == 1 ? axis(x) : axis(x) + ng * axis(y);
        G.pcell[p] = (unsigned short)c;
        atomicAdd(G.start + c + 1, 1);
    }
    __syncthreads();
    if (wave == 0) { // inclusive prefix: start[c] = members of the cells before c
        int base = 0;
        for (int c0 = 0; c0 <= ncell; c0 += 64) {
            const int c = c0 + lane;
            const int v = c <= ncell ? G.start[c] : 0;
            int tot;
            const int pre = wave_prefix(v, lane, tot);
            if (c <= ncell) G.start[c] = base + pre + v;
            base += tot;
        }
    }
    __syncthreads();
    for (int p = tid; p < N; p += nt) {
        const int at = atomicAdd(G.start + G.pcell[p], 1); // start[c] ends at cell c's end
        G.mem[at] = (unsigned short)p;
    }
    __syncthreads();
}

// Slots of the chunk's particles pa, pa + 1, ... (at most kc <= kChunk): for
// each the partners q with cand(p, q), by index.  Fills as many particles as
// the pool holds (at least one: cap >= N).  Returns the number of particles filled; X.misc[1] is the
// slot count.  All threads of the workgroup.
template <int DIM = 2, typename T, typename Cand>
__device__ inline int chunk_slots(ChunkPtr<T> &X, int pa, int N, int cap, int kc, Cand cand,
                                  const WinGrid *G = nullptr)
{
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6, nwv = nt >> 6;
    const int nq = (N + 63) / 64, nc = min(kc, N - pa);
    if (G && G->ng > 0) {
        // candidates from the grid stencil (one wave per particle, the stencil's members
        // flattened over the lanes), set as bits of the particle's mask
        const int ng = G->ng;
        constexpr int NS = DIM == 1 ? 3 : 9;
        for (int i = tid; i < nc * nq; i += nt) X.mask[i] = 0ull;
        __syncthreads();
        for (int l = wave; l < nc; l += nwv) {
            const int p = pa + l, pc = G->pcell[p];
            const int ix = DIM == 1 ? pc : pc % ng, iy = DIM == 1 ? 0 : pc / ng;
            int b[NS], pre[NS + 1];
            pre[0] = 0;
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                int cx = ix + k % 3 - 1, cy = iy + (DIM == 1 ? 0 : k / 3 - 1);
                cx += cx < 0 ? ng : (cx >= ng ? -ng : 0);
                cy += cy < 0 ? ng : (cy >= ng ? -ng : 0);
                const int cc = cx + ng * cy;
                b[k] = cc ? G->start[cc - 1] : 0;
                pre[k + 1] = pre[k] + G->start[cc] - b[k];
            }
            for (int i = lane; i < pre[NS]; i += 64) {
                int at = b[0] + i;
#pragma unroll
                for (int k = 1; k < NS; ++k) at = i >= pre[k] ? b[k] + (i - pre[k]) : at;
                const int q = G->mem[at];
                if (cand(p, q)) atomicOr(X.mask + l * nq + (q >> 6), 1ull << (q & 63));
            }
        }
        __syncthreads();
        for (int l = tid; l < nc; l += nt) {
            int cnt = 0;
            for (int k = 0; k < nq; ++k) cnt += __popcll(X.mask[l * nq + k]);
            X.cnt[l] = cnt;
        }
    } else {
        for (int l = wave; l < nc; l += nwv) {
            const int p = pa + l;
            int cnt = 0;
            for (int q0 = 0; q0 < N; q0 += 64) {
                const int q = q0 + lane;
                const bool ok = q < N && cand(p, q);
                const unsigned long long bal = __ballot(ok);
                if (lane == 0) X.mask[l * nq + (q0 >> 6)] = bal;
                cnt += __popcll(bal);
            }
            if (lane == 0) X.cnt[l] = cnt;
        }
    }
    __syncthreads();
    if (wave == 0) {
        const int v = lane < nc ? X.cnt[lane] : 0;
        int tot;
        const int pre = wave_prefix(v, lane, tot);
        const bool fits = lane < nc && pre + v <= cap;
        const int nf = __popcll(__ballot(fits));
        X.off[lane] = pre;
        if (lane == nf - 1) {
            X.misc[0] = nf;
            X.misc[1] = pre + v;
        }
    }
    __syncthreads();
    const int nf = X.misc[0];
    for (int it = tid; it < nf * nq; it += nt) {
        const int l = it / nq, wd = it - l * nq;
        const unsigned long long *mk = X.mask + (size_t)l * nq;
        int s = X.off[l];
        for (int k = 0; k < wd; ++k) s += __popcll(mk[k]);
        unsigned long long b = mk[wd];
        while (b) {
            const int bit = __ffsll((long long)b) - 1;
            b &= b - 1;
            X.pq[s] = (unsigned short)(wd * 64 + bit);
            X.po[s] = (unsigned char)l;
            ++s;
        }
    }
    __syncthreads();
    return nf;
}

// Decisions of the chunk's tries pa .. pa + nf - 1 in index order (wave 0).
// Slots X.off[l] .. X.nb[l] hold try l's earlier partners by index with their
// C values; acc[q] holds the decisions of the tries before the chunk.  A try
// waits for its earlier partners inside the chunk that have a term within the
// cut-off (the others add exactly zero).  Each round decides every try whose
// such partners are decided; a try's sum takes the partners before the chunk
// in index order, then the partners inside the chunk as they are accepted
// (by round, in index order within a round), a fixed order.  Returns the
// acceptance mask.
template <typename T, typename Unif>
__device__ inline unsigned long long chunk_decide(const ChunkPtr<T> &X, int pa, int nf, const unsigned char *acc,
                                                  Unif unif)
{
    const int l = threadIdx.x & 63;
    const bool in = l < nf;
    const int o0 = in ? X.off[l] : 0, o1 = in ? X.nb[l] : 0;
    T tot = T(0);
    unsigned long long dep = 0ull, pm = 0ull; // in-chunk partners with a term within the cut-off, all
    int si = o1;                              // first in-chunk partner's slot
    for (int i = o0; i < o1; ++i) {
        const int qq = X.pq[i], q = qq & 0x7fff;
        if (q < pa) {
            if (acc[q]) tot += X.pc[i];
        } else {
            if (si == o1) si = i;
            pm |= 1ull << (q - pa);
            if (qq & 0x8000) dep |= 1ull << (q - pa);
        }
    }
    unsigned long long decided = nf >= 64 ? 0ull : (~0ull << nf), am = 0ull;
    const double sp = in ? X.spec[l] : 0.0;
    const T u = in ? unif(pa + l) : T(0);
    while (decided != ~0ull) {
        const bool ready = in && !((decided >> l) & 1ull) && (dep & ~decided) == 0ull;
        bool ok = false;
        if (ready) {
            const T lw = (T)(sp + (double)tot);
            ok = metropolis(ex(lw), u);
        }
        decided |= __ballot(ready);
        const unsigned long long na = __ballot(ok);
        am |= na;
        if (!((decided >> l) & 1ull)) {
            unsigned long long b = na & pm;
            while (b) {
                const int k = __ffsll((long long)b) - 1;
                b &= b - 1;
                tot += X.pc[si + __popcll(pm & ((1ull << k) - 1ull))];
            }
        }
    }
    return am;
}

// Clear the per-launch outputs of the listed chains (the kernels of a launch
// step add into them).
__global__ void k_clear_chain_meas(double *chain_meas, const int *ids, int n)
{
    const int i = blockIdx.x;
    if (i >= n) return;
    double *o = chain_meas + (size_t)ids[i] * (kNQ + kCnt);
    for (int k = threadIdx.x; k < kNQ + kCnt; k += blockDim.x) o[k] = 0.0;
}

// --- staging ---

// LDS of k_stage_win: old beads of the window's ell + 1 slices, new interior
// beads, bounds, Delta V and Metropolis uniform per particle, decisions, then
// the chunk machinery with pool capacity cap.
struct StageLds {
    size_t ob, nb, bnd, dv, su, acc;
    GridLds gr;
    ChunkLds ck;
};
__host__ __device__ inline StageLds stage_lds(int N, int ell, int dim, int tsize, int cap)
{
    StageLds L;
    size_t o = 0;
    L.ob = o;
    o = align16(o + (size_t)(ell + 1) * N * dim * tsize);
    L.nb = o;
    o = align16(o + (size_t)(ell > 1 ? ell - 1 : 1) * N * dim * tsize);
    L.bnd = o;
    o = align16(o + (size_t)N * 16);
    L.dv = o;
    o = align16(o + (size_t)N * tsize);
    L.su = o;
    o = align16(o + (size_t)N * tsize);
    L.acc = o;
    o = align16(o + (size_t)N);
    L.gr = grid_lds(o, N, dim);
    L.ck = chunk_lds(L.gr.end, N, cap, tsize, 2);
    return L;
}

// One staging window of one chain per workgroup: grid (windows, chains of the
// list a.whole), one sweep (a.rs->sweep + a.sweep_off).
template <typename T, int DIM> __global__ void __launch_bounds__(512) k_stage_win(SweepArgs a, int cap)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int c = a.whole[blockIdx.y];
    const ChainDev ch = a.chains[c];
    const int w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x, wave = tid >> 6;
    const int M = ch.M, N = ch.N, ell = ch.ell;
    if (w >= ch.nseg) return;
    const int m = min(ell, M - w * ell); // links of this window
    if (m < 2) return;                   // no interior bead
    const uint32_t sw = a.rs->sweep + (uint32_t)a.sweep_off;
    if (!pass_on(ch.every_st, sw)) return; // no staging pass in this sweep (move weights)
    if (cap < N) { // the host sizes the pool for every particle's partners
        if (tid == 0) atomicOr(a.status + c, 2);
        return;
    }
    const Prm<T> P = reinterpret_cast<const Prm<T> *>(a.prm)[c];
    const U4 xo = draw(a.seed, 0u, sw, (uint32_t)c, kOffset); // the chain's offset of this sweep
    const int off = (int)(((uint64_t)xo.x * (uint64_t)ell) >> 32);
    int s0 = off + w * ell;
    if (s0 >= M) s0 -= M;
    auto slice = [&](int t) { const int j = s0 + t; return j >= M ? j - M : j; };

    const StageLds L = stage_lds(N, ell, DIM, (int)sizeof(T), cap);
    T *ob = reinterpret_cast<T *>(smem + L.ob); // [t][k][d], t = 0..m
    T *nb = reinterpret_cast<T *>(smem + L.nb); // [t - 1][k][d], t = 1..m-1
    float4 *bnd = reinterpret_cast<float4 *>(smem + L.bnd);
    T *dvs = reinterpret_cast<T *>(smem + L.dv);
    T *su = reinterpret_cast<T *>(smem + L.su);
    unsigned char *acc = smem + L.acc;
    ChunkPtr<T> X(smem, L.ck);

    T *const gpath = reinterpret_cast<T *>(a.pos) + ch.pos_off;
    for (int t = 0; t <= m; ++t) {
        const T *src = gpath + (size_t)slice(t) * N * DIM;
        T *dst = ob + (size_t)t * N * DIM;
        for (int i = tid; i < N * DIM; i += nt) dst[i] = src[i];
    }
    __syncthreads();

    // proposals: the closed form of levy!'s recursion (src/updates/helper.jl:118-139): bead i
    // of a segment of m links from c0 to e is c0 + (i/m)(e - c0) + (m - i) S_i with
    // S_i = sum_{k <= i} sd_k xi_k / (m - k), sd_k = sqrt(2 lambda tau alpha_k),
    // alpha_k = (m - k)/(m - k + 1); Delta V of the interior beads; the window bound
    for (int p = tid; p < N; p += nt) {
        T c0[DIM], em[DIM], S[DIM], g1[2] = {T(0), T(0)};
        float ref[DIM];
#pragma unroll
        for (int e = 0; e < DIM; ++e) {
            c0[e] = ob[(size_t)p * DIM + e];
            em[e] = shift_end(c0[e], ob[((size_t)m * N + p) * DIM + e], P);
            S[e] = T(0);
            ref[e] = (float)c0[e];
        }
        T dv = T(0);
        for (int i = 1; i < m; ++i) {
            T xi[2];
            if (i & 1) {
                const U4 xr = draw(a.seed, (uint32_t)(p * M + slice(i)), sw, (uint32_t)c, kLevy);
                box_muller(xr.x, xr.y, xi[0], xi[1]);
                box_muller(xr.z, xr.w, g1[0], g1[1]);
            } else {
                xi[0] = g1[0];
                xi[1] = g1[1];
            }
            const T wi = bridge_sd(bridge_alpha<T>(m, i), P) / T(m - i);
            const T fi = T(i) / T(m);
            T rn[DIM], ro[DIM];
#pragma unroll
            for (int e = 0; e < DIM; ++e) {
                S[e] += wi * xi[e];
                rn[e] = teleport((c0[e] + fi * (em[e] - c0[e])) + T(m - i) * S[e], P);
                nb[((size_t)(i - 1) * N + p) * DIM + e] = rn[e];
                ro[e] = ob[((size_t)i * N + p) * DIM + e];
            }
            dv += -P.tau * (potV<T, DIM>(rn, P) - potV<T, DIM>(ro, P));
        }
        // bound of the old beads 0..m and the new interior beads, as offsets from bead 0
        auto offset = [&](int k, float *o) {
            const T *r = k <= m ? ob + ((size_t)k * N + p) * DIM : nb + ((size_t)(k - m - 1) * N + p) * DIM;
#pragma unroll
            for (int e = 0; e < DIM; ++e) o[e] = (float)signed_mindist(r[e], c0[e], P);
        };
        const int npts = 2 * m; // m + 1 old, m - 1 new
        float mean[DIM] = {};
        for (int k = 0; k < npts; ++k) {
            float o[DIM];
            offset(k, o);
#pragma unroll
            for (int e = 0; e < DIM; ++e) mean[e] += o[e];
        }
#pragma unroll
        for (int e = 0; e < DIM; ++e) mean[e] /= (float)npts;
        float r2 = 0.f;
        for (int k = 0; k < npts; ++k) {
            float o[DIM], d2 = 0.f;
            offset(k, o);
#pragma unroll
            for (int e = 0; e < DIM; ++e) d2 += (o[e] - mean[e]) * (o[e] - mean[e]);
            r2 = fmaxf(r2, d2);
        }
        bnd[p] = make_float4(ref[0] + mean[0], DIM > 1 ? ref[DIM - 1] + mean[DIM - 1] : 0.f, sqrtf(r2), 0.f);
        dvs[p] = dv;
        su[p] = Uniform<T>::get(draw(a.seed, (uint32_t)(p * M + s0), sw, (uint32_t)c, kLevyAccept).x);
        acc[p] = 0;
    }
    __syncthreads();

    const float h = (float)ch.h, twoL = (float)(2.0 * ch.L), inv2L = 1.0f / twoL;
    auto bead = [&](int p, int q) {
        return [=](int t, int who, T *ro, T *rn) {
            const int k = who ? q : p;
            const T *so = ob + ((size_t)t * N + k) * DIM;
            const T *sn = (t > 0 && t < m) ? nb + ((size_t)(t - 1) * N + k) * DIM : so;
#pragma unroll
            for (int e = 0; e < DIM; ++e) {
                ro[e] = so[e];
                rn[e] = sn[e];
            }
        };
    };
    WinGrid G(smem, L.gr);
    grid_build<DIM>(G, N, h, twoL, inv2L, [&](int p, float &x, float &y, float &r) {
        const float4 b = bnd[p];
        x = b.x;
        y = b.y;
        r = b.z;
    });
    int lacc = 0;
    for (int pa = 0; pa < N;) {
        const int nf = chunk_slots<DIM>(X, pa, N, cap, kChunk, [&](int p, int q) {
            if (q == p) return false;
            const float4 A = bnd[p], B = bnd[q];
            return bnd_near<DIM>(A.x, A.y, A.z, B.x, B.y, B.z, h, twoL, inv2L);
        }, &G);
        const int total = X.misc[1];
        for (int i = tid; i < total; i += nt) {
            const int p = pa + X.po[i], q = X.pq[i];
            T A, C;
            bool hit;
            pair_window<T, DIM, 1, true>(bead(p, q), m, q < p, P, ch.nbins, A, C, hit);
            X.pa[i] = A;
            X.pc[i] = C;
            if (q < p && hit) X.pq[i] = (unsigned short)(q | 0x8000);
        }
        __syncthreads();
        if (tid < nf) {
            const int p = pa + tid, o0 = X.off[tid], o1 = o0 + X.cnt[tid];
            double sp = (double)dvs[p];
            int e = o1;
            for (int i = o0; i < o1; ++i) {
                sp += (double)X.pa[i];
                if (e == o1 && (X.pq[i] & 0x7fff) > p) e = i;
            }
            X.spec[tid] = sp;
            X.nb[tid] = e;
        }
        __syncthreads();
        if (wave == 0) {
            const unsigned long long am = chunk_decide<T>(X, pa, nf, acc, [&](int p) { return su[p]; });
            const int l = tid;
            if (l < nf) acc[pa + l] = (unsigned char)((am >> l) & 1ull);
            lacc += __popcll(am);
        }
        __syncthreads();
        pa += nf;
    }
    // accepted proposals to HBM
    for (int i = 1; i < m; ++i) {
        T *dst = gpath + (size_t)slice(i) * N * DIM;
        const T *src = nb + (size_t)(i - 1) * N * DIM;
        for (int k = tid; k < N; k += nt)
            if (acc[k]) {
#pragma unroll
                for (int e = 0; e < DIM; ++e) dst[k * DIM + e] = src[k * DIM + e];
            }
    }
    if (tid == 0) atomicAdd(a.chain_meas + (size_t)c * (kNQ + kCnt) + kNQ, (double)lacc);
}

// --- centre of mass ---

// The try of particle p in this sweep: shift (words 0..dim-1) and uniform (word 2).
template <typename T, int DIM>
__device__ __forceinline__ void com_try(const SweepArgs &a, int c, uint32_t sw, int p, T *d, T &u)
{
    const U4 y = draw(a.seed, (uint32_t)p, sw, (uint32_t)c, kCom);
    const T s2 = (T)a.com_step[c] * T(2);
    d[0] = s2 * (Uniform<T>::get(y.x) - T(0.5));
    if (DIM > 1) d[1] = s2 * (Uniform<T>::get(y.y) - T(0.5));
    u = Uniform<T>::get(y.z);
}

struct ComSpecLds {
    size_t ob, sd, bnd, dv;
    GridLds gr;
    ChunkLds ck;
};
__host__ __device__ inline ComSpecLds com_spec_lds(int N, int lc, int dim, int tsize, int cap)
{
    ComSpecLds L;
    size_t o = 0;
    L.ob = o;
    o = align16(o + (size_t)(lc + 1) * N * dim * tsize);
    L.sd = o;
    o = align16(o + (size_t)N * dim * tsize);
    L.bnd = o;
    o = align16(o + (size_t)N * 16);
    L.dv = o;
    o = align16(o + (size_t)N * tsize);
    L.gr = grid_lds(o, N, dim);
    L.ck = chunk_lds(L.gr.end, N, cap, tsize, 1);
    return L;
}

// Phase 1 of the centre-of-mass moves: one workgroup per (window, chain),
// windows of com_win_len(M) slices from slice 0.  spec[part * nwin + w]: the
// window's Delta V + sum_q A of the particle's try against the configuration
// at the start of the phase; wbnd[part * nwin + w]: the bound of the
// particle's old beads in the window.
template <typename T, int DIM>
__global__ void __launch_bounds__(512) k_com_spec(SweepArgs a, double *spec, float4 *wbnd, int nwin, int cap, T *pm)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int c = a.whole[blockIdx.y];
    const ChainDev ch = a.chains[c];
    const int w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int M = ch.M, N = ch.N, lc = com_win_len(M);
    if (w * lc >= M) return;
    const int m = min(lc, M - w * lc), s0 = w * lc;
    const uint32_t sw = a.rs->sweep + (uint32_t)a.sweep_off;
    if (!pass_on(ch.every_com, sw)) return; // no centre-of-mass pass in this sweep
    if (cap < N) {
        if (tid == 0) atomicOr(a.status + c, 2);
        return;
    }
    const Prm<T> P = reinterpret_cast<const Prm<T> *>(a.prm)[c];
    const ComSpecLds L = com_spec_lds(N, lc, DIM, (int)sizeof(T), cap);
    T *ob = reinterpret_cast<T *>(smem + L.ob);
    T *sd = reinterpret_cast<T *>(smem + L.sd);
    float4 *bnd = reinterpret_cast<float4 *>(smem + L.bnd);
    T *dvs = reinterpret_cast<T *>(smem + L.dv);
    ChunkPtr<T> X(smem, L.ck);
    const T *const gpath = reinterpret_cast<const T *>(a.pos) + ch.pos_off;
    for (int t = 0; t <= m; ++t) {
        const int j = (s0 + t) % M;
        const T *src = gpath + (size_t)j * N * DIM;
        T *dst = ob + (size_t)t * N * DIM;
        for (int i = tid; i < N * DIM; i += nt) dst[i] = src[i];
    }
    __syncthreads();
    // the window's slices particle-major for the decision kernel: a particle's
    // consecutive beads share cache lines there
    {
        T *dst = pm + ch.pos_off;
        for (int it = tid; it < N * m; it += nt) {
            const int p = it / m, t = it - p * m;
#pragma unroll
            for (int e = 0; e < DIM; ++e) dst[((size_t)p * M + s0 + t) * DIM + e] = ob[((size_t)t * N + p) * DIM + e];
        }
    }
    for (int p = tid; p < N; p += nt) {
        T d[DIM], u;
        com_try<T, DIM>(a, c, sw, p, d, u);
        T dv = T(0);
        float mean[DIM] = {};
        for (int t = 0; t < m; ++t) {
            T ro[DIM], rn[DIM];
#pragma unroll
            for (int e = 0; e < DIM; ++e) {
                ro[e] = ob[((size_t)t * N + p) * DIM + e];
                rn[e] = teleport(ro[e] + d[e], P);
            }
            dv += -P.tau * (potV<T, DIM>(rn, P) - potV<T, DIM>(ro, P));
        }
        for (int t = 0; t <= m; ++t) {
#pragma unroll
            for (int e = 0; e < DIM; ++e)
                mean[e] += (float)signed_mindist(ob[((size_t)t * N + p) * DIM + e], ob[(size_t)p * DIM + e], P);
        }
#pragma unroll
        for (int e = 0; e < DIM; ++e) mean[e] /= (float)(m + 1);
        float r2 = 0.f;
        for (int t = 0; t <= m; ++t) {
            float d2 = 0.f;
#pragma unroll
            for (int e = 0; e < DIM; ++e) {
                const float o = (float)signed_mindist(ob[((size_t)t * N + p) * DIM + e], ob[(size_t)p * DIM + e], P) - mean[e];
                d2 += o * o;
            }
            r2 = fmaxf(r2, d2);
        }
        const float4 b = make_float4((float)ob[(size_t)p * DIM] + mean[0],
                                     DIM > 1 ? (float)ob[(size_t)p * DIM + DIM - 1] + mean[DIM - 1] : 0.f, sqrtf(r2), 0.f);
        bnd[p] = b;
        wbnd[((size_t)ch.part_base + p) * nwin + w] = b;
#pragma unroll
        for (int e = 0; e < DIM; ++e) sd[p * DIM + e] = d[e];
        dvs[p] = dv;
    }
    __syncthreads();
    const float h = (float)ch.h, twoL = (float)(2.0 * ch.L), inv2L = 1.0f / twoL;
    // grid of the bounds of old and shifted beads together (centre halfway, radius
    // grown by half the shift): both of a try's pairings lie within them
    WinGrid G(smem, L.gr);
    grid_build<DIM>(G, N, h, twoL, inv2L, [&](int p, float &x, float &y, float &r) {
        const float4 b = bnd[p];
        const float dx = (float)sd[p * DIM], dy = DIM > 1 ? (float)sd[p * DIM + DIM - 1] : 0.f;
        x = b.x + 0.5f * dx;
        y = b.y + 0.5f * dy;
        r = (b.z + 0.5f * sqrtf(dx * dx + dy * dy)) * 1.00001f;
    });
    for (int pa = 0; pa < N;) {
        const int nf = chunk_slots<DIM>(X, pa, N, cap, kChunk, [&](int p, int q) {
            if (q == p) return false;
            const float4 A = bnd[p], B = bnd[q];
            const float dx = (float)sd[p * DIM], dy = DIM > 1 ? (float)sd[p * DIM + DIM - 1] : 0.f;
            return bnd_near<DIM>(A.x, A.y, A.z, B.x, B.y, B.z, h, twoL, inv2L) ||
                   bnd_near<DIM>(A.x + dx, A.y + dy, A.z, B.x, B.y, B.z, h, twoL, inv2L);
        }, &G);
        const int total = X.misc[1];
        for (int i = tid; i < total; i += nt) {
            const int p = pa + X.po[i], q = X.pq[i];
            auto bead = [&](int t, int who, T *ro, T *rn) {
                const int k = who ? q : p;
#pragma unroll
                for (int e = 0; e < DIM; ++e) {
                    ro[e] = ob[((size_t)t * N + k) * DIM + e];
                    rn[e] = teleport(ro[e] + sd[k * DIM + e], P);
                }
            };
            T A, C;
            bool hit;
            pair_window<T, DIM, 1>(bead, m, false, P, ch.nbins, A, C, hit);
            X.pa[i] = A;
        }
        __syncthreads();
        if (tid < nf) {
            const int p = pa + tid, o0 = X.off[tid], o1 = o0 + X.cnt[tid];
            double sp = (double)dvs[p];
            for (int i = o0; i < o1; ++i) sp += (double)X.pa[i];
            spec[((size_t)ch.part_base + p) * nwin + w] = sp;
        }
        __syncthreads();
        pa += nf;
    }
}

// Items of the decision kernel: (slot, window) pairs whose bounds come within
// reach, taken kItemBatch at a time.
constexpr int kItemBatch = 2048;

struct ComDecLds {
    size_t spec, su, sd, rb, acc, pwb, wm, ioff, hitf, cw, hw;
    ChunkLds ck;
};
__host__ __device__ inline ComDecLds com_dec_lds(int N, int dim, int tsize, int cap)
{
    ComDecLds L;
    size_t o = 0;
    L.spec = o;
    o = align16(o + (size_t)N * 8);
    L.su = o;
    o = align16(o + (size_t)N * tsize);
    L.sd = o;
    o = align16(o + (size_t)N * dim * tsize);
    L.rb = o;
    o = align16(o + (size_t)N * 16);
    L.acc = o;
    o = align16(o + (size_t)N);
    L.pwb = o; // window bounds of the chunk's particles, kChunk x 32
    o = align16(o + (size_t)kChunk * 32 * 16);
    L.wm = o; // per slot: mask of the windows to evaluate
    o = align16(o + (size_t)cap * 4);
    L.ioff = o; // per slot: first item
    o = align16(o + (size_t)cap * 4);
    L.hitf = o; // per slot: a term within the cut-off
    o = align16(o + (size_t)cap);
    L.cw = o; // per item of a batch: C of the window
    o = align16(o + (size_t)kItemBatch * tsize);
    L.hw = o;
    o = align16(o + (size_t)kItemBatch);
    L.ck = chunk_lds(o, N, cap, tsize, 1);
    return L;
}

// Phase 2 of the centre-of-mass moves: one workgroup per chain.  Per try, the
// C terms of its accepted earlier partners whose ring bounds (old or shifted)
// come within the cut-off, summed over the windows where their window bounds
// do; the tries are decided in chunks of rounds.  cacc[part]: 1 when the
// particle's try was accepted, cshift[part * 2 + e] its shift.
template <typename T, int DIM>
__global__ void __launch_bounds__(1024) k_com_dec(SweepArgs a, const double *spec, const float4 *wbnd, int nwin,
                                                   int cap, int kc, unsigned char *cacc, T *cshift, const T *pm)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int c = a.whole[blockIdx.x];
    const ChainDev ch = a.chains[c];
    const int N = ch.N, M = ch.M, tid = threadIdx.x, nt = blockDim.x, wave = tid >> 6;
    const int lc = com_win_len(M), nw = com_nwin(M);
    const uint32_t sw = a.rs->sweep + (uint32_t)a.sweep_off;
    if (!pass_on(ch.every_com, sw)) return; // no centre-of-mass pass in this sweep: no tries
    if (cap < N) {
        if (tid == 0) atomicOr(a.status + c, 2);
        return;
    }
    const Prm<T> P = reinterpret_cast<const Prm<T> *>(a.prm)[c];
    const ComDecLds L = com_dec_lds(N, DIM, (int)sizeof(T), cap);
    double *s_spec = reinterpret_cast<double *>(smem + L.spec);
    T *su = reinterpret_cast<T *>(smem + L.su);
    T *sd = reinterpret_cast<T *>(smem + L.sd);
    float4 *rb = reinterpret_cast<float4 *>(smem + L.rb);
    unsigned char *acc = smem + L.acc;
    float4 *pwb = reinterpret_cast<float4 *>(smem + L.pwb);
    unsigned *s_wm = reinterpret_cast<unsigned *>(smem + L.wm);
    int *s_ioff = reinterpret_cast<int *>(smem + L.ioff);
    unsigned char *s_hitf = smem + L.hitf;
    T *s_cw = reinterpret_cast<T *>(smem + L.cw);
    unsigned char *s_hw = smem + L.hw;
    ChunkPtr<T> X(smem, L.ck);
    const T *ring = pm + ch.pos_off; // particle-major copy of the path: [particle][slice][d]
    const float h = (float)ch.h, twoL = (float)(2.0 * ch.L), inv2L = 1.0f / twoL;
    const float4 *wb = wbnd + (size_t)ch.part_base * nwin;
    // per try: phase 1's value (windows in order), shift and uniform, and the ring bound
    // from the window bounds (centres unwrapped against window 0's)
    for (int p = tid; p < N; p += nt) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += spec[((size_t)ch.part_base + p) * nwin + w];
        s_spec[p] = s;
        T d[DIM], u;
        com_try<T, DIM>(a, c, sw, p, d, u);
#pragma unroll
        for (int e = 0; e < DIM; ++e) sd[p * DIM + e] = d[e];
        su[p] = u;
        acc[p] = 0;
        const float4 b0 = wb[(size_t)p * nwin];
        float sx = 0.f, sy = 0.f;
        for (int w = 0; w < nw; ++w) {
            const float4 b = wb[(size_t)p * nwin + w];
            sx += mi_f(b.x - b0.x, twoL, inv2L);
            if (DIM > 1) sy += mi_f(b.y - b0.y, twoL, inv2L);
        }
        sx /= (float)nw;
        sy /= (float)nw;
        float r = 0.f;
        for (int w = 0; w < nw; ++w) {
            const float4 b = wb[(size_t)p * nwin + w];
            const float dx = mi_f(b.x - b0.x, twoL, inv2L) - sx, dy = DIM > 1 ? mi_f(b.y - b0.y, twoL, inv2L) - sy : 0.f;
            r = fmaxf(r, sqrtf(dx * dx + dy * dy) + b.z);
        }
        rb[p] = make_float4(b0.x + sx, b0.y + sy, r * 1.00001f, 0.f);
    }
    __syncthreads();
    auto fshift = [&](int k, float *f) {
        f[0] = (float)sd[k * DIM];
        if (DIM > 1) f[1] = (float)sd[k * DIM + DIM - 1];
    };
    int nacc = 0;
    for (int pa = 0; pa < N;) {
        const int nf = chunk_slots(X, pa, N, cap, kc, [&](int p, int q) {
            if (q >= p) return false;
            if (q < pa && !acc[q]) return false;
            float dp[2], dq[2];
            fshift(p, dp);
            fshift(q, dq);
            return bnd_near4<DIM>(rb[p], dp, rb[q], dq, h, twoL, inv2L);
        });
        const int total = X.misc[1];
        for (int it = tid; it < nf * nw; it += nt) {
            const int l = it / nw, w = it - l * nw;
            pwb[l * 32 + w] = wb[(size_t)(pa + l) * nwin + w];
        }
        __syncthreads();
        // the windows of each slot whose bounds (old or shifted) come within reach:
        // one thread per (slot, 4 windows), q's bounds fetched together
        for (int i = tid; i < total; i += nt) {
            s_wm[i] = 0u;
            s_hitf[i] = 0;
            X.pc[i] = T(0);
        }
        __syncthreads();
        const int ngrp = (nw + 3) / 4;
        for (int it = tid; it < total * ngrp; it += nt) {
            const int i = it / ngrp, w0 = (it - i * ngrp) * 4;
            const int l = X.po[i], p = pa + l, q = X.pq[i];
            float dp[2], dq[2];
            fshift(p, dp);
            fshift(q, dq);
            float4 bq[4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (w0 + k < nw) bq[k] = wb[(size_t)q * nwin + w0 + k];
            unsigned bits = 0u;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (w0 + k < nw && bnd_near4<DIM>(pwb[l * 32 + w0 + k], dp, bq[k], dq, h, twoL, inv2L))
                    bits |= 1u << (w0 + k);
            if (bits) atomicOr(s_wm + i, bits);
        }
        __syncthreads();
        // items = (slot, window) in slot order, windows ascending within a slot
        if (wave == 0) {
            const int lane = tid;
            int base = 0;
            for (int i0 = 0; i0 < total; i0 += 64) {
                const int i = i0 + lane;
                const int v = i < total ? __popc(s_wm[i]) : 0;
                int tot;
                const int pre = wave_prefix(v, lane, tot);
                if (i < total) s_ioff[i] = base + pre;
                base += tot;
            }
            if (lane == 0) X.misc[2] = base;
        }
        __syncthreads();
        const int nitems = X.misc[2];
        for (int k0 = 0; k0 < nitems; k0 += kItemBatch) {
            const int kn = min(kItemBatch, nitems - k0);
            for (int kk = tid; kk < kn; kk += nt) {
                const int k = k0 + kk;
                int lo = 0, hi = total - 1; // the last slot whose first item is <= k
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_ioff[mid] <= k) lo = mid;
                    else hi = mid - 1;
                }
                const int i = lo;
                unsigned wm = s_wm[i];
                for (int r = k - s_ioff[i]; r > 0; --r) wm &= wm - 1u;
                const int w = __ffs((int)wm) - 1;
                const int p = pa + X.po[i], q = X.pq[i];
                const int s0 = w * lc, m = min(lc, M - s0);
                auto bead = [&](int t, int who, T *ro, T *rn) {
                    const int kp = who ? q : p;
                    int j = s0 + t;
                    if (j >= M) j -= M;
                    const T *r = ring + ((size_t)kp * M + j) * DIM;
                    if constexpr (DIM == 2) {
                        using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
                        const V2 v = *reinterpret_cast<const V2 *>(r);
                        ro[0] = v.x;
                        ro[1] = v.y;
                    } else {
                        ro[0] = r[0];
                    }
#pragma unroll
                    for (int e = 0; e < DIM; ++e) rn[e] = teleport(ro[e] + sd[kp * DIM + e], P);
                };
                T A, Cw;
                bool hw;
                pair_window<T, DIM, 4>(bead, m, true, P, ch.nbins, A, Cw, hw);
                s_cw[kk] = Cw;
                s_hw[kk] = hw ? 1 : 0;
            }
            __syncthreads();
            // each slot adds its items of this batch in window order
            for (int i = tid; i < total; i += nt) {
                const int b0 = max(s_ioff[i], k0), b1 = min(s_ioff[i] + __popc(s_wm[i]), k0 + kn);
                T acc_c = X.pc[i];
                unsigned char hf = s_hitf[i];
                for (int k = b0; k < b1; ++k) {
                    acc_c += s_cw[k - k0];
                    hf |= s_hw[k - k0];
                }
                X.pc[i] = acc_c;
                s_hitf[i] = hf;
            }
            __syncthreads();
        }
        for (int i = tid; i < total; i += nt)
            if (s_hitf[i]) X.pq[i] = (unsigned short)(X.pq[i] | 0x8000);
        __syncthreads();
        if (tid < nf) {
            X.spec[tid] = s_spec[pa + tid];
            X.nb[tid] = X.off[tid] + X.cnt[tid];
        }
        __syncthreads();
        if (wave == 0) {
            const unsigned long long am = chunk_decide<T>(X, pa, nf, acc, [&](int p) { return su[p]; });
            const int l = tid;
            if (l < nf) acc[pa + l] = (unsigned char)((am >> l) & 1ull);
            nacc += __popcll(am);
        }
        __syncthreads();
        pa += nf;
    }
    for (int p = tid; p < N; p += nt) {
        const size_t q = (size_t)ch.part_base + p;
        cacc[q] = acc[p];
#pragma unroll
        for (int e = 0; e < DIM; ++e) cshift[q * 2 + e] = sd[p * DIM + e];
    }
    if (tid == 0) {
        double *out = a.chain_meas + (size_t)c * (kNQ + kCnt);
        out[kNQ + 1] += (double)nacc;
        out[kNQ + 2] += (double)N;
    }
}

// The accepted centre-of-mass tries move their rings: grid (bead blocks,
// chains), one thread per (slice, particle).
template <typename T, int DIM>
__global__ void __launch_bounds__(256) k_com_apply(SweepArgs a, const unsigned char *cacc, const T *cshift)
{
    const int c = a.whole[blockIdx.y];
    const ChainDev ch = a.chains[c];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)ch.M * ch.N) return;
    const int j = (int)(i / ch.N), k = (int)(i - (long long)j * ch.N);
    const size_t q = (size_t)ch.part_base + k;
    // the flags are those of this sweep's k_com_dec only when the pass ran
    if (!pass_on(ch.every_com, a.rs->sweep + (uint32_t)a.sweep_off) || !cacc[q]) return;
    const Prm<T> P = reinterpret_cast<const Prm<T> *>(a.prm)[c];
    const Path<T, DIM, true> path{reinterpret_cast<T *>(a.pos) + ch.pos_off, ch.M, ch.N};
    T r[DIM];
    path.get(k, j, r);
#pragma unroll
    for (int e = 0; e < DIM; ++e) r[e] = teleport(r[e] + cshift[q * 2 + e], P);
    path.put(k, j, r);
}

// --- estimators ---

struct MeasWinLds {
    size_t ob, bnd;
    GridLds gr;
    ChunkLds ck;
};
__host__ __device__ inline MeasWinLds meas_win_lds(int N, int lc, int dim, int tsize, int cap)
{
    MeasWinLds L;
    size_t o = 0;
    L.ob = o;
    o = align16(o + (size_t)(lc + 1) * N * dim * tsize);
    L.bnd = o;
    o = align16(o + (size_t)N * 16);
    L.gr = grid_lds(o, N, dim);
    L.ck = chunk_lds(L.gr.end, N, cap, tsize, 1);
    return L;
}

// Estimators of Gaussian-pair chains after a sweep, by the centre-of-mass
// windows: one 512-thread workgroup per (window, chain of a.ids).  The terms of
// measure_chain for the beads of slices s0 .. s0 + m - 1: kinetic link to the
// next slice, V, r.grad V, dV/d omega, the radial and Cartesian histograms,
// and the pair energy and pair virial of each bead as pair_energy_bead takes
// them (every neighbour within the cut-off at the bead's slice, weight
// tau/4 times its stencil multiplicity), the neighbours found from window
// bounds.  part[(y * nwin + w) * 6]: the window's sums (link, V, r.grad V,
// S_U/g, pair virial, dV/d omega), folded in window order by k_meas_fold.
template <typename T, int DIM> __global__ void __launch_bounds__(512) k_meas_win(MeasArgs a, int nwin, int cap)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double wbuf[32];
    if (!split_measured(a)) return;
    const int y = blockIdx.y, c = a.ids[y], w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const ChainDev ch = a.chains[c];
    const int N = ch.N, M = ch.M, lc = com_win_len(M);
    double *out = a.part + ((size_t)y * nwin + w) * 6;
    if (w * lc >= M || cap < N) { // no such window for this chain (the host sizes cap >= N)
        if (tid < 6) out[tid] = 0.0;
        return;
    }
    const int m = min(lc, M - w * lc), s0 = w * lc;
    const Prm<T> P = reinterpret_cast<const Prm<T> *>(a.prm)[c];
    const MeasWinLds L = meas_win_lds(N, lc, DIM, (int)sizeof(T), cap);
    T *ob = reinterpret_cast<T *>(smem + L.ob);
    float4 *bnd = reinterpret_cast<float4 *>(smem + L.bnd);
    ChunkPtr<T> X(smem, L.ck);
    const T *const gpath = reinterpret_cast<const T *>(a.pos) + ch.pos_off;
    for (int t = 0; t <= m; ++t) {
        const int j = (s0 + t) % M;
        const T *src = gpath + (size_t)j * N * DIM;
        T *dst = ob + (size_t)t * N * DIM;
        for (int i = tid; i < N * DIM; i += nt) dst[i] = src[i];
    }
    __syncthreads();
    for (int p = tid; p < N; p += nt) {
        float mean[DIM] = {};
        for (int t = 0; t <= m; ++t) {
#pragma unroll
            for (int e = 0; e < DIM; ++e)
                mean[e] += (float)signed_mindist(ob[((size_t)t * N + p) * DIM + e], ob[(size_t)p * DIM + e], P);
        }
#pragma unroll
        for (int e = 0; e < DIM; ++e) mean[e] /= (float)(m + 1);
        float r2 = 0.f;
        for (int t = 0; t <= m; ++t) {
            float d2 = 0.f;
#pragma unroll
            for (int e = 0; e < DIM; ++e) {
                const float o = (float)signed_mindist(ob[((size_t)t * N + p) * DIM + e], ob[(size_t)p * DIM + e], P) - mean[e];
                d2 += o * o;
            }
            r2 = fmaxf(r2, d2);
        }
        bnd[p] = make_float4((float)ob[(size_t)p * DIM] + mean[0],
                             DIM > 1 ? (float)ob[(size_t)p * DIM + DIM - 1] + mean[DIM - 1] : 0.f, sqrtf(r2), 0.f);
    }
    // single-bead terms, one thread per (particle, slice)
    double l = 0.0, v = 0.0, g = 0.0, r2w = 0.0, su = 0.0, rgp = 0.0;
    for (int it = tid; it < N * m; it += nt) {
        const int p = it / m, t = it - p * m;
        T r0[DIM], r1[DIM];
#pragma unroll
        for (int e = 0; e < DIM; ++e) {
            r0[e] = ob[((size_t)t * N + p) * DIM + e];
            r1[e] = ob[((size_t)(t + 1) * N + p) * DIM + e];
        }
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
        if (a.radial_bins > 0) {
            T rr = r0[0] * r0[0];
            if (DIM > 1) rr += r0[1] * r0[1];
            const double qb = (double)sq(rr) * a.radial_winv;
            if (qb < (double)a.radial_bins) atomicAdd(&a.radial[(size_t)c * a.radial_bins + (int)qb], 1ull);
        }
        if (a.cart_bins > 0) {
            const double bw = (2.0 * ch.L) / a.cart_bins;
            long long idx = 0, stride = 1;
            bool ok = true;
#pragma unroll
            for (int e = 0; e < DIM; ++e) {
                const long long ib = (long long)floor(((double)r0[e] + ch.L) / bw);
                ok = ok && ib > 0 && ib < a.cart_bins + 1;
                idx += (ib - 1) * stride;
                stride *= a.cart_bins;
            }
            if (ok) atomicAdd(&a.cart[(size_t)c * stride + idx], 1ull);
        }
    }
    __syncthreads();
    // pair terms: bead (p, t) with each neighbour q within the cut-off at slice t.  The term is
    // symmetric in p and q (distances, stencil multiplicity, the link to the next slice), so each
    // unordered pair is taken once, from its lower index, with the weight of both ends
    const float h = (float)ch.h, twoL = (float)(2.0 * ch.L), inv2L = 1.0f / twoL;
    const double inv_s2 = 1.0 / (ch.sigma * ch.sigma), wq = 2.0 * (0.5 * (0.5 * (double)P.tau));
    WinGrid G(smem, L.gr);
    grid_build<DIM>(G, N, h, twoL, inv2L, [&](int p, float &x, float &y, float &r) {
        const float4 b = bnd[p];
        x = b.x;
        y = b.y;
        r = b.z;
    });
    for (int pa = 0; pa < N;) {
        const int nf = chunk_slots<DIM>(X, pa, N, cap, kChunk, [&](int p, int q) {
            if (q <= p) return false;
            const float4 A = bnd[p], B = bnd[q];
            return bnd_near<DIM>(A.x, A.y, A.z, B.x, B.y, B.z, h, twoL, inv2L);
        }, &G);
        const int total = X.misc[1];
        for (int i = tid; i < total; i += nt) {
            const int p = pa + X.po[i], q = X.pq[i];
            for (int t = 0; t < m; ++t) {
                const T *rp = ob + ((size_t)t * N + p) * DIM, *rq = ob + ((size_t)t * N + q) * DIM;
                const T r2 = dist2<T, DIM>(rq, rp, P);
                if (!(r2 <= P.h2)) continue;
                const T mult = stencil_mult<T, DIM>(rp, rq, P, ch.nbins);
                const T s2 = dist2<T, DIM>(rq + (size_t)N * DIM, rp + (size_t)N * DIM, P);
                const double ea = (double)ex(-r2 * P.inv2s2), eb = (double)ex(-s2 * P.inv2s2);
                const double wt = wq * (double)mult;
                su += wt * (ea + eb);
                rgp += wt * (-((double)r2 * inv_s2) * ea - ((double)s2 * inv_s2) * eb);
            }
        }
        __syncthreads();
        pa += nf;
    }
    const double sl = block_sum(l, wbuf), sv = block_sum(v, wbuf), sg = block_sum(g, wbuf);
    const double ssu = block_sum(su, wbuf), srg = block_sum(rgp, wbuf), sw = block_sum(r2w, wbuf);
    if (tid == 0) {
        out[0] = sl;
        out[1] = sv;
        out[2] = sg;
        out[3] = ssu;
        out[4] = srg;
        out[5] = sw;
    }
}

// --- cell lists ---

// Rebuild the cell lists of the listed chains (those with cell lists): one
// wave per slice, grid (slice blocks, chains), LDS of ncell counts per wave.
// Members of each cell by particle index; a cell beyond cap is reported
// through the status word as in the incremental updates.
template <typename T, int DIM>
__global__ void k_cells_win(const ChainDev *chains, const void *prm, const int *ids, int n, const void *pos,
                            int *cell_cnt, unsigned short *cell_mem, int *bins, int *status, int ncell_max)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int c = ids[blockIdx.y];
    const ChainDev ch = chains[c];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const int j = blockIdx.x * nwv + wave;
    if (ch.ncell <= 0 || j >= ch.M) return;
    int *cnt = reinterpret_cast<int *>(smem) + (size_t)wave * ncell_max;
    const Prm<T> P = reinterpret_cast<const Prm<T> *>(prm)[c];
    const Path<T, DIM, true> path{const_cast<T *>(reinterpret_cast<const T *>(pos)) + ch.pos_off, ch.M, ch.N};
    const Cells C{cell_cnt + ch.cell_off, cell_mem + ch.mem_off, bins + ch.bin_off, ch.M, ch.ncell, ch.cap,
                  ch.nbins};
    for (int cc = lane; cc < ch.ncell; cc += 64) cnt[cc] = 0;
    int nbits = 1;
    while ((1 << nbits) < ch.ncell) ++nbits;
    bool over = false;
    for (int k0 = 0; k0 < ch.N; k0 += 64) {
        const int k = k0 + lane;
        const bool act = k < ch.N;
        int key = 0;
        if (act) {
            T r[DIM];
            path.get(k, j, r);
            key = cell_of<T, DIM>(r, P, ch.nbins);
        }
        // lanes with the same cell
        unsigned long long peers = __ballot(act);
        for (int b = 0; b < nbits; ++b) {
            const bool bit = (key >> b) & 1;
            const unsigned long long bal = __ballot(act && bit);
            peers &= bit ? bal : ~bal;
        }
        const int rank = __popcll(peers & ((1ull << lane) - 1ull));
        const int base = act ? cnt[key] : 0;
        const int pos = base + rank;
        if (act && rank == 0) cnt[key] = base + __popcll(peers);
        if (act) {
            if (pos < ch.cap) {
                C.members(j, key)[pos] = (unsigned short)k;
                C.bin[(long long)k * ch.M + j] = key;
            } else {
                C.bin[(long long)k * ch.M + j] = -1;
                over = true;
            }
        }
    }
    for (int cc = lane; cc < ch.ncell; cc += 64) C.count(j, cc) = min(cnt[cc], ch.cap);
    if (__ballot(over) && lane == 0) atomicOr(status + c, 1);
}

} // namespace pimc
