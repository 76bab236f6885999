// Device-side physics of the primitive-action PIMC sampling path.
//
// Every function here is shared by the batched sweep kernels and the
// single-move kernels behind the C ABI, so the parity checks of the single
// moves also cover the arithmetic of the sweeps.  Reference citations are
// file:line in the PIMC.jl tree.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "philox.hpp"

namespace pimc {

constexpr int kPotNone = 0, kPotHarmonic = 1, kPotLattice = 2, kPotOptical = 3;
constexpr int kMaxBeams = 16;
constexpr int kPairNone = 0, kPairGaussian = 1, kPairTable = 2;

// One chain as the kernels see it (device-resident table, 8-byte fields
// first).  Positions of particle n, coordinate d, slice j sit at
// pos[pos_off + (n * dim + d) * M + j]: each particle an M x dim column-major
// block, so consecutive slices are consecutive addresses.
struct ChainDev {
    double beta, tau, lambda, L;
    double omega, V0, a, g;
    double sigma, h;
    long long pos_off;  // element offset of the chain in the position buffer
    long long cell_off; // offset of slice 0, cell 0 in the cell-count array (pair chains)
    long long mem_off;  // offset of slice 0, cell 0, slot 0 in the cell-member array
    long long bin_off;  // offset of the chain's per-bead cell index array (pair chains)
    const double *tab;  // kPairTable: integral terms of the pair-product propagator, tab_n x tab_n
    double tab_r0, tab_h, mu; // grid start and spacing; lnU where the propagator is negative
    double hs_a;        // hard-sphere radius of hardspherelevy!/move_polymer! (0: off)
    const void *beam;   // kPotOptical: nbeam triples (kx, ky, phase) in revolutions, in the arithmetic type
    int N, M, dim, pot;
    int pair, pair_in_accept, ell, nbins;
    int cap, ncell, part_base, nseg;
    int sm;                   // positions stored slice-major (chains run whole per workgroup)
    int swaps, swap_m, swap_rule; // swap tries per sweep, links per try, acceptance rule
    int tab_n;
    int nbeam;
    int every_st, every_com, every_sw; // passes run on the sweeps s with s % every == 0
};

// Whether a pass with move weight 1/every runs in sweep sw (every >= 1).
__host__ __device__ __forceinline__ bool pass_on(int every, uint32_t sw)
{
    return every == 1 || sw % (uint32_t)every == 0u;
}

// Sweeps s in [s0, s0 + n) with s % every == 0 (every >= 1).
__host__ __device__ inline long long every_count(unsigned s0, int n, int every)
{
    const unsigned long long e = (unsigned long long)every, a = s0, b = (unsigned long long)s0 + (unsigned)n;
    return (long long)((b + e - 1) / e - (a + e - 1) / e);
}

// The chain constants converted to the arithmetic type (built once per chain
// on the host, so the kernels load them instead of dividing per item).
template <typename T> struct Prm {
    T tau, lambda, L, twoL, inv2L;
    T halfw2, w2;        // omega^2 / 2, omega^2
    T V0, inva;          // lattice depth, 1/a
    T g, inv2s2, h2;     // pair strength, 1/(2 sigma^2), cut-off radius^2
    T sdl;               // sqrt(2 lambda tau): free-particle displacement of one link
    const double *tab;   // tabulated pair-product propagator (kPairTable), FP64 throughout
    double tab_r0, tab_h, mu, tau64;
    T hs2;               // hard-sphere radius^2
    T cellw;             // cell side 2L / nbins (pair chains)
    const T *beam;       // optical lattice: (kx, ky, phase) per beam, revolutions
    T nbeamT;            // beam count
    int nbeam;
    int tab_n;
    int pot, pair;

    __host__ __device__ __forceinline__ static Prm make(const ChainDev &c)
    {
        Prm p;
        p.tau = (T)c.tau;
        p.lambda = (T)c.lambda;
        p.L = (T)c.L;
        p.twoL = (T)(2.0 * c.L);
        p.inv2L = (T)(1.0 / (2.0 * c.L));
        p.halfw2 = (T)(0.5 * c.omega * c.omega);
        p.w2 = (T)(c.omega * c.omega);
        p.V0 = (T)c.V0;
        p.inva = (T)(1.0 / c.a);
        p.g = (T)c.g;
        p.inv2s2 = (T)(1.0 / (2.0 * c.sigma * c.sigma));
        p.h2 = (T)(c.h * c.h);
        p.sdl = (T)sqrt(2.0 * c.lambda * c.tau);
        p.tab = c.tab;
        p.tab_r0 = c.tab_r0;
        p.tab_h = c.tab_h;
        p.mu = c.mu;
        p.tau64 = c.tau;
        p.hs2 = (T)(c.hs_a * c.hs_a);
        p.cellw = p.twoL / (T)c.nbins;
        p.beam = static_cast<const T *>(c.beam);
        p.nbeam = c.nbeam;
        p.nbeamT = (T)c.nbeam;
        p.tab_n = c.tab_n;
        p.pot = c.pot;
        p.pair = c.pair;
        return p;
    }
};

__device__ __forceinline__ double fl(double x) { return floor(x); }
__device__ __forceinline__ float fl(float x) { return floorf(x); }
__device__ __forceinline__ double cpi(double x) { return cospi(x); }
__device__ __forceinline__ float cpi(float x) { return cospif(x); }
__device__ __forceinline__ double spi(double x) { return sinpi(x); }
__device__ __forceinline__ float spi(float x) { return sinpif(x); }
__device__ __forceinline__ double ex(double x) { return exp(x); }
__device__ __forceinline__ float ex(float x) { return __expf(x); }
__device__ __forceinline__ double sq(double x) { return sqrt(x); }
__device__ __forceinline__ float sq(float x) { return __builtin_amdgcn_sqrtf(x); } // v_sqrt_f32, 1 ulp
// sin and cos of 2 pi t: FP64 through sinpi/cospi, FP32 through v_sin/v_cos (revolutions) of
// the fractional part (the hardware range is |t| <= 256)
__device__ __forceinline__ void sincos2pi(double t, double &s, double &c)
{
    s = sinpi(2.0 * t);
    c = cospi(2.0 * t);
}
__device__ __forceinline__ void sincos2pi(float t, float &s, float &c)
{
    const float f = __builtin_amdgcn_fractf(t);
    s = __builtin_amdgcn_sinf(f);
    c = __builtin_amdgcn_cosf(f);
}

// The optical-lattice intensity normalized_intensity (examples/tools/
// potentialtools.jl:1-16): S = sum_b sin(2 pi t_b), C = sum_b cos(2 pi t_b),
// t_b = scale (x sin a_b + y cos a_b) (+ a_b / 2 pi for helical beams);
// V = V0 ((S/n)^2 + (C/n)^2).  With grad, also dS/dx_d and dC/dx_d.
template <typename T, bool GRAD>
__device__ __forceinline__ void optical_sums(const T *r, const Prm<T> &p, T &S, T &Cc, T *dS, T *dC)
{
    S = T(0);
    Cc = T(0);
    if (GRAD) dS[0] = dS[1] = dC[0] = dC[1] = T(0);
    for (int b = 0; b < p.nbeam; ++b) {
        const T kx = p.beam[3 * b], ky = p.beam[3 * b + 1];
        const T t = kx * r[0] + ky * r[1] + p.beam[3 * b + 2];
        T sn, cs;
        sincos2pi(t, sn, cs);
        S += sn;
        Cc += cs;
        if (GRAD) {
            dS[0] += cs * kx;
            dS[1] += cs * ky;
            dC[0] -= sn * kx;
            dC[1] -= sn * ky;
        }
    }
}

// src/propagator.jl:30-32, teleport(x) = (x + L) - floor(x / 2L + 1/2) 2L - L
template <typename T> __device__ __forceinline__ T teleport(T x, const Prm<T> &p)
{
    return ((x + p.L) - fl(x * p.inv2L + T(0.5)) * p.twoL) - p.L;
}

// src/propagator.jl:6-9, per-component |minimum-image| offset
template <typename T> __device__ __forceinline__ T mindist(T x1, T x2, const Prm<T> &p)
{
    const T dx = fabs(x1 - x2);
    const T a = p.twoL - dx;
    return a < dx ? a : dx;
}

// External potential V(r) (the `potential` closure of System, src/system.jl:129-131),
// kind fixed at compile time (the tile kernel is instantiated per kind).
template <typename T, int DIM, int POT> __device__ __forceinline__ T potVk(const T *r, const Prm<T> &p)
{
    if constexpr (POT == kPotHarmonic) {
        T rr = r[0] * r[0];
        if (DIM > 1) rr += r[1] * r[1];
        return p.halfw2 * rr;
    } else if constexpr (POT == kPotLattice) {
        T c0 = cpi(r[0] * p.inva);
        T v = c0 * c0;
        if (DIM > 1) {
            T c1 = cpi(r[1] * p.inva);
            v += c1 * c1;
        }
        return p.V0 * v;
    } else if constexpr (POT == kPotOptical && DIM == 2) {
        T S, Cc;
        optical_sums<T, false>(r, p, S, Cc, nullptr, nullptr);
        S /= p.nbeamT;
        Cc /= p.nbeamT;
        return p.V0 * (S * S + Cc * Cc);
    } else {
        return T(0);
    }
}

// r . grad V (the virial term of src/measurement.jl:107)
template <typename T, int DIM, int POT> __device__ __forceinline__ T rgradVk(const T *r, const Prm<T> &p)
{
    if constexpr (POT == kPotHarmonic) {
        T rr = r[0] * r[0];
        if (DIM > 1) rr += r[1] * r[1];
        return p.w2 * rr;
    } else if constexpr (POT == kPotLattice) {
        // d/dx V0 cos^2(pi x / a) = -V0 (pi/a) sin(2 pi x / a)
        const T pia = T(3.14159265358979323846) * p.inva;
        T s = r[0] * spi(T(2) * r[0] * p.inva);
        if (DIM > 1) s += r[1] * spi(T(2) * r[1] * p.inva);
        return -p.V0 * pia * s;
    } else if constexpr (POT == kPotOptical && DIM == 2) {
        // dV/dx_d = 2 V0 / n^2 (S dS/dx_d + C dC/dx_d), the k_b carrying the factor 2 pi
        T S, Cc, dS[2], dC[2];
        optical_sums<T, true>(r, p, S, Cc, dS, dC);
        const T f = T(2) * T(6.283185307179586476925286766559) * p.V0 / (p.nbeamT * p.nbeamT);
        return f * (r[0] * (S * dS[0] + Cc * dC[0]) + r[1] * (S * dS[1] + Cc * dC[1]));
    } else {
        return T(0);
    }
}

// Runtime-dispatched forms for kernels whose workgroup owns one chain (the
// kind is then uniform and the branch is scalar).
template <typename T, int DIM> __device__ __forceinline__ T potV(const T *r, const Prm<T> &p)
{
    if (p.pot == kPotHarmonic) return potVk<T, DIM, kPotHarmonic>(r, p);
    if (p.pot == kPotLattice) return potVk<T, DIM, kPotLattice>(r, p);
    if (p.pot == kPotOptical) return potVk<T, DIM, kPotOptical>(r, p);
    return T(0);
}

template <typename T, int DIM> __device__ __forceinline__ T rgradV(const T *r, const Prm<T> &p)
{
    if (p.pot == kPotHarmonic) return rgradVk<T, DIM, kPotHarmonic>(r, p);
    if (p.pot == kPotLattice) return rgradVk<T, DIM, kPotLattice>(r, p);
    if (p.pot == kPotOptical) return rgradVk<T, DIM, kPotOptical>(r, p);
    return T(0);
}

template <typename T, int DIM> __device__ __forceinline__ void gradV(const T *r, const Prm<T> &p, T *g)
{
    g[0] = T(0);
    if (DIM > 1) g[1] = T(0);
    if (p.pot == kPotHarmonic) {
        g[0] = p.w2 * r[0];
        if (DIM > 1) g[1] = p.w2 * r[1];
    } else if (p.pot == kPotLattice) {
        const T pia = T(3.14159265358979323846) * p.inva;
        g[0] = -p.V0 * pia * spi(T(2) * r[0] * p.inva);
        if (DIM > 1) g[1] = -p.V0 * pia * spi(T(2) * r[1] * p.inva);
    } else if (p.pot == kPotOptical && DIM == 2) {
        T S, Cc, dS[2], dC[2];
        optical_sums<T, true>(r, p, S, Cc, dS, dC);
        const T f = T(2) * T(6.283185307179586476925286766559) * p.V0 / (p.nbeamT * p.nbeamT);
        g[0] = f * (S * dS[0] + Cc * dC[0]);
        if (DIM > 1) g[DIM - 1] = f * (S * dS[1] + Cc * dC[1]);
    }
}

// dV/d omega of the harmonic trap: omega |r|^2
template <typename T, int DIM> __device__ __forceinline__ T dVdomega(const T *r, double omega)
{
    T rr = r[0] * r[0];
    if (DIM > 1) rr += r[1] * r[1];
    return (T)omega * rr;
}

// Gaussian pair potential v as a function of r^2.
template <typename T> __device__ __forceinline__ T vpair_r2(T r2, const Prm<T> &p)
{
    return p.g * ex(-r2 * p.inv2s2);
}

// Bilinear interpolation of the tabulated integral terms at (x, y) (the
// scale(interpolate(A, BSpline(Linear())), r, r) of src/propagator.jl:67-68),
// arguments clamped to the grid.  tab[i * n + k] = terms(r_i, r_k).
template <typename T> __device__ inline double table_eval(const Prm<T> &p, double x, double y)
{
    const int n = p.tab_n;
    double px = (x - p.tab_r0) / p.tab_h, py = (y - p.tab_r0) / p.tab_h;
    px = fmin(fmax(px, 0.0), (double)(n - 1));
    py = fmin(fmax(py, 0.0), (double)(n - 1));
    const int ix = min((int)floor(px), n - 2), iy = min((int)floor(py), n - 2);
    const double fx = px - ix, fy = py - iy;
    const double *t0 = p.tab + (size_t)ix * n + iy, *t1 = t0 + n;
    const double a0 = (1.0 - fy) * t0[0] + fy * t0[1];
    const double a1 = (1.0 - fy) * t1[0] + fy * t1[1];
    return (1.0 - fx) * a0 + fx * a1;
}

// lnU of the tabulated pair-product propagator (src/system.jl:26-28 with
// build_prop_int, src/propagator.jl:72-89): propint = 1 + terms(|r1|, |r2|) /
// prop_rel0(r1, r2, tau), prop_rel0 = exp(-|r1 - r2|^2 / (4 tau)) / (4 pi tau),
// lnU = propint < 0 ? -mu : log(propint).  c1, c2 are the per-component
// |minimum-image| offsets of the pair at slices j and j+1 (the reference's
// r1_rel, r2_rel, so r1 - r2 is taken between absolute values).
template <typename T, int DIM> __device__ inline T lnu_table(const T *c1, const T *c2, const Prm<T> &p)
{
    double n1 = 0.0, n2 = 0.0, dd = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const double a = (double)c1[d], b = (double)c2[d];
        n1 += a * a;
        n2 += b * b;
        dd += (a - b) * (a - b);
    }
    const double p0 = exp(-dd / (4.0 * p.tau64)) / (4.0 * 3.14159265358979323846 * p.tau64);
    const double pr = 1.0 + table_eval(p, sqrt(n1), sqrt(n2)) / p0;
    return (T)(pr < 0.0 ? -p.mu : log(pr));
}

// lnU of one pair over one link, given the pair's offsets at both slices:
// rk, rkn are the neighbour's beads at j and j+1, q, qn the particle's, r2 and
// s2 the squared minimum-image distances at j and j+1.  Gaussian pair: the
// primitive form -tau/2 (v(r_j) + v(r_j+1)).  TAB: 1 tabulated propagator,
// 0 Gaussian, -1 chosen at run time from p.pair.
template <typename T, int DIM, int TAB = -1>
__device__ __forceinline__ T pair_lnU(const T *rk, const T *q, const T *rkn, const T *qn, T r2, T s2, const Prm<T> &p)
{
    if (TAB == 1 || (TAB < 0 && p.pair == kPairTable)) {
        T c1[DIM], c2[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            c1[d] = mindist(rk[d], q[d], p);
            c2[d] = mindist(rkn[d], qn[d], p);
        }
        return lnu_table<T, DIM>(c1, c2, p);
    }
    return (T(-0.5) * p.tau) * (vpair_r2(r2, p) + vpair_r2(s2, p));
}

// src/updates/helper.jl:3-5
template <typename T> __device__ __forceinline__ bool metropolis(T delta, T u)
{
    return (delta >= T(1)) || (delta > u);
}

// End-bead shift of levy! (src/updates/helper.jl:120-125): if the ends are
// more than L apart in a coordinate, move the far end by sign(start) 2L.
template <typename T> __device__ __forceinline__ T shift_end(T e0, T em, const Prm<T> &p)
{
    if (fabs(e0 - em) > p.L) {
        const T sg = e0 > T(0) ? T(1) : (e0 < T(0) ? T(-1) : T(0));
        em += sg * p.twoL;
    }
    return em;
}

// One step of the Levy bridge (src/updates/helper.jl:129-135): interior bead
// i of a segment of m links, alpha = (m - i) / (m - i + 1), from the
// (untelported) previous bead `prev` towards the end bead `em`.
template <typename T> __device__ __forceinline__ T bridge_alpha(int m, int i) { return T(m - i) / T(m - i + 1); }
// FP32: the hardware reciprocal (1 ulp) instead of the correctly rounded divide.
template <> __device__ __forceinline__ float bridge_alpha<float>(int m, int i)
{
    return float(m - i) * __builtin_amdgcn_rcpf(float(m - i + 1));
}

template <typename T> __device__ __forceinline__ T bridge_sd(T alpha, const Prm<T> &p)
{
    return sq(((T(2) * p.lambda) * alpha) * p.tau);
}

// Compile-time bridge coefficients of a segment of exactly MAXL links:
// alpha_i = (MAXL - i)/(MAXL - i + 1) and sqrt(alpha_i), i = 1..MAXL-1, rounded
// once from double (sd_i = sqrt(2 lambda tau) sqrt(alpha_i)).
constexpr double csqrt(double x)
{
    double r = 1.0;
    for (int k = 0; k < 60; ++k) r = 0.5 * (r + x / r);
    return r;
}
template <int MAXL> struct BridgeConst {
    float alpha[MAXL + 1] = {}, root[MAXL + 1] = {};
    constexpr BridgeConst()
    {
        for (int i = 1; i < MAXL; ++i) {
            const double al = double(MAXL - i) / double(MAXL - i + 1);
            alpha[i] = float(al);
            root[i] = float(csqrt(al));
        }
    }
};
template <int MAXL> constexpr BridgeConst<MAXL> kBridge{};

template <typename T> __device__ __forceinline__ T bridge_step(T prev, T em, T alpha, T xi, T sd)
{
    return (alpha * prev + (T(1) - alpha) * em) + xi * sd;
}

// Bin coordinates of a point (src/nearest_neighbours.jl:26-33), 0-based;
// nbins is the chain's, whose cell side p.cellw = 2L / nbins is rounded in T
// as the quotient would be here.
template <typename T, int DIM>
__device__ __forceinline__ void bin_xy(const T *r, const Prm<T> &p, int nbins, int &bx, int &by)
{
    bx = (int)fl((r[0] + p.L) / p.cellw);
    bx = bx < 0 ? 0 : (bx >= nbins ? nbins - 1 : bx);
    by = 0;
    if (DIM > 1) {
        by = (int)fl((r[1] + p.L) / p.cellw);
        by = by < 0 ? 0 : (by >= nbins ? nbins - 1 : by);
    }
}

// Cell index of a point, x fastest.
template <typename T, int DIM> __device__ __forceinline__ int cell_of(const T *r, const Prm<T> &p, int nbins)
{
    int bx, by;
    bin_xy<T, DIM>(r, p, nbins, bx, by);
    return bx + nbins * by;
}

// x + d wrapped into [0, n) for |d| <= 1 and 0 <= x < n: (x + d + n) % n
// without an integer divide.
__device__ __forceinline__ int wrap_bin(int x, int n) { return x < 0 ? x + n : (x >= n ? x - n : x); }

// Stencil of bin_neighbors (src/nearest_neighbours.jl:55-65), in its order,
// around the cell with bin coordinates (bx, by).
__device__ __forceinline__ int stencil_xy(int bx, int by, int s, int nbins, int dim)
{
    if (dim == 1) {
        const int dx = (s == 0) ? 0 : (s == 1 ? -1 : 1);
        return wrap_bin(bx + dx, nbins);
    }
    constexpr int sx[9] = {0, -1, 0, 1, -1, 1, -1, 0, 1};
    constexpr int sy[9] = {0, 1, 1, 1, 0, 0, -1, -1, -1};
    return wrap_bin(bx + sx[s], nbins) + nbins * wrap_bin(by + sy[s], nbins);
}
__device__ __forceinline__ int stencil_cell(int c, int s, int nbins, int dim)
{
    return stencil_xy(c % nbins, c / nbins, s, nbins, dim);
}

// Accessor for one chain's positions.  Particle-major (SM = false): the
// reference's layout, particle n, coordinate d, slice j at (n*DIM + d)*M + j,
// used by the tile kernel (a ring is contiguous).  Slice-major (SM = true):
// (j*N + n)*DIM + d, used for chains with a pair action, where the beads of
// all particles at one slice are gathered together and one bead is one
// vector load.
template <typename T, int DIM, bool SM = false> struct Path {
    T *base;
    int M, N;
    __device__ __forceinline__ T &at(int n, int d, int j) const
    {
        if constexpr (SM) return base[((long long)j * N + n) * DIM + d];
        else return base[((long long)n * DIM + d) * M + j];
    }
    __device__ __forceinline__ void get(int n, int j, T *r) const
    {
        if constexpr (SM && DIM == 2) {
            using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
            const V2 v = *reinterpret_cast<const V2 *>(&at(n, 0, j));
            r[0] = v.x;
            r[1] = v.y;
        } else {
            r[0] = at(n, 0, j);
            if (DIM > 1) r[1] = at(n, 1, j);
        }
    }
    __device__ __forceinline__ void put(int n, int j, const T *r) const
    {
        at(n, 0, j) = r[0];
        if (DIM > 1) at(n, 1, j) = r[1];
    }
};

// Cell lists of one chain: per slice j and cell c a count and up to `cap`
// particle ids, plus the cell of every bead (the reference's nn grid and
// Particle.bins, src/system.jl:80-91).
struct Cells {
    int *cnt;            // [M][ncell]
    unsigned short *mem; // [M][ncell][cap]
    int *bin;            // [N][M]
    int M, ncell, cap, nbins;
    const int *next = nullptr; // permutation (chains with swap moves): bead M-1 of k links to next[k]'s bead 0
    __device__ __forceinline__ int &count(int j, int c) const { return cnt[(long long)j * ncell + c]; }
    // the particle holding the bead after particle k's bead at slice jn - 1 (jn == 0: across the ring's end)
    __device__ __forceinline__ int after(int k, int jn) const { return (next != nullptr && jn == 0) ? next[k] : k; }
    __device__ __forceinline__ unsigned short *members(int j, int c) const
    {
        return mem + ((long long)j * ncell + c) * cap;
    }
};

// File particle n's bead at slice j, listed in cell oc (-1: in no list, after
// an overflow), under cell nc != oc instead.  One thread per slice may move
// beads at once: the lists of a slice are touched by its thread only.
__device__ __forceinline__ bool cells_refile(const Cells &C, int n, int j, int oc, int nc)
{
    int &b = C.bin[(long long)n * C.M + j];
    int &ncnt = C.count(j, nc);
    const int nc0 = ncnt; // loaded with the old cell's list: nc != oc
    if (oc >= 0) {
        // find n in the old cell's list, read in 16-byte chunks (cap is a multiple
        // of 8 and every list starts 16-byte aligned) issued before any is tested
        unsigned short *om = C.members(j, oc);
        int &ocnt = C.count(j, oc);
        const int cnt = ocnt;
        int at = -1;
        for (int q0 = 0; q0 < C.cap && at < 0; q0 += 32) {
            uint4 w[4];
#pragma unroll
            for (int h = 0; h < 4; ++h)
                w[h] = (q0 + 8 * h < C.cap) ? reinterpret_cast<const uint4 *>(om + q0)[h] : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const unsigned ww[4] = {w[h].x, w[h].y, w[h].z, w[h].w};
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const int qi = q0 + 8 * h + t;
                    const unsigned v = (ww[t >> 1] >> (16 * (t & 1))) & 0xffffu;
                    if (at < 0 && qi < cnt && v == (unsigned)n) at = qi;
                }
            }
        }
        if (at >= 0) {
            om[at] = om[cnt - 1];
            ocnt = cnt - 1;
        }
    }
    if (nc0 >= C.cap) {
        b = -1; // overflow: reported through the status word
        return false;
    }
    C.members(j, nc)[nc0] = (unsigned short)n;
    ncnt = nc0 + 1;
    b = nc;
    return true;
}

// Particles n1 and n2 exchange their beads at slice j (a swap move's tails):
// their ids change places in the lists of slice j.
__device__ __forceinline__ void cells_exchange(const Cells &C, int n1, int n2, int j)
{
    int &b1 = C.bin[(long long)n1 * C.M + j], &b2 = C.bin[(long long)n2 * C.M + j];
    const int c1 = b1, c2 = b2;
    if (c1 != c2) {
        if (c1 >= 0) {
            unsigned short *m = C.members(j, c1);
            for (int i = 0, n = C.count(j, c1); i < n; ++i)
                if (m[i] == (unsigned short)n1) m[i] = (unsigned short)n2;
        }
        if (c2 >= 0) {
            unsigned short *m = C.members(j, c2);
            for (int i = 0, n = C.count(j, c2); i < n; ++i)
                if (m[i] == (unsigned short)n2) m[i] = (unsigned short)n1;
        }
    }
    b1 = c2;
    b2 = c1;
}

// Move particle n's bead at slice j from its stored cell to the cell of r.
template <typename T, int DIM>
__device__ __forceinline__ bool cells_move(const Cells &C, const Prm<T> &p, int n, int j, const T *r)
{
    const int nc = cell_of<T, DIM>(r, p, C.nbins);
    const int oc = C.bin[(long long)n * C.M + j];
    if (nc == oc) return true;
    return cells_refile(C, n, j, oc, nc);
}

// The same move when the bead's old position rold is at hand: its cell is the
// one the bead was filed under, so the stored cell need not be read first.
template <typename T, int DIM>
__device__ __forceinline__ bool cells_move(const Cells &C, const Prm<T> &p, int n, int j, const T *r, const T *rold)
{
    const int nc = cell_of<T, DIM>(r, p, C.nbins);
    const int oc = cell_of<T, DIM>(rold, p, C.nbins);
    if (nc == oc) return true;
    return cells_refile(C, n, j, oc, nc);
}

} // namespace pimc
