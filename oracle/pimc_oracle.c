/*
 * pimc_oracle.c -- CPU FP64 restatement of PIMC.jl's sampling path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark load
 * this library; the product never does.
 *
 * The reference is Julia and cannot run in this environment (no Julia
 * toolchain), so this file restates its algorithm line by line.  Each
 * function cites the reference file:line it follows.  Pinning of each part
 * (closed forms, brute force, known-answer vectors) lives in tests/.
 */
/* j0 and y0 (POSIX Bessel functions of libm) */
#define _DEFAULT_SOURCE
#include "pimc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11.  Multipliers and Weyl constants are the published
 * ones.                                                                    */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * x0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * x2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        x0 = hi1 ^ x1 ^ k0;
        x1 = lo1;
        x2 = hi0 ^ x3 ^ k1;
        x3 = lo0;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

double o_u01(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }

void o_box_muller(uint32_t x0, uint32_t x1, double *n0, double *n1)
{
    const double two_pi = 6.283185307179586476925286766559;
    double u0 = o_u01(x0), u1 = o_u01(x1);
    double rad = sqrt(-2.0 * log(u0));
    *n0 = rad * cos(two_pi * u1);
    *n1 = rad * sin(two_pi * u1);
}

/* draw purposes: the fourth counter word of every keyed draw */
enum { PUR_LEVY = 1, PUR_LEVY_ACC = 2, PUR_COM = 3, PUR_OFFSET = 4, PUR_INIT_START = 5, PUR_INIT_LEVY = 6,
       PUR_SWAP = 7, PUR_SWAP_LEVY = 8, PUR_HS = 9, PUR_HS_COM = 10, PUR_HS_SWAP = 11 };

static void keyed(uint64_t seed, uint32_t c0, uint32_t sweep, uint32_t chain, uint32_t purpose,
                  uint32_t out[4])
{
    uint32_t ctr[4] = {c0, sweep, chain, purpose};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    o_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------------------ */
/* primitives, src/propagator.jl                                            */
/* ------------------------------------------------------------------------ */

/* src/propagator.jl:6-9: per-component absolute minimum-image offset */
double o_distance(double x1, double x2, double L)
{
    double dx = fabs(x1 - x2);
    double a = (2.0 * L) - dx;
    return a < dx ? a : dx;
}

/* src/propagator.jl:30-32 */
double o_teleport(double x, double L)
{
    return ((x + L) - floor(x / (2.0 * L) + 0.5) * (2.0 * L)) - L;
}

/* src/propagator.jl:16-19 (dot of the distance vector, then / (4 lambda tau)) */
double o_lnK(const double *r1, const double *r2, int dim, double tau, double lambda, double L)
{
    double dd = 0.0;
    for (int d = 0; d < dim; ++d) {
        double dr = o_distance(r1[d], r2[d], L);
        dd += dr * dr;
    }
    return -dd / ((4.0 * lambda) * tau);
}

double o_V(const o_pot *p, const double *r, int dim)
{
    switch (p->kind) {
    case O_POT_HARMONIC: {
        double rr = 0.0;
        for (int d = 0; d < dim; ++d) rr += r[d] * r[d];
        return 0.5 * (p->omega * p->omega) * rr;
    }
    case O_POT_LATTICE: {
        const double pi = 3.14159265358979323846264338327950288;
        double v = 0.0;
        for (int d = 0; d < dim; ++d) {
            double c = cos(pi * r[d] / p->a);
            v += c * c;
        }
        return p->V0 * v;
    }
    case O_POT_OPTICAL: {
        /* normalized_intensity (examples/tools/potentialtools.jl:1-16): plane waves of the
         * beams summed, each sum divided by the beam count, squared and added */
        const double pi = 3.14159265358979323846264338327950288;
        double sn = 0.0, cs = 0.0;
        for (int b = 0; b < p->nbeam; ++b) {
            double rb = r[0] * sin(p->angle[b]) + r[1] * cos(p->angle[b]);
            double arg = 2.0 * pi * rb * p->scale + (p->helical ? p->angle[b] : 0.0);
            sn += sin(arg);
            cs += cos(arg);
        }
        sn /= p->nbeam;
        cs /= p->nbeam;
        return p->V0 * (sn * sn + cs * cs);
    }
    default:
        return 0.0;
    }
}

void o_dV(const o_pot *p, const double *r, int dim, double *out)
{
    for (int d = 0; d < dim; ++d) out[d] = 0.0;
    if (p->kind == O_POT_HARMONIC) {
        for (int d = 0; d < dim; ++d) out[d] = (p->omega * p->omega) * r[d];
    } else if (p->kind == O_POT_LATTICE) {
        const double pi = 3.14159265358979323846264338327950288;
        /* d/dx cos^2(pi x/a) = -(pi/a) sin(2 pi x / a) */
        for (int d = 0; d < dim; ++d) out[d] = -p->V0 * (pi / p->a) * sin(2.0 * pi * r[d] / p->a);
    } else if (p->kind == O_POT_OPTICAL) {
        /* V = V0 (S^2 + C^2) / n^2: dV/dx_d = 2 V0 / n^2 (S dS/dx_d + C dC/dx_d) */
        const double pi = 3.14159265358979323846264338327950288;
        double sn = 0.0, cs = 0.0, dsn[2] = {0, 0}, dcs[2] = {0, 0};
        for (int b = 0; b < p->nbeam; ++b) {
            double k[2] = {2.0 * pi * p->scale * sin(p->angle[b]), 2.0 * pi * p->scale * cos(p->angle[b])};
            double arg = k[0] * r[0] + k[1] * r[1] + (p->helical ? p->angle[b] : 0.0);
            double sa = sin(arg), ca = cos(arg);
            sn += sa;
            cs += ca;
            for (int d = 0; d < 2; ++d) {
                dsn[d] += ca * k[d];
                dcs[d] -= sa * k[d];
            }
        }
        const double f = 2.0 * p->V0 / ((double)p->nbeam * p->nbeam);
        for (int d = 0; d < 2 && d < dim; ++d) out[d] = f * (sn * dsn[d] + cs * dcs[d]);
    }
}

/* src/propagator.jl:26-28 */
double o_lnV(const double *r1, const double *r2, int dim, double tau, const o_pot *p)
{
    return (-0.5 * tau) * (o_V(p, r1, dim) + o_V(p, r2, dim));
}

/* ------------------------------------------------------------------------ */
/* pair-product propagator (src/propagator.jl:35-89)                         */
/* ------------------------------------------------------------------------ */
#define GL_N 16
#define GEO_PANELS 40

/* nodes and weights of the n-point Gauss-Legendre rule on [-1, 1] by Newton's
 * iteration on P_n (three-term recurrence), x_i descending */
static void gauss_legendre(int n, double *x, double *w)
{
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < n; ++i) {
        double z = cos(pi * (i + 0.75) / (n + 0.5)), dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int l = 2; l <= n; ++l) {
                double p2 = ((2.0 * l - 1.0) * z * p1 - (l - 1.0) * p0) / l;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = z;
        for (int l = 2; l <= n; ++l) {
            double p2 = ((2.0 * l - 1.0) * z * p1 - (l - 1.0) * p0) / l;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1.0);
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

static int add_panel(double a, double b, const double *gx, const double *gw, double *k, double *w, int cnt, int cap)
{
    for (int i = 0; i < GL_N; ++i, ++cnt) {
        if (cnt < cap) {
            k[cnt] = 0.5 * (b - a) * gx[i] + 0.5 * (b + a);
            w[cnt] = 0.5 * (b - a) * gw[i];
        }
    }
    return cnt;
}

int o_prop_quad_nodes(double rmax, double tau, double *k, double *w, int cap)
{
    double gx[GL_N], gw[GL_N];
    gauss_legendre(GL_N, gx, gw);
    double h = 4.0 / rmax;
    if (h > 0.5) h = 0.5;
    int cnt = 0;
    for (int p = GEO_PANELS - 1; p >= 0; --p) cnt = add_panel(h * ldexp(1.0, -(p + 1)), h * ldexp(1.0, -p), gx, gw, k, w, cnt, cap);
    double K = sqrt(40.0 / tau);
    if (K < 2.0 * h) K = 2.0 * h;
    int np = (int)ceil((K - h) / h);
    double hw = (K - h) / np;
    for (int p = 0; p < np; ++p) cnt = add_panel(h + p * hw, h + (p + 1) * hw, gx, gw, k, w, cnt, cap);
    return cnt;
}

/* the k-dependent factors of the integrand at node k with weight w:
 * a = w k e^{-tau k^2} / (2 pi) * tk^2 / (1 + tk^2), b = ... * tk / (1 + tk^2),
 * tk = 1 / d, d = (2/pi)(eulergamma + log(k/2)) - 4/g0 (src/propagator.jl:45) */
static void prop_coef(double k, double w, double g0, double tau, double *a, double *b)
{
    const double pi = 3.14159265358979323846, eg = 0.57721566490153286061;
    double d = (2.0 / pi) * (eg + log(k / 2.0)) - 4.0 / g0;
    double c = w * k * exp(-tau * k * k) / (2.0 * pi);
    *a = c / (1.0 + d * d);
    *b = c * d / (1.0 + d * d);
}

double o_prop_int_term(double r1, double r2, double g0, double tau, double rmax)
{
    int n = o_prop_quad_nodes(rmax, tau, NULL, NULL, 0);
    double *k = (double *)malloc(sizeof(double) * n), *w = (double *)malloc(sizeof(double) * n);
    o_prop_quad_nodes(rmax, tau, k, w, n);
    double t = 0.0;
    for (int i = 0; i < n; ++i) {
        double a, b;
        prop_coef(k[i], w[i], g0, tau, &a, &b);
        double j1 = j0(k[i] * r1), j2 = j0(k[i] * r2), y1 = y0(k[i] * r1), y2 = y0(k[i] * r2);
        t += -a * (j1 * j2) - b * (j1 * y2 + j2 * y1) + a * (y1 * y2);
    }
    free(k);
    free(w);
    return t;
}

void o_prop_int_terms(int n, double r0, double r1, double g0, double tau, double *out)
{
    int nk = o_prop_quad_nodes(r1, tau, NULL, NULL, 0);
    double *k = (double *)malloc(sizeof(double) * nk), *w = (double *)malloc(sizeof(double) * nk);
    double *a = (double *)malloc(sizeof(double) * nk), *b = (double *)malloc(sizeof(double) * nk);
    double *J = (double *)malloc(sizeof(double) * nk * n), *Y = (double *)malloc(sizeof(double) * nk * n);
    o_prop_quad_nodes(r1, tau, k, w, nk);
    for (int q = 0; q < nk; ++q) prop_coef(k[q], w[q], g0, tau, &a[q], &b[q]);
    for (int i = 0; i < n; ++i) {
        double r = (n > 1) ? r0 + i * ((r1 - r0) / (n - 1)) : r0;
        for (int q = 0; q < nk; ++q) {
            J[(size_t)i * nk + q] = j0(k[q] * r);
            Y[(size_t)i * nk + q] = y0(k[q] * r);
        }
    }
    for (int i = 0; i < n; ++i)
        for (int l = 0; l < n; ++l) {
            const double *Ji = J + (size_t)i * nk, *Yi = Y + (size_t)i * nk;
            const double *Jl = J + (size_t)l * nk, *Yl = Y + (size_t)l * nk;
            double t = 0.0;
            for (int q = 0; q < nk; ++q) t += -a[q] * (Ji[q] * Jl[q]) - b[q] * (Ji[q] * Yl[q] + Jl[q] * Yi[q]) + a[q] * (Yi[q] * Yl[q]);
            out[(size_t)i * n + l] = t;
        }
    free(k); free(w); free(a); free(b); free(J); free(Y);
}

double o_prop_table_eval(const o_pairpot *pp, double x, double y)
{
    const int n = pp->tab_n;
    const double h = (pp->tab_r1 - pp->tab_r0) / (n - 1);
    double px = (x - pp->tab_r0) / h, py = (y - pp->tab_r0) / h;
    if (px < 0.0) px = 0.0;
    if (px > n - 1) px = n - 1;
    if (py < 0.0) py = 0.0;
    if (py > n - 1) py = n - 1;
    int ix = (int)floor(px), iy = (int)floor(py);
    if (ix > n - 2) ix = n - 2;
    if (iy > n - 2) iy = n - 2;
    const double fx = px - ix, fy = py - iy;
    const double *t = pp->tab;
    const double a0 = (1.0 - fy) * t[(size_t)ix * n + iy] + fy * t[(size_t)ix * n + iy + 1];
    const double a1 = (1.0 - fy) * t[(size_t)(ix + 1) * n + iy] + fy * t[(size_t)(ix + 1) * n + iy + 1];
    return (1.0 - fx) * a0 + fx * a1;
}

double o_prop_int(const o_pairpot *pp, const double *r1rel, const double *r2rel, int dim, double tau)
{
    const double pi = 3.14159265358979323846;
    double n1 = 0.0, n2 = 0.0, dd = 0.0;
    for (int d = 0; d < dim; ++d) {
        n1 += r1rel[d] * r1rel[d];
        n2 += r2rel[d] * r2rel[d];
        const double x = r1rel[d] - r2rel[d];
        dd += x * x;
    }
    const double p0 = exp(-dd / (4.0 * tau)) / (4.0 * pi * tau);
    return 1.0 + o_prop_table_eval(pp, sqrt(n1), sqrt(n2)) / p0;
}

double o_vpair(const o_pairpot *pp, double r)
{
    if (pp->kind != O_PAIR_GAUSSIAN) return 0.0;
    return pp->g * exp(-(r * r) / (2.0 * pp->sigma * pp->sigma));
}

/* src/system.jl:26-28, lnU = log(propint(r1_rel, r2_rel, tau)).  The
 * primitive Gaussian pair action used by the interacting workloads enters as
 * propint = exp(-tau/2 (v(|r1_rel|) + v(|r2_rel|))), so its log is taken in
 * closed form.  r1rel, r2rel are the per-component |minimum-image| offsets of
 * the pair at slices j and j+1. */
double o_lnU(const double *r1rel, const double *r2rel, int dim, double tau, const o_pairpot *pp)
{
    if (pp->kind == O_PAIR_TABLE) {
        /* src/system.jl:26-28: propint < 0 ? -mu : log(propint) */
        const double p = o_prop_int(pp, r1rel, r2rel, dim, tau);
        return p < 0.0 ? -pp->mu : log(p);
    }
    double a = 0.0, b = 0.0;
    for (int d = 0; d < dim; ++d) {
        a += r1rel[d] * r1rel[d];
        b += r2rel[d] * r2rel[d];
    }
    return (-0.5 * tau) * (o_vpair(pp, sqrt(a)) + o_vpair(pp, sqrt(b)));
}

/* src/updates/helper.jl:3-5 */
int o_metropolis(double delta, double u) { return (delta >= 1.0) || (delta > u); }

/* src/updates/helper.jl:118-139 */
void o_levy(double *rp, int nrows, int dim, double tau, double L, double lambda, const double *normals)
{
    const int last = nrows - 1;
    for (int d = 0; d < dim; ++d) {
        if (fabs(rp[d * nrows] - rp[d * nrows + last]) > L) {
            double sg = rp[d * nrows] > 0.0 ? 1.0 : (rp[d * nrows] < 0.0 ? -1.0 : 0.0);
            rp[d * nrows + last] += sg * (2.0 * L);
        }
    }
    const int m = nrows - 2;
    for (int j = 1; j <= m; ++j) {
        double alpha = (double)(m + 1 - j) / (double)(m + 2 - j);
        double sd = sqrt(((2.0 * lambda) * alpha) * tau);
        for (int d = 0; d < dim; ++d) {
            double a = alpha * rp[d * nrows + (j - 1)];
            double b = (1.0 - alpha) * rp[d * nrows + last];
            double c = normals[(j - 1) * dim + d] * sd;
            rp[d * nrows + j] = (a + b) + c;
        }
    }
    for (int j = 0; j < nrows; ++j)
        for (int d = 0; d < dim; ++d) rp[d * nrows + j] = o_teleport(rp[d * nrows + j], L);
}

/* ------------------------------------------------------------------------ */
/* system                                                                   */
/* ------------------------------------------------------------------------ */
static inline double R(const o_sys *s, int n, int d, int j) { return s->r[((size_t)n * s->dim + d) * s->M + j]; }
static inline void bead(const o_sys *s, int n, int j, double *out)
{
    for (int d = 0; d < s->dim; ++d) out[d] = R(s, n, d, j);
}
static inline int mod1(int j, int M) { return ((j % M) + M) % M; } /* 0-based wrap */

/* ---- per-slice index of the beads by cell (speeds up find_nns and find_nn) ----
 * For each slice j: the particle indices sorted by (o_bin of their bead,
 * index) and the start of each cell's run, so a query visits its stencil
 * cells' members in index order, the order of a scan over all particles.
 * Bucket ncell holds beads whose o_bin lies outside 0..ncell-1.  Every
 * position write goes through set_r, which moves the bead to the bucket of
 * its new position; o_sys_refresh_V (called after positions are replaced
 * from outside) drops the index, which is then rebuilt on the next query. */
typedef struct {
    int ncell;   /* nbins^dim */
    int *start;  /* M x (ncell + 2) */
    int *ord;    /* M x N */
    int *bucket; /* N x M: bucket of bead (n, j) */
    int valid;
} o_cells;

static int bucket_of(const o_sys *s, int ncell, int n, int j)
{
    double a[3];
    bead(s, n, j, a);
    int b = o_bin(s, a);
    return (b >= 0 && b < ncell) ? b : ncell;
}

static o_cells *cells_get(o_sys *s)
{
    o_cells *c = (o_cells *)s->cells;
    const int N = s->N, M = s->M;
    if (!c) {
        c = (o_cells *)calloc(1, sizeof(o_cells));
        long nc = 1;
        for (int d = 0; d < s->dim; ++d) nc *= s->nbins;
        c->ncell = (int)nc;
        c->start = (int *)malloc(sizeof(int) * (size_t)M * (size_t)(c->ncell + 2));
        c->ord = (int *)malloc(sizeof(int) * (size_t)M * N);
        c->bucket = (int *)malloc(sizeof(int) * (size_t)N * M);
        s->cells = c;
    }
    if (!c->valid) {
        const int nc = c->ncell;
        int *fill = (int *)malloc(sizeof(int) * (size_t)(nc + 1));
        for (int j = 0; j < M; ++j) {
            int *st = c->start + (size_t)j * (nc + 2), *ord = c->ord + (size_t)j * N;
            memset(st, 0, sizeof(int) * (size_t)(nc + 2));
            for (int n = 0; n < N; ++n) {
                const int b = bucket_of(s, nc, n, j);
                c->bucket[(size_t)n * M + j] = b;
                st[b + 1]++;
            }
            for (int b = 0; b <= nc; ++b) st[b + 1] += st[b];
            memcpy(fill, st, sizeof(int) * (size_t)(nc + 1));
            for (int n = 0; n < N; ++n) ord[fill[c->bucket[(size_t)n * M + j]]++] = n;
        }
        free(fill);
        c->valid = 1;
    }
    return c;
}

static void cells_free(o_sys *s)
{
    o_cells *c = (o_cells *)s->cells;
    if (!c) return;
    free(c->start); free(c->ord); free(c->bucket); free(c);
    s->cells = NULL;
}

/* bead (n, j) was written: move it to the bucket of its new position */
static void cells_update(o_sys *s, int n, int j)
{
    o_cells *c = (o_cells *)s->cells;
    if (!c || !c->valid) return;
    const int nc = c->ncell, N = s->N, M = s->M;
    const int b = bucket_of(s, nc, n, j), a = c->bucket[(size_t)n * M + j];
    if (a == b) return;
    int *st = c->start + (size_t)j * (nc + 2), *ord = c->ord + (size_t)j * N;
    int p = st[a];
    while (ord[p] != n) ++p;
    memmove(ord + p, ord + p + 1, sizeof(int) * (size_t)(N - 1 - p));
    for (int k = a + 1; k <= nc + 1; ++k) st[k]--;
    int q = st[b];
    while (q < st[b + 1] && ord[q] < n) ++q;
    memmove(ord + q + 1, ord + q, sizeof(int) * (size_t)(N - 1 - q));
    ord[q] = n;
    for (int k = b + 1; k <= nc + 1; ++k) st[k]++;
    c->bucket[(size_t)n * M + j] = b;
}

/* the one way positions change inside the oracle */
static inline void set_r(o_sys *s, int n, int d, int j, double v)
{
    s->r[((size_t)n * s->dim + d) * s->M + j] = v;
    cells_update(s, n, j);
}

void o_sys_refresh_V(o_sys *s)
{
    if (s->cells) ((o_cells *)s->cells)->valid = 0;
    double a[3], b[3];
    for (int n = 0; n < s->N; ++n)
        for (int j = 0; j < s->M; ++j) {
            bead(s, n, j, a);
            bead(s, n, (j + 1) % s->M, b);
            s->V[(size_t)n * s->M + j] = o_lnV(a, b, s->dim, s->tau, &s->pot);
        }
}

int o_sys_init(o_sys *s, int dim, int M, int N, double lambda, double L, double beta,
               o_pot pot, o_pairpot pair, double r_a, int pair_in_accept, const double *r)
{
    memset(s, 0, sizeof(*s));
    if (dim < 1 || dim > 2 || M < 3 || N < 1 || !(L > 0) || !(beta > 0) || !(lambda > 0)) return -1;
    s->dim = dim; s->M = M; s->N = N;
    s->lambda = lambda; s->L = L; s->beta = beta; s->tau = beta / M;
    s->pot = pot; s->pair = pair; s->pair_in_accept = pair_in_accept;
    /* src/system.jl:22-24: without interactions r_a defaults to L/4 */
    if (r_a <= 0.0) r_a = L / 4.0;
    s->nbins = (int)floor((2.0 * L) / r_a);          /* src/system.jl:81 */
    if (s->nbins < 1) s->nbins = 1;
    s->h = (2.0 * L) / s->nbins;                       /* src/nearest_neighbours.jl:127 */
    s->r = (double *)malloc(sizeof(double) * (size_t)N * dim * M);
    s->V = (double *)malloc(sizeof(double) * (size_t)N * M);
    s->next = (int64_t *)malloc(sizeof(int64_t) * (size_t)N);
    if (!s->r || !s->V || !s->next) return -2;
    memcpy(s->r, r, sizeof(double) * (size_t)N * dim * M);
    for (int n = 0; n < N; ++n) s->next[n] = n;
    s->hs_a = 0.0;
    s->hs_ctr = 10000;
    o_sys_refresh_V(s);
    return 0;
}

void o_sys_free(o_sys *s)
{
    cells_free(s);
    free(s->r); free(s->V); free(s->next);
    s->r = NULL; s->V = NULL; s->next = NULL;
}

/* src/nearest_neighbours.jl:8-33 (bin_index, bin), returned 0-based */
int o_bin(const o_sys *s, const double *r)
{
    double w = (2.0 * s->L) / s->nbins;
    int b = 0, stride = 1;
    for (int d = 0; d < s->dim; ++d) {
        long ib = (long)floor((r[d] + s->L) / w);
        b += (int)ib * stride;
        stride *= s->nbins;
    }
    return b;
}

/* periodic Euclidean distance of the BallTree metric (src/nearest_neighbours.jl:126-127) */
static double periodic_norm(const o_sys *s, const double *a, const double *b)
{
    double q = 0.0;
    for (int d = 0; d < s->dim; ++d) {
        double dd = o_distance(a[d], b[d], s->L);
        q += dd * dd;
    }
    return sqrt(q);
}

/* src/nearest_neighbours.jl:55-65 (bin_neighbors) + :72-104 (within_neigbourhood)
 * + :106-154 (find_nns).  Candidate order: stencil order, then particle index
 * (the reference's per-cell order depends on its update history; only the
 * summation order of the pair terms depends on it). */
int o_find_nn(const o_sys *s, const double *target, int j, int exclude)
{
    const int nb = s->nbins, dim = s->dim;
    int b = o_bin(s, target);
    int bx = b % nb, by = (dim >= 2) ? (b / nb) % nb : 0;
    int best = -1;
    double bd = 0.0, rk[3];
    const o_cells *c = cells_get((o_sys *)s); /* an index of the positions, not state */
    const int *st = c->start + (size_t)j * (c->ncell + 2), *ord = c->ord + (size_t)j * s->N;
    /* the nearest of the stencil candidates, the lowest index among equal distances (the
     * result of a scan over all particles in index order keeping the first minimum) */
    for (int cc = 0; cc < ((dim == 1) ? 3 : 9); ++cc) {
        const int cx = mod1(bx + (dim == 1 ? cc - 1 : cc % 3 - 1), nb);
        const int cy = (dim == 1) ? 0 : mod1(by + cc / 3 - 1, nb);
        const int cell = cx + nb * cy;
        for (int p = st[cell]; p < st[cell + 1]; ++p) {
            const int k = ord[p];
            if (k == exclude) continue;
            bead(s, k, j, rk);
            double d = periodic_norm(s, rk, target);
            if (best < 0 || d < bd || (d == bd && k < best)) { best = k; bd = d; }
        }
    }
    /* beads whose o_bin lies outside the grid take the cell their index names */
    for (int p = st[c->ncell]; p < st[c->ncell + 1]; ++p) {
        const int k = ord[p];
        if (k == exclude) continue;
        bead(s, k, j, rk);
        int ck = o_bin(s, rk), kx = ck % nb, ky = (dim >= 2) ? (ck / nb) % nb : 0;
        int dx = mod1(kx - bx, nb), dy = mod1(ky - by, nb);
        int inx = dx == 0 || dx == 1 || dx == nb - 1, iny = (dim < 2) || dy == 0 || dy == 1 || dy == nb - 1;
        if (!(inx && iny)) continue;
        double d = periodic_norm(s, rk, target);
        if (best < 0 || d < bd || (d == bd && k < best)) { best = k; bd = d; }
    }
    return best;
}

int o_find_nns(const o_sys *s, const double *target, int j, int exclude, int *out, int cap)
{
    const int nb = s->nbins, dim = s->dim;
    static const int st2[9][2] = {{0, 0}, {-1, 1}, {0, 1}, {1, 1}, {-1, 0}, {1, 0}, {-1, -1}, {0, -1}, {1, -1}};
    static const int st1[3] = {0, -1, 1};
    int b = o_bin(s, target);
    int bx = b % nb, by = (dim >= 2) ? (b / nb) % nb : 0;
    int ncells = (dim == 1) ? 3 : 9;
    int cnt = 0;
    double rk[3];
    const o_cells *c = cells_get((o_sys *)s); /* an index of the positions, not state */
    const int *st = c->start + (size_t)j * (c->ncell + 2), *ord = c->ord + (size_t)j * s->N;
    for (int cc = 0; cc < ncells; ++cc) {
        int cell;
        if (dim == 1) {
            cell = mod1(bx + st1[cc], nb);
        } else {
            int cx = mod1(bx + st2[cc][0], nb), cy = mod1(by + st2[cc][1], nb);
            cell = cx + nb * cy;
            /* 3D is not part of the reference's stencil; dim 3 is rejected at init */
        }
        /* the members of the cell in index order */
        for (int p = st[cell]; p < st[cell + 1]; ++p) {
            const int k = ord[p];
            if (k == exclude) continue;
            bead(s, k, j, rk);
            if (periodic_norm(s, rk, target) <= s->h) {
                if (cnt < cap) out[cnt] = k;
                ++cnt;
            }
        }
    }
    return cnt;
}

/* next(): src/updates/helper.jl:287-300, 0-based (slice M-1 -> next particle, slice 0) */
static void next_bead(const o_sys *s, int n, int j, int *nn, int *jn)
{
    *nn = (j == s->M - 1) ? (int)s->next[n] : n;
    *jn = (j + 1) % s->M;
}

#define NNCAP 4096

/* src/updates/helper.jl:306-324 for a 1-cycle (fpcycle(j) == n) */
double o_interaction_old(const o_sys *s, int n, int j)
{
    int list[NNCAP];
    double rn[3], rk[3], rkn[3], rnn[3], r1[3], r2[3];
    bead(s, n, j, rn);
    int cnt = o_find_nns(s, rn, j, n, list, NNCAP);
    if (cnt > NNCAP) cnt = NNCAP;
    double w = 0.0;
    for (int q = 0; q < cnt; ++q) {
        int k = list[q], kn, kj, pn, pj;
        next_bead(s, k, j, &kn, &kj);
        next_bead(s, n, j, &pn, &pj);
        bead(s, k, j, rk);
        bead(s, kn, kj, rkn);
        bead(s, pn, pj, rnn);
        for (int d = 0; d < s->dim; ++d) {
            r1[d] = o_distance(rk[d], rn[d], s->L);
            r2[d] = o_distance(rkn[d], rnn[d], s->L);
        }
        w += o_lnU(r1, r2, s->dim, s->tau, &s->pair);
    }
    return w;
}

/* src/updates/helper.jl:326-347: rows jp, jp+1 of r' (0-based row jp), slice
 * j0 + jp, the moving particle n excluded */
double o_interaction_new(const o_sys *s, const double *rp, int nrows, int jp, int j0, int n)
{
    int list[NNCAP];
    double t[3], t1[3], rk[3], rkn[3], r1[3], r2[3];
    int j = mod1(j0 + jp, s->M);
    for (int d = 0; d < s->dim; ++d) {
        t[d] = rp[d * nrows + jp];
        t1[d] = rp[d * nrows + jp + 1];
    }
    int cnt = o_find_nns(s, t, j, n, list, NNCAP);
    if (cnt > NNCAP) cnt = NNCAP;
    double w = 0.0;
    for (int q = 0; q < cnt; ++q) {
        int k = list[q], kn, kj;
        next_bead(s, k, j, &kn, &kj);
        if (kn == n) continue; /* nn_nextbead[1] in exceptions */
        bead(s, k, j, rk);
        bead(s, kn, kj, rkn);
        for (int d = 0; d < s->dim; ++d) {
            r1[d] = o_distance(rk[d], t[d], s->L);
            r2[d] = o_distance(rkn[d], t1[d], s->L);
        }
        w += o_lnU(r1, r2, s->dim, s->tau, &s->pair);
    }
    return w;
}

/* src/updates/helper.jl:349-366 for a 1-cycle: rp_all is M x dim (the
 * displaced worldline), link j -> j+1 (mnext), neighbour set at slice j */
double o_interaction_com(const o_sys *s, const double *rp_all, int j, int n)
{
    int list[NNCAP];
    double t[3], t1[3], rk[3], rkn[3], r1[3], r2[3];
    const int M = s->M;
    int jn = (j + 1) % M;
    for (int d = 0; d < s->dim; ++d) {
        t[d] = rp_all[d * M + j];
        t1[d] = rp_all[d * M + jn];
    }
    int cnt = o_find_nns(s, t, j, n, list, NNCAP);
    if (cnt > NNCAP) cnt = NNCAP;
    double w = 0.0;
    for (int q = 0; q < cnt; ++q) {
        int k = list[q], kn, kj;
        next_bead(s, k, j, &kn, &kj);
        bead(s, k, j, rk);
        bead(s, kn, kj, rkn);
        for (int d = 0; d < s->dim; ++d) {
            r1[d] = o_distance(rk[d], t[d], s->L);
            r2[d] = o_distance(rkn[d], t1[d], s->L);
        }
        w += o_lnU(r1, r2, s->dim, s->tau, &s->pair);
    }
    return w;
}

/* ------------------------------------------------------------------------ */
/* ReshapeLinear, src/updates/reshape.jl:31-91                              */
/* ------------------------------------------------------------------------ */
int o_reshape_propose(const o_sys *s, int n, int j0, int m, const double *normals, double *rp, o_weights *w)
{
    const int M = s->M, dim = s->dim, nrows = m + 1;
    if (m < 2 || m > M) return 0;
    /* :56-58, 1-cycle: fpcycle(j) == n */
    int jm = j0 + m;
    for (int d = 0; d < dim; ++d) {
        rp[d * nrows + 0] = R(s, n, d, j0);
        rp[d * nrows + m] = R(s, n, d, mod1(jm, M));
    }
    /* :60 hardspherelevy! with a = 0 never re-draws: it equals levy! */
    o_levy(rp, nrows, dim, s->tau, s->L, s->lambda, normals);
    /* :66-77 */
    double wVi = 0.0, wUi = 0.0, wUu = 0.0;
    double Vp[4096 + 1];
    double a[3], b[3];
    for (int j = j0; j < jm; ++j) {
        int jp = j - j0;
        wVi += s->V[(size_t)n * M + mod1(j, M)];
        if (s->pair.kind != O_PAIR_NONE) wUi += o_interaction_old(s, n, mod1(j, M));
        for (int d = 0; d < dim; ++d) {
            a[d] = rp[d * nrows + jp];
            b[d] = rp[d * nrows + jp + 1];
        }
        Vp[jp] = o_lnV(a, b, dim, s->tau, &s->pot);
        if (s->pair.kind != O_PAIR_NONE) wUu += o_interaction_new(s, rp, nrows, jp, j0, n);
    }
    double wVu = 0.0;
    for (int jp = 0; jp < m; ++jp) wVu += Vp[jp];
    w->wV_initial = wVi; w->wV_updated = wVu;
    w->wU_initial = wUi; w->wU_updated = wUu;
    return 1;
}

int o_reshape_apply(o_sys *s, int n, int j0, int m, const double *normals, double u, int *accepted, o_weights *w)
{
    const int M = s->M, dim = s->dim, nrows = m + 1;
    double *rp = (double *)malloc(sizeof(double) * nrows * dim);
    o_weights ww;
    if (!o_reshape_propose(s, n, j0, m, normals, rp, &ww)) { free(rp); *accepted = 0; return 0; }
    if (w) *w = ww;
    double lw = ww.wV_updated - ww.wV_initial;
    if (s->pair_in_accept) lw += ww.wU_updated - ww.wU_initial;
    *accepted = o_metropolis(exp(lw), u);
    if (*accepted) {
        /* :82-86 rows 0..m-1 are written back (the start bead as teleported) */
        double a[3], b[3];
        for (int jp = 0; jp < m; ++jp) {
            int j = mod1(j0 + jp, M);
            for (int d = 0; d < dim; ++d) {
                set_r(s, n, d, j, rp[d * nrows + jp]);
                a[d] = rp[d * nrows + jp];
                b[d] = rp[d * nrows + jp + 1];
            }
            s->V[(size_t)n * M + j] = o_lnV(a, b, dim, s->tau, &s->pot);
        }
    }
    free(rp);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* SingleCenterOfMass, src/updates/com.jl:136-224                           */
/* ------------------------------------------------------------------------ */
int o_com_propose(const o_sys *s, int n, const double *dsp, double *rp, o_weights *w)
{
    const int M = s->M, dim = s->dim;
    /* :168-175 */
    double wVi = 0.0, wUi = 0.0;
    for (int j = 0; j < M; ++j) wVi += s->V[(size_t)n * M + j];
    if (s->pair.kind != O_PAIR_NONE)
        for (int j = 0; j < M; ++j) wUi += o_interaction_old(s, n, j);
    /* move_polymer!, src/updates/helper.jl:368-395 (a = 0: no re-draw) */
    for (int j = 0; j < M; ++j)
        for (int d = 0; d < dim; ++d) rp[d * M + j] = o_teleport(R(s, n, d, j) + dsp[d], s->L);
    /* :189-201 */
    double wVu = 0.0, wUu = 0.0, a[3], b[3];
    for (int j = 0; j < M; ++j) {
        int jn = (j + 1) % M;
        for (int d = 0; d < dim; ++d) { a[d] = rp[d * M + j]; b[d] = rp[d * M + jn]; }
        wVu += o_lnV(a, b, dim, s->tau, &s->pot);
        if (s->pair.kind != O_PAIR_NONE) wUu += o_interaction_com(s, rp, j, n);
    }
    w->wV_initial = wVi; w->wV_updated = wVu;
    w->wU_initial = wUi; w->wU_updated = wUu;
    return 1;
}

int o_com_apply(o_sys *s, int n, const double *dsp, double u, int *accepted, o_weights *w)
{
    const int M = s->M, dim = s->dim;
    double *rp = (double *)malloc(sizeof(double) * M * dim);
    o_weights ww;
    o_com_propose(s, n, dsp, rp, &ww);
    if (w) *w = ww;
    double lw = ww.wV_updated - ww.wV_initial;
    if (s->pair_in_accept) lw += ww.wU_updated - ww.wU_initial;
    *accepted = o_metropolis(exp(lw), u);
    if (*accepted) {
        for (int j = 0; j < M; ++j)
            for (int d = 0; d < dim; ++d) set_r(s, n, d, j, rp[d * M + j]);
        o_sys_refresh_V(s); /* :211-219 store V' per link: same values */
    }
    free(rp);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* estimators, src/measurement.jl                                           */
/* ------------------------------------------------------------------------ */
/* :92-122 */
void o_energy(const o_sys *s, double *E, double *Ev)
{
    const int M = s->M, dim = s->dim;
    double link = 0.0, pot = 0.0, vkin = 0.0, vpot = 0.0;
    double a[3], b[3], g[3];
    for (int i = 0; i < s->N; ++i)
        for (int j = 0; j < M; ++j) {
            int inext = (j == M - 1) ? (int)s->next[i] : i;
            int jnext = (j + 1) % M;
            bead(s, i, j, a);
            bead(s, inext, jnext, b);
            double dd = 0.0;
            for (int d = 0; d < dim; ++d) {
                double dr = o_distance(a[d], b[d], s->L);
                dd += dr * dr;
            }
            link += dd;
            pot += o_V(&s->pot, a, dim) + o_V(&s->pot, b, dim);
            o_dV(&s->pot, a, dim, g);
            double rg = 0.0;
            for (int d = 0; d < dim; ++d) rg += a[d] * g[d];
            vkin += rg;
            vpot += o_V(&s->pot, a, dim) + o_V(&s->pot, b, dim);
        }
    double tau = s->tau;
    *E = ((double)(dim * s->N) / (2.0 * tau) - (1.0 / (((4.0 * s->lambda) * (tau * tau)) * M)) * link) +
         (1.0 / (2.0 * M)) * pot;
    *Ev = (1.0 / (2.0 * M)) * vkin + (1.0 / (2.0 * M)) * vpot;
}

/* Pair energy of the link action S_U = -sum lnU (build addition; the
 * reference's Energy has no pair term).  Thermodynamic: <S_U>/beta.  Virial:
 * <S_U>/beta + (1/(2 beta)) sum r . dS_U/dr, the gradient taken with the
 * neighbour indicator held fixed. */
void o_pair_energy(const o_sys *s, double *Epair, double *Evir_pair)
{
    *Epair = 0.0; *Evir_pair = 0.0;
    if (s->pair.kind != O_PAIR_GAUSSIAN) return; /* no pair energy for the tabulated propagator */
    const int M = s->M, dim = s->dim;
    int list[NNCAP];
    double ri[3], rk[3], rin[3], rkn[3];
    double SU = 0.0, rgrad = 0.0;
    const double s2 = s->pair.sigma * s->pair.sigma;
    for (int i = 0; i < s->N; ++i)
        for (int j = 0; j < M; ++j) {
            bead(s, i, j, ri);
            int cnt = o_find_nns(s, ri, j, i, list, NNCAP);
            if (cnt > NNCAP) cnt = NNCAP;
            for (int q = 0; q < cnt; ++q) {
                int k = list[q], in_, ij, kn, kj;
                next_bead(s, i, j, &in_, &ij); /* next beads along the cycles */
                next_bead(s, k, j, &kn, &kj);
                bead(s, k, j, rk);
                bead(s, in_, ij, rin);
                bead(s, kn, kj, rkn);
                double a2 = 0.0, b2 = 0.0;
                double da[3], db[3];
                for (int d = 0; d < dim; ++d) {
                    /* signed minimum-image separations for the gradient */
                    double x = ri[d] - rk[d];
                    x -= (2.0 * s->L) * floor(x / (2.0 * s->L) + 0.5);
                    double y = rin[d] - rkn[d];
                    y -= (2.0 * s->L) * floor(y / (2.0 * s->L) + 0.5);
                    da[d] = x; db[d] = y;
                    a2 += x * x; b2 += y * y;
                }
                double va = o_vpair(&s->pair, sqrt(a2)), vb = o_vpair(&s->pair, sqrt(b2));
                /* each unordered pair is visited from both ends: weight 1/2 */
                SU += 0.5 * (0.5 * s->tau) * (va + vb);
                /* grad of v(|x|) wrt x_i is -x/s2 v; ri.grad_i + rk.grad_k = (ri-rk).grad */
                double ga = 0.0, gb = 0.0;
                for (int d = 0; d < dim; ++d) { ga += da[d] * da[d]; gb += db[d] * db[d]; }
                rgrad += 0.5 * (0.5 * s->tau) * (-(ga / s2) * va - (gb / s2) * vb);
            }
        }
    *Epair = SU / s->beta;
    *Evir_pair = SU / s->beta + rgrad / (2.0 * s->beta);
}

/* :45-55 */
void o_density(const o_sys *s, int nbins, double *dens, int64_t *ndata)
{
    const double binw = (2.0 * s->L) / nbins;
    for (int n = 0; n < s->N; ++n)
        for (int m = 0; m < s->M; ++m) {
            long ib[3];
            int ok = 1;
            for (int d = 0; d < s->dim; ++d) {
                ib[d] = (long)floor((R(s, n, d, m) + s->L) / binw);
                if (!(ib[d] > 0 && ib[d] < nbins + 1)) ok = 0;
            }
            if (ok) {
                /* Julia stores at 1-based index ibin, i.e. 0-based ibin-1;
                 * dens is nbins^dim column-major like the Julia Array */
                size_t idx = 0, stride = 1;
                for (int d = 0; d < s->dim; ++d) { idx += (size_t)(ib[d] - 1) * stride; stride *= (size_t)nbins; }
                dens[idx] += 1.0;
            }
        }
    *ndata += s->M;
}

void o_radial(const o_sys *s, int nb, double rmax, double *hist, int64_t *ndata)
{
    const double w = rmax / nb;
    for (int n = 0; n < s->N; ++n)
        for (int m = 0; m < s->M; ++m) {
            double rr = 0.0;
            for (int d = 0; d < s->dim; ++d) rr += R(s, n, d, m) * R(s, n, d, m);
            double q = sqrt(rr) / w;
            if (q < nb) hist[(int)q] += 1.0;
        }
    *ndata += s->M;
}

/* ------------------------------------------------------------------------ */
/* brute-force full actions                                                  */
/* ------------------------------------------------------------------------ */
double o_total_lnV(const o_sys *s)
{
    double a[3], b[3], w = 0.0;
    for (int n = 0; n < s->N; ++n)
        for (int j = 0; j < s->M; ++j) {
            bead(s, n, j, a);
            bead(s, n, (j + 1) % s->M, b);
            w += o_lnV(a, b, s->dim, s->tau, &s->pot);
        }
    return w;
}

double o_total_lnK(const o_sys *s)
{
    double a[3], b[3], w = 0.0;
    for (int n = 0; n < s->N; ++n)
        for (int j = 0; j < s->M; ++j) {
            bead(s, n, j, a);
            bead(s, n, (j + 1) % s->M, b);
            w += o_lnK(a, b, s->dim, s->tau, s->lambda, s->L);
        }
    return w;
}

/* sum over links of sum over unordered pairs, by O(N^2) brute force with the
 * slice-j radius test (no cell list): pins the cell-list enumeration */
double o_total_lnU(const o_sys *s)
{
    if (s->pair.kind == O_PAIR_NONE) return 0.0;
    double ri[3], rk[3], rin[3], rkn[3], r1[3], r2[3], w = 0.0;
    for (int j = 0; j < s->M; ++j) {
        for (int i = 0; i < s->N; ++i)
            for (int k = i + 1; k < s->N; ++k) {
                bead(s, i, j, ri);
                bead(s, k, j, rk);
                if (periodic_norm(s, ri, rk) > s->h) continue;
                int in_, ij, kn, kj; /* next beads along the cycles */
                next_bead(s, i, j, &in_, &ij);
                next_bead(s, k, j, &kn, &kj);
                bead(s, in_, ij, rin);
                bead(s, kn, kj, rkn);
                for (int d = 0; d < s->dim; ++d) {
                    r1[d] = o_distance(rk[d], ri[d], s->L);
                    r2[d] = o_distance(rkn[d], rin[d], s->L);
                }
                w += o_lnU(r1, r2, s->dim, s->tau, &s->pair);
            }
    }
    return w;
}

/* ------------------------------------------------------------------------ */
/* the reference schedule, src/simulation.jl:29-42                          */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint64_t seed, ctr;
    uint32_t buf[4];
    int have;
} stream_t;

static uint32_t stream_next(stream_t *st)
{
    if (st->have == 0) {
        uint32_t c[4] = {(uint32_t)st->ctr, (uint32_t)(st->ctr >> 32), 0xC0FFEEu, 0u};
        uint32_t k[2] = {(uint32_t)st->seed, (uint32_t)(st->seed >> 32)};
        o_philox4x32_10(c, k, st->buf);
        st->ctr++;
        st->have = 4;
    }
    return st->buf[4 - st->have--];
}
static double stream_u(stream_t *st) { return o_u01(stream_next(st)); }
static int stream_int(stream_t *st, int lo, int hi) /* uniform in lo..hi */
{
    uint64_t span = (uint64_t)(hi - lo + 1);
    return lo + (int)(((uint64_t)stream_next(st) * span) >> 32);
}
static double stream_normal(stream_t *st)
{
    double n0, n1;
    uint32_t a = stream_next(st), b = stream_next(st);
    o_box_muller(a, b, &n0, &n1);
    return n0;
}

/* Counter: sliding window of accept flags (src/updates/helper.jl:6-11) */
typedef struct {
    unsigned char *q;
    int range, head, len, sum;
    int64_t tries;
} window_t;
static void win_init(window_t *w, int range) { w->q = (unsigned char *)calloc(range, 1); w->range = range; w->head = 0; w->len = 0; w->sum = 0; w->tries = 0; }
static void win_free(window_t *w) { free(w->q); }
/* queue! (src/simulation.jl:3-10) */
static void win_push(window_t *w, int acc)
{
    w->tries++;
    if (w->len == w->range) { w->sum -= w->q[w->head]; w->q[w->head] = (unsigned char)acc; w->head = (w->head + 1) % w->range; }
    else { w->q[(w->head + w->len) % w->range] = (unsigned char)acc; w->len++; }
    w->sum += acc;
}
static double win_acceptance(const window_t *w) { return w->len ? (double)w->sum / w->len : 0.0; }

void o_refstate_init(o_refstate *st, const o_sys *s, int cap, double step)
{
    memset(st, 0, sizeof(*st));
    int maxslices = s->M - 2;
    st->cap = cap < maxslices ? cap : maxslices;       /* src/updates/reshape.jl:25 */
    st->step = step;
    st->win_range = 10000;                               /* range = 10_000 */
    st->adj = 10;                                        /* adj = 10 */
}

int o_run_reference(o_sys *s, o_refstate *st, int64_t n_iter, uint64_t seed, int ncycle,
                    double *E, double *Ev, int64_t nE)
{
    stream_t rs = {seed, st->rng_counter, {0, 0, 0, 0}, 0};
    window_t wl, wc;
    win_init(&wl, st->win_range);
    win_init(&wc, st->win_range);
    const int M = s->M, dim = s->dim;
    int maxslices = M - 2, nctr = 0;
    double *normals = (double *)malloc(sizeof(double) * (size_t)(M + 1) * dim);
    for (int64_t it = 0; it < n_iter; ++it) {
        /* sample(fs, Weights(1 ./ every)) with every = (1, 1): COM first */
        int is_com = stream_u(&rs) < 0.5;
        if (is_com) {
            /* SingleCenterOfMass: all particles are 1-cycles */
            int n = stream_int(&rs, 0, s->N - 1);
            double dsp[3];
            for (int d = 0; d < dim; ++d) dsp[d] = (st->step * 2.0) * (stream_u(&rs) - 0.5);
            double *rp = (double *)malloc(sizeof(double) * M * dim);
            o_weights w;
            o_com_propose(s, n, dsp, rp, &w);
            double lw = w.wV_updated - w.wV_initial;
            if (s->pair_in_accept) lw += w.wU_updated - w.wU_initial;
            double delta = exp(lw);
            int acc = (delta >= 1.0) || (delta > stream_u(&rs));
            if (acc) {
                for (int j = 0; j < M; ++j)
                    for (int d = 0; d < dim; ++d) set_r(s, n, d, j, rp[d * M + j]);
                double a[3], b[3];
                for (int j = 0; j < M; ++j) {
                    for (int d = 0; d < dim; ++d) { a[d] = rp[d * M + j]; b[d] = rp[d * M + (j + 1) % M]; }
                    s->V[(size_t)n * M + j] = o_lnV(a, b, dim, s->tau, &s->pot);
                }
            }
            free(rp);
            win_push(&wc, acc);
            st->com_tries++; st->com_acc += acc; st->bead_moves += M;
            /* apply!: adjust every adj tries (src/simulation.jl:19-27, helper.jl:22-32) */
            if (wc.tries % st->adj == 0) {
                double a = win_acceptance(&wc);
                if (a < 0.4) st->step *= 0.9;
                else if (a > 0.6) st->step *= 1.1;
                if (st->step < 0.1) st->step = 0.1;
                if (st->step > s->L / 2) st->step = s->L / 2;
            }
        } else {
            int n = stream_int(&rs, 0, s->N - 1);
            int j0 = stream_int(&rs, 0, M - 1);
            int m = stream_int(&rs, 2, st->cap);
            if (m > maxslices) m = maxslices;
            for (int q = 0; q < (m - 1) * dim; ++q) normals[q] = stream_normal(&rs);
            double *rp = (double *)malloc(sizeof(double) * (m + 1) * dim);
            o_weights w;
            o_reshape_propose(s, n, j0, m, normals, rp, &w);
            double lw = w.wV_updated - w.wV_initial;
            if (s->pair_in_accept) lw += w.wU_updated - w.wU_initial;
            double delta = exp(lw);
            int acc = (delta >= 1.0) || (delta > stream_u(&rs));
            if (acc) {
                double a[3], b[3];
                for (int jp = 0; jp < m; ++jp) {
                    int j = mod1(j0 + jp, M);
                    for (int d = 0; d < dim; ++d) {
                        set_r(s, n, d, j, rp[d * (m + 1) + jp]);
                        a[d] = rp[d * (m + 1) + jp];
                        b[d] = rp[d * (m + 1) + jp + 1];
                    }
                    s->V[(size_t)n * M + j] = o_lnV(a, b, dim, s->tau, &s->pot);
                }
            }
            free(rp);
            win_push(&wl, acc);
            st->levy_tries++; st->levy_acc += acc; st->bead_moves += m - 1;
            if (wl.tries % st->adj == 0) {
                double a = win_acceptance(&wl);
                if (a < 0.6) st->cap -= 1;
                else if (a > 0.8) st->cap += 1;
                if (st->cap < 2) st->cap = 2;
                if (st->cap > maxslices) st->cap = maxslices;
            }
        }
        /* measurement_Z_sector, src/measurement.jl:1-17 */
        if (E) {
            if (++nctr == ncycle) {
                nctr = 0;
                if (st->n_meas < nE) o_energy(s, &E[st->n_meas], &Ev[st->n_meas]);
                st->n_meas++;
            }
        }
    }
    st->rng_counter = rs.ctr;
    free(normals);
    win_free(&wl);
    win_free(&wc);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* the sweep schedule of the HIP kernels                                     */
/* ------------------------------------------------------------------------ */
int64_t o_sweep_bead_moves(int N, int M, int ell)
{
    int64_t per = 0;
    int nseg = (M + ell - 1) / ell;
    for (int k = 0; k < nseg; ++k) {
        int m = (M - k * ell) < ell ? (M - k * ell) : ell;
        if (m >= 2) per += m - 1;
    }
    return (int64_t)N * (per + M);
}

double o_adapt_step(double step, int64_t acc, int64_t tries, double L)
{
    if (tries <= 0) return step;
    double a = (double)acc / (double)tries;
    if (a < 0.4) step *= 0.9;
    else if (a > 0.6) step *= 1.1;
    if (step < 0.1) step = 0.1;
    if (step > L / 2) step = L / 2;
    return step;
}

/* adjust!(::NumbOfSlices) (src/updates/helper.jl:42-52) with ReshapeLinear's
 * bounds minslices = 2, maxslices = M - 2 (src/updates/reshape.jl:12-21), a cap
 * of 32 links, and 2 as the floor when M - 2 < 2 */
int o_adapt_ell(int ell, int64_t acc, int64_t tries, int M)
{
    if (tries <= 0) return ell;
    double a = (double)acc / (double)tries;
    if (a < 0.6) ell -= 1;
    else if (a > 0.8) ell += 1;
    int hi = M - 2 < 32 ? M - 2 : 32;
    if (ell > hi) ell = hi;
    if (ell < 2) ell = 2;
    return ell;
}

/* pair link sum for particle n over link j -> j+1 with its beads at slices j
 * and j+1 replaced by pj, pj1 (neighbour set from slice j) */
static double link_pair(const o_sys *s, int n, int j, const double *pj, const double *pj1)
{
    int list[NNCAP];
    double rk[3], rkn[3], r1[3], r2[3];
    int jn = (j + 1) % s->M;
    int cnt = o_find_nns(s, pj, j, n, list, NNCAP);
    if (cnt > NNCAP) cnt = NNCAP;
    double w = 0.0;
    for (int q = 0; q < cnt; ++q) {
        int k = list[q], kn, kj;
        next_bead(s, k, j, &kn, &kj); /* the neighbour's next bead along its cycle */
        bead(s, k, j, rk);
        bead(s, kn, kj, rkn);
        for (int d = 0; d < s->dim; ++d) {
            r1[d] = o_distance(rk[d], pj[d], s->L);
            r2[d] = o_distance(rkn[d], pj1[d], s->L);
        }
        w += o_lnU(r1, r2, s->dim, s->tau, &s->pair);
    }
    (void)jn;
    return w;
}

/* bead u of the polymer that starts at (p, slice 0): u < M on p, M <= u < 2M on next[p] */
static inline int owner(const o_sys *s, int p, int u) { return u < s->M ? p : (int)s->next[p]; }

static void o_swaps(o_sys *s, uint32_t chain, uint32_t sweep, uint64_t seed, int n_swap, int m, int rule,
                    o_sweepstats *st);

/* whether the bead r of particle n at slice j is closer than a to its nearest
 * neighbour (the test of hardspherelevy! and move_polymer!, src/updates/helper.jl:166-170, 385-389) */
static int hs_near(const o_sys *s, const double *r, int j, int n)
{
    int nn = o_find_nn(s, r, j, n);
    if (nn < 0) return 0;
    double rk[3];
    bead(s, nn, j, rk);
    return periodic_norm(s, rk, r) < s->hs_a;
}

/* hardspherelevy! (src/updates/helper.jl:141-181), the rows of rp holding the
 * segment from slice `start`: the end shift of levy!, then bead by bead the
 * bridge step, redrawn while the teleported bead lies closer than a to its
 * nearest neighbour other than particle `excl` (the reference's exception,
 * fpcycle(mod1(j0 + j, M)): the first particle of the cycle the move names).
 * Draw 1 of interior bead i is normals[i - 1]; draw t >= 2 is Box-Muller of
 * words 0, 1 of keyed(seed, item, sweep, chain, pur | t << 8), item = base +
 * slice (by_slice) or base + i.  Returns 0 when some bead needs more than
 * hs_ctr draws. */
static int hs_levy_x(const o_sys *s, int excl, int start, double *rp, int nrows, const double *normals,
                     uint32_t chain, uint32_t sweep, uint64_t seed, uint32_t base, int by_slice, uint32_t pur);

/* staging: particle p, redraws keyed by p M + slice and PUR_HS */
static int hs_levy(const o_sys *s, int p, int start, double *rp, int nrows, const double *normals, uint32_t chain,
                   uint32_t sweep, uint64_t seed)
{
    return hs_levy_x(s, p, start, rp, nrows, normals, chain, sweep, seed, (uint32_t)(p * s->M), 1, PUR_HS);
}

static int hs_levy_x(const o_sys *s, int excl, int start, double *rp, int nrows, const double *normals,
                     uint32_t chain, uint32_t sweep, uint64_t seed, uint32_t base, int by_slice, uint32_t pur)
{
    const int dim = s->dim, last = nrows - 1, M = s->M;
    const double L = s->L;
    for (int d = 0; d < dim; ++d) {
        if (fabs(rp[d * nrows] - rp[d * nrows + last]) > L) {
            double sg = rp[d * nrows] > 0.0 ? 1.0 : (rp[d * nrows] < 0.0 ? -1.0 : 0.0);
            rp[d * nrows + last] += sg * (2.0 * L);
        }
    }
    const int m = nrows - 2;
    uint32_t x[4];
    for (int j = 1; j <= m; ++j) {
        double alpha = (double)(m + 1 - j) / (double)(m + 2 - j);
        double sd = sqrt(((2.0 * s->lambda) * alpha) * s->tau);
        const int slice = (start + j) % M;
        double nr[3] = {0, 0, 0}, t[3];
        for (int d = 0; d < dim; ++d) nr[d] = normals[(j - 1) * dim + d];
        for (int attempt = 1;; ++attempt) {
            if (attempt > s->hs_ctr) return 0;
            if (attempt > 1) {
                keyed(seed, base + (uint32_t)(by_slice ? slice : j), sweep, chain, pur | ((uint32_t)attempt << 8), x);
                o_box_muller(x[0], x[1], &nr[0], &nr[1]);
            }
            for (int d = 0; d < dim; ++d) {
                double a = alpha * rp[d * nrows + (j - 1)];
                double b = (1.0 - alpha) * rp[d * nrows + last];
                double c = nr[d] * sd;
                rp[d * nrows + j] = (a + b) + c;
                t[d] = o_teleport(rp[d * nrows + j], L);
            }
            if (!hs_near(s, t, slice, excl)) break;
        }
    }
    for (int j = 0; j < nrows; ++j)
        for (int d = 0; d < dim; ++d) rp[d * nrows + j] = o_teleport(rp[d * nrows + j], L);
    return 1;
}

static void sweep_core(o_sys *s, uint32_t chain, uint32_t sweep, uint64_t seed, int ell, double com_step,
                       int n_swap, int swap_m, int rule, const int every[3], o_sweepstats *st)
{
    /* move weights: the staging pass, the COM pass and the swap tries run on the sweeps with
     * sweep % every == 0 (run! draws each update with weight 1/every, src/simulation.jl:33-36) */
    const int st_on = sweep % (uint32_t)every[0] == 0, com_on = sweep % (uint32_t)every[1] == 0;
    const int sw_on = sweep % (uint32_t)every[2] == 0;
    const int M = s->M, dim = s->dim, N = s->N;
    const int nseg = (M + ell - 1) / ell;
    const int use_pair = (s->pair.kind != O_PAIR_NONE) && s->pair_in_accept;
    double *rp = (double *)malloc(sizeof(double) * (size_t)(M + 1) * dim);
    double *nrm = (double *)malloc(sizeof(double) * (size_t)(M + 1) * dim);
    uint32_t x[4];
    /* Levy phase: particles in order; the segments of one particle have
     * disjoint interiors and are independent given their end beads.  One
     * staging offset per chain and sweep: the segments of all particles then
     * share the same slice windows, and moves in different windows touch
     * disjoint beads, so any order of the windows gives this result. */
    keyed(seed, 0u, sweep, chain, PUR_OFFSET, x);
    const int off = (int)(((uint64_t)x[0] * (uint64_t)ell) >> 32);
    for (int p = 0; p < (st_on ? N : 0); ++p) {
        for (int k = 0; k < nseg; ++k) {
            int m = (M - k * ell) < ell ? (M - k * ell) : ell;
            if (m < 2) continue;
            const int us = off + k * ell; /* window position of the start bead (< 2M) */
            int start = us % M;
            const int nrows = m + 1;
            for (int d = 0; d < dim; ++d) {
                rp[d * nrows + 0] = R(s, owner(s, p, us), d, start);
                rp[d * nrows + m] = R(s, owner(s, p, us + m), d, (start + m) % M);
            }
            for (int i = 1; i < m; ++i) {
                int b = (start + i) % M;
                /* one Philox block serves the interior beads (i, i+1), i odd,
                 * keyed by the slice of bead i: words 0,1 for bead i, 2,3 for i+1 */
                if (i & 1) keyed(seed, (uint32_t)(p * M + b), sweep, chain, PUR_LEVY, x);
                double n0, n1;
                if (i & 1) o_box_muller(x[0], x[1], &n0, &n1);
                else o_box_muller(x[2], x[3], &n0, &n1);
                nrm[(i - 1) * dim + 0] = n0;
                if (dim > 1) nrm[(i - 1) * dim + 1] = n1;
            }
            /* the bridge of levy! (hardspherelevy! with a > 0), with the fixed ends kept as stored */
            double e0[3], em[3];
            for (int d = 0; d < dim; ++d) { e0[d] = rp[d * nrows]; em[d] = rp[d * nrows + m]; }
            int bridged = 1;
            if (s->hs_a > 0.0) bridged = hs_levy(s, p, start, rp, nrows, nrm, chain, sweep, seed);
            else o_levy(rp, nrows, dim, s->tau, s->L, s->lambda, nrm);
            for (int d = 0; d < dim; ++d) { rp[d * nrows] = e0[d]; rp[d * nrows + m] = em[d]; }
            if (!bridged) { /* no position outside the hard spheres: a rejected try */
                st->levy_tries++;
                st->bead_moves += m - 1;
                continue;
            }
            /* Delta ln W = -tau sum_interior (V' - V) + Delta lnU */
            double dv = 0.0, a[3], b[3];
            for (int i = 1; i < m; ++i) {
                int j = (start + i) % M;
                for (int d = 0; d < dim; ++d) { a[d] = rp[d * nrows + i]; b[d] = R(s, owner(s, p, us + i), d, j); }
                dv += o_V(&s->pot, a, dim) - o_V(&s->pot, b, dim);
            }
            double lw = -s->tau * dv;
            if (use_pair) {
                double du = 0.0, c0[3], c1[3];
                for (int t = 0; t < m; ++t) {
                    /* link t of the window: bead u = us + t of the polymer (its owner at slice j)
                     * to bead u + 1 (along the cycle) */
                    int j = (start + t) % M, jn = (j + 1) % M;
                    const int q0 = owner(s, p, us + t), q1 = owner(s, p, us + t + 1);
                    for (int d = 0; d < dim; ++d) {
                        a[d] = rp[d * nrows + t]; b[d] = rp[d * nrows + t + 1];
                        c0[d] = R(s, q0, d, j); c1[d] = R(s, q1, d, jn);
                    }
                    du += link_pair(s, q0, j, a, b) - link_pair(s, q0, j, c0, c1);
                }
                lw += du;
            }
            keyed(seed, (uint32_t)(p * M + start), sweep, chain, PUR_LEVY_ACC, x);
            int acc = o_metropolis(exp(lw), o_u01(x[0]));
            st->levy_tries++;
            st->bead_moves += m - 1;
            if (acc) {
                st->levy_acc++;
                for (int i = 1; i < m; ++i) {
                    int j = (start + i) % M, q = owner(s, p, us + i);
                    for (int d = 0; d < dim; ++d) set_r(s, q, d, j, rp[d * nrows + i]);
                }
            }
        }
    }
    /* COM phase: 1-cycles only (SingleCenterOfMass picks among next[i] == i) */
    for (int p = 0; p < (com_on ? N : 0); ++p) {
        if (s->next[p] != p) continue;
        keyed(seed, (uint32_t)p, sweep, chain, PUR_COM, x);
        double dsp[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) dsp[d] = (com_step * 2.0) * (o_u01(x[d]) - 0.5);
        double dv = 0.0, a[3], b[3];
        /* move_polymer! (src/updates/helper.jl:368-395): with a > 0 the shift is redrawn
         * while a moved bead lies closer than a to its nearest neighbour; draw t >= 2 is
         * words 0..dim-1 of keyed(seed, p, sweep, chain, PUR_HS_COM | t << 8) */
        int placed = 1;
        for (int attempt = 1;; ++attempt) {
            if (attempt > 1) {
                if (attempt > s->hs_ctr) { placed = 0; break; }
                uint32_t y[4];
                keyed(seed, (uint32_t)p, sweep, chain, PUR_HS_COM | ((uint32_t)attempt << 8), y);
                for (int d = 0; d < dim; ++d) dsp[d] = (com_step * 2.0) * (o_u01(y[d]) - 0.5);
            }
            for (int j = 0; j < M; ++j)
                for (int d = 0; d < dim; ++d) rp[d * M + j] = o_teleport(R(s, p, d, j) + dsp[d], s->L);
            if (!(s->hs_a > 0.0)) break;
            int near = 0;
            for (int j = 0; j < M && !near; ++j) {
                for (int d = 0; d < dim; ++d) a[d] = rp[d * M + j];
                near = hs_near(s, a, j, p);
            }
            if (!near) break;
        }
        if (!placed) { /* a rejected try */
            st->com_tries++;
            st->bead_moves += M;
            continue;
        }
        for (int j = 0; j < M; ++j) {
            for (int d = 0; d < dim; ++d) { a[d] = rp[d * M + j]; b[d] = R(s, p, d, j); }
            dv += o_V(&s->pot, a, dim) - o_V(&s->pot, b, dim);
        }
        double lw = -s->tau * dv;
        if (use_pair) {
            double du = 0.0, c0[3], c1[3];
            for (int j = 0; j < M; ++j) {
                int jn = (j + 1) % M;
                for (int d = 0; d < dim; ++d) {
                    a[d] = rp[d * M + j]; b[d] = rp[d * M + jn];
                    c0[d] = R(s, p, d, j); c1[d] = R(s, p, d, jn);
                }
                du += link_pair(s, p, j, a, b) - link_pair(s, p, j, c0, c1);
            }
            lw += du;
        }
        int acc = o_metropolis(exp(lw), o_u01(x[2]));
        st->com_tries++;
        st->bead_moves += M;
        if (acc) {
            st->com_acc++;
            for (int j = 0; j < M; ++j)
                for (int d = 0; d < dim; ++d) set_r(s, p, d, j, rp[d * M + j]);
        }
    }
    if (n_swap > 0 && sw_on) o_swaps(s, chain, sweep, seed, n_swap, swap_m, rule, st);
    o_sys_refresh_V(s);
    free(rp);
    free(nrm);
}

static const int every1[3] = {1, 1, 1};

void o_sweep(o_sys *s, uint32_t chain, uint32_t sweep, uint64_t seed, int ell, double com_step,
             o_sweepstats *st)
{
    sweep_core(s, chain, sweep, seed, ell, com_step, 0, 0, 0, every1, st);
}

void o_sweep_perm(o_sys *s, uint32_t chain, uint32_t sweep, uint64_t seed, int ell, double com_step, int n_swap,
                  int swap_m, int rule, o_sweepstats *st)
{
    sweep_core(s, chain, sweep, seed, ell, com_step, n_swap, swap_m, rule, every1, st);
}

void o_sweep_every(o_sys *s, uint32_t chain, uint32_t sweep, uint64_t seed, int ell, double com_step, int n_swap,
                   int swap_m, int rule, int staging_every, int com_every, int swap_every, o_sweepstats *st)
{
    const int every[3] = {staging_every > 1 ? staging_every : 1, com_every > 1 ? com_every : 1,
                          swap_every > 1 ? swap_every : 1};
    sweep_core(s, chain, sweep, seed, ell, com_step, n_swap, swap_m, rule, every, st);
}

/* free-particle propagator over m links, up to its normalisation:
 * exp(lnK) = exp(-|x - y|^2 / (4 lambda m tau)), minimum image (src/propagator.jl:16-19) */
static double rho0(const o_sys *s, const double *x, const double *y, int m)
{
    double dd = 0.0;
    for (int d = 0; d < s->dim; ++d) {
        double r = o_distance(x[d], y[d], s->L);
        dd += r * r;
    }
    return exp(-dd / ((4.0 * s->lambda) * ((double)m * s->tau)));
}

/* Swap tries of one sweep.  Try a: words of keyed(seed, a, sweep, chain,
 * PUR_SWAP) give the slice j0 = floor(x0 M / 2^32), the first particle n1 =
 * floor(x1 N / 2^32), the selection uniform u01(x2) and the Metropolis uniform
 * u01(x3).  E_i is the bead m links after (i, j0) along i's cycle.  n2 is drawn
 * with weights w_i = rho0(r_n1(j0), E_i) rho0(r_i(j0), E_n1) (the exchanged
 * product of sampleparticles), scanning i in order; n2 == n1 rejects.  Paths:
 * n1 from A = r_n1(j0) to D = E_n2 and n2 from B = r_n2(j0) to C = E_n1, by levy!
 * with normals keyed(seed, (2a + w) M + i, sweep, chain, PUR_SWAP_LEVY) paired as
 * in the staging move (w = 0 for n1, 1 for n2).  Delta lnW = -tau sum over the
 * interior beads of both paths of (V' - V).  rule 0 multiplies exp(Delta lnW)
 * by (1/W'(n1) + 1/W'(n2)) / (1/W(n1) + 1/W(n2)), W(n) = sum_i of n's selection
 * weights (primed: after the swap): the ratio of the reverse and forward
 * selection probabilities, either particle drawn first, which with free
 * bridges gives detailed balance.  rule 1 is the reference's exp(Delta lnW).
 * Accepted: new interiors written along the new cycles; for j0 + m < M the
 * beads j0 + m .. M - 1 of n1 and n2 are exchanged; next[n1] <-> next[n2]. */
/* The number of times find_nns, queried at a, lists a bead at b: the stencil cells of a's cell that
 * are b's cell (several when nbins < 3), if b lies within h (src/nearest_neighbours.jl:106-154) */
static int pair_mult(const o_sys *s, const double *a, const double *b)
{
    if (periodic_norm(s, a, b) > s->h) return 0;
    const int nb = s->nbins, dim = s->dim;
    static const int st2[9][2] = {{0, 0}, {-1, 1}, {0, 1}, {1, 1}, {-1, 0}, {1, 0}, {-1, -1}, {0, -1}, {1, -1}};
    static const int st1[3] = {0, -1, 1};
    const int ba = o_bin(s, a), bb = o_bin(s, b);
    const int bx = ba % nb, by = (dim >= 2) ? (ba / nb) % nb : 0;
    int mult = 0;
    for (int cc = 0; cc < ((dim == 1) ? 3 : 9); ++cc) {
        const int cell = (dim == 1) ? mod1(bx + st1[cc], nb)
                                    : mod1(bx + st2[cc][0], nb) + nb * mod1(by + st2[cc][1], nb);
        mult += cell == bb;
    }
    return mult;
}

/* lnU of the link (x at slice j, its next bead xn) with neighbour k's beads at slice j and along k's cycle */
static double link_term(const o_sys *s, int k, int j, const double *x, const double *xn)
{
    double rk[3], rkn[3], r1[3], r2[3];
    int kn, kj;
    next_bead(s, k, j, &kn, &kj);
    bead(s, k, j, rk);
    bead(s, kn, kj, rkn);
    for (int d = 0; d < s->dim; ++d) {
        r1[d] = o_distance(rk[d], x[d], s->L);
        r2[d] = o_distance(rkn[d], xn[d], s->L);
    }
    return o_lnU(r1, r2, s->dim, s->tau, &s->pair);
}

/* The pair part of a swap try's log acceptance ratio, before the swap (the old cycles in s).
 * Links t = 0..m-1 of the window from slice j0: the old paths run along the cycles of n1 and n2
 * (beads of x1 = owner(n1, j0 + t), x2 = owner(n2, j0 + t) at slice j), the new paths are rp1,
 * rp2 ((m + 1) x dim, column-major).
 * reference = 1: ReshapeSwapLinear's sums as written (src/updates/reshape.jl:178-245): every lnU
 * lands in w_initial, the old ones with the neighbour lists of x1 and x2 (each excluding itself
 * only, so the x1-x2 pair twice), the new ones with both x1 and x2 excluded (no pair between the
 * two new paths); the ratio is exp(-(those sums)).
 * reference = 0: lnU(new) - lnU(old) over the same links, each pair once: the neighbours other
 * than x1, x2 for each path, plus the pair of the two paths themselves. */
static double swap_pair_lw(const o_sys *s, int n1, int n2, int j0, int m, const double *rp1, const double *rp2,
                           int reference)
{
    const int M = s->M, dim = s->dim, nrows = m + 1;
    int list[NNCAP];
    double w_old = 0.0, w_new = 0.0;
    for (int t = 0; t < m; ++t) {
        const int u = j0 + t, j = u % M, jn = (j + 1) % M;
        const int x[2] = {owner(s, n1, u), owner(s, n2, u)};
        const int y[2] = {owner(s, n1, u + 1), owner(s, n2, u + 1)};
        double o0[2][3], o1[2][3], a0[2][3], a1[2][3];
        for (int w = 0; w < 2; ++w) {
            const double *rp = w == 0 ? rp1 : rp2;
            for (int d = 0; d < dim; ++d) {
                o0[w][d] = R(s, x[w], d, j);
                o1[w][d] = R(s, y[w], d, jn);
                a0[w][d] = rp[d * nrows + t];
                a1[w][d] = rp[d * nrows + t + 1];
            }
        }
        for (int w = 0; w < 2; ++w) {
            /* old: the neighbours of x_w's bead, x_w excluded */
            int cnt = o_find_nns(s, o0[w], j, x[w], list, NNCAP);
            if (cnt > NNCAP) cnt = NNCAP;
            for (int q = 0; q < cnt; ++q) {
                const int k = list[q];
                if (!reference && k == x[1 - w] && w == 1) continue; /* the x1-x2 pair once (from x1's list) */
                w_old += link_term(s, k, j, o0[w], o1[w]);
            }
            /* new: the neighbours of the new bead, x1 and x2 excluded */
            cnt = o_find_nns(s, a0[w], j, x[0], list, NNCAP);
            if (cnt > NNCAP) cnt = NNCAP;
            for (int q = 0; q < cnt; ++q) {
                const int k = list[q];
                if (k == x[1]) continue;
                w_new += link_term(s, k, j, a0[w], a1[w]);
            }
        }
        if (!reference) {
            /* the pair of the two new paths, as find_nns at the first path's bead would list it */
            const int mult = pair_mult(s, a0[0], a0[1]);
            if (mult > 0) {
                double r1[3], r2[3];
                for (int d = 0; d < dim; ++d) {
                    r1[d] = o_distance(a0[1][d], a0[0][d], s->L);
                    r2[d] = o_distance(a1[1][d], a1[0][d], s->L);
                }
                w_new += mult * o_lnU(r1, r2, dim, s->tau, &s->pair);
            }
        }
    }
    return reference ? -(w_old + w_new) : w_new - w_old;
}

double o_swap_pair_lw(const o_sys *s, int n1, int n2, int j0, int m, const double *rp1, const double *rp2,
                      int reference)
{
    return swap_pair_lw(s, n1, n2, j0, m, rp1, rp2, reference);
}

static void o_swaps(o_sys *s, uint32_t chain, uint32_t sweep, uint64_t seed, int n_swap, int m, int rule,
                    o_sweepstats *st)
{
    const int M = s->M, N = s->N, dim = s->dim;
    double *E = (double *)malloc(sizeof(double) * (size_t)N * dim);
    double *B0 = (double *)malloc(sizeof(double) * (size_t)N * dim);
    double *rp1 = (double *)malloc(sizeof(double) * (size_t)(m + 1) * dim);
    double *rp2 = (double *)malloc(sizeof(double) * (size_t)(m + 1) * dim);
    double *nrm = (double *)malloc(sizeof(double) * (size_t)(m + 1) * dim);
    uint32_t x[4];
    for (int a = 0; a < n_swap; ++a) {
        keyed(seed, (uint32_t)a, sweep, chain, PUR_SWAP, x);
        const int j0 = (int)(((uint64_t)x[0] * (uint64_t)M) >> 32);
        const int n1 = (int)(((uint64_t)x[1] * (uint64_t)N) >> 32);
        const double usel = o_u01(x[2]), uacc = o_u01(x[3]);
        st->swap_tries++;
        st->bead_moves += 2 * (m - 1);
        if (N < 2) continue;
        const int jm = j0 + m, sm = jm % M;
        for (int i = 0; i < N; ++i) {
            const int oi = jm < M ? i : (int)s->next[i];
            for (int d = 0; d < dim; ++d) {
                E[i * dim + d] = R(s, oi, d, sm);
                B0[i * dim + d] = R(s, i, d, j0);
            }
        }
        const double *A = B0 + n1 * dim, *C = E + n1 * dim;
        double W1 = 0.0;
        for (int i = 0; i < N; ++i) W1 += rho0(s, A, E + i * dim, m) * rho0(s, B0 + i * dim, C, m);
        const double target = usel * W1;
        double cum = 0.0;
        int n2 = N - 1;
        for (int i = 0; i < N; ++i) {
            cum += rho0(s, A, E + i * dim, m) * rho0(s, B0 + i * dim, C, m);
            if (cum > target) { n2 = i; break; }
        }
        if (n2 == n1) continue;
        const double *Bp = B0 + n2 * dim, *D = E + n2 * dim;
        /* the two bridges; with a > 0 hardspherelevy!(r1, s, j0, fpcycle2) and (r2, s, j0, fpcycle1)
         * (src/updates/reshape.jl:157-162): path 1's beads keep clear of every particle but n2, path 2's
         * of every particle but n1; redraws keyed by (2a + w) M + i and PUR_HS_SWAP.  A bead out of draws
         * rejects the try */
        int bridged = 1;
        for (int w = 0; w < 2; ++w) {
            double *rp = w == 0 ? rp1 : rp2;
            const double *st0 = w == 0 ? A : Bp, *en = w == 0 ? D : C;
            const int nrows = m + 1;
            for (int d = 0; d < dim; ++d) { rp[d * nrows] = st0[d]; rp[d * nrows + m] = en[d]; }
            for (int i = 1; i < m; ++i) {
                double n0, nn1;
                if (i & 1) keyed(seed, (uint32_t)((2 * a + w) * M + i), sweep, chain, PUR_SWAP_LEVY, x);
                if (i & 1) o_box_muller(x[0], x[1], &n0, &nn1);
                else o_box_muller(x[2], x[3], &n0, &nn1);
                nrm[(i - 1) * dim] = n0;
                if (dim > 1) nrm[(i - 1) * dim + 1] = nn1;
            }
            if (s->hs_a > 0.0)
                bridged &= hs_levy_x(s, w == 0 ? n2 : n1, j0, rp, nrows, nrm, chain, sweep, seed,
                                     (uint32_t)((2 * a + w) * M), 0, PUR_HS_SWAP);
            else
                o_levy(rp, nrows, dim, s->tau, s->L, s->lambda, nrm);
            for (int d = 0; d < dim; ++d) { rp[d * nrows] = st0[d]; rp[d * nrows + m] = en[d]; }
        }
        if (!bridged) continue;
        /* Delta lnW: interior beads of both paths, new minus old along the old cycles */
        double dv = 0.0, na[3], ob[3];
        for (int w = 0; w < 2; ++w) {
            const int nw = w == 0 ? n1 : n2;
            const double *rp = w == 0 ? rp1 : rp2;
            for (int i = 1; i < m; ++i) {
                const int u = j0 + i, j = u % M;
                for (int d = 0; d < dim; ++d) { na[d] = rp[d * (m + 1) + i]; ob[d] = R(s, owner(s, nw, u), d, j); }
                dv += o_V(&s->pot, na, dim) - o_V(&s->pot, ob, dim);
            }
        }
        double lw = -s->tau * dv;
        /* pair sums: as the reference writes them when the pair sums of the other moves are left
         * out of the ratio (pair_in_accept = 0), the physical difference otherwise */
        if (s->pair.kind != O_PAIR_NONE) lw += swap_pair_lw(s, n1, n2, j0, m, rp1, rp2, !s->pair_in_accept);
        double ratio = exp(lw);
        if (rule == 0) {
            /* selection normalisations before (E) and after (E' = E with n1, n2's ends exchanged) */
            double W2 = 0.0, W1p = 0.0, W2p = 0.0;
            for (int i = 0; i < N; ++i) {
                const double *Ei = E + i * dim, *Bi = B0 + i * dim;
                const double *Epi = i == n1 ? D : (i == n2 ? C : Ei);
                W2 += rho0(s, Bp, Ei, m) * rho0(s, Bi, D, m);
                W1p += rho0(s, A, Epi, m) * rho0(s, Bi, D, m);
                W2p += rho0(s, Bp, Epi, m) * rho0(s, Bi, C, m);
            }
            ratio *= (1.0 / W1p + 1.0 / W2p) / (1.0 / W1 + 1.0 / W2);
        }
        if (!o_metropolis(ratio, uacc)) continue;
        st->swap_acc++;
        const int q1 = (int)s->next[n1], q2 = (int)s->next[n2];
        for (int w = 0; w < 2; ++w) {
            const int nw = w == 0 ? n1 : n2, qn = w == 0 ? q2 : q1; /* the new next of nw */
            const double *rp = w == 0 ? rp1 : rp2;
            for (int i = 1; i < m; ++i) {
                const int u = j0 + i, j = u % M, q = u < M ? nw : qn;
                for (int d = 0; d < dim; ++d) set_r(s, q, d, j, rp[d * (m + 1) + i]);
            }
        }
        if (jm < M) {
            for (int j = jm; j < M; ++j)
                for (int d = 0; d < dim; ++d) {
                    double t = s->r[((size_t)n1 * dim + d) * M + j];
                    set_r(s, n1, d, j, s->r[((size_t)n2 * dim + d) * M + j]);
                    set_r(s, n2, d, j, t);
                }
        }
        s->next[n1] = q2;
        s->next[n2] = q1;
    }
    free(E);
    free(B0);
    free(rp1);
    free(rp2);
    free(nrm);
}

/* ------------------------------------------------------------------------ */
/* init_world (src/system.jl:36-78), Philox-keyed                            */
/* ------------------------------------------------------------------------ */
/* One try of particle p: the first bead uniform in [-L, L)^dim (2L (u - 1/2)),
 * the last bead equal to it, and levy! over the M - 1 links in between.
 * Draws of try t: start = keyed(seed, p, t, chain, PUR_INIT_START) words
 * 0..dim-1; the normals of interior bead i (1 <= i <= M - 2) come in pairs
 * (i, i+1), i odd, from keyed(seed, p*M + i, t, chain, PUR_INIT_LEVY):
 * Box-Muller of words (x, y) for bead i and (z, w) for bead i + 1.
 * rp: M x dim, column-major by coordinate (rp[d*M + j]). */
static void init_try(uint64_t seed, uint32_t chain, int p, uint32_t t, int M, int dim, double L, double tau,
                     double lambda, double *rp, double *nrm)
{
    uint32_t x[4];
    keyed(seed, (uint32_t)p, t, chain, PUR_INIT_START, x);
    for (int d = 0; d < dim; ++d) {
        const double r0 = 2.0 * L * (o_u01(x[d]) - 0.5);
        rp[d * M] = r0;
        rp[d * M + (M - 1)] = r0;
    }
    for (int i = 1; i <= M - 2; ++i) {
        double n0, n1;
        if (i & 1) keyed(seed, (uint32_t)(p * M + i), t, chain, PUR_INIT_LEVY, x);
        if (i & 1) o_box_muller(x[0], x[1], &n0, &n1);
        else o_box_muller(x[2], x[3], &n0, &n1);
        nrm[(i - 1) * dim] = n0;
        if (dim > 1) nrm[(i - 1) * dim + 1] = n1;
    }
    o_levy(rp, M, dim, tau, L, lambda, nrm);
}

/* init_world with hard spheres of radius a (a = 0: every particle's first
 * try).  Particles in order; particle p > 0 is drawn again (try t = 0, 1, ...)
 * while any of its beads lies closer than a to an earlier particle's bead at
 * the same slice (norm of the per-component |minimum-image| offsets, d < a),
 * at most 10 000 tries (src/system.jl:44-64).  The reference's extra levy! on
 * a found overlap only consumes draws: the try is discarded either way.
 * out: N x dim x M; tries[p] (may be NULL): the accepted try + 1.  Returns 0,
 * or -1 when a particle found no place. */
int o_init_world_hs(uint64_t seed, uint32_t chain, int N, int M, int dim, double L, double tau, double lambda,
                    double a, double *out, int *tries)
{
    double *rp = (double *)malloc(sizeof(double) * (size_t)M * dim);
    double *nrm = (double *)malloc(sizeof(double) * (size_t)(M > 2 ? M - 2 : 1) * dim);
    int rc = 0;
    for (int p = 0; p < N && rc == 0; ++p) {
        uint32_t t = 0;
        for (;; ++t) {
            if (t >= 10000u) {
                rc = -1;
                break;
            }
            init_try(seed, chain, p, t, M, dim, L, tau, lambda, rp, nrm);
            int clash = 0;
            if (a > 0.0)
                for (int j = 0; j < M && !clash; ++j)
                    for (int i = 0; i < p && !clash; ++i) {
                        double d2 = 0.0;
                        for (int d = 0; d < dim; ++d) {
                            const double dd = o_distance(out[((size_t)i * dim + d) * M + j], rp[d * M + j], L);
                            d2 += dd * dd;
                        }
                        clash = sqrt(d2) < a;
                    }
            if (!clash) break;
        }
        if (tries) tries[p] = (int)t + 1;
        for (int d = 0; d < dim; ++d)
            for (int j = 0; j < M; ++j) out[((size_t)p * dim + d) * M + j] = rp[d * M + j];
    }
    free(rp);
    free(nrm);
    return rc;
}

/* init_world with a = 0 */
void o_init_world(uint64_t seed, uint32_t chain, int N, int M, int dim, double L, double tau, double lambda,
                  double *out)
{
    (void)o_init_world_hs(seed, chain, N, M, dim, L, tau, lambda, 0.0, out, NULL);
}
