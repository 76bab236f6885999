/*
 * pimc_oracle.h -- CPU FP64 restatement of PIMC.jl's sampling path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only the test-suite, the smoke check and the CPU-baseline
 * leg of the benchmark load it.  The product never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the PIMC.jl source tree).  Indices are 0-based here; the reference is
 * 1-based.  Positions use the reference's `Particle.r` layout: for particle n,
 * coordinate d, slice j the value sits at r[(n*dim + d)*M + j] (an M x dim
 * column-major matrix per particle, particles concatenated).
 *
 * Built with -ffp-contract=off so that a*b+c is rounded twice, as Julia does
 * without @fastmath/@muladd.
 */
#ifndef PIMC_ORACLE_H
#define PIMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- potentials (the Julia closures become an enum + parameters) ---- */
enum { O_POT_NONE = 0, O_POT_HARMONIC = 1, O_POT_LATTICE = 2, O_POT_OPTICAL = 3 };
#define O_MAX_BEAMS 16
enum { O_PAIR_NONE = 0, O_PAIR_GAUSSIAN = 1, O_PAIR_TABLE = 2 };

typedef struct {
    int kind;        /* O_POT_* */
    double omega;    /* harmonic: V = 1/2 omega^2 |r|^2 (m = 1) */
    double V0, a;    /* lattice: V = V0 * sum_d cos^2(pi x_d / a) */
    /* optical: V = V0 * normalized_intensity(r, angles, scale; helical)
     * (examples/tools/potentialtools.jl:1-16), 2D */
    int nbeam, helical;
    double scale;
    double angle[O_MAX_BEAMS];
} o_pot;

typedef struct {
    int kind;        /* O_PAIR_* */
    double g, sigma; /* v(r) = g exp(-r^2 / (2 sigma^2)) */
    /* O_PAIR_TABLE: the integral terms of the pair-product propagator
     * (prop_rel_interpolate_terms, src/propagator.jl:36-70) on the n x n grid
     * r_i = r0 + i (r1 - r0) / (n - 1), row-major tab[i * n + k] = terms(r_i, r_k);
     * mu: the value of lnU where the propagator is negative (src/system.jl:27) */
    const double *tab;
    int tab_n;
    double tab_r0, tab_r1, mu;
} o_pairpot;

/* ---- system state (src/system.jl:93-126 restricted to the hot path) ---- */
typedef struct {
    int dim, M, N;
    double lambda, L, beta, tau;
    o_pot pot;
    o_pairpot pair;
    int nbins;          /* cell-list bins per dimension: floor(2L / r_a) (src/system.jl:81) */
    double h;           /* neighbour radius = cell side 2L/nbins (src/nearest_neighbours.jl:127) */
    int pair_in_accept; /* 0: pair sums computed and dropped, as the reference does
                           (src/updates/helper.jl:306-366 take w by value); 1: they enter the
                           Metropolis ratio */
    double *r;          /* N*dim*M */
    double *V;          /* N*M cached potential link actions (src/system.jl:72-74) */
    int64_t *next;      /* N, 0-based; identity for the hot-path configs */
    double hs_a;        /* hard-sphere radius a = exp(-2 pi / g) of interacting systems (src/system.jl:148);
                           0 switches the rejection off */
    int hs_ctr;         /* redraws allowed per bead or COM shift (s.ctr: 10000 thermalising, 1000 measuring,
                           src/simulation.jl:31-32) */
    void *cells;        /* per-slice index of the beads by cell (built on first use, kept in step with every
                           position write); speeds up find_nns / find_nn without changing their results */
} o_sys;

/* ---- Philox4x32-10 (Salmon et al., SC'11) and the draw helpers ---- */
void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double o_u01(uint32_t x);                              /* (x + 0.5) * 2^-32 */
void o_box_muller(uint32_t x0, uint32_t x1, double *n0, double *n1);

/* ---- primitives, src/propagator.jl ---- */
double o_distance(double x1, double x2, double L);    /* :6-9   */
double o_teleport(double x, double L);                /* :30-32 */
double o_lnK(const double *r1, const double *r2, int dim, double tau, double lambda, double L); /* :16-19 */
double o_V(const o_pot *p, const double *r, int dim);
void   o_dV(const o_pot *p, const double *r, int dim, double *out);
double o_lnV(const double *r1, const double *r2, int dim, double tau, const o_pot *p);  /* :26-28 */
double o_vpair(const o_pairpot *pp, double r);
double o_lnU(const double *r1rel, const double *r2rel, int dim, double tau, const o_pairpot *pp); /* src/system.jl:26-28 */
int    o_metropolis(double delta, double u);           /* src/updates/helper.jl:3-5 */

/* ---- pair-product propagator, src/propagator.jl:35-89 ----
 * int_terms(r1, r2) = -term1 - term2 + term3 of prop_rel_interpolate_terms
 * (:36-70) by a fixed Gauss-Legendre rule over k in place of quadgk: 16-point
 * panels, 40 geometric panels [h 2^-(p+1), h 2^-p] below h = min(1/2, 4/rmax),
 * uniform panels of width <= h from h to K = sqrt(40/tau) (exp(-tau K^2) =
 * 4e-18), so that (r1 + r2) * width <= 8 on every panel. */
int    o_prop_quad_nodes(double rmax, double tau, double *k, double *w, int cap); /* node count */
double o_prop_int_term(double r1, double r2, double g0, double tau, double rmax);
void   o_prop_int_terms(int n, double r0, double r1, double g0, double tau, double *out); /* n x n, grid r0..r1 */
/* bilinear interpolation on the grid (scale(interpolate(A, BSpline(Linear())), ...)),
 * arguments clamped to [r0, r1] */
double o_prop_table_eval(const o_pairpot *pp, double x, double y);
/* prop_int(r1_rel, r2_rel, tau) = 1 + terms(|r1|, |r2|) / prop_rel0(r1_rel, r2_rel, tau) (:72-89) */
double o_prop_int(const o_pairpot *pp, const double *r1rel, const double *r2rel, int dim, double tau);

/* levy! (src/updates/helper.jl:118-139) with injected normals.  rp holds
 * nrows x dim column-major (rp[d*nrows + j]); rows 0 and nrows-1 are the
 * fixed ends.  normals[(j-1)*dim + d] is the draw for interior row j. */
void o_levy(double *rp, int nrows, int dim, double tau, double L, double lambda, const double *normals);

/* ---- system set-up ---- */
int  o_sys_init(o_sys *s, int dim, int M, int N, double lambda, double L, double beta,
                o_pot pot, o_pairpot pair, double r_a, int pair_in_accept, const double *r);
void o_sys_free(o_sys *s);
void o_sys_refresh_V(o_sys *s);                        /* src/system.jl:72-74 */
int  o_bin(const o_sys *s, const double *r);           /* src/nearest_neighbours.jl:26-33, 0-based */

/* find_nns (src/nearest_neighbours.jl:106-154): neighbours of the point
 * `target` at slice j, excluding particle `exclude`; candidates come from the
 * 9-cell stencil (with the reference's repeats when nbins < 3) and are kept
 * when their periodic distance is <= h.  Returns the count (<= cap). */
int o_find_nns(const o_sys *s, const double *target, int j, int exclude, int *out, int cap);
/* find_nn (src/nearest_neighbours.jl:157-180): the nearest bead at slice j among the
 * particles of the 9 stencil cells around target, other than exclude; -1 if none */
int o_find_nn(const o_sys *s, const double *target, int j, int exclude);

/* interaction_action! values (src/updates/helper.jl:306-366), i.e. what the
 * reference would add to w if w were passed by reference. */
double o_interaction_old(const o_sys *s, int n, int j);                   /* :306-324 */
double o_interaction_new(const o_sys *s, const double *rp, int nrows, int jp, int j0, int n); /* :326-347 */
double o_interaction_com(const o_sys *s, const double *rp_all, int j, int n); /* :349-366 */

/* ---- ReshapeLinear (src/updates/reshape.jl:31-91) ----
 * n, j0 0-based, m = number of links (>= 2), normals as for o_levy (m-1 rows).
 * Fills rp ((m+1) x dim) and the four weights; returns 1 if the bridge was
 * built.  u is the Metropolis uniform; *accepted gets the decision and the
 * state is committed on accept. */
typedef struct { double wV_initial, wV_updated, wU_initial, wU_updated; } o_weights;
int o_reshape_propose(const o_sys *s, int n, int j0, int m, const double *normals, double *rp, o_weights *w);
int o_reshape_apply(o_sys *s, int n, int j0, int m, const double *normals, double u, int *accepted, o_weights *w);

/* ---- SingleCenterOfMass (src/updates/com.jl:136-224, move_polymer! helper.jl:368-395) ---- */
int o_com_propose(const o_sys *s, int n, const double *d, double *rp /* M x dim */, o_weights *w);
int o_com_apply(o_sys *s, int n, const double *d, double u, int *accepted, o_weights *w);

/* ---- estimators, src/measurement.jl ---- */
void o_energy(const o_sys *s, double *E, double *Ev);  /* :92-122 */
/* pair energy added by this build: thermodynamic (link form with the
 * slice-j cut-off indicator) and virial parts. */
void o_pair_energy(const o_sys *s, double *Epair, double *Evir_pair);
/* Density (src/measurement.jl:45-55): nbins^dim histogram, reference's bin
 * shift (first bin dropped).  dens is accumulated, *ndata += M. */
void o_density(const o_sys *s, int nbins, double *dens, int64_t *ndata);
/* radial histogram (build addition): bins of width rmax/nb on |r| */
void o_radial(const o_sys *s, int nb, double rmax, double *hist, int64_t *ndata);

/* ---- full action, used to pin the Delta S computations by brute force ---- */
double o_total_lnV(const o_sys *s);          /* sum over all links of lnV */
double o_total_lnU(const o_sys *s);          /* sum over all links of the pair link action */
double o_total_lnK(const o_sys *s);

/* ---- the reference schedule: run! (src/simulation.jl:29-42) with
 * ReshapeLinear(s, cap) and SingleCenterOfMass(s, step), weights 1:1,
 * adaptivity (src/updates/helper.jl:6-52) and a Philox stream in place of
 * Julia's default RNG.  Counts proposed beads (m-1 per Levy try, M per COM
 * try).  If E/Ev are non-null an Energy measurement is taken every ncycle
 * attempts (src/measurement.jl:1-17) and written there (cap nE). */
typedef struct {
    int cap;              /* staging length cap (NumbOfSlices.m) */
    double step;          /* COM step (Step.size) */
    int64_t bead_moves;
    int64_t levy_tries, levy_acc, com_tries, com_acc;
    uint64_t rng_counter;
    int64_t n_meas;
    /* sliding windows of the Counter structs */
    int win_range, adj;
} o_refstate;
void o_refstate_init(o_refstate *st, const o_sys *s, int cap, double step);
int  o_run_reference(o_sys *s, o_refstate *st, int64_t n_iter, uint64_t seed, int ncycle,
                     double *E, double *Ev, int64_t nE);

/* ---- the batched sweep schedule of the HIP kernels, restated in FP64 ----
 * One sweep of chain `chain`: staging segments of length ell tiling each ring
 * at a Philox-drawn offset (particles in order), then one centre-of-mass try
 * per particle.  Draws are keyed by (chain, sweep, bead).  Returns the number
 * of accepted Levy segments and COM moves. */
typedef struct { int64_t levy_acc, levy_tries, com_acc, com_tries, bead_moves, swap_acc, swap_tries; } o_sweepstats;
void o_sweep(o_sys *s, uint32_t chain, uint32_t sweep, uint64_t seed, int ell, double com_step,
             o_sweepstats *st);
/* The sweep with permutations: the staging window of particle p runs along
 * its permutation cycle (beads past slice M-1 belong to next[p]), the
 * centre-of-mass try moves 1-cycles only (SingleCenterOfMass,
 * src/updates/com.jl:136-160), and n_swap swap tries of m links follow
 * (ReshapeSwapLinear, src/updates/reshape.jl:99-283, with sampleparticles,
 * src/updates/helper.jl:224-267).  rule 0: acceptance with the heat-bath
 * normalisations (detailed balance); rule 1: the reference's
 * exp(w_updated - w_initial) only.  With a pair action every pair link follows
 * the cycles (next()), and the swap adds its pair sums: as the reference
 * writes them when pair_in_accept = 0, the pair action's difference when 1
 * (o_swap_pair_lw).  With hard spheres the swap's bridges are drawn as
 * hardspherelevy!(r1, s, j0, fpcycle2) and (r2, s, j0, fpcycle1) draw them
 * (src/updates/reshape.jl:157-162): redraws keyed by (2a + w) M + i and
 * PUR_HS_SWAP, path 1 clear of every particle but n2, path 2 of every particle
 * but n1; a bead out of draws rejects the try. */
void o_sweep_perm(o_sys *s, uint32_t chain, uint32_t sweep, uint64_t seed, int ell, double com_step, int n_swap,
                  int swap_m, int rule, o_sweepstats *st);
/* o_sweep_perm with move weights: the staging pass, the centre-of-mass pass and the swap tries run
 * only when sweep % every == 0 (every <= 1: always), the sweep-granular form of run!'s draw of one
 * update with weight 1/every (src/simulation.jl:33-36) */
void o_sweep_every(o_sys *s, uint32_t chain, uint32_t sweep, uint64_t seed, int ell, double com_step, int n_swap,
                   int swap_m, int rule, int staging_every, int com_every, int swap_every, o_sweepstats *st);
/* The pair part of the log acceptance ratio of a swap try (window of m links from slice j0, the
 * paths of n1 and n2 along their cycles replaced by rp1, rp2: (m + 1) x dim, rp[d * (m + 1) + i]):
 * reference = 1 the sums of ReshapeSwapLinear as written (src/updates/reshape.jl:178-245, all in
 * w_initial), 0 the difference of the pair action with each pair once */
double o_swap_pair_lw(const o_sys *s, int n1, int n2, int j0, int m, const double *rp1, const double *rp2,
                      int reference);
/* the COM step adaptation applied between launches while thermalising */
double o_adapt_step(double step, int64_t acc, int64_t tries, double L);
/* the staging-length adaptation (NumbOfSlices) applied between thermalising launches */
int o_adapt_ell(int ell, int64_t acc, int64_t tries, int M);
/* bead-moves of one sweep (Levy m-1 per segment, M per COM try) */
int64_t o_sweep_bead_moves(int N, int M, int ell);

/* init_world (src/system.jl:36-78) with a = 0 and Philox-keyed draws, the
 * device initialisation pimc_init_paths restates this; out N x dim x M */
void o_init_world(uint64_t seed, uint32_t chain, int N, int M, int dim, double L, double tau, double lambda,
                  double *out);
/* init_world with hard spheres of radius a: particle p > 0 redrawn (try t keyed
 * by sweep word t) while a bead lies closer than a to an earlier particle's at
 * the same slice, at most 10 000 tries; tries[p] = accepted try + 1 (or NULL).
 * Returns 0, or -1 when a particle found no place. */
int o_init_world_hs(uint64_t seed, uint32_t chain, int N, int M, int dim, double L, double tau, double lambda,
                    double a, double *out, int *tries);

#ifdef __cplusplus
}
#endif
#endif
