/*
 * pimc_hip.h -- C ABI of the MI355X (gfx950) sampling path of PIMC.jl.
 *
 * The library owns device memory, its HIP streams (one per context, and a
 * second that a context's window-kernel chains may run on, joined back into
 * the first within every launch) and the kernels; the caller owns every host
 * buffer (copies are synchronous).  No exception or abort
 * crosses this boundary: every entry point returns PIMC_OK (0) or a negative
 * PIMC_ERR_* code, and pimc_last_error() describes the last failure.  A
 * context is used by one host thread at a time; separate contexts may run
 * concurrently on their own streams.  The library never calls back into the
 * caller.
 *
 * Reference interfaces replaced (paths relative to the PIMC.jl tree):
 *   System(potential; dim, M, N, L, T, lambda, ...)   src/system.jl:129-167   -> pimc_create
 *   Particle.r (M x dim per particle)                  src/system.jl:1-6       -> pimc_set_paths / pimc_get_paths
 *   run!(s, n, updates; Zmeasurements)                 src/simulation.jl:29-42 -> pimc_run
 *   (u::ReshapeLinear)(s) with levy!                   src/updates/reshape.jl:31-91,
 *                                                      src/updates/helper.jl:118-181 -> pimc_levy_propose / pimc_reshape_linear
 *   (u::SingleCenterOfMass)(s), move_polymer!          src/updates/com.jl:136-224,
 *                                                      src/updates/helper.jl:368-395 -> pimc_com_propose / pimc_center_of_mass
 *   interaction_action!(w, s, fpcycle, j)              src/updates/helper.jl:306-324 -> pimc_interaction
 *   (u::Energy)(s)                                     src/measurement.jl:92-122 -> pimc_energy
 *   (d::Density)(s)                                    src/measurement.jl:45-55 -> pimc_density
 *   build_prop_int(L, g0, tau)                         src/propagator.jl:35-89 -> pimc_prop_int_terms +
 *                                                      pimc_set_pair_table (PIMC_PAIR_TABLE chains)
 */
#ifndef PIMC_HIP_H
#define PIMC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIMC_OK 0
#define PIMC_ERR_ARG (-1)      /* invalid argument or descriptor */
#define PIMC_ERR_HIP (-2)      /* a HIP runtime call failed */
#define PIMC_ERR_STATE (-3)    /* call not valid in the current state */
#define PIMC_ERR_CAPACITY (-4) /* a fixed-size device structure overflowed (e.g. a cell list) */
#define PIMC_ERR_NOMEM (-5)

/* external potentials: the reference passes a Julia closure `potential` and
 * `dV`; a device needs an enum plus parameters */
#define PIMC_POT_NONE 0
#define PIMC_POT_HARMONIC 1 /* V = 1/2 omega^2 |r|^2, dV = omega^2 r */
#define PIMC_POT_LATTICE 2  /* V = V0 sum_d cos^2(pi x_d / a)         */
#define PIMC_POT_OPTICAL 3  /* V = V0 normalized_intensity(r, angles, scale; helical) of
                               examples/tools/potentialtools.jl:1-38 (2D): the squared mean of the
                               beams' plane waves; V0 < 0 is the attractive lattice of generate_V */
#define PIMC_MAX_BEAMS 16

/* pair actions, entering through the reference's lnU hook (src/system.jl:26-28) */
#define PIMC_PAIR_NONE 0
#define PIMC_PAIR_GAUSSIAN 1 /* primitive link action of v(r) = g exp(-r^2 / (2 sigma^2)) */
#define PIMC_PAIR_TABLE 2    /* pair-product propagator of build_prop_int (src/propagator.jl:35-89):
                                lnU = log(propint) (-mu where propint < 0) with propint = 1 + terms(|r1|, |r2|)
                                / prop_rel0(r1, r2, tau), terms bilinear in a table set by pimc_set_pair_table.
                                Needs r_a > 0; no pair energy, no adjoint, no derivatives */

#define PIMC_FP32 0
#define PIMC_FP64 1

/* One independent chain (the reference's System).  hbar = k_B = 1. */
typedef struct pimc_chain_desc {
    int32_t N, M, dim;       /* particles, time slices, dimension (1 or 2) */
    int32_t potential;       /* PIMC_POT_* */
    int32_t pair;            /* PIMC_PAIR_* */
    int32_t pair_in_accept;  /* 1: pair action enters the Metropolis ratio; 0: the
                                reference's behaviour, where the pair sums of
                                interaction_action! are computed and dropped */
    int32_t ell;             /* staging segment length >= 2 (0: library default) */
    int32_t swaps;           /* swap tries per sweep (ReshapeSwapLinear, src/updates/reshape.jl:99-283);
                                chains with swaps run whole per workgroup (with hard spheres, the bridges of
                                a swap are redrawn as hardspherelevy! draws them, :157-162).  With a pair
                                action the swap's pair sums follow pair_in_accept: 0 the reference's
                                sums (old and new lnU both in w_initial, the pair of the two moved paths
                                listed twice among the old and not among the new, :178-245); 1 the
                                difference of the pair action, each pair once */
    double beta, lambda, L;  /* L is the half-width: box [-L, L)^dim (src/system.jl:101) */
    double omega;            /* harmonic trap */
    double V0, a;            /* square lattice */
    double g, sigma;         /* Gaussian pair repulsion */
    double r_a;              /* cell-list range: nbins = floor(2L / r_a); 0 -> L/4 (src/system.jl:22-24) */
    double com_step;         /* initial centre-of-mass step (SingleCenterOfMass(s, step)) */
    int32_t swap_m;          /* links of a swap try, 2 .. M - 1 (0: ell) */
    int32_t swap_rule;       /* 0: acceptance with the heat-bath normalisations of the partner choice
                                (detailed balance); 1: the reference's exp(w_updated - w_initial) */
    double hs_a;             /* hard-sphere radius a (System's exp(-2 pi / g) with interactions, src/system.jl:148):
                                bridge beads and COM shifts are redrawn while a moved bead lies closer than a to
                                another particle's bead at its slice (hardspherelevy!, move_polymer!,
                                src/updates/helper.jl:141-181, 368-395), at most 10000 draws thermalising and
                                1000 measuring; 0: off.  PIMC_PAIR_TABLE chains only, a <= 2L / nbins; their
                                single-move hooks are refused */
    double beam_scale;       /* PIMC_POT_OPTICAL: wave number of the beams (generate_V's scale) */
    int32_t n_beams;         /* 1 .. PIMC_MAX_BEAMS */
    int32_t helical;         /* 1: beam b carries the phase a_b (normalized_intensity's helical) */
    double beam_angle[PIMC_MAX_BEAMS]; /* beam directions a_b: the wave vector is scale (sin a_b, cos a_b) */
    /* the move weights of run!(s, n, updates) (src/simulation.jl:33-36), which draws each update with
     * weight 1/every: here the staging pass (ReshapeLinear), the centre-of-mass pass (SingleCenterOfMass)
     * and the swap tries (ReshapeSwapLinear) of a sweep run on the global sweeps s (the Philox sweep word)
     * with s % every == 0, so each runs on a fraction 1/every of the sweeps.  0 or 1: every sweep (the
     * default); at most 65536.  Each pass keeps detailed balance on its own, so the weights change the
     * mixing rate, not the sampled distribution */
    int32_t staging_every, com_every, swap_every;
} pimc_chain_desc;

typedef struct pimc_options {
    int32_t precision;         /* PIMC_FP32 (positions and Delta S in FP32) or PIMC_FP64 */
    int32_t sweeps_per_launch; /* sweeps per kernel launch = one estimator block (0: 50) */
    int32_t radial_bins;       /* radial density histogram per chain (0: off) */
    int32_t cart_bins;         /* the reference's Cartesian Density per chain (0: off) */
    double radial_rmax;        /* radial histogram range [0, rmax) */
    int32_t tile_beads;        /* beads staged per workgroup (0: autotuned) */
    int32_t block_threads;     /* tile-kernel workgroup size, a multiple of 64 and <= 256 (0: 256) */
    int32_t use_graph;         /* replay launch sequences from a captured hipGraph */
    int32_t derivatives;       /* run every chain whole per workgroup and record the
                                  reweighting sums for d<E>/d omega, d<E>/d g */
    uint64_t seed;             /* Philox key */
    int32_t measure_every;     /* estimators after every k-th sweep (0: 1), at global sweeps s with
                                  (s + 1) % k == 0; the measurement cycle Ncycle of
                                  measurement_Z_sector (src/measurement.jl:1-17), counted in sweeps.
                                  sweeps_per_launch must be a multiple of it, and so must the sweep
                                  count of a measuring pimc_run */
    int32_t adapt_ell;         /* 1: every chain adapts its staging length after every thermalising launch by
                                  adjust!(::NumbOfSlices) (src/updates/helper.jl:42-52) on that launch's staging
                                  acceptance: -1 below 0.6, +1 above 0.8, within [2, min(32, M - 2)]
                                  (ReshapeLinear's minslices and maxslices, src/updates/reshape.jl:12-21).  The
                                  tiles of chains without a pair action are rebuilt between launches when a length
                                  changes (a host synchronisation per thermalising launch; a length that would give
                                  a particle more segments than a workgroup holds is refused); Gaussian-pair chains
                                  then run on the whole-chain kernel.  pimc_get_ell reads the current value */
    int32_t measure_kernel;    /* estimators of chains run whole per workgroup, in measuring runs: 1 inside the
                                  sweep kernel; 2 in a wide kernel of their own after each sweep (the sweep
                                  kernel then runs one sweep per launch); 0: 2 when such a chain has a pair
                                  action and at least 8192 beads, else 1.  Same samples either way */
    int32_t pair_kernel;       /* chains with a Gaussian pair action in the acceptance (pair_in_accept 1) and no
                                  swaps, hard spheres or staging-length adaptation: 1 the whole-chain kernel (one workgroup per chain, tries in
                                  order); 2 the window kernels (staging by slice windows, centre-of-mass tries
                                  decided in dependency rounds) wherever their LDS state fits; 0: 2 for chains
                                  with cell lists (more than 96 particles), else 1.  Same sampling either way */
} pimc_options;

typedef struct pimc_counters {
    int64_t sweeps;            /* sweeps run by this call */
    int64_t bead_moves;        /* proposed beads: m-1 per staging try, M per COM try */
    int64_t levy_tries, levy_acc, com_tries, com_acc;
    double wall_ms;            /* device time of the whole call (events on the library stream) */
    double kernel_ms;          /* summed device time of the sweep kernel launches */
    int64_t kernel_launches;
    int64_t swap_tries, swap_acc;
} pimc_counters;

/* Estimates of one chain from the measured blocks (block means, standard
 * error from their spread).  E: thermodynamic energy (src/measurement.jl:111)
 * plus the pair energy; Ev: virial energy (src/measurement.jl:112) plus the
 * pair virial. */
typedef struct pimc_estimates {
    int64_t n_blocks, n_meas;
    double E, E_err, Ev, Ev_err;
    double Epair, Epair_err;
    double dSdomega, dSdg;             /* <dS/d omega>, <dS/d g> */
    double dE_domega, dE_domega_err;   /* d<E>/d omega by reweighting */
    double dE_dg, dE_dg_err;           /* d<E>/d g by reweighting */
} pimc_estimates;

typedef struct pimc_ctx pimc_ctx;

/* System(...) for a batch of chains (src/system.jl:129-167).  Initial paths
 * are zero; set them with pimc_set_paths. */
int pimc_create(pimc_ctx **ctx, const pimc_chain_desc *chains, int32_t n_chains, const pimc_options *opt);
int pimc_destroy(pimc_ctx *ctx);
const char *pimc_last_error(const pimc_ctx *ctx);
int pimc_device_info(char *name, int32_t len, int32_t *cu_count);

/* Particle.r of every particle of one chain: N x dim x M doubles (each
 * particle an M x dim column-major matrix, src/system.jl:1-6). */
int pimc_set_paths(pimc_ctx *ctx, int32_t chain, const double *r);
int pimc_get_paths(pimc_ctx *ctx, int32_t chain, double *r);
/* initial paths on the device, replacing init_world (src/system.jl:36-78): per
 * particle the first bead uniform in [-L, L)^dim, the last bead equal to it and
 * levy! over the M - 1 links in between; Philox-keyed by (seed, chain, particle,
 * try).  For a chain with hard spheres (hs_a = a > 0) the particles are placed
 * in order and a particle is drawn again while one of its beads lies closer than
 * a to an earlier particle's bead at the same slice, at most 10 000 times
 * (src/system.jl:44-64), else PIMC_ERR_CAPACITY.  chain = -1 initialises every
 * chain.  Cell lists are rebuilt. */
int pimc_init_paths(pimc_ctx *ctx, int32_t chain, uint64_t seed);
/* permutation of one chain (Worldline.next, src/system.jl:1-6; next(), src/updates/helper.jl:287-300):
 * next[p] is the particle whose slice 0 follows slice M - 1 of particle p (0-based).  Identity
 * unless the chain has swap moves; set_next checks that it is a permutation */
int pimc_set_next(pimc_ctx *ctx, int32_t chain, const int32_t *next);
int pimc_get_next(pimc_ctx *ctx, int32_t chain, int32_t *next);
int pimc_get_com_step(pimc_ctx *ctx, int32_t chain, double *step);
int pimc_set_com_step(pimc_ctx *ctx, int32_t chain, double step);
/* staging length of one chain (NumbOfSlices.m of ReshapeLinear, src/updates/reshape.jl:7-28);
 * changes only under pimc_options.adapt_ell */
int pimc_get_ell(pimc_ctx *ctx, int32_t chain, int32_t *ell);

/* Philox sweep counter of a context (the sweep word of every draw; pimc_run advances it).  A run
 * resumed in a new context from saved paths (pimc_get_paths / pimc_set_paths), permutations
 * (pimc_get_next / pimc_set_next), COM steps (pimc_get_com_step / pimc_set_com_step) and this
 * counter continues the saved run's random streams exactly.  There is no reference counterpart (the
 * reference has no checkpoint, only CSV output, examples/tools/savetools.jl:4-69); the counter is
 * below 2^32. */
int pimc_get_sweep(pimc_ctx *ctx, int64_t *sweep);
int pimc_set_sweep(pimc_ctx *ctx, int64_t sweep);

/* prop_rel_interpolate_terms (src/propagator.jl:36-70) on the device: terms[i * n + k] =
 * -term1 - term2 + term3 at (r_i, r_k), r_i = r0 + i (r1 - r0) / (n - 1), for contact strength g0
 * and time step tau; the k-integrals by a fixed Gauss-Legendre rule in place of quadgk.  The
 * reference's grid is n = 600, r0 = 1e-20, r1 = L.  Takes no context: the current device, errors
 * in pimc_last_error(NULL) */
int pimc_prop_int_terms(int32_t n, double r0, double r1, double g0, double tau, double *terms);
/* the table of a PIMC_PAIR_TABLE chain (build_prop_int's interpolation, src/propagator.jl:67-89):
 * n x n terms on the grid r0..r1 (as pimc_prop_int_terms fills it), arguments clamped to the grid;
 * mu is lnU where propint < 0 (System's mu, src/system.jl:26-28).  Must precede pimc_run and the
 * single-move hooks of that chain; may be replaced between runs */
int pimc_set_pair_table(pimc_ctx *ctx, int32_t chain, int32_t n, double r0, double r1, double mu,
                        const double *terms);

/* run!(s, n, [(1, SingleCenterOfMass), (1, ReshapeLinear)]; Zmeasurements)
 * as batched sweeps (src/simulation.jl:29-42).  One sweep: per particle,
 * staging segments tiling the ring at a random offset, then one centre-of-mass
 * try.  measure = 0 thermalises (COM step adapted between launches, as
 * adjust!(::Step) does); measure = 1 records one estimator block per launch. */
int pimc_run(pimc_ctx *ctx, int32_t n_sweeps, int32_t measure, pimc_counters *out);
int pimc_get_estimates(pimc_ctx *ctx, int32_t chain, pimc_estimates *out);
/* accept counters of one chain since pimc_create: {staging accepted, staging
 * tries, COM accepted, COM tries} (the Counter totals of src/updates/helper.jl:6-11);
 * chain = -1 fills 4 values for every chain, in chain order */
int pimc_get_counters(pimc_ctx *ctx, int32_t chain, int64_t *out);
/* per-block means of one chain: E, Ev, Epair per block (each n_blocks long) */
int pimc_get_blocks(pimc_ctx *ctx, int32_t chain, double *E, double *Ev, double *Epair, int64_t cap, int64_t *n_blocks);
int pimc_get_radial(pimc_ctx *ctx, int32_t chain, double *hist, int64_t *ndata);
int pimc_get_cartesian(pimc_ctx *ctx, int32_t chain, double *hist, int64_t *ndata);
int pimc_reset_estimators(pimc_ctx *ctx);

/* ---- one reference move or estimator on the device (parity hooks) ----
 * The four single moves below treat each particle as its own ring and return
 * PIMC_ERR_ARG for chains with swap moves (which sample through pimc_run). */

/* levy! + the ReshapeLinear weights for particle n, start slice j0 (0-based),
 * m links, with injected normals ((m-1) x dim, row-major per bead).  Writes
 * the (m+1) x dim bridge (column-major like r') and
 * w = {wV_initial, wV_updated, wU_initial, wU_updated}. */
int pimc_levy_propose(pimc_ctx *ctx, int32_t chain, int32_t n, int32_t j0, int32_t m, const double *normals,
                      double *r_out, double *w_out);
/* the same move followed by metropolis(exp(dlnW)) with uniform u; commits on accept */
int pimc_reshape_linear(pimc_ctx *ctx, int32_t chain, int32_t n, int32_t j0, int32_t m, const double *normals,
                        double u, int32_t *accepted);
/* SingleCenterOfMass with displacement d[dim]: M x dim moved worldline and weights */
int pimc_com_propose(pimc_ctx *ctx, int32_t chain, int32_t n, const double *d, double *r_out, double *w_out);
int pimc_center_of_mass(pimc_ctx *ctx, int32_t chain, int32_t n, const double *d, double u, int32_t *accepted);
/* interaction_action!(w, s, fpcycle, j): pair link sum of particle n at slice j */
int pimc_interaction(pimc_ctx *ctx, int32_t chain, int32_t n, int32_t j, double *w);
/* Energy(s) on the current paths: out = {E, Ev, Epair, Evpair} */
int pimc_energy(pimc_ctx *ctx, int32_t chain, double *out);
/* Density(s) on the current paths, accumulated into dens (nbins^dim,
 * column-major); *ndata += M */
int pimc_density(pimc_ctx *ctx, int32_t chain, int32_t nbins, double *dens, int64_t *ndata);

/* Density(s) as pimc_density, without allocating device memory.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight serving loops).  The
 * counts go to a fixed device array of the library (one per GPU, 2 MiB, kept until the process
 * ends) instead of a per-call allocation, so nbins^dim is at most 262144 (512 x 512 in 2D), else
 * PIMC_ERR_ARG.  Every call on a GPU shares that GPU's array: calls that overlap on one GPU (two
 * host threads, two contexts) race. */
int pimc_density_noalloc(pimc_ctx *ctx, int32_t chain, int32_t nbins, double *dens, int64_t *ndata);

/* Adjoint of the potential action S_pot = sum_links tau/2 (V + V') + S_pair:
 * dS/d(beads) (N x dim x M), dS/domega and dS/dg on the current paths.
 * The adjoint kernel recomputes bead distances instead of storing them. */
int pimc_adjoint(pimc_ctx *ctx, int32_t chain, double *dS_dr, double *dS_dparams, double *S_out);

/* potential action S_pot of one chain (for finite-difference checks) */
int pimc_action(pimc_ctx *ctx, int32_t chain, double *S_out);

#ifdef __cplusplus
}
#endif
#endif
