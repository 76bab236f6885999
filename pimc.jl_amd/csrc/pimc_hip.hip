// C ABI of the MI355X sampling path (declarations and reference citations in
// include/pimc_hip.h).  Host side: chain layout in HBM, tiles for the LDS
// sweep kernel, launch sequencing, estimator blocks, error codes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "hook_kernels.hpp"
#include "pimc_hip.h"
#include "propagator_kernels.hpp"
#include "sweep_kernels.hpp"
#include "window_kernels.hpp"

using namespace pimc;

// Histogram counts of pimc_density_noalloc: a fixed device array instead of a
// per-call allocation.  Each GPU holds its own copy until the process ends;
// host code uses the address hipGetSymbolAddress returns for the current
// device (kept per device in g_scratch), never the array's name, which on the
// host is only a stand-in for the symbol.
constexpr long long kScratchBins = 1ll << 18; // 2 MiB: nbins^dim up to 512 x 512
static __device__ __align__(256) unsigned long long scratch[kScratchBins];

namespace {

std::mutex g_scratch_mu;
constexpr int kMaxDevices = 256;
unsigned long long *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on device dev, which must be current;
// resolved once per device.
hipError_t scratch_address(int dev, unsigned long long **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = static_cast<unsigned long long *>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

__global__ void k_zero_u64(unsigned long long *p, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

struct HipError {
    hipError_t e;
    const char *what;
};

#define HIPCHK(call)                                                                                                   \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) throw HipError{e_, #call};                                                               \
    } while (0)

template <typename T> T *dalloc(size_t n)
{
    void *p = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc(&p, n * sizeof(T)));
    return reinterpret_cast<T *>(p);
}

void dfree(void *p)
{
    if (p) (void)hipFree(p);
}

struct Launch {
    hipEvent_t a, b;
};

// A run of consecutive tiles whose chains share one external potential (one launch).
struct TileGroup {
    int pot, maxl, first, count; // maxl: register slots of the bridge (8, 16 or 32)
};

} // namespace

struct pimc_ctx {
    int dev = 0;
    hipStream_t stream = nullptr;
    std::string err;
    pimc_options opt{};
    int prec = PIMC_FP32;
    int dim = 2;
    int n_chains = 0;
    std::vector<pimc_chain_desc> desc;
    std::vector<ChainDev> ch;
    std::vector<int> is_whole;   // chain -> 1 if run by the whole-chain kernel
    std::vector<int> whole_ids;  // whole chains
    std::vector<int> whole_cells_ids;  // whole chains with cell lists
    std::vector<int> whole_brute_ids;  // whole chains with brute-force neighbour sums
    std::vector<int> whole_lds_ids;    // ... whose path fits in LDS next to the kernel's buffers
    // Gaussian-pair chains on the window staging and centre-of-mass kernels, with and without
    // cell lists for the estimators (index 0: cells, 1: all-pairs); win_all: both
    std::vector<int> win_ids[2], win_all;
    int *d_win_all = nullptr;
    int win_seg = 0, win_maxN = 0, win_maxM = 0, win_maxell = 0, win_maxM0 = 0;
    int cap_stage = 0, cap_cspec = 0, cap_cdec = 0;   // pair-slot pools of the three kernels
    size_t lds_stage = 0, lds_cspec = 0, lds_cdec = 0, lds_cells = 0;
    int cells_wpg = 4, win_ncell = 0;                 // k_cells_win: waves per workgroup, largest cell count
    double *d_spec = nullptr;                         // centre of mass: Delta V + sum A per (particle, window)
    void *d_wbnd = nullptr;                           // window bounds (float4) per (particle, window)
    int spec_nwin = 1;
    unsigned char *d_cacc = nullptr;                  // centre of mass: accepted flag per particle
    void *d_cshift = nullptr;                         // and its shift (2 per particle, arithmetic type)
    void *d_pm = nullptr;                             // particle-major copy of the window chains' paths
    double *d_win_part = nullptr;                     // estimator partials of the window chains [chain][window][6]
    size_t lds_meas = 0;
    // The window chains in one group, or in two of about equal beads when there are two or more
    // and they fill at most half of the CUs: group g is win_all[wg_off[g] .. wg_off[g] + wg_n[g]);
    // group 1 runs on stream2, which each launch step forks from stream (ev_fork) and joins back
    // into it (ev_join).  Chains share no buffer, so the groups overlap freely: one group's
    // decision kernel (a workgroup per chain) runs beside the other group's staging and
    // speculative sums, half a sweep apart (ev_stagger).
    int n_wgroup = 0;
    int wg_off[2] = {0, 0}, wg_n[2] = {0, 0};
    std::vector<int> wg_cells[2];                     // the group's chains with cell lists
    int *d_wg_cells[2] = {nullptr, nullptr};
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // group 0's speculative sums of the current sweep done; group 1's sweep starts after it, so the
    // groups run half a sweep apart instead of in step (recorded and waited on once per sweep, in
    // that order, so one event serves every sweep)
    hipEvent_t ev_stagger = nullptr;
    int cap_meas = 0;
    std::vector<int> cells_all_ids;               // every chain with cell lists
    int *d_cells_all = nullptr;
    int n_meas_cells = 0;                         // meas_ids: the chains with cell lists come first
    // each of the three lists holds its Gaussian/pairless chains first, then from
    // this index the chains with the tabulated propagator (a kernel of their own)
    int cells_tab0 = 0, brute_tab0 = 0, lds_tab0 = 0;
    int cells_tperm0 = 0, brute_tperm0 = 0, lds_tperm0 = 0; // tabulated-propagator chains with swaps
    // measuring steps with the estimators of whole chains in their own wide kernel
    bool split_meas = false;
    int meas_nblk = 1;
    std::vector<int> meas_ids;      // whole chains: cell-list chains first, then all-pairs chains
    int *d_meas_ids = nullptr;
    double *d_meas_part = nullptr;  // [id][blk][6] partial sums
    std::vector<int> order;         // memory order of the chains
    std::vector<TileDev> tiles;
    std::vector<TileGroup> tile_groups;
    size_t tiles_cap = 0;           // TileDev slots allocated at d_tiles
    std::vector<int> part_chain, part_idx;
    long long n_pos = 0, n_cell_cnt = 0, n_bins = 0;
    int n_part = 0;
    int maxl = 8;
    int threads_tile = 256;
    size_t lds_tile = 0, lds_whole = 0;
    uint32_t sweep = 0;
    int nblk = 0, nblk_cap = 0;
    // device
    ChainDev *d_ch = nullptr;
    void *d_prm = nullptr;
    TileDev *d_tiles = nullptr;
    int *d_pchain = nullptr, *d_pidx = nullptr;
    void *d_pos = nullptr;
    double *d_part_acc = nullptr;
    int *d_placc = nullptr, *d_pcacc = nullptr;
    double *d_step = nullptr;
    long long *d_cnt = nullptr;
    double *d_chain_meas = nullptr;
    double *d_blocks = nullptr;
    unsigned long long *d_radial = nullptr, *d_cart = nullptr;
    int *d_whole = nullptr, *d_whole_cells = nullptr, *d_whole_brute = nullptr, *d_is_whole = nullptr;
    int *d_whole_lds = nullptr;
    int *d_next = nullptr; // per particle: the particle that follows it across slice M - 1
    int *d_cell_cnt = nullptr, *d_bins = nullptr;
    unsigned short *d_cell_mem = nullptr;
    int *d_status = nullptr;
    double *d_work = nullptr, *d_io = nullptr;
    size_t work_elems = 0, io_elems = 0;
    std::vector<Launch> events;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    long long cart_per_chain = 0;
    unsigned long long *d_scratch = nullptr;   // this device's `scratch` (pimc_density_noalloc)
    RunState *d_rs = nullptr;                   // device sweep/block counters
    hipGraphExec_t graph[2] = {nullptr, nullptr}; // captured launch step (thermalising, measuring)
    std::vector<double> hblocks; // host copy of the estimator blocks
    int hblk_n = -1;             // blocks held in hblocks (-1: stale)
    std::vector<double *> d_tab; // per chain: pair-product propagator table (PIMC_PAIR_TABLE chains)
    void *d_beams = nullptr;     // per chain: kMaxBeams (kx, ky, phase) triples of PIMC_POT_OPTICAL, arithmetic type
};

namespace {

int set_err(pimc_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

int hip_err(pimc_ctx *c, const HipError &h)
{
    return set_err(c, PIMC_ERR_HIP, std::string(h.what) + ": " + hipGetErrorString(h.e));
}

size_t tsize(const pimc_ctx *c) { return c->prec == PIMC_FP64 ? sizeof(double) : sizeof(float); }

int next_pow2_maxl(int ell)
{
    int m = 8;
    while (m < ell) m <<= 1;
    return m;
}

// ---- kernel dispatch over (T, DIM, MAXL) ----
template <typename T, int DIM, int MAXL, int POT> void launch_tiles_p(pimc_ctx *c, SweepArgs a, const TileGroup &g)
{
    auto k = k_sweep_tiles<T, DIM, MAXL, POT>;
    if (c->lds_tile > 65536) HIPCHK(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->lds_tile));
    a.tile0 = g.first;
    hipLaunchKernelGGL(k, dim3((unsigned)g.count), dim3(c->threads_tile), c->lds_tile, c->stream, a);
    HIPCHK(hipGetLastError());
}

template <typename T, int DIM, int MAXL> void launch_tiles_t(pimc_ctx *c, const SweepArgs &a, const TileGroup &g)
{
    if (g.pot == PIMC_POT_HARMONIC) launch_tiles_p<T, DIM, MAXL, kPotHarmonic>(c, a, g);
    else if (g.pot == PIMC_POT_LATTICE) launch_tiles_p<T, DIM, MAXL, kPotLattice>(c, a, g);
    else if (g.pot == PIMC_POT_OPTICAL) launch_tiles_p<T, DIM, MAXL, kPotOptical>(c, a, g);
    else launch_tiles_p<T, DIM, MAXL, kPotNone>(c, a, g);
}

template <typename T, int DIM> void launch_tiles_d(pimc_ctx *c, const SweepArgs &a)
{
    for (const TileGroup &g : c->tile_groups) {
        switch (g.maxl) {
        case 8: launch_tiles_t<T, DIM, 8>(c, a, g); break;
        case 16: launch_tiles_t<T, DIM, 16>(c, a, g); break;
        default: launch_tiles_t<T, DIM, 32>(c, a, g); break;
        }
    }
}

void launch_tiles(pimc_ctx *c, const SweepArgs &a)
{
    if (c->prec == PIMC_FP64) {
        if (c->dim == 1) launch_tiles_d<double, 1>(c, a);
        else launch_tiles_d<double, 2>(c, a);
    } else {
        if (c->dim == 1) launch_tiles_d<float, 1>(c, a);
        else launch_tiles_d<float, 2>(c, a);
    }
}

template <typename T, int DIM, bool CELLS, bool LDSR, bool TAB, bool TPERM = false>
void launch_whole_t(pimc_ctx *c, SweepArgs a, const std::vector<int> &all, int lo, int hi, const int *d_ids)
{
    if (hi <= lo) return;
    a.whole = d_ids + lo;
    int maxM = 0;
    size_t lds = c->lds_whole;
    for (int x = lo; x < hi; ++x) {
        const int id = all[x];
        maxM = std::max(maxM, c->ch[id].M);
        if (LDSR) lds = std::max(lds, whole_lds_bytes(c->ch[id], sizeof(T)) + whole_path_bytes(c->ch[id], sizeof(T)));
    }
    // lanes per link: as many as the workgroup holds for one pass over the links (2..16)
    const int nt = whole_threads<TAB>();
    int G = 1;
    while (G < 16 && 2 * G * maxM <= nt) G <<= 1;
    a.link_group = std::max(2, G);
    auto k = k_sweep_whole<T, DIM, CELLS, LDSR, TAB, TPERM>;
    if (lds > 65536) HIPCHK(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k, dim3((unsigned)(hi - lo)), dim3(nt), lds, c->stream, a);
    HIPCHK(hipGetLastError());
}

template <typename T, int DIM> void launch_whole_d(pimc_ctx *c, const SweepArgs &a)
{
    const auto &wc = c->whole_cells_ids, &wb = c->whole_brute_ids, &wl = c->whole_lds_ids;
    // per list: analytic pairs, then tabulated-propagator chains, then those with swap moves
    launch_whole_t<T, DIM, true, false, false>(c, a, wc, 0, c->cells_tab0, c->d_whole_cells);
    launch_whole_t<T, DIM, true, false, true>(c, a, wc, c->cells_tab0, c->cells_tperm0, c->d_whole_cells);
    launch_whole_t<T, DIM, true, false, true, true>(c, a, wc, c->cells_tperm0, (int)wc.size(), c->d_whole_cells);
    launch_whole_t<T, DIM, false, false, false>(c, a, wb, 0, c->brute_tab0, c->d_whole_brute);
    launch_whole_t<T, DIM, false, false, true>(c, a, wb, c->brute_tab0, c->brute_tperm0, c->d_whole_brute);
    launch_whole_t<T, DIM, false, false, true, true>(c, a, wb, c->brute_tperm0, (int)wb.size(), c->d_whole_brute);
    launch_whole_t<T, DIM, false, true, false>(c, a, wl, 0, c->lds_tab0, c->d_whole_lds);
    launch_whole_t<T, DIM, false, true, true>(c, a, wl, c->lds_tab0, c->lds_tperm0, c->d_whole_lds);
    launch_whole_t<T, DIM, false, true, true, true>(c, a, wl, c->lds_tperm0, (int)wl.size(), c->d_whole_lds);
}

void launch_whole(pimc_ctx *c, const SweepArgs &a)
{
    if (c->prec == PIMC_FP64) {
        if (c->dim == 1) launch_whole_d<double, 1>(c, a);
        else launch_whole_d<double, 2>(c, a);
    } else {
        if (c->dim == 1) launch_whole_d<float, 1>(c, a);
        else launch_whole_d<float, 2>(c, a);
    }
}

// Rebuild the cell lists of one chain (only >= 0) or of all cell-list chains.
void build_cells(pimc_ctx *c, int only)
{
    if (c->cells_all_ids.empty()) return;
    int maxM = 0;
    for (int id : c->cells_all_ids) maxM = std::max(maxM, c->ch[id].M);
    dim3 grid((maxM + 63) / 64, (unsigned)c->cells_all_ids.size());
#define BC(T, D)                                                                                                       \
    hipLaunchKernelGGL((k_build_cells<T, D>), grid, dim3(64), 0, c->stream, c->d_ch, c->d_cells_all,                   \
                       (int)c->cells_all_ids.size(), c->d_pos, c->d_cell_cnt, c->d_cell_mem, c->d_bins, c->d_status, only)
    if (c->prec == PIMC_FP64) {
        if (c->dim == 1) BC(double, 1);
        else BC(double, 2);
    } else {
        if (c->dim == 1) BC(float, 1);
        else BC(float, 2);
    }
#undef BC
    HIPCHK(hipGetLastError());
}

SweepArgs sweep_args(pimc_ctx *c)
{
    SweepArgs a{};
    a.chains = c->d_ch;
    a.prm = c->d_prm;
    a.tiles = c->d_tiles;
    a.part_chain = c->d_pchain;
    a.part_idx = c->d_pidx;
    a.next = c->d_next;
    a.pos = c->d_pos;
    a.part_acc = c->d_part_acc;
    a.part_lacc = c->d_placc;
    a.part_cacc = c->d_pcacc;
    a.com_step = c->d_step;
    a.radial = c->d_radial;
    a.cart = c->d_cart;
    a.radial_bins = c->opt.radial_bins;
    a.radial_winv = c->opt.radial_bins > 0 ? c->opt.radial_bins / c->opt.radial_rmax : 0.0;
    a.cart_bins = c->opt.cart_bins;
    a.seed = c->opt.seed;
    a.rs = c->d_rs;
    a.cell_cnt = c->d_cell_cnt;
    a.cell_mem = c->d_cell_mem;
    a.bins = c->d_bins;
    a.chain_meas = c->d_chain_meas;
    a.status = c->d_status;
    return a;
}

// Ensure room for `need` blocks; returns true if the buffer was reallocated.
bool grow_blocks(pimc_ctx *c, int need)
{
    if (need <= c->nblk_cap) return false;
    int cap = std::max(need, std::max(64, c->nblk_cap * 2));
    double *nb = dalloc<double>((size_t)cap * c->n_chains * kNQ);
    if (c->nblk > 0)
        HIPCHK(hipMemcpyAsync(nb, c->d_blocks, sizeof(double) * (size_t)c->nblk * c->n_chains * kNQ,
                              hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    dfree(c->d_blocks);
    c->d_blocks = nb;
    c->nblk_cap = cap;
    return true;
}

// Block b of one chain from the host copy, refreshed after each run.
// Host to device, ordered on the library stream: that stream is non-blocking, so a plain
// hipMemcpy (null stream) is not ordered before its kernels.  Returns when the copy is complete.
void h2d(pimc_ctx *c, void *dst, const void *src, size_t bytes)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
}

const double *host_block(pimc_ctx *c, int b, int chain)
{
    if (c->hblk_n != c->nblk) {
        HIPCHK(hipStreamSynchronize(c->stream));
        c->hblocks.resize((size_t)c->nblk * c->n_chains * kNQ);
        if (c->nblk > 0)
            HIPCHK(hipMemcpy(c->hblocks.data(), c->d_blocks, sizeof(double) * c->hblocks.size(), hipMemcpyDeviceToHost));
        c->hblk_n = c->nblk;
    }
    return &c->hblocks[((size_t)b * c->n_chains + chain) * kNQ];
}

int check_chain(pimc_ctx *c, int chain)
{
    if (!c) return PIMC_ERR_ARG;
    if (chain < 0 || chain >= c->n_chains) return set_err(c, PIMC_ERR_ARG, "chain index out of range");
    return PIMC_OK;
}

// A chain with the tabulated pair propagator samples only once its table is set.
int check_table(pimc_ctx *c, int chain)
{
    const int lo = chain < 0 ? 0 : chain, hi = chain < 0 ? c->n_chains : chain + 1;
    for (int i = lo; i < hi; ++i)
        if (c->ch[i].pair == kPairTable && !c->ch[i].tab)
            return set_err(c, PIMC_ERR_STATE, "chain " + std::to_string(i) +
                                                  ": the pair-product propagator table is not set (pimc_set_pair_table)");
    return PIMC_OK;
}

// Cell-list overflow flags are per chain and stay set until the chain's paths are set again.
int check_status(pimc_ctx *c)
{
    std::vector<int> st(c->n_chains);
    HIPCHK(hipMemcpy(st.data(), c->d_status, sizeof(int) * c->n_chains, hipMemcpyDeviceToHost));
    for (int i = 0; i < c->n_chains; ++i)
        if (st[i])
            return set_err(c, PIMC_ERR_CAPACITY, "chain " + std::to_string(i) +
                                                     ": a cell list overflowed its capacity; raise r_a or lower the density");
    return PIMC_OK;
}

void ensure_io(pimc_ctx *c, size_t n)
{
    if (n <= c->io_elems) return;
    dfree(c->d_io);
    c->d_io = dalloc<double>(n);
    c->io_elems = n;
}

HookArgs hook_args(pimc_ctx *c, int chain)
{
    HookArgs h{};
    h.chains = c->d_ch;
    h.pos = c->d_pos;
    h.cell_cnt = c->d_cell_cnt;
    h.cell_mem = c->d_cell_mem;
    h.bins = c->d_bins;
    h.status = c->d_status;
    h.chain = chain;
    h.work = c->d_work;
    h.use_cells = (c->ch[chain].ncell > 0) ? 1 : 0;
    h.next = c->d_next + c->ch[chain].part_base;
    return h;
}

// Launch a hook kernel templated on <T, DIM, CELLS, SM>: cell lists exist only
// for slice-major (pair) chains, so three layouts occur.
#define HOOK_T(KERNEL, T, D, grid, block, lds, ...)                                                                   \
    do {                                                                                                               \
        if (cells_) hipLaunchKernelGGL((KERNEL<T, D, true, true>), grid, block, lds, c->stream, __VA_ARGS__);          \
        else if (sm_) hipLaunchKernelGGL((KERNEL<T, D, false, true>), grid, block, lds, c->stream, __VA_ARGS__);       \
        else hipLaunchKernelGGL((KERNEL<T, D, false, false>), grid, block, lds, c->stream, __VA_ARGS__);               \
    } while (0)
#define HOOK_DISPATCH(KERNEL, grid, block, lds, ...)                                                                   \
    do {                                                                                                               \
        const bool cells_ = h.use_cells != 0;                                                                          \
        const bool sm_ = c->ch[h.chain].sm != 0;                                                                       \
        if (c->prec == PIMC_FP64) {                                                                                    \
            if (c->dim == 1) HOOK_T(KERNEL, double, 1, grid, block, lds, __VA_ARGS__);                                 \
            else HOOK_T(KERNEL, double, 2, grid, block, lds, __VA_ARGS__);                                             \
        } else {                                                                                                       \
            if (c->dim == 1) HOOK_T(KERNEL, float, 1, grid, block, lds, __VA_ARGS__);                                  \
            else HOOK_T(KERNEL, float, 2, grid, block, lds, __VA_ARGS__);                                              \
        }                                                                                                              \
        HIPCHK(hipGetLastError());                                                                                     \
    } while (0)

// The estimators of one sweep in their own kernels: the whole chains of the
// wide estimator kernel (k_meas_beads, with or without cell lists) and the
// window chains (k_meas_win), each folded into its chain's sums by k_meas_fold.
template <typename T, int DIM> void launch_meas_d(pimc_ctx *c, int sweep_off)
{
    MeasArgs m{};
    m.chains = c->d_ch;
    m.prm = c->d_prm;
    m.nblk = c->meas_nblk;
    m.pos = c->d_pos;
    m.cell_cnt = c->d_cell_cnt;
    m.cell_mem = c->d_cell_mem;
    m.bins = c->d_bins;
    m.next = c->d_next;
    m.radial = c->d_radial;
    m.cart = c->d_cart;
    m.radial_bins = c->opt.radial_bins;
    m.radial_winv = c->opt.radial_bins > 0 ? c->opt.radial_bins / c->opt.radial_rmax : 0.0;
    m.cart_bins = c->opt.cart_bins;
    m.chain_meas = c->d_chain_meas;
    m.rs = c->d_rs;
    m.sweep_off = sweep_off;
    m.measure_every = c->opt.measure_every;
    const int ncell = c->n_meas_cells, nall = (int)c->meas_ids.size();
    if (ncell > 0) {
        MeasArgs mc = m;
        mc.ids = c->d_meas_ids;
        mc.n_ids = ncell;
        mc.part = c->d_meas_part;
        hipLaunchKernelGGL((k_meas_beads<T, DIM, true>), dim3((unsigned)c->meas_nblk, (unsigned)ncell), dim3(256), 0,
                           c->stream, mc);
        HIPCHK(hipGetLastError());
    }
    if (nall > ncell) {
        MeasArgs mb = m;
        mb.ids = c->d_meas_ids + ncell;
        mb.n_ids = nall - ncell;
        mb.part = c->d_meas_part + (size_t)ncell * c->meas_nblk * 6;
        hipLaunchKernelGGL((k_meas_beads<T, DIM, false>), dim3((unsigned)c->meas_nblk, (unsigned)(nall - ncell)),
                           dim3(256), 0, c->stream, mb);
        HIPCHK(hipGetLastError());
    }
    if (nall > 0) {
        m.ids = c->d_meas_ids;
        m.n_ids = nall;
        m.part = c->d_meas_part;
        hipLaunchKernelGGL(k_meas_fold, dim3((unsigned)((nall + 127) / 128)), dim3(128), 0, c->stream, m);
        HIPCHK(hipGetLastError());
    }
}

void launch_meas(pimc_ctx *c, int sweep_off)
{
    if (c->prec == PIMC_FP64) {
        if (c->dim == 1) launch_meas_d<double, 1>(c, sweep_off);
        else launch_meas_d<double, 2>(c, sweep_off);
    } else {
        if (c->dim == 1) launch_meas_d<float, 1>(c, sweep_off);
        else launch_meas_d<float, 2>(c, sweep_off);
    }
}

// The stream of window group g.
hipStream_t win_stream(const pimc_ctx *c, int g) { return g == 0 ? c->stream : c->stream2; }

// The estimators of one sweep of window group g: by centre-of-mass windows, then the fold in
// window order (the partials of the group's chains start at its offset in win_all).
template <typename T, int DIM> void launch_meas_win_d(pimc_ctx *c, int sweep_off, int g)
{
    const int nw = c->wg_n[g];
    if (nw == 0) return;
    MeasArgs mw{};
    mw.chains = c->d_ch;
    mw.prm = c->d_prm;
    mw.pos = c->d_pos;
    mw.cell_cnt = c->d_cell_cnt;
    mw.cell_mem = c->d_cell_mem;
    mw.bins = c->d_bins;
    mw.next = c->d_next;
    mw.radial = c->d_radial;
    mw.cart = c->d_cart;
    mw.radial_bins = c->opt.radial_bins;
    mw.radial_winv = c->opt.radial_bins > 0 ? c->opt.radial_bins / c->opt.radial_rmax : 0.0;
    mw.cart_bins = c->opt.cart_bins;
    mw.chain_meas = c->d_chain_meas;
    mw.rs = c->d_rs;
    mw.sweep_off = sweep_off;
    mw.measure_every = c->opt.measure_every;
    mw.ids = c->d_win_all + c->wg_off[g];
    mw.n_ids = nw;
    mw.nblk = c->spec_nwin;
    mw.part = c->d_win_part + (size_t)c->wg_off[g] * c->spec_nwin * 6;
    const hipStream_t st = win_stream(c, g);
    hipLaunchKernelGGL((k_meas_win<T, DIM>), dim3((unsigned)c->spec_nwin, (unsigned)nw), dim3(512), c->lds_meas, st,
                       mw, c->spec_nwin, c->cap_meas);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_meas_fold, dim3((unsigned)((nw + 127) / 128)), dim3(128), 0, st, mw);
    HIPCHK(hipGetLastError());
}

void launch_meas_win(pimc_ctx *c, int sweep_off, int g)
{
    if (c->prec == PIMC_FP64) {
        if (c->dim == 1) launch_meas_win_d<double, 1>(c, sweep_off, g);
        else launch_meas_win_d<double, 2>(c, sweep_off, g);
    } else {
        if (c->dim == 1) launch_meas_win_d<float, 1>(c, sweep_off, g);
        else launch_meas_win_d<float, 2>(c, sweep_off, g);
    }
}

// One sweep of window group g on the window kernels: the staging windows, then the
// centre-of-mass speculation, decisions and moves.
template <typename T, int DIM> void launch_win_d(pimc_ctx *c, const SweepArgs &b, int g, hipEvent_t stagger)
{
    const int n = c->wg_n[g];
    if (n == 0) return;
    SweepArgs w = b;
    w.whole = c->d_win_all + c->wg_off[g];
    const hipStream_t st = win_stream(c, g);
    if (stagger && g == 1) HIPCHK(hipStreamWaitEvent(st, stagger, 0));
    hipLaunchKernelGGL((k_stage_win<T, DIM>), dim3((unsigned)c->win_seg, (unsigned)n), dim3(512), c->lds_stage, st,
                       w, c->cap_stage);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL((k_com_spec<T, DIM>), dim3((unsigned)c->spec_nwin, (unsigned)n), dim3(512), c->lds_cspec, st,
                       w, c->d_spec, static_cast<float4 *>(c->d_wbnd), c->spec_nwin, c->cap_cspec,
                       static_cast<T *>(c->d_pm));
    HIPCHK(hipGetLastError());
    if (stagger && g == 0) HIPCHK(hipEventRecord(stagger, st));
    hipLaunchKernelGGL((k_com_dec<T, DIM>), dim3((unsigned)n), dim3(1024), c->lds_cdec, st, w,
                       (const double *)c->d_spec, static_cast<const float4 *>(c->d_wbnd), c->spec_nwin, c->cap_cdec,
                       kChunk, c->d_cacc, static_cast<T *>(c->d_cshift), static_cast<const T *>(c->d_pm));
    HIPCHK(hipGetLastError());
    const long long beads = (long long)c->win_maxM * c->win_maxN;
    hipLaunchKernelGGL((k_com_apply<T, DIM>), dim3((unsigned)((beads + 255) / 256), (unsigned)n), dim3(256), 0, st, w,
                       (const unsigned char *)c->d_cacc, static_cast<const T *>(c->d_cshift));
    HIPCHK(hipGetLastError());
}

void launch_win(pimc_ctx *c, const SweepArgs &b, int g)
{
    // with two groups, group 0 records the sweep's event and group 1 (launched next) waits on it
    hipEvent_t stagger = c->n_wgroup == 2 ? c->ev_stagger : nullptr;
    if (c->prec == PIMC_FP64) {
        if (c->dim == 1) launch_win_d<double, 1>(c, b, g, stagger);
        else launch_win_d<double, 2>(c, b, g, stagger);
    } else {
        if (c->dim == 1) launch_win_d<float, 1>(c, b, g, stagger);
        else launch_win_d<float, 2>(c, b, g, stagger);
    }
}

// Cell lists of window group g's chains that have them, rebuilt from the positions.
template <typename T, int DIM> void launch_cells_win_d(pimc_ctx *c, int g)
{
    const int n = (int)c->wg_cells[g].size();
    if (n == 0) return;
    const int wpg = c->cells_wpg;
    hipLaunchKernelGGL((k_cells_win<T, DIM>), dim3((unsigned)((c->win_maxM0 + wpg - 1) / wpg), (unsigned)n),
                       dim3(64 * wpg), c->lds_cells, win_stream(c, g), c->d_ch, c->d_prm, c->d_wg_cells[g], n,
                       c->d_pos, c->d_cell_cnt, c->d_cell_mem, c->d_bins, c->d_status, c->win_ncell);
    HIPCHK(hipGetLastError());
}

void launch_cells_win(pimc_ctx *c, int g)
{
    if (c->prec == PIMC_FP64) {
        if (c->dim == 1) launch_cells_win_d<double, 1>(c, g);
        else launch_cells_win_d<double, 2>(c, g);
    } else {
        if (c->dim == 1) launch_cells_win_d<float, 1>(c, g);
        else launch_cells_win_d<float, 2>(c, g);
    }
}

// Dynamic LDS above 64 KiB for the window kernels of this context's type.
template <typename T, int DIM> void win_attrs_d(pimc_ctx *c)
{
    auto set = [](const void *k, size_t lds) {
        if (lds > 65536) HIPCHK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    };
    set((const void *)k_stage_win<T, DIM>, c->lds_stage);
    set((const void *)k_com_spec<T, DIM>, c->lds_cspec);
    set((const void *)k_com_dec<T, DIM>, c->lds_cdec);
    set((const void *)k_cells_win<T, DIM>, c->lds_cells);
    set((const void *)k_meas_win<T, DIM>, c->lds_meas);
}

void win_attrs(pimc_ctx *c)
{
    if (c->win_all.empty()) return;
    if (c->prec == PIMC_FP64) {
        if (c->dim == 1) win_attrs_d<double, 1>(c);
        else win_attrs_d<double, 2>(c);
    } else {
        if (c->dim == 1) win_attrs_d<float, 1>(c);
        else win_attrs_d<float, 2>(c);
    }
}

void launch_step(pimc_ctx *c, int k, int measure)
{
    SweepArgs a = sweep_args(c);
    a.nsweeps = k;
    a.measure = measure ? 1 : 0;
    a.measure_every = c->opt.measure_every;
    const bool win = !c->win_all.empty();
    const int ng = win ? c->n_wgroup : 0;
    if (ng == 2) {
        // window group 1 starts after the previous launch step (sweep word advanced)
        HIPCHK(hipEventRecord(c->ev_fork, c->stream));
        HIPCHK(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
    }
    if (!c->tiles.empty()) launch_tiles(c, a);
    // the window kernels add their counts into the launch's per-chain outputs
    for (int g = 0; g < ng; ++g)
        if (c->wg_n[g] > 0) {
            hipLaunchKernelGGL(k_clear_chain_meas, dim3((unsigned)c->wg_n[g]), dim3(32), 0, win_stream(c, g),
                               c->d_chain_meas, c->d_win_all + c->wg_off[g], c->wg_n[g]);
            HIPCHK(hipGetLastError());
        }
    if (!c->whole_ids.empty() || win) {
        const bool old = !c->whole_ids.empty();
        if (measure && (c->split_meas || win)) {
            // one sweep per launch, estimators in their own wide kernel (the chains on the
            // whole-chain kernel with in-kernel estimators run all k sweeps in one launch)
            if (old && !c->split_meas) launch_whole(c, a);
            for (int s = 0; s < k; ++s) {
                SweepArgs b = a;
                b.nsweeps = 1;
                b.sweep_off = s;
                b.split = 1;
                if (old && c->split_meas) launch_whole(c, b);
                for (int g = 0; g < ng; ++g) launch_win(c, b, g);
                launch_meas(c, s);
                for (int g = 0; g < ng; ++g) launch_meas_win(c, s, g);
            }
        } else {
            if (old) launch_whole(c, a);
            for (int s = 0; win && s < k; ++s) {
                SweepArgs b = a;
                b.nsweeps = 1;
                b.sweep_off = s;
                b.split = 1;
                for (int g = 0; g < ng; ++g) launch_win(c, b, g);
            }
        }
        for (int g = 0; g < ng; ++g) launch_cells_win(c, g);
    }
    if (ng == 2) {
        // every launch step ends on stream: k_post reads the window chains' outputs
        HIPCHK(hipEventRecord(c->ev_join, c->stream2));
        HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    }
    PostArgs pa{};
    pa.chains = c->d_ch;
    pa.n_chains = c->n_chains;
    pa.is_whole = c->d_is_whole;
    pa.part_acc = c->d_part_acc;
    pa.part_lacc = c->d_placc;
    pa.part_cacc = c->d_pcacc;
    pa.chain_meas = c->d_chain_meas;
    pa.com_step = c->d_step;
    pa.counters = c->d_cnt;
    pa.blocks = c->d_blocks;
    pa.rs = c->d_rs;
    pa.measure = measure ? 1 : 0;
    pa.nsweeps = k;
    pa.adapt_ell = c->opt.adapt_ell;
    hipLaunchKernelGGL(k_post, dim3((c->n_chains + 127) / 128), dim3(128), 0, c->stream, pa);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, c->stream, c->d_rs, k, measure ? 1 : 0);
    HIPCHK(hipGetLastError());
}

void drop_graphs(pimc_ctx *c)
{
    for (auto &g : c->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

int validate(const pimc_chain_desc &d, std::string &why)
{
    if (d.dim != 1 && d.dim != 2) return why = "dim must be 1 or 2", -1;
    if (d.N < 1 || d.N > 65535) return why = "N must be in [1, 65535]", -1;
    if (d.M < 4) return why = "M must be >= 4", -1;
    if (!(d.beta > 0) || !(d.lambda > 0) || !(d.L > 0)) return why = "beta, lambda and L must be positive", -1;
    if (d.potential < PIMC_POT_NONE || d.potential > PIMC_POT_OPTICAL) return why = "unknown potential", -1;
    if (d.potential == PIMC_POT_OPTICAL) {
        if (d.dim != 2) return why = "the optical lattice potential is two-dimensional", -1;
        if (d.n_beams < 1 || d.n_beams > PIMC_MAX_BEAMS) return why = "n_beams must be in [1, 16]", -1;
        if (!(d.beam_scale > 0)) return why = "beam_scale must be positive", -1;
    }
    if (d.pair < PIMC_PAIR_NONE || d.pair > PIMC_PAIR_TABLE) return why = "unknown pair potential", -1;
    if (d.pair == PIMC_PAIR_GAUSSIAN && !(d.sigma > 0)) return why = "sigma must be positive", -1;
    if (d.pair == PIMC_PAIR_TABLE && !(d.r_a > 0))
        return why = "a tabulated pair propagator needs the neighbour range r_a > 0", -1;
    if (!(d.hs_a >= 0)) return why = "hs_a must be >= 0", -1;
    if (d.hs_a > 0) {
        if (d.pair != PIMC_PAIR_TABLE) return why = "hard spheres (hs_a > 0) come with the tabulated pair propagator", -1;
        const double ra = d.r_a > 0 ? d.r_a : d.L / 4.0;
        const int nb = std::max(1, (int)std::floor((2.0 * d.L) / ra));
        if (d.hs_a > (2.0 * d.L) / nb) return why = "hs_a must not exceed the cell side 2L / nbins", -1;
    }
    if (d.potential == PIMC_POT_LATTICE && !(d.a > 0)) return why = "lattice constant a must be positive", -1;
    if (d.ell != 0 && (d.ell < 2 || d.ell > 32 || d.ell > d.M)) return why = "ell must be in [2, min(32, M)]", -1;
    if (!(d.com_step >= 0)) return why = "com_step must be >= 0", -1;
    if (d.swaps < 0) return why = "swaps must be >= 0", -1;
    for (int e : {d.staging_every, d.com_every, d.swap_every})
        if (e < 0 || e > 65536) return why = "staging_every, com_every and swap_every must be in [0, 65536]", -1;
    if (d.swaps > 0) {
        if (d.N > 4096) return why = "swap moves need N <= 4096", -1;
        const int sm = d.swap_m ? d.swap_m : (d.ell ? d.ell : std::min(8, d.M));
        if (sm < 2 || sm > d.M - 1) return why = "swap_m must be in [2, M - 1]", -1;
        if (d.swap_rule != 0 && d.swap_rule != 1) return why = "swap_rule must be 0 or 1", -1;
    }
    return 0;
}

// Nodes and weights of the k-integral of the pair-product propagator's terms:
// 16-point Gauss-Legendre panels, 40 geometric panels [h 2^-(p+1), h 2^-p]
// below h = min(1/2, 4/rmax) for the logarithmic behaviour near k = 0, then
// uniform panels of width <= h up to K = sqrt(40/tau), where exp(-tau K^2) =
// 4e-18; (r1 + r2) * width <= 8 keeps the Bessel products resolved.
void prop_quad_nodes(double rmax, double tau, std::vector<double> &k, std::vector<double> &w)
{
    constexpr int ng = 16, npanel_geo = 40;
    double gx[ng], gw[ng];
    // Legendre roots by Newton's iteration on P_16 (three-term recurrence)
    for (int i = 0; i < ng; ++i) {
        double z = std::cos(3.14159265358979323846 * (i + 0.75) / (ng + 0.5)), dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int l = 2; l <= ng; ++l) {
                const double p2 = ((2.0 * l - 1.0) * z * p1 - (l - 1.0) * p0) / l;
                p0 = p1;
                p1 = p2;
            }
            dp = ng * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = z;
        for (int l = 2; l <= ng; ++l) {
            const double p2 = ((2.0 * l - 1.0) * z * p1 - (l - 1.0) * p0) / l;
            p0 = p1;
            p1 = p2;
        }
        dp = ng * (z * p1 - p0) / (z * z - 1.0);
        gx[i] = z;
        gw[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
    auto panel = [&](double a, double b) {
        for (int i = 0; i < ng; ++i) {
            k.push_back(0.5 * (b - a) * gx[i] + 0.5 * (b + a));
            w.push_back(0.5 * (b - a) * gw[i]);
        }
    };
    const double h = std::min(0.5, 4.0 / rmax);
    k.clear();
    w.clear();
    for (int p = npanel_geo - 1; p >= 0; --p) panel(h * std::ldexp(1.0, -(p + 1)), h * std::ldexp(1.0, -p));
    const double K = std::max(2.0 * h, std::sqrt(40.0 / tau));
    const int np = (int)std::ceil((K - h) / h);
    const double hw = (K - h) / np;
    for (int p = 0; p < np; ++p) panel(h + p * hw, h + (p + 1) * hw);
}

template <int DIM, int MAXL>
void launch_tile_fast_t(pimc_ctx *c, const HookArgs &h, int pot, size_t lds)
{
    switch (pot) {
    case PIMC_POT_HARMONIC: hipLaunchKernelGGL((k_hook_tile_fast<DIM, MAXL, kPotHarmonic>), dim3(1), dim3(64), lds, c->stream, h); break;
    case PIMC_POT_LATTICE: hipLaunchKernelGGL((k_hook_tile_fast<DIM, MAXL, kPotLattice>), dim3(1), dim3(64), lds, c->stream, h); break;
    case PIMC_POT_OPTICAL: hipLaunchKernelGGL((k_hook_tile_fast<DIM, MAXL, kPotOptical>), dim3(1), dim3(64), lds, c->stream, h); break;
    default: hipLaunchKernelGGL((k_hook_tile_fast<DIM, MAXL, kPotNone>), dim3(1), dim3(64), lds, c->stream, h); break;
    }
    HIPCHK(hipGetLastError());
}

void launch_tile_fast(pimc_ctx *c, const HookArgs &h, int maxl, int pot, size_t lds)
{
    if (c->dim == 1) {
        if (maxl == 8) launch_tile_fast_t<1, 8>(c, h, pot, lds);
        else if (maxl == 16) launch_tile_fast_t<1, 16>(c, h, pot, lds);
        else launch_tile_fast_t<1, 32>(c, h, pot, lds);
    } else {
        if (maxl == 8) launch_tile_fast_t<2, 8>(c, h, pot, lds);
        else if (maxl == 16) launch_tile_fast_t<2, 16>(c, h, pot, lds);
        else launch_tile_fast_t<2, 32>(c, h, pot, lds);
    }
}

} // namespace

// ============================================================================
extern "C" {

// message of the last failed call that takes no context (pimc_prop_int_terms), per thread
static thread_local std::string g_err;

const char *pimc_last_error(const pimc_ctx *ctx)
{
    if (ctx) return ctx->err.c_str();
    return g_err.empty() ? "null context" : g_err.c_str();
}

int pimc_device_info(char *name, int32_t len, int32_t *cu_count)
{
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess) return PIMC_ERR_HIP;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return PIMC_ERR_HIP;
    if (name && len > 0) {
        std::snprintf(name, (size_t)len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    return PIMC_OK;
}

// Tiles of the LDS kernel over the tile chains in memory order: consecutive particles of chains
// with one potential, MAXL, M, ell and pair of move weights share a tile, within the LDS share of 4
// workgroups per CU; tiles form launch groups by (potential, MAXL).  At create time, and again
// when a tile chain's staging length changes (adapt_ell).
void build_tiles(pimc_ctx *c)
{
    const int ts = (int)tsize(c);
    // built aside and kept only when every check passes, so a failure leaves the context as it was
    std::vector<TileDev> tiles;
    std::vector<TileGroup> tile_groups;
    const int threads_tile = c->opt.block_threads > 0 ? c->opt.block_threads : 256;
    if (threads_tile % 64 != 0 || threads_tile > kTileMaxThreads)
        throw HipError{hipErrorInvalidValue, "block_threads must be a multiple of 64, <= 256"};
    int tile_beads = c->opt.tile_beads > 0 ? c->opt.tile_beads : (ts == 4 ? 4096 : 2048);
    const int max_items = kItemsPerThread * threads_tile;
    size_t lds_max = 0;
    for (int id : c->order) {
        if (c->is_whole[id]) continue;
        const ChainDev &h = c->ch[id];
        if (h.nseg > max_items) throw HipError{hipErrorInvalidValue, "a particle has more segments than a workgroup holds; raise block_threads or ell"};
        for (int p = 0; p < h.N; ++p) {
            const int gq = h.part_base + p;
            bool fresh = tiles.empty();
            if (!fresh) {
                TileDev &t = tiles.back();
                fresh = t.M != h.M || t.ell != h.ell || t.every_st != h.every_st || t.every_com != h.every_com ||
                        tile_groups.back().pot != h.pot ||
                        tile_groups.back().maxl != next_pow2_maxl(h.ell) ||
                        (t.np + 1) * h.M > tile_beads || (t.np + 1) * h.nseg > max_items;
                if (!fresh) {
                    // keep every tile within the LDS share of 4 workgroups per CU
                    const int slots = t.nslot + (p == 0 ? 1 : 0);
                    fresh = tile_lds_bytes(t.np + 1, c->dim, t.pst, (t.np + 1) * h.nseg, ts, slots,
                                           c->opt.radial_bins) > kTileLdsBudget;
                }
            }
            if (fresh) {
                TileDev t{};
                t.pos_off = h.pos_off + (long long)p * h.dim * h.M;
                t.q0 = gq;
                t.np = 0;
                t.M = h.M;
                t.ell = h.ell;
                t.nseg = h.nseg;
                t.every_st = h.every_st;
                t.every_com = h.every_com;
                t.sh = 0;
                while ((2 << t.sh) <= t.ell) ++t.sh; // floor(log2(ell))
                t.pst = tile_particle_stride(t.M, t.ell, t.sh, c->dim);
                t.nslot = 0;
                {
                    const int maxl = next_pow2_maxl(h.ell);
                    t.grouped = ts == 4 && h.ell == maxl && h.M % maxl == 0 && h.nseg <= 64 &&
                                (h.nseg & (h.nseg - 1)) == 0;
                }
                if (tile_groups.empty() || tile_groups.back().pot != h.pot ||
                    tile_groups.back().maxl != next_pow2_maxl(h.ell))
                    tile_groups.push_back(TileGroup{h.pot, next_pow2_maxl(h.ell), (int)tiles.size(), 0});
                tile_groups.back().count += 1;
                tiles.push_back(t);
            }
            TileDev &t = tiles.back();
            if (t.np == 0 || p == 0) t.nslot += 1; // chains are contiguous in the particle order
            t.np += 1;
            t.items = t.np * t.nseg;
        }
    }
    for (TileDev &t : tiles) {
        int nslot = 1;
        for (int q = 1; q < t.np; ++q)
            if (c->part_chain[t.q0 + q] != c->part_chain[t.q0 + q - 1]) ++nslot;
        if (nslot != t.nslot) throw HipError{hipErrorUnknown, "tile chain-slot count mismatch"};
        lds_max = std::max(lds_max, tile_lds_bytes(t.np, c->dim, t.pst, t.items, ts, t.nslot, c->opt.radial_bins));
    }
    if (lds_max > 160 * 1024) throw HipError{hipErrorInvalidValue, "tile exceeds LDS; lower tile_beads"};
    c->tiles.swap(tiles);
    c->tile_groups.swap(tile_groups);
    c->threads_tile = threads_tile;
    c->lds_tile = lds_max;
}

// After a thermalising launch under adapt_ell: the staging lengths k_post chose for the tile
// chains.  Where one changed, the host copy follows and the tiles are rebuilt before the next
// launch (captured launch sequences hold the old tiles, so they are dropped).
void retile_after_adapt(pimc_ctx *c)
{
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<ChainDev> dev(c->n_chains);
    HIPCHK(hipMemcpy(dev.data(), c->d_ch, sizeof(ChainDev) * c->n_chains, hipMemcpyDeviceToHost));
    std::vector<int> changed;
    for (int i = 0; i < c->n_chains; ++i)
        if (!c->is_whole[i] && dev[i].ell != c->ch[i].ell) changed.push_back(i);
    if (changed.empty()) return;
    std::vector<ChainDev> before(c->ch);
    for (int i : changed) {
        c->ch[i].ell = dev[i].ell;
        c->ch[i].nseg = dev[i].nseg;
    }
    try {
        build_tiles(c);
    } catch (const HipError &) {
        // a length whose tiles the workgroup cannot hold (more segments than its items): the
        // chains keep their lengths, on the host and on the device, and the tiles stay
        c->ch.swap(before);
        for (int i : changed)
            h2d(c, c->d_ch + i, &c->ch[i], sizeof(ChainDev));
        return;
    }
    if (c->tiles.size() > c->tiles_cap) {
        HIPCHK(hipFree(c->d_tiles));
        c->d_tiles = nullptr;
        c->d_tiles = dalloc<TileDev>(c->tiles.size());
        c->tiles_cap = c->tiles.size();
    }
    h2d(c, c->d_tiles, c->tiles.data(), sizeof(TileDev) * c->tiles.size());
    drop_graphs(c);
}

int pimc_create(pimc_ctx **out, const pimc_chain_desc *chains, int32_t n_chains, const pimc_options *opt)
{
    if (!out || !chains || n_chains < 1 || !opt) return PIMC_ERR_ARG;
    *out = nullptr;
    pimc_ctx *c = new (std::nothrow) pimc_ctx();
    if (!c) return PIMC_ERR_NOMEM;
    std::string why;
    for (int i = 0; i < n_chains; ++i) {
        if (validate(chains[i], why) != 0) {
            c->err = "chain " + std::to_string(i) + ": " + why;
            *out = c;
            return PIMC_ERR_ARG;
        }
        if (chains[i].dim != chains[0].dim) {
            c->err = "all chains of a context share one dimension";
            *out = c;
            return PIMC_ERR_ARG;
        }
    }
    if (opt->precision != PIMC_FP32 && opt->precision != PIMC_FP64) {
        c->err = "precision must be PIMC_FP32 or PIMC_FP64";
        *out = c;
        return PIMC_ERR_ARG;
    }
    if (opt->radial_bins < 0 || opt->cart_bins < 0 || (opt->radial_bins > 0 && !(opt->radial_rmax > 0))) {
        c->err = "bad histogram options";
        *out = c;
        return PIMC_ERR_ARG;
    }
    if (opt->derivatives) {
        for (int i = 0; i < n_chains; ++i)
            if (chains[i].pair == PIMC_PAIR_TABLE) {
                c->err = "chain " + std::to_string(i) + ": derivatives need a Gaussian pair or none";
                *out = c;
                return PIMC_ERR_ARG;
            }
    }
    if (opt->pair_kernel < 0 || opt->pair_kernel > 2) {
        c->err = "pair_kernel must be 0, 1 or 2";
        *out = c;
        return PIMC_ERR_ARG;
    }
    if (opt->measure_kernel < 0 || opt->measure_kernel > 2) {
        c->err = "measure_kernel must be 0, 1 or 2";
        *out = c;
        return PIMC_ERR_ARG;
    }
    {
        const int K = opt->sweeps_per_launch > 0 ? opt->sweeps_per_launch : 50;
        const int me = opt->measure_every > 0 ? opt->measure_every : 1;
        if (K % me != 0) {
            c->err = "sweeps_per_launch must be a multiple of measure_every";
            *out = c;
            return PIMC_ERR_ARG;
        }
    }
    *out = c;
    try {
        c->opt = *opt;
        if (c->opt.sweeps_per_launch <= 0) c->opt.sweeps_per_launch = 50;
        if (c->opt.measure_every <= 0) c->opt.measure_every = 1;
        c->prec = opt->precision;
        c->dim = chains[0].dim;
        c->n_chains = n_chains;
        c->desc.assign(chains, chains + n_chains);
        HIPCHK(hipGetDevice(&c->dev));
        HIPCHK(scratch_address(c->dev, &c->d_scratch));
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&c->ev0));
        HIPCHK(hipEventCreate(&c->ev1));
        int cu = 256;
        {
            hipDeviceProp_t p;
            HIPCHK(hipGetDeviceProperties(&p, c->dev));
            cu = p.multiProcessorCount;
        }

        // ---- chain table; whole-chain kernel for pair chains (and all chains with derivatives)
        c->ch.resize(n_chains);
        c->d_tab.assign(n_chains, nullptr);
        c->is_whole.assign(n_chains, 0);
        int ellmax = 2;
        for (int i = 0; i < n_chains; ++i) {
            const pimc_chain_desc &d = chains[i];
            ChainDev &h = c->ch[i];
            std::memset(&h, 0, sizeof(h));
            h.beta = d.beta;
            h.tau = d.beta / d.M;
            h.lambda = d.lambda;
            h.L = d.L;
            h.omega = d.omega;
            h.V0 = d.V0;
            h.a = d.a > 0 ? d.a : 1.0;
            h.g = d.g;
            h.sigma = d.sigma > 0 ? d.sigma : 1.0;
            h.N = d.N;
            h.M = d.M;
            h.dim = d.dim;
            h.pot = d.potential;
            h.pair = d.pair;
            h.pair_in_accept = d.pair_in_accept;
            h.ell = d.ell ? d.ell : std::min(8, d.M);
            h.nseg = (d.M + h.ell - 1) / h.ell;
            const double ra = d.r_a > 0 ? d.r_a : d.L / 4.0;                // src/system.jl:22-24
            h.nbins = std::max(1, (int)std::floor((2.0 * d.L) / ra));     // src/system.jl:81
            h.h = (2.0 * d.L) / h.nbins;                                  // src/nearest_neighbours.jl:127
            h.swaps = d.swaps;
            h.swap_m = d.swaps > 0 ? (d.swap_m ? d.swap_m : h.ell) : 0;
            h.swap_rule = d.swap_rule;
            h.hs_a = d.hs_a;
            h.nbeam = d.potential == PIMC_POT_OPTICAL ? d.n_beams : 0;
            h.every_st = std::max(1, d.staging_every);
            h.every_com = std::max(1, d.com_every);
            h.every_sw = std::max(1, d.swap_every);
            ellmax = std::max(ellmax, h.ell);
            c->is_whole[i] = (d.pair != PIMC_PAIR_NONE || opt->derivatives || d.swaps > 0) ? 1 : 0;
        }
        c->maxl = next_pow2_maxl(ellmax);
        if (c->maxl > 32) throw HipError{hipErrorInvalidValue, "ell > 32"};
        // optical-lattice beams in revolutions: t_b = scale (x sin a_b + y cos a_b) + [helical] a_b / 2 pi
        {
            const size_t per = 3 * kMaxBeams, n = (size_t)n_chains * per;
            std::vector<double> bd(n, 0.0);
            for (int i = 0; i < n_chains; ++i) {
                const pimc_chain_desc &d = chains[i];
                for (int b = 0; b < c->ch[i].nbeam; ++b) {
                    bd[i * per + 3 * b] = d.beam_scale * std::sin(d.beam_angle[b]);
                    bd[i * per + 3 * b + 1] = d.beam_scale * std::cos(d.beam_angle[b]);
                    bd[i * per + 3 * b + 2] = d.helical ? d.beam_angle[b] / (2.0 * 3.14159265358979323846) : 0.0;
                }
            }
            if (c->prec == PIMC_FP64) {
                double *p = dalloc<double>(n);
                c->d_beams = p;
                HIPCHK(hipMemcpy(p, bd.data(), sizeof(double) * n, hipMemcpyHostToDevice));
            } else {
                std::vector<float> bf(bd.begin(), bd.end());
                float *p = dalloc<float>(n);
                c->d_beams = p;
                HIPCHK(hipMemcpy(p, bf.data(), sizeof(float) * n, hipMemcpyHostToDevice));
            }
            for (int i = 0; i < n_chains; ++i)
                c->ch[i].beam = static_cast<const char *>(c->d_beams) + (size_t)i * per * tsize(c);
        }

        // ---- memory order: tile chains sorted by (M, ell), then whole chains
        std::vector<int> &order = c->order;
        order.resize(n_chains);
        for (int i = 0; i < n_chains; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            if (c->is_whole[x] != c->is_whole[y]) return c->is_whole[x] < c->is_whole[y];
            if (c->ch[x].pot != c->ch[y].pot) return c->ch[x].pot < c->ch[y].pot;
            const int lx = next_pow2_maxl(c->ch[x].ell), ly = next_pow2_maxl(c->ch[y].ell);
            if (lx != ly) return lx < ly;
            if (c->ch[x].M != c->ch[y].M) return c->ch[x].M > c->ch[y].M;
            if (c->ch[x].ell != c->ch[y].ell) return c->ch[x].ell < c->ch[y].ell;
            // a tile holds chains of one pair of move weights
            if (c->ch[x].every_st != c->ch[y].every_st) return c->ch[x].every_st < c->ch[y].every_st;
            return c->ch[x].every_com < c->ch[y].every_com;
        });
        long long off = 0, celloff = 0, binoff = 0;
        int pbase = 0;
        for (int id : order) {
            ChainDev &h = c->ch[id];
            h.pos_off = off;
            off += (long long)h.N * h.dim * h.M;
            h.part_base = pbase;
            pbase += h.N;
            if (c->is_whole[id]) {
                h.sm = 1; // slice-major: the beads of one slice together
                // cell lists when they beat the all-pairs scan and stay small
                const long long ncell = (h.dim == 1) ? h.nbins : (long long)h.nbins * h.nbins;
                const double mean = (double)h.N / (double)ncell;
                int cap = (int)std::min<long long>(h.N, (long long)std::ceil(4.0 * mean + 16.0));
                cap = (cap + 7) & ~7; // whole 16-byte chunks per member list
                const long long bytes = (long long)h.M * ncell * (cap * 2 + 4);
                const bool cells = h.pair != PIMC_PAIR_NONE && h.N > 96 && h.nbins >= 3 && bytes < (256ll << 20);
                // Gaussian-pair chains (the pair action in the acceptance; no swaps, hard spheres or
                // staging-length adaptation) whose window state, with every window chain so far, fits
                // in LDS run on the window kernels
                const int tsz = (int)tsize(c);
                bool win = h.pair == kPairGaussian && h.pair_in_accept && h.swaps == 0 && !(h.hs_a > 0.0) &&
                           !opt->adapt_ell &&
                           h.N <= 4096 && h.M >= 3 && (!cells || ncell <= 16384) &&
                           (opt->pair_kernel == 2 || (opt->pair_kernel == 0 && cells));
                if (win) {
                    const int N = std::max(c->win_maxN, h.N), ell = std::max(c->win_maxell, h.ell);
                    const int lc = std::max(com_win_len(std::max(c->win_maxM, h.M)), 8);
                    win = stage_lds(N, ell, h.dim, tsz, N).ck.end <= 160 * 1024 &&
                          com_spec_lds(N, lc, h.dim, tsz, N).ck.end <= 160 * 1024 &&
                          com_dec_lds(N, h.dim, tsz, N).ck.end <= 160 * 1024;
                }
                if (cells) {
                    h.ncell = (int)ncell;
                    h.cap = cap;
                    h.cell_off = celloff;
                    celloff += (long long)h.M * ncell;
                    h.bin_off = binoff;
                    binoff += (long long)h.N * h.M;
                    c->cells_all_ids.push_back(id);
                } else {
                    h.ncell = 0;
                    h.cap = 0;
                }
                if (win) {
                    const int v = cells ? 0 : 1;
                    c->win_ids[v].push_back(id);
                    c->win_all.push_back(id);
                    c->spec_nwin = std::max(c->spec_nwin, com_nwin(h.M));
                    c->win_seg = std::max(c->win_seg, h.nseg);
                    c->win_maxN = std::max(c->win_maxN, h.N);
                    c->win_maxM = std::max(c->win_maxM, h.M);
                    c->win_maxell = std::max(c->win_maxell, h.ell);
                    if (cells) {
                        c->win_maxM0 = std::max(c->win_maxM0, h.M);
                        c->win_ncell = std::max(c->win_ncell, (int)ncell);
                    }
                } else {
                    c->whole_ids.push_back(id);
                    if (cells) c->whole_cells_ids.push_back(id);
                    else c->whole_brute_ids.push_back(id);
                }
            }
        }
        c->n_pos = off;
        c->n_part = pbase;
        c->n_cell_cnt = celloff;
        c->n_bins = binoff;

        // ---- particle table and tiles of the LDS kernel
        c->part_chain.assign(c->n_part, 0);
        c->part_idx.assign(c->n_part, 0);
        for (int i = 0; i < n_chains; ++i)
            for (int p = 0; p < c->ch[i].N; ++p) {
                c->part_chain[c->ch[i].part_base + p] = i;
                c->part_idx[c->ch[i].part_base + p] = p;
            }
        const int ts = (int)tsize(c);
        build_tiles(c);
        // whole-chain kernel
        int maxMw = 0;
        for (int id : c->whole_ids) maxMw = std::max(maxMw, c->ch[id].M);
        if (maxMw > 0) {
            c->lds_whole = 0;
            for (int id : c->whole_ids) c->lds_whole = std::max(c->lds_whole, whole_lds_bytes(c->ch[id], ts));
            if (c->lds_whole > 160 * 1024) throw HipError{hipErrorInvalidValue, "M too large for the whole-chain kernel"};
        }
        // all-pairs chains whose positions fit in LDS next to the buffers keep them there
        {
            std::vector<int> keep;
            for (int id : c->whole_brute_ids) {
                const ChainDev &h = c->ch[id];
                if (whole_lds_bytes(h, ts) + whole_path_bytes(h, ts) <= 160 * 1024) c->whole_lds_ids.push_back(id);
                else keep.push_back(id);
            }
            c->whole_brute_ids.swap(keep);
        }
        // tabulated-propagator chains last in each list (their own kernel instantiation)
        {
            auto part = [&](std::vector<int> &v) {
                return (int)(std::stable_partition(v.begin(), v.end(), [&](int id) { return c->ch[id].pair != kPairTable; }) -
                             v.begin());
            };
            c->cells_tab0 = part(c->whole_cells_ids);
            c->brute_tab0 = part(c->whole_brute_ids);
            c->lds_tab0 = part(c->whole_lds_ids);
            // ... and among those, the chains with swap moves last (the TPERM instantiation)
            auto part_perm = [&](std::vector<int> &v, int lo) {
                return (int)(std::stable_partition(v.begin() + lo, v.end(), [&](int id) { return c->ch[id].swaps == 0; }) -
                             v.begin());
            };
            c->cells_tperm0 = part_perm(c->whole_cells_ids, c->cells_tab0);
            c->brute_tperm0 = part_perm(c->whole_brute_ids, c->brute_tab0);
            c->lds_tperm0 = part_perm(c->whole_lds_ids, c->lds_tab0);
        }
        // estimators of whole chains: in the sweep kernel, or (split) a wide kernel after each sweep;
        // automatic: split when a whole chain with a pair action has >= 8192 beads
        {
            bool big_pair = false;
            for (int id : c->whole_ids) {
                const long long nb = (long long)c->ch[id].N * c->ch[id].M;
                big_pair = big_pair || (c->ch[id].pair != kPairNone && nb >= 8192);
            }
            const int mk = c->opt.measure_kernel;
            c->split_meas = !c->whole_ids.empty() && (mk == 2 || (mk == 0 && big_pair));
            // the wide estimator kernel: chains on the window kernels always, the others when split;
            // chains with cell lists first
            c->meas_ids.clear();
            if (c->split_meas) c->meas_ids = c->whole_cells_ids;
            c->n_meas_cells = (int)c->meas_ids.size();
            if (c->split_meas) {
                c->meas_ids.insert(c->meas_ids.end(), c->whole_brute_ids.begin(), c->whole_brute_ids.end());
                c->meas_ids.insert(c->meas_ids.end(), c->whole_lds_ids.begin(), c->whole_lds_ids.end());
            }
            long long most = 0;
            for (int id : c->meas_ids) most = std::max(most, (long long)c->ch[id].N * c->ch[id].M);
            c->meas_nblk = (int)std::max(1LL, (most + 1023) / 1024);
        }
        // ---- device allocations
        c->d_ch = dalloc<ChainDev>(n_chains);
        HIPCHK(hipMemcpy(c->d_ch, c->ch.data(), sizeof(ChainDev) * n_chains, hipMemcpyHostToDevice));
        if (c->prec == PIMC_FP64) {
            std::vector<Prm<double>> pr(n_chains);
            for (int i = 0; i < n_chains; ++i) pr[i] = Prm<double>::make(c->ch[i]);
            c->d_prm = dalloc<Prm<double>>(n_chains);
            HIPCHK(hipMemcpy(c->d_prm, pr.data(), sizeof(Prm<double>) * n_chains, hipMemcpyHostToDevice));
        } else {
            std::vector<Prm<float>> pr(n_chains);
            for (int i = 0; i < n_chains; ++i) pr[i] = Prm<float>::make(c->ch[i]);
            c->d_prm = dalloc<Prm<float>>(n_chains);
            HIPCHK(hipMemcpy(c->d_prm, pr.data(), sizeof(Prm<float>) * n_chains, hipMemcpyHostToDevice));
        }
        c->d_tiles = dalloc<TileDev>(c->tiles.size());
        c->tiles_cap = c->tiles.size();
        if (!c->tiles.empty())
            HIPCHK(hipMemcpy(c->d_tiles, c->tiles.data(), sizeof(TileDev) * c->tiles.size(), hipMemcpyHostToDevice));
        c->d_pchain = dalloc<int>(c->n_part);
        c->d_pidx = dalloc<int>(c->n_part);
        HIPCHK(hipMemcpy(c->d_pchain, c->part_chain.data(), sizeof(int) * c->n_part, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_pidx, c->part_idx.data(), sizeof(int) * c->n_part, hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&c->d_pos, std::max<size_t>(1, (size_t)c->n_pos * ts)));
        HIPCHK(hipMemset(c->d_pos, 0, std::max<size_t>(1, (size_t)c->n_pos * ts)));
        c->d_part_acc = dalloc<double>((size_t)c->n_part * 4);
        c->d_placc = dalloc<int>(c->n_part);
        c->d_pcacc = dalloc<int>(c->n_part);
        HIPCHK(hipMemset(c->d_placc, 0, sizeof(int) * c->n_part));
        HIPCHK(hipMemset(c->d_pcacc, 0, sizeof(int) * c->n_part));
        c->d_step = dalloc<double>(n_chains);
        {
            std::vector<double> st(n_chains);
            for (int i = 0; i < n_chains; ++i) st[i] = chains[i].com_step;
            HIPCHK(hipMemcpy(c->d_step, st.data(), sizeof(double) * n_chains, hipMemcpyHostToDevice));
        }
        {
            std::vector<int> ident(c->n_part);
            for (int i = 0; i < n_chains; ++i)
                for (int p = 0; p < c->ch[i].N; ++p) ident[(size_t)c->ch[i].part_base + p] = p;
            c->d_next = dalloc<int>(c->n_part);
            HIPCHK(hipMemcpy(c->d_next, ident.data(), sizeof(int) * c->n_part, hipMemcpyHostToDevice));
        }
        c->d_cnt = dalloc<long long>((size_t)n_chains * kCnt);
        HIPCHK(hipMemset(c->d_cnt, 0, sizeof(long long) * n_chains * kCnt));
        c->d_chain_meas = dalloc<double>((size_t)n_chains * (kNQ + kCnt));
        HIPCHK(hipMemset(c->d_chain_meas, 0, sizeof(double) * n_chains * (kNQ + kCnt)));
        c->d_is_whole = dalloc<int>(n_chains);
        HIPCHK(hipMemcpy(c->d_is_whole, c->is_whole.data(), sizeof(int) * n_chains, hipMemcpyHostToDevice));
        c->d_whole = dalloc<int>(c->whole_ids.size());
        c->d_whole_cells = dalloc<int>(c->whole_cells_ids.size());
        c->d_whole_brute = dalloc<int>(c->whole_brute_ids.size());
        c->d_whole_lds = dalloc<int>(c->whole_lds_ids.size());
        c->d_meas_ids = dalloc<int>(c->meas_ids.size());
        if (!c->meas_ids.empty())
            HIPCHK(hipMemcpy(c->d_meas_ids, c->meas_ids.data(), sizeof(int) * c->meas_ids.size(), hipMemcpyHostToDevice));
        c->d_meas_part = dalloc<double>(!c->meas_ids.empty() ? c->meas_ids.size() * (size_t)c->meas_nblk * 6 : 1);
        c->d_win_all = dalloc<int>(c->win_all.size());
        if (!c->win_all.empty()) {
            HIPCHK(hipMemcpy(c->d_win_all, c->win_all.data(), sizeof(int) * c->win_all.size(), hipMemcpyHostToDevice));
            // pools: at least N slots (one particle's partners), at most a whole chunk's, and
            // otherwise what keeps the staging kernel within 64 KiB (two 512-thread workgroups
            // per CU), the phase-1 kernel within 48 KiB (three) and the decision kernel (one
            // workgroup per chain) within 128 KiB; measured on gauss_box and lattice
            const int N = c->win_maxN, ts_ = (int)ts, d = c->dim;
            const int lc = std::max(com_win_len(c->win_maxM), 8);
            auto grow = [&](auto bytes, size_t limit) {
                int cap = std::min(8192, kChunk * std::max(1, N - 1));
                while (cap > N && bytes(cap) > limit) cap -= 64;
                return std::max(cap, N);
            };
            c->cap_stage = grow([&](int k) { return stage_lds(N, c->win_maxell, d, ts_, k).ck.end; }, 64 * 1024);
            c->cap_cspec = grow([&](int k) { return com_spec_lds(N, lc, d, ts_, k).ck.end; }, 48 * 1024);
            c->cap_cdec = grow([&](int k) { return com_dec_lds(N, d, ts_, k).ck.end; }, 128 * 1024);
            c->cap_meas = grow([&](int k) { return meas_win_lds(N, lc, d, ts_, k).ck.end; }, 48 * 1024);
            c->lds_meas = meas_win_lds(N, lc, d, ts_, c->cap_meas).ck.end;
            c->d_win_part = dalloc<double>(c->win_all.size() * (size_t)c->spec_nwin * 6);
            c->lds_stage = stage_lds(N, c->win_maxell, d, ts_, c->cap_stage).ck.end;
            c->lds_cspec = com_spec_lds(N, lc, d, ts_, c->cap_cspec).ck.end;
            c->lds_cdec = com_dec_lds(N, d, ts_, c->cap_cdec).ck.end;
            c->cells_wpg = 4;
            while (c->cells_wpg > 1 && (size_t)c->cells_wpg * c->win_ncell * 4 > 64 * 1024) c->cells_wpg >>= 1;
            c->lds_cells = (size_t)c->cells_wpg * std::max(c->win_ncell, 1) * 4;
            c->d_spec = dalloc<double>((size_t)c->n_part * c->spec_nwin);
            HIPCHK(hipMalloc(&c->d_wbnd, (size_t)c->n_part * c->spec_nwin * sizeof(float4)));
            c->d_cacc = dalloc<unsigned char>((size_t)c->n_part);
            HIPCHK(hipMalloc(&c->d_cshift, (size_t)c->n_part * 2 * ts));
            HIPCHK(hipMalloc(&c->d_pm, std::max<size_t>(1, (size_t)c->n_pos * ts)));
            win_attrs(c);
            // groups: the shortest prefix of win_all holding at least half of the beads, when the
            // window chains fill at most half of the CUs (one decision workgroup a chain: with
            // more, the decision kernels already occupy the chip and a second group only contends
            // for it; measured on lattice, 64 chains, and gauss_box, 256)
            const int nwc = (int)c->win_all.size();
            const bool two = nwc >= 2 && 2 * nwc <= cu;
            int split = nwc;
            if (two) {
                long long tot = 0, acc = 0;
                for (int id : c->win_all) tot += (long long)c->ch[id].N * c->ch[id].M;
                split = nwc - 1;
                for (int i = 0; i < nwc - 1; ++i) {
                    acc += (long long)c->ch[c->win_all[i]].N * c->ch[c->win_all[i]].M;
                    if (2 * acc >= tot) {
                        split = i + 1;
                        break;
                    }
                }
            }
            c->n_wgroup = two ? 2 : 1;
            c->wg_off[0] = 0;
            c->wg_n[0] = split;
            c->wg_off[1] = split;
            c->wg_n[1] = nwc - split;
            std::vector<char> has_cells(n_chains, 0);
            for (int id : c->win_ids[0]) has_cells[id] = 1;
            for (int g = 0; g < c->n_wgroup; ++g) {
                for (int i = c->wg_off[g]; i < c->wg_off[g] + c->wg_n[g]; ++i)
                    if (has_cells[c->win_all[i]]) c->wg_cells[g].push_back(c->win_all[i]);
                c->d_wg_cells[g] = dalloc<int>(c->wg_cells[g].size());
                if (!c->wg_cells[g].empty())
                    HIPCHK(hipMemcpy(c->d_wg_cells[g], c->wg_cells[g].data(), sizeof(int) * c->wg_cells[g].size(),
                                     hipMemcpyHostToDevice));
            }
            if (c->n_wgroup == 2) {
                HIPCHK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
                HIPCHK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
                HIPCHK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
                HIPCHK(hipEventCreateWithFlags(&c->ev_stagger, hipEventDisableTiming));
            }
        }
        c->d_cells_all = dalloc<int>(c->cells_all_ids.size());
        if (!c->cells_all_ids.empty())
            HIPCHK(hipMemcpy(c->d_cells_all, c->cells_all_ids.data(), sizeof(int) * c->cells_all_ids.size(), hipMemcpyHostToDevice));
        if (!c->whole_lds_ids.empty())
            HIPCHK(hipMemcpy(c->d_whole_lds, c->whole_lds_ids.data(), sizeof(int) * c->whole_lds_ids.size(), hipMemcpyHostToDevice));
        if (!c->whole_ids.empty())
            HIPCHK(hipMemcpy(c->d_whole, c->whole_ids.data(), sizeof(int) * c->whole_ids.size(), hipMemcpyHostToDevice));
        if (!c->whole_cells_ids.empty())
            HIPCHK(hipMemcpy(c->d_whole_cells, c->whole_cells_ids.data(), sizeof(int) * c->whole_cells_ids.size(), hipMemcpyHostToDevice));
        if (!c->whole_brute_ids.empty())
            HIPCHK(hipMemcpy(c->d_whole_brute, c->whole_brute_ids.data(), sizeof(int) * c->whole_brute_ids.size(), hipMemcpyHostToDevice));
        long long mem_slots = 0;
        for (int id : c->cells_all_ids) {
            ChainDev &h = c->ch[id];
            h.mem_off = mem_slots;
            mem_slots += (long long)h.M * h.ncell * h.cap;
        }
        HIPCHK(hipMemcpy(c->d_ch, c->ch.data(), sizeof(ChainDev) * n_chains, hipMemcpyHostToDevice));
        c->d_cell_cnt = dalloc<int>((size_t)c->n_cell_cnt);
        c->d_cell_mem = dalloc<unsigned short>((size_t)mem_slots);
        c->d_bins = dalloc<int>((size_t)c->n_bins);
        c->d_rs = dalloc<RunState>(1);
        HIPCHK(hipMemset(c->d_rs, 0, sizeof(RunState)));
        c->d_status = dalloc<int>(n_chains);
        HIPCHK(hipMemset(c->d_status, 0, sizeof(int) * n_chains));
        int maxM = 0;
        for (int i = 0; i < n_chains; ++i) maxM = std::max(maxM, c->ch[i].M);
        c->work_elems = (size_t)(maxM + 1) * c->dim + 16;
        c->d_work = dalloc<double>(c->work_elems);
        if (opt->radial_bins > 0) {
            c->d_radial = dalloc<unsigned long long>((size_t)n_chains * opt->radial_bins);
            HIPCHK(hipMemset(c->d_radial, 0, sizeof(unsigned long long) * n_chains * opt->radial_bins));
        }
        if (opt->cart_bins > 0) {
            c->cart_per_chain = c->dim == 1 ? opt->cart_bins : (long long)opt->cart_bins * opt->cart_bins;
            c->d_cart = dalloc<unsigned long long>((size_t)n_chains * c->cart_per_chain);
            HIPCHK(hipMemset(c->d_cart, 0, sizeof(unsigned long long) * n_chains * c->cart_per_chain));
        }
        grow_blocks(c, 64);
        HIPCHK(hipDeviceSynchronize());
    } catch (const HipError &h) {
        return hip_err(c, h);
    } catch (const std::bad_alloc &) {
        return set_err(c, PIMC_ERR_NOMEM, "host allocation failed");
    }
    return PIMC_OK;
}

int pimc_destroy(pimc_ctx *c)
{
    if (!c) return PIMC_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    dfree(c->d_rs);
    void *ptrs[] = {c->d_ch, c->d_prm, c->d_tiles, c->d_pchain, c->d_pidx, c->d_pos, c->d_part_acc, c->d_placc, c->d_pcacc,
                    c->d_step, c->d_cnt, c->d_chain_meas, c->d_blocks, c->d_radial, c->d_cart, c->d_whole,
                    c->d_whole_cells, c->d_whole_brute, c->d_is_whole, c->d_cell_cnt, c->d_bins, c->d_cell_mem,
                    c->d_status, c->d_work, c->d_io, c->d_whole_lds, c->d_next, c->d_meas_ids, c->d_meas_part,
                    c->d_cells_all, c->d_spec, c->d_wbnd, c->d_win_all, c->d_cacc, c->d_cshift, c->d_pm, c->d_win_part,
                    c->d_beams, c->d_wg_cells[0], c->d_wg_cells[1]};
    for (void *p : ptrs) dfree(p);
    for (double *p : c->d_tab) dfree(p);
    for (auto &e : c->events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev_stagger) (void)hipEventDestroy(c->ev_stagger);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PIMC_OK;
}

int pimc_set_paths(pimc_ctx *c, int32_t chain, const double *r)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    if (!r) return set_err(c, PIMC_ERR_ARG, "null path buffer");
    const ChainDev &h = c->ch[chain];
    const size_t n = (size_t)h.N * h.dim * h.M;
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(r[i]) || r[i] < -h.L || r[i] >= h.L)
            return set_err(c, PIMC_ERR_ARG, "bead outside the box [-L, L) or not finite");
    try {
        // caller layout N x dim x M (Particle.r blocks) -> device layout of the chain
        auto dst_index = [&](size_t n_, int d_, int j_) -> size_t {
            return h.sm ? ((size_t)j_ * h.N + n_) * h.dim + d_ : ((size_t)n_ * h.dim + d_) * h.M + j_;
        };
        if (c->prec == PIMC_FP64) {
            std::vector<double> f(n);
            for (int p = 0; p < h.N; ++p)
                for (int d = 0; d < h.dim; ++d)
                    for (int j = 0; j < h.M; ++j) f[dst_index(p, d, j)] = r[((size_t)p * h.dim + d) * h.M + j];
            h2d(c, (double *)c->d_pos + h.pos_off, f.data(), n * sizeof(double));
        } else {
            std::vector<float> f(n);
            for (int p = 0; p < h.N; ++p)
                for (int d = 0; d < h.dim; ++d)
                    for (int j = 0; j < h.M; ++j) {
                        float v = (float)r[((size_t)p * h.dim + d) * h.M + j];
                        if (v >= (float)h.L) v = std::nextafter((float)h.L, -INFINITY);
                        f[dst_index(p, d, j)] = v;
                    }
            h2d(c, (float *)c->d_pos + h.pos_off, f.data(), n * sizeof(float));
        }
        if (h.ncell > 0) {
            // on the library stream, ahead of the cell build (not in a captured graph)
            HIPCHK(hipMemsetAsync(c->d_status + chain, 0, sizeof(int), c->stream));
            build_cells(c, chain);
            HIPCHK(hipStreamSynchronize(c->stream));
            return check_status(c);
        }
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_set_next(pimc_ctx *c, int32_t chain, const int32_t *next)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    if (!next) return set_err(c, PIMC_ERR_ARG, "null next buffer");
    const ChainDev &h = c->ch[chain];
    std::vector<char> seen(h.N, 0);
    bool ident = true;
    for (int p = 0; p < h.N; ++p) {
        if (next[p] < 0 || next[p] >= h.N || seen[next[p]]) return set_err(c, PIMC_ERR_ARG, "next is not a permutation");
        seen[next[p]] = 1;
        ident = ident && next[p] == p;
    }
    if (!ident && h.swaps == 0) return set_err(c, PIMC_ERR_ARG, "permutation cycles need a chain with swap moves");
    try {
        HIPCHK(hipStreamSynchronize(c->stream));
        h2d(c, c->d_next + h.part_base, next, sizeof(int) * h.N);
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_get_next(pimc_ctx *c, int32_t chain, int32_t *next)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    if (!next) return set_err(c, PIMC_ERR_ARG, "null next buffer");
    try {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(next, c->d_next + c->ch[chain].part_base, sizeof(int) * c->ch[chain].N, hipMemcpyDeviceToHost));
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_init_paths(pimc_ctx *c, int32_t chain, uint64_t seed)
{
    if (!c) return PIMC_ERR_ARG;
    if (chain != -1) {
        int rc = check_chain(c, chain);
        if (rc) return rc;
    }
    try {
        int q0 = 0, nq = (int)c->part_chain.size();
        if (chain >= 0) {
            q0 = c->ch[chain].part_base;
            nq = c->ch[chain].N;
        }
        const int threads = 128, blocks = (nq + threads - 1) / threads;
#define INIT(T, D)                                                                                                     \
    hipLaunchKernelGGL((k_init_paths<T, D>), dim3((unsigned)blocks), dim3(threads), 0, c->stream, c->d_ch, c->d_pchain, \
                       c->d_pidx, q0, nq, c->d_pos, seed)
        if (c->prec == PIMC_FP64) {
            if (c->dim == 1) INIT(double, 1);
            else INIT(double, 2);
        } else {
            if (c->dim == 1) INIT(float, 1);
            else INIT(float, 2);
        }
#undef INIT
        HIPCHK(hipGetLastError());
        // chains with hard spheres: particles placed in order, redrawn while they overlap
        std::vector<int> hs;
        for (int id = 0; id < c->n_chains; ++id)
            if ((chain < 0 || id == chain) && c->ch[id].hs_a > 0.0) hs.push_back(id);
        if (!hs.empty()) {
            int maxM = 0;
            for (int id : hs) maxM = std::max(maxM, c->ch[id].M);
            int *d_ids = dalloc<int>(hs.size());
            double *d_hs64 = dalloc<double>((size_t)std::max<long long>(1, c->n_pos));
            HIPCHK(hipMemcpyAsync(d_ids, hs.data(), sizeof(int) * hs.size(), hipMemcpyHostToDevice, c->stream));
            for (int id : hs) HIPCHK(hipMemsetAsync(c->d_status + id, 0, sizeof(int), c->stream));
            const size_t lds = (size_t)maxM * c->dim * sizeof(double) + 16;
            if (lds > 160 * 1024) throw HipError{hipErrorInvalidValue, "M too large for the hard-sphere init_world"};
#define INITHS(T, D)                                                                                                   \
    do {                                                                                                               \
        if (lds > 65536)                                                                                               \
            HIPCHK(hipFuncSetAttribute((const void *)k_init_paths_hs<T, D>,                                            \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));                         \
        hipLaunchKernelGGL((k_init_paths_hs<T, D>), dim3((unsigned)hs.size()), dim3(256), lds, c->stream, c->d_ch,     \
                           d_ids, (int)hs.size(), c->d_pos, d_hs64, seed, c->d_status);                                \
    } while (0)
            if (c->prec == PIMC_FP64) {
                if (c->dim == 1) INITHS(double, 1);
                else INITHS(double, 2);
            } else {
                if (c->dim == 1) INITHS(float, 1);
                else INITHS(float, 2);
            }
#undef INITHS
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(c->stream));
            std::vector<int> st(c->n_chains);
            HIPCHK(hipMemcpy(st.data(), c->d_status, sizeof(int) * c->n_chains, hipMemcpyDeviceToHost));
            dfree(d_ids);
            dfree(d_hs64);
            for (int id : hs)
                if (st[id] & 4)
                    return set_err(c, PIMC_ERR_CAPACITY, "chain " + std::to_string(id) +
                                                             ": init_world placed no particle clear of the hard "
                                                             "spheres within 10000 tries; lower hs_a or the density");
        }
        if (chain >= 0) {
            if (c->ch[chain].ncell > 0) {
                HIPCHK(hipMemsetAsync(c->d_status + chain, 0, sizeof(int), c->stream));
                build_cells(c, chain);
            }
        } else if (!c->cells_all_ids.empty()) {
            HIPCHK(hipMemsetAsync(c->d_status, 0, sizeof(int) * c->n_chains, c->stream));
            build_cells(c, -1);
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        return check_status(c);
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
}

int pimc_get_paths(pimc_ctx *c, int32_t chain, double *r)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    if (!r) return set_err(c, PIMC_ERR_ARG, "null path buffer");
    const ChainDev &h = c->ch[chain];
    const size_t n = (size_t)h.N * h.dim * h.M;
    try {
        HIPCHK(hipStreamSynchronize(c->stream));
        auto src_index = [&](size_t n_, int d_, int j_) -> size_t {
            return h.sm ? ((size_t)j_ * h.N + n_) * h.dim + d_ : ((size_t)n_ * h.dim + d_) * h.M + j_;
        };
        if (c->prec == PIMC_FP64) {
            std::vector<double> f(n);
            HIPCHK(hipMemcpy(f.data(), (double *)c->d_pos + h.pos_off, n * sizeof(double), hipMemcpyDeviceToHost));
            for (int p = 0; p < h.N; ++p)
                for (int d = 0; d < h.dim; ++d)
                    for (int j = 0; j < h.M; ++j) r[((size_t)p * h.dim + d) * h.M + j] = f[src_index(p, d, j)];
        } else {
            std::vector<float> f(n);
            HIPCHK(hipMemcpy(f.data(), (float *)c->d_pos + h.pos_off, n * sizeof(float), hipMemcpyDeviceToHost));
            for (int p = 0; p < h.N; ++p)
                for (int d = 0; d < h.dim; ++d)
                    for (int j = 0; j < h.M; ++j) r[((size_t)p * h.dim + d) * h.M + j] = f[src_index(p, d, j)];
        }
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_get_com_step(pimc_ctx *c, int32_t chain, double *step)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    try {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(step, c->d_step + chain, sizeof(double), hipMemcpyDeviceToHost));
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_set_com_step(pimc_ctx *c, int32_t chain, double step)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    if (!(step >= 0)) return set_err(c, PIMC_ERR_ARG, "step must be >= 0");
    try {
        HIPCHK(hipStreamSynchronize(c->stream));
        h2d(c, c->d_step + chain, &step, sizeof(double));
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_get_ell(pimc_ctx *c, int32_t chain, int32_t *ell)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    if (!ell) return set_err(c, PIMC_ERR_ARG, "ell is null");
    try {
        HIPCHK(hipStreamSynchronize(c->stream));
        int v = 0;
        HIPCHK(hipMemcpy(&v, reinterpret_cast<const char *>(c->d_ch + chain) + offsetof(ChainDev, ell), sizeof(int),
                         hipMemcpyDeviceToHost));
        *ell = v;
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_set_pair_table(pimc_ctx *c, int32_t chain, int32_t n, double r0, double r1, double mu,
                        const double *terms)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    ChainDev &h = c->ch[chain];
    if (h.pair != kPairTable) return set_err(c, PIMC_ERR_ARG, "the chain's pair kind is not PIMC_PAIR_TABLE");
    if (n < 2 || !(r1 > r0) || !(r0 >= 0) || !terms) return set_err(c, PIMC_ERR_ARG, "need n >= 2, 0 <= r0 < r1, terms");
    for (size_t i = 0; i < (size_t)n * n; ++i)
        if (!std::isfinite(terms[i])) return set_err(c, PIMC_ERR_ARG, "non-finite table entry");
    try {
        HIPCHK(hipStreamSynchronize(c->stream));
        double *t = dalloc<double>((size_t)n * n);
        if (hipMemcpyAsync(t, terms, sizeof(double) * n * n, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            dfree(t);
            throw HipError{hipErrorInvalidValue, "hipMemcpy(pair table)"};
        }
        dfree(c->d_tab[chain]);
        c->d_tab[chain] = t;
        h.tab = t;
        h.tab_n = n;
        h.tab_r0 = r0;
        h.tab_h = (r1 - r0) / (n - 1);
        h.mu = mu;
        // only the table fields: the device copy's staging length may have been adapted (adapt_ell)
        char *dc = reinterpret_cast<char *>(c->d_ch + chain);
        h2d(c, dc + offsetof(ChainDev, tab), &h.tab, sizeof(h.tab));
        h2d(c, dc + offsetof(ChainDev, tab_n), &h.tab_n, sizeof(h.tab_n));
        h2d(c, dc + offsetof(ChainDev, tab_r0), &h.tab_r0, sizeof(h.tab_r0));
        h2d(c, dc + offsetof(ChainDev, tab_h), &h.tab_h, sizeof(h.tab_h));
        h2d(c, dc + offsetof(ChainDev, mu), &h.mu, sizeof(h.mu));
        if (c->prec == PIMC_FP64) {
            const Prm<double> pr = Prm<double>::make(h);
            h2d(c, reinterpret_cast<Prm<double> *>(c->d_prm) + chain, &pr, sizeof(pr));
        } else {
            const Prm<float> pr = Prm<float>::make(h);
            h2d(c, reinterpret_cast<Prm<float> *>(c->d_prm) + chain, &pr, sizeof(pr));
        }
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_get_sweep(pimc_ctx *c, int64_t *sweep)
{
    if (!c) return PIMC_ERR_ARG;
    if (!sweep) return set_err(c, PIMC_ERR_ARG, "sweep is null");
    *sweep = (int64_t)c->sweep;
    return PIMC_OK;
}

int pimc_set_sweep(pimc_ctx *c, int64_t sweep)
{
    if (!c) return PIMC_ERR_ARG;
    if (sweep < 0 || sweep > 0xffffffffll) return set_err(c, PIMC_ERR_ARG, "sweep must be in [0, 2^32)");
    try {
        HIPCHK(hipStreamSynchronize(c->stream));
        const unsigned v = (unsigned)sweep;
        // the kernels read the counter from the device run state, so captured graphs stay valid
        h2d(c, reinterpret_cast<char *>(c->d_rs) + offsetof(RunState, sweep), &v, sizeof(v));
        c->sweep = v;
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_run(pimc_ctx *c, int32_t n_sweeps, int32_t measure, pimc_counters *out)
{
    if (!c) return PIMC_ERR_ARG;
    if (n_sweeps < 0) return set_err(c, PIMC_ERR_ARG, "n_sweeps < 0");
    if (measure && n_sweeps % c->opt.measure_every != 0)
        return set_err(c, PIMC_ERR_ARG, "a measuring run must be a multiple of measure_every sweeps");
    if (int rc = check_table(c, -1)) return rc;
    try {
        std::vector<long long> before((size_t)c->n_chains * kCnt);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(before.data(), c->d_cnt, sizeof(long long) * before.size(), hipMemcpyDeviceToHost));
        const int K = c->opt.sweeps_per_launch;
        const int nl = (n_sweeps + K - 1) / K;
        c->hblk_n = -1;
        if (measure && grow_blocks(c, c->nblk + nl)) drop_graphs(c); // the block buffer moved
        while ((int)c->events.size() < nl) {
            Launch l;
            HIPCHK(hipEventCreate(&l.a));
            HIPCHK(hipEventCreate(&l.b));
            c->events.push_back(l);
        }
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        int left = n_sweeps, li = 0;
        while (left > 0) {
            const int k = std::min(K, left);
            HIPCHK(hipEventRecord(c->events[li].a, c->stream));
            hipGraphExec_t &ge = c->graph[measure ? 1 : 0];
            if (c->opt.use_graph && k == K && ge) {
                HIPCHK(hipGraphLaunch(ge, c->stream));
            } else if (c->opt.use_graph && k == K && li > 0) {
                // capture one full launch step once (after an eager step has run) and replay it
                hipGraph_t g = nullptr;
                HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                launch_step(c, k, measure);
                HIPCHK(hipStreamEndCapture(c->stream, &g));
                HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
                HIPCHK(hipGraphDestroy(g));
                HIPCHK(hipGraphLaunch(ge, c->stream));
            } else {
                launch_step(c, k, measure);
            }
            HIPCHK(hipEventRecord(c->events[li].b, c->stream));
            // tile chains take their adapted staging lengths before the next launch
            if (!measure && c->opt.adapt_ell && !c->tiles.empty()) retile_after_adapt(c);
            if (measure) c->nblk++;
            c->sweep += (uint32_t)k;
            left -= k;
            ++li;
        }
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        int rc = check_status(c);
        if (out) {
            std::memset(out, 0, sizeof(*out));
            std::vector<long long> after(before.size());
            HIPCHK(hipMemcpy(after.data(), c->d_cnt, sizeof(long long) * after.size(), hipMemcpyDeviceToHost));
            for (int i = 0; i < c->n_chains; ++i) {
                const long long *b = before.data() + (size_t)i * kCnt, *e = after.data() + (size_t)i * kCnt;
                out->levy_acc += e[0] - b[0];
                out->com_acc += e[1] - b[1];
                out->com_tries += e[2] - b[2];
                out->swap_acc += e[3] - b[3];
                out->swap_tries += e[4] - b[4];
                out->levy_tries += e[5] - b[5];
                out->bead_moves += (e[6] - b[6]) + (int64_t)c->ch[i].M * (e[2] - b[2]) +
                                   2 * (int64_t)(c->ch[i].swap_m - 1) * (e[4] - b[4]);
            }
            out->sweeps = n_sweeps;
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            out->wall_ms = ms;
            double kms = 0.0;
            for (int i = 0; i < li; ++i) {
                float t = 0.f;
                HIPCHK(hipEventElapsedTime(&t, c->events[i].a, c->events[i].b));
                kms += t;
            }
            out->kernel_ms = kms;
            out->kernel_launches = li;
        }
        return rc;
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
}

int pimc_get_counters(pimc_ctx *c, int32_t chain, int64_t *out)
{
    if (!c) return PIMC_ERR_ARG;
    if (chain != -1) {
        int rc = check_chain(c, chain);
        if (rc) return rc;
    }
    if (!out) return set_err(c, PIMC_ERR_ARG, "null output");
    try {
        const int n = (int)c->ch.size();
        const int first = chain < 0 ? 0 : chain, count = chain < 0 ? n : 1;
        std::vector<long long> acc((size_t)count * kCnt);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(acc.data(), c->d_cnt + (size_t)first * kCnt, sizeof(long long) * acc.size(), hipMemcpyDeviceToHost));
        for (int i = 0; i < count; ++i) {
            out[4 * i + 0] = acc[kCnt * i];
            out[4 * i + 1] = acc[kCnt * i + 5];
            out[4 * i + 2] = acc[kCnt * i + 1];
            out[4 * i + 3] = acc[kCnt * i + 2];
        }
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_get_blocks(pimc_ctx *c, int32_t chain, double *E, double *Ev, double *Epair, int64_t cap, int64_t *n_blocks)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    try {
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<double> q((size_t)c->nblk * kNQ);
        for (int b = 0; b < c->nblk; ++b)
            std::memcpy(&q[(size_t)b * kNQ], host_block(c, b, chain), sizeof(double) * kNQ);
        if (n_blocks) *n_blocks = c->nblk;
        for (int b = 0; b < c->nblk && b < cap; ++b) {
            const double *x = &q[(size_t)b * kNQ];
            const double n = x[qN] > 0 ? x[qN] : 1.0;
            if (E) E[b] = x[qE] / n;
            if (Ev) Ev[b] = x[qEv] / n;
            if (Epair) Epair[b] = x[qEpair] / n;
        }
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_get_estimates(pimc_ctx *c, int32_t chain, pimc_estimates *out)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    if (!out) return set_err(c, PIMC_ERR_ARG, "null output");
    std::memset(out, 0, sizeof(*out));
    try {
        HIPCHK(hipStreamSynchronize(c->stream));
        const int nb = c->nblk;
        std::vector<double> q((size_t)nb * kNQ);
        for (int b = 0; b < nb; ++b)
            std::memcpy(&q[(size_t)b * kNQ], host_block(c, b, chain), sizeof(double) * kNQ);
        out->n_blocks = nb;
        if (nb == 0) return PIMC_OK;
        // totals and block means
        std::vector<double> tot(kNQ, 0.0);
        for (int b = 0; b < nb; ++b)
            for (int k = 0; k < kNQ; ++k) tot[k] += q[(size_t)b * kNQ + k];
        const double n = tot[qN];
        out->n_meas = (int64_t)n;
        if (n <= 0) return PIMC_OK;
        // Standard error by blocking (Flyvbjerg & Petersen, J. Chem. Phys. 91, 461 (1989)):
        // the launch blocks are merged pairwise level by level, and the largest error over
        // the levels that keep at least 8 blocks is kept, so correlations longer than one
        // launch do not shrink the error bar.
        auto mean_err = [&](int k, double &m, double &e) {
            m = tot[k] / n;
            e = 0.0;
            if (nb < 2) return;
            std::vector<double> sum(nb), cnt(nb);
            for (int b = 0; b < nb; ++b) {
                sum[b] = q[(size_t)b * kNQ + k];
                cnt[b] = q[(size_t)b * kNQ + qN];
            }
            int len = nb;
            while (len >= 2) {
                double s = 0.0;
                int used = 0;
                for (int b = 0; b < len; ++b) {
                    if (cnt[b] <= 0) continue;
                    const double x = sum[b] / cnt[b];
                    s += (x - m) * (x - m);
                    ++used;
                }
                if (used >= 2) e = std::max(e, std::sqrt(s / (used - 1) / used));
                if (len / 2 < 8) break;
                for (int b = 0; b < len / 2; ++b) {
                    sum[b] = sum[2 * b] + sum[2 * b + 1];
                    cnt[b] = cnt[2 * b] + cnt[2 * b + 1];
                }
                len /= 2;
            }
        };
        mean_err(qE, out->E, out->E_err);
        mean_err(qEv, out->Ev, out->Ev_err);
        mean_err(qEpair, out->Epair, out->Epair_err);
        out->dSdomega = tot[qSw] / n;
        out->dSdg = tot[qSg] / n;
        // reweighting derivatives with jackknife errors over blocks:
        // d<E>/dx = <dE/dx> - (<E dS/dx> - <E><dS/dx>)
        const bool whole = tot[qWhole] > 0;
        if (whole) {
            auto deriv = [&](const std::vector<double> &t, int kdE, int kES, int kS) {
                const double m = t[qN];
                return t[kdE] / m - (t[kES] / m - (t[qE] / m) * (t[kS] / m));
            };
            out->dE_domega = deriv(tot, qdEw, qESw, qSw);
            out->dE_dg = deriv(tot, qdEg, qESg, qSg);
            if (nb >= 2) {
                double s1 = 0.0, s2 = 0.0;
                std::vector<double> loo(kNQ);
                for (int b = 0; b < nb; ++b) {
                    for (int k = 0; k < kNQ; ++k) loo[k] = tot[k] - q[(size_t)b * kNQ + k];
                    if (loo[qN] <= 0) continue;
                    const double a1 = deriv(loo, qdEw, qESw, qSw) - out->dE_domega;
                    const double a2 = deriv(loo, qdEg, qESg, qSg) - out->dE_dg;
                    s1 += a1 * a1;
                    s2 += a2 * a2;
                }
                out->dE_domega_err = std::sqrt((double)(nb - 1) / nb * s1);
                out->dE_dg_err = std::sqrt((double)(nb - 1) / nb * s2);
            }
        } else {
            out->dE_domega = NAN;
            out->dE_dg = NAN;
        }
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_get_radial(pimc_ctx *c, int32_t chain, double *hist, int64_t *ndata)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    if (!c->d_radial) return set_err(c, PIMC_ERR_STATE, "radial histogram not enabled");
    try {
        const int nb = c->opt.radial_bins;
        std::vector<unsigned long long> h(nb);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(h.data(), c->d_radial + (size_t)chain * nb, sizeof(unsigned long long) * nb, hipMemcpyDeviceToHost));
        for (int i = 0; i < nb; ++i) hist[i] = (double)h[i];
        double nm = 0.0;
        for (int b = 0; b < c->nblk; ++b) {
            nm += host_block(c, b, chain)[qN];
        }
        if (ndata) *ndata = (int64_t)nm * c->ch[chain].M;
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_get_cartesian(pimc_ctx *c, int32_t chain, double *hist, int64_t *ndata)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    if (!c->d_cart) return set_err(c, PIMC_ERR_STATE, "Cartesian histogram not enabled");
    try {
        const long long nb = c->cart_per_chain;
        std::vector<unsigned long long> h((size_t)nb);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(h.data(), c->d_cart + (size_t)chain * nb, sizeof(unsigned long long) * nb, hipMemcpyDeviceToHost));
        for (long long i = 0; i < nb; ++i) hist[i] = (double)h[i];
        double nm = 0.0;
        for (int b = 0; b < c->nblk; ++b) {
            nm += host_block(c, b, chain)[qN];
        }
        if (ndata) *ndata = (int64_t)nm * c->ch[chain].M;
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_reset_estimators(pimc_ctx *c)
{
    if (!c) return PIMC_ERR_ARG;
    try {
        HIPCHK(hipStreamSynchronize(c->stream));
        c->nblk = 0;
        c->hblk_n = -1;
        {
            RunState rs{};
            HIPCHK(hipMemcpy(&rs, c->d_rs, sizeof(RunState), hipMemcpyDeviceToHost));
            rs.blk = 0;
            h2d(c, c->d_rs, &rs, sizeof(RunState));
        }
        // the histograms are cleared on the library stream, which the next run's kernels follow
        auto zero = [&](unsigned long long *p, long long n) {
            hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)std::min<long long>(1024, (n + 255) / 256)), dim3(256), 0,
                               c->stream, p, n);
            HIPCHK(hipGetLastError());
        };
        if (c->d_radial) zero(c->d_radial, (long long)c->n_chains * c->opt.radial_bins);
        if (c->d_cart) zero(c->d_cart, (long long)c->n_chains * c->cart_per_chain);
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

// ---------------------------------------------------------------- hooks
static int hook_move(pimc_ctx *c, int chain, int n, int j0, int m, const double *in, int nin, int op, double u,
                     double *r_out, double *w_out, int32_t *accepted, bool com)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    const ChainDev &h = c->ch[chain];
    if (n < 0 || n >= h.N) return set_err(c, PIMC_ERR_ARG, "particle out of range");
    if (!com && (j0 < 0 || j0 >= h.M || m < 2 || m > h.M)) return set_err(c, PIMC_ERR_ARG, "bad j0 or m");
    if (h.swaps > 0)
        return set_err(c, PIMC_ERR_ARG, "single moves treat each particle as its own ring; chains with swap moves "
                                        "sample through pimc_run");
    if (h.hs_a > 0)
        return set_err(c, PIMC_ERR_ARG, "single moves take injected draws; chains with hard spheres (hs_a > 0) "
                                        "sample through pimc_run");
    if (!in) return set_err(c, PIMC_ERR_ARG, "null input");
    if ((rc = check_table(c, chain))) return rc;
    // FP32 tile chains whose segments and COM chunks are one skew block of the LDS rows: the
    // batched kernel's fast-path code (levy_full_propose, com_chunk_full) with the given draws
    const int maxl = next_pow2_maxl(h.ell);
    const bool fast = c->prec == PIMC_FP32 && !c->is_whole[chain] && h.ell == maxl && h.M % maxl == 0 &&
                      (com || m == h.ell);
    const int pot = h.pot, nseg = h.M / maxl;
    int sh = 0;
    while ((2 << sh) <= maxl) ++sh;
    const size_t lds_fast = ((size_t)h.dim * tile_row_len(h.M, sh) + (size_t)nseg) * sizeof(float);
    try {
        const int rows = com ? h.M : m + 1;
        const size_t nout = (size_t)rows * h.dim + 5;
        ensure_io(c, nout + (size_t)nin + 8);
        double *d_in = c->d_io + nout;
        HIPCHK(hipStreamSynchronize(c->stream));
        h2d(c, d_in, in, sizeof(double) * nin);
        HookArgs h = hook_args(c, chain);
        h.n = n;
        h.j0 = j0;
        h.m = m;
        h.op = op;
        h.in = d_in;
        h.u = u;
        h.out = c->d_io;
        if (fast) {
            h.m = com ? 0 : m;
            launch_tile_fast(c, h, maxl, pot, lds_fast);
        } else if (com) {
            HOOK_DISPATCH(k_hook_com, dim3(1), dim3(64), 0, h);
        } else {
            HOOK_DISPATCH(k_hook_reshape, dim3(1), dim3(64), 0, h);
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<double> o(nout);
        HIPCHK(hipMemcpy(o.data(), c->d_io, sizeof(double) * nout, hipMemcpyDeviceToHost));
        if (r_out) std::memcpy(r_out, o.data(), sizeof(double) * (size_t)rows * c->dim);
        if (w_out) std::memcpy(w_out, o.data() + (size_t)rows * c->dim, sizeof(double) * 4);
        if (accepted) *accepted = o[(size_t)rows * c->dim + 4] != 0.0;
        return check_status(c);
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
}

int pimc_levy_propose(pimc_ctx *c, int32_t chain, int32_t n, int32_t j0, int32_t m, const double *normals,
                      double *r_out, double *w_out)
{
    if (!c) return PIMC_ERR_ARG;
    return hook_move(c, chain, n, j0, m, normals, (m - 1) * c->dim, 0, 0.0, r_out, w_out, nullptr, false);
}

int pimc_reshape_linear(pimc_ctx *c, int32_t chain, int32_t n, int32_t j0, int32_t m, const double *normals,
                        double u, int32_t *accepted)
{
    if (!c) return PIMC_ERR_ARG;
    return hook_move(c, chain, n, j0, m, normals, (m - 1) * c->dim, 1, u, nullptr, nullptr, accepted, false);
}

int pimc_com_propose(pimc_ctx *c, int32_t chain, int32_t n, const double *d, double *r_out, double *w_out)
{
    if (!c) return PIMC_ERR_ARG;
    return hook_move(c, chain, n, 0, 0, d, c->dim, 0, 0.0, r_out, w_out, nullptr, true);
}

int pimc_center_of_mass(pimc_ctx *c, int32_t chain, int32_t n, const double *d, double u, int32_t *accepted)
{
    if (!c) return PIMC_ERR_ARG;
    return hook_move(c, chain, n, 0, 0, d, c->dim, 1, u, nullptr, nullptr, accepted, true);
}

int pimc_interaction(pimc_ctx *c, int32_t chain, int32_t n, int32_t j, double *w)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    const ChainDev &hc = c->ch[chain];
    if (n < 0 || n >= hc.N || j < 0 || j >= hc.M) return set_err(c, PIMC_ERR_ARG, "bad particle or slice");
    if ((rc = check_table(c, chain))) return rc;
    try {
        ensure_io(c, 8);
        HIPCHK(hipStreamSynchronize(c->stream));
        HookArgs h = hook_args(c, chain);
        h.n = n;
        h.j0 = j;
        h.out = c->d_io;
        HOOK_DISPATCH(k_hook_interaction, dim3(1), dim3(64), 0, h);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(w, c->d_io, sizeof(double), hipMemcpyDeviceToHost));
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_energy(pimc_ctx *c, int32_t chain, double *out)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    try {
        ensure_io(c, 8);
        HIPCHK(hipStreamSynchronize(c->stream));
        HookArgs h = hook_args(c, chain);
        h.out = c->d_io;
        HOOK_DISPATCH(k_hook_energy, dim3(1), dim3(256), 256 * sizeof(double), h);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(out, c->d_io, sizeof(double) * 4, hipMemcpyDeviceToHost));
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_density(pimc_ctx *c, int32_t chain, int32_t nbins, double *dens, int64_t *ndata)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    if (nbins < 1 || !dens) return set_err(c, PIMC_ERR_ARG, "bad histogram");
    const ChainDev &hc = c->ch[chain];
    const size_t nh = hc.dim == 1 ? (size_t)nbins : (size_t)nbins * nbins;
    unsigned long long *d_h = nullptr;
    try {
        d_h = dalloc<unsigned long long>(nh);
        // cleared on the library stream, ahead of the counting kernel: a plain hipMemset runs on the
        // null stream, which the library's non-blocking stream does not wait for
        const unsigned fb = (unsigned)std::min<size_t>(1024, (nh + 255) / 256);
        hipLaunchKernelGGL(k_zero_u64, dim3(fb), dim3(256), 0, c->stream, d_h, (long long)nh);
        HIPCHK(hipGetLastError());
        HookArgs h = hook_args(c, chain);
        const int nb = std::max(1, std::min(1024, (hc.N * hc.M + 255) / 256));
#define DENS(T, D, S) hipLaunchKernelGGL((k_hook_density<T, D, S>), dim3(nb), dim3(256), 0, c->stream, h, nbins, d_h)
        const bool sm = hc.sm != 0;
        if (c->prec == PIMC_FP64) {
            if (c->dim == 1) { if (sm) DENS(double, 1, true); else DENS(double, 1, false); }
            else { if (sm) DENS(double, 2, true); else DENS(double, 2, false); }
        } else {
            if (c->dim == 1) { if (sm) DENS(float, 1, true); else DENS(float, 1, false); }
            else { if (sm) DENS(float, 2, true); else DENS(float, 2, false); }
        }
#undef DENS
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<unsigned long long> hh(nh);
        HIPCHK(hipMemcpy(hh.data(), d_h, sizeof(unsigned long long) * nh, hipMemcpyDeviceToHost));
        dfree(d_h);
        d_h = nullptr;
        for (size_t i = 0; i < nh; ++i) dens[i] += (double)hh[i];
        if (ndata) *ndata += hc.M;
    } catch (const HipError &e) {
        dfree(d_h);
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_density_noalloc(pimc_ctx *c, int32_t chain, int32_t nbins, double *dens, int64_t *ndata)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    if (nbins < 1 || !dens) return set_err(c, PIMC_ERR_ARG, "bad histogram");
    const ChainDev &hc = c->ch[chain];
    const long long nh = hc.dim == 1 ? (long long)nbins : (long long)nbins * nbins;
    if (nh > kScratchBins)
        return set_err(c, PIMC_ERR_ARG, "pimc_density_noalloc holds at most 262144 bins (nbins^dim; 512 x 512 in 2D): "
                                        "use pimc_density for larger histograms");
    if (!c->d_scratch) return set_err(c, PIMC_ERR_STATE, "no scratch address for this device");
    try {
        unsigned long long *d_h = c->d_scratch;
        // the counts start at zero: cleared by a kernel (the array keeps the last call's counts)
        const unsigned fb = (unsigned)std::min<long long>(1024, (nh + 255) / 256);
        hipLaunchKernelGGL(k_zero_u64, dim3(fb), dim3(256), 0, c->stream, d_h, nh);
        HIPCHK(hipGetLastError());
        HookArgs h = hook_args(c, chain);
        const int nb = std::max(1, std::min(1024, (hc.N * hc.M + 255) / 256));
#define DENS(T, D, S) hipLaunchKernelGGL((k_hook_density<T, D, S>), dim3(nb), dim3(256), 0, c->stream, h, nbins, d_h)
        const bool sm = hc.sm != 0;
        if (c->prec == PIMC_FP64) {
            if (c->dim == 1) { if (sm) DENS(double, 1, true); else DENS(double, 1, false); }
            else { if (sm) DENS(double, 2, true); else DENS(double, 2, false); }
        } else {
            if (c->dim == 1) { if (sm) DENS(float, 1, true); else DENS(float, 1, false); }
            else { if (sm) DENS(float, 2, true); else DENS(float, 2, false); }
        }
#undef DENS
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        // counts to the caller's doubles through a fixed host buffer
        unsigned long long buf[2048];
        for (long long i0 = 0; i0 < nh; i0 += 2048) {
            const long long cnt = std::min<long long>(2048, nh - i0);
            HIPCHK(hipMemcpy(buf, d_h + i0, sizeof(unsigned long long) * (size_t)cnt, hipMemcpyDeviceToHost));
            for (long long i = 0; i < cnt; ++i) dens[i0 + i] += (double)buf[i];
        }
        if (ndata) *ndata += hc.M;
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_adjoint(pimc_ctx *c, int32_t chain, double *dS_dr, double *dS_dparams, double *S_out)
{
    int rc = check_chain(c, chain);
    if (rc) return rc;
    const ChainDev &hc = c->ch[chain];
    if (hc.pair == kPairTable)
        return set_err(c, PIMC_ERR_ARG, "the adjoint is defined for a Gaussian pair action or none");
    const size_t nb = (size_t)hc.N * hc.dim * hc.M;
    try {
        ensure_io(c, nb + 8);
        HIPCHK(hipStreamSynchronize(c->stream));
        HookArgs h = hook_args(c, chain);
        h.out = c->d_io;
        HOOK_DISPATCH(k_hook_adjoint, dim3(1), dim3(256), 256 * sizeof(double), h);
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<double> o(nb + 3);
        HIPCHK(hipMemcpy(o.data(), c->d_io, sizeof(double) * (nb + 3), hipMemcpyDeviceToHost));
        if (dS_dr) std::memcpy(dS_dr, o.data(), sizeof(double) * nb);
        if (dS_dparams) {
            dS_dparams[0] = o[nb + 1];
            dS_dparams[1] = o[nb + 2];
        }
        if (S_out) *S_out = o[nb];
    } catch (const HipError &e) {
        return hip_err(c, e);
    }
    return PIMC_OK;
}

int pimc_action(pimc_ctx *c, int32_t chain, double *S_out)
{
    return pimc_adjoint(c, chain, nullptr, nullptr, S_out);
}

int pimc_prop_int_terms(int32_t n, double r0, double r1, double g0, double tau, double *terms)
{
    g_err.clear();
    if (n < 2 || !(r1 > r0) || !(r0 >= 0) || !(tau > 0) || !(g0 != 0) || !terms) {
        g_err = "need n >= 2, 0 <= r0 < r1, tau > 0, g0 != 0 and an output buffer";
        return PIMC_ERR_ARG;
    }
    std::vector<double> k, w;
    prop_quad_nodes(r1, tau, k, w);
    const int nk = (int)k.size();
    std::vector<double> ca(nk), cb(nk);
    for (int q = 0; q < nk; ++q) {
        // tk = 1 / d (src/propagator.jl:45); tk^2 / (1 + tk^2) = 1 / (1 + d^2), tk / (1 + tk^2) = d / (1 + d^2)
        const double pi = 3.14159265358979323846, eg = 0.57721566490153286061;
        const double d = (2.0 / pi) * (eg + std::log(k[q] / 2.0)) - 4.0 / g0;
        const double cq = w[q] * k[q] * std::exp(-tau * k[q] * k[q]) / (2.0 * pi);
        ca[q] = cq / (1.0 + d * d);
        cb[q] = cq * d / (1.0 + d * d);
    }
    double *dk = nullptr, *dca = nullptr, *dcb = nullptr, *dJ = nullptr, *dY = nullptr, *dout = nullptr;
    int rc = PIMC_OK;
    try {
        dk = dalloc<double>(nk);
        dca = dalloc<double>(nk);
        dcb = dalloc<double>(nk);
        dJ = dalloc<double>((size_t)n * nk);
        dY = dalloc<double>((size_t)n * nk);
        dout = dalloc<double>((size_t)n * n);
        HIPCHK(hipMemcpy(dk, k.data(), sizeof(double) * nk, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dca, ca.data(), sizeof(double) * nk, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dcb, cb.data(), sizeof(double) * nk, hipMemcpyHostToDevice));
        const double h = (r1 - r0) / (n - 1);
        const long long tot = (long long)n * nk;
        hipLaunchKernelGGL(k_prop_bessel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, nullptr, dk, nk, r0, h, n,
                           dJ, dY);
        HIPCHK(hipGetLastError());
        const unsigned nt = (unsigned)((n + kPropTile - 1) / kPropTile);
        hipLaunchKernelGGL(k_prop_terms, dim3(nt, nt), dim3(kPropThreads), 0, nullptr, dJ, dY, dca, dcb, nk, n, dout);
        HIPCHK(hipGetLastError());
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(terms, dout, sizeof(double) * n * n, hipMemcpyDeviceToHost));
    } catch (const HipError &e) {
        g_err = std::string(e.what) + ": " + hipGetErrorString(e.e);
        rc = PIMC_ERR_HIP;
    }
    for (double *p : {dk, dca, dcb, dJ, dY, dout}) dfree(p);
    return rc;
}

} // extern "C"
