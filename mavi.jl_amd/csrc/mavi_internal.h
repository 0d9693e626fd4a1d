// mavi_internal.h -- shared types and device helpers of libmavi_hip.so.
//
// Device data layout (per system, HBM, allocated once at creation).
// The particle state is kept SORTED BY CHUNK between steps: every step's
// binning permutes it into the new chunk order, so all kernels read and
// write it with unit stride.  Two copies ping-pong (index `cur`).
//   P[2], V[2]   double2[na]   positions / velocities, sorted slots
//   ID[2]        int32[na]     slot -> particle id (id = slot of state.pos)
//   F            double2[na]   forces of the last evaluation, sorted slots
//   P2, F1       double2[na]   drifted positions x' and first-evaluation
//                              forces inside a Verlet step
//   key, rank    int32[na]     new chunk of each old slot / rank in it
//   src, scell   int32[na]     new slot -> old slot / chunk of new slot
//   sid          int32[na]     id of each new slot in arrival order
//   slot_of      int32[n]      particle id -> sorted slot (rings)
//   cnt, start   int32[ncell+1] particles per chunk / exclusive scan
//   scan_state   u64[ntiles]   published tile states of the one-pass scan
//                              (4096 or 16384 chunks per tile), scan_ticket its counter
//   pos_id, vel_id, force_id   double2[n] id-order staging for the host
// na = number of active ids (get_particles_ids); the chunk index is
// row0*cols + col0 with row0 = reference row - 1 (row 1 = top of the box,
// src/chunks.jl:129) and col0 = reference col - 1.  Row-major chunks make the
// three chunks (row, col-1..col+1) one contiguous range of sorted slots.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mavi_hip.h"

#define MAVI_FLAG_OVERFLOW 1u
#define MAVI_FLAG_OUT_OF_BOX 2u
#define MAVI_FLAG_NONFINITE 4u
#define MAVI_FLAG_SCAN_STALL 8u
#define MAVI_FLAG_NEIGH_OVERFLOW 16u
#define MAVI_FLAG_RING_OUT_OF_BOX 32u
#define MAVI_FLAG_RING_CHUNK_OVERFLOW 64u

// Lifetimes.  Systems, batches and networks hold a reference to their
// context, and a learned propagator to its system and network.  Destroying
// an object that something still references only marks it; the last release
// frees it.  Callers (and garbage collectors, which finalize in any order)
// may therefore destroy objects in any order.
struct mavi_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int refs = 0;
    bool closing = false;
};
#define MAVI_HIDDEN __attribute__((visibility("hidden")))
extern "C" {
MAVI_HIDDEN void ctx_retain(mavi_ctx *ctx);
MAVI_HIDDEN void ctx_release(mavi_ctx *ctx);  // frees a destroyed context at its last release
}

struct DevWall;
struct DevLine;

// Parameters every kernel of a system reads (passed by value).
struct SysParams {
    int32_t n;         // slots
    int32_t na;        // active ids
    int32_t cols, rows, ncell, nc;
    int32_t periodic;
    int32_t potential;
    double blx, bly, L, H;
    double cl, ch;     // chunk_length, chunk_height
    double halfL, halfH;
    double dt, term, hdt;  // dt, dt^2/2, dt/2
    double radius;
    // pair potential (LJ or HarmTrunc; rings use the tables below)
    double p0, p1, p2, p3;
    double cut2;       // exact threshold: dist > rcut  <=>  r2 > cut2
    double eq2;        // exact threshold: dist < dist_eq <=> r2 < eq2
    double lj_c12, lj_c6;  // 48 eps sigma^12, 24 eps sigma^6
    // rings
    int32_t nmax, nrings, ntypes;
    int32_t typed;
    // self-propelled particles (SzaboCfg / RunTumbleCfg)
    double sz_rep, sz_adh;    // k_rep / (r_max - r_eq), k_adh / r_eq
    double ap_vo, ap_mu;      // vo, mobility
    double ap_inv_tau;        // 1 / relax_time
    double ap_noise;          // sqrt(2 rot_diff dt)
    double ap_tumble;         // tumble_rate * dt
    // walls list (nwalls > 0 replaces the periodic / rigid rectangle)
    int32_t nwalls, pad_w;
    const DevWall *walls;
    const DevLine *lines;
    // rings with a variable population: 1 per active ring slot (null: all)
    const int32_t *rmask;
};

// Per ring type parameters (wradius: get_particle_radius of the type, the
// slippery walls' particle radius, src/rings/rings.jl:44-53).
struct RingType {
    double p0, relax, vo, mu, rot, karea, kspring, lspring;
    double wradius;
    int32_t np;
    int32_t pad;
};

// Device copy of a mavi_wall (lines in a separate table, with the unit
// normal, tangent and length of Line2D, src/configs.jl:99-117).
struct DevWall {
    int32_t kind, mode, pot_kind, first, nlines, pad;
    double g[4];
    double pot[4];
};
struct DevLine {
    double p1x, p1y, p2x, p2y, nx, ny, tx, ty, len, pad;
};

// Interaction (HarmTrunc) for one pair of ring types, with exact thresholds.
struct RingInter {
    double k_rep, k_atr, dist_eq, dist_max;
    double max2;   // dist > dist_max <=> r2 > max2
    double eq2;    // dist < dist_eq  <=> r2 < eq2
    double nb2;    // contact neighbours: dist < 2 particle_radius * tol <=> r2 < nb2
};

// Chunk grid of the rings' centres of mass (r_chunks, src/rings/rings.jl:
// 188-202) over the system's chunk space; cols == 0: no grid (all pairs).
struct InvGrid {
    int32_t rows, cols, nc, pad;
    double cl, ch;
};

// Sources and sinks of rings (src/rings/sources.jl) on the device: the
// sinks' geometries and the sources' spawn boxes (bbox grown by pad is the
// emptiness test; chunked: only the particles of the box's ChunksChecker
// chunks count, else every active particle).
struct DevSink {
    int32_t geom, pad;
    double g[4];
};
struct DevBox {
    double x, y, len, h, pad;
    int32_t chunked, pad2;
};
constexpr int kMaxSinks = 32;     // sink bits per ring
constexpr int kMaxBoxes = 64;     // spawn boxes over all sources (one 64-bit word)
constexpr int kSinkFetch = 512;   // sink hits read back with the step's first copy

// Host copy of one source or sink (mavi_ring_source).
struct HostSource {
    int32_t kind = 0, geom = 0;
    double g[4] = {0, 0, 0, 0};
    int32_t nbox = 0, random_pol = 0;
    double pad = 0.0, pol = 0.0;
    int32_t box0 = 0;              // first global box index
    int32_t sink_bit = 0;          // sinks: bit in the per-ring sink word
    std::vector<double> spawn;     // nbox x nmax x 2
};

struct ProfSlot {
    double ms = 0.0;
    int64_t launches = 0;
};

// Outputs of a force-only ring pass for the learned propagator's features.
struct RingRates {
    double2 *U = nullptr;      // per slot: the overdamped velocity v0 p + mu F
    double *rate = nullptr;    // per ring: d(area)/dt over the target area
    double *omega = nullptr;   // per ring: asin of the polarisation x mean-velocity alignment
    uint32_t *gate = nullptr;  // cleared
};

constexpr int32_t kNlBlockCap = 60;  // most steps between host checks of the list validity

// Verlet neighbour lists of a finite-range plain system (LJ with a cutoff,
// HarmTrunc) whose chunks are at least the interaction range wide (mavi_nlist.hip).
struct NList {
    bool on = true;           // option MAVI_OPT_NLIST
    bool valid = false;       // built for the current slot order and still complete
    bool disabled = false;    // a list overflowed: the chunk gather takes over
    double range = 0.0;       // interaction range (rcut / dist_max)
    double skin_frac = 0.15;  // skin = skin_frac * range (option, thousandths)
    double rl2 = 0.0;         // list radius squared (range + skin)^2, with margin
    double thr2 = 0.0;        // displacement bound squared (skin / 2)^2, with margin
    double edge = 0.0;        // within this distance of a periodic edge: minimum image
    int32_t kmax = 48;        // entries per particle
    int32_t block = 0;        // steps between host checks (option; 0: as many as the bound allows, at most kNlBlockCap)
    int32_t block_now = kNlBlockCap;  // current largest block (halved after a rollback)
    double step_hist[3] = {0.0, 0.0, 0.0};  // the largest drift step of each of the last three blocks
    int32_t step_i = 0;
    SysParams G{};            // internal grid for sorting and building (cells >= range + skin)
    int32_t *nbr = nullptr;   // (kmax / 4) x na groups of 4 16-bit partner offsets (int2), ascending partner id, padded with the particle
    int32_t *ncount = nullptr;// entries per slot; bit 31: minimum image needed
    int32_t *key[3] = {nullptr, nullptr, nullptr};  // reference chunk (row << 16 | col) of each slot: K_n, K_{n+1}, K_{n+2} rotating
    int ph = 0;               // steps since the build mod 6: key[ph % 3] holds K_n; chg, work, px ping-pong by ph % 2
    int32_t *ccnt = nullptr;  // particles per reference chunk
    uint8_t *chg[2] = {nullptr, nullptr};   // a step's drift marks: 2 = changed chunk, 1 = a list partner did (stale test)
    double2 *fk = nullptr;                  // the next first evaluation where it differs from the last second one
    uint8_t *fk_on = nullptr;               // fk holds it
    double2 *px[2] = {nullptr, nullptr};    // drifted positions x_{n+1}, x_{n+2} (before the periodic wrap)
    bool unsettled = false;                 // the state's positions lag the drifted ones (nl_settle)
    double2 *xb = nullptr;    // positions at build
    double2 *vsnap = nullptr;  // block-start velocities (rollback; positions are written after the check)
    uint32_t *fsnap = nullptr;
    unsigned long long *dmax = nullptr;  // reduced per block (bits): [0] largest drift step^2, [1] largest pending displacement^2, [2] status
    double *bmax = nullptr;              // per workgroup of the evaluations: largest drift step^2 since the block began
    double *dpend = nullptr;             // per workgroup: largest displacement^2 of the pending positions (block end)
    int32_t nbmax = 0;
    uint32_t *status = nullptr;          // bit 0: a list overflowed kmax; bit 1: a chunk count may have passed the capacity
    unsigned long long *scan_ticket = nullptr, *scan_state = nullptr;  // the list grid's scan (cleared per build)
    int32_t scan_words = 0;
    unsigned long long *host = nullptr;  // pinned read back: dmax[0..1], status
    double d_start = 0.0;     // largest displacement from the build of the next block's first positions
    int64_t rebuilds = 0, rollbacks = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // six steps from ph = 0 on state copy 0 / 1 (a rebuild's sort flips it)
};

struct mavi_sys {
    mavi_ctx *ctx = nullptr;
    int refs = 0;          // learned propagators bound to this system
    bool closing = false;  // destroyed while referenced: freed at the last release
    bool ctx_held = false; // holds a reference to ctx (set once creation succeeded)
    SysParams P{};
    int32_t wg = 64;   // pair-force workgroup size (one wave: waves of unequal length free CU slots one at a time)
    bool two_phase = false; // candidate list for finite-range interactions
    bool lds_tiles = false; // LDS-staged chunk-row segments for the pair kernels
    double list2 = INFINITY; // rings: widest interaction range squared
    int cur = 0;       // which ping-pong copy holds the state
    // device buffers
    double2 *P_[2] = {nullptr, nullptr}, *V_[2] = {nullptr, nullptr};
    int32_t *ID_[2] = {nullptr, nullptr};
    double2 *F = nullptr, *P2 = nullptr, *F1 = nullptr;
    int32_t *key = nullptr, *rank = nullptr, *src = nullptr, *scell = nullptr, *sid = nullptr, *slot_of = nullptr;
    int32_t *cnt = nullptr, *start = nullptr;
    unsigned long long *scan_state = nullptr, *scan_ticket = nullptr;  // one-pass scan
    uint32_t *flags = nullptr;
    unsigned long long *counter = nullptr;  // device step counter (rotational noise)
    int32_t *active_ids = nullptr;          // na ids ascending
    double2 *pos_id = nullptr, *vel_id = nullptr, *force_id = nullptr;
    int32_t nscan_blocks = 0;
    // rings
    double *pol = nullptr, *areas = nullptr;
    double2 *cms = nullptr, *cms_next = nullptr, *cpos = nullptr;
    int32_t *rtype = nullptr;
    RingType *types = nullptr;
    RingInter *inter = nullptr;
    double *noise = nullptr;
    int64_t noise_cap = 0;
    uint64_t seed = 0;
    // observables scratch
    double *red = nullptr;
    int32_t red_cap = 0;
    bool has_state = false;
    bool sorted = false;        // a binning has put the state in chunk order
    // reuse of the second force evaluation of a Verlet step as the first one
    // of the next step where a particle's 3x3 chunk block kept its members
    int32_t *stamp = nullptr;   // per chunk: binning epoch in which its membership last changed
    // contact neighbours of rings (ParticleNeighbors), 15 per particle
    bool neigh_on = false;
    int neigh_only_rings = 0;
    double neigh_tol = 1.1;
    int32_t *neigh_cnt = nullptr, *neigh_list = nullptr;  // by particle id
    std::vector<RingInter> host_inter;
    std::vector<int32_t> host_ring_np;  // particles per ring slot (its type's num_particles)
    // invasions of rings (InvasionsCfg, r_chunks)
    bool inv_sched = false;     // InvasionsCfg given: check every inv_every steps
    int64_t inv_every = 0, inv_since = 0;  // steps_to_update; num_steps - last_check
    InvGrid inv_grid{};
    int32_t *inv_cnt = nullptr, *inv_cell = nullptr, *inv_chunk = nullptr;  // ring grid
    double4 *inv_box = nullptr; // per ring: bounding box of its ring points
    int32_t *inv_rec = nullptr; // records (invasor, invaded, p_id), 1-based
    uint32_t *inv_n = nullptr;  // records found by the last check
    int64_t inv_cap = 0, inv_count = 0;
    int32_t inv_cnt_cap = 0, inv_cell_cap = 0;  // allocated ring chunks / chunk slots
    // rings with a variable population (VarRingsIds) and their sources/sinks
    bool var = false;
    std::vector<int32_t> vmask;         // host copy, authoritative
    std::vector<int64_t> vids, vuids;   // VarRingsIds.ids (1-based, stale tail kept), uids
    int64_t vnum_active = 0;
    std::vector<HostSource> vsrc;
    int32_t nsinks = 0, nboxes = 0;
    std::vector<double> pol_draws;      // uniform draws for random spawn polarisations
    int64_t pol_used = 0;
    int64_t pol_stream = 0;             // draws taken from the built-in stream once the pool is empty
    int32_t *rmask = nullptr;           // device mask (P.rmask)
    DevSink *dsinks = nullptr;
    DevBox *dboxes = nullptr;
    unsigned long long *box_mask = nullptr;  // per chunk: boxes whose ChunksChecker lists it
    // per-step read back, one buffer: header (sink hits, occupied spawn
    // boxes) then the (ring, sink bits) pairs
    void *vrb = nullptr;
    uint32_t *sink_n = nullptr;         // -> vrb header
    unsigned long long *nonempty = nullptr;
    int32_t *sink_list = nullptr;       // -> after the header, 2 x nrings
    std::vector<uint32_t> vbits;        // host scratch: sink bits per ring slot (kept zero between steps)
    std::vector<char> vadded;           // host scratch: spawned this step
    std::vector<double> vacm;           // host scratch: centres of mass of this step's spawns
    void *vstage = nullptr;             // pinned host staging (read back and uploads)
    size_t vstage_bytes = 0;
    DevWall *walls = nullptr;   // walls list (P.walls / P.lines point here)
    DevLine *lines = nullptr;
    int32_t *needc = nullptr;   // per chunk: binning epoch in which a chunk of its 3x3 block changed
    int32_t *perm = nullptr;    // sorted slot -> slot before the last binning
    uint32_t *epoch = nullptr;  // binnings so far (device)
    bool reuse = true;          // option
    bool f_valid = false;       // F holds the last step's second evaluation
    bool keys_ready = false;    // the last second evaluation left key/rank/cnt for the next binning
    NList *nl = nullptr;        // neighbour lists (finite-range plain systems), null if not eligible
    // graphs of two steps (state starting in copy 0)
    hipGraphExec_t graph2 = nullptr;
    bool use_graphs = true;
    // profiling
    bool prof = false;
    ProfSlot profs[MAVI_PROF_NCLASS];
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

// --------------------------------------------------------------------------
// Device helpers
// --------------------------------------------------------------------------

// Exact truncated quotient of a/b, equal to Julia's Base div(a, b) for
// Float64 (src/chunks.jl:129-130).  q0 = trunc(a * rcp(b)) is within one of
// trunc(a/b) while |a/b| < 2^20 (v_rcp_f64 is good to 5e-8); the FMA
// residual r = a - q0 b is exact in sign and in its comparison with +-b, and
// moves q0 to the quotient whose residual lies in [0, b) (a >= 0) or (-b, 0]
// (a < 0).  Farther quotients (positions far off the grid) stay far off it.
__device__ __forceinline__ double exact_trunc_div(double a, double b) {
    double q = trunc(a * __builtin_amdgcn_rcp(b));
    const double r = __builtin_fma(-q, b, a);
    if (a >= 0.0) {
        if (r < 0.0) q -= 1.0;
        else if (r >= b) q += 1.0;
    } else {
        if (r > 0.0) q += 1.0;
        else if (r <= -b) q -= 1.0;
    }
    return q;
}

// calc_diff for a periodic rectangle, src/integration.jl:43-48.
__device__ __forceinline__ double min_image(double dr, double size, double half) {
    double m = (fabs(dr) > half) ? 1.0 : 0.0;
    return dr - m * copysign(size, dr);
}

extern "C" {
MAVI_HIDDEN void sys_retain(mavi_sys *s);
MAVI_HIDDEN void sys_release(mavi_sys *s);  // frees a destroyed system at its last release
}

// Host-side launchers (mavi_kernels.hip).  Each works on the state copy
// s->cur and leaves the result in s->cur (binning flips it).
namespace mavi {
// device address of mavi_mlp_forward_noalloc's staging array (mavi_mlp.hip),
// resolved once per device with that device current
hipError_t noalloc_scratch_setup(int device);
hipError_t launch_iota_state(mavi_sys *s, hipStream_t st);          // id-order staging -> state copy cur
hipError_t launch_unpermute(mavi_sys *s, bool vel, bool force, hipStream_t st);  // state -> staging
// binning + permutation (cur flips); count_step advances the step counter
// of the device rotational noise (a step's binning, not a lone update_chunks!)
// stamp: mark chunks whose membership changed (force reuse); defer: leave
// velocities in the old order and write perm (new slot -> old slot) for the
// next pair launch to move them
// over: another grid over the same space, with its own scan ticket and
// states (the one-pass scan's epochs assume one tile count per ticket)
hipError_t launch_bin(mavi_sys *s, hipStream_t st, bool count_step = false, bool stamp = false,
                      bool defer = false, const SysParams *over = nullptr, unsigned long long *over_ticket = nullptr,
                      unsigned long long *over_state = nullptr);
// Neighbour lists (mavi_nlist.hip): eligibility and allocation at creation,
// the rebuild (sort on the list grid, lists sorted by partner id, reference
// chunk keys and counts), one newton_step! (two launches), the block-end
// settling of the state's positions, and release.
bool nl_eligible(const mavi_sys *s);
hipError_t nl_alloc(mavi_sys *s, hipStream_t st);
void nl_free(mavi_sys *s);
hipError_t nl_rebuild(mavi_sys *s, hipStream_t st);
hipError_t nl_step(mavi_sys *s, hipStream_t st);
// block start: step maxima cleared (clear) and the flags word kept
hipError_t nl_block_start(mavi_sys *s, hipStream_t st, bool clear);
// block end: dmax <- the largest drift step, the largest displacement of the
// pending positions from the build, the status word
hipError_t nl_measure(mavi_sys *s, hipStream_t st);
// after an accepted block: the state's positions <- the wrap of the last
// evaluated drift
hipError_t nl_settle(mavi_sys *s, hipStream_t st);
void nl_set_skin(mavi_sys *s, double skin_frac);
hipError_t launch_export_chunks(mavi_sys *s, int64_t *d_cp, int64_t *d_counts, hipStream_t st);
hipError_t launch_forces_only(mavi_sys *s, hipStream_t st);
// One newton_step!; reuse = the previous call was a Verlet step of this
// system (F holds its second evaluation), so unchanged blocks reuse it.
hipError_t launch_verlet_step(mavi_sys *s, hipStream_t st, bool reuse = false);
hipError_t launch_rings_step(mavi_sys *s, const double *noise_dev, hipStream_t st);
// One szabo_step! / rtp_step!; noise_dev = this step's host draws, or null
// for the device Philox stream.
hipError_t launch_active_step(mavi_sys *s, const double *noise_dev, hipStream_t st);
// The ring pass of forces! can also give the learned propagator's per-ring
// inputs (RingRates): per slot u = v0 p + mu F, per ring the rate of its area
// (in units of the target area) and the alignment angle; it clears *gate.
hipError_t launch_rings_forces(mavi_sys *s, hipStream_t st, const RingRates *rates = nullptr);
hipError_t launch_rings_init(mavi_sys *s, hipStream_t st);
hipError_t launch_ring_neigh(mavi_sys *s, hipStream_t st);
// check_invasions!: ring boxes and the ring grid, then the crossing tests;
// prep = false reruns only the tests (after the record buffer grew)
hipError_t launch_invasions(mavi_sys *s, bool prep, hipStream_t st);
// update_cms! over the active rings and the sinks' membership of every ring
// slot, then the spawn boxes' emptiness (start of a variable-population step)
hipError_t launch_var_pre(mavi_sys *s, hipStream_t st);
// zero cms / cms_next of the inactive ring slots (RingsInfo.cms starts at zero)
hipError_t launch_var_clear(mavi_sys *s, hipStream_t st);

// The update kernel of a rings step (fixed ring population, rings of up to
// 32 particles) or of a self-propelled step does the next binning's first
// pass itself on a chunk grid: the binning then starts from its keys
// (keys_ready).
inline bool update_fuses_bin(const mavi_sys *s) {
    if (s->P.cols <= 0 || s->P.na <= 0) return false;
    if (s->P.potential == MAVI_POT_RINGS) return !s->var && s->P.nmax <= 32;
    return s->P.potential == MAVI_POT_SZABO || s->P.potential == MAVI_POT_RTP;
}

// The state was replaced or moved outside the kernels that keep the chunk
// keys: drop the keys and counts a previous pass left for the next binning.
hipError_t drop_keys(mavi_sys *s, hipStream_t st);
hipError_t launch_observables(mavi_sys *s, double *d_out, hipStream_t st);
hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t st);
}  // namespace mavi

// Host-side result of validating a descriptor (mavi_capi.hip build_params).
struct HostSys {
    SysParams P{};
    double list2 = INFINITY;
    uint64_t seed = 0;
    std::vector<int32_t> active;   // active ids (rings)
    std::vector<int32_t> rtype;    // per ring, 0-based type
    std::vector<RingType> types;
    std::vector<RingInter> inter;
    std::vector<DevWall> walls;    // walls list (device copies are set into P)
    std::vector<DevLine> lines;
};
int build_params(const mavi_sys_desc *d, HostSys &h, std::string &err);

// Host helpers (mavi_capi.hip)
double exact_gt_threshold(double rc);   // largest r2 with sqrt(r2) <= rc
double exact_lt_threshold(double deq);  // smallest r2 with sqrt(r2) >= deq
