// mavi_nlist.hip -- Verlet neighbour lists for newton_step! of finite-range
// plain systems (LenJonesCfg with a cutoff, HarmTruncCfg) whose chunks are at
// least the interaction range wide (calc_forces!, src/integration.jl:112-157;
// update_verlet!, :415-431; newton_step!, :507-515).
//
// Which pairs interact.  The reference's calc_forces! visits the pairs of
// particles in the same or adjacent chunks of the last update_chunks! and the
// range drops the distant ones.  With chunks at least the range wide, every
// pair within range at the binned positions is chunk-adjacent, so the first
// evaluation of a step is "every pair within range".  The second evaluation
// keeps the step's chunks at the drifted positions: a pair within range whose
// particles' step-start chunks are not adjacent is left out.  That can only
// happen to a pair with a particle whose chunk changed in the drift; such a
// particle marks the list partners whose step-start chunk is not adjacent to
// its own, and the marked particles test chunk adjacency pair by pair.
//
// Lists.  Every particle lists the others within range + skin, found on a
// grid of cells at least that wide after sorting the state by those cells.
// The state keeps that order until the next rebuild, so a step is two
// launch and no binning: the second evaluation of step n also makes the
// first evaluation of step n + 1 and its drift.  Where the pair set did not
// change (no chunk change in or around the particle's list) the next first
// evaluation is the second one, bit for bit; the marked particles sum their
// list a second time, every pair within range at the wrapped positions, as
// the next step's calc_forces! would.  Between the host's checks positions
// live only as the drifted ones before the periodic wrap (two copies,
// x_{n+1} and x_{n+2}); the state's positions, their wrap, are written once
// per block.  A list stays complete while no particle has moved
// more than skin / 2 since the build (no two particles can then close the
// gap).  Each drift records the largest step length of its workgroup; at a
// block's end the exact largest displacement of the next positions is
// measured, and the host checks D_start + T * (largest step) against skin /
// 2 (a bound for every position the block's evaluations used), rolling the
// block back and rebuilding when it was exceeded.
//
// Order.  A list is sorted by partner id, so a particle's force is the sum of
// its in-range pairs in ascending partner id whatever the slot order and
// whenever the list was built.  Runs do not depend on when lists were
// rebuilt, and the first evaluation of a step takes the previous second
// evaluation bit for bit wherever the pair set is the same (no chunk change
// in or around the particle's list).
//
// Errors.  The reference chunk of every final position is computed every
// step (out of box, non-finite), and chunk counts follow the particles that
// change chunk (the capacity check of update_particle_chunk!,
// src/chunks.jl:144-146): a count that passes the capacity while a step's
// moves are applied (at least the count after them) hands the block to the
// chunk gather, whose binning checks it exactly.
#include <math.h>

#include "mavi_device.h"
#include "mavi_host.h"

namespace {

constexpr uint32_t kEdgeBit = 0x80000000u;  // ncount: minimum image needed
constexpr uint32_t kStatusOverflow = 1u;    // a list held more than kmax entries, or a partner more
                                            // than 32767 slots away (modulo na)
constexpr uint32_t kStatusCapacity = 2u;    // a chunk count passed the capacity during a step's moves
constexpr int kNlMax = 48;                  // list entries per particle (12 groups of 4)
constexpr int kNeedBlock = 64;              // workgroup of the evaluation kernels: one wave, so waves of
                                            // unequal length free and take CU slots one at a time
constexpr int kGhost = 32768;               // drifted-position copies each side of the slots (see nl_sum)

struct NLArgs {
    const int2 *nbr;      // groups of 4 partner offsets ((slot - s) mod na in [-32768, 32767]): nbr[g * na + s]
    const int32_t *ncount;
    const int32_t *kcur;  // K_n: reference chunk (row << 16 | col) of the step's start positions
    int32_t *knext;       // K_{n+1}: of the drifted positions
    int32_t *knn;         // K_{n+2}: of the next step's drift
    int32_t *ccnt;
    uint8_t *chg;         // marks of this step's drift: 2 = changed chunk, 1 = a list partner did
    uint8_t *chg_next;    // marks of the next step's drift
    double *bmax;         // per workgroup: the largest squared drift step since the last reset
    const double2 *Pst;   // the state's positions (x_n after a build)
    double2 *V;
    double2 *F;           // the last second evaluation (the reference's forces)
    double2 *Fk;          // where fk_on: the next first evaluation, when it differs from F
    uint8_t *fk_on;
    double2 *Px;          // drifted positions x_{n+1} (before the periodic wrap)
    double2 *Px_next;     // x_{n+2}
    uint32_t *flags;
    uint32_t *status;
    int32_t reuse;        // 0: every particle sums its next first evaluation afresh
};

__device__ __forceinline__ int32_t pack_rc(int32_t r, int32_t c) { return (r << 16) | c; }

// A mark bit on slot j's byte (atomic on its 32-bit word).
__device__ __forceinline__ void mark(uint8_t *chg, int j, uint32_t bit) {
    atomicOr((uint32_t *)(chg + (j & ~3)), bit << (8 * (j & 3)));
}

// The state's position of a drifted one: walls! wraps it into a periodic box.
__device__ __forceinline__ double2 state_pos(const SysParams &P, double2 p) {
    if (P.periodic) periodic_wrap(P, p.x, p.y);
    return p;
}

// Chunks of the reference grid adjacent (the same chunk or one of its eight
// neighbours, across the periodic edges for a periodic box): the pairs
// calc_forces! visits (src/chunks.jl:61-118 on grids of at least 3 x 3).
__device__ __forceinline__ bool chunks_adjacent(const SysParams &P, int ri, int ci, int32_t kj) {
    const int rj = kj >> 16, cj = kj & 0xffff;
    int dr = abs(ri - rj), dc = abs(ci - cj);
    if (P.periodic) {
        dr = min(dr, P.rows - dr);
        dc = min(dc, P.cols - dc);
    }
    return dr <= 1 && dc <= 1;
}

// One pair.  A list pad is the particle itself (its offsets are 0 and the LJ
// factor is masked), so it adds nothing.  STALE: with `test`, the pair counts
// only when the step-start chunks are adjacent.
// Slots of a group's two entries packed in one word: 16-bit offsets from s
// taken modulo na (periodic rows put partners about na slots away).
__device__ __forceinline__ int wrap_slot(int j, int na) { return j < 0 ? j + na : (j >= na ? j - na : j); }
__device__ __forceinline__ int lo_slot(int s, int32_t w, int na) {
    return wrap_slot(s + (int32_t)(int16_t)(w & 0xffff), na);
}
__device__ __forceinline__ int hi_slot(int s, int32_t w, int na) { return wrap_slot(s + (w >> 16), na); }

// (j - s) mod na folded into [-na/2, na/2]; the caller checks it fits 16 bits.
__device__ __forceinline__ int32_t slot_offset(int s, int j, int na) {
    int32_t d = j - s;
    if (d > na / 2) d -= na;
    if (d < -(na / 2)) d += na;
    return d;
}

template <int POT, bool WRAP, bool STALE>
__device__ __forceinline__ void nl_pair(const SysParams &P, const double2 pj, const int32_t j, const int s,
                                        double xi, double yi, bool test, int ri, int ci,
                                        const int32_t *__restrict__ kcur, PairAcc<POT> &acc) {
    double dx = xi - pj.x, dy = yi - pj.y;
    if constexpr (WRAP) {
        dx = min_image(dx, P.L, P.halfL);
        dy = min_image(dy, P.H, P.halfH);
    }
    bool take = j != s;
    if constexpr (STALE) take = take && (!test || chunks_adjacent(P, ri, ci, kcur[wrap_slot(j, P.na)]));
    if constexpr (POT == POT_LJ) {
        const double f = lj_scale_masked(P, dx * dx + dy * dy, take);
        acc.fx = __builtin_fma(f, dx, acc.fx);
        acc.fy = __builtin_fma(f, dy, acc.fy);
    } else {
        RingCtx R{};
        if (take) acc.add(P, R, dx, dy, true, 0);
    }
}

// The sum over one particle's list in list order (ascending partner id),
// four entries per iteration; the next group is loaded one iteration ahead,
// so the list and position loads of consecutive groups overlap.  WRAPIN: pin
// holds drifted positions, taken through state_pos.  GHOST: pin holds
// kGhost copies of the other end's positions on each side of the slots
// (pin[j - na] = pin[j] for j >= na - kGhost, pin[j + na] = pin[j] for
// j < kGhost), so a partner is read at s + offset without the wrap modulo na
// (|offset| <= min(32768, na / 2)).
template <int POT, bool WRAP, bool STALE, bool WRAPIN = false, bool GHOST = false>
__device__ __forceinline__ void nl_sum(const SysParams &P, const double2 *__restrict__ pin,
                                       const int2 *__restrict__ nbr, int s, int groups, double xi, double yi,
                                       bool test, int ri, int ci, const int32_t *__restrict__ kcur,
                                       PairAcc<POT> &acc) {
    if (groups <= 0) return;
    const int na = P.na;
    const double2 *pb = GHOST ? pin - kGhost : pin;  // indexed from 0 with kGhost added
    int2 q = nbr[s];
    for (int g = 0; g < groups; g++) {
        const int2 c = q;
        if (g + 1 < groups) q = nbr[(g + 1) * na + s];
        int j[4];
        if constexpr (GHOST) {
            j[0] = s + (int32_t)(int16_t)(c.x & 0xffff);
            j[1] = s + (c.x >> 16);
            j[2] = s + (int32_t)(int16_t)(c.y & 0xffff);
            j[3] = s + (c.y >> 16);
        } else {
            j[0] = lo_slot(s, c.x, na);
            j[1] = hi_slot(s, c.x, na);
            j[2] = lo_slot(s, c.y, na);
            j[3] = hi_slot(s, c.y, na);
        }
        double2 p[4];
#pragma unroll
        for (int e = 0; e < 4; e++) p[e] = load_slot(pb, GHOST ? j[e] + kGhost : j[e]);
        if constexpr (WRAPIN) {  // after all four loads are in flight
#pragma unroll
            for (int e = 0; e < 4; e++) p[e] = state_pos(P, p[e]);
        }
#pragma unroll
        for (int e = 0; e < 4; e++) nl_pair<POT, WRAP, STALE>(P, p[e], j[e], s, xi, yi, test, ri, ci, kcur, acc);
    }
}

// The columns [ca, cb] of one grid row, as one or two runs of consecutive
// cells (a periodic grid's stencil wraps at the edges); every column at most
// once, also on grids of fewer than 3 columns.
struct ColRuns {
    int n, a0, b0, a1, b1;
};
__device__ __forceinline__ ColRuns stencil_cols(const SysParams &G, int col) {
    ColRuns R{1, 0, 0, 0, 0};
    if (!G.periodic) {
        R.a0 = max(col - 1, 0);
        R.b0 = min(col + 1, G.cols - 1);
    } else if (G.cols <= 3) {
        R.b0 = G.cols - 1;
    } else if (col == 0) {
        R = ColRuns{2, 0, 1, G.cols - 1, G.cols - 1};
    } else if (col == G.cols - 1) {
        R = ColRuns{2, col - 1, col, 0, 0};
    } else {
        R.a0 = col - 1;
        R.b0 = col + 1;
    }
    return R;
}

// Rows of the 3x3 stencil around grid row `row`, each at most once (-1:
// outside a walled box).
__device__ __forceinline__ void stencil_rows(const SysParams &G, int row, int rr[3]) {
    for (int d = 0; d < 3; d++) {
        int r;
        if (G.periodic && G.rows <= 3) {
            r = d < G.rows ? d : -1;
        } else {
            r = row + d - 1;
            if (G.periodic) r = r < 0 ? r + G.rows : (r >= G.rows ? r - G.rows : r);
            else if (r < 0 || r >= G.rows) r = -1;
        }
        rr[d] = r;
    }
}

constexpr int kBuildCand = 128;  // stencil candidates one build wave holds (2 per lane)

// Bitonic sort, ascending, of the wave's 64 * H keys, element e = H * lane + h.
template <int H>
__device__ __forceinline__ void wave_bitonic(uint32_t (&key)[H]) {
    const int lane = (int)lane_id();
#pragma unroll
    for (int k = 2; k <= 64 * H; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (H == 2 && j == 1) {
                // partners in one lane
                const int e = 2 * lane;
                const bool asc = (e & k) == 0;
                const uint32_t lo = min(key[0], key[1]), hi = max(key[0], key[1]);
                key[0] = asc ? lo : hi;
                key[1] = asc ? hi : lo;
            } else {
                const int lj = H == 2 ? j >> 1 : j;
#pragma unroll
                for (int h = 0; h < H; h++) {
                    const uint32_t other = (uint32_t)__shfl_xor((int)key[h], lj, 64);
                    const int e = H * lane + h;
                    const bool keep_min = ((e & j) == 0) == ((e & k) == 0);
                    key[h] = keep_min ? min(key[h], other) : max(key[h], other);
                }
            }
        }
    }
}

constexpr int kBuildRows = 16;  // particles whose lists one build wave stages in LDS at a time
constexpr int kBuildWaves = 4;  // cells (waves) per build workgroup

// Wave-scope ordering of the wave's own LDS writes before its reads.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One cell's lists.  The wave's candidates (the cell's 3x3 stencil, at most
// kBuildCand, staged in LDS by candidate index, the cell's own particles
// among them at candidate ec0 + (i - t0); NEAR: periodic images already the
// nearest to every particle of the cell, or a walled box) are sorted by id;
// then, for each particle of the cell, every lane tests its candidates
// against the list radius and the passing ones, ranked by ballot in id order,
// go to the particle's LDS row as 16-bit offsets, the last group of 4 padded
// with offset 0 (the particle itself).  kBuildRows rows at a time are written
// out as whole groups (8 bytes, consecutive particles of a group row
// adjacent).
template <int H, bool NEAR>
__device__ __forceinline__ void build_cell(const SysParams &P, const int32_t *__restrict__ ID, int t0, int t1,
                                           int M, int ec0, const int32_t *w_slot, const double2 *w_pos,
                                           int16_t (*w_list)[kNlMax], int16_t *w_dump, int32_t *w_groups,
                                           int2 *__restrict__ nbr,
                                           int32_t *__restrict__ ncount, double rl2, double edge,
                                           uint32_t *__restrict__ status) {
    const int lane = (int)lane_id();
    const int na = P.na;
    uint32_t key[H];
#pragma unroll
    for (int h = 0; h < H; h++) {
        const int e = H * lane + h;
        key[h] = 0xffffffffu;
        if (e < M) key[h] = ((uint32_t)ID[w_slot[e]] << 7) | (uint32_t)e;
    }
    wave_bitonic<H>(key);
    // per candidate: its slot, the slot's offset from the cell's first slot
    // (modulo na, in [-na/2, na/2]) and position (no candidate: slot -1,
    // infinitely far)
    int js[H], d0[H];
    double2 pj[H];
#pragma unroll
    for (int h = 0; h < H; h++) {
        const bool ok = key[h] != 0xffffffffu;
        const int e = ok ? (int)(key[h] & 127u) : 0;
        js[h] = ok ? w_slot[e] : -1;
        d0[h] = slot_offset(t0, js[h], na);
        pj[h] = ok ? w_pos[e] : make_double2(INFINITY, INFINITY);
    }
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (int b0 = t0; b0 < t1; b0 += kBuildRows) {
        const int nb = min(kBuildRows, t1 - b0);
        int my_count = 0;  // lane p < nb: entries of particle b0 + p
        int ng_max = 0;    // groups of the longest staged list
        for (int p = 0; p < nb; p++) {
            const int i = b0 + p, q = i - t0;
            const double2 pi = w_pos[ec0 + q];
            bool pass[H];
            unsigned long long m[H];
#pragma unroll
            for (int h = 0; h < H; h++) {
                double dx = pi.x - pj[h].x, dy = pi.y - pj[h].y;
                if (!NEAR) {
                    dx = min_image(dx, P.L, P.halfL);
                    dy = min_image(dy, P.H, P.halfH);
                }
                pass[h] = js[h] != i && dx * dx + dy * dy <= rl2;
                m[h] = __ballot(pass[h]);
            }
            int cnt = 0, rank[H];
            if (H == 1) {
                rank[0] = __popcll(m[0] & lt);
                cnt = __popcll(m[0]);
            } else {
                rank[0] = __popcll(m[0] & lt) + __popcll(m[H - 1] & lt);
                rank[H - 1] = rank[0] + (pass[0] ? 1 : 0);
                cnt = __popcll(m[0]) + __popcll(m[H - 1]);
            }
            const bool over = cnt > kNlMax;
            const int cw = over ? kNlMax : cnt;
            // every lane writes (no branches): an entry at its rank, else its
            // dump slot; then a pad of the last group (offset 0, the particle
            // itself), else the dump slot
            bool far = false;
#pragma unroll
            for (int h = 0; h < H; h++) {
                // slot_offset(i, js, na): d0 - q is at most na / 2
                int32_t d = d0[h] - q;
                d = d < -(na / 2) ? d + na : d;
                const bool put = pass[h] && rank[h] < kNlMax;
                far = far || (put && (d > 32767 || d < -32768));
                *(put ? &w_list[p][rank[h]] : &w_dump[lane]) = (int16_t)d;
            }
            const int pad = (((cw + 3) >> 2) << 2) - cw;
            *(lane < pad ? &w_list[p][cw + lane] : &w_dump[lane]) = 0;
            if (__any(far) || over) {
                if (lane == 0) atomicOr(status, kStatusOverflow);
            }
            my_count = lane == p ? cw : my_count;
            ng_max = max(ng_max, (cw + 3) >> 2);
        }
        if (lane < nb) w_groups[lane] = (my_count + 3) >> 2;
        wave_lds_sync();
        if (lane < nb) {
            bool near_edge = false;
            if (P.periodic) {
                const double2 pi = w_pos[ec0 + (b0 - t0) + lane];
                const double ex = fmin(pi.x - P.blx, (P.blx + P.L) - pi.x);
                const double ey = fmin(pi.y - P.bly, (P.bly + P.H) - pi.y);
                near_edge = !(ex >= edge && ey >= edge);
            }
            ncount[b0 + lane] = (int32_t)((uint32_t)my_count | (near_edge ? kEdgeBit : 0u));
        }
        // group g of particle p from lane g * kBuildRows + p
        for (int l = lane; l < ng_max * kBuildRows; l += 64) {
            const int g = l / kBuildRows, p = l % kBuildRows;
            if (p < nb && g < w_groups[p]) nbr[(int64_t)g * na + b0 + p] = *(const int2 *)&w_list[p][4 * g];
        }
        wave_lds_sync();  // the rows are staged again for the next particles
    }
}

// Build: the state is sorted on the list grid G (cells >= range + skin, a
// row's cells consecutive slots, each cell in ascending id).  One wave per
// cell: the slot runs of the 3x3 cell stencil give the candidates (coalesced
// loads into LDS); more than kBuildCand of them flag an overflow (the lists
// switch off) and leave the cell's lists empty.  A particle within `edge` of
// a periodic edge sums with the minimum image (its partners may wrap before
// the next build).
__global__ __launch_bounds__(64 * kBuildWaves) void k_nl_build(SysParams G, SysParams P, const double2 *__restrict__ pos,
                                                  const int32_t *__restrict__ ID,
                                                  const int32_t *__restrict__ start, int2 *__restrict__ nbr,
                                                  int32_t *__restrict__ ncount, double rl2, double edge,
                                                  uint32_t *__restrict__ status) {
    __shared__ int32_t w_slot[kBuildWaves][kBuildCand];
    __shared__ double2 w_pos[kBuildWaves][kBuildCand];
    __shared__ __align__(16) int16_t w_list[kBuildWaves][kBuildRows][kNlMax];
    __shared__ int32_t w_groups[kBuildWaves][kBuildRows];
    __shared__ int16_t w_dump[kBuildWaves][64];
    // the cell (row blockIdx.y, column 4 blockIdx.x + the wave's index):
    // uniform, so the stencil is scalar work
    const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = (int)lane_id();
    const int row = (int)blockIdx.y, col = (int)blockIdx.x * kBuildWaves + wid;
    if (col >= G.cols) return;  // whole wave; no workgroup barrier below
    const int cell = row * G.cols + col;
    const int t0 = start[cell], t1 = start[cell + 1];
    if (t0 == t1) return;
    int rb[6] = {0, 0, 0, 0, 0, 0}, rn[6] = {0, 0, 0, 0, 0, 0}, M = 0, ec0 = 0;
    if (row >= 1 && row + 1 < G.rows && col >= 1 && col + 1 < G.cols) {
        // an interior cell: three runs of three cells, the cell itself in
        // the middle one
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const int32_t *sr = start + (row - 1 + d) * G.cols + (col - 1);
            rb[2 * d] = sr[0];
            rn[2 * d] = sr[3] - sr[0];
        }
        M = rn[0] + rn[2] + rn[4];
        ec0 = rn[0] + (t0 - rb[2]);
    } else {
        int rr[3];
        stencil_rows(G, row, rr);
        const ColRuns R = stencil_cols(G, col);
#pragma unroll
        for (int d = 0; d < 3; d++) {
#pragma unroll
            for (int run = 0; run < 2; run++) {
                const int k = 2 * d + run;
                if (rr[d] < 0 || run >= R.n) continue;
                const int ca = run ? R.a1 : R.a0, cb = run ? R.b1 : R.b0;
                rb[k] = start[rr[d] * G.cols + ca];
                rn[k] = start[rr[d] * G.cols + cb + 1] - rb[k];
                // the run holding the cell itself (every column and row of
                // the stencil at most once): its first particle's candidate
                // index
                if (t0 >= rb[k] && t0 < rb[k] + rn[k]) ec0 = M + (t0 - rb[k]);
                M += rn[k];
            }
        }
    }
    if (M > kBuildCand) {
        for (int i = t0 + lane; i < t1; i += 64) ncount[i] = 0;
        if (lane == 0) atomicOr(status, kStatusOverflow);
        return;
    }
    // On a periodic list grid of at least 7 x 7 cells the image of a
    // candidate nearest the cell's first particle is the nearest to every
    // particle of the cell (each is within a cell width of the first, each
    // candidate within two of it, the box at least 7 wide): the candidates
    // are staged as those images and the tests need no minimum image.
    const bool near = !P.periodic || (G.cols >= 7 && G.rows >= 7);
    const double2 ref = pos[t0];
    // candidate e -> slot, position (runs concatenated in order)
    for (int e = lane; e < M; e += 64) {
        int rest = e, j = 0;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            if (rest >= 0 && rest < rn[k]) j = rb[k] + rest;
            rest -= rn[k];
        }
        double2 pj = pos[j];
        if (P.periodic && near) {
            const double dx = pj.x - ref.x, dy = pj.y - ref.y;
            if (dx > P.halfL) pj.x -= P.L;
            else if (dx < -P.halfL) pj.x += P.L;
            if (dy > P.halfH) pj.y -= P.H;
            else if (dy < -P.halfH) pj.y += P.H;
        }
        w_slot[wid][e] = j;
        w_pos[wid][e] = pj;
    }
    wave_lds_sync();
#define MAVI_BUILD_CELL(H, NEAR)                                                                               \
    build_cell<H, NEAR>(P, ID, t0, t1, M, ec0, w_slot[wid], w_pos[wid], w_list[wid], w_dump[wid], w_groups[wid], nbr, \
                        ncount, rl2, edge, status)
    if (near) {
        if (M <= 64) MAVI_BUILD_CELL(1, true);
        else MAVI_BUILD_CELL(2, true);
    } else {
        if (M <= 64) MAVI_BUILD_CELL(1, false);
        else MAVI_BUILD_CELL(2, false);
    }
#undef MAVI_BUILD_CELL
}

// After a build: the reference chunk of every particle (K_n), chunk counts
// with the capacity check (one atomic per run of equal chunks in a wave), no
// marks, the build positions.
__global__ void k_nl_keys(SysParams P, const double2 *__restrict__ pos, int32_t *__restrict__ key,
                          int32_t *__restrict__ ccnt, uint8_t *__restrict__ chg0, uint8_t *__restrict__ chg1,
                          double2 *__restrict__ xb, uint32_t *__restrict__ flags) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = s < P.na;
    uint32_t flag = 0;
    int32_t r = 0, c = 0, k = -1;
    double2 p = make_double2(0.0, 0.0);
    if (valid) {
        p = pos[s];
        cell_rc(P, P.rows, P.cols, P.cl, P.ch, p.x, p.y, &flag, MAVI_FLAG_OUT_OF_BOX, &r, &c);
        k = r * P.cols + c;
    }
    const int32_t rank = wave_arrival_rank(valid, k, ccnt);
    if (valid && rank + 1 > P.nc) flag |= MAVI_FLAG_OVERFLOW;
    if (flag) atomicOr(flags, flag);
    if (!valid) return;
    key[s] = pack_rc(r, c);
    chg0[s] = 0;
    chg1[s] = 0;
    xb[s] = p;
}

// The drift of one particle (update_verlet!, src/integration.jl:424) from
// start position pi with velocity v and force f into pout, the reference
// chunk of its final position (after walls!'s periodic wrap) into knew and,
// if it differs from kold (the chunks the next second evaluation keeps), the
// chunk counts and the marks in chg of the pairs that evaluation could treat
// differently from the first evaluation after it: a list partner whose kold
// chunk is not adjacent to this particle's (1) and then the particle (2);
// all partners if the wrap moved the particle (the first evaluation sums the
// wrapped positions).
__device__ __forceinline__ double nl_drift(const SysParams &P, const NLArgs &A, int s, double2 pi, double2 v,
                                           double fx, double fy, double2 *__restrict__ pout,
                                           const int32_t *__restrict__ kold, int32_t *__restrict__ knew,
                                           uint8_t *__restrict__ chg) {
    const double px = pi.x + (v.x * P.dt + fx * P.term), py = pi.y + (v.y * P.dt + fy * P.term);
    pout[s] = make_double2(px, py);
    // the copies the list sums read across the ends of the slot range
    if (s < kGhost) pout[s + P.na] = make_double2(px, py);
    if (s >= P.na - kGhost) pout[s - P.na] = make_double2(px, py);
    double wx = px, wy = py;
    if (P.periodic) periodic_wrap(P, wx, wy);  // walls! moves it there after the kick
    const int32_t ko = kold[s];
    // still clear inside its chunk (by far more than the rounding of the
    // chunk edges): the same chunk, without the exact division
    const int r0 = ko >> 16, c0 = ko & 0xffff;
    const double ax = wx - P.blx, ay = (-wy + P.bly) + P.H;
    const double m = 1e-12 * (P.L + P.H);
    int32_t kn = ko;
    if (!(ax >= (double)c0 * P.cl + m && ax <= (double)(c0 + 1) * P.cl - m && ay >= (double)r0 * P.ch + m &&
          ay <= (double)(r0 + 1) * P.ch - m)) {
        uint32_t flag = 0;
        int32_t r, c;
        cell_rc(P, P.rows, P.cols, P.cl, P.ch, wx, wy, &flag, MAVI_FLAG_OUT_OF_BOX, &r, &c);
        if (flag) atomicOr(A.flags, flag);
        kn = pack_rc(r, c);
    }
    knew[s] = kn;
    if (kn != ko) {
        const bool wrapped = wx != px || wy != py;
        bool any = false;
        const int groups = (int)(((uint32_t)A.ncount[s] & ~kEdgeBit) + 3u) >> 2;
        for (int g = 0; g < groups; g++) {
            const int2 q = A.nbr[g * P.na + s];
            const int j4[4] = {lo_slot(s, q.x, P.na), hi_slot(s, q.x, P.na), lo_slot(s, q.y, P.na),
                               hi_slot(s, q.y, P.na)};
            for (int e = 0; e < 4; e++) {
                const int j = j4[e];
                if (j == s) continue;  // list pad
                if (wrapped || !chunks_adjacent(P, r0, c0, kold[j])) {
                    mark(chg, j, 1u);
                    any = true;
                }
            }
        }
        if (any) mark(chg, s, 2u);
        atomicSub(&A.ccnt[r0 * P.cols + c0], 1);
        // while one step's moves are applied a count runs through values at
        // least as large as its final one
        if (atomicAdd(&A.ccnt[(kn >> 16) * P.cols + (kn & 0xffff)], 1) + 1 > P.nc)
            atomicOr(A.status, kStatusCapacity);
    }
    const double lx = px - pi.x, ly = py - pi.y;  // the step's length, squared
    return lx * lx + ly * ly;
}

// max of v over the wave into slot[blockIdx.x] (kept across launches: the
// larger of the stored and the new value; an atomic per wave, so a wave
// retires without waiting for the rest of its workgroup; squared lengths are
// non-negative doubles, ordered as their bit patterns).
__device__ __forceinline__ void block_max_into(double v, double *__restrict__ slot) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if (lane_id() == 0)
        atomicMax((unsigned long long *)&slot[blockIdx.x], (unsigned long long)__double_as_longlong(v));
}

// After a build: the first evaluation of newton_step! (forces at x_n over
// every pair within range) and the drift, for every particle; the largest
// step per workgroup.
template <int POT>
__global__ __launch_bounds__(kNeedBlock) void k_nl_start(SysParams P, NLArgs A) {
    const int s = xcd_block() * blockDim.x + threadIdx.x;
    const bool valid = s < P.na;
    const int sv = valid ? s : 0;
    const double2 pi = A.Pst[sv];
    const uint32_t m = valid ? (uint32_t)A.ncount[s] : 0u;
    const int groups = (int)((m & ~kEdgeBit) + 3u) >> 2;
    PairAcc<POT> acc;
    if (__any((m & kEdgeBit) != 0u))
        nl_sum<POT, true, false>(P, A.Pst, A.nbr, sv, groups, pi.x, pi.y, false, 0, 0, nullptr, acc);
    else
        nl_sum<POT, false, false>(P, A.Pst, A.nbr, sv, groups, pi.x, pi.y, false, 0, 0, nullptr, acc);
    double l2 = 0.0;
    if (valid) {
        A.F[s] = make_double2(acc.fx, acc.fy);
        A.fk_on[s] = 0;
        l2 = nl_drift(P, A, s, pi, A.V[s], acc.fx, acc.fy, A.Px, A.kcur, A.knext, A.chg);
    }
    block_max_into(l2, A.bmax);
}

// Second evaluation at the drifted positions with the step-start chunks,
// then the kick and walls! (src/integration.jl:425-431, 271-285, 309-324),
// then the next step: its first evaluation (this one, or for the marked
// particles every pair within range at the wrapped positions) and its drift.
// Records the largest drift step.
template <int POT>
__global__ __launch_bounds__(kNeedBlock) void k_nl_second(SysParams P, NLArgs A) {
    const int s = xcd_block() * blockDim.x + threadIdx.x;
    const bool valid = s < P.na;
    const int sv = valid ? s : 0;
    const double2 pu = A.Px[sv];
    const double2 pw = state_pos(P, pu);
    const uint32_t c = valid ? A.chg[s] : 0u;
    const uint32_t m = valid ? (uint32_t)A.ncount[s] : 0u;
    const int groups = (int)((m & ~kEdgeBit) + 3u) >> 2;
    const bool stale = c != 0u;
    const bool fresh = stale || !A.reuse;  // the next first evaluation is summed afresh
    const bool wrap = __any((m & kEdgeBit) != 0u);
    PairAcc<POT> acc, acc1;
    if (__any(fresh)) {
        const int32_t ki = A.kcur[sv];
        const int ri = ki >> 16, ci = ki & 0xffff;
        const int g1 = fresh ? groups : 0;
        if (wrap) {
            nl_sum<POT, true, true, false, true>(P, A.Px, A.nbr, sv, groups, pu.x, pu.y, stale, ri, ci, A.kcur, acc);
            nl_sum<POT, true, false, true, true>(P, A.Px, A.nbr, sv, g1, pw.x, pw.y, false, 0, 0, nullptr, acc1);
        } else {
            nl_sum<POT, false, true, false, true>(P, A.Px, A.nbr, sv, groups, pu.x, pu.y, stale, ri, ci, A.kcur,
                                                  acc);
            nl_sum<POT, false, false, true, true>(P, A.Px, A.nbr, sv, g1, pw.x, pw.y, false, 0, 0, nullptr, acc1);
        }
    } else {
        if (wrap)
            nl_sum<POT, true, false, false, true>(P, A.Px, A.nbr, sv, groups, pu.x, pu.y, false, 0, 0, nullptr, acc);
        else
            nl_sum<POT, false, false, false, true>(P, A.Px, A.nbr, sv, groups, pu.x, pu.y, false, 0, 0, nullptr, acc);
    }
    double l2 = 0.0;
    if (valid) {
        const double fx = acc.fx, fy = acc.fy;                     // this second evaluation
        const double gx = fresh ? acc1.fx : fx, gy = fresh ? acc1.fy : fy;  // the next first one
        const double2 v = A.V[s];
        const double2 fo = A.fk_on[s] ? A.Fk[s] : A.F[s];           // this step's first one
        // vel += dt/2 * (forces + old_forces) (src/integration.jl:430), then walls!
        double vx = v.x + P.hdt * (fx + fo.x);
        double vy = v.y + P.hdt * (fy + fo.y);
        const double x = pw.x, y = pw.y;
        if (!P.periodic) {
            // src/integration.jl:271-285
            const double rx = x - P.blx, ry = y - P.bly;
            const bool ox = ((rx + P.radius) > P.L) || ((rx - P.radius) < 0.0);
            const bool oy = ((ry + P.radius) > P.H) || ((ry - P.radius) < 0.0);
            if (ox) vx = -vx;
            if (oy) vy = -vy;
        }
        A.V[s] = make_double2(vx, vy);
        A.F[s] = make_double2(fx, fy);
        if (fresh) A.Fk[s] = make_double2(gx, gy);
        A.fk_on[s] = fresh ? 1 : 0;
        A.chg[s] = 0;
        l2 = nl_drift(P, A, s, pw, make_double2(vx, vy), gx, gy, A.Px_next, A.knext, A.knn, A.chg_next);
    }
    // the largest step since the block began, per workgroup slot
    block_max_into(l2, A.bmax);
}

// out[0] <- max of bmax, out[1] <- max of dpend (the doubles' bits), out[2]
// <- the status word (one workgroup of 1024 threads).
__global__ __launch_bounds__(1024) void k_nl_reduce(const double *__restrict__ bmax, const double *__restrict__ dpend,
                                                    int32_t n, const uint32_t *__restrict__ status,
                                                    unsigned long long *__restrict__ out) {
    __shared__ double s_max[2][16];
    double m[2] = {0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += 1024) {
        m[0] = fmax(m[0], bmax[i]);
        m[1] = fmax(m[1], dpend[i]);
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m[k] = fmax(m[k], __shfl_xor(m[k], o, 64));
        if (lane_id() == 0) s_max[k][threadIdx.x >> 6] = m[k];
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int k = threadIdx.x;
        double v = s_max[k][0];
        for (int w = 1; w < 16; w++) v = fmax(v, s_max[k][w]);
        out[k] = (unsigned long long)__double_as_longlong(v);
    }
    if (threadIdx.x == 2) out[2] = *status;
}

// At a block's start: the step maxima cleared (not after a build, whose drift
// counts among the block's steps) and the flags word kept for a rollback.
__global__ void k_nl_block_start(double *__restrict__ bmax, int32_t n, int32_t clear,
                                 const uint32_t *__restrict__ flags, uint32_t *__restrict__ fsnap) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (clear && i < n) bmax[i] = 0.0;
    if (i == 0) *fsnap = *flags;
}

// At a block's end, per workgroup: the largest squared displacement of the
// next positions (pending, not yet evaluated) from the build positions.
__global__ __launch_bounds__(kNeedBlock) void k_nl_measure(SysParams P, const double2 *__restrict__ pending,
                                                          const double2 *__restrict__ xb,
                                                          double *__restrict__ dpend) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0;
    if (s < P.na) {
        const double2 q = state_pos(P, pending[s]), b = xb[s];
        double dx = q.x - b.x, dy = q.y - b.y;
        if (P.periodic) {
            dx = min_image(dx, P.L, P.halfL);
            dy = min_image(dy, P.H, P.halfH);
        }
        d2 = dx * dx + dy * dy;
    }
    __shared__ double s_max[kNeedBlock / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) d2 = fmax(d2, __shfl_xor(d2, o, 64));
    if (lane_id() == 0) s_max[threadIdx.x >> 6] = d2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = s_max[0];
        for (int w = 1; w < kNeedBlock / 64; w++) mx = fmax(mx, s_max[w]);
        dpend[blockIdx.x] = mx;
    }
}

// After an accepted block: the state's positions, the wrap of the last
// evaluated drift px.
__global__ void k_nl_settle(SysParams P, const double2 *__restrict__ px, double2 *__restrict__ pst) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < P.na) pst[s] = state_pos(P, px[s]);
}

NLArgs nl_args(mavi_sys *s) {
    NList &N = *s->nl;
    const int n = N.ph;
    NLArgs A{};
    A.nbr = (const int2 *)N.nbr;
    A.ncount = N.ncount;
    A.kcur = N.key[n % 3];
    A.knext = N.key[(n + 1) % 3];
    A.knn = N.key[(n + 2) % 3];
    A.ccnt = N.ccnt;
    A.chg = N.chg[n % 2];
    A.chg_next = N.chg[(n + 1) % 2];
    A.bmax = N.bmax;
    A.Pst = s->P_[s->cur];
    A.V = s->V_[s->cur];
    A.F = s->F;
    A.Fk = N.fk;
    A.fk_on = N.fk_on;
    A.Px = N.px[n % 2];
    A.Px_next = N.px[(n + 1) % 2];
    A.flags = s->flags;
    A.status = N.status;
    A.reuse = s->reuse ? 1 : 0;
    return A;
}

double nl_range(const mavi_sys *s) {
    const SysParams &P = s->P;
    if (P.potential == MAVI_POT_LJ && P.cut2 < INFINITY && P.p2 > 0.0) return P.p2;
    if (P.potential == MAVI_POT_HARMTRUNC) return P.p3;
    return 0.0;
}

}  // namespace

namespace mavi {

bool nl_eligible(const mavi_sys *s) {
    const SysParams &P = s->P;
    const double range = nl_range(s);
    if (!(range > 0.0) || P.nwalls > 0 || s->var || P.na <= 0) return false;
    // the build sorts 32-bit keys id << 7 | candidate
    if (P.na >= (1 << 25)) return false;
    if (P.cols < 3 || P.rows < 3) return false;
    // every pair within range is in adjacent chunks at the binned positions
    return P.cl >= range * (1.0 + 1e-12) && P.ch >= range * (1.0 + 1e-12);
}

void nl_set_skin(mavi_sys *s, double skin_frac) {
    NList &N = *s->nl;
    const SysParams &P = s->P;
    N.skin_frac = skin_frac;
    const double skin = skin_frac * N.range;
    const double rl = (N.range + skin) * (1.0 + 1e-9);
    N.rl2 = rl * rl;
    const double half = 0.5 * skin * (1.0 - 1e-9);
    N.thr2 = half * half;
    N.edge = rl + 2.0 * skin;
    // list grid: cells at least the list radius wide (the reference grid when
    // its chunks already are)
    SysParams G = P;
    G.cols = P.cl >= rl ? P.cols : (int32_t)floor(P.L / rl);
    G.rows = P.ch >= rl ? P.rows : (int32_t)floor(P.H / rl);
    if (G.cols < 1) G.cols = 1;
    if (G.rows < 1) G.rows = 1;
    G.cl = P.L / (double)G.cols;
    G.ch = P.H / (double)G.rows;
    G.ncell = G.cols * G.rows;
    G.nc = P.na;  // no capacity check on the list grid
    N.G = G;
    N.valid = false;
}

hipError_t nl_alloc(mavi_sys *s, hipStream_t st) {
    s->nl = new NList();
    NList &N = *s->nl;
    N.range = nl_range(s);
    nl_set_skin(s, N.skin_frac);
    const int64_t na = s->P.na;
    N.kmax = kNlMax;
    hipError_t e = dalloc(&N.nbr, (int64_t)(N.kmax / 2) * na, st);  // 16-bit entries
    if (e == hipSuccess) e = dalloc(&N.ncount, na, st);
    for (int k = 0; k < 3 && e == hipSuccess; k++) e = dalloc(&N.key[k], na, st);
    for (int k = 0; k < 2 && e == hipSuccess; k++) e = dalloc(&N.chg[k], (na + 3) & ~int64_t(3), st);
    // drifted positions with kGhost copies on each side (nl_sum's GHOST form)
    for (int k = 0; k < 2 && e == hipSuccess; k++) {
        e = dalloc(&N.px[k], na + 2 * (int64_t)kGhost, st);
        if (e == hipSuccess) N.px[k] += kGhost;
    }
    if (e == hipSuccess) e = dalloc(&N.fk, na, st);
    if (e == hipSuccess) e = dalloc(&N.fk_on, na, st);
    if (e == hipSuccess) e = dalloc(&N.ccnt, (int64_t)s->P.ncell, st);
    if (e == hipSuccess) e = dalloc(&N.xb, na, st);
    if (e == hipSuccess) e = dalloc(&N.vsnap, na, st);
    if (e == hipSuccess) e = dalloc(&N.fsnap, 1, st);
    if (e == hipSuccess) e = dalloc(&N.dmax, 3, st);
    N.nbmax = nblk(s->P.na, kNeedBlock);  // workgroups of the second evaluation
    if (e == hipSuccess) e = dalloc(&N.bmax, N.nbmax, st);
    if (e == hipSuccess) e = dalloc(&N.dpend, N.nbmax, st);
    if (e == hipSuccess) e = dalloc(&N.status, 1, st);
    if (e == hipSuccess) e = hipHostMalloc((void **)&N.host, 3 * sizeof(unsigned long long), hipHostMallocDefault);
    // scan states of the list grid (no more cells than the chunk grid)
    N.scan_words = (s->P.ncell + 1 + kScanTileMin - 1) / kScanTileMin + 1;
    if (e == hipSuccess) e = dalloc(&N.scan_ticket, 1, st);
    if (e == hipSuccess) e = dalloc(&N.scan_state, N.scan_words, st);
    return e;
}

void nl_free(mavi_sys *s) {
    if (!s->nl) return;
    NList &N = *s->nl;
    for (auto &g : N.graph)
        if (g) (void)hipGraphExecDestroy(g);
    dfree(N.nbr); dfree(N.ncount); dfree(N.ccnt);
    for (auto &k : N.key) dfree(k);
    for (auto &c : N.chg) dfree(c);
    for (auto &p : N.px) {
        if (!p) continue;
        double2 *base = p - kGhost;  // allocated with kGhost copies before the slots
        dfree(base);
        p = nullptr;
    }
    dfree(N.fk); dfree(N.fk_on); dfree(N.xb); dfree(N.vsnap); dfree(N.fsnap);
    dfree(N.dmax); dfree(N.bmax); dfree(N.dpend); dfree(N.status); dfree(N.scan_ticket); dfree(N.scan_state);
    if (N.host) (void)hipHostFree(N.host);
    delete s->nl;
    s->nl = nullptr;
}

hipError_t nl_rebuild(mavi_sys *s, hipStream_t st) {
    NList &N = *s->nl;
    hipError_t e = drop_keys(s, st);  // counts a fused first pass left on the chunk grid
    // the list grid's scan starts from a cleared ticket and states (its tile
    // count changes with the skin)
    if (e == hipSuccess) e = launch_fill_i32((int32_t *)N.scan_ticket, 0, 2, st);
    if (e == hipSuccess) e = launch_fill_i32((int32_t *)N.scan_state, 0, 2 * (int64_t)N.scan_words, st);
    if (e == hipSuccess) e = launch_bin(s, st, false, false, false, &N.G, N.scan_ticket, N.scan_state);
    if (e != hipSuccess) return e;
    s->sorted = false;   // not in reference chunk order: the chunk gather's stamps start over
    s->f_valid = false;
    const SysParams &P = s->P;
    const int c = s->cur;
    {
        ProfScope ps(s, MAVI_PROF_BIN, st);
        hipLaunchKernelGGL(k_nl_build, dim3(nblk(N.G.cols, kBuildWaves), N.G.rows), dim3(64 * kBuildWaves), 0, st, N.G, P, s->P_[c], s->ID_[c],
                           s->start, (int2 *)N.nbr, N.ncount, N.rl2, N.edge, N.status);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = launch_fill_i32(N.ccnt, 0, P.ncell, st)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_nl_keys, dim3(nblk(P.na, 256)), dim3(256), 0, st, P, s->P_[c], N.key[0], N.ccnt,
                           N.chg[0], N.chg[1], N.xb, s->flags);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = launch_fill_i32((int32_t *)N.bmax, 0, 2 * (int64_t)N.nbmax, st)) != hipSuccess) return e;
    }
    // the first evaluation and drift of the first step
    N.ph = 0;
    {
        ProfScope ps(s, MAVI_PROF_FORCE_REUSE, st);
        const NLArgs A = nl_args(s);
        const int nb = nblk(P.na, kNeedBlock);
        if (P.potential == MAVI_POT_LJ) hipLaunchKernelGGL(k_nl_start<POT_LJ>, dim3(nb), dim3(kNeedBlock), 0, st, P, A);
        else hipLaunchKernelGGL(k_nl_start<POT_HT>, dim3(nb), dim3(kNeedBlock), 0, st, P, A);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    N.unsettled = false;
    N.valid = true;
    N.d_start = 0.0;  // the first block's positions: the build's plus the drift steps
    N.rebuilds++;
    return hipSuccess;
}

hipError_t nl_block_start(mavi_sys *s, hipStream_t st, bool clear) {
    NList &N = *s->nl;
    hipLaunchKernelGGL(k_nl_block_start, dim3(nblk(N.nbmax, 256)), dim3(256), 0, st, N.bmax, N.nbmax,
                       clear ? 1 : 0, s->flags, N.fsnap);
    return hipGetLastError();
}

hipError_t nl_measure(mavi_sys *s, hipStream_t st) {
    NList &N = *s->nl;
    // the last step drifted into px[(ph - 1) % 2] what it evaluated and into
    // px[ph % 2] the next positions
    hipLaunchKernelGGL(k_nl_measure, dim3(N.nbmax), dim3(kNeedBlock), 0, st, s->P, N.px[N.ph % 2], N.xb, N.dpend);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_nl_reduce, dim3(1), dim3(1024), 0, st, N.bmax, N.dpend, N.nbmax, N.status, N.dmax);
    return hipGetLastError();
}

hipError_t nl_settle(mavi_sys *s, hipStream_t st) {
    NList &N = *s->nl;
    if (!N.unsettled) return hipSuccess;
    hipLaunchKernelGGL(k_nl_settle, dim3(nblk(s->P.na, 256)), dim3(256), 0, st, s->P, N.px[(N.ph + 1) % 2],
                       s->P_[s->cur]);
    N.unsettled = false;
    return hipGetLastError();
}

hipError_t nl_step(mavi_sys *s, hipStream_t st) {
    NList &N = *s->nl;
    const SysParams &P = s->P;
    const NLArgs A = nl_args(s);
    const int nb = nblk(P.na, kNeedBlock);
    {
        ProfScope ps(s, MAVI_PROF_FORCE, st);
        if (P.potential == MAVI_POT_LJ) hipLaunchKernelGGL(k_nl_second<POT_LJ>, dim3(nb), dim3(kNeedBlock), 0, st, P, A);
        else hipLaunchKernelGGL(k_nl_second<POT_HT>, dim3(nb), dim3(kNeedBlock), 0, st, P, A);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    N.ph = (N.ph + 1) % 6;
    N.unsettled = true;
    return hipSuccess;
}

}  // namespace mavi
