// mavi_kernels.hip -- gfx950 kernels of the step hot path.
//
//   K1 binning      k_bin_count (wave-aggregated chunk counts) ->
//                   k_scan_lookback (one-pass scan) -> k_scatter -> k_finalize
//                   (ids ascending inside a chunk, as update_chunks! appends
//                   them, src/chunks.jl:150-161; the state is permuted into
//                   the new chunk order)
//   K2 pair forces  k_pair<MODE, POT>: each particle gathers over the 3x3
//                   chunk block around its chunk.  The stencil of
//                   src/chunks.jl:61-118 visits each adjacent chunk pair once
//                   and applies f and -f; the gather visits every pair from
//                   both sides, so the interaction sets are equal for grids
//                   >= 3x3 (src/integration.jl:112-157).
//   K3 Verlet+walls MODE_VERLET1/2 epilogues of k_pair (src/integration.jl:
//                   415-431 with walls! :271-285 / :309-324).
//   K4 rings        k_rings_lanes<UPDATE> (floor(64 / nmax) rings of nmax
//                   lanes per wave, rings of up to 32 particles) and
//                   k_rings<UPDATE> (one ring per
//                   thread, any size): unwrap, shoelace area, springs, area
//                   force and overdamped update (src/rings/integration.jl:79-351).
//   K7 observables  k_observables (src/quantities.jl:12-18, 46-66).
//
// The library is compiled with -ffp-contract=off: a*b+c is only fused where
// the code calls __builtin_fma, so the arithmetic that follows Julia's
// rounding (positions, velocities, HarmTrunc forces, ring updates) does.
#include "mavi_device.h"
#include "mavi_host.h"

namespace {

// ---------------------------------------------------------------------------
// K1 binning
// ---------------------------------------------------------------------------

// Chunk of every particle and its arrival rank in that chunk.  With stamp:
// a particle whose chunk differs from the previous binning's (prev_cell,
// indexed by the current slot) stamps both chunks with this binning's epoch.
__global__ void k_bin_count(SysParams P, const double2 *__restrict__ pos, int32_t *__restrict__ key,
                            int32_t *__restrict__ rank, int32_t *__restrict__ cnt, uint32_t *__restrict__ flags,
                            const int32_t *__restrict__ prev_cell, int32_t *__restrict__ stamp,
                            const uint32_t *__restrict__ epoch) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = s < P.na;
    uint32_t flag = 0;
    int32_t k = -1;
    if (valid) {
        const double2 p = pos[s];
        k = cell_of(P, p.x, p.y, &flag);
    }
    if (flag) atomicOr(flags, flag);
    if (stamp && valid) {
        const int32_t pc = prev_cell[s];
        if (pc != k) {
            const int32_t e = (int32_t)(epoch[0] + 1u);
            stamp[k] = e;
            stamp[pc] = e;
        }
    }
    const int32_t rk = wave_arrival_rank(valid, k, cnt);
    if (valid) {
        key[s] = k;
        rank[s] = rk;
    }
}

// One thread per new slot: the final slot inside the chunk is the number of
// chunk members with a smaller id (update_chunks! appends ids ascending);
// move the particle's position and id there, and its velocity (Vnew) or
// its old slot (perm: the next pair launch moves velocity and forces).
__global__ void k_finalize(SysParams P, const int32_t *__restrict__ start, const int32_t *__restrict__ src,
                           const int32_t *__restrict__ scell, const int32_t *__restrict__ sid,
                           const double2 *__restrict__ Pold,
                           const double2 *__restrict__ Vold, const int32_t *__restrict__ IDold,
                           double2 *__restrict__ Pnew, double2 *__restrict__ Vnew, int32_t *__restrict__ IDnew,
                           int32_t *__restrict__ slot_of, uint32_t *__restrict__ flags,
                           int32_t *__restrict__ perm, uint32_t *__restrict__ epoch) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) epoch[0] += 1u;  // binnings so far
    if (t >= P.na) return;
    const int c = scell[t];
    const int b = start[c], e = start[c + 1];
    const int so = src[t];
    const int myid = sid[t];
    int r = 0;
    for (int q = b; q < e; q++) r += (sid[q] < myid) ? 1 : 0;
    const int t2 = b + r;
    Pnew[t2] = Pold[so];
    if (Vnew) Vnew[t2] = Vold[so];
    if (perm) perm[t2] = so;
    IDnew[t2] = myid;
    if (slot_of) slot_of[myid] = t2;
    if (t == b && e - b > P.nc) atomicOr(flags, MAVI_FLAG_OVERFLOW);
}

// chunk_particles (nc, rows, cols) and num_particles_in_chunk (rows, cols),
// column-major, 1-based: the reference layout of src/chunks.jl:23-24.
__global__ void k_export_chunks(SysParams P, const int32_t *__restrict__ start, const int32_t *__restrict__ ID,
                                int64_t *__restrict__ cp, int64_t *__restrict__ counts) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.ncell) return;
    int row0 = c / P.cols, col0 = c - row0 * P.cols;
    int64_t k = (int64_t)col0 * P.rows + row0;
    int b = start[c], m = start[c + 1] - b;
    counts[k] = m;
    int lim = m < P.nc ? m : P.nc;
    for (int q = 0; q < lim; q++) cp[(int64_t)P.nc * k + q] = (int64_t)ID[b + q] + 1;
}

// ---------------------------------------------------------------------------
// K2/K3 pair forces with fused integrator epilogues
// ---------------------------------------------------------------------------

// With perm (first evaluation of a Verlet step after a binning that left
// velocities in the old order): velocities move to the new order here,
// Vst[s] = Vold[perm[s]].  With needc (needs perm): a particle whose chunk
// is not marked in this binning's epoch (no chunk of its 3x3 block changed
// members) takes its forces from Fold[perm[s]], the previous step's second
// evaluation.  Those forces were computed at the same positions over the
// same candidates in the same order, so they are bit-identical to a
// recomputation.  (A compacted worklist of the particles to recompute was
// slower: it loses the sorted order's neighbour locality.)
// The next binning's first pass (k_bin_count) done by the second
// evaluation on the final positions it writes: chunk, arrival rank, flags,
// and the force-reuse stamps of the particles that change chunk.
struct NextBin {
    int32_t *key, *rank, *cnt;
    uint32_t *flags;
    int32_t *stamp;
};

template <int MODE, int POT, bool ALLPAIRS, bool WALLS = false>
__global__ __attribute__((amdgpu_waves_per_eu(WALLS ? 1 : 8))) void k_pair(SysParams P, RingCtx R, const double2 *__restrict__ pin, const int32_t *__restrict__ scell,
                       const int32_t *__restrict__ start, double2 *__restrict__ Pst, double2 *__restrict__ Vst,
                       double2 *__restrict__ Fout, double2 *__restrict__ P2, double2 *__restrict__ F1,
                       double list2, const int32_t *__restrict__ needc, const uint32_t *__restrict__ epoch,
                       const int32_t *__restrict__ perm, const double2 *__restrict__ Vold,
                       const double2 *__restrict__ Fold, int32_t *__restrict__ wstamp, NextBin nb) {
    extern __shared__ uint16_t s_list[];
    int s = xcd_block() * blockDim.x + threadIdx.x;
    bool valid = s < P.na;
    int sv = valid ? s : 0;
    double2 pi = pin[sv];
    if constexpr (MODE == MODE_VERLET1 && !ALLPAIRS) {
        const int so = (perm && valid) ? perm[s] : 0;
        if (perm && valid) Vst[s] = Vold[so];
        if (needc) {
            // chunks whose 3x3 block changed members in this binning
            const bool need = valid && needc[scell[sv]] == (int32_t)epoch[0];
            if (valid && !need) {
                const double2 f = Fold[so];
                pair_epilogue<MODE, WALLS>(P, s, pi, f.x, f.y, Pst, Vst, Fout, P2, F1);
            }
            // the block's particles to recompute go to its first lanes (a
            // particle's forces do not depend on the lane or the wave it is
            // computed in); waves left without one exit
            __shared__ int32_t s_redo[1024];
            __shared__ int32_t s_nredo;
            if (threadIdx.x == 0) s_nredo = 0;
            __syncthreads();
            const unsigned long long m = __ballot(need);
            if (m) {
                const int leader = __ffsll(m) - 1;
                int base = 0;
                if ((int)lane_id() == leader) base = atomicAdd(&s_nredo, __popcll(m));
                base = __shfl(base, leader, 64);
                if (need) s_redo[base + __popcll(m & ((1ull << lane_id()) - 1ull))] = s;
            }
            __syncthreads();
            const int nredo = s_nredo;
            if ((int)(threadIdx.x & ~63u) >= nredo) return;  // whole wave
            valid = (int)threadIdx.x < nredo;
            s = s_redo[valid ? threadIdx.x : 0];
            sv = s;
            pi = pin[sv];
        }
    }
    PairAcc<POT> acc;
    if constexpr (POT == POT_RINGS) {
        acc.id_i = R.ids[sv] - R.id_base;
        acc.ring_i = ring_of((uint32_t)acc.id_i, R);
        acc.type_i = P.typed ? R.rtype[acc.ring_i] : 0;
    }
    if constexpr (ALLPAIRS) {
        if (valid) scan_range<POT>(P, R, acc, pin, s, pi.x, pi.y, P.periodic != 0, 0, P.na);
    } else {
        pair_gather_chunks<POT>(P, R, acc, pin, scell[sv], start, s, valid, pi, list2, s_list);
    }
    if constexpr (MODE == MODE_VERLET2 && !ALLPAIRS) {
        double2 xf = make_double2(0.0, 0.0);
        const bool moved = valid && pair_epilogue<MODE, WALLS>(P, s, pi, acc.fx, acc.fy, Pst, Vst, Fout, P2, F1, &xf);
        if constexpr (WALLS) {
            // forces of a particle walls! moved cannot be reused: mark its
            // chunk for the next binning (epoch + 1), as a chunk change would
            if (moved && wstamp) wstamp[scell[s]] = (int32_t)(epoch[0] + 1u);
        }
        if (nb.key) {  // every lane of the wave takes part in the arrival ranks
            uint32_t flag = 0;
            int32_t k = -1;
            if (valid) k = cell_of(P, xf.x, xf.y, &flag);
            if (flag) atomicOr(nb.flags, flag);
            if (nb.stamp && valid) {
                const int32_t pc = scell[s];
                if (pc != k) {
                    const int32_t e = (int32_t)(epoch[0] + 1u);
                    nb.stamp[k] = e;
                    nb.stamp[pc] = e;
                }
            }
            const int32_t rk = wave_arrival_rank(valid, k, nb.cnt);
            if (valid) {
                nb.key[s] = k;
                nb.rank[s] = rk;
            }
        }
        return;
    }
    if (!valid) return;
    pair_epilogue<MODE, WALLS>(P, s, pi, acc.fx, acc.fy, Pst, Vst, Fout, P2, F1);
}

// ---------------------------------------------------------------------------
// Contact neighbours of rings: ParticleNeighbors (src/rings/neighbors.jl:
// 117-137) as calc_interaction fills them (src/rings/integration.jl:32-45)
// in forces!.  One thread per slot gathers over its 3x3 chunk block (all
// slots without chunks): the particles closer than tol * dist_eq of the
// pair's interaction (of another ring only, with only_rings), as 1-based ids
// in ascending order, -1 padded.  More than 15 raises a flag (the reference's
// BoundsError).  Runs between the pair kernel and the update, on the
// positions forces! sees.
// ---------------------------------------------------------------------------
constexpr int kNeighMax = 15;

__global__ void k_ring_neigh(SysParams P, RingCtx R, const double2 *__restrict__ pin,
                             const int32_t *__restrict__ start, const int32_t *__restrict__ scell, int only_rings,
                             int32_t *__restrict__ cnt_out, int32_t *__restrict__ list_out,
                             uint32_t *__restrict__ flags) {
    const int s = xcd_block() * blockDim.x + threadIdx.x;
    if (s >= P.na) return;
    const int id = R.ids[s] - R.id_base;
    const int ri = ring_of((uint32_t)id, R);
    const int ti = P.typed ? R.rtype[ri] : 0;
    const double2 pi = pin[s];
    int list[kNeighMax];
    int cnt = 0;
    auto visit = [&](int a, int b) {
        for (int j = a; j < b; j++) {
            if (j == s) continue;
            const int idj = R.ids[j] - R.id_base;
            const int rj = ring_of((uint32_t)idj, R);
            if (only_rings && rj == ri) continue;
            const RingInter *ic = R.inter + (P.typed ? ti + P.ntypes * R.rtype[rj] : 0);
            const double2 d = cdiff(P, pi, pin[j]);
            if (d.x * d.x + d.y * d.y < ic->nb2) {
                if (cnt < kNeighMax) list[cnt] = idj + 1;
                cnt++;
            }
        }
    };
    if (P.cols == 0) {
        visit(0, P.na);
    } else {
        const int c = scell[s];
        const int row = c / P.cols, col = c - row * P.cols;
        for (int dr = -1; dr <= 1; dr++) {
            int r = row + dr;
            if (P.periodic) r = r < 0 ? r + P.rows : (r >= P.rows ? r - P.rows : r);
            else if (r < 0 || r >= P.rows) continue;
            for (int dc = -1; dc <= 1; dc++) {
                int cc = col + dc;
                if (P.periodic) cc = cc < 0 ? cc + P.cols : (cc >= P.cols ? cc - P.cols : cc);
                else if (cc < 0 || cc >= P.cols) continue;
                visit(start[r * P.cols + cc], start[r * P.cols + cc + 1]);
            }
        }
    }
    if (cnt > kNeighMax) {
        atomicOr(flags, MAVI_FLAG_NEIGH_OVERFLOW);
        cnt = kNeighMax;
    }
    for (int a = 1; a < cnt; a++) {  // ascending ids
        const int v = list[a];
        int b = a - 1;
        while (b >= 0 && list[b] > v) {
            list[b + 1] = list[b];
            b--;
        }
        list[b + 1] = v;
    }
    cnt_out[id] = cnt;
    for (int k = 0; k < kNeighMax; k++) list_out[(int64_t)id * kNeighMax + k] = k < cnt ? list[k] : -1;
}

// ---------------------------------------------------------------------------
// Start of a step of a variable ring population (src/rings/integration.jl:
// 366-372, src/rings/sources.jl:187-260).  k_var_pre: one thread per ring
// slot: update_cms! for the active rings (cms_next holds the mean of the
// last continuous positions), then the sinks that hold the slot's centre of
// mass (stale for inactive slots, as the reference's cms vector is), listed
// as (ring, sink bits).  k_src_empty: one thread per active particle: the
// spawn boxes (grown by pad) it lies in, among the boxes whose ChunksChecker
// lists its chunk of the last binning (every box without chunks).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool inside_rect(const double x, const double y, const double bx, const double by,
                                            const double len, const double h, const double pad) {
    // is_inside(point, ::RectangleCfg; pad), src/configs.jl:89-93
    return (bx - pad <= x && x <= bx + len + pad) && (by - pad <= y && y <= by + h + pad);
}

__global__ void k_var_pre(SysParams P, double2 *__restrict__ cms, const double2 *__restrict__ cms_next,
                          const DevSink *__restrict__ sinks, int nsinks, uint32_t *__restrict__ sink_n,
                          int32_t *__restrict__ sink_list) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= P.nrings) return;
    double2 c = cms[r];
    if (P.rmask[r]) {
        c = cms_next[r];
        cms[r] = c;
    }
    uint32_t bits = 0;
    for (int j = 0; j < nsinks; j++) {
        const DevSink S = sinks[j];
        bool in;
        if (S.geom == MAVI_GEOM_RECT) {
            in = inside_rect(c.x, c.y, S.g[0], S.g[1], S.g[2], S.g[3], 0.0);
        } else {  // is_inside(point, ::CircleCfg), src/configs.jl:165-168
            const double dx = c.x - S.g[0], dy = c.y - S.g[1];
            in = dx * dx + dy * dy <= S.g[2] * S.g[2];
        }
        if (in) bits |= 1u << j;
    }
    if (bits) {
        const uint32_t k = atomicAdd(sink_n, 1u);
        sink_list[2 * k] = r;
        sink_list[2 * k + 1] = (int32_t)bits;
    }
}

__global__ void k_src_empty(SysParams P, const double2 *__restrict__ Pst, const int32_t *__restrict__ scell,
                            const DevBox *__restrict__ boxes, int nbox,
                            const unsigned long long *__restrict__ box_mask,
                            unsigned long long *__restrict__ nonempty) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.na) return;
    const double2 p = Pst[s];
    const unsigned long long listed = box_mask ? box_mask[scell[s]] : 0ull;
    unsigned long long hit = 0;
    for (int b = 0; b < nbox; b++) {
        const DevBox B = boxes[b];
        if (B.chunked && !((listed >> b) & 1ull)) continue;
        if (inside_rect(p.x, p.y, B.x, B.y, B.len, B.h, B.pad)) hit |= 1ull << b;
    }
    if (hit) atomicOr(nonempty, hit);
}

// ---------------------------------------------------------------------------
// Invasions of rings: check_invasions! (src/rings/integration.jl:379-509) at
// the ring points (continuous positions cpos) and centres of mass cms the
// last forces! step saw.
// k_inv_prep: one thread per ring: the bounding box of its points and, with
// a ring grid, its chunk (update_particle_chunk! of the centre of mass)
// appended to that chunk's list.  The order inside a chunk does not matter:
// the host sorts the records.
// k_invasions: one wave per ring r1, lanes over r1's points.  Every other
// ring r2 of the 3x3 block of r1's chunk (every ring without a grid) is
// tested: the reference visits each unordered pair of rings in the same or
// adjacent chunks once and tests both directions, the gather visits each
// ordered pair once.  A point can count a crossing of r2 only when its
// height is strictly inside r2's box and it lies left of the box's right
// side (the height tests of point_line_intersect are strict, and an
// intersection lies between the segment's end abscissas), so boxes skip
// pairs and points without changing any count.  An odd count is an
// invasion record (invasor r1, invaded r2, p_id), 1-based.
// ---------------------------------------------------------------------------
__global__ void k_inv_prep(SysParams P, InvGrid G, const int32_t *__restrict__ rtype,
                           const RingType *__restrict__ types, const double2 *__restrict__ cpos,
                           const double2 *__restrict__ cms, double4 *__restrict__ box,
                           int32_t *__restrict__ gcnt, int32_t *__restrict__ gcell, int32_t *__restrict__ gchunk,
                           uint32_t *__restrict__ flags) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= P.nrings) return;
    const int np = types[P.typed ? rtype[r] : 0].np;
    const double2 *c = cpos + (int64_t)r * P.nmax;
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (int i = 0; i < np; i++) {
        const double2 q = c[i];
        x0 = fmin(x0, q.x);
        x1 = fmax(x1, q.x);
        y0 = fmin(y0, q.y);
        y1 = fmax(y1, q.y);
    }
    box[r] = make_double4(x0, x1, y0, y1);
    if (G.cols == 0 || (P.rmask && !P.rmask[r])) return;  // r_chunks bin get_rings_ids
    uint32_t f = 0;
    const double2 m = cms[r];
    const int k = cell_in_grid(P, G.rows, G.cols, G.cl, G.ch, m.x, m.y, &f, MAVI_FLAG_RING_OUT_OF_BOX);
    gchunk[r] = k;
    const int slot = atomicAdd(gcnt + k, 1);
    if (slot < G.nc) gcell[(int64_t)k * G.nc + slot] = r;
    else f |= MAVI_FLAG_RING_CHUNK_OVERFLOW;  // BoundsError on chunk_particles
    if (f) atomicOr(flags, f);
}

// point_line_intersect(p, (a, b)), src/rings/integration.jl:379-419, with
// the reference's operations in its order (the library is built without
// FP contraction; the division is correctly rounded).
__device__ __forceinline__ int ray_crosses(const double2 p, const double2 a, const double2 b) {
    const double dx = b.x - a.x;
    const double dy = b.y - a.y;
    const double ylo = dy < 0 ? b.y : a.y, yhi = dy < 0 ? a.y : b.y;
    if (!(ylo < p.y && p.y < yhi)) return 0;
    if (dx == 0) return a.x > p.x ? 1 : 0;
    if (dy == 0) return 0;
    const double c = dy * a.x - dx * a.y;
    const double x_inter = (c + dx * p.y) / dy;
    const double xlo = dx < 0 ? b.x : a.x, xhi = dx < 0 ? a.x : b.x;
    return (x_inter > p.x && xlo < x_inter && x_inter < xhi) ? 1 : 0;
}

constexpr int kInvWaves = 4;

__global__ __launch_bounds__(64 * kInvWaves) void k_invasions(
    SysParams P, InvGrid G, const int32_t *__restrict__ rtype, const RingType *__restrict__ types,
    const double2 *__restrict__ cpos, const double4 *__restrict__ box, const int32_t *__restrict__ gcnt,
    const int32_t *__restrict__ gcell, const int32_t *__restrict__ gchunk, int32_t *__restrict__ rec,
    uint32_t *__restrict__ nrec, const int64_t cap) {
    const int r1 = blockIdx.x * kInvWaves + (int)(threadIdx.x >> 6);
    if (r1 >= P.nrings) return;  // whole waves leave together
    const unsigned lane = threadIdx.x & 63;
    const int np1 = types[P.typed ? rtype[r1] : 0].np;
    const double4 b1 = box[r1];
    const double2 *c1 = cpos + (int64_t)r1 * P.nmax;
    auto test = [&](const int r2, const double2 p, const bool act, const int i) {
        if (r2 == r1) return;
        const double4 b2 = box[r2];
        if (b1.w <= b2.z || b1.z >= b2.w || b1.x >= b2.y) return;  // no point of r1 can count
        const bool cand = act && p.y > b2.z && p.y < b2.w && p.x < b2.y;
        if (__ballot(cand) == 0) return;
        const int np2 = types[P.typed ? rtype[r2] : 0].np;
        const double2 *c2 = cpos + (int64_t)r2 * P.nmax;
        int cross = 0;
        double2 a = c2[0];
        for (int e = 0; e < np2; e++) {
            const double2 b = c2[e + 1 < np2 ? e + 1 : 0];
            cross += ray_crosses(p, a, b);
            a = b;
        }
        const bool odd = cand && (cross & 1);
        const unsigned long long m = __ballot(odd);
        if (m == 0) return;
        uint32_t base = 0;
        if (lane == (unsigned)(__ffsll((long long)m) - 1)) base = atomicAdd(nrec, (uint32_t)__popcll(m));
        base = __shfl(base, __ffsll((long long)m) - 1, 64);
        if (odd) {
            const int64_t k = (int64_t)base + __popcll(m & ((1ull << lane) - 1));
            if (k < cap) {
                rec[3 * k + 0] = r1 + 1;
                rec[3 * k + 1] = r2 + 1;
                rec[3 * k + 2] = r1 * P.nmax + i + 1;  // to_scalar_idx(state, r1, i)
            }
        }
    };
    for (int i0 = 0; i0 < np1; i0 += 64) {
        const int i = i0 + (int)lane;
        const bool act = i < np1;
        const double2 p = act ? c1[i] : make_double2(0.0, 0.0);
        if (G.cols == 0) {
            // ring ids 1 .. length(get_rings_ids) (every ring of a fixed population)
            if (r1 >= G.pad) return;
            for (int r2 = 0; r2 < G.pad; r2++) test(r2, p, act, i);
            continue;
        }
        if (P.rmask && !P.rmask[r1]) return;
        const int k = gchunk[r1];
        const int row = k / G.cols, col = k - row * G.cols;
        for (int dr = -1; dr <= 1; dr++) {
            int rr = row + dr;
            if (P.periodic) rr = rr < 0 ? rr + G.rows : (rr >= G.rows ? rr - G.rows : rr);
            else if (rr < 0 || rr >= G.rows) continue;
            for (int dc = -1; dc <= 1; dc++) {
                int cc = col + dc;
                if (P.periodic) cc = cc < 0 ? cc + G.cols : (cc >= G.cols ? cc - G.cols : cc);
                else if (cc < 0 || cc >= G.cols) continue;
                const int kk = rr * G.cols + cc;
                const int n = min(gcnt[kk], G.nc);
                for (int j = 0; j < n; j++) test(gcell[(int64_t)kk * G.nc + j], p, act, i);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Self-propelled particles: szabo_step! / rtp_step! (src/integration.jl:
// 517-535).  One launch after the binning: pair forces at the binned
// positions (Szabo, or the truncated LJ of run-and-tumble), then, fused,
// update_szabo! / update_rtp! (:433-489) and the periodic walls!.  The
// polarisation angle travels in the velocity slot (x component).  The new
// state goes to the other state copy in the same slot order, because
// neighbouring lanes still read pin.  Noise: this step's host draws by
// particle id, or the Philox stream keyed by (seed, step, id).
// ---------------------------------------------------------------------------
constexpr double kTwoPi = 2.0 * 3.141592653589793;  // 2*pi in Julia

template <int MODE, int POT, bool ALLPAIRS, bool WALLS>
__global__ void k_active(SysParams P, const double2 *__restrict__ pin, const double2 *__restrict__ ain,
                         const int32_t *__restrict__ idin, const int32_t *__restrict__ scell,
                         const int32_t *__restrict__ start, double2 *__restrict__ pout, double2 *__restrict__ aout,
                         int32_t *__restrict__ idout, double2 *__restrict__ Fout, const double *__restrict__ noise,
                         uint64_t seed, const unsigned long long *__restrict__ counter, NextBin nb) {
    const int s = xcd_block() * blockDim.x + threadIdx.x;
    const bool valid = s < P.na;
    const int sv = valid ? s : 0;
    const double2 pi = pin[sv];
    const RingCtx R{};
    PairAcc<POT> acc;
    if constexpr (ALLPAIRS) {
        if (valid) scan_range<POT>(P, R, acc, pin, s, pi.x, pi.y, P.periodic != 0, 0, P.na);
    } else {
        pair_gather_chunks<POT>(P, R, acc, pin, scell[sv], start, s, valid, pi, INFINITY, nullptr);
    }
    double xo = 0.0, yo = 0.0;  // final position
    if (valid) {
    double fx = acc.fx, fy = acc.fy;
    if constexpr (WALLS) wall_forces(P, pi.x, pi.y, fx, fy);  // calc_walls_forces!
    Fout[s] = make_double2(fx, fy);
    const int id = idin[s];
    double th = ain[s].x;
    double sn, cs;
    sincos(th, &sn, &cs);
    double x = pi.x, y = pi.y;
    if constexpr (MODE == MODE_SZABO) {
        // vel = vo * pol + mu * F; speed = sqrt(sum(abs, vel)) (:448-452)
        const double vx = P.ap_vo * cs + P.ap_mu * fx;
        const double vy = P.ap_vo * sn + P.ap_mu * fy;
        const double speed = sqrt(fabs(vx) + fabs(vy));
        double cr = (speed > 0.0) ? (cs * vy - sn * vx) / speed : 0.0;
        if (fabs(cr) > 1.0) cr = (cr > 0.0) ? 1.0 : -1.0;
        const double xi = noise ? noise[id] : philox_normal(seed, counter[0], (uint32_t)id);
        const double dth = P.ap_inv_tau * asin(cr) * P.dt + P.ap_noise * xi;
        x = x + vx * P.dt;
        y = y + vy * P.dt;
        th = th + dth;
    } else {
        // vel = vo * pol + F; tumble to 2 pi rand() when rand() < rate dt (:476-488)
        const double vx = P.ap_vo * cs + fx;
        const double vy = P.ap_vo * sn + fy;
        x = x + vx * P.dt;
        y = y + vy * P.dt;
        double u, w;
        if (noise) {
            u = noise[2 * (int64_t)id];
            w = noise[2 * (int64_t)id + 1];
        } else {
            philox_uniform2(seed, counter[0], (uint32_t)id, &u, &w);
        }
        if (u < P.ap_tumble) th = kTwoPi * w;
    }
    // walls! (self-propelled states have no velocity: no rigid walls)
    if constexpr (WALLS) walls_apply(P, x, y, nullptr, nullptr, P.radius);
    else periodic_wrap(P, x, y);
    pout[s] = make_double2(x, y);
    aout[s] = make_double2(th, 0.0);
    idout[s] = id;
    xo = x;
    yo = y;
    }
    if (nb.key) {  // the next binning's first pass; every lane takes part in the arrival ranks
        uint32_t flag = 0;
        int32_t k = -1;
        if (valid) k = cell_of(P, xo, yo, &flag);
        if (flag) atomicOr(nb.flags, flag);
        const int32_t rk = wave_arrival_rank(valid, k, nb.cnt);
        if (valid) {
            nb.key[s] = k;
            nb.rank[s] = rk;
        }
    }
}

// ---------------------------------------------------------------------------
// K2/K3, LDS-tiled form.  One wave owns a segment of kSegCells consecutive
// chunks of one chunk row.  It stages the particles of the three neighbouring
// chunk rows over the segment's columns +-1 (about 3 * 14 * 4.4 = 185 at
// density 0.7) into LDS, column by column in unwrapped column order, with an
// offset table per (row, column).  A lane's 3x3 chunk block is then exactly
// three contiguous LDS ranges, at the periodic edges too.  Global memory is
// read in bulk; the candidate loops only touch LDS.
// ---------------------------------------------------------------------------
constexpr int kSegCells = 12;
constexpr int kSegWaves = 4;
constexpr int kStageCap = 416;
constexpr int kSegCols = kSegCells + 2;        // unwrapped columns staged per row
constexpr int kSegOff = 3 * kSegCols + 1;      // offset table entries

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int POT, bool LISTED>
__device__ __forceinline__ bool seg_accumulate(const SysParams &P, const RingCtx &R, PairAcc<POT> &acc,
                                               const double2 *__restrict__ stg, const int32_t *__restrict__ off,
                                               int iu, int me, double xi, double yi, bool wrap, double list2,
                                               uint16_t *__restrict__ lst, int lane, const int32_t *__restrict__ sid) {
    // ranges: row d covers unwrapped columns iu-1..iu+1
    const int a0 = off[iu - 1], e0 = off[iu + 2];
    const int a1 = off[kSegCols + iu - 1], e1 = off[kSegCols + iu + 2];
    const int a2 = off[2 * kSegCols + iu - 1], e2 = off[2 * kSegCols + iu + 2];
    const int l0 = e0 - a0, l1 = e1 - a1, l2 = e2 - a2;
    const int tot = l0 + l1 + l2;
    const int b1 = l0, b2 = l0 + l1;
    const int o0 = a0, o1 = a1 - l0, o2 = a2 - l0 - l1;
    if constexpr (POT == POT_LJ && !LISTED) {
        // the three staged rows one after another, as the gather form (wrap
        // is the same for the whole wave: it owns one segment)
        if (wrap) {
            lj_range<true, false>(P, acc, stg, xi, yi, a0, l0, 0);
            lj_range<true, true>(P, acc, stg, xi, yi, a1, l1, me - a1);
            lj_range<true, false>(P, acc, stg, xi, yi, a2, l2, 0);
        } else {
            lj_range<false, false>(P, acc, stg, xi, yi, a0, l0, 0);
            lj_range<false, true>(P, acc, stg, xi, yi, a1, l1, me - a1);
            lj_range<false, false>(P, acc, stg, xi, yi, a2, l2, 0);
        }
        return false;
    } else if constexpr (!LISTED) {
        for (int k = 0; k < tot; k++) {
            const int j = k + ((k < b1) ? o0 : ((k < b2) ? o1 : o2));
            const double2 pj = stg[j];
            double dx = xi - pj.x, dy = yi - pj.y;
            if (wrap) {
                dx = min_image(dx, P.L, P.halfL);
                dy = min_image(dy, P.H, P.halfH);
            }
            acc.add(P, R, dx, dy, j != me, POT == POT_RINGS ? sid[j] : 0);
        }
        return false;
    } else {
        int cnt = 0;
        for (int k = 0; k < tot; k++) {
            const int j = k + ((k < b1) ? o0 : ((k < b2) ? o1 : o2));
            const double2 pj = stg[j];
            double dx = xi - pj.x, dy = yi - pj.y;
            if (wrap) {
                dx = min_image(dx, P.L, P.halfL);
                dy = min_image(dy, P.H, P.halfH);
            }
            const double r2 = dx * dx + dy * dy;
            if (j != me && r2 <= list2) {
                if (cnt < kListCap) lst[cnt * 64 + lane] = (uint16_t)j;
                cnt++;
            }
        }
        const int m = cnt < kListCap ? cnt : kListCap;
        for (int q = 0; q < m; q++) {
            const int j = lst[q * 64 + lane];
            const double2 pj = stg[j];
            double dx = xi - pj.x, dy = yi - pj.y;
            if (wrap) {
                dx = min_image(dx, P.L, P.halfL);
                dy = min_image(dy, P.H, P.halfH);
            }
            acc.add(P, R, dx, dy, true, POT == POT_RINGS ? sid[j] : 0);
        }
        return cnt > kListCap;
    }
}

template <int MODE, int POT, bool LISTED>
__global__ __launch_bounds__(64 * kSegWaves) void k_pair_seg(
    SysParams P, RingCtx R, const double2 *__restrict__ pin, const int32_t *__restrict__ scell,
    const int32_t *__restrict__ start, double2 *__restrict__ Pst, double2 *__restrict__ Vst,
    double2 *__restrict__ Fout, double2 *__restrict__ P2, double2 *__restrict__ F1, double list2, int nsegc,
    int nseg) {
    __shared__ double2 s_stage[kSegWaves][kStageCap];
    __shared__ int32_t s_off[kSegWaves][kSegOff];
    __shared__ int32_t s_sid[kSegWaves][POT == POT_RINGS ? kStageCap : 1];  // staged particle ids (rings)
    __shared__ uint16_t s_lst[kSegWaves][LISTED ? kListCap * 64 : 1];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = xcd_block() * kSegWaves + w;
    if (g >= nseg) return;  // whole wave; no workgroup barrier in this kernel
    const int R0 = g / nsegc;
    const int c0 = (g - R0 * nsegc) * kSegCells;
    const int c1 = min(c0 + kSegCells, P.cols) - 1;
    const int nu = c1 - c0 + 3;  // unwrapped columns c0-1 .. c1+1
    double2 *stg = s_stage[w];
    int32_t *off = s_off[w];

    // per (row d, unwrapped column iu) chunk: global start and count
    int gstart = 0, gcount = 0;
    const int li = lane;  // li = d * kSegCols + iu
    const int d = li / kSegCols, iu = li - d * kSegCols;
    if (li < 3 * kSegCols && iu < nu) {
        int rr = R0 + d - 1;
        bool ok = true;
        if (P.periodic) rr = (rr < 0) ? rr + P.rows : (rr >= P.rows ? rr - P.rows : rr);
        else ok = rr >= 0 && rr < P.rows;
        int u = c0 - 1 + iu;
        if (P.periodic) u = (u < 0) ? u + P.cols : (u >= P.cols ? u - P.cols : u);
        else ok = ok && u >= 0 && u < P.cols;
        if (ok) {
            const int cell = rr * P.cols + u;
            gstart = start[cell];
            gcount = start[cell + 1] - gstart;
        }
    }
    // exclusive scan of the counts over the 3*kSegCols table slots
    int incl = gcount;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    const int excl = incl - gcount;
    const int total = __shfl(incl, 3 * kSegCols - 1, 64);
    if (lane < 3 * kSegCols) off[lane] = excl;
    if (lane == 0) off[3 * kSegCols] = total;
    const bool staged = total <= kStageCap;
    if (staged) {
        for (int q = 0; q < gcount; q++) {
            stg[excl + q] = pin[gstart + q];
            if constexpr (POT == POT_RINGS) s_sid[w][excl + q] = R.ids[gstart + q] - R.id_base;
        }
    }
    wave_sync();

    const int p0 = start[R0 * P.cols + c0], p1 = start[R0 * P.cols + c1 + 1];
    const bool edge_rows = P.periodic && (R0 == 0 || R0 == P.rows - 1);
    const bool edge_cols = P.periodic && (c0 == 0 || c1 == P.cols - 1);
    for (int base = p0; base < p1; base += 64) {
        const int s = base + lane;
        const bool valid = s < p1;
        const int sv = valid ? s : p0;
        const int c = scell[sv];
        const int col = c - R0 * P.cols;
        const int iu_self = col - c0 + 1;
        const int me = off[kSegCols + iu_self] + (sv - start[c]);
        const double2 pi = pin[sv];
        PairAcc<POT> acc;
        if constexpr (POT == POT_RINGS) {
            acc.id_i = R.ids[sv] - R.id_base;
            acc.ring_i = ring_of((uint32_t)acc.id_i, R);
            acc.type_i = P.typed ? R.rtype[acc.ring_i] : 0;
        }
        const bool wrap = edge_rows || edge_cols;
        if (staged) {
            if (valid) {
                const int32_t *sid = s_sid[w];
                const bool over = seg_accumulate<POT, LISTED>(P, R, acc, stg, off, iu_self, me, pi.x, pi.y, wrap,
                                                             list2, s_lst[w], lane, sid);
                if (over) {  // more in-range candidates than the list holds: evaluate all
                    acc.fx = 0.0;
                    acc.fy = 0.0;
                    seg_accumulate<POT, false>(P, R, acc, stg, off, iu_self, me, pi.x, pi.y, wrap, list2, s_lst[w],
                                               lane, sid);
                }
            }
        } else if (valid) {
            // staging overflow (unusually dense region): read candidates from global memory
            for (int dd = -1; dd <= 1; dd++) {
                int r = R0 + dd;
                if (P.periodic) r = (r < 0) ? r + P.rows : (r >= P.rows ? r - P.rows : r);
                else if (r < 0 || r >= P.rows) continue;
                for (int du = -1; du <= 1; du++) {
                    int u = col + du;
                    if (P.periodic) u = (u < 0) ? u + P.cols : (u >= P.cols ? u - P.cols : u);
                    else if (u < 0 || u >= P.cols) continue;
                    const int cell = r * P.cols + u;
                    scan_range<POT>(P, R, acc, pin, s, pi.x, pi.y, wrap, start[cell], start[cell + 1]);
                }
            }
        }
        if (!valid) continue;
        const double fx = acc.fx, fy = acc.fy;
        if constexpr (MODE == MODE_FORCES) {
            Fout[s] = make_double2(fx, fy);
        } else if constexpr (MODE == MODE_VERLET1) {
            const double2 v = Vst[s];
            P2[s] = make_double2(pi.x + (v.x * P.dt + fx * P.term), pi.y + (v.y * P.dt + fy * P.term));
            F1[s] = make_double2(fx, fy);
        } else {
            const double2 v = Vst[s];
            const double2 fo = F1[s];
            double vx = v.x + P.hdt * (fx + fo.x);
            double vy = v.y + P.hdt * (fy + fo.y);
            double x = pi.x, y = pi.y;
            if (P.periodic) {
                const double dxc = x - (P.blx + P.halfL), dyc = y - (P.bly + P.halfH);
                const bool ox = fabs(dxc) > P.halfL, oy = fabs(dyc) > P.halfH;
                if (ox) x = x - copysign(P.halfL * 2.0, dxc);
                if (oy) y = y - copysign(P.halfH * 2.0, dyc);
            } else {
                const double rx = x - P.blx, ry = y - P.bly;
                const bool ox = ((rx + P.radius) > P.L) || ((rx - P.radius) < 0.0);
                const bool oy = ((ry + P.radius) > P.H) || ((ry - P.radius) < 0.0);
                if (ox) vx = -vx;
                if (oy) vy = -vy;
            }
            Pst[s] = make_double2(x, y);
            Vst[s] = make_double2(vx, vy);
            Fout[s] = make_double2(fx, fy);
        }
    }
}

// ---------------------------------------------------------------------------
// K4 rings: one thread per ring
// ---------------------------------------------------------------------------

template <bool UPDATE>
__global__ void k_rings(SysParams P, const int32_t *__restrict__ rtype, const RingType *__restrict__ types,
                        const int32_t *__restrict__ slot_of, double2 *__restrict__ Pst, double2 *__restrict__ F,
                        double *__restrict__ pol, double *__restrict__ areas, double2 *__restrict__ cms,
                        double2 *__restrict__ cms_next, double2 *__restrict__ cpos, const double *__restrict__ noise,
                        uint64_t seed, const unsigned long long *__restrict__ counter) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= P.nrings || (P.rmask && !P.rmask[r])) return;  // get_rings_ids
    const RingType T = types[P.typed ? rtype[r] : 0];
    const int np = T.np;
    const int32_t *slot = slot_of + (int64_t)r * P.nmax;
    auto X = [&](int i) { return Pst[slot[i]]; };
    double2 *c = cpos + (int64_t)r * P.nmax;

    // update_cms! (src/rings/integration.jl:366-372) reads the continuous
    // positions of the previous update_continuos_pos!
    if constexpr (UPDATE) cms[r] = cms_next[r];
    double2 cm;
    // slots past np (padding of typed rings) keep their stored positions
    const double area = ring_geometry(
        P, np, [&](int i) { return i < np ? X(i) : c[i]; }, c, &cm);
    cms_next[r] = cm;
    areas[r] = area;

    // springs (forces!, :205-223) in the reference's accumulation order:
    // particle 1 gets +f_1 then -f_n, particle i > 1 gets -f_{i-1} then +f_i.
    auto spring = [&](int a, int b) {
        const double2 d = cdiff(P, X(a), X(b));
        const double dist = sqrt(d.x * d.x + d.y * d.y);
        const double fmod = -T.kspring * (dist - T.lspring);
        const double q = fmod / dist;
        return make_double2(q * d.x, q * d.y);
    };
    double area0 = (double)np * T.lspring / T.p0;
    area0 = area0 * area0;
    const double fa = T.karea * (area - area0);
    const double2 f_last = spring(np - 1, 0);
    double2 f_prev = f_last;
    for (int i = 0; i < np; i++) {
        const double2 fi = (i == np - 1) ? f_last : spring(i, i + 1);
        double2 Fi = F[slot[i]];
        if (i == 0) {
            Fi.x += fi.x;
            Fi.y += fi.y;
            Fi.x -= f_last.x;
            Fi.y -= f_last.y;
        } else {
            Fi.x -= f_prev.x;
            Fi.y -= f_prev.y;
            Fi.x += fi.x;
            Fi.y += fi.y;
        }
        f_prev = fi;
        F[slot[i]] = Fi;
    }
    // area_forces!, :157-194 (after all springs, as in the reference)
    for (int i = 0; i < np; i++) {
        const int i1 = (i == 0) ? np - 1 : i - 1;
        const int i2 = (i == np - 1) ? 0 : i + 1;
        const double2 d = cdiff(P, X(i2), X(i1));
        double2 Fi = F[slot[i]];
        Fi.x -= fa * (d.y / 2.0);
        Fi.y -= fa * (-d.x / 2.0);
        F[slot[i]] = Fi;
    }
    if constexpr (!UPDATE) return;

    // update!, :300-351
    const double theta = pol[r];
    double sn, cs;
    sincos(theta, &sn, &cs);
    double vcx = 0.0, vcy = 0.0;
    for (int i = 0; i < np; i++) {
        const double2 Fi = F[slot[i]];
        const double vx = T.vo * cs + T.mu * Fi.x;
        const double vy = T.vo * sn + T.mu * Fi.y;
        vcx += vx;
        vcy += vy;
        const double2 xi = X(i);
        double px = xi.x + vx * P.dt, py = xi.y + vy * P.dt;
        if (P.periodic) {  // walls! after update!, applied per particle
            const double dxc = px - (P.blx + P.halfL), dyc = py - (P.bly + P.halfH);
            const bool ox = fabs(dxc) > P.halfL, oy = fabs(dyc) > P.halfH;
            if (ox) px = px - copysign(P.halfL * 2.0, dxc);
            if (oy) py = py - copysign(P.halfH * 2.0, dyc);
        }
        Pst[slot[i]] = make_double2(px, py);
    }
    vcx = vcx / (double)np;
    vcy = vcy / (double)np;
    const double speed = sqrt(vcx * vcx + vcy * vcy);
    double cr = 0.0;
    if (speed != 0.0) {
        cr = (cs * vcy - sn * vcx) / speed;
        if (fabs(cr) > 1.0) cr = cr > 0.0 ? 1.0 : -1.0;
    }
    const double z = noise ? noise[r] : (T.rot != 0.0 ? philox_normal(seed, counter[0], (uint32_t)r) : 0.0);
    const double d_theta = 1.0 / T.relax * asin(cr) * P.dt + sqrt(2.0 * T.rot * P.dt) * z;
    pol[r] = theta + d_theta;
}

// ---------------------------------------------------------------------------
// K7 observables: per-block partial sums (fixed reduction tree), summed on
// the host in block order, so the result is reproducible run to run.
// ---------------------------------------------------------------------------
constexpr int kObsThreads = 256;

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[kObsThreads];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kObsThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

__global__ void k_observables(SysParams P, const double2 *__restrict__ pos, const double2 *__restrict__ vel,
                              int with_pe, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0, pe = 0.0;
    if (i < P.na) {
        if (vel) {
            const double2 v = vel[i];
            ke = v.x * v.x + v.y * v.y;
        }
        if (with_pe) {
            const double2 pi = pos[i];
            const double sig = P.p0;
            for (int j = i + 1; j < P.na; j++) {
                const double2 pj = pos[j];
                double dx = pi.x - pj.x, dy = pi.y - pj.y;
                if (P.periodic) {
                    dx = min_image(dx, P.L, P.halfL);
                    dy = min_image(dy, P.H, P.halfH);
                }
                const double dist = sqrt(dx * dx + dy * dy);
                const double t = sig / dist;
                const double t2 = t * t, t6 = t2 * t2 * t2;
                pe += (t6 * t6 - t6);
            }
        }
    }
    const double bke = block_sum(ke);
    const double bpe = block_sum(pe);
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = bke;
        out[2 * blockIdx.x + 1] = bpe;
    }
}


RingCtx ring_ctx(mavi_sys *s) {
    RingCtx R{};
    R.ids = s->ID_[s->cur];
    R.id_base = 0;
    R.rtype = s->rtype;
    R.types = s->types;
    R.inter = s->inter;
    R.nmax = (uint32_t)(s->P.nmax > 0 ? s->P.nmax : 1);
    R.inv_nmax = 1.0f / (float)R.nmax;
    return R;
}

// Squared range of the candidate list: every pair that can interact lies
// within it (rings: the widest interaction of the type matrix).  Infinite
// for the reference's LJ without cutoff: every stencil pair interacts.
double list_range2(const mavi_sys *s) {
    if (s->P.potential != MAVI_POT_RINGS) return s->P.cut2;
    return s->list2;
}

template <int MODE, int POT, bool ALL>
hipError_t launch_pair_t(mavi_sys *s, const double2 *pin, hipStream_t st, bool reuse, bool defer) {
    const int wg = s->wg;
    const int c = s->cur;
    const double l2 = s->two_phase ? list_range2(s) : INFINITY;
    const bool walls = (MODE == MODE_VERLET1 || MODE == MODE_VERLET2) && s->P.nwalls > 0;
    if constexpr (!ALL) {
        if (s->lds_tiles && !walls) {  // the LDS-tiled form carries no walls list
            const int nsegc = (s->P.cols + kSegCells - 1) / kSegCells;
            const int nseg = nsegc * s->P.rows;
            const int nb = (nseg + kSegWaves - 1) / kSegWaves;
            if (l2 < INFINITY)
                hipLaunchKernelGGL((k_pair_seg<MODE, POT, true>), dim3(nb), dim3(64 * kSegWaves), 0, st, s->P,
                                   ring_ctx(s), pin, s->scell, s->start, s->P_[c], s->V_[c], s->F, s->P2, s->F1, l2,
                                   nsegc, nseg);
            else
                hipLaunchKernelGGL((k_pair_seg<MODE, POT, false>), dim3(nb), dim3(64 * kSegWaves), 0, st, s->P,
                                   ring_ctx(s), pin, s->scell, s->start, s->P_[c], s->V_[c], s->F, s->P2, s->F1, l2,
                                   nsegc, nseg);
            return hipGetLastError();
        }
    }
    // a first evaluation that reuses forces compacts its recomputations per
    // block: large blocks leave fewer partly filled waves
    const int wgk = (MODE == MODE_VERLET1 && reuse) ? 1024 : wg;
    const size_t lds = (l2 < INFINITY) ? sizeof(uint16_t) * kListCap * (size_t)wgk : 0;
    const int32_t *needc = reuse ? s->needc : nullptr;
    const int32_t *perm = defer ? s->perm : nullptr;
    // a second evaluation on chunks does the next binning's first pass (with
    // the stamps force reuse reads, as k_bin_count would make them)
    NextBin nb{};
    const bool fuse_stamp = s->reuse && (!s->lds_tiles || s->P.nwalls > 0);
    if (MODE == MODE_VERLET2 && !ALL && s->P.cols > 0) {
        nb = NextBin{s->key, s->rank, s->cnt, s->flags, fuse_stamp ? s->stamp : nullptr};
        s->keys_ready = true;
    }
    if constexpr (MODE == MODE_VERLET1 || MODE == MODE_VERLET2) {
        if (walls) {  // the walls list's forces and walls! in the epilogue
            hipLaunchKernelGGL((k_pair<MODE, POT, ALL, true>), dim3(nblk(s->P.na, wgk)), dim3(wgk), lds, st, s->P,
                               ring_ctx(s), pin, s->scell, s->start, s->P_[c], s->V_[c], s->F, s->P2, s->F1, l2,
                               needc, s->epoch, perm, s->V_[1 - c], s->F,
                               (MODE == MODE_VERLET2 && !ALL) ? s->stamp : nullptr, nb);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_pair<MODE, POT, ALL>), dim3(nblk(s->P.na, wgk)), dim3(wgk), lds, st, s->P, ring_ctx(s), pin,
                       s->scell, s->start, s->P_[c], s->V_[c], s->F, s->P2, s->F1, l2, needc, s->epoch, perm,
                       s->V_[1 - c], s->F, nullptr, nb);
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_pair(mavi_sys *s, const double2 *pin, hipStream_t st, bool reuse = false, bool defer = false) {
    const bool all = s->P.cols == 0;
    switch (s->P.potential) {
        case MAVI_POT_LJ:
            return all ? launch_pair_t<MODE, POT_LJ, true>(s, pin, st, false, false)
                       : launch_pair_t<MODE, POT_LJ, false>(s, pin, st, reuse, defer);
        case MAVI_POT_HARMTRUNC:
            return all ? launch_pair_t<MODE, POT_HT, true>(s, pin, st, false, false)
                       : launch_pair_t<MODE, POT_HT, false>(s, pin, st, reuse, defer);
        case MAVI_POT_RTP:  // truncated LJ (cut2 = the 2^(1/6) sigma threshold)
            if constexpr (MODE != MODE_FORCES) return hipErrorInvalidValue;
            return all ? launch_pair_t<MODE, POT_LJ, true>(s, pin, st, false, false)
                       : launch_pair_t<MODE, POT_LJ, false>(s, pin, st, false, false);
        case MAVI_POT_SZABO:
            if constexpr (MODE != MODE_FORCES) return hipErrorInvalidValue;
            else
                return all ? launch_pair_t<MODE, POT_SZABO, true>(s, pin, st, false, false)
                           : launch_pair_t<MODE, POT_SZABO, false>(s, pin, st, false, false);
        default:
            return all ? launch_pair_t<MODE, POT_RINGS, true>(s, pin, st, false, false)
                       : launch_pair_t<MODE, POT_RINGS, false>(s, pin, st, false, false);
    }
}

template <int MODE, int POT, bool ALL, bool WALLS>
void launch_active_k(mavi_sys *s, const double *noise_dev, hipStream_t st) {
    const int wg = s->wg;
    const int c = s->cur, o = 1 - c;
    // the next binning's first pass on the positions the step writes
    NextBin nb{};
    if (mavi::update_fuses_bin(s)) {
        nb = NextBin{s->key, s->rank, s->cnt, s->flags, nullptr};
        s->keys_ready = true;
    }
    hipLaunchKernelGGL((k_active<MODE, POT, ALL, WALLS>), dim3(nblk(s->P.na, wg)), dim3(wg), 0, st, s->P, s->P_[c],
                       s->V_[c], s->ID_[c], s->scell, s->start, s->P_[o], s->V_[o], s->ID_[o], s->F, noise_dev,
                       s->seed, s->counter, nb);
}

template <int MODE, int POT>
hipError_t launch_active_t(mavi_sys *s, const double *noise_dev, hipStream_t st) {
    const bool all = s->P.cols == 0, walls = s->P.nwalls > 0;
    if (all) {
        if (walls) launch_active_k<MODE, POT, true, true>(s, noise_dev, st);
        else launch_active_k<MODE, POT, true, false>(s, noise_dev, st);
    } else {
        if (walls) launch_active_k<MODE, POT, false, true>(s, noise_dev, st);
        else launch_active_k<MODE, POT, false, false>(s, noise_dev, st);
    }
    s->cur = 1 - s->cur;
    return hipGetLastError();
}

}  // namespace

namespace mavi {

hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n, 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_iota_state(mavi_sys *s, hipStream_t st) {
    if (s->P.na == 0) return hipSuccess;
    const int c = s->cur;
    const bool has_vel = s->P.potential != MAVI_POT_RINGS;
    hipLaunchKernelGGL(k_iota_state, dim3(nblk(s->P.na, 256)), dim3(256), 0, st, s->P.na, s->active_ids, s->pos_id,
                       has_vel ? s->vel_id : nullptr, s->P_[c], has_vel ? s->V_[c] : nullptr, s->ID_[c], s->slot_of, 0, 0);
    return hipGetLastError();
}

hipError_t launch_unpermute(mavi_sys *s, bool vel, bool force, hipStream_t st) {
    if (force && s->var) {  // clean_forces! zeroes the forces of inactive particles too
        hipError_t e = launch_fill_i32((int32_t *)s->force_id, 0, 4 * (int64_t)s->P.n, st);
        if (e != hipSuccess) return e;
    }
    if (s->P.na == 0) return hipSuccess;
    const int c = s->cur;
    const bool has_vel = vel && s->P.potential != MAVI_POT_RINGS;
    hipLaunchKernelGGL(k_unpermute, dim3(nblk(s->P.na, 256)), dim3(256), 0, st, s->P.na, s->ID_[c], s->P_[c],
                       has_vel ? s->V_[c] : nullptr, s->F, s->pos_id, has_vel ? s->vel_id : nullptr,
                       force ? s->force_id : nullptr);
    return hipGetLastError();
}

hipError_t launch_bin(mavi_sys *s, hipStream_t st, bool count_step, bool stamp, bool defer, const SysParams *over,
                      unsigned long long *over_ticket, unsigned long long *over_state) {
    // over: another grid over the same space (the neighbour lists' sort); the
    // default is the reference chunk grid
    const SysParams &P = over ? *over : s->P;
    if (P.cols == 0) return hipSuccess;  // chunks = nothing: the state order stays the active id order
    if (s->nl) s->nl->valid = false;     // a new slot order: lists index slots
    ProfScope ps(s, MAVI_PROF_BIN, st);
    const int c = s->cur, nc = 1 - c;
    // stamps need the previous binning's chunk of every slot (scell)
    const bool stamping = stamp && s->sorted;
    // (a variable ring population can be empty; a second evaluation may have
    // done this pass already)
    if (P.na > 0 && !s->keys_ready)
        hipLaunchKernelGGL(k_bin_count, dim3(nblk(P.na, 256)), dim3(256), 0, st, P, s->P_[c], s->key, s->rank,
                           s->cnt, s->flags, s->scell, stamping ? s->stamp : nullptr, s->epoch);
    s->keys_ready = false;
    const int n = P.ncell + 1;
    launch_scan(s->cnt, n, s->start, over ? over_ticket : s->scan_ticket, over ? over_state : s->scan_state, s->flags,
                st);
    const int grid = nblk(P.na > n ? P.na : n, 256);
    hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(256), 0, st, P, s->key, s->rank, s->start, s->src, s->scell,
                       s->ID_[c], s->sid, s->cnt, count_step ? s->counter : nullptr, stamping ? s->stamp : nullptr, s->epoch, s->needc);
    const bool has_vel = P.potential != MAVI_POT_RINGS && !defer;
    if (P.na > 0) {
        hipLaunchKernelGGL(k_finalize, dim3(nblk(P.na, 256)), dim3(256), 0, st, P, s->start, s->src, s->scell,
                           s->sid, s->P_[c], has_vel ? s->V_[c] : nullptr, s->ID_[c], s->P_[nc],
                           has_vel ? s->V_[nc] : nullptr, s->ID_[nc],
                           P.potential == MAVI_POT_RINGS ? s->slot_of : nullptr, s->flags, defer ? s->perm : nullptr,
                           s->epoch);
    }
    s->cur = nc;
    s->sorted = true;
    return hipGetLastError();
}

hipError_t launch_export_chunks(mavi_sys *s, int64_t *d_cp, int64_t *d_counts, hipStream_t st) {
    hipLaunchKernelGGL(k_export_chunks, dim3(nblk(s->P.ncell, 256)), dim3(256), 0, st, s->P, s->start,
                       s->ID_[s->cur], d_cp, d_counts);
    return hipGetLastError();
}

hipError_t launch_forces_only(mavi_sys *s, hipStream_t st) {
    ProfScope ps(s, MAVI_PROF_FORCE, st);
    return launch_pair<MODE_FORCES>(s, s->P_[s->cur], st);
}

hipError_t launch_verlet_step(mavi_sys *s, hipStream_t st, bool reuse) {
    // the gather kernel's first evaluation moves velocities (and reused
    // forces) into the new order, so the binning leaves them where they are
    const bool defer = (!s->lds_tiles || s->P.nwalls > 0) && s->P.cols > 0;
    const bool can = s->reuse && defer;
    reuse = reuse && can && s->sorted;
    hipError_t e = launch_bin(s, st, true, can, defer);
    if (e != hipSuccess) return e;
    {
        ProfScope ps(s, reuse ? MAVI_PROF_FORCE_REUSE : MAVI_PROF_FORCE, st);
        if ((e = launch_pair<MODE_VERLET1>(s, s->P_[s->cur], st, reuse, defer)) != hipSuccess) return e;
    }
    ProfScope ps(s, MAVI_PROF_FORCE, st);
    e = launch_pair<MODE_VERLET2>(s, s->P2, st);
    s->f_valid = can && e == hipSuccess;
    return e;
}

hipError_t drop_keys(mavi_sys *s, hipStream_t st) {
    if (!s->keys_ready) return hipSuccess;
    s->keys_ready = false;
    return launch_fill_i32(s->cnt, 0, (int64_t)s->P.ncell + 1, st);
}

static hipError_t launch_rings_kernel(mavi_sys *s, bool update, const double *noise_dev, hipStream_t st,
                                      const RingRates *rates = nullptr) {
    ProfScope ps(s, MAVI_PROF_RINGS, st);
    const int c = s->cur;
    const double *nz = update ? noise_dev : nullptr;
    if (s->P.nmax <= 32) {
        const int per = 64 / s->P.nmax;  // rings per wave
        const int nb = nblk((int64_t)((s->P.nrings + per - 1) / per) * 64, 256);
        // an update also does the next binning's first pass (chunk keys,
        // arrival ranks, counts) on the positions it writes
        const bool fuse = update && update_fuses_bin(s);
        int32_t *nk = fuse ? s->key : nullptr, *nr = fuse ? s->rank : nullptr, *nc = fuse ? s->cnt : nullptr;
        uint32_t *nf = fuse ? s->flags : nullptr;
        if (fuse) s->keys_ready = true;
        if (update && s->P.nwalls > 0)
            hipLaunchKernelGGL((k_rings_lanes<true, true>), dim3(nb), dim3(256), 0, st, s->P, s->rtype, s->types,
                               s->slot_of, s->P_[c], s->F, s->pol, s->areas, s->cms, s->cms_next, s->cpos, nz,
                               s->seed, s->counter, nk, nr, nc, nf, RingRates{});
        else if (update)
            hipLaunchKernelGGL(k_rings_lanes<true>, dim3(nb), dim3(256), 0, st, s->P, s->rtype, s->types, s->slot_of,
                               s->P_[c], s->F, s->pol, s->areas, s->cms, s->cms_next, s->cpos, nz, s->seed,
                               s->counter, nk, nr, nc, nf, RingRates{});
        else if (rates)
            hipLaunchKernelGGL((k_rings_lanes<false, false, true>), dim3(nb), dim3(256), 0, st, s->P, s->rtype,
                               s->types, s->slot_of, s->P_[c], s->F, s->pol, s->areas, s->cms, s->cms_next, s->cpos,
                               nz, s->seed, s->counter, nk, nr, nc, nf, *rates);
        else
            hipLaunchKernelGGL(k_rings_lanes<false>, dim3(nb), dim3(256), 0, st, s->P, s->rtype, s->types,
                               s->slot_of, s->P_[c], s->F, s->pol, s->areas, s->cms, s->cms_next, s->cpos, nz,
                               s->seed, s->counter, nk, nr, nc, nf, RingRates{});
        return hipGetLastError();
    }
    if (update)
        hipLaunchKernelGGL(k_rings<true>, dim3(nblk(s->P.nrings, 64)), dim3(64), 0, st, s->P, s->rtype, s->types,
                           s->slot_of, s->P_[c], s->F, s->pol, s->areas, s->cms, s->cms_next, s->cpos, nz, s->seed,
                           s->counter);
    else
        hipLaunchKernelGGL(k_rings<false>, dim3(nblk(s->P.nrings, 64)), dim3(64), 0, st, s->P, s->rtype, s->types,
                           s->slot_of, s->P_[c], s->F, s->pol, s->areas, s->cms, s->cms_next, s->cpos, nullptr,
                           s->seed, s->counter);
    return hipGetLastError();
}

hipError_t launch_ring_neigh(mavi_sys *s, hipStream_t st) {
    hipLaunchKernelGGL(k_ring_neigh, dim3(nblk(s->P.na, 256)), dim3(256), 0, st, s->P, ring_ctx(s), s->P_[s->cur],
                       s->start, s->scell, s->neigh_only_rings, s->neigh_cnt, s->neigh_list, s->flags);
    return hipGetLastError();
}

hipError_t launch_invasions(mavi_sys *s, bool prep, hipStream_t st) {
    InvGrid G = s->inv_grid;
    G.pad = s->var ? (int32_t)s->vnum_active : s->P.nrings;  // length(get_rings_ids)
    if (prep) {
        if (G.cols > 0) {
            hipError_t e = launch_fill_i32(s->inv_cnt, 0, (int64_t)G.rows * G.cols, st);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_inv_prep, dim3(nblk(s->P.nrings, 256)), dim3(256), 0, st, s->P, G, s->rtype, s->types,
                           s->cpos, s->cms, s->inv_box, s->inv_cnt, s->inv_cell, s->inv_chunk, s->flags);
    }
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, (int32_t *)s->inv_n, 0, (int64_t)1);
    hipLaunchKernelGGL(k_invasions, dim3(nblk(s->P.nrings, kInvWaves)), dim3(64 * kInvWaves), 0, st, s->P, G,
                       s->rtype, s->types, s->cpos, s->inv_box, s->inv_cnt, s->inv_cell, s->inv_chunk, s->inv_rec,
                       s->inv_n, s->inv_cap);
    return hipGetLastError();
}

__global__ void k_var_clear(SysParams P, double2 *__restrict__ cms, double2 *__restrict__ cms_next) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= P.nrings || P.rmask[r]) return;
    cms[r] = make_double2(0.0, 0.0);
    cms_next[r] = make_double2(0.0, 0.0);
}

hipError_t launch_var_clear(mavi_sys *s, hipStream_t st) {
    hipLaunchKernelGGL(k_var_clear, dim3(nblk(s->P.nrings, 256)), dim3(256), 0, st, s->P, s->cms, s->cms_next);
    return hipGetLastError();
}

hipError_t launch_var_pre(mavi_sys *s, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, (int32_t *)s->sink_n, 0, (int64_t)1);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, (int32_t *)s->nonempty, 0, (int64_t)2);
    hipLaunchKernelGGL(k_var_pre, dim3(nblk(s->P.nrings, 256)), dim3(256), 0, st, s->P, s->cms, s->cms_next,
                       s->dsinks, s->nsinks, s->sink_n, s->sink_list);
    if (s->nboxes > 0 && s->P.na > 0)
        hipLaunchKernelGGL(k_src_empty, dim3(nblk(s->P.na, 256)), dim3(256), 0, st, s->P, s->P_[s->cur], s->scell,
                           s->dboxes, s->nboxes, s->box_mask, s->nonempty);
    return hipGetLastError();
}

hipError_t launch_rings_forces(mavi_sys *s, hipStream_t st, const RingRates *rates) {
    if (s->P.na == 0) return hipSuccess;
    hipError_t e = launch_bin(s, st, false, false, false);
    if (e != hipSuccess) return e;
    {
        ProfScope ps(s, MAVI_PROF_FORCE, st);
        if ((e = launch_pair<MODE_FORCES>(s, s->P_[s->cur], st)) != hipSuccess) return e;
    }
    if (s->neigh_on && (e = launch_ring_neigh(s, st)) != hipSuccess) return e;
    return launch_rings_kernel(s, false, nullptr, st, rates);
}

hipError_t launch_rings_step(mavi_sys *s, const double *noise_dev, hipStream_t st) {
    if (s->P.na == 0) {  // no active ring: the step only advances the noise counter
        hipLaunchKernelGGL(k_counter_inc, dim3(1), dim3(64), 0, st, s->counter);
        return hipGetLastError();
    }
    hipError_t e = launch_bin(s, st, true, false, false);
    if (e != hipSuccess) return e;
    if (s->P.cols == 0) {  // no binning kernel to advance the noise step counter
        hipLaunchKernelGGL(k_counter_inc, dim3(1), dim3(64), 0, st, s->counter);
    }
    {
        ProfScope ps(s, MAVI_PROF_FORCE, st);
        if ((e = launch_pair<MODE_FORCES>(s, s->P_[s->cur], st)) != hipSuccess) return e;
    }
    if (s->neigh_on && (e = launch_ring_neigh(s, st)) != hipSuccess) return e;
    return launch_rings_kernel(s, true, noise_dev, st);
}

hipError_t launch_active_step(mavi_sys *s, const double *noise_dev, hipStream_t st) {
    hipError_t e = launch_bin(s, st, true, false, false);  // angles move with the velocity slot
    if (e != hipSuccess) return e;
    if (s->P.cols == 0) hipLaunchKernelGGL(k_counter_inc, dim3(1), dim3(64), 0, st, s->counter);
    ProfScope ps(s, MAVI_PROF_FORCE, st);
    if (s->P.potential == MAVI_POT_SZABO) return launch_active_t<MODE_SZABO, POT_SZABO>(s, noise_dev, st);
    return launch_active_t<MODE_RTP, POT_LJ>(s, noise_dev, st);
}

hipError_t launch_rings_init(mavi_sys *s, hipStream_t st) {
    hipLaunchKernelGGL(k_rings_init, dim3(nblk(s->P.nrings, 64)), dim3(64), 0, st, s->P, s->rtype, s->types,
                       s->pos_id, s->cms, s->cms_next, s->cpos, s->areas);
    return hipGetLastError();
}

hipError_t launch_observables(mavi_sys *s, double *d_out, hipStream_t st) {
    const int with_pe = (s->P.potential == MAVI_POT_LJ) ? 1 : 0;
    const bool has_vel = s->P.potential != MAVI_POT_RINGS;
    hipLaunchKernelGGL(k_observables, dim3(nblk(s->P.na, kObsThreads)), dim3(kObsThreads), 0, st, s->P,
                       s->P_[s->cur], has_vel ? s->V_[s->cur] : nullptr, with_pe, d_out);
    return hipGetLastError();
}

}  // namespace mavi
