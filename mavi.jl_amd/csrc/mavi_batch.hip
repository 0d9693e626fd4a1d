// mavi_batch.hip -- replica batches: many independent systems of uneven size
// stepped together (the device side of run_experiment_batch,
// src/experiments.jl:246-300).
//
// Layout.  The replicas' particles share one set of sorted-state arrays.
// Replica r owns the sorted slots [pbase, pbase + npad), npad = na rounded up
// to a multiple of 64, so that every wavefront works on one replica and the
// wave-uniform decisions of the pair gather (wrap, split, list) are the
// single-system ones.  Its chunks are [cbase, cbase + ncell) followed by one
// pad chunk that collects the npad - na padding slots: one counting sort over
// the concatenated chunk space then keeps every replica inside its slot
// range, in the same order a single-system binning gives.  Particle ids are
// concatenated too (nbase + local id), and ring arrays (rbase, even); a
// wavefront of the ring kernel holds floor(64 / nmax) rings of one replica.
//
// A step is: binning of all replicas, one pair launch per kind and
// evaluation (VERLET1 / VERLET2 for LJ and HarmTrunc replicas, FORCES for
// rings), and the per-ring kernel for rings replicas.  Each replica runs the
// same device functions as a single system (mavi_device.h), so its results
// are bit-identical to mavi_step on the same descriptor.
//
// Force reuse, as for a single system: the second evaluation of an LJ or
// HarmTrunc replica does the next binning's first pass on the positions it
// writes (chunk, arrival rank) and marks, in the next binning's epoch, the
// chunks of the 3x3 blocks around a particle's old and new chunk when it
// changes chunk; the ring kernel does the same first pass for the rings
// replicas (padding slots included).  The binning then starts from the
// scan, writes the permutation instead of moving velocities, and the first
// evaluation moves velocities through it and takes the previous second
// evaluation wherever its chunk is not marked (same positions, same
// candidates, same order: the same bits).  Particles to recompute are
// handled per wave (a workgroup may hold waves of other replicas).  A
// replica whose state was replaced recomputes everything once (rvalid).
#include <math.h>

#include <algorithm>
#include <string>
#include <vector>

#include "mavi_device.h"
#include "mavi_host.h"

namespace {

struct Replica {
    SysParams P;
    int32_t pbase, npad;  // first sorted slot (multiple of 64), padded slot count
    int32_t nbase;        // first particle id
    int32_t cbase;        // first chunk; cbase + P.ncell is the pad chunk
    int32_t rbase;        // first ring (even; rings only)
    int32_t tbase, ibase; // first ring type / interaction entry
    int32_t abase;        // first entry of the active id list (rings only)
    int64_t cpbase;       // first continuous position (rings only)
    double list2;         // squared candidate-list range
    uint64_t seed;
    float inv_nmax;
    int32_t kind;         // 0 LJ, 1 HarmTrunc, 2 rings
    int32_t rwbase;       // first wave of the ring kernel (rings only)
};

enum { KIND_LJ = 0, KIND_HT = 1, KIND_RINGS = 2, NKIND = 3 };

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Chunk and arrival rank of every padded slot.  Padding slots go to their
// replica's pad chunk.  nslots is a multiple of 64: whole waves leave.
__global__ void kb_bin_count(const Replica *__restrict__ reps, const int32_t *__restrict__ wave_rep, int32_t nslots,
                             const double2 *__restrict__ pos, int32_t *__restrict__ key, int32_t *__restrict__ rank,
                             int32_t *__restrict__ cnt, uint32_t *__restrict__ flags) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const int r = uniform(wave_rep[s >> 6]);
    const Replica &Q = reps[r];
    uint32_t flag = 0;
    int32_t k;
    if (s - Q.pbase < Q.P.na) {
        const double2 p = pos[s];
        k = Q.cbase + cell_of(Q.P, p.x, p.y, &flag);
    } else {
        k = Q.cbase + Q.P.ncell;
    }
    if (flag) atomicOr(&flags[r], flag);
    const int32_t rk = wave_arrival_rank(true, k, cnt);
    key[s] = k;
    rank[s] = rk;
}

// k_finalize over all replicas; padding slots (id -1) never move, so they
// keep id -1 in both state copies.  slot_of is kept for the rings replicas
// only (the ring update is its one reader; an LJ replica's id order is
// scattered over its slots, so those writes would be scattered too).
__global__ void kb_finalize(const Replica *__restrict__ reps, const int32_t *__restrict__ wave_rep, int32_t nslots,
                            const int32_t *__restrict__ start, const int32_t *__restrict__ src,
                            const int32_t *__restrict__ scell, const int32_t *__restrict__ sid,
                            const double2 *__restrict__ Pold, const int32_t *__restrict__ IDold,
                            double2 *__restrict__ Pnew, int32_t *__restrict__ IDnew,
                            int32_t *__restrict__ slot_of, uint32_t *__restrict__ flags, int32_t *__restrict__ perm,
                            uint32_t *__restrict__ epoch) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) epoch[0] += 1u;  // binnings so far
    if (t >= nslots) return;
    const int r = uniform(wave_rep[t >> 6]);
    const int so = src[t];
    const int myid = sid[t];
    if (myid < 0) return;
    const Replica &Q = reps[r];
    const int c = scell[t];
    const int b = start[c], e = start[c + 1];
    int rk = 0;
    for (int q = b; q < e; q++) rk += (sid[q] < myid) ? 1 : 0;
    const int t2 = b + rk;
    Pnew[t2] = Pold[so];
    perm[t2] = so;  // velocities move in the first evaluation (rings replicas use none)
    IDnew[t2] = myid;
    if (Q.kind == KIND_RINGS) slot_of[myid] = t2;
    if (t == b && e - b > Q.P.nc) atomicOr(&flags[r], MAVI_FLAG_OVERFLOW);
}

__global__ void kb_unpermute(int32_t nslots, const int32_t *__restrict__ ID, const double2 *__restrict__ P,
                             const double2 *__restrict__ V, const double2 *__restrict__ F,
                             double2 *__restrict__ pos_id, double2 *__restrict__ vel_id,
                             double2 *__restrict__ force_id) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const int id = ID[s];
    if (id < 0) return;
    pos_id[id] = P[s];
    vel_id[id] = V[s];
    force_id[id] = F[s];
}

// Reuse state of the Verlet kernels (LJ / HarmTrunc replicas).
struct BatchReuse {
    const int32_t *perm;   // new slot -> old slot of the last binning
    const double2 *Vold;   // velocities in the old slot order
    const double2 *Fold;   // the previous second evaluation, old slot order
    int32_t *needc;        // per chunk: the binning epoch in which a chunk of its 3x3 block changed members
    const uint32_t *epoch; // binnings so far
    uint32_t *rvalid;      // per replica: 1 once a second evaluation ran on its current state
    // the next binning's first pass
    int32_t *key, *rank, *cnt;
    uint32_t *flags;
};

// needc <- e on the chunks of the 3x3 block around chunk (row, col) of a
// replica's grid (periodic wrap or clipped), as the single system's stamps
// and k_scatter's block test mark them.
__device__ __forceinline__ void mark_block(const SysParams &P, int32_t *needc, int cbase, int c, int32_t e) {
    const int row = c / P.cols, col = c - row * P.cols;
    for (int dr = -1; dr <= 1; dr++) {
        int r = row + dr;
        if (P.periodic) r = r < 0 ? r + P.rows : (r >= P.rows ? r - P.rows : r);
        else if (r < 0 || r >= P.rows) continue;
        for (int dc = -1; dc <= 1; dc++) {
            int cc = col + dc;
            if (P.periodic) cc = cc < 0 ? cc + P.cols : (cc >= P.cols ? cc - P.cols : cc);
            else if (cc < 0 || cc >= P.cols) continue;
            needc[cbase + r * P.cols + cc] = e;
        }
    }
}

// Pair forces + integrator epilogue for the waves of one kind: items[w] =
// (global wave, replica).  The per-particle work is pair_gather_chunks /
// pair_epilogue with the replica's parameters, as in k_pair; with U (Verlet
// kinds), the force reuse and fused binning pass described at the top.
template <int MODE, int POT>
__global__ void kb_pair(const Replica *__restrict__ reps, const int2 *__restrict__ items, int32_t nitems,
                        const int32_t *__restrict__ ids, const int32_t *__restrict__ rtype,
                        const RingType *__restrict__ types, const RingInter *__restrict__ inter,
                        const double2 *__restrict__ pin, const int32_t *__restrict__ scell,
                        const int32_t *__restrict__ start, double2 *__restrict__ Pst, double2 *__restrict__ Vst,
                        double2 *__restrict__ Fout, double2 *__restrict__ P2, double2 *__restrict__ F1,
                        int32_t listed, BatchReuse U) {
    extern __shared__ uint16_t s_list[];
    const int wl = xcd_block() * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
    if (wl >= nitems) return;
    const int2 it = items[wl];
    const int w = uniform(it.x), r = uniform(it.y);
    const Replica &Q = reps[r];
    const SysParams P = Q.P;
    const int s = w * 64 + (int)lane_id();
    bool valid = s - Q.pbase < P.na;
    const int sv = valid ? s : Q.pbase;
    const double2 pi = pin[sv];
    if constexpr (MODE == MODE_VERLET1) {
        if (U.perm) {
            const int so = valid ? U.perm[s] : 0;
            if (valid) Vst[s] = U.Vold[so];
            // a chunk whose 3x3 block kept its members: the previous second
            // evaluation, bit for bit
            const bool need =
                valid && (uniform((int)U.rvalid[r]) == 0 || U.needc[scell[sv]] == (int32_t)U.epoch[0]);
            if (valid && !need) {
                const double2 f = U.Fold[so];
                pair_epilogue<MODE>(P, s, pi, f.x, f.y, Pst, Vst, Fout, P2, F1);
            }
            if (!__any(need)) return;  // whole wave
            valid = need;
        }
    }
    RingCtx R{};
    PairAcc<POT> acc;
    if constexpr (POT == POT_RINGS) {
        R.ids = ids;
        R.id_base = Q.nbase;
        R.rtype = rtype + Q.rbase;
        R.types = types + Q.tbase;
        R.inter = inter + Q.ibase;
        R.nmax = (uint32_t)P.nmax;
        R.inv_nmax = Q.inv_nmax;
        acc.id_i = ids[sv] - Q.nbase;
        acc.ring_i = ring_of((uint32_t)acc.id_i, R);
        acc.type_i = P.typed ? R.rtype[acc.ring_i] : 0;
    }
    pair_gather_chunks<POT>(P, R, acc, pin, scell[sv] - Q.cbase, start + Q.cbase, s, valid, pi,
                            listed ? Q.list2 : INFINITY, s_list);
    if constexpr (MODE == MODE_VERLET2) {
        double2 xf = make_double2(0.0, 0.0);
        if (valid) pair_epilogue<MODE>(P, s, pi, acc.fx, acc.fy, Pst, Vst, Fout, P2, F1, &xf);
        if (U.key) {  // every lane of the wave takes part in the arrival ranks
            uint32_t flag = 0;
            int32_t k = Q.cbase + P.ncell;  // padding slots: the replica's pad chunk
            if (valid) {
                const int kl = cell_of(P, xf.x, xf.y, &flag);
                const int pc = scell[s] - Q.cbase;
                if (kl != pc) {
                    const int32_t e = (int32_t)(U.epoch[0] + 1u);
                    mark_block(P, U.needc, Q.cbase, kl, e);
                    mark_block(P, U.needc, Q.cbase, pc, e);
                }
                k = Q.cbase + kl;
            }
            if (flag) atomicOr(&U.flags[r], flag);
            const int32_t rk = wave_arrival_rank(true, k, U.cnt);
            U.key[s] = k;
            U.rank[s] = rk;
            if (lane_id() == 0) U.rvalid[r] = 1u;
        }
        return;
    }
    if (!valid) return;
    pair_epilogue<MODE>(P, s, pi, acc.fx, acc.fy, Pst, Vst, Fout, P2, F1);
}

// Springs, area force and overdamped update of every ring of the rings
// replicas: wave wl holds rings of replica ring_wave_rep[wl] only, floor(64 /
// nmax) of them, from ring (wl - rwbase) * floor(64 / nmax) of the replica.
// Then the next binning's first pass for the replica's slots, as kb_bin_count
// makes it: the chunk and arrival rank of every final position, and the
// replica's first ring wave puts its padding slots in its pad chunk.
__global__ __launch_bounds__(256) void kb_rings(const Replica *__restrict__ reps,
                                                const int32_t *__restrict__ ring_wave_rep, int32_t nwaves,
                                                const int32_t *__restrict__ rtype, const RingType *__restrict__ types,
                                                const int32_t *__restrict__ slot_of, double2 *__restrict__ Pst,
                                                double2 *__restrict__ F, double *__restrict__ pol,
                                                double *__restrict__ areas, double2 *__restrict__ cms,
                                                double2 *__restrict__ cms_next, double2 *__restrict__ cpos,
                                                const unsigned long long *__restrict__ counter,
                                                int32_t *__restrict__ key, int32_t *__restrict__ rank,
                                                int32_t *__restrict__ cnt, uint32_t *__restrict__ flags) {
    const int wl = blockIdx.x * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
    if (wl >= nwaves) return;
    const int r = uniform(ring_wave_rep[wl]);
    const Replica &Q = reps[r];
    const SysParams P = Q.P;
    // floor(64 / nmax) rings of nmax lanes per wave, as the single-system kernel
    const int w = (int)(threadIdx.x & 63);
    const int per = 64 / P.nmax;
    const int grp = w / P.nmax, i = w - grp * P.nmax;
    const int rr = (wl - Q.rwbase) * per + grp;
    const bool ring_ok = grp < per && rr < P.nrings;
    const int rc = ring_ok ? rr : 0;
    const RingType T = types[Q.tbase + (P.typed ? rtype[Q.rbase + rc] : 0)];
    ring_lanes<true>(P, T, ring_ok, rc, grp * P.nmax, i, slot_of + Q.nbase, Pst, F, pol + Q.rbase,
                     areas + Q.rbase, cms + Q.rbase, cms_next + Q.rbase, cpos + Q.cpbase, nullptr, Q.seed,
                     counter[0], key, rank, cnt, flags + r, RingRates{}, Q.cbase);
    if (wl == Q.rwbase) {
        // padding slots [pbase + na, pbase + npad): all in the last wave of
        // the replica's slots
        const int tail = P.na & 63;
        const bool pad = tail != 0 && w >= tail;
        const int s = Q.pbase + (P.na - tail) + w;
        const int32_t k = Q.cbase + P.ncell;
        const int32_t rk = wave_arrival_rank(pad, k, cnt);
        if (pad) {
            key[s] = k;
            rank[s] = rk;
        }
    }
}

}  // namespace

struct mavi_batch {
    mavi_ctx *ctx = nullptr;
    std::vector<HostSys> hs;
    std::vector<Replica> reps;
    bool finalized = false;
    int64_t nids = 0, nactive = 0, nslots = 0, ncells = 0, nrings = 0;
    SysParams G{};  // na = nslots, ncell = ncells: the sort over all replicas
    // per-kind wave lists
    int2 *items[NKIND] = {nullptr, nullptr, nullptr};
    int32_t nitems[NKIND] = {0, 0, 0};
    int32_t *ring_wave_rep = nullptr;
    int32_t nring_waves = 0;
    Replica *d_reps = nullptr;
    int32_t *wave_rep = nullptr;
    // state (same roles as in mavi_sys)
    double2 *P_[2] = {nullptr, nullptr}, *V_[2] = {nullptr, nullptr};
    int32_t *ID_[2] = {nullptr, nullptr};
    double2 *F = nullptr, *P2 = nullptr, *F1 = nullptr;
    int32_t *key = nullptr, *rank = nullptr, *src = nullptr, *scell = nullptr, *sid = nullptr, *slot_of = nullptr;
    int32_t *cnt = nullptr, *start = nullptr;
    unsigned long long *scan_state = nullptr, *scan_ticket = nullptr;
    uint32_t *flags = nullptr;
    // force reuse
    int32_t *perm = nullptr, *needc = nullptr;
    uint32_t *epoch = nullptr, *rvalid = nullptr;
    bool keys_ready = false;  // the last second evaluation did the next binning's first pass
    unsigned long long *counter = nullptr;
    int32_t *active_ids = nullptr;
    double2 *pos_id = nullptr, *vel_id = nullptr, *force_id = nullptr;
    double *pol = nullptr, *areas = nullptr;
    double2 *cms = nullptr, *cms_next = nullptr, *cpos = nullptr;
    int32_t *rtype = nullptr;
    RingType *types = nullptr;
    RingInter *inter = nullptr;
    int cur = 0;
    std::vector<char> uploaded;
    // execution shape
    int32_t wg[2] = {256, 256};      // plain, rings
    int32_t listed[2] = {0, 0};
    int32_t graph_steps = 100;
    bool use_graphs = true;
    hipGraphExec_t graph = nullptr;
    // profiling
    bool prof = false;
    ProfSlot profs[MAVI_PROF_NCLASS];
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

struct BatchProf {
    mavi_batch *b;
    int which;
    hipStream_t st;
    BatchProf(mavi_batch *b_, int w, hipStream_t st_) : b(b_), which(w), st(st_) {
        if (b->prof) (void)hipEventRecord(b->ev0, st);
    }
    ~BatchProf() {
        if (!b->prof) return;
        (void)hipEventRecord(b->ev1, st);
        if (hipEventSynchronize(b->ev1) == hipSuccess) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess) {
                b->profs[which].ms += ms;
                b->profs[which].launches += 1;
            }
        }
    }
};

hipError_t batch_bin(mavi_batch *b, hipStream_t st) {
    BatchProf ps(b, MAVI_PROF_BIN, st);
    const int c = b->cur, nc = 1 - c;
    const int ns = (int)b->nslots;
    const int n = (int)b->ncells + 1;
    if (!b->keys_ready) {
        // counts a fused pass may have left are dropped: every slot afresh
        hipError_t e = mavi::launch_fill_i32(b->cnt, 0, n, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(kb_bin_count, dim3(nblk(ns, 256)), dim3(256), 0, st, b->d_reps, b->wave_rep, ns,
                           b->P_[c], b->key, b->rank, b->cnt, b->flags);
    }
    // (else the last step did the first pass: the Verlet replicas' second
    // evaluation and the ring update, padding slots included)
    // a stalled scan is reported through replica 0's flags
    launch_scan(b->cnt, n, b->start, b->scan_ticket, b->scan_state, b->flags, st);
    hipLaunchKernelGGL(k_scatter, dim3(nblk(std::max(ns, n), 256)), dim3(256), 0, st, b->G, b->key, b->rank,
                       b->start, b->src, b->scell, b->ID_[c], b->sid, b->cnt, b->counter, nullptr, nullptr, nullptr);
    hipLaunchKernelGGL(kb_finalize, dim3(nblk(ns, 256)), dim3(256), 0, st, b->d_reps, b->wave_rep, ns, b->start,
                       b->src, b->scell, b->sid, b->P_[c], b->ID_[c], b->P_[nc], b->ID_[nc], b->slot_of, b->flags,
                       b->perm, b->epoch);
    b->cur = nc;
    return hipGetLastError();
}

template <int MODE, int POT>
void batch_pair_kind(mavi_batch *b, int kind, const double2 *pin, hipStream_t st) {
    const int ni = b->nitems[kind];
    if (ni == 0) return;
    const int slot = kind == KIND_RINGS ? 1 : 0;
    const int wg = b->wg[slot];
    const int listed = b->listed[slot];
    const size_t lds = listed ? sizeof(uint16_t) * kListCap * (size_t)wg : 0;
    const int c = b->cur;
    BatchReuse U{};
    if (MODE == MODE_VERLET1) {
        U.perm = b->perm;
        U.Vold = b->V_[1 - c];
        U.Fold = b->F;
        U.needc = b->needc;
        U.epoch = b->epoch;
        U.rvalid = b->rvalid;
    } else if (MODE == MODE_VERLET2) {
        U.needc = b->needc;
        U.epoch = b->epoch;
        U.rvalid = b->rvalid;
        U.key = b->key;
        U.rank = b->rank;
        U.cnt = b->cnt;
        U.flags = b->flags;
    }
    hipLaunchKernelGGL((kb_pair<MODE, POT>), dim3(nblk(ni, wg / 64)), dim3(wg), lds, st, b->d_reps, b->items[kind],
                       ni, b->ID_[c], b->rtype, b->types, b->inter, pin, b->scell, b->start, b->P_[c], b->V_[c],
                       b->F, b->P2, b->F1, listed, U);
}

hipError_t batch_step(mavi_batch *b, hipStream_t st) {
    hipError_t e = batch_bin(b, st);
    if (e != hipSuccess) return e;
    {
        BatchProf ps(b, MAVI_PROF_FORCE, st);
        batch_pair_kind<MODE_VERLET1, POT_LJ>(b, KIND_LJ, b->P_[b->cur], st);
        batch_pair_kind<MODE_VERLET1, POT_HT>(b, KIND_HT, b->P_[b->cur], st);
        batch_pair_kind<MODE_FORCES, POT_RINGS>(b, KIND_RINGS, b->P_[b->cur], st);
    }
    {
        BatchProf ps(b, MAVI_PROF_FORCE, st);
        batch_pair_kind<MODE_VERLET2, POT_LJ>(b, KIND_LJ, b->P2, st);
        batch_pair_kind<MODE_VERLET2, POT_HT>(b, KIND_HT, b->P2, st);
    }
    if (b->nring_waves > 0) {
        BatchProf ps(b, MAVI_PROF_RINGS, st);
        hipLaunchKernelGGL(kb_rings, dim3(nblk(b->nring_waves, 4)), dim3(256), 0, st, b->d_reps, b->ring_wave_rep,
                           b->nring_waves, b->rtype, b->types, b->slot_of, b->P_[b->cur], b->F, b->pol, b->areas,
                           b->cms, b->cms_next, b->cpos, b->counter, b->key, b->rank, b->cnt, b->flags);
    }
    b->keys_ready = true;
    return hipGetLastError();
}

hipError_t batch_graph(mavi_batch *b, hipStream_t st) {
    if (b->graph) return hipSuccess;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    const int cur0 = b->cur;
    hipError_t e1 = hipSuccess;
    for (int k = 0; k < b->graph_steps && e1 == hipSuccess; k++) e1 = batch_step(b, st);
    e = hipStreamEndCapture(st, &g);
    b->cur = cur0;
    if (e == hipSuccess && e1 != hipSuccess) e = e1;
    if (e == hipSuccess) e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    return e;
}

void drop_graph(mavi_batch *b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
}

int batch_check_flags(mavi_batch *b) {
    mavi_ctx *ctx = b->ctx;
    const size_t nr = b->reps.size();
    std::vector<uint32_t> h(nr, 0);
    HIPCHK(ctx, hipMemcpyAsync(h.data(), b->flags, sizeof(uint32_t) * nr, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t r = 0; r < nr; r++) {
        if (!h[r]) continue;
        HIPCHK(ctx, hipMemsetAsync(b->flags, 0, sizeof(uint32_t) * nr, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        const std::string who = "replica " + std::to_string(r) + ": ";
        if (h[r] & MAVI_FLAG_SCAN_STALL) return set_err(ctx, MAVI_E_HIP, "chunk-count scan: look-back did not complete");
        if (h[r] & MAVI_FLAG_NONFINITE)
            return set_err(ctx, MAVI_E_NONFINITE, who + "non-finite particle position met while binning");
        if (h[r] & MAVI_FLAG_OUT_OF_BOX)
            return set_err(ctx, MAVI_E_OUT_OF_BOX, who + "particle outside the chunk grid (BoundsError in update_chunks!)");
        return set_err(ctx, MAVI_E_CELL_OVERFLOW,
                       who + "a chunk holds more particles than its capacity nc (BoundsError in update_chunks!)");
    }
    return MAVI_OK;
}

// Id-order staging of one replica -> its sorted slots in copy cur (+ ring
// geometry), as mavi_sys_upload does for a single system.
hipError_t batch_stage_in(mavi_batch *b, int r, hipStream_t st) {
    const Replica &Q = b->reps[r];
    const int c = b->cur;
    // a replaced state: its forces are recomputed and every slot binned afresh
    b->keys_ready = false;
    hipError_t e0 = hipMemsetAsync(b->rvalid + r, 0, sizeof(uint32_t), st);
    if (e0 != hipSuccess) return e0;
    const bool rings = Q.kind == KIND_RINGS;
    hipLaunchKernelGGL(k_iota_state, dim3(nblk(Q.P.na, 256)), dim3(256), 0, st, Q.P.na,
                       rings ? b->active_ids + Q.abase : nullptr, b->pos_id + Q.nbase, b->vel_id + Q.nbase,
                       b->P_[c] + Q.pbase, b->V_[c] + Q.pbase, b->ID_[c] + Q.pbase, b->slot_of + Q.nbase, Q.nbase,
                       Q.pbase);
    if (rings)
        hipLaunchKernelGGL(k_rings_init, dim3(nblk(Q.P.nrings, 64)), dim3(64), 0, st, Q.P, b->rtype + Q.rbase,
                           b->types + Q.tbase, b->pos_id + Q.nbase, b->cms + Q.rbase, b->cms_next + Q.rbase,
                           b->cpos + Q.cpbase, b->areas + Q.rbase);
    return hipGetLastError();
}

int run_impl(mavi_batch *b, int64_t nsteps) {
    mavi_ctx *ctx = b->ctx;
    hipStream_t st = ctx->stream;
    int64_t k = 0;
    const bool graphs = b->use_graphs && !b->prof && nsteps >= b->graph_steps;
    // the graph holds steps that follow a step of these kernels (fused
    // binning pass) from state copy 0
    while (graphs && k < nsteps && (!b->keys_ready || b->cur != 0)) {
        HIPCHK(ctx, batch_step(b, st));
        k++;
    }
    if (graphs) {
        HIPCHK(ctx, batch_graph(b, st));
        for (; k + b->graph_steps <= nsteps; k += b->graph_steps) HIPCHK(ctx, hipGraphLaunch(b->graph, st));
    }
    for (; k < nsteps; k++) HIPCHK(ctx, batch_step(b, st));
    const int ns = (int)b->nslots;
    hipLaunchKernelGGL(kb_unpermute, dim3(nblk(ns, 256)), dim3(256), 0, st, ns, b->ID_[b->cur], b->P_[b->cur],
                       b->V_[b->cur], b->F, b->pos_id, b->vel_id, b->force_id);
    HIPCHK(ctx, hipGetLastError());
    return batch_check_flags(b);
}

}  // namespace

extern "C" {

int mavi_batch_create(mavi_ctx *ctx, mavi_batch **out) {
    if (!ctx || !out) return MAVI_E_BADARG;
    *out = new mavi_batch();
    (*out)->ctx = ctx;
    ctx_retain(ctx);
    return MAVI_OK;
}

int mavi_batch_destroy(mavi_batch *b) {
    if (!b) return MAVI_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    for (int c = 0; c < 2; c++) {
        dfree(b->P_[c]);
        dfree(b->V_[c]);
        dfree(b->ID_[c]);
    }
    for (int k = 0; k < NKIND; k++) dfree(b->items[k]);
    dfree(b->ring_wave_rep); dfree(b->d_reps); dfree(b->wave_rep);
    dfree(b->F); dfree(b->P2); dfree(b->F1);
    dfree(b->perm); dfree(b->needc); dfree(b->epoch); dfree(b->rvalid);
    dfree(b->key); dfree(b->rank); dfree(b->src); dfree(b->scell); dfree(b->sid); dfree(b->slot_of);
    dfree(b->cnt); dfree(b->start); dfree(b->scan_state); dfree(b->scan_ticket); dfree(b->flags); dfree(b->counter);
    dfree(b->active_ids); dfree(b->pos_id); dfree(b->vel_id); dfree(b->force_id);
    dfree(b->pol); dfree(b->areas); dfree(b->cms); dfree(b->cms_next); dfree(b->cpos);
    dfree(b->rtype); dfree(b->types); dfree(b->inter);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    mavi_ctx *ctx = b->ctx;
    delete b;
    ctx_release(ctx);
    return MAVI_OK;
}

int mavi_batch_add_replica(mavi_batch *b, const mavi_sys_desc *d, int32_t *index) {
    if (!b || !d) return MAVI_E_BADARG;
    mavi_ctx *ctx = b->ctx;
    if (b->finalized) return set_err(ctx, MAVI_E_BADARG, "batch already finalized");
    HostSys h;
    std::string err;
    int st = build_params(d, h, err);
    if (st != MAVI_OK) return set_err(ctx, st, err);
    if (h.P.cols == 0) return set_err(ctx, MAVI_E_UNSUPPORTED, "batch replicas need chunks (cols, rows > 0)");
    if (h.P.potential == MAVI_POT_SZABO || h.P.potential == MAVI_POT_RTP)
        return set_err(ctx, MAVI_E_UNSUPPORTED, "batch replicas are LJ / HarmTrunc gases or rings");
    if (h.P.nwalls > 0) return set_err(ctx, MAVI_E_UNSUPPORTED, "batch replicas take a single rectangle wall");
    if (h.P.potential == MAVI_POT_RINGS && h.P.nmax > 32)
        return set_err(ctx, MAVI_E_UNSUPPORTED, "batch rings replicas need nmax <= 32");
    b->hs.push_back(std::move(h));
    if (index) *index = (int32_t)(b->hs.size() - 1);
    return MAVI_OK;
}

int mavi_batch_finalize(mavi_batch *b) {
    if (!b) return MAVI_E_BADARG;
    mavi_ctx *ctx = b->ctx;
    if (b->finalized) return set_err(ctx, MAVI_E_BADARG, "batch already finalized");
    if (b->hs.empty()) return set_err(ctx, MAVI_E_BADARG, "empty batch");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int nr = (int)b->hs.size();
    std::vector<int2> items[NKIND];
    std::vector<int32_t> wave_rep, ring_wave_rep, active, rtype;
    std::vector<RingType> types;
    std::vector<RingInter> inter;
    int64_t nslots = 0, nids = 0, nactive = 0, ncells = 0, nrings = 0, ncpos = 0, ntypes = 0, ninter = 0;
    b->reps.resize(nr);
    for (int r = 0; r < nr; r++) {
        const HostSys &h = b->hs[r];
        Replica &Q = b->reps[r];
        Q = Replica{};
        Q.P = h.P;
        Q.list2 = h.list2;
        Q.seed = h.seed;
        Q.kind = h.P.potential == MAVI_POT_LJ ? KIND_LJ : (h.P.potential == MAVI_POT_HARMTRUNC ? KIND_HT : KIND_RINGS);
        const int64_t npad = ((int64_t)h.P.na + 63) / 64 * 64;
        Q.pbase = (int32_t)nslots;
        Q.npad = (int32_t)npad;
        Q.nbase = (int32_t)nids;
        Q.cbase = (int32_t)ncells;
        Q.inv_nmax = 1.0f / (float)(h.P.nmax > 0 ? h.P.nmax : 1);
        for (int64_t w = 0; w < npad / 64; w++) {
            wave_rep.push_back(r);
            items[Q.kind].push_back(make_int2((int)(nslots / 64 + w), r));
        }
        if (Q.kind == KIND_RINGS) {
            Q.rbase = (int32_t)nrings;
            Q.tbase = (int32_t)ntypes;
            Q.ibase = (int32_t)ninter;
            Q.abase = (int32_t)nactive;
            Q.cpbase = ncpos;
            const int64_t nrpad = ((int64_t)h.P.nrings + 1) / 2 * 2;
            const int per = 64 / std::max(h.P.nmax, 1);  // rings per wave of the ring kernel
            Q.rwbase = (int32_t)ring_wave_rep.size();
            for (int64_t w = 0; w < ((int64_t)h.P.nrings + per - 1) / per; w++) ring_wave_rep.push_back(r);
            for (int q = 0; q < h.P.nrings; q++) rtype.push_back(h.rtype[q]);
            for (int64_t q = h.P.nrings; q < nrpad; q++) rtype.push_back(0);
            types.insert(types.end(), h.types.begin(), h.types.end());
            inter.insert(inter.end(), h.inter.begin(), h.inter.end());
            active.insert(active.end(), h.active.begin(), h.active.end());
            nrings += nrpad;
            ntypes += (int64_t)h.types.size();
            ninter += (int64_t)h.inter.size();
            nactive += h.P.na;
            ncpos += (int64_t)h.P.nrings * h.P.nmax;
        }
        nslots += npad;
        nids += h.P.n;
        ncells += (int64_t)h.P.ncell + 1;
        if (nslots > 0x3fffffff || nids > 0x3fffffff || ncells > 0x3fffffff)
            return set_err(ctx, MAVI_E_BADARG, "batch too large (2^30 slots, ids or chunks)");
    }
    b->nslots = nslots;
    b->nids = nids;
    b->ncells = ncells;
    b->nrings = nrings;
    b->nactive = 0;
    for (int r = 0; r < nr; r++) b->nactive += b->reps[r].P.na;
    b->G = SysParams{};
    b->G.na = (int32_t)nslots;
    b->G.ncell = (int32_t)ncells;
    b->uploaded.assign(nr, 0);

    hipStream_t st = ctx->stream;
    hipError_t e = hipSuccess;
#define ALLOC(ptr, cnt) \
    if (e == hipSuccess) e = dalloc(&(ptr), (cnt), st)
    for (int c = 0; c < 2; c++) {
        ALLOC(b->P_[c], nslots);
        ALLOC(b->V_[c], nslots);
        ALLOC(b->ID_[c], nslots);
    }
    ALLOC(b->F, nslots);
    ALLOC(b->P2, nslots);
    ALLOC(b->F1, nslots);
    ALLOC(b->key, nslots);
    ALLOC(b->rank, nslots);
    ALLOC(b->src, nslots);
    ALLOC(b->scell, nslots);
    ALLOC(b->sid, nslots);
    ALLOC(b->slot_of, nids);
    ALLOC(b->cnt, ncells + 1);
    ALLOC(b->start, ncells + 1);
    ALLOC(b->scan_state, nblk(ncells + 1, kScanTileMin));
    ALLOC(b->scan_ticket, 1);
    ALLOC(b->flags, nr);
    ALLOC(b->counter, 1);
    ALLOC(b->perm, nslots);
    ALLOC(b->needc, ncells + 1);
    ALLOC(b->epoch, 1);
    ALLOC(b->rvalid, nr);
    ALLOC(b->pos_id, nids);
    ALLOC(b->vel_id, nids);
    ALLOC(b->force_id, nids);
    ALLOC(b->d_reps, nr);
    ALLOC(b->wave_rep, (int64_t)wave_rep.size());
    for (int k = 0; k < NKIND; k++) {
        b->nitems[k] = (int32_t)items[k].size();
        if (!items[k].empty()) ALLOC(b->items[k], (int64_t)items[k].size());
    }
    b->nring_waves = (int32_t)ring_wave_rep.size();
    if (nrings > 0) {
        ALLOC(b->ring_wave_rep, (int64_t)ring_wave_rep.size());
        ALLOC(b->active_ids, (int64_t)active.size());
        ALLOC(b->pol, nrings);
        ALLOC(b->areas, nrings);
        ALLOC(b->cms, nrings);
        ALLOC(b->cms_next, nrings);
        ALLOC(b->cpos, ncpos);
        ALLOC(b->rtype, nrings);
        ALLOC(b->types, (int64_t)types.size());
        ALLOC(b->inter, (int64_t)inter.size());
    }
#undef ALLOC
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    auto up = [&](void *dst, const void *srcp, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(dst, srcp, bytes, hipMemcpyHostToDevice, st);
    };
    up(b->d_reps, b->reps.data(), sizeof(Replica) * b->reps.size());
    up(b->wave_rep, wave_rep.data(), sizeof(int32_t) * wave_rep.size());
    for (int k = 0; k < NKIND; k++) up(b->items[k], items[k].data(), sizeof(int2) * items[k].size());
    if (nrings > 0) {
        up(b->ring_wave_rep, ring_wave_rep.data(), sizeof(int32_t) * ring_wave_rep.size());
        up(b->active_ids, active.data(), sizeof(int32_t) * active.size());
        up(b->rtype, rtype.data(), sizeof(int32_t) * rtype.size());
        up(b->types, types.data(), sizeof(RingType) * types.size());
        up(b->inter, inter.data(), sizeof(RingInter) * inter.size());
    }
    // padding slots carry id -1 in both state copies for good
    for (int c = 0; c < 2 && e == hipSuccess; c++) e = mavi::launch_fill_i32(b->ID_[c], -1, nslots, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return set_err(ctx, MAVI_E_HIP, std::string("batch allocation/upload failed: ") +
                                                             hipGetErrorString(e));
    b->hs.clear();
    b->hs.shrink_to_fit();
    b->finalized = true;
    return MAVI_OK;
}

static int check_index(mavi_batch *b, int32_t i) {
    if (!b->finalized) return set_err(b->ctx, MAVI_E_BADARG, "batch not finalized");
    if (i < 0 || i >= (int32_t)b->reps.size()) return set_err(b->ctx, MAVI_E_BADARG, "replica index out of range");
    return MAVI_OK;
}

int mavi_batch_upload(mavi_batch *b, int32_t i, const double *pos, const double *vel, const double *pol) {
    if (!b || !pos) return MAVI_E_BADARG;
    int rc = check_index(b, i);
    if (rc != MAVI_OK) return rc;
    mavi_ctx *ctx = b->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const Replica &Q = b->reps[i];
    const size_t bytes = sizeof(double2) * (size_t)Q.P.n;
    HIPCHK(ctx, hipMemcpyAsync(b->pos_id + Q.nbase, pos, bytes, hipMemcpyHostToDevice, st));
    if (vel && Q.kind != KIND_RINGS)
        HIPCHK(ctx, hipMemcpyAsync(b->vel_id + Q.nbase, vel, bytes, hipMemcpyHostToDevice, st));
    if (pol && Q.kind == KIND_RINGS)
        HIPCHK(ctx, hipMemcpyAsync(b->pol + Q.rbase, pol, sizeof(double) * Q.P.nrings, hipMemcpyHostToDevice, st));
    HIPCHK(ctx, batch_stage_in(b, i, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    b->uploaded[i] = 1;
    return MAVI_OK;
}

int mavi_batch_download(mavi_batch *b, int32_t i, double *pos, double *vel, double *pol) {
    if (!b) return MAVI_E_BADARG;
    int rc = check_index(b, i);
    if (rc != MAVI_OK) return rc;
    mavi_ctx *ctx = b->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const Replica &Q = b->reps[i];
    const size_t bytes = sizeof(double2) * (size_t)Q.P.n;
    if (pos) HIPCHK(ctx, hipMemcpyAsync(pos, b->pos_id + Q.nbase, bytes, hipMemcpyDeviceToHost, st));
    if (vel && Q.kind != KIND_RINGS)
        HIPCHK(ctx, hipMemcpyAsync(vel, b->vel_id + Q.nbase, bytes, hipMemcpyDeviceToHost, st));
    if (pol && Q.kind == KIND_RINGS)
        HIPCHK(ctx, hipMemcpyAsync(pol, b->pol + Q.rbase, sizeof(double) * Q.P.nrings, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    return MAVI_OK;
}

int mavi_batch_get_forces(mavi_batch *b, int32_t i, double *forces) {
    if (!b || !forces) return MAVI_E_BADARG;
    int rc = check_index(b, i);
    if (rc != MAVI_OK) return rc;
    mavi_ctx *ctx = b->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const Replica &Q = b->reps[i];
    HIPCHK(ctx, hipMemcpyAsync(forces, b->force_id + Q.nbase, sizeof(double2) * (size_t)Q.P.n,
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MAVI_OK;
}

int mavi_batch_run(mavi_batch *b, int64_t nsteps) {
    if (!b) return MAVI_E_BADARG;
    mavi_ctx *ctx = b->ctx;
    if (!b->finalized) return set_err(ctx, MAVI_E_BADARG, "batch not finalized");
    if (nsteps < 0) return set_err(ctx, MAVI_E_BADARG, "nsteps < 0");
    for (size_t r = 0; r < b->uploaded.size(); r++)
        if (!b->uploaded[r]) return set_err(ctx, MAVI_E_BADARG, "replica " + std::to_string(r) + " has no state");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    return run_impl(b, nsteps);
}

int mavi_batch_noise_get_counter(mavi_batch *b, uint64_t *counter) {
    if (!b || !counter) return MAVI_E_BADARG;
    mavi_ctx *ctx = b->ctx;
    if (!b->finalized) return set_err(ctx, MAVI_E_BADARG, "batch not finalized");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    unsigned long long c = 0;
    HIPCHK(ctx, hipMemcpyAsync(&c, b->counter, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    *counter = c;
    return MAVI_OK;
}

int mavi_batch_noise_set_counter(mavi_batch *b, uint64_t counter) {
    if (!b) return MAVI_E_BADARG;
    mavi_ctx *ctx = b->ctx;
    if (!b->finalized) return set_err(ctx, MAVI_E_BADARG, "batch not finalized");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    const unsigned long long c = counter;
    HIPCHK(ctx, hipMemcpy(b->counter, &c, sizeof(c), hipMemcpyHostToDevice));
    return MAVI_OK;
}

// A changed pair-kernel shape may visit candidates in another order: the
// reused forces would not be its bits, so every replica recomputes once.
static int batch_forget_forces(mavi_batch *b) {
    if (!b->finalized || !b->rvalid) return MAVI_OK;
    HIPCHK(b->ctx, hipSetDevice(b->ctx->device));
    HIPCHK(b->ctx, hipMemsetAsync(b->rvalid, 0, sizeof(uint32_t) * b->reps.size(), b->ctx->stream));
    return MAVI_OK;
}

int mavi_batch_set_option(mavi_batch *b, int32_t option, int32_t value) {
    if (!b) return MAVI_E_BADARG;
    mavi_ctx *ctx = b->ctx;
    switch (option) {
        case MAVI_BOPT_WG_PLAIN:
        case MAVI_BOPT_WG_RINGS:
            if (value == 0) value = 256;
            if (!is_pow2_wg(value)) return set_err(ctx, MAVI_E_BADARG, "workgroup must be 64..1024, a power of 2");
            if (b->wg[option - MAVI_BOPT_WG_PLAIN] != value) drop_graph(b);
            b->wg[option - MAVI_BOPT_WG_PLAIN] = value;
            return batch_forget_forces(b);
        case MAVI_BOPT_LIST_PLAIN:
        case MAVI_BOPT_LIST_RINGS:
            if (b->listed[option - MAVI_BOPT_LIST_PLAIN] != (value != 0)) drop_graph(b);
            b->listed[option - MAVI_BOPT_LIST_PLAIN] = value != 0;
            return batch_forget_forces(b);
        case MAVI_BOPT_GRAPH_STEPS:
            if (value < 2 || (value & 1)) return set_err(ctx, MAVI_E_BADARG, "graph steps must be even and >= 2");
            if (value != b->graph_steps) drop_graph(b);
            b->graph_steps = value;
            return MAVI_OK;
        case MAVI_BOPT_GRAPHS:
            b->use_graphs = value != 0;
            return MAVI_OK;
        default:
            return set_err(ctx, MAVI_E_BADARG, "unknown batch option");
    }
}

int mavi_batch_info(mavi_batch *b, int64_t *info) {
    if (!b || !info) return MAVI_E_BADARG;
    if (!b->finalized) return set_err(b->ctx, MAVI_E_BADARG, "batch not finalized");
    info[0] = (int64_t)b->reps.size();
    info[1] = b->nids;
    info[2] = b->nactive;
    info[3] = b->nslots;
    info[4] = b->ncells;
    int64_t nr = 0;
    for (const Replica &Q : b->reps) nr += Q.kind == KIND_RINGS ? Q.P.nrings : 0;
    info[5] = nr;
    return MAVI_OK;
}

int mavi_batch_prof_enable(mavi_batch *b, int on) {
    if (!b) return MAVI_E_BADARG;
    b->prof = on != 0;
    for (auto &p : b->profs) p = ProfSlot{};
    return MAVI_OK;
}

int mavi_batch_prof_get(mavi_batch *b, int which, double *ms_total, int64_t *launches) {
    if (!b || which < 0 || which >= MAVI_PROF_NCLASS) return MAVI_E_BADARG;
    if (ms_total) *ms_total = b->profs[which].ms;
    if (launches) *launches = b->profs[which].launches;
    return MAVI_OK;
}

int mavi_batch_autotune(mavi_batch *b, int32_t trial_steps, double *best_us, int32_t *shape) {
    if (!b || trial_steps <= 0) return MAVI_E_BADARG;
    mavi_ctx *ctx = b->ctx;
    if (!b->finalized) return set_err(ctx, MAVI_E_BADARG, "batch not finalized");
    for (size_t r = 0; r < b->uploaded.size(); r++)
        if (!b->uploaded[r]) return set_err(ctx, MAVI_E_BADARG, "replica " + std::to_string(r) + " has no state");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // snapshot: id-order staging of every replica, polarisations, step counter
    const int ns = (int)b->nslots;
    hipLaunchKernelGGL(kb_unpermute, dim3(nblk(ns, 256)), dim3(256), 0, st, ns, b->ID_[b->cur], b->P_[b->cur],
                       b->V_[b->cur], b->F, b->pos_id, b->vel_id, b->force_id);
    HIPCHK(ctx, hipGetLastError());
    double2 *pos0 = nullptr, *vel0 = nullptr;
    double *pol0 = nullptr;
    unsigned long long *cnt0 = nullptr;
    const size_t bytes = sizeof(double2) * (size_t)b->nids;
    hipError_t e = dalloc(&pos0, b->nids, st);
    if (e == hipSuccess) e = dalloc(&vel0, b->nids, st);
    if (e == hipSuccess) e = dalloc(&cnt0, 1, st);
    if (e == hipSuccess && b->nrings) e = dalloc(&pol0, b->nrings, st);
    if (e == hipSuccess) e = hipMemcpyAsync(pos0, b->pos_id, bytes, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(vel0, b->vel_id, bytes, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(cnt0, b->counter, sizeof(unsigned long long), hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess && pol0) e = hipMemcpyAsync(pol0, b->pol, sizeof(double) * b->nrings, hipMemcpyDeviceToDevice, st);
    auto restore = [&]() -> hipError_t {
        hipError_t r = hipMemcpyAsync(b->pos_id, pos0, bytes, hipMemcpyDeviceToDevice, st);
        if (r == hipSuccess) r = hipMemcpyAsync(b->vel_id, vel0, bytes, hipMemcpyDeviceToDevice, st);
        if (r == hipSuccess)
            r = hipMemcpyAsync(b->counter, cnt0, sizeof(unsigned long long), hipMemcpyDeviceToDevice, st);
        if (r == hipSuccess && pol0)
            r = hipMemcpyAsync(b->pol, pol0, sizeof(double) * b->nrings, hipMemcpyDeviceToDevice, st);
        b->cur = 0;
        for (int q = 0; q < (int)b->reps.size() && r == hipSuccess; q++) r = batch_stage_in(b, q, st);
        if (r == hipSuccess) r = hipStreamSynchronize(st);
        return r;
    };
    const int wgs[4] = {64, 128, 256, 512};
    const bool saved_prof = b->prof;
    b->prof = false;
    double best_total = 1e300;
    for (int slot = 0; slot < 2 && e == hipSuccess; slot++) {
        const bool present = slot == 0 ? (b->nitems[KIND_LJ] + b->nitems[KIND_HT] > 0) : b->nitems[KIND_RINGS] > 0;
        if (!present) continue;
        double best = 1e300;
        int bw = b->wg[slot], bl = b->listed[slot];
        for (int tp = 0; tp < 2 && e == hipSuccess; tp++) {
            for (int wi = 0; wi < 4 && e == hipSuccess; wi++) {
                b->wg[slot] = wgs[wi];
                b->listed[slot] = tp;
                if ((e = restore()) != hipSuccess) break;
                if ((e = batch_step(b, st)) != hipSuccess) break;  // warm-up
                if ((e = hipEventRecord(b->ev0, st)) != hipSuccess) break;
                for (int k = 0; k < trial_steps && e == hipSuccess; k++) e = batch_step(b, st);
                if (e == hipSuccess) e = hipEventRecord(b->ev1, st);
                if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
                float ms = 0.f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, b->ev0, b->ev1);
                const double us = 1e3 * ms / trial_steps;
                if (e == hipSuccess && us < best) {
                    best = us;
                    bw = wgs[wi];
                    bl = tp;
                }
            }
        }
        b->wg[slot] = bw;
        b->listed[slot] = bl;
    }
    // the chosen shapes of both slots together: a mixed batch steps both
    // kinds, so the time of one slot's search is not the step's time
    if (e == hipSuccess) e = restore();
    if (e == hipSuccess) e = batch_step(b, st);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(b->ev0, st);
    for (int k = 0; k < trial_steps && e == hipSuccess; k++) e = batch_step(b, st);
    if (e == hipSuccess) e = hipEventRecord(b->ev1, st);
    if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
    if (e == hipSuccess) {
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, b->ev0, b->ev1);
        best_total = 1e3 * ms / trial_steps;
    }
    b->prof = saved_prof;
    drop_graph(b);
    if (e == hipSuccess) e = restore();
    if (e == hipSuccess) (void)hipMemsetAsync(b->flags, 0, sizeof(uint32_t) * b->reps.size(), st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dfree(pos0);
    dfree(vel0);
    dfree(pol0);
    dfree(cnt0);
    HIPCHK(ctx, e);
    if (best_us) *best_us = best_total;
    if (shape) {
        shape[0] = b->wg[0];
        shape[1] = b->listed[0];
        shape[2] = b->wg[1];
        shape[3] = b->listed[1];
    }
    return MAVI_OK;
}

}  // extern "C"
