// mavi_device.h -- device helpers shared by the single-system kernels
// (mavi_kernels.hip) and the replica-batch kernels (mavi_batch.hip).  Both
// use the same functions, so a replica in a batch computes bit-identical
// results to the same system stepped alone.
#pragma once
#include "mavi_internal.h"

#include "mavi_philox.h"

namespace {

constexpr int kXCD = 8;  // MI355X: workgroups go round-robin over 8 XCDs

// Map the hardware block id so that every XCD (and its L2) works on one
// contiguous stretch of the sorted particles: neighbouring chunk rows are
// then read from the same L2.
__device__ __forceinline__ int xcd_block() {
    const int b = blockIdx.x, nb = gridDim.x;
    const int full = (nb / kXCD) * kXCD;
    if (b >= full) return b;
    return (b % kXCD) * (nb / kXCD) + b / kXCD;
}

__device__ __forceinline__ double rcp_nr(double x) {
    // v_rcp_f64 (relative error <= 4.6e-8, measured on gfx950 over 4M values
    // of r^2 in [0.3, 6.3]) and one Newton step: <= 2.3e-15 relative, so
    // the LJ force 4 eps (12 s^12 inv^7 - 6 s^6 inv^4) stays within ~2e-14,
    // far inside the path's 1e-12 force tolerance.
    const double y = __builtin_amdgcn_rcp(x);
    const double e = __builtin_fma(-x, y, 1.0);
    return __builtin_fma(y, e, y);
}

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// Cell of one position, update_particle_chunk!, src/chunks.jl:120-142, on a
// rows x cols grid of cl x ch chunks over the chunk space.  A position off
// the grid raises out_flag and is clamped so that memory accesses stay in
// bounds.
// Row and column (0-based) of the cell.
__device__ __forceinline__ void cell_rc(const SysParams &P, int32_t rows, int32_t cols, double cl, double ch, double x,
                                        double y, uint32_t *flag, uint32_t out_flag, int32_t *row, int32_t *col) {
    double qr = exact_trunc_div((-y + P.bly) + P.H, ch);
    double qc = exact_trunc_div(x - P.blx, cl);
    if (!(fabs(qr) < 1e15) || !(fabs(qc) < 1e15)) {  // NaN / Inf / absurd
        *flag |= MAVI_FLAG_NONFINITE;
        *row = 0;
        *col = 0;
        return;
    }
    int64_t r = (int64_t)qr;  // 0-based (reference row - 1)
    int64_t c = (int64_t)qc;
    r -= (r == rows) ? 1 : 0;  // the N+1 -> N clamp
    c -= (c == cols) ? 1 : 0;
    if (r < 0 || r >= rows || c < 0 || c >= cols) {
        *flag |= out_flag;
        r = r < 0 ? 0 : (r >= rows ? rows - 1 : r);
        c = c < 0 ? 0 : (c >= cols ? cols - 1 : c);
    }
    *row = (int32_t)r;
    *col = (int32_t)c;
}

__device__ __forceinline__ int32_t cell_in_grid(const SysParams &P, int32_t rows, int32_t cols, double cl,
                                                double ch, double x, double y, uint32_t *flag, uint32_t out_flag) {
    int32_t r, c;
    cell_rc(P, rows, cols, cl, ch, x, y, flag, out_flag, &r, &c);
    return r * cols + c;
}

__device__ __forceinline__ int32_t cell_of(const SysParams &P, double x, double y, uint32_t *flag) {
    return cell_in_grid(P, P.rows, P.cols, P.cl, P.ch, x, y, flag, MAVI_FLAG_OUT_OF_BOX);
}

// Arrival rank of each lane's particle in its chunk k.  Consecutive lanes
// with the same chunk form a run (the state is sorted by the previous step's
// chunks, so a wave holds ~15 runs); the first lane of a run reserves the
// run's ranks with one atomic.  Ranks are unique for any input: a chunk
// met in several runs gets one atomic per run.
__device__ __forceinline__ int32_t wave_arrival_rank(const bool valid, const int32_t k, int32_t *__restrict__ cnt) {
    const unsigned lane = lane_id();
    const int32_t kprev = __shfl_up(k, 1, 64);
    const int vprev = __shfl_up(valid ? 1 : 0, 1, 64);
    const bool head = lane == 0 || k != kprev || (valid ? 1 : 0) != vprev;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane (all 64 for lane 63)
    const int first = 63 - __clzll((long long)(heads & upto));
    const unsigned long long after = heads & ~upto;
    const int end = after ? __ffsll((long long)after) - 1 : 64;
    int base = 0;
    if (valid && (int)lane == first) base = atomicAdd(&cnt[k], end - first);
    base = __shfl(base, first, 64);
    return base + (int)lane - first;
}

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

// Block-wide exclusive scan of one int per thread (256 threads, wave64).
__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t *total) {
    __shared__ int32_t wsum[kScanThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int32_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int32_t t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int32_t woff = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kScanThreads / 64; k++) {
        if (k < w) woff += wsum[k];
        tot += wsum[k];
    }
    *total = tot;
    return woff + incl - v;
}

enum { MODE_FORCES = 0, MODE_VERLET1 = 1, MODE_VERLET2 = 2, MODE_SZABO = 3, MODE_RTP = 4 };
enum { POT_LJ = 0, POT_HT = 1, POT_RINGS = 2, POT_SZABO = 3 };

struct RingCtx {
    const int32_t *ids;  // sorted slot -> id
    int32_t id_base;     // first id of the system (replica of a batch), 0 for a single system
    const int32_t *rtype;
    const RingType *types;
    const RingInter *inter;
    uint32_t nmax;
    float inv_nmax;
};

__device__ __forceinline__ int32_t ring_of(uint32_t id, const RingCtx &R) {
    // id / nmax via a float estimate and an exact integer correction
    int32_t q = (int32_t)((float)id * R.inv_nmax);
    int32_t r = (int32_t)id - q * (int32_t)R.nmax;
    if (r < 0) q -= 1;
    else if (r >= (int32_t)R.nmax) q += 1;
    r = (int32_t)id - q * (int32_t)R.nmax;
    if (r < 0) q -= 1;
    else if (r >= (int32_t)R.nmax) q += 1;
    return q;
}

// LJ force over distance: fmod/dist = 4 eps (12 s^12/d^14 - 6 s^6/d^8)
// = inv^4 (c12 inv^3 - c6), inv = 1/d^2; 0 outside the cutoff or for self.
__device__ __forceinline__ double lj_scale(const SysParams &P, double r2, bool self_ok) {
    const bool ok = self_ok && (r2 <= P.cut2);
    const double inv = rcp_nr(r2);
    const double inv3 = inv * inv * inv;
    const double fr = inv * inv3 * __builtin_fma(P.lj_c12, inv3, -P.lj_c6);
    return ok ? fr : 0.0;
}

template <int POT>
struct PairAcc {
    double fx = 0.0, fy = 0.0;
    int32_t id_i = 0, ring_i = 0, type_i = 0;  // rings only

    // id_j: particle id of the candidate (read by the rings interaction only)
    __device__ __forceinline__ void add(const SysParams &P, const RingCtx &R, double dx, double dy, bool self_ok,
                                        int id_j) {
        const double r2 = dx * dx + dy * dy;  // sum(dr.^2): two roundings, no FMA
        if constexpr (POT == POT_LJ) {
            const double fr = lj_scale(P, r2, self_ok);
            fx = __builtin_fma(fr, dx, fx);
            fy = __builtin_fma(fr, dy, fy);
        } else if constexpr (POT == POT_SZABO) {
            // calc_interaction(i, j, ::SzaboCfg), src/integration.jl:68-86:
            // ((-f_mod) * (dist - r_eq)) * dr, zero beyond r_max
            if (self_ok && r2 <= P.cut2) {
                const double dist = sqrt(r2);
                const double f_mod = (dist > P.p2) ? P.sz_adh : P.sz_rep;
                const double q = -f_mod * (dist - P.p2);
                fx += q * dx;
                fy += q * dy;
            }
        } else if constexpr (POT == POT_HT) {
            // potential_force(::HarmTruncCfg), src/configs.jl:354-368, term by term
            if (self_ok && r2 <= P.cut2) {
                const double dist = sqrt(r2);
                const double fmod = (r2 < P.eq2) ? -P.p0 * (dist / P.p2 - 1.0) : -P.p1 * (dist / P.p2 - 1.0);
                const double q = fmod / dist;
                fx += q * dx;
                fy += q * dy;
            }
        } else {
            // rings calc_interaction_force, src/rings/integration.jl:47-77
            if (!self_ok) return;
            const int32_t ring_j = ring_of((uint32_t)id_j, R);
            const RingInter *ic = R.inter;
            if (P.typed) ic += type_i + P.ntypes * R.rtype[ring_j];
            if (r2 > ic->max2) return;
            const bool same = ring_j == ring_i;
            if (same) {
                const int32_t diff = abs(id_i - id_j);
                const int32_t np = R.types[type_i].np;
                if (diff == 1 || diff == np - 1) return;
            }
            const double dist = sqrt(r2);
            double fmod;
            if (r2 < ic->eq2) fmod = -ic->k_rep * (dist / ic->dist_eq - 1.0);
            else fmod = same ? 0.0 : -ic->k_atr * (dist / ic->dist_eq - 1.0);
            const double q = fmod / dist;
            fx += q * dx;
            fy += q * dy;
        }
    }
};

__device__ __forceinline__ double2 cdiff(const SysParams &P, double2 a, double2 b) {
    double dx = a.x - b.x, dy = a.y - b.y;
    if (P.periodic) {
        dx = min_image(dx, P.L, P.halfL);
        dy = min_image(dy, P.H, P.halfH);
    }
    return make_double2(dx, dy);
}

__device__ double philox_normal(uint64_t seed, uint64_t step, uint32_t ring) {
    uint32_t c0 = (uint32_t)step, c1 = (uint32_t)(step >> 32), c2 = ring, c3 = 0x6d617669u;
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; r++) {
        philox_round(c0, c1, c2, c3, k0, k1);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    const double u1 = ((double)(((uint64_t)c0 << 21) ^ ((uint64_t)c1 >> 11)) + 1.0) * 0x1p-53;  // (0,1]
    const double u2 = (double)((((uint64_t)c2 << 21) ^ ((uint64_t)c3 >> 11)) & ((1ull << 53) - 1)) * 0x1p-53;
    return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}


// Candidates of one particle as up to three contiguous slot ranges, walked
// as one loop (fewer idle lanes than one loop per range) with the next
// position loaded before the current one is used.
template <int POT, bool WRAP>
__device__ __forceinline__ void scan3(const SysParams &P, const RingCtx &R, PairAcc<POT> &acc,
                                      const double2 *__restrict__ pin, int s, double xi, double yi, int a0, int l0,
                                      int a1, int l1, int a2, int l2) {
    const int tot = l0 + l1 + l2;
    if (tot <= 0) return;
    const int b1 = l0, b2 = l0 + l1;
    const int o0 = a0, o1 = a1 - l0, o2 = a2 - l0 - l1;
    int j = (0 < b1) ? o0 : ((0 < b2) ? o1 : o2);
    double2 pj = pin[j];
    for (int k = 0; k < tot; k++) {
        const int kn = k + 1;
        const int jn = (kn < tot) ? kn + ((kn < b1) ? o0 : ((kn < b2) ? o1 : o2)) : j;
        const double2 pn = pin[jn];
        double dx = xi - pj.x, dy = yi - pj.y;
        if constexpr (WRAP) {
            dx = min_image(dx, P.L, P.halfL);
            dy = min_image(dy, P.H, P.halfH);
        }
        acc.add(P, R, dx, dy, j != s, POT == POT_RINGS ? R.ids[j] - R.id_base : 0);
        j = jn;
        pj = pn;
    }
}

// Position of sorted slot j through a 32-bit byte offset from the array
// base: the load keeps the base in scalar registers and costs one shift per
// candidate (slots < 2^28).
__device__ __forceinline__ double2 load_slot(const double2 *__restrict__ pin, int j) {
    return *(const double2 *)((const char *)pin + ((uint32_t)j << 4));
}

// LJ factor of one candidate with the range test folded into the operand:
// a candidate that is beyond the cutoff (or the lane itself) gets the high
// word of 2^1023 in place of that of r^2, so its reciprocal is ~2^-1023
// and inv^3 underflows to 0: the factor is +-0 and adds nothing (one 32-bit
// select instead of two on the result; in-range factors are unchanged).
__device__ __forceinline__ double lj_scale_masked(const SysParams &P, double r2, bool take) {
    const bool ok = take && (r2 <= P.cut2);
    const double rs = __hiloint2double(ok ? __double2hiint(r2) : 0x7FE00000, __double2loint(r2));
    const double inv = rcp_nr(rs);
    const double inv3 = inv * inv * inv;
    return inv * inv3 * __builtin_fma(P.lj_c12, inv3, -P.lj_c6);
}

// LJ sum over one slot range [a, a + l) of a lane (one chunk row of its 3x3
// block), two candidates per iteration: their loads are issued together and
// their force chains are independent; the sums run in slot order.  The
// iteration count k is the same for every lane still in the loop, so the
// loads take a scalar base (pin + k) plus the lane's fixed byte offset of a
// and need no per-candidate address arithmetic.  SELF: the range holds the
// lane's own slot, as candidate ks.
template <bool WRAP, bool SELF>
__device__ __forceinline__ void lj_range(const SysParams &P, PairAcc<POT_LJ> &acc, const double2 *__restrict__ pin,
                                         double xi, double yi, int a, int l, int ks) {
    const uint32_t aoff = (uint32_t)a << 4;
    auto force = [&](const double2 p, bool take) {
        double dx = xi - p.x, dy = yi - p.y;
        if constexpr (WRAP) {
            dx = min_image(dx, P.L, P.halfL);
            dy = min_image(dy, P.H, P.halfH);
        }
        struct FDr { double f, dx, dy; };
        return FDr{lj_scale_masked(P, dx * dx + dy * dy, take), dx, dy};
    };
    for (int k = 0; k + 2 <= l; k += 2) {
        const char *base = (const char *)pin + ((size_t)(uint32_t)k << 4);
        const double2 p0 = *(const double2 *)(base + aoff);
        const double2 p1 = *(const double2 *)(base + aoff + 16);
        const auto f0 = force(p0, !SELF || k != ks);
        const auto f1 = force(p1, !SELF || k + 1 != ks);
        acc.fx = __builtin_fma(f0.f, f0.dx, acc.fx);
        acc.fy = __builtin_fma(f0.f, f0.dy, acc.fy);
        acc.fx = __builtin_fma(f1.f, f1.dx, acc.fx);
        acc.fy = __builtin_fma(f1.f, f1.dy, acc.fy);
    }
    if (l & 1) {  // the last candidate, l - 1, of an odd range
        const int k = l - 1;
        const auto f0 = force(load_slot(pin, a + k), !SELF || k != ks);
        acc.fx = __builtin_fma(f0.f, f0.dx, acc.fx);
        acc.fy = __builtin_fma(f0.f, f0.dy, acc.fy);
    }
}

// LJ form of scan3: the three chunk rows of the block one after another
// (candidate order as scan3), the middle one holding the lane's own slot s.
template <bool WRAP>
__device__ __forceinline__ void scan3_lj(const SysParams &P, PairAcc<POT_LJ> &acc, const double2 *__restrict__ pin,
                                         int s, double xi, double yi, int a0, int l0, int a1, int l1, int a2,
                                         int l2) {
    lj_range<WRAP, false>(P, acc, pin, xi, yi, a0, l0, 0);
    lj_range<WRAP, true>(P, acc, pin, xi, yi, a1, l1, s - a1);
    lj_range<WRAP, false>(P, acc, pin, xi, yi, a2, l2, 0);
}

// Two-phase form for finite-range interactions: a first pass over the
// candidates only computes r^2 and keeps the in-range candidate numbers
// (uint16, in LDS, column per lane); the FP64 force math then runs on the
// kept candidates only, so lanes of a wave stay busy on pairs that matter.
constexpr int kListCap = 32;

template <int POT, bool WRAP>
__device__ __forceinline__ bool scan3_listed(const SysParams &P, const RingCtx &R, PairAcc<POT> &acc,
                                             const double2 *__restrict__ pin, int s, double xi, double yi, int a0,
                                             int l0, int a1, int l1, int a2, int l2, double list2,
                                             uint16_t *__restrict__ lst, int stride) {
    const int tot = l0 + l1 + l2;
    const int b1 = l0, b2 = l0 + l1;
    const int o0 = a0, o1 = a1 - l0, o2 = a2 - l0 - l1;
    int cnt = 0;
    for (int k = 0; k < tot; k++) {
        const int j = k + ((k < b1) ? o0 : ((k < b2) ? o1 : o2));
        const double2 pj = pin[j];
        double dx = xi - pj.x, dy = yi - pj.y;
        if constexpr (WRAP) {
            dx = min_image(dx, P.L, P.halfL);
            dy = min_image(dy, P.H, P.halfH);
        }
        const double r2 = dx * dx + dy * dy;
        if (j != s && r2 <= list2) {
            if (cnt < kListCap) lst[cnt * stride] = (uint16_t)k;
            cnt++;
        }
    }
    const int m = cnt < kListCap ? cnt : kListCap;
    for (int q = 0; q < m; q++) {
        const int k = lst[q * stride];
        const int j = k + ((k < b1) ? o0 : ((k < b2) ? o1 : o2));
        const double2 pj = pin[j];
        double dx = xi - pj.x, dy = yi - pj.y;
        if constexpr (WRAP) {
            dx = min_image(dx, P.L, P.halfL);
            dy = min_image(dy, P.H, P.halfH);
        }
        acc.add(P, R, dx, dy, true, POT == POT_RINGS ? R.ids[j] - R.id_base : 0);
    }
    return cnt > kListCap;
}

template <int POT>
__device__ __forceinline__ void scan_range(const SysParams &P, const RingCtx &R, PairAcc<POT> &acc,
                                           const double2 *__restrict__ pin, int s, double xi, double yi, bool wrap,
                                           int a, int b) {
    for (int j = a; j < b; j++) {
        const double2 pj = pin[j];
        double dx = xi - pj.x, dy = yi - pj.y;
        if (wrap) {
            dx = min_image(dx, P.L, P.halfL);
            dy = min_image(dy, P.H, P.halfH);
        }
        acc.add(P, R, dx, dy, j != s, POT == POT_RINGS ? R.ids[j] - R.id_base : 0);
    }
}


// Pair forces of sorted slot s over the 3x3 chunk block of chunk c (local
// chunk index; `start` is the scan of this system's chunks, holding global
// sorted slots).  Shared by the single-system and batch kernels.
template <int POT>
__device__ __forceinline__ void pair_gather_chunks(const SysParams &P, const RingCtx &R, PairAcc<POT> &acc,
                                                   const double2 *__restrict__ pin, const int c,
                                                   const int32_t *__restrict__ start, const int s, const bool valid,
                                                   const double2 pi, const double list2, uint16_t *s_list) {
    const int row = c / P.cols, col = c - row * P.cols;
    const bool edge_col = col == 0 || col == P.cols - 1;
    const bool wrap = valid && P.periodic && (row == 0 || row == P.rows - 1 || edge_col);
    const bool split = valid && P.periodic && edge_col;
    if (!__any(split)) {
        // every lane: one slot range per chunk row
        const int c0 = col > 0 ? col - 1 : 0, c1 = col < P.cols - 1 ? col + 1 : P.cols - 1;
        auto row_range = [&](int d, int &a, int &l) {
            int r = row + d;
            bool ok = valid;
            if (P.periodic) r = (r < 0) ? r + P.rows : (r >= P.rows ? r - P.rows : r);
            else ok = ok && r >= 0 && r < P.rows;
            const int base = (ok ? r : 0) * P.cols;
            a = ok ? start[base + c0] : 0;
            l = ok ? start[base + c1 + 1] - a : 0;
        };
        int a0, l0, a1, l1, a2, l2;
        row_range(-1, a0, l0);
        row_range(0, a1, l1);
        row_range(1, a2, l2);
        const bool wrapw = __any(wrap);
        // list entries are 16-bit candidate numbers
        bool direct = !(list2 < INFINITY) || __any(l0 + l1 + l2 > 65535);
        if (!direct) {
            uint16_t *lst = s_list + threadIdx.x;
            const bool over =
                wrapw ? scan3_listed<POT, true>(P, R, acc, pin, s, pi.x, pi.y, a0, l0, a1, l1, a2, l2, list2, lst,
                                                blockDim.x)
                      : scan3_listed<POT, false>(P, R, acc, pin, s, pi.x, pi.y, a0, l0, a1, l1, a2, l2, list2, lst,
                                                 blockDim.x);
            if (__any(over)) {  // a lane had more in-range candidates than the list holds
                acc.fx = 0.0;
                acc.fy = 0.0;
                direct = true;
            }
        }
        if (direct) {
            if constexpr (POT == POT_LJ) {
                if (wrapw) scan3_lj<true>(P, acc, pin, s, pi.x, pi.y, a0, l0, a1, l1, a2, l2);
                else scan3_lj<false>(P, acc, pin, s, pi.x, pi.y, a0, l0, a1, l1, a2, l2);
            } else {
                if (wrapw) scan3<POT, true>(P, R, acc, pin, s, pi.x, pi.y, a0, l0, a1, l1, a2, l2);
                else scan3<POT, false>(P, R, acc, pin, s, pi.x, pi.y, a0, l0, a1, l1, a2, l2);
            }
        }
    } else if (valid) {
        // waves touching the periodic left/right edge: split row windows
        for (int d = -1; d <= 1; d++) {
            int r = row + d;
            r = (r < 0) ? r + P.rows : (r >= P.rows ? r - P.rows : r);
            const int base = r * P.cols;
            int a0, a1, b0 = 0, b1 = 0;
            if (col == 0) {
                a0 = start[base];
                a1 = start[base + 2];
                b0 = start[base + P.cols - 1];
                b1 = start[base + P.cols];
            } else if (col == P.cols - 1) {
                a0 = start[base + P.cols - 2];
                a1 = start[base + P.cols];
                b0 = start[base];
                b1 = start[base + 1];
            } else {
                a0 = start[base + col - 1];
                a1 = start[base + col + 2];
            }
            scan_range<POT>(P, R, acc, pin, s, pi.x, pi.y, wrap, a0, a1);
            if (b1 > b0) scan_range<POT>(P, R, acc, pin, s, pi.x, pi.y, wrap, b0, b1);
        }
    }

}

// walls! for a periodic rectangle, src/integration.jl:309-324.
__device__ __forceinline__ void periodic_wrap(const SysParams &P, double &x, double &y) {
    const double dxc = x - (P.blx + P.halfL), dyc = y - (P.bly + P.halfH);
    const bool ox = fabs(dxc) > P.halfL, oy = fabs(dyc) > P.halfH;
    if (ox) x = x - copysign(P.halfL * 2.0, dxc);
    if (oy) y = y - copysign(P.halfH * 2.0, dyc);
}

// ---------------------------------------------------------------------------
// Walls list (ManyWalls and the non-rectangle geometries), in list order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double jl_sign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x); }

// potential_force(dr, dist, potential) of a potential wall at the processed
// distance (which may be negative), src/configs.jl:354-368, 389-397.
__device__ __forceinline__ void wall_potential_force(const DevWall &w, double drx, double dry, double dist,
                                                     double &fx, double &fy) {
    double q;
    if (w.pot_kind == MAVI_POT_LJ) {
        const double sigma = w.pot[0], eps = w.pot[1];
        const double fmod =
            4.0 * eps * (12.0 * pow(sigma, 12.0) / pow(dist, 13.0) - 6.0 * pow(sigma, 6.0) / pow(dist, 7.0));
        q = fmod / dist;
    } else {
        if (dist > w.pot[3]) return;
        const double fmod = (dist < w.pot[2]) ? -w.pot[0] * (dist / w.pot[2] - 1.0) : -w.pot[1] * (dist / w.pot[2] - 1.0);
        q = fmod / dist;
    }
    fx += q * drx;
    fy += q * dry;
}

__device__ __forceinline__ double process_dist(int mode, double dist, double flag) {
    // src/configs.jl:259-261
    if (mode == MAVI_WMODE_OUTSIDE) return flag * dist;
    if (mode == MAVI_WMODE_INSIDE) return -flag * dist;
    return dist;
}

// calc_walls_forces!, src/integration.jl:230-255: potential walls only.
__device__ __forceinline__ void wall_forces(const SysParams &P, double x, double y, double &fx, double &fy) {
    for (int k = 0; k < P.nwalls; k++) {
        const DevWall w = P.walls[k];
        if (w.kind == MAVI_WK_POTENTIAL_CIRCLE) {
            // signed_pos(point, ::CircleCfg), src/configs.jl:154-162
            const double d0 = x - w.g[0], d1 = y - w.g[1];
            const double dist = sqrt(d0 * d0 + d1 * d1);
            const double h0 = d0 / dist, h1 = d1 / dist;
            const double sd = dist - w.g[2];
            wall_potential_force(w, d0 - h0 * w.g[2], d1 - h1 * w.g[2], process_dist(w.mode, fabs(sd), jl_sign(sd)),
                                 fx, fy);
        } else if (w.kind == MAVI_WK_POTENTIAL_LINES) {
            for (int l = 0; l < w.nlines; l++) {
                // signed_pos(point, ::Line2D), src/configs.jl:119-134
                const DevLine L = P.lines[w.first + l];
                double drx = x - L.p1x, dry = y - L.p1y;
                const double delta_t = drx * L.tx + dry * L.ty;
                if (delta_t > 0.0) {
                    if (delta_t < L.len) {
                        drx = x - (L.p1x + L.tx * delta_t);
                        dry = y - (L.p1y + L.ty * delta_t);
                    } else {
                        drx = x - L.p2x;
                        dry = y - L.p2y;
                    }
                }
                const double dist = sqrt(drx * drx + dry * dry);
                wall_potential_force(w, drx, dry, process_dist(w.mode, dist, 1.0), fx, fy);
            }
        }
    }
}

// walls! over the list, src/integration.jl:271-403.  vx/vy may be null when
// no rigid wall is listed (rings, self-propelled states).
__device__ __forceinline__ void walls_apply(const SysParams &P, double &x, double &y, double *vx, double *vy,
                                            double radius) {
    for (int k = 0; k < P.nwalls; k++) {
        const DevWall w = P.walls[k];
        if (w.kind == MAVI_WK_RIGID_RECT) {
            const double rx = x - w.g[0], ry = y - w.g[1];
            const bool ox = ((rx + radius) > w.g[2]) || ((rx - radius) < 0.0);
            const bool oy = ((ry + radius) > w.g[3]) || ((ry - radius) < 0.0);
            if (ox) *vx = -*vx;
            if (oy) *vy = -*vy;
        } else if (w.kind == MAVI_WK_PERIODIC_RECT) {
            const double hx = w.g[2] / 2.0, hy = w.g[3] / 2.0;
            const double dxc = x - (w.g[0] + hx), dyc = y - (w.g[1] + hy);
            if (fabs(dxc) > hx) x = x - copysign(hx * 2.0, dxc);
            if (fabs(dyc) > hy) y = y - copysign(hy * 2.0, dyc);
        } else if (w.kind == MAVI_WK_RIGID_CIRCLE) {
            // the reflection uses pos.x * pos.y as the reference writes it
            const double m = w.g[2] - radius;
            const double dx = x - w.g[0], dy = y - w.g[1];
            const double dx2 = dx * dx, dy2 = dy * dy;
            const double r2 = dx2 + dy2;
            if (r2 > m * m) {
                const double ux = *vx, uy = *vy;
                *vx = (ux * (dy2 - dx2) - 2.0 * uy * x * y) / r2;
                *vy = (-uy * (dy2 - dx2) - 2.0 * ux * x * y) / r2;
            }
        } else if (w.kind == MAVI_WK_SLIPPERY_LINES) {
            // corrections from the position at the start of this wall, added
            const double x0 = x, y0 = y;
            for (int l = 0; l < w.nlines; l++) {
                const DevLine L = P.lines[w.first + l];
                const double drx = x0 - L.p1x, dry = y0 - L.p1y;
                const double delta_s = drx * L.nx + dry * L.ny;
                if (fabs(delta_s) > radius) continue;
                const double delta_t = drx * L.tx + dry * L.ty;
                bool corner = false;
                double cpx = 0.0, cpy = 0.0;
                if (delta_t > 0.0) {
                    if (delta_t > L.len) {
                        if (delta_t > (L.len + radius)) continue;
                        corner = true;
                        cpx = L.p2x;
                        cpy = L.p2y;
                    }
                } else if (delta_t > -radius) {
                    corner = true;
                    cpx = L.p1x;
                    cpy = L.p1y;
                } else {
                    continue;
                }
                if (corner) {
                    const double c0 = x0 - cpx, c1 = y0 - cpy;
                    const double norm = sqrt(c0 * c0 + c1 * c1);
                    if (norm > radius) continue;
                    const double alpha = radius / norm - 1.0;
                    x = x + alpha * c0;
                    y = y + alpha * c1;
                } else {
                    const double sgn = jl_sign(delta_s);
                    const double alpha = sgn * (radius - sgn * delta_s);
                    x = x + alpha * L.nx;
                    y = y + alpha * L.ny;
                }
            }
        } else if (w.kind == MAVI_WK_SLIPPERY_CIRCLE) {
            // dr through the space's calc_diff (the main wall's)
            const double m = w.g[2] + radius;
            const double2 dr = cdiff(P, make_double2(x, y), make_double2(w.g[0], w.g[1]));
            const double dr_2 = dr.x * dr.x + dr.y * dr.y;
            if (dr_2 > m * m) continue;
            const double dr_norm = sqrt(dr_2);
            const double kk = w.g[2] + radius - dr_norm;
            x = x + kk * dr.x / dr_norm;
            y = y + kk * dr.y / dr_norm;
        }
    }
}

// Integrator epilogue of a pair-force launch for sorted slot s (WALLS: the
// system has a walls list; only kernels instantiated for it carry its code).
// Returns whether walls! moved the particle (second evaluation only).
template <int MODE, bool WALLS = false>
__device__ __forceinline__ bool pair_epilogue(const SysParams &P, const int s, const double2 pi, const double fx_,
                                              const double fy_, double2 *__restrict__ Pst, double2 *__restrict__ Vst,
                                              double2 *__restrict__ Fout, double2 *__restrict__ P2,
                                              double2 *__restrict__ F1, double2 *xout = nullptr) {
    bool moved = false;
    if constexpr (MODE == MODE_FORCES) {
        Fout[s] = make_double2(fx_, fy_);
    } else if constexpr (MODE == MODE_VERLET1) {
        // calc_walls_forces! joins the first evaluation only (src/integration.jl:512)
        double fx = fx_, fy = fy_;
        if constexpr (WALLS) wall_forces(P, pi.x, pi.y, fx, fy);
        // update_verlet!: pos += vel*dt + forces*term (src/integration.jl:424)
        const double2 v = Vst[s];
        P2[s] = make_double2(pi.x + (v.x * P.dt + fx * P.term), pi.y + (v.y * P.dt + fy * P.term));
        F1[s] = make_double2(fx, fy);
    } else {
        // vel += dt/2 * (forces + old_forces) (src/integration.jl:430), then walls!
        const double fx = fx_, fy = fy_;
        const double2 v = Vst[s];
        const double2 fo = F1[s];
        double vx = v.x + P.hdt * (fx + fo.x);
        double vy = v.y + P.hdt * (fy + fo.y);
        double x = pi.x, y = pi.y;
        if constexpr (WALLS) {
            walls_apply(P, x, y, &vx, &vy, P.radius);
            moved = x != pi.x || y != pi.y;
        } else if (P.periodic) {
            periodic_wrap(P, x, y);
        } else {
            // src/integration.jl:271-285
            const double rx = x - P.blx, ry = y - P.bly;
            const bool ox = ((rx + P.radius) > P.L) || ((rx - P.radius) < 0.0);
            const bool oy = ((ry + P.radius) > P.H) || ((ry - P.radius) < 0.0);
            if (ox) vx = -vx;
            if (oy) vy = -vy;
        }
        Pst[s] = make_double2(x, y);
        Vst[s] = make_double2(vx, vy);
        Fout[s] = make_double2(fx, fy);
        if (xout) *xout = make_double2(x, y);
    }
    return moved;
}

__global__ void k_counter_inc(unsigned long long *counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] += 1ull;
}

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// State staging (host id order <-> sorted state)
// ---------------------------------------------------------------------------

// State copy of one system from its id-order staging (pointers start at the
// system's first id / first slot; ids and slots are stored with their bases).
__global__ void k_iota_state(int32_t na, const int32_t *__restrict__ active, const double2 *__restrict__ pos_id,
                             const double2 *__restrict__ vel_id, double2 *__restrict__ P, double2 *__restrict__ V,
                             int32_t *__restrict__ ID, int32_t *__restrict__ slot_of, int32_t id_base,
                             int32_t slot_base) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= na) return;
    int id = active ? active[i] : i;
    P[i] = pos_id[id];
    if (V) V[i] = vel_id[id];
    ID[i] = id_base + id;
    if (slot_of) slot_of[id] = slot_base + i;
}

// Sorted state back to id-order staging (ids are global in a batch).
__global__ void k_unpermute(int32_t na, const int32_t *__restrict__ ID, const double2 *__restrict__ P,
                            const double2 *__restrict__ V, const double2 *__restrict__ F, double2 *__restrict__ pos_id,
                            double2 *__restrict__ vel_id, double2 *__restrict__ force_id) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na) return;
    int id = ID[t];
    if (pos_id) pos_id[id] = P[t];
    if (vel_id) vel_id[id] = V[t];
    if (force_id) force_id[id] = F[t];
}

// Exclusive scan of cnt[0..n) in one pass (decoupled look-back: Merrill and
// Garland, "Single-pass Parallel Prefix Scan with Decoupled Look-back", 2016).
// A block takes its tile from a 64-bit ticket counter that is never reset, so
// tiles are started in order and a block only waits on tiles already
// running.  The ticket's quotient by the tile count is the launch's epoch;
// every published tile state carries it, so no state is cleared between
// launches (and a captured graph needs no memset node).  State word:
// bits 63..34 epoch, 33..32 kind (1 = tile total, 2 = inclusive prefix),
// 31..0 value.  A tile holds kScanThreads * ITEMS counts: 4096 up to 1M
// counts, 16384 above (fewer tiles: fewer block starts and polling rounds on
// large grids, more tiles in flight on small ones).  A look-back that
// does not see its predecessors within a bounded number of polls raises
// MAVI_FLAG_SCAN_STALL and takes 0 for the missing part instead of waiting
// for ever.
constexpr int kScanItemsSmall = 16, kScanItemsLarge = 64;
constexpr int kScanTileMin = kScanThreads * kScanItemsSmall;  // bounds the tile count from above
constexpr int kScanLargeFrom = 1 << 20;
constexpr uint32_t kScanMaxPolls = 1u << 20;
constexpr int kScanLook = 4;  // tiles per lane in one look-back polling round

__device__ __forceinline__ unsigned long long scan_word(uint32_t epoch, uint32_t kind, int32_t v) {
    return ((unsigned long long)epoch << 34) | ((unsigned long long)kind << 32) | (uint32_t)v;
}

template <int ITEMS>
__global__ __launch_bounds__(kScanThreads) void k_scan_lookback(const int32_t *__restrict__ in, int32_t n,
                                                                int32_t *__restrict__ out,
                                                                unsigned long long *__restrict__ ticket,
                                                                unsigned long long *__restrict__ state,
                                                                int32_t ntiles, uint32_t *__restrict__ flags) {
    __shared__ int32_t s_tile, s_prefix;
    __shared__ uint32_t s_epoch;
    if (threadIdx.x == 0) {
        const unsigned long long t = atomicAdd(ticket, 1ull);
        s_tile = (int32_t)(t % (unsigned long long)ntiles);
        s_epoch = (uint32_t)((t / (unsigned long long)ntiles) & 0x3fffffffull);
    }
    __syncthreads();
    const int tile = s_tile;
    const uint32_t epoch = s_epoch;
    const int base = tile * (kScanThreads * ITEMS) + threadIdx.x * ITEMS;
    int32_t v[ITEMS];
    if (base + ITEMS <= n) {
        const int4 *q = reinterpret_cast<const int4 *>(in + base);
#pragma unroll
        for (int k = 0; k < ITEMS / 4; k++) {
            const int4 w = q[k];
            v[4 * k] = w.x;
            v[4 * k + 1] = w.y;
            v[4 * k + 2] = w.z;
            v[4 * k + 3] = w.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) v[k] = (base + k < n) ? in[base + k] : 0;
    }
    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) sum += v[k];
    int32_t total;
    const int32_t local = block_excl_scan(sum, &total);
    if (threadIdx.x == 0)
        __hip_atomic_store(&state[tile], scan_word(epoch, tile == 0 ? 2u : 1u, total), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x < 64) {
        // one wave looks back over windows of 64 * kScanLook tiles (lane l
        // holds tiles j - kScanLook * l - q, q < kScanLook, nearest first): a
        // tile far from the start needs 1 / kScanLook of the polling rounds
        // one tile per lane would, and each round waits on memory once (the
        // lane's loads are all issued before any result is used; a tile
        // before tile 0 loads tile 0's word and ignores it)
        const int lane = threadIdx.x;
        int32_t excl = 0;
        int j = tile - 1;  // nearest tile not yet accounted for
        uint32_t polls = 0;
        bool stalled = false;
        while (j >= 0) {
            unsigned long long w[kScanLook];
#pragma unroll
            for (int q = 0; q < kScanLook; q++) {
                const int t = j - kScanLook * lane - q;
                w[q] = __hip_atomic_load(&state[t < 0 ? 0 : t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            uint32_t kind[kScanLook];
            int32_t val[kScanLook];
            int first_pre = kScanLook;  // nearest tile of this lane holding an inclusive prefix
#pragma unroll
            for (int q = 0; q < kScanLook; q++) {
                const int t = j - kScanLook * lane - q;
                kind[q] = 3u;  // before tile 0: contributes nothing
                val[q] = 0;
                if (t >= 0) {
                    kind[q] = 0u;
                    if ((uint32_t)(w[q] >> 34) == epoch) {
                        kind[q] = (uint32_t)(w[q] >> 32) & 3u;
                        val[q] = (int32_t)(uint32_t)w[q];
                    }
                }
                if (kind[q] == 2u && first_pre == kScanLook) first_pre = q;
            }
            // stop at the nearest inclusive prefix, and never step over a
            // tile that has published nothing yet
            const unsigned long long pre = __ballot(first_pre < kScanLook);
            const int stop = pre ? __ffsll(pre) - 1 : 64;
            const int last_q = lane < stop ? kScanLook - 1 : (lane == stop ? first_pre : -1);
            bool missing = false;
            int32_t part = 0;
#pragma unroll
            for (int q = 0; q < kScanLook; q++) {
                if (q <= last_q) {
                    missing = missing || kind[q] == 0u;
                    part += val[q];
                }
            }
            if (__ballot(missing)) {
                if (++polls >= kScanMaxPolls) {
                    stalled = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
#pragma unroll
            for (int d = 32; d > 0; d >>= 1) part += __shfl_xor(part, d, 64);
            excl += part;
            if (pre) break;
            j -= 64 * kScanLook;
        }
        if (lane == 0) {
            if (tile > 0)
                __hip_atomic_store(&state[tile], scan_word(epoch, 2u, excl + total), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            if (stalled) atomicOr(flags, MAVI_FLAG_SCAN_STALL);
            s_prefix = excl;
        }
    }
    __syncthreads();
    int32_t off = s_prefix + local;
    if (base + ITEMS <= n) {
        int4 *q = reinterpret_cast<int4 *>(out + base);
#pragma unroll
        for (int k = 0; k < ITEMS / 4; k++) {
            int4 w;
            w.x = off;
            w.y = w.x + v[4 * k];
            w.z = w.y + v[4 * k + 1];
            w.w = w.z + v[4 * k + 2];
            off = w.w + v[4 * k + 3];
            q[k] = w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            if (base + k < n) out[base + k] = off;
            off += v[k];
        }
    }
}

// Exclusive scan of cnt[0..n) into start on stream st (tile size by n; one
// ticket counter and state array per scanned array, with at least
// ceil(n / kScanTileMin) state words).
inline void launch_scan(const int32_t *cnt, int32_t n, int32_t *start, unsigned long long *ticket,
                        unsigned long long *state, uint32_t *flags, hipStream_t st) {
    if (n > kScanLargeFrom) {
        const int ntiles = (n + kScanThreads * kScanItemsLarge - 1) / (kScanThreads * kScanItemsLarge);
        hipLaunchKernelGGL(k_scan_lookback<kScanItemsLarge>, dim3(ntiles), dim3(kScanThreads), 0, st, cnt, n, start,
                           ticket, state, ntiles, flags);
    } else {
        const int ntiles = (n + kScanTileMin - 1) / kScanTileMin;
        hipLaunchKernelGGL(k_scan_lookback<kScanItemsSmall>, dim3(ntiles), dim3(kScanThreads), 0, st, cnt, n, start,
                           ticket, state, ntiles, flags);
    }
}

// Place every particle at start[chunk] + arrival rank (old slot src, chunk
// scell and id sid of each arrival-order slot); reset the counts for
// the next binning; advance the step counter of the device rotational noise
// when it is given.  With stamp: mark every chunk whose 3x3 block holds a
// chunk stamped in this binning (stamp == epoch + 1) with needc = epoch + 1;
// k_finalize advances the epoch afterwards.
__global__ void k_scatter(SysParams P, const int32_t *__restrict__ key, const int32_t *__restrict__ rank,
                          const int32_t *__restrict__ start, int32_t *__restrict__ src, int32_t *__restrict__ scell,
                          const int32_t *__restrict__ ID, int32_t *__restrict__ sid,
                          int32_t *__restrict__ cnt, unsigned long long *__restrict__ counter,
                          const int32_t *__restrict__ stamp, const uint32_t *__restrict__ epoch,
                          int32_t *__restrict__ needc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P.na) {
        const int k = key[i];
        const int t = start[k] + rank[i];
        src[t] = i;
        scell[t] = k;
        sid[t] = ID[i];
    }
    if (i <= P.ncell) cnt[i] = 0;
    if (i == 0 && counter) counter[0] += 1ull;
    if (stamp && i < P.ncell) {
        const int32_t e = (int32_t)(epoch[0] + 1u);
        const int row = i / P.cols, col = i - row * P.cols;
        bool need = false;
        for (int dr = -1; dr <= 1; dr++) {
            int r = row + dr;
            if (P.periodic) r = r < 0 ? r + P.rows : (r >= P.rows ? r - P.rows : r);
            else if (r < 0 || r >= P.rows) continue;
            for (int dc = -1; dc <= 1; dc++) {
                int cc = col + dc;
                if (P.periodic) cc = cc < 0 ? cc + P.cols : (cc >= P.cols ? cc - P.cols : cc);
                else if (cc < 0 || cc >= P.cols) continue;
                need = need || stamp[r * P.cols + cc] == e;
            }
        }
        if (need) needc[i] = e;
    }
}

// Unwrap (update_continuos_pos!, :118-136), centre of mass and shoelace area
// of ring r from positions x(i) (calc_area, :103-116).
template <typename PosFn>
__device__ __forceinline__ double ring_geometry(const SysParams &P, int np, PosFn x, double2 *__restrict__ c,
                                                double2 *cm) {
    for (int i = 0; i < P.nmax; i++) c[i] = x(i);
    if (P.periodic) {
        for (int i = 1; i < np; i++) {
            const double2 d = cdiff(P, x(i), x(i - 1));
            c[i] = make_double2(c[i - 1].x + d.x, c[i - 1].y + d.y);
        }
    }
    double sx = 0.0, sy = 0.0;
    for (int i = 0; i < np; i++) {
        sx += c[i].x;
        sy += c[i].y;
    }
    *cm = make_double2(sx / (double)np, sy / (double)np);
    double area = 0.0;
    for (int i = 0; i < np - 1; i++) area += c[i].x * c[i + 1].y - c[i].y * c[i + 1].x;
    area += c[np - 1].x * c[0].y - c[np - 1].y * c[0].x;
    return area / 2.0;
}

__global__ void k_rings_init(SysParams P, const int32_t *__restrict__ rtype, const RingType *__restrict__ types,
                             const double2 *__restrict__ pos_id, double2 *__restrict__ cms,
                             double2 *__restrict__ cms_next, double2 *__restrict__ cpos, double *__restrict__ areas) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= P.nrings) return;
    if (P.rmask && !P.rmask[r]) return;  // inactive rings keep their info
    const int np = types[P.typed ? rtype[r] : 0].np;
    const double2 *x = pos_id + (int64_t)r * P.nmax;
    double2 cm;
    areas[r] = ring_geometry(P, np, [&](int i) { return x[i]; }, cpos + (int64_t)r * P.nmax, &cm);
    cms[r] = cms_next[r] = cm;
}

// One ring per group of lanes of a wave (lane0 = the group's first lane in
// the wave, i = particle index inside the ring).  Every reduction the
// reference does as a sequential loop (unwrap, centre of mass, shoelace
// area, mean velocity) is done here as the same left-to-right loop over
// values broadcast from the owning lanes, so the result is bit-identical to
// the one-thread-per-ring form.  slot_of and cpos start at the system's
// first particle id, pol/areas/cms/cms_next at its first ring; noise (host
// draws) or the Philox stream (seed, step) give the rotational noise of ring rr.
template <bool UPDATE, bool WALLS = false, bool RATES = false>
__device__ __forceinline__ void ring_lanes(const SysParams &P, const RingType &T, const bool ring_ok, const int rr,
                                           const int lane0, const int i, const int32_t *__restrict__ slot_of,
                                           double2 *__restrict__ Pst, double2 *__restrict__ F,
                                           double *__restrict__ pol, double *__restrict__ areas,
                                           double2 *__restrict__ cms, double2 *__restrict__ cms_next,
                                           double2 *__restrict__ cpos, const double *__restrict__ noise,
                                           const uint64_t seed, const uint64_t step,
                                           int32_t *__restrict__ nkey = nullptr, int32_t *__restrict__ nrank = nullptr,
                                           int32_t *__restrict__ ncnt = nullptr, uint32_t *__restrict__ nflags = nullptr,
                                           const RingRates RR = RingRates{}, const int32_t kbase = 0) {
    const int r = rr;
    const int np = T.np;
    const bool act = ring_ok && i < np;
    const int64_t base = (int64_t)rr * P.nmax;
    const int slot = act ? slot_of[base + i] : 0;
    const double2 x = act ? Pst[slot] : make_double2(0.0, 0.0);
    (void)lane0;  // the ring's lanes start at thread tb below

    if constexpr (UPDATE) {
        if (ring_ok && i == 0) cms[r] = cms_next[r];
    }
    // The ring's values that other lanes need (neighbours' positions and
    // spring forces, and each sequential sum's terms) go through LDS: every
    // lane writes its own, then reads those it needs (broadcast reads for
    // the sums, in ring order), in place of cross-lane shuffles.  A ring's
    // lanes lie in one wave, so a wave-scope fence orders the writes before
    // the reads; each array is written again only after its last reads.
    __shared__ double2 s_a[256 + 32], s_b[256 + 32], s_c[256 + 32];
    __shared__ double s_t[256 + 32];
    const int me = (int)threadIdx.x, tb = me - i;  // tb: the ring's first lane
    auto lds_sync = []() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };

    // neighbours along the ring
    const int inext = (i == np - 1) ? 0 : i + 1;
    const int iprev = (i == 0) ? np - 1 : i - 1;
    s_a[me] = x;
    lds_sync();
    const double2 xn = s_a[tb + inext];
    const double2 xp = s_a[tb + iprev];

    // update_continuos_pos!: c_0 = x_0, c_i = c_{i-1} + calc_diff(x_i, x_{i-1})
    const double2 d = (i == 0) ? x : cdiff(P, x, xp);
    double2 *const s_d = s_b;
    s_d[me] = d;
    lds_sync();
    double cx = 0.0, cy = 0.0;
#pragma unroll 4
    for (int k = 0; k < np; k++) {
        const double2 dk = s_d[tb + k];
        if (k == 0) {
            cx = dk.x;
            cy = dk.y;
        } else if (k <= i) {
            cx = cx + dk.x;
            cy = cy + dk.y;
        }
    }
    if (!P.periodic) {  // update_continuos_pos! is a no-op: continuous = stored positions
        cx = x.x;
        cy = x.y;
    }
    if (act) cpos[base + i] = make_double2(cx, cy);
    // centre of mass and shoelace area, sums in ring order
    s_c[me] = make_double2(cx, cy);
    lds_sync();
    const double2 cn = s_c[tb + inext];
    const double term = cx * cn.y - cy * cn.x;
    s_t[me] = term;
    lds_sync();
    double sx = 0.0, sy = 0.0, area = 0.0;
#pragma unroll 4
    for (int k = 0; k < np; k++) {
        const double2 ck = s_c[tb + k];
        sx += ck.x;
        sy += ck.y;
        area += s_t[tb + k];
    }
    area = area / 2.0;
    if (ring_ok && i == 0) {
        cms_next[r] = make_double2(sx / (double)np, sy / (double)np);
        areas[r] = area;
    }

    // springs: lane i holds f_i = spring(i, i+1); particle i gets +f_1 then
    // -f_n (i = 0) or -f_{i-1} then +f_i, then the area term
    const double2 ds = cdiff(P, x, xn);
    const double dist = sqrt(ds.x * ds.x + ds.y * ds.y);
    const double q = -T.kspring * (dist - T.lspring) / dist;
    const double2 fi = make_double2(q * ds.x, q * ds.y);
    s_b[me] = fi;  // the unwrap's reads of s_b are done
    lds_sync();
    const double2 fprev = s_b[tb + iprev];
    double area0 = (double)np * T.lspring / T.p0;
    area0 = area0 * area0;
    const double fa = T.karea * (area - area0);
    const double2 da = cdiff(P, xn, xp);
    double2 Fi = act ? F[slot] : make_double2(0.0, 0.0);
    if (i == 0) {
        Fi.x += fi.x;
        Fi.y += fi.y;
        Fi.x -= fprev.x;
        Fi.y -= fprev.y;
    } else {
        Fi.x -= fprev.x;
        Fi.y -= fprev.y;
        Fi.x += fi.x;
        Fi.y += fi.y;
    }
    Fi.x -= fa * (da.y / 2.0);
    Fi.y -= fa * (-da.x / 2.0);
    if constexpr (UPDATE && WALLS) {
        if (act) wall_forces(P, x.x, x.y, Fi.x, Fi.y);  // calc_walls_forces!
    }
    if (act) F[slot] = Fi;
    if constexpr (RATES) {
        // the learned propagator's inputs: u = vo p + mu F per particle, and
        // per ring the rate of its area, 1/2 sum u x (x_next - x_prev), over
        // the target area and the alignment of p with the mean of u
        const double th = ring_ok ? pol[rr] : 0.0;
        double sn, cs;
        sincos(th, &sn, &cs);
        const double2 u = make_double2(T.vo * cs + T.mu * Fi.x, T.vo * sn + T.mu * Fi.y);
        if (act && ring_ok) RR.U[slot] = u;
        double2 *const s_u = s_a;  // the neighbour reads of s_a are done
        s_u[me] = u;
        s_t[me] = u.x * da.y - u.y * da.x;  // the area sums' reads of s_t are done
        lds_sync();
        if (ring_ok && i == 0) {
            double ux = 0.0, uy = 0.0, ar = 0.0;
            for (int k = 0; k < np; k++) {
                const double2 uk = s_u[tb + k];
                ux += uk.x;
                uy += uk.y;
                ar += s_t[tb + k];
            }
            RR.rate[rr] = 0.5 * ar / area0;
            const double vx = ux / (double)np, vy = uy / (double)np;
            const double speed = sqrt(vx * vx + vy * vy);
            double cr = 0.0;
            if (speed != 0.0) {
                cr = (cs * vy - sn * vx) / speed;
                if (fabs(cr) > 1.0) cr = cr > 0.0 ? 1.0 : -1.0;
            }
            RR.omega[rr] = asin(cr);
        }
    }
    if constexpr (!UPDATE) return;

    // update!: v = vo pol + mu F, x += v dt (then walls!), polarisation relaxation
    const double theta = ring_ok ? pol[rr] : 0.0;
    double sn, cs;
    sincos(theta, &sn, &cs);
    const double vx = T.vo * cs + T.mu * Fi.x;
    const double vy = T.vo * sn + T.mu * Fi.y;
    double2 *const s_v = s_a;  // the neighbour reads of s_a are done
    s_v[me] = make_double2(vx, vy);
    lds_sync();
    double vcx = 0.0, vcy = 0.0;
    if (ring_ok && i == 0) {  // only the ring's first lane uses the mean velocity
#pragma unroll 4
        for (int k = 0; k < np; k++) {
            const double2 vk = s_v[tb + k];
            vcx += vk.x;
            vcy += vk.y;
        }
    }
    double px = x.x + vx * P.dt, py = x.y + vy * P.dt;
    if constexpr (WALLS) {
        if (act) walls_apply(P, px, py, nullptr, nullptr, T.wradius);
    } else if (P.periodic) {
        periodic_wrap(P, px, py);
    }
    if (act) Pst[slot] = make_double2(px, py);
    if (ncnt) {
        // the next binning's first pass (chunk and arrival rank of the final
        // position, as k_bin_count makes them; chunks numbered from kbase);
        // every lane of the wave takes part in the arrival ranks
        uint32_t flag = 0;
        int32_t k = -1;
        if (act) k = kbase + cell_of(P, px, py, &flag);
        if (flag) atomicOr(nflags, flag);
        const int32_t rk = wave_arrival_rank(act, k, ncnt);
        if (act) {
            nkey[slot] = k;
            nrank[slot] = rk;
        }
    }
    if (ring_ok && i == 0) {
        vcx = vcx / (double)np;
        vcy = vcy / (double)np;
        const double speed = sqrt(vcx * vcx + vcy * vcy);
        double cr = 0.0;
        if (speed != 0.0) {
            cr = (cs * vcy - sn * vcx) / speed;
            if (fabs(cr) > 1.0) cr = cr > 0.0 ? 1.0 : -1.0;
        }
        const double z = noise ? noise[r] : (T.rot != 0.0 ? philox_normal(seed, step, (uint32_t)r) : 0.0);
        const double d_theta = 1.0 / T.relax * asin(cr) * P.dt + sqrt(2.0 * T.rot * P.dt) * z;
        pol[r] = theta + d_theta;
    }
}

template <bool UPDATE, bool WALLS = false, bool RATES = false>
__global__ __launch_bounds__(256) void k_rings_lanes(
    SysParams P, const int32_t *__restrict__ rtype, const RingType *__restrict__ types,
    const int32_t *__restrict__ slot_of, double2 *__restrict__ Pst, double2 *__restrict__ F, double *__restrict__ pol,
    double *__restrict__ areas, double2 *__restrict__ cms, double2 *__restrict__ cms_next, double2 *__restrict__ cpos,
    const double *__restrict__ noise, uint64_t seed, const unsigned long long *__restrict__ counter,
    int32_t *__restrict__ nkey, int32_t *__restrict__ nrank, int32_t *__restrict__ ncnt, uint32_t *__restrict__ nflags,
    RingRates RR) {
    if (RATES && blockIdx.x == 0 && threadIdx.x == 0) RR.gate[0] = 0u;
    // floor(64 / nmax) rings per wave, nmax lanes each (3 rings of 20 per
    // wave: 60 busy lanes instead of 40 with one ring per 32 lanes)
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int w = (int)(threadIdx.x & 63);
    const int per = 64 / P.nmax;
    const int grp = w / P.nmax, i = w - grp * P.nmax;
    const int r = wave * per + grp;
    const bool ring_ok = grp < per && r < P.nrings && (!P.rmask || P.rmask[r]);  // get_rings_ids
    const int rr = ring_ok ? r : 0;
    const RingType T = types[P.typed ? rtype[rr] : 0];
    ring_lanes<UPDATE, WALLS, RATES>(P, T, ring_ok, rr, grp * P.nmax, i, slot_of, Pst, F, pol, areas, cms, cms_next,
                                     cpos, noise, seed, counter[0], nkey, nrank, ncnt, nflags, RR);
}

}  // namespace
