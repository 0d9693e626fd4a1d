// mavi_learned.hip -- the learned propagator for rings systems: one network
// evaluation stands for k plain steps, gated by an overlap/energy check with
// the plain FP64 steps as fallback.
//
// A learned step from the state x(t) (one hipGraph, then one 4-byte read of
// the gate word):
//   1. binning + forces!  (plain kernels: pair, springs, area; FP64; the
//      ring kernel also gives u = v0 p + mu F and the per-ring rates, or
//      k_ring_rates for rings of more than 32 particles)
//   2. k_features: per particle, in the frame of its ring's
//      polarisation (rotation-canonical), 32 self features (ring type, area
//      deviation and its rate, the alignment term, the two ring neighbours
//      and their velocities relative to it, the overdamped velocity
//      u = v0 p + mu F) and the 32 nearest particles within r_feat (offset,
//      ring membership +1 same / -1 other, rate of change of the distance),
//      sorted by (distance, id); rows in particle-id order, BF16
//   3. k_mlp_fused: correction y = MLP(features)
//   4. k_apply: displacement = k dt (v0 p + mu F) + s_t R^T y  (the explicit
//      k dt step plus the learned correction), candidate positions x'
//   5. gate: |displacement| <= max_disp, every listed pair (all pairs within
//      r_feat; more than 32 fails) stays >= d_min apart (default 0.3 dist_eq,
//      below what the plain dynamics reaches), and no ring's spring + area
//      energy rises by more than e_tol.  Pairs that were farther than
//      r_feat cannot come closer than d_min when max_disp <= (r_feat - d_min)/2
//   6. accepted: x <- x', theta += k dt/tau asin(p x v_cm/|v_cm|) +
//      sqrt(2 D_r k dt) z (one normal per ring from the device noise stream);
//      rejected: the host runs k plain steps instead.
//
// Training data (mavi_learned_collect): features at x(t), k plain steps with
// the device noise, target = R (x(t + k dt) - x(t) - k dt (v0 p + mu F)) / s_t.
#include <math.h>

#include <string>
#include <vector>

#include "mavi_device.h"
#include "mavi_host.h"
#include "mavi_mlp.h"

namespace mavi {
hipError_t mlp_forward_fused(mavi_mlp *p, const uint16_t *X, int M, float *Y, hipStream_t st);
hipError_t mlp_dataset_reserve(mavi_mlp *p, int64_t cap);
}  // namespace mavi

namespace {

constexpr int kNb = 32;    // neighbour slots
constexpr int kSelf = 32;  // self features (8 used)
static_assert(kSelf + 4 * kNb == kMlpIn, "feature width");

enum : uint32_t { GATE_DISP = 1, GATE_OVERLAP = 2, GATE_ENERGY = 4, GATE_NEIGH = 8, GATE_NONFINITE = 16 };

struct LearnParams {
    int32_t k;
    int32_t use_net;
    double rf2;        // exact: dist <= r_feat <=> r2 <= rf2
    double kdt;        // k dt
    double s_t;        // target scale
    double dmin2;      // exact: dist < d_min <=> r2 < dmin2
    double max_disp2;  // |displacement|^2 limit
    double e_tol;      // ring energy rise limit
    uint64_t seed;     // noise stream of learned steps
};

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ uint16_t d2bf(double v) { return f2bf((float)v); }

__device__ __forceinline__ double2 rot(double cs, double sn, double2 v) {
    return make_double2(cs * v.x + sn * v.y, -sn * v.x + cs * v.y);
}
__device__ __forceinline__ double2 rot_back(double cs, double sn, double2 v) {
    return make_double2(cs * v.x - sn * v.y, sn * v.x + cs * v.y);
}

__device__ __forceinline__ bool key_less(double da, int ia, double db, int ib) {
    return da < db || (da == db && ia < ib);
}

// Sum over the 32 lanes of a half-wave (tree order).
__device__ __forceinline__ double half_sum(double v) {
#pragma unroll
    for (int d = 16; d > 0; d >>= 1) v += __shfl_xor(v, d, 32);
    return v;
}

// Per ring (32 lanes, lane i = particle i): the overdamped velocity
// u = v0 (cos th, sin th) + mu F of each particle (update!,
// src/rings/integration.jl:300-351, kept by sorted slot for the features),
// the area rate dA/dt = 1/2 sum_i cross(u_i, x_{i+1} - x_{i-1}) over A0, and
// the alignment term asin(p x v_cm / |v_cm|) of update!.  Every slot is a
// ring particle (full rings).  Also clears the gate word of the step (no
// kernel before it sets it).
__global__ __launch_bounds__(256) void k_ring_rates(SysParams P, const int32_t *__restrict__ slot_of,
                                                    const double2 *__restrict__ Pst, const double2 *__restrict__ F,
                                                    double2 *__restrict__ U, const double *__restrict__ pol,
                                                    const int32_t *__restrict__ rtype,
                                                    const RingType *__restrict__ types, double *__restrict__ rate,
                                                    double *__restrict__ omega, uint32_t *__restrict__ gate) {
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    if (gt == 0) gate[0] = 0u;
    const int r = gt >> 5, i = gt & 31;
    const bool ring_ok = r < P.nrings;
    const int rr = ring_ok ? r : 0;
    const RingType T = types[P.typed ? rtype[rr] : 0];
    const int np = T.np;
    const bool act = i < np;
    const int slot = act ? slot_of[(int64_t)rr * P.nmax + i] : 0;
    double sn, cs;
    sincos(pol[rr], &sn, &cs);
    double2 u = make_double2(0.0, 0.0);
    if (act) {
        const double2 f = F[slot];
        u = make_double2(T.vo * cs + T.mu * f.x, T.vo * sn + T.mu * f.y);
        if (ring_ok) U[slot] = u;
    }
    const double2 x = act ? Pst[slot] : make_double2(0.0, 0.0);
    const int inext = (i == np - 1) ? 0 : i + 1, iprev = (i == 0) ? np - 1 : i - 1;
    const double2 xn = make_double2(__shfl(x.x, inext, 32), __shfl(x.y, inext, 32));
    const double2 xp = make_double2(__shfl(x.x, iprev, 32), __shfl(x.y, iprev, 32));
    const double2 d = cdiff(P, xn, xp);
    const double da = 0.5 * half_sum(act ? u.x * d.y - u.y * d.x : 0.0);
    const double vx = half_sum(u.x) / (double)np, vy = half_sum(u.y) / (double)np;
    if (!ring_ok || i != 0) return;
    double area0 = (double)np * T.lspring / T.p0;
    area0 = area0 * area0;
    rate[rr] = da / area0;
    const double speed = sqrt(vx * vx + vy * vy);
    double cr = 0.0;
    if (speed != 0.0) {
        cr = (cs * vy - sn * vx) / speed;
        if (fabs(cr) > 1.0) cr = cr > 0.0 ? 1.0 : -1.0;
    }
    omega[rr] = asin(cr);
}

__device__ __forceinline__ uint32_t pack2(double a, double b) {
    return (uint32_t)d2bf(a) | ((uint32_t)d2bf(b) << 16);
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
    return v;
}

// Features of every particle (one wavefront per block, one lane per sorted
// slot, row = particle id).  The wave's 64 consecutive slots sit in one or
// two chunk rows, so the candidates of all its lanes in chunk rows r-1, r,
// r+1 form three contiguous slot windows: they are staged in LDS (positions
// and ids) and scanned from there; ranges outside a window (periodic wrap)
// are read from global memory.  One pass keeps the 8 nearest in a sorted
// register list (insertion by compare-exchange, no dynamic indexing); a
// particle with more than 8 neighbours within r_feat continues with passes
// that each select the next (distance, id) key.  Rows are written with
// 16-byte stores.
constexpr int kFeatWG = 64;
constexpr int kWin = 192;  // staged slots per window
constexpr int kReg = 8;    // neighbours kept in registers by the first pass

struct Cand {
    double d2;
    int id, j;
};

__device__ __forceinline__ void cand_insert(Cand (&k)[kReg], Cand c) {
#pragma unroll
    for (int t = 0; t < kReg; t++) {
        const bool lt = key_less(c.d2, c.id, k[t].d2, k[t].id);
        const Cand o = k[t];
        k[t].d2 = lt ? c.d2 : o.d2;
        k[t].id = lt ? c.id : o.id;
        k[t].j = lt ? c.j : o.j;
        c.d2 = lt ? o.d2 : c.d2;
        c.id = lt ? o.id : c.id;
        c.j = lt ? o.j : c.j;
    }
}

__global__ __launch_bounds__(kFeatWG) void k_features(
    SysParams P, LearnParams L, const double2 *__restrict__ Pst, const int32_t *__restrict__ ID,
    const int32_t *__restrict__ slot_of, const int32_t *__restrict__ scell, const int32_t *__restrict__ start,
    const double2 *__restrict__ U, const double *__restrict__ pol, const double *__restrict__ areas,
    const int32_t *__restrict__ rtype, const RingType *__restrict__ types, const double *__restrict__ rate,
    const double *__restrict__ omega, uint16_t *__restrict__ X, int32_t *__restrict__ nbr,
    uint8_t *__restrict__ qprev, double2 *__restrict__ euler, uint32_t *__restrict__ gate) {
    __shared__ double2 spos[3][kWin];
    __shared__ int32_t sid[3][kWin];
    const int lane = threadIdx.x;
    const int s = blockIdx.x * kFeatWG + lane;
    const bool valid = s < P.na;
    const int sv = valid ? s : P.na - 1;
    const int c = scell[sv];
    const int row = c / P.cols, col = c - row * P.cols;
    int wa[3], wb[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int rr = row + d - 1;
        const bool ok = valid && rr >= 0 && rr < P.rows;
        const int a = ok ? start[rr * P.cols + (col > 0 ? col - 1 : 0)] : 0x7fffffff;
        const int b = ok ? start[rr * P.cols + (col < P.cols - 1 ? col + 1 : P.cols - 1) + 1] : -1;
        wa[d] = wave_min(a);
        wb[d] = wave_max(b);
        if (wb[d] - wa[d] > kWin || wb[d] <= wa[d]) wb[d] = wa[d] = 0;  // not staged
    }
    {  // all window loads in flight, then the LDS stores
        constexpr int kPer = kWin / kFeatWG;
        // plain doubles: an array of double2 stays in scratch memory here
        double vx[3][kPer], vy[3][kPer];
        int32_t vi[3][kPer];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int u = 0; u < kPer; u++) {
                const int q = lane + u * kFeatWG;
                const int j = q < wb[d] - wa[d] ? wa[d] + q : 0;
                const double2 p = Pst[j];
                vx[d][u] = p.x;
                vy[d][u] = p.y;
                vi[d][u] = ID[j];
            }
        // q < kWin always; slots past a window's length get slot 0's values
        // and are never read
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int u = 0; u < kPer; u++) {
                const int q = lane + u * kFeatWG;
                spos[d][q] = make_double2(vx[d][u], vy[d][u]);
                sid[d][q] = vi[d][u];
            }
    }
    __syncthreads();
    if (!valid) return;

    const int id = ID[s];
    const int ring = id / P.nmax, i = id - ring * P.nmax;
    const RingType T = types[P.typed ? rtype[ring] : 0];
    const int np = T.np;
    const double2 xi = Pst[s];
    double sn, cs;
    sincos(pol[ring], &sn, &cs);

    // candidate ranges of the 3x3 chunk block (periodic grid >= 3x3): per
    // block row the three chunks' slots are consecutive, one run (two where
    // the columns wrap), each either inside its row's staged window (LDS) or
    // not (global).  The kept neighbours are ordered by (distance, id), so
    // the visiting order does not matter.
    int rb[6], re[6];
    bool inl[6];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int r = row + d - 1;
        r = r < 0 ? r + P.rows : (r >= P.rows ? r - P.rows : r);
        const int *sr = start + r * P.cols;
        int a0 = col - 1, b0 = col + 1, a1 = 0, b1 = -1;  // columns [a0, b0] and [a1, b1]
        if (col == 0) {
            a0 = 0;
            b0 = 1;
            a1 = b1 = P.cols - 1;
        } else if (col == P.cols - 1) {
            a0 = col - 1;
            b0 = col;
            a1 = b1 = 0;
        }
        rb[2 * d] = sr[a0];
        re[2 * d] = sr[b0 + 1];
        rb[2 * d + 1] = b1 >= a1 ? sr[a1] : 0;
        re[2 * d + 1] = b1 >= a1 ? sr[b1 + 1] : 0;
#pragma unroll
        for (int u = 0; u < 2; u++)
            inl[2 * d + u] = rb[2 * d + u] >= wa[d] && re[2 * d + u] <= wb[d];
    }
    // pass 1: the kReg nearest within r_feat, sorted
    Cand kn[kReg];
#pragma unroll
    for (int t = 0; t < kReg; t++) kn[t] = Cand{INFINITY, 0x7fffffff, -1};
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int d = k / 2;
        for (int j = rb[k]; j < re[k]; j++) {
            if (j == s) continue;
            double2 xj;
            int idj;
            if (inl[k]) {
                xj = spos[d][j - wa[d]];
                idj = sid[d][j - wa[d]];
            } else {
                xj = Pst[j];
                idj = ID[j];
            }
            const double dx = min_image(xj.x - xi.x, P.L, P.halfL);
            const double dy = min_image(xj.y - xi.y, P.H, P.halfH);
            const double d2 = dx * dx + dy * dy;
            if (d2 <= L.rf2) {
                cnt++;
                cand_insert(kn, Cand{d2, idj, j});
            }
        }
    }

    // self features
    uint4 *xrow = reinterpret_cast<uint4 *>(X + (int64_t)id * kMlpIn);
    double area0 = (double)np * T.lspring / T.p0;
    area0 = area0 * area0;
    const int base = ring * P.nmax;
    const int sp = slot_of[base + (i == 0 ? np - 1 : i - 1)], sx = slot_of[base + (i == np - 1 ? 0 : i + 1)];
    const double2 dp = rot(cs, sn, cdiff(P, Pst[sp], xi));
    const double2 dn = rot(cs, sn, cdiff(P, Pst[sx], xi));
    const double2 u = U[s];
    euler[id] = make_double2(L.kdt * u.x, L.kdt * u.y);
    const double2 ur = rot(cs, sn, u);
    const double2 up = U[sp], un = U[sx];
    const double2 dup = rot(cs, sn, make_double2(up.x - u.x, up.y - u.y));
    const double2 dun = rot(cs, sn, make_double2(un.x - u.x, un.y - u.y));
    const double tflag = P.typed ? (double)rtype[ring] : 0.0;
    xrow[0] = make_uint4(pack2(tflag, (areas[ring] - area0) / area0), pack2(dp.x, dp.y), pack2(dn.x, dn.y),
                         pack2(ur.x, ur.y));
    xrow[1] = make_uint4(pack2(dup.x, dup.y), pack2(dun.x, dun.y), pack2(rate[ring], omega[ring]), 0u);
    // self features 16..31 stay the zeros of the allocation

    // neighbour features; the row's slots past the list are zero (those this
    // row's previous list used past the new one are cleared below)
    int2 *nb2 = reinterpret_cast<int2 *>(nbr + (int64_t)id * kNb);
    int jprev = -1;
    auto emit = [&](int q, int j, int idj, double d2, uint32_t &w0, uint32_t &w1) {
        const double2 dl = cdiff(P, Pst[j], xi);
        const double2 dd = rot(cs, sn, dl);
        const double2 uj = U[j];
        const double rdot = ((uj.x - u.x) * dl.x + (uj.y - u.y) * dl.y) / sqrt(d2);
        const uint32_t a0 = pack2(dd.x, dd.y);
        const uint32_t a1 = (idj / P.nmax == ring ? 0x3f80u : 0xbf80u) | ((uint32_t)d2bf(rdot) << 16);
        if (q & 1) {
            xrow[4 + q / 2] = make_uint4(w0, w1, a0, a1);
            nb2[q / 2] = make_int2(jprev, j);
        } else {
            w0 = a0;
            w1 = a1;
            jprev = j;
        }
    };
    uint32_t w0 = 0u, w1 = 0u;
    const int nreg = cnt < kReg ? cnt : kReg;
#pragma unroll
    for (int t = 0; t < kReg; t++)
        if (t < nreg) emit(t, kn[t].j, kn[t].id, kn[t].d2, w0, w1);
    int q = nreg;
    if (cnt > kReg) {
        // rare: further neighbours by selection passes above the last key
        double pd = kn[kReg - 1].d2;
        int pid = kn[kReg - 1].id;
        for (; q < cnt && q < kNb; q++) {
            double bd = INFINITY;
            int bid = 0x7fffffff, bj = -1;
            for (int k = 0; k < 6; k++) {
                for (int j = rb[k]; j < re[k]; j++) {
                    if (j == s) continue;
                    const double2 xj = Pst[j];
                    const double dx = min_image(xj.x - xi.x, P.L, P.halfL);
                    const double dy = min_image(xj.y - xi.y, P.H, P.halfH);
                    const double d2 = dx * dx + dy * dy;
                    if (!(d2 <= L.rf2)) continue;
                    const int idj = ID[j];
                    if (key_less(pd, pid, d2, idj) && key_less(d2, idj, bd, bid)) {
                        bd = d2;
                        bid = idj;
                        bj = j;
                    }
                }
            }
            emit(q, bj, bid, bd, w0, w1);
            pd = bd;
            pid = bid;
        }
        if (cnt > kNb) atomicOr(gate, GATE_NEIGH);
    }
    if (q & 1) xrow[4 + q / 2] = make_uint4(w0, w1, 0u, 0u);
    const int qp = qprev[id];
    for (int h = (q + 1) / 2; h < (qp + 1) / 2; h++) xrow[4 + h] = make_uint4(0u, 0u, 0u, 0u);
    qprev[id] = (uint8_t)q;
    // the list ends at the first -1
    if (q & 1) nb2[q / 2] = make_int2(jprev, -1);
    else if (q < kNb) nb2[q / 2] = make_int2(-1, -1);
}

// Candidate positions x' = x + displacement (then the periodic wrap of
// walls!, src/integration.jl:309-324).  Advances the step counter by k ahead
// of k_commit, which draws at counter - k and takes the advance back when
// the step is not committed.
__global__ void k_apply(SysParams P, LearnParams L, const double2 *__restrict__ Pst, const int32_t *__restrict__ ID,
                        const double *__restrict__ pol, const double2 *__restrict__ euler,
                        const float *__restrict__ Y, double2 *__restrict__ P2, double2 *__restrict__ disp,
                        unsigned long long *__restrict__ counter, uint32_t *__restrict__ gate) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) counter[0] += (unsigned long long)L.k;
    if (s >= P.na) return;
    const int id = ID[s];
    const int ring = id / P.nmax;
    double sn, cs;
    sincos(pol[ring], &sn, &cs);
    double2 d = euler[id];
    if (L.use_net) {
        const double2 y = make_double2((double)Y[2 * (int64_t)id] * L.s_t, (double)Y[2 * (int64_t)id + 1] * L.s_t);
        const double2 c = rot_back(cs, sn, y);
        d.x += c.x;
        d.y += c.y;
    }
    const double d2 = d.x * d.x + d.y * d.y;
    uint32_t flag = 0;
    if (!(d2 == d2) || isinf(d2)) flag |= GATE_NONFINITE;
    else if (d2 > L.max_disp2) flag |= GATE_DISP;
    if (flag) {
        atomicOr(gate, flag);
        d = make_double2(0.0, 0.0);
    }
    const double2 x = Pst[s];
    double px = x.x + d.x, py = x.y + d.y;
    if (P.periodic) {
        const double dxc = px - (P.blx + P.halfL), dyc = py - (P.bly + P.halfH);
        if (fabs(dxc) > P.halfL) px = px - copysign(P.halfL * 2.0, dxc);
        if (fabs(dyc) > P.halfH) py = py - copysign(P.halfH * 2.0, dyc);
    }
    P2[s] = make_double2(px, py);
    disp[id] = d;
}

// Overlap gate over the listed pairs (each pair once, from its smaller slot:
// a list holds every particle within r_feat, so the two particles list each
// other, and more than 32 fails the gate anyway; a list ends at the first -1
// or after 32 entries).
__device__ __forceinline__ void gate_pairs(const SysParams &P, const LearnParams &L, int s,
                                           const int32_t *__restrict__ ID, const int32_t *__restrict__ nbr,
                                           const double2 *__restrict__ P2, uint32_t *__restrict__ gate) {
    if (s >= P.na) return;
    const int id = ID[s];
    const double2 xi = P2[s];
    const int4 *nb4 = reinterpret_cast<const int4 *>(nbr + (int64_t)id * kNb);
    bool bad = false, done = false;
    for (int q4 = 0; q4 < kNb / 4 && !done; q4++) {
        const int4 v = nb4[q4];
        const int js[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int j = js[u];
            done = done || j < 0;
            if (done) continue;
            if (j < s) continue;  // the pair is tested from its smaller slot
            const double2 xj = P2[j];
            const double dx = min_image(xj.x - xi.x, P.L, P.halfL);
            const double dy = min_image(xj.y - xi.y, P.H, P.halfH);
            if (dx * dx + dy * dy < L.dmin2) bad = true;
        }
    }
    if (bad) atomicOr(gate, GATE_OVERLAP);
}

// Spring + area energy of ring rr from positions X (32 lanes, lane i =
// particle i): unwrap by an inclusive scan of the minimum-image steps, area
// by the shoelace sum, both as tree operations over the half-wave.
__device__ __forceinline__ double ring_energy(const SysParams &P, const RingType &T, const int32_t *slot_of, int rr,
                                              int i, const double2 *__restrict__ X) {
    const int np = T.np;
    const bool act = i < np;
    const double2 x = act ? X[slot_of[(int64_t)rr * P.nmax + i]] : make_double2(0.0, 0.0);
    const int inext = (i == np - 1) ? 0 : i + 1;
    const int iprev = (i == 0) ? np - 1 : i - 1;
    const double2 xn = make_double2(__shfl(x.x, inext, 32), __shfl(x.y, inext, 32));
    const double2 xp = make_double2(__shfl(x.x, iprev, 32), __shfl(x.y, iprev, 32));
    const double2 ds = cdiff(P, x, xn);
    const double dl = sqrt(ds.x * ds.x + ds.y * ds.y) - T.lspring;
    const double es = act ? 0.5 * T.kspring * dl * dl : 0.0;
    // c_i = x_0 + sum_{k=1..i} calc_diff(x_k, x_{k-1})
    const double2 d0 = (i == 0) ? x : (act ? cdiff(P, x, xp) : make_double2(0.0, 0.0));
    double cx = d0.x, cy = d0.y;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const double tx = __shfl_up(cx, o, 32), ty = __shfl_up(cy, o, 32);
        if (i >= o) {
            cx += tx;
            cy += ty;
        }
    }
    const double cnx = __shfl(cx, inext, 32), cny = __shfl(cy, inext, 32);
    const double area = half_sum(act ? cx * cny - cy * cnx : 0.0) / 2.0;
    double area0 = (double)np * T.lspring / T.p0;
    area0 = area0 * area0;
    return half_sum(es) + 0.5 * T.karea * (area - area0) * (area - area0);
}

// The gates of the candidate positions in one launch: workgroups [0, nbp)
// test the listed pairs for overlap (one thread per slot), the rest the
// spring + area energy of each ring (32 lanes per ring).
__global__ __launch_bounds__(256) void k_gate(SysParams P, LearnParams L, int nbp, const int32_t *__restrict__ ID,
                                              const int32_t *__restrict__ nbr, const int32_t *__restrict__ rtype,
                                              const RingType *__restrict__ types,
                                              const int32_t *__restrict__ slot_of,
                                              const double2 *__restrict__ Pst, const double2 *__restrict__ P2,
                                              uint32_t *__restrict__ gate) {
    if ((int)blockIdx.x < nbp) {
        gate_pairs(P, L, blockIdx.x * blockDim.x + threadIdx.x, ID, nbr, P2, gate);
        return;
    }
    const int gt = (blockIdx.x - nbp) * blockDim.x + threadIdx.x;
    const int r = gt >> 5;
    const bool ring_ok = r < P.nrings;
    const int rr = ring_ok ? r : 0;
    const RingType T = types[P.typed ? rtype[rr] : 0];
    const int i = gt & 31;
    const double e0 = ring_energy(P, T, slot_of, rr, i, Pst);
    const double e1 = ring_energy(P, T, slot_of, rr, i, P2);
    if (ring_ok && i == 0 && !(e1 - e0 <= L.e_tol)) atomicOr(gate, GATE_ENERGY);
}

// Accepted step: positions (blocks [0, nbp), one thread per slot) and
// polarisations (the remaining blocks, one thread per ring; one normal per
// ring from the device stream at the step counter before k_apply's advance),
// and ctl[1] (accepted steps) up by one.  A rejected step stores its gate
// word in ctl[0]; while ctl[0] is set, later steps of the same batch commit
// nothing (the host runs the plain fallback there).  A step not committed
// takes k_apply's counter advance back (one thread; no workgroup reads the
// counter then).
__device__ __forceinline__ void commit_pol(const SysParams &P, const LearnParams &L, int r,
                                           const int32_t *__restrict__ rtype, const RingType *__restrict__ types,
                                           const double2 *__restrict__ disp, double *__restrict__ pol,
                                           unsigned long long step) {
    const RingType T = types[P.typed ? rtype[r] : 0];
    double vx = 0.0, vy = 0.0;
    for (int i = 0; i < T.np; i++) {
        const double2 d = disp[(int64_t)r * P.nmax + i];
        vx += d.x;
        vy += d.y;
    }
    vx = vx / ((double)T.np * L.kdt);
    vy = vy / ((double)T.np * L.kdt);
    const double theta = pol[r];
    double sn, cs;
    sincos(theta, &sn, &cs);
    const double speed = sqrt(vx * vx + vy * vy);
    double cr = 0.0;
    if (speed != 0.0) {
        cr = (cs * vy - sn * vx) / speed;
        if (fabs(cr) > 1.0) cr = cr > 0.0 ? 1.0 : -1.0;
    }
    const double z = T.rot != 0.0 ? philox_normal(L.seed, step, (uint32_t)r) : 0.0;
    pol[r] = theta + (L.kdt / T.relax * asin(cr) + sqrt(2.0 * T.rot * L.kdt) * z);
}

__global__ __launch_bounds__(256) void k_commit(SysParams P, LearnParams L, int nbp,
                                                const int32_t *__restrict__ rtype,
                                                const RingType *__restrict__ types, const double2 *__restrict__ P2,
                                                double2 *__restrict__ Pst, const double2 *__restrict__ disp,
                                                double *__restrict__ pol, unsigned long long *__restrict__ counter,
                                                uint32_t *__restrict__ ctl, const uint32_t *__restrict__ gate) {
    const bool first = blockIdx.x == 0 && threadIdx.x == 0;
    if (ctl[0] != 0) {  // an earlier step of this batch was rejected
        if (first) counter[0] -= (unsigned long long)L.k;
        return;
    }
    const uint32_t g = gate[0];
    if (g != 0) {  // one word: every workgroup leaves here
        if (first) {
            ctl[0] = g;
            counter[0] -= (unsigned long long)L.k;
        }
        return;
    }
    if (first) ctl[1] += 1u;
    if ((int)blockIdx.x < nbp) {
        const int s = blockIdx.x * blockDim.x + threadIdx.x;
        if (s < P.na) Pst[s] = P2[s];
    } else {
        const int r = (blockIdx.x - nbp) * blockDim.x + threadIdx.x;
        if (r < P.nrings) commit_pol(P, L, r, rtype, types, disp, pol, counter[0] - (unsigned long long)L.k);
    }
}

// Data collection: positions by id at the snapshot, then targets after k steps.
__global__ void k_snapshot(int32_t na, const int32_t *__restrict__ ID, const double2 *__restrict__ Pst,
                           double2 *__restrict__ pos0) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= na) return;
    pos0[ID[s]] = Pst[s];
}

__global__ void k_targets(SysParams P, LearnParams L, const int32_t *__restrict__ slot_of,
                          const double2 *__restrict__ Pst, const double2 *__restrict__ pos0,
                          const double *__restrict__ pol0, const double2 *__restrict__ euler, float *__restrict__ T) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= P.na) return;
    const double2 x1 = Pst[slot_of[id]];
    const double2 x0 = pos0[id];
    const double dx = min_image(x1.x - x0.x, P.L, P.halfL), dy = min_image(x1.y - x0.y, P.H, P.halfH);
    const double2 e = euler[id];
    double sn, cs;
    sincos(pol0[id / P.nmax], &sn, &cs);
    const double2 r = rot(cs, sn, make_double2(dx - e.x, dy - e.y));
    T[2 * (int64_t)id] = (float)(r.x / L.s_t);
    T[2 * (int64_t)id + 1] = (float)(r.y / L.s_t);
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// rows < n: row j is particle j; otherwise rows drawn by (seed, snap, j).
__global__ void k_append_rows(const uint16_t *__restrict__ X, const float *__restrict__ T, int32_t n, int32_t rows,
                              uint64_t seed, int32_t snap, uint16_t *__restrict__ dsX, float *__restrict__ dsT) {
    constexpr int kChunks = kMlpIn / 8;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)rows * kChunks) return;
    const int j = (int)(q / kChunks), c = (int)(q % kChunks);
    const int src = rows >= n ? j : (int)(mix64(seed ^ mix64(((uint64_t)snap << 32) ^ (uint64_t)j)) % (uint64_t)n);
    *reinterpret_cast<uint4 *>(dsX + (int64_t)j * kMlpIn + c * 8) =
        *reinterpret_cast<const uint4 *>(X + (int64_t)src * kMlpIn + c * 8);
    if (c == 0) {
        dsT[2 * (int64_t)j] = T[2 * (int64_t)src];
        dsT[2 * (int64_t)j + 1] = T[2 * (int64_t)src + 1];
    }
}

}  // namespace

struct mavi_learned {
    mavi_sys *s = nullptr;
    mavi_mlp *mlp = nullptr;
    bool held = false;  // holds references to s and mlp (set once creation succeeded)
    LearnParams L{};
    uint16_t *X = nullptr;
    float *Y = nullptr, *T = nullptr;
    int32_t *nbr = nullptr;
    uint8_t *qprev = nullptr;  // per particle id: neighbour slots its feature row holds (past them: zeros)
    double2 *euler = nullptr, *disp = nullptr, *pos0 = nullptr;
    double *pol0 = nullptr;
    double *rate = nullptr, *omega = nullptr;
    double2 *U = nullptr;
    uint32_t *gate = nullptr;
    uint32_t *ctl = nullptr;        // per batch of learned steps: [0] first rejection's gate word, [1] accepted
    uint32_t *ctl_host = nullptr;   // pinned copy of ctl
    int batch = 8;                  // learned steps per host check (halved after a rejection, up to 16)
    uint32_t *gate_host = nullptr;  // pinned
    hipGraphExec_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [start copy][use_net]
    int64_t accepted = 0, rejected = 0;
    int64_t reasons[5] = {0, 0, 0, 0, 0};  // rejections by gate bit
};

namespace {

hipError_t launch_features(mavi_learned *l, hipStream_t st) {
    mavi_sys *s = l->s;
    // forces! with the per-ring inputs (u, area rate, alignment) from the
    // same ring pass (nmax <= 32 packs rings into waves), else k_ring_rates
    RingRates rr;
    rr.U = l->U;
    rr.rate = l->rate;
    rr.omega = l->omega;
    rr.gate = l->gate;
    const bool fused = s->P.nmax <= 32;
    hipError_t e = mavi::launch_rings_forces(s, st, fused ? &rr : nullptr);
    if (e != hipSuccess) return e;
    const int c = s->cur;
    if (!fused)
        hipLaunchKernelGGL(k_ring_rates, dim3(nblk((int64_t)s->P.nrings * 32, 256)), dim3(256), 0, st, s->P,
                           s->slot_of, s->P_[c], s->F, l->U, s->pol, s->rtype, s->types, l->rate, l->omega, l->gate);
    hipLaunchKernelGGL(k_features, dim3(nblk(s->P.na, kFeatWG)), dim3(kFeatWG), 0, st, s->P, l->L, s->P_[c],
                       s->ID_[c], s->slot_of, s->scell, s->start, l->U, s->pol, s->areas, s->rtype, s->types, l->rate,
                       l->omega, l->X, l->nbr, l->qprev, l->euler, l->gate);
    return hipGetLastError();
}

// Everything of a learned step up to and including the predicated commit.
hipError_t launch_learned_step(mavi_learned *l, hipStream_t st) {
    mavi_sys *s = l->s;
    hipError_t e = launch_features(l, st);
    if (e != hipSuccess) return e;
    if (l->L.use_net && (e = mavi::mlp_forward_fused(l->mlp, l->X, s->P.na, l->Y, st)) != hipSuccess) return e;
    const int c = s->cur;
    const int nb = nblk(s->P.na, 256);
    hipLaunchKernelGGL(k_apply, dim3(nb), dim3(256), 0, st, s->P, l->L, s->P_[c], s->ID_[c], s->pol, l->euler, l->Y,
                       s->P2, l->disp, s->counter, l->gate);
    hipLaunchKernelGGL(k_gate, dim3(nb + nblk((int64_t)s->P.nrings * 32, 256)), dim3(256), 0, st, s->P, l->L, nb,
                       s->ID_[c], l->nbr, s->rtype, s->types, s->slot_of, s->P_[c], s->P2, l->gate);
    hipLaunchKernelGGL(k_commit, dim3(nb + nblk(s->P.nrings, 256)), dim3(256), 0, st, s->P, l->L, nb, s->rtype,
                       s->types, s->P2, s->P_[c], l->disp, s->pol, s->counter, l->ctl, l->gate);
    return hipGetLastError();
}

hipError_t learned_graph(mavi_learned *l, hipStream_t st, hipGraphExec_t *out) {
    mavi_sys *s = l->s;
    const int cur0 = s->cur;
    const bool prof = s->prof;
    s->prof = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        s->prof = prof;
        return e;
    }
    hipError_t e1 = launch_learned_step(l, st);
    e = hipStreamEndCapture(st, &g);
    s->cur = cur0;
    s->prof = prof;
    if (e == hipSuccess && e1 != hipSuccess) e = e1;
    if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    return e;
}

int check_learned(mavi_learned *l) {
    if (!l || !l->s) return MAVI_E_BADARG;
    if (!l->s->has_state) return set_err(l->s->ctx, MAVI_E_BADARG, "no state uploaded");
    return MAVI_OK;
}

}  // namespace

extern "C" {

int mavi_learned_create(mavi_sys *s, mavi_mlp *mlp, const mavi_learned_desc *d, mavi_learned **out) {
    if (!s || !d || !out) return MAVI_E_BADARG;
    *out = nullptr;
    mavi_ctx *ctx = s->ctx;
    const SysParams &P = s->P;
    if (P.potential != MAVI_POT_RINGS || !P.periodic || P.cols < 3 || P.rows < 3 || P.nmax > 32)
        return set_err(ctx, MAVI_E_UNSUPPORTED,
                       "the learned propagator runs periodic rings systems with chunks (>= 3x3) and nmax <= 32");
    if (P.na != P.n) return set_err(ctx, MAVI_E_UNSUPPORTED, "the learned propagator needs full rings (np == nmax)");
    if (P.nwalls > 0) return set_err(ctx, MAVI_E_UNSUPPORTED, "the learned propagator runs without a walls list");
    if (d->k < 1 || !(d->r_feat > 0.0) || !(d->target_scale > 0.0))
        return set_err(ctx, MAVI_E_BADARG, "learned descriptor: k >= 1, r_feat > 0, target_scale > 0");
    if (P.cl < d->r_feat || P.ch < d->r_feat)
        return set_err(ctx, MAVI_E_BADARG, "chunks must be at least r_feat wide (the feature gather is 3x3)");
    if (mlp && mlp->ctx != ctx) return set_err(ctx, MAVI_E_BADARG, "network and system on different contexts");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    // defaults of the gate
    double dmin = d->d_min;
    if (!(dmin > 0.0)) {
        std::vector<RingInter> hi((size_t)P.ntypes * P.ntypes);
        HIPCHK(ctx, hipMemcpy(hi.data(), s->inter, sizeof(RingInter) * hi.size(), hipMemcpyDeviceToHost));
        double m = INFINITY;
        for (const RingInter &I : hi) m = fmin(m, I.dist_eq);
        dmin = 0.3 * m;  // the plain dynamics of dense active rings reaches ~0.4 dist_eq in jams
    }
    double maxd = d->max_disp;
    if (!(maxd > 0.0)) maxd = 0.5 * (d->r_feat - dmin);
    if (!(maxd > 0.0) || maxd > 0.5 * (d->r_feat - dmin))
        return set_err(ctx, MAVI_E_BADARG, "max_disp must be in (0, (r_feat - d_min)/2]");
    mavi_learned *l = new mavi_learned();
    l->s = s;
    l->mlp = mlp;
    LearnParams &L = l->L;
    L.k = d->k;
    L.use_net = mlp != nullptr;
    L.rf2 = exact_gt_threshold(d->r_feat);
    L.kdt = (double)d->k * P.dt;
    L.s_t = d->target_scale;
    L.dmin2 = exact_lt_threshold(dmin);
    L.max_disp2 = maxd * maxd;
    L.e_tol = d->e_tol > 0.0 ? d->e_tol : 1.0;
    L.seed = d->seed;
    const int64_t n = P.n;
    hipStream_t st = ctx->stream;
    hipError_t e = dalloc(&l->X, n * kMlpIn, st);
    if (e == hipSuccess) e = dalloc(&l->Y, n * kMlpOut, st);
    if (e == hipSuccess) e = dalloc(&l->T, n * kMlpOut, st);
    if (e == hipSuccess) e = dalloc(&l->nbr, n * kNb, st);
    if (e == hipSuccess) e = dalloc(&l->qprev, n, st);
    if (e == hipSuccess) e = dalloc(&l->euler, n, st);
    if (e == hipSuccess) e = dalloc(&l->disp, n, st);
    if (e == hipSuccess) e = dalloc(&l->pos0, n, st);
    if (e == hipSuccess) e = dalloc(&l->pol0, P.nrings, st);
    if (e == hipSuccess) e = dalloc(&l->rate, P.nrings, st);
    if (e == hipSuccess) e = dalloc(&l->U, n, st);
    if (e == hipSuccess) e = dalloc(&l->omega, P.nrings, st);
    if (e == hipSuccess) e = dalloc(&l->gate, 1, st);
    if (e == hipSuccess) e = dalloc(&l->ctl, 2, st);
    if (e == hipSuccess) e = hipHostMalloc((void **)&l->ctl_host, 2 * sizeof(uint32_t), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&l->gate_host, sizeof(uint32_t), hipHostMallocDefault);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        std::string m = std::string("learned allocation failed: ") + hipGetErrorString(e);
        mavi_learned_destroy(l);
        return set_err(ctx, MAVI_E_HIP, m);
    }
    sys_retain(s);
    if (mlp) mlp_retain(mlp);
    l->held = true;
    *out = l;
    return MAVI_OK;
}

int mavi_learned_destroy(mavi_learned *l) {
    if (!l) return MAVI_OK;
    (void)hipSetDevice(l->s->ctx->device);
    (void)hipStreamSynchronize(l->s->ctx->stream);
    for (auto &gr : l->graph)
        for (auto &g : gr)
            if (g) (void)hipGraphExecDestroy(g);
    dfree(l->X); dfree(l->Y); dfree(l->T); dfree(l->nbr); dfree(l->qprev); dfree(l->euler); dfree(l->disp); dfree(l->pos0);
    dfree(l->pol0); dfree(l->rate); dfree(l->omega); dfree(l->U); dfree(l->gate); dfree(l->ctl);
    if (l->gate_host) (void)hipHostFree(l->gate_host);
    if (l->ctl_host) (void)hipHostFree(l->ctl_host);
    mavi_sys *s = l->s;
    mavi_mlp *mlp = l->mlp;
    const bool held = l->held;
    delete l;
    if (held) {
        if (mlp) mlp_release(mlp);
        sys_release(s);
    }
    return MAVI_OK;
}

int mavi_learned_features(mavi_learned *l, uint16_t *x, double *euler) {
    int rc = check_learned(l);
    if (rc != MAVI_OK) return rc;
    mavi_sys *s = l->s;
    mavi_ctx *ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    HIPCHK(ctx, launch_features(l, st));
    HIPCHK(ctx, mavi::launch_unpermute(s, true, true, st));
    const int64_t n = s->P.n;
    if (x) HIPCHK(ctx, hipMemcpyAsync(x, l->X, sizeof(uint16_t) * n * kMlpIn, hipMemcpyDeviceToHost, st));
    if (euler) HIPCHK(ctx, hipMemcpyAsync(euler, l->euler, sizeof(double2) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipMemcpyAsync(l->gate_host, l->gate, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    if (*l->gate_host & GATE_NEIGH)
        return set_err(ctx, MAVI_E_BADARG, "a particle has more than 32 neighbours within r_feat");
    return MAVI_OK;
}

int mavi_learned_collect(mavi_learned *l, int32_t nsnap, int32_t stride, int32_t rows_per_snap, uint64_t seed) {
    int rc = check_learned(l);
    if (rc != MAVI_OK) return rc;
    mavi_sys *s = l->s;
    mavi_ctx *ctx = s->ctx;
    if (!l->mlp) return set_err(ctx, MAVI_E_BADARG, "collect needs a network (its data set receives the rows)");
    if (nsnap < 0 || stride < 0 || rows_per_snap <= 0) return MAVI_E_BADARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    mavi_mlp *p = l->mlp;
    const int n = s->P.n;
    const int rows = rows_per_snap >= n ? n : rows_per_snap;
    HIPCHK(ctx, mavi::mlp_dataset_reserve(p, p->ds_count + (int64_t)nsnap * rows));
    for (int q = 0; q < nsnap; q++) {
        HIPCHK(ctx, launch_features(l, st));
        const int c = s->cur;
        hipLaunchKernelGGL(k_snapshot, dim3(nblk(n, 256)), dim3(256), 0, st, n, s->ID_[c], s->P_[c], l->pos0);
        HIPCHK(ctx, hipMemcpyAsync(l->pol0, s->pol, sizeof(double) * s->P.nrings, hipMemcpyDeviceToDevice, st));
        for (int k = 0; k < l->L.k; k++) HIPCHK(ctx, mavi::launch_rings_step(s, nullptr, st));
        hipLaunchKernelGGL(k_targets, dim3(nblk(n, 256)), dim3(256), 0, st, s->P, l->L, s->slot_of, s->P_[s->cur],
                           l->pos0, l->pol0, l->euler, l->T);
        const int64_t nchunk = (int64_t)rows * (kMlpIn / 8);
        hipLaunchKernelGGL(k_append_rows, dim3((unsigned)((nchunk + 255) / 256)), dim3(256), 0, st, l->X, l->T, n,
                           rows, seed, q, p->dsX + p->ds_count * kMlpIn, p->dsT + p->ds_count * kMlpOut);
        HIPCHK(ctx, hipGetLastError());
        p->ds_count += rows;
        for (int k = 0; k < stride; k++) HIPCHK(ctx, mavi::launch_rings_step(s, nullptr, st));
    }
    HIPCHK(ctx, mavi::launch_unpermute(s, true, true, st));
    HIPCHK(ctx, hipMemcpyAsync(l->gate_host, l->gate, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    if (*l->gate_host & GATE_NEIGH)
        return set_err(ctx, MAVI_E_BADARG, "a particle has more than 32 neighbours within r_feat");
    return MAVI_OK;
}

int mavi_learned_run(mavi_learned *l, int64_t nmacro, int32_t use_net, int64_t *accepted) {
    int rc = check_learned(l);
    if (rc != MAVI_OK) return rc;
    mavi_sys *s = l->s;
    mavi_ctx *ctx = s->ctx;
    if (nmacro < 0) return MAVI_E_BADARG;
    if (use_net && !l->mlp) return set_err(ctx, MAVI_E_BADARG, "no network attached");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    l->L.use_net = use_net != 0;
    const int un = use_net ? 1 : 0;
    // Learned steps run in batches with one read back each: a rejected step
    // leaves its gate word in ctl[0] and the batch's later steps commit
    // nothing, so the plain fallback runs exactly where the step-by-step
    // loop (src/rings/integration.jl step! x k on rejection) would run it.
    int64_t acc = 0, q = 0;
    while (q < nmacro) {
        const int64_t m = std::min<int64_t>(l->batch, nmacro - q);
        // a learned step moves the particles outside the kernels that keep
        // the chunk keys: its binning starts from scratch (also when captured)
        HIPCHK(ctx, mavi::drop_keys(s, st));
        HIPCHK(ctx, mavi::launch_fill_i32((int32_t *)l->ctl, 0, 2, st));
        for (int64_t j = 0; j < m; j++) {
            if (s->use_graphs && !s->prof) {
                hipGraphExec_t &g = l->graph[s->cur][un];
                if (!g) HIPCHK(ctx, learned_graph(l, st, &g));
                HIPCHK(ctx, hipGraphLaunch(g, st));
                s->cur = 1 - s->cur;  // the step's binning
            } else {
                HIPCHK(ctx, launch_learned_step(l, st));
            }
        }
        HIPCHK(ctx, hipMemcpyAsync(l->ctl_host, l->ctl, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
        const uint32_t gw = l->ctl_host[0];
        const int64_t na = (int64_t)l->ctl_host[1];
        acc += na;
        q += na;
        if (gw == 0) {
            l->batch = std::min(16, 2 * l->batch);
            continue;
        }
        for (int b = 0; b < 5; b++)
            if (gw & (1u << b)) l->reasons[b]++;
        for (int k = 0; k < l->L.k; k++) HIPCHK(ctx, mavi::launch_rings_step(s, nullptr, st));
        q += 1;
        l->batch = std::max(1, l->batch / 2);
    }
    l->accepted += acc;
    l->rejected += nmacro - acc;
    if (accepted) *accepted = acc;
    HIPCHK(ctx, mavi::launch_unpermute(s, true, true, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    return MAVI_OK;
}

int mavi_learned_stats(mavi_learned *l, int64_t *stats) {
    if (!l || !stats) return MAVI_E_BADARG;
    stats[0] = l->accepted;
    stats[1] = l->rejected;
    for (int b = 0; b < 5; b++) stats[2 + b] = l->reasons[b];
    return MAVI_OK;
}

}  // extern "C"
